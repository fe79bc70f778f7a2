// C-ABI layer of natrium_amd: context set-up, operator assembly dispatch,
// population upload/download and the step loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "nat_internal.hpp"

// kernel launchers (k_*.hip)
int nat_launch_assemble_general(nat_ctx* ctx, const double* d_disp, const double* d_dirw, int* d_cell, double* d_xi,
                                unsigned* d_flipmask, double* d_bias);
int nat_launch_support_points(nat_ctx* ctx);
int nat_launch_wall_compact(nat_ctx* ctx, const int* d_list, const unsigned* d_flipmask, const double* d_bias);
int nat_launch_assemble_axis(nat_ctx* ctx, const int* d_N, const double* d_vaxis, int maxc, const double* d_disp,
                             const int* d_nshift, double* d_w, int* d_off);
int nat_launch_records_f32(nat_ctx* ctx, int64_t nrec);
int nat_launch_sl_struct(nat_ctx* ctx, int write_macro);
int nat_launch_sl_struct_tile(nat_ctx* ctx, int* tile);
int nat_internal_pack_key(nat_ctx* ctx, int64_t* key);
int nat_internal_pack_step(nat_ctx* const* ctxs, int n, int nsteps, hipStream_t stream);
int nat_internal_autotune(nat_ctx* ctx, int nsteps, int* tile, double* best_us);
int nat_internal_autotune_general(nat_ctx* ctx, int nsteps, int* tile, double* best_us);
int nat_launch_sl_general(nat_ctx* ctx, int write_macro);
int nat_launch_collide(nat_ctx* ctx, int init);
int nat_launch_stabilize(nat_ctx* ctx);
int nat_launch_sums(nat_ctx* ctx, int layout);
int nat_launch_sums_noalloc(nat_ctx* ctx, int layout);
int nat_sums_scratch_setup(nat_ctx* ctx);
int nat_sedg_setup(nat_ctx* ctx);
int nat_sedg_step(nat_ctx* ctx, int collide_after, int write_macro);
int nat_comp_setup(nat_ctx* ctx);
int nat_launch_sl_comp(nat_ctx* ctx, int write_macro);
int nat_launch_collide_comp(nat_ctx* ctx);

namespace nat {

bool make_stencil(int id, double s, StencilHost& st) {
  st = StencilHost();
  st.id = id;
  st.s = s;
  auto put = [&](int i, double x, double y, double z) {
    st.dir[i][0] = x;
    st.dir[i][1] = y;
    st.dir[i][2] = z;
  };
  if (id == NAT_D2Q9 || id == NAT_D3Q19 || id == NAT_D3Q27) {
    int Q = id == NAT_D2Q9 ? 9 : (id == NAT_D3Q19 ? 19 : 27);
    st.D = id == NAT_D2Q9 ? 2 : 3;
    st.Q = Q;
    for (int i = 0; i < Q; i++) {
      const int* e = Q == 9 ? Lattice<9>::e[i] : (Q == 19 ? Lattice<19>::e[i] : Lattice<27>::e[i]);
      put(i, e[0] * s, e[1] * s, e[2] * s);
      st.w[i] = Q == 9 ? Lattice<9>::w[i] : (Q == 19 ? Lattice<19>::w[i] : Lattice<27>::w[i]);
    }
  } else if (id == NAT_D2Q25H) {
    // stencils/D2Q25H.cpp:42-82: tensor product of the 5-velocity Hermite set
    st.D = 2;
    st.Q = 25;
    const double r = (std::sqrt(5.) - std::sqrt(2.)) / std::sqrt(3.);
    const double w0 = (-3 - 3 * r * r * r * r + 54 * r * r) / (75 * r * r);
    const double wm = (9 * r * r * r * r - 6 - 27 * r * r) / (300 * r * r * (r * r - 1));
    const double wn = (9 - 6 * r * r * r * r - 27 * r * r) / (300 * (1 - r * r));
    const double cm = s * std::sqrt(5. - std::sqrt(10.)) / std::sqrt(3.);
    const double cn = std::sqrt(5. + std::sqrt(10.)) * s / std::sqrt(3.);
    const double c1[25][2] = {{0, 0},   {cm, 0},    {0, cm},   {-cm, 0},  {0, -cm},  {cm, cm},   {-cm, cm},
                              {-cm, -cm}, {cm, -cm}, {cn, 0},   {0, cn},   {-cn, 0},  {0, -cn},   {cn, cn},
                              {-cn, cn}, {-cn, -cn}, {cn, -cn}, {cm, cn},  {cm, -cn}, {-cm, -cn}, {-cm, cn},
                              {cn, cm},  {cn, -cm}, {-cn, -cm}, {-cn, cm}};
    const double ww[25] = {w0 * w0, w0 * wm, w0 * wm, w0 * wm, w0 * wm, wm * wm, wm * wm, wm * wm, wm * wm,
                           w0 * wn, w0 * wn, w0 * wn, w0 * wn, wn * wn, wn * wn, wn * wn, wn * wn, wm * wn,
                           wm * wn, wm * wn, wm * wn, wm * wn, wm * wn, wm * wn, wm * wn};
    for (int i = 0; i < 25; i++) {
      put(i, c1[i][0], c1[i][1], 0);
      st.w[i] = ww[i];
    }
  } else {
    return false;
  }
  st.cs2 = s * s / 3.;
  st.max_speed = s * std::sqrt(2.);
  return true;
}

void gll_points(int p, double* x, double* w) {
  // Gauss-Lobatto on [-1,1]: endpoints and the roots of P_p', by Newton from
  // the Chebyshev-Gauss-Lobatto points; mapped to [0,1].
  const int n = p + 1;
  std::vector<long double> xs(n), ws(n);
  for (int j = 0; j < n; j++) {
    long double t = -std::cos(3.14159265358979323846264338327950288L * j / p);
    auto leg = [&](long double z, long double& pn, long double& pm) {
      long double a = 1, b = z;
      if (p == 0) {
        pn = 1;
        pm = 0;
        return;
      }
      for (int k = 2; k <= p; k++) {
        long double c = ((2 * k - 1) * z * b - (k - 1) * a) / k;
        a = b;
        b = c;
      }
      pn = b;
      pm = a;
    };
    if (j == 0) t = -1;
    if (j == p) t = 1;
    if (j > 0 && j < p)
      for (int it = 0; it < 100; it++) {
        long double pn, pm;
        leg(t, pn, pm);
        long double dt = (t * pn - pm) / ((p + 1) * pn);
        t -= dt;
        if (std::fabs((double)dt) < 1e-19) break;
      }
    long double pn, pm;
    leg(t, pn, pm);
    xs[j] = t;
    ws[j] = 2.0L / (p * (p + 1) * pn * pn);
  }
  for (int j = 0; j < n; j++) {
    long double a = (xs[j] + 1) / 2, b = (1 - xs[p - j]) / 2;
    x[j] = (double)((a + b) / 2);
    w[j] = (double)(ws[j] / 2);
  }
  x[0] = 0.0;
  x[p] = 1.0;
  if (p % 2 == 0) x[p / 2] = 0.5;
}

}  // namespace nat

using namespace nat;

namespace {

const double* vtx(const nat_ctx* c, int i, int j, int k) {
  return &c->verts[3 * ((size_t)i + (size_t)(c->desc.ncells[0] + 1) * ((size_t)j + (size_t)(c->desc.ncells[1] + 1) * k))];
}

void cell_vertex(const nat_ctx* c, const int cc[3], int v, double out[3]) {
  const double* p = vtx(c, cc[0] + (v & 1), cc[1] + ((v >> 1) & 1), c->D == 3 ? cc[2] + ((v >> 2) & 1) : 0);
  out[0] = p[0];
  out[1] = p[1];
  out[2] = p[2];
}

void map_cell(const nat_ctx* c, const int cc[3], const double xi[3], bool cartesian, double x[3]) {
  const int nv = c->D == 3 ? 8 : 4;
  if (cartesian) {
    double a[3], b[3];
    cell_vertex(c, cc, 0, a);
    cell_vertex(c, cc, nv - 1, b);
    for (int d = 0; d < 3; d++) x[d] = d < c->D ? a[d] + xi[d] * (b[d] - a[d]) : 0.0;
    return;
  }
  x[0] = x[1] = x[2] = 0;
  for (int v = 0; v < nv; v++) {
    double Nv = 1, X[3];
    for (int a = 0; a < c->D; a++) Nv *= ((v >> a) & 1) ? xi[a] : (1 - xi[a]);
    cell_vertex(c, cc, v, X);
    for (int d = 0; d < c->D; d++) x[d] += Nv * X[d];
  }
}

double min_vertex_distance(const nat_ctx* c) {
  double mind = 1e300;
  const int nv = c->D == 3 ? 8 : 4;
  const int nz = c->D == 3 ? c->desc.ncells[2] : 1;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < c->desc.ncells[1]; j++)
      for (int i = 0; i < c->desc.ncells[0]; i++) {
        int cc[3] = {i, j, k};
        double V[8][3];
        for (int v = 0; v < nv; v++) cell_vertex(c, cc, v, V[v]);
        for (int a = 0; a < nv; a++)
          for (int b = a + 1; b < nv; b++) {
            double s = 0;
            for (int d = 0; d < c->D; d++) s += (V[a][d] - V[b][d]) * (V[a][d] - V[b][d]);
            mind = std::min(mind, std::sqrt(s));
          }
      }
  return mind;
}

// uniform axis-aligned lattice (a prerequisite of the structured backend)
bool is_uniform_cartesian(const nat_ctx* c) {
  const int D = c->D;
  const int N[3] = {c->desc.ncells[0], c->desc.ncells[1], D == 3 ? c->desc.ncells[2] : 1};
  for (int a = 0; a < D; a++) {
    const double h0 = vtx(c, a == 0 ? 1 : 0, a == 1 ? 1 : 0, a == 2 ? 1 : 0)[a] - vtx(c, 0, 0, 0)[a];
    if (!(h0 > 0)) return false;
    for (int k = 0; k <= (D == 3 ? N[2] : 0); k++)
      for (int j = 0; j <= N[1]; j++)
        for (int i = 0; i <= N[0]; i++) {
          const double* p = vtx(c, i, j, k);
          const int idx[3] = {i, j, k};
          // axis a coordinate depends only on idx[a], with spacing h0
          const double xa = p[a];
          const double ref = vtx(c, a == 0 ? i : 0, a == 1 ? j : 0, a == 2 ? k : 0)[a];
          if (std::fabs(xa - ref) > 1e-12 * std::fabs(h0)) return false;
          if (idx[a] > 0) {
            const double prev = vtx(c, a == 0 ? i - 1 : 0, a == 1 ? j - 1 : 0, a == 2 ? k - 1 : 0)[a];
            if (std::fabs((ref - prev) - h0) > 1e-10 * h0) return false;
          }
        }
  }
  return true;
}

// every cell an axis-aligned box (what MappingCartesian assumes)
bool is_axis_aligned(const nat_ctx* c) {
  const int nv = c->D == 3 ? 8 : 4;
  const int nz = c->D == 3 ? c->desc.ncells[2] : 1;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < c->desc.ncells[1]; j++)
      for (int i = 0; i < c->desc.ncells[0]; i++) {
        int cc[3] = {i, j, k};
        double lo[3], hi[3], X[3];
        cell_vertex(c, cc, 0, lo);
        cell_vertex(c, cc, nv - 1, hi);
        for (int v = 0; v < nv; v++) {
          cell_vertex(c, cc, v, X);
          for (int a = 0; a < c->D; a++) {
            const double want = ((v >> a) & 1) ? hi[a] : lo[a];
            if (std::fabs(X[a] - want) > 1e-12 * (std::fabs(hi[a] - lo[a]) + 1e-300)) return false;
          }
        }
        for (int a = 0; a < c->D; a++)
          if (!(hi[a] > lo[a])) return false;
      }
  return true;
}

template <typename T>
int alloc(nat_ctx* ctx, T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) {
    ctx->err = std::string("hipMalloc: ") + hipGetErrorString(e);
    return NAT_ERR_NOMEM;
  }
  return 0;
}

int fail(nat_ctx* ctx, int code, const std::string& msg) {
  ctx->err = msg;
  return code;
}

#define TRY(x)              \
  do {                      \
    int _rc = (x);          \
    if (_rc) return _rc;    \
  } while (0)

int check_status(nat_ctx* ctx) {
  // the copy lands in pinned host memory: no staging buffer, one wait for the whole call
  NAT_HIP(hipMemcpyAsync(ctx->h_status, ctx->d_status, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  NAT_HIP(hipStreamSynchronize(ctx->stream));
  const int st = *ctx->h_status;
  if (st) NAT_HIP(hipMemsetAsync(ctx->d_status, 0, sizeof(int), ctx->stream));
  if (st & 1) return fail(ctx, NAT_ERR_DENSITY, "Densities too small (< 1e-10) for collisions. Decrease time step size.");
  if (st & 2) return fail(ctx, NAT_ERR_UNSUPPORTED, "departure path reached a wall but no wall table was built");
  if (st & 4) return fail(ctx, NAT_ERR_ARG, "departure point could not be located in its cell");
  if (st & 16) return fail(ctx, NAT_ERR_ARG, "structured departure walk left the axis");
  return 0;
}

// DG nodes: cell-major, local lexicographic
void dg_points(const nat_ctx* c, std::vector<double>& xyz) {
  double gll[MAXP + 1], gw[MAXP + 1];
  gll_points(c->p, gll, gw);
  const int np = c->p + 1, nl = c->D == 3 ? np * np * np : np * np;
  xyz.assign(3 * c->n_dofs, 0.0);
  const int nz = c->D == 3 ? c->desc.ncells[2] : 1;
  int64_t t = 0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < c->desc.ncells[1]; j++)
      for (int i = 0; i < c->desc.ncells[0]; i++) {
        int cc[3] = {i, j, k};
        for (int l = 0; l < nl; l++, t++) {
          const double xi[3] = {gll[l % np], gll[(l / np) % np], c->D == 3 ? gll[l / (np * np)] : 0.0};
          double x[3];
          map_cell(c, cc, xi, false, x);
          for (int d = 0; d < 3; d++) xyz[3 * t + d] = x[d];
        }
      }
}

// device storage index of FE_Q DoF (i,j,k)
inline int64_t store_index(const nat_ctx* c, int64_t dof) {
  if (c->backend != NAT_BACKEND_STRUCTURED || c->desc.scheme != NAT_SL) return dof;
  const int i = (int)(dof % c->nfe[0]);
  const int j = (int)((dof / c->nfe[0]) % c->nfe[1]);
  const int k = (int)(dof / ((int64_t)c->nfe[0] * c->nfe[1]));
  const int ii = i % c->nu[0], jj = j % c->nu[1], kk = k % c->nu[2];
  return ii + (int64_t)c->nu[0] * (jj + (int64_t)c->nu[1] * kk);
}

// Device scratch freed when the set-up function returns, on every path.
struct DevScratch {
  std::vector<void*> bufs;
  ~DevScratch() {
    for (void* b : bufs)
      if (b) hipFree(b);
  }
  template <typename T>
  int get(nat_ctx* ctx, T** p, size_t count) {
    int rc = alloc(ctx, p, count);
    if (!rc) bufs.push_back(*p);
    return rc;
  }
};

// Structured 1-D tables of a candidate operator. They are built beside the tables in
// use and only installed (commit_structured_tables) once every check has passed, so a
// time step or mesh the structured backend cannot take leaves the current operator intact.
struct StructTables {
  double* w = nullptr;  // [3][MAXSHIFT][MAXP][MAXP+1] FP64 weights
  int* off = nullptr;   // [3][MAXSHIFT][MAXP] source cell offsets
  bool sym = false;
  int off_lo[3][MAXSHIFT] = {}, off_hi[3][MAXSHIFT] = {};
  void release() {
    if (w) hipFree(w);
    if (off) hipFree(off);
    w = nullptr;
    off = nullptr;
  }
};

void commit_structured_tables(nat_ctx* ctx, StructTables& t) {
  if (ctx->ax_w) hipFree(ctx->ax_w);
  if (ctx->ax_off) hipFree(ctx->ax_off);
  ctx->ax_w = t.w;
  ctx->ax_off = t.off;
  ctx->ax_sym = t.sym;
  std::memcpy(ctx->off_lo, t.off_lo, sizeof(t.off_lo));
  std::memcpy(ctx->off_hi, t.off_hi, sizeof(t.off_hi));
  t.w = nullptr;
  t.off = nullptr;
}

// Structured 1-D tables on the GPU (FP64) for the context's current dt and vertices,
// then host checks: offsets within one cell layer and no tensor-product weight the
// reference would drop (|w| < 1e-10). *ok = false (and no tables) when a check fails.
int build_structured_tables(nat_ctx* ctx, StructTables& out, bool* ok) {
  *ok = false;
  const int D = ctx->D, p = ctx->p;
  // shift id 0: component +s (departure at -dt s), 1: component -s
  double disp[3 * MAXSHIFT] = {};
  int nshift[3] = {0, 0, 0};
  for (int a = 0; a < D; a++) {
    nshift[a] = 2;
    disp[a * MAXSHIFT + 0] = ctx->st.s * (-ctx->dt);
    disp[a * MAXSHIFT + 1] = -ctx->st.s * (-ctx->dt);
  }
  int maxc = 0;
  for (int a = 0; a < D; a++) maxc = std::max(maxc, ctx->desc.ncells[a]);
  std::vector<double> vaxis(3 * (maxc + 1), 0.0);
  for (int a = 0; a < D; a++)
    for (int i = 0; i <= ctx->desc.ncells[a]; i++)
      vaxis[a * (maxc + 1) + i] = vtx(ctx, a == 0 ? i : 0, a == 1 ? i : 0, a == 2 ? i : 0)[a];
  int N[3] = {ctx->desc.ncells[0], ctx->desc.ncells[1], D == 3 ? ctx->desc.ncells[2] : 1};
  const size_t nw = 3 * MAXSHIFT * MAXP * (MAXP + 1), no = 3 * MAXSHIFT * MAXP;
  std::vector<double> w(nw);
  std::vector<int> off(no);
  {
    DevScratch tmp;
    int *dN = nullptr, *dns = nullptr, *doff = nullptr;
    double *dva = nullptr, *ddisp = nullptr, *dw = nullptr;
    TRY(tmp.get(ctx, &dN, 3));
    TRY(tmp.get(ctx, &dns, 3));
    TRY(tmp.get(ctx, &doff, no));
    TRY(tmp.get(ctx, &dva, vaxis.size()));
    TRY(tmp.get(ctx, &ddisp, 3 * MAXSHIFT));
    TRY(tmp.get(ctx, &dw, nw));
    NAT_HIP(hipMemcpyAsync(dN, N, sizeof(N), hipMemcpyHostToDevice, ctx->stream));
    NAT_HIP(hipMemcpyAsync(dns, nshift, sizeof(nshift), hipMemcpyHostToDevice, ctx->stream));
    NAT_HIP(hipMemcpyAsync(dva, vaxis.data(), vaxis.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    NAT_HIP(hipMemcpyAsync(ddisp, disp, sizeof(disp), hipMemcpyHostToDevice, ctx->stream));
    NAT_HIP(hipMemsetAsync(dw, 0, nw * sizeof(double), ctx->stream));
    NAT_HIP(hipMemsetAsync(doff, 0, no * sizeof(int), ctx->stream));
    TRY(nat_launch_assemble_axis(ctx, dN, dva, maxc, ddisp, dns, dw, doff));
    NAT_HIP(hipMemcpyAsync(w.data(), dw, nw * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    NAT_HIP(hipMemcpyAsync(off.data(), doff, no * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    NAT_HIP(hipStreamSynchronize(ctx->stream));
  }
  int st = 0;
  NAT_HIP(hipMemcpy(&st, ctx->d_status, sizeof(int), hipMemcpyDeviceToHost));
  if (st) {
    NAT_HIP(hipMemset(ctx->d_status, 0, sizeof(int)));
    return 0;
  }
  auto W = [&](int a, int s, int j, int m) { return w[((a * MAXSHIFT + s) * MAXP + j) * (MAXP + 1) + m]; };
  StructTables t;
  for (int a = 0; a < D; a++)
    for (int s = 0; s < 2; s++) {
      int lo = 100, hi = -100;
      for (int j = 0; j < p; j++) {
        const int o = off[(a * MAXSHIFT + s) * MAXP + j];
        lo = std::min(lo, o);
        hi = std::max(hi, o);
      }
      if (lo < -1 || hi > 1 || hi - lo > 1) return 0;
      t.off_lo[a][s] = lo;
      t.off_hi[a][s] = hi;
      // the kernel's fixed window pattern: for a positive component only the
      // vertex node (j = 0) reads the cell below; otherwise every node reads its own cell
      for (int j = 0; j < p; j++) {
        const int o = off[(a * MAXSHIFT + s) * MAXP + j];
        const int want = (s == 0 && j == 0) ? -1 : 0;
        if (o != want) return 0;
      }
    }
  // cut-off check over all directions and node classes
  const int np = p + 1;
  for (int al = 1; al < ctx->Q; al++) {
    int sid[3] = {-1, -1, -1};
    for (int a = 0; a < D; a++) {
      const double c = ctx->st.dir[al][a];
      sid[a] = c > 0 ? 0 : (c < 0 ? 1 : -1);
    }
    const int jz_n = D == 3 ? p : 1, mz_n = D == 3 ? np : 1;
    for (int jz = 0; jz < jz_n; jz++)
      for (int jy = 0; jy < p; jy++)
        for (int jx = 0; jx < p; jx++)
          for (int mz = 0; mz < mz_n; mz++)
            for (int my = 0; my < np; my++)
              for (int mx = 0; mx < np; mx++) {
                const int jj[3] = {jx, jy, jz}, mm[3] = {mx, my, mz};
                double prod = 1;
                bool zero = false;
                for (int a = 0; a < D; a++) {
                  double v;
                  if (sid[a] < 0) {
                    if (mm[a] != jj[a]) zero = true;
                    v = 1.0;
                  } else {
                    v = W(a, sid[a], jj[a], mm[a]);
                  }
                  prod *= v;
                }
                if (zero || prod == 0) continue;
                if (std::fabs(prod) < 1e-10) return 0;
              }
  }
  // compact tables: every axis and both signs follow from the positive-shift table of
  // axis 0 by the GLL mirror symmetry W(a, -, j, m) = W(0, +, (p - j) mod p, p - m)
  {
    bool sym = true;
    for (int a = 0; a < D && sym; a++)
      for (int j = 0; j < p && sym; j++)
        for (int m = 0; m <= p && sym; m++) {
          const double wp = W(0, 0, j, m);
          const double tol = 1e-14 * (1.0 + std::fabs(wp));
          if (std::fabs(W(a, 0, j, m) - wp) > tol) sym = false;
          if (std::fabs(W(a, 1, (p - j) % p, p - m) - wp) > tol) sym = false;
        }
    t.sym = sym;
  }
  // FP64 weights for both storage precisions (the kernel interpolates in FP64)
  {
    int rc = alloc(ctx, &t.w, nw);
    if (!rc) rc = alloc(ctx, &t.off, no);
    if (rc) {
      t.release();
      return rc;
    }
    hipError_t e = hipMemcpy(t.w, w.data(), nw * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(t.off, off.data(), no * sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      t.release();
      return fail(ctx, NAT_ERR_HIP, std::string("structured tables: ") + hipGetErrorString(e));
    }
  }
  out = t;
  *ok = true;
  return 0;
}

bool structured_instantiated(int D, int Q, int p) {
  const bool pq = (p == 1 || p == 2 || p == 3 || p == 4 || p == 6);
  if (D == 3) return (Q == 19 || Q == 27) && pq;
  return Q == 9 && pq;
}

int assemble_general(nat_ctx* ctx) {
  const int Qm1 = ctx->Q - 1;
  std::vector<double> disp(3 * Qm1, 0.0), dirw(4 * Qm1, 0.0);
  for (int a = 1; a < ctx->Q; a++) {
    for (int d = 0; d < 3; d++) {
      disp[3 * (a - 1) + d] = ctx->st.dir[a][d] * (-ctx->dt);
      dirw[4 * (a - 1) + d] = ctx->st.dir[a][d];
    }
    dirw[4 * (a - 1) + 3] = ctx->st.w[a];
  }
  DevScratch tmp;
  double *ddisp, *ddirw, *dbias = nullptr;
  unsigned* dmask = nullptr;
  TRY(tmp.get(ctx, &ddisp, disp.size()));
  TRY(tmp.get(ctx, &ddirw, dirw.size()));
  NAT_HIP(hipMemcpyAsync(ddisp, disp.data(), disp.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  NAT_HIP(hipMemcpyAsync(ddirw, dirw.data(), dirw.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  const int64_t nrec = ctx->n_dofs * Qm1;
  if (!ctx->dep_cell) TRY(alloc(ctx, &ctx->dep_cell, nrec));
  if (!ctx->dep_xi64) TRY(alloc(ctx, &ctx->dep_xi64, nrec * ctx->D));
  if (ctx->has_walls) {
    TRY(tmp.get(ctx, &dmask, ctx->n_dofs));
    TRY(tmp.get(ctx, &dbias, nrec));
    NAT_HIP(hipMemsetAsync(dmask, 0, sizeof(unsigned) * ctx->n_dofs, ctx->stream));
  }
  TRY(nat_launch_assemble_general(ctx, ddisp, ddirw, ctx->dep_cell, ctx->dep_xi64, dmask, dbias));
  if (ctx->fp32) {
    if (!ctx->dep_rec32) TRY(alloc(ctx, &ctx->dep_rec32, nrec * (ctx->D + 1)));
    const int64_t nl =
        ctx->D == 3 ? (int64_t)(ctx->p + 1) * (ctx->p + 1) * (ctx->p + 1) : (int64_t)(ctx->p + 1) * (ctx->p + 1);
    if ((int64_t)ctx->desc.ncells[0] * ctx->desc.ncells[1] * ctx->desc.ncells[2] * nl >= (int64_t)1 << 31)
      return fail(ctx, NAT_ERR_UNSUPPORTED, "mesh too large for 32-bit FP32 departure records");
    TRY(nat_launch_records_f32(ctx, nrec));
  }
  if (ctx->has_walls) {
    // wall table: every node with a hit in some direction gets an entry
    std::vector<unsigned> mask(ctx->n_dofs);
    NAT_HIP(hipMemcpyAsync(mask.data(), dmask, sizeof(unsigned) * ctx->n_dofs, hipMemcpyDeviceToHost, ctx->stream));
    NAT_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<int> idx(ctx->n_dofs, -1), list;
    for (int64_t i = 0; i < ctx->n_dofs; i++)
      if (mask[i]) {
        idx[i] = (int)list.size();
        list.push_back((int)i);
      }
    for (void* b : {(void*)ctx->wall_idx, (void*)ctx->wall_flip, (void*)ctx->wall_bias})
      if (b) hipFree(b);
    ctx->wall_idx = nullptr;
    ctx->wall_flip = nullptr;
    ctx->wall_bias = nullptr;
    ctx->n_wall = (int64_t)list.size();
    TRY(alloc(ctx, &ctx->wall_idx, ctx->n_dofs));
    NAT_HIP(hipMemcpyAsync(ctx->wall_idx, idx.data(), sizeof(int) * ctx->n_dofs, hipMemcpyHostToDevice, ctx->stream));
    if (ctx->n_wall) {
      int* dlist;
      TRY(tmp.get(ctx, &dlist, list.size()));
      TRY(alloc(ctx, &ctx->wall_flip, list.size()));
      TRY(alloc(ctx, &ctx->wall_bias, list.size() * ctx->Q));
      NAT_HIP(hipMemcpyAsync(dlist, list.data(), sizeof(int) * list.size(), hipMemcpyHostToDevice, ctx->stream));
      TRY(nat_launch_wall_compact(ctx, dlist, dmask, dbias));
    }
  }
  NAT_HIP(hipStreamSynchronize(ctx->stream));
  return check_status(ctx);
}

void nat_drop_graphs(nat_ctx* ctx) {
  for (int b = 0; b < 2; b++) {
    if (ctx->gexec[b]) hipGraphExecDestroy(ctx->gexec[b]);
    if (ctx->graph[b]) hipGraphDestroy(ctx->graph[b]);
    ctx->gexec[b] = nullptr;
    ctx->graph[b] = nullptr;
  }
}

}  // namespace

extern "C" {

const char* nat_last_error(const nat_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }
int64_t nat_n_dofs(const nat_ctx* ctx) { return ctx->n_dofs; }
int nat_n_stored_nodes(const nat_ctx* ctx, int64_t* n) {
  *n = ctx->n_store;
  return 0;
}
int nat_Q(const nat_ctx* ctx) { return ctx->Q; }
double nat_dt(const nat_ctx* ctx) { return ctx->dt; }
double nat_tau(const nat_ctx* ctx) { return ctx->tau; }
double nat_time(const nat_ctx* ctx) { return ctx->time; }
int nat_backend(const nat_ctx* ctx) { return ctx->backend; }

void nat_destroy(nat_ctx* ctx) {
  if (!ctx) return;
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  nat_drop_graphs(ctx);
  void* bufs[] = {ctx->f[0],     ctx->f[1],        ctx->g[0],     ctx->g[1],       ctx->rho,       ctx->u,
                  ctx->T,        ctx->mask,        ctx->d_vert,   ctx->d_gll,      ctx->d_coords,  ctx->d_owner,
                  ctx->d_status, ctx->d_sums,      ctx->dep_cell, ctx->dep_xi64,   ctx->dep_rec32, ctx->ax_w,
                  ctx->ax_off,   ctx->sedg_df,     ctx->sedg_geo, ctx->comp_tab,    ctx->wall_idx,
                  ctx->wall_flip, ctx->wall_bias};
  for (void* b : bufs)
    if (b) hipFree(b);
  if (ctx->d_partials) hipFree(ctx->d_partials);
  if (ctx->h_status) hipHostFree(ctx->h_status);
  for (hipEvent_t e : ctx->ev)
    if (e) hipEventDestroy(e);
  if (ctx->stream) hipStreamDestroy(ctx->stream);
  if (ctx->cap_stream) hipStreamDestroy(ctx->cap_stream);
  delete ctx;
}

int nat_create(const nat_desc* desc, nat_ctx** out) {
  *out = nullptr;
  nat_ctx* ctx = new (std::nothrow) nat_ctx;
  if (!ctx) return NAT_ERR_NOMEM;
  *out = ctx;
  if (!desc) return fail(ctx, NAT_ERR_ARG, "null descriptor");
  ctx->desc = *desc;
  nat_desc& d = ctx->desc;
  if (d.dim != 2 && d.dim != 3) return fail(ctx, NAT_ERR_ARG, "dim must be 2 or 3");
  if (!make_stencil(d.stencil, d.scaling, ctx->st)) return fail(ctx, NAT_ERR_ARG, "unknown stencil");
  if (ctx->st.D != d.dim) return fail(ctx, NAT_ERR_ARG, "stencil dimension does not match dim");
  if (d.order < 1 || d.order > MAXP) return fail(ctx, NAT_ERR_ARG, "order must be in [1, 8]");
  if (!(d.scaling > 0) || !(d.viscosity > 0) || !(d.cfl > 0)) return fail(ctx, NAT_ERR_ARG, "scaling, viscosity and cfl must be > 0");
  if (!d.vertices) return fail(ctx, NAT_ERR_ARG, "vertices required");
  if (d.precision != NAT_FP64 && d.precision != NAT_FP32_SHIFTED) return fail(ctx, NAT_ERR_ARG, "unknown precision");
  if (d.scheme != NAT_SL && d.scheme != NAT_SEDG && d.scheme != NAT_SEDG_CENTRAL) return fail(ctx, NAT_ERR_ARG, "unknown scheme");
  if (d.equilibrium != NAT_EQ_BGK && !d.compressible)
    return fail(ctx, NAT_ERR_UNSUPPORTED, "incompressible collision supports the BGK equilibrium only");
  if (d.compressible) {
    // selectCollision: compressible (2, 25, BGK_STANDARD, QUARTIC_EQUILIBRIUM) (CollisionSelection.h:143)
    if (d.stencil != NAT_D2Q25H) return fail(ctx, NAT_ERR_UNSUPPORTED, "compressible populations need the D2Q25H stencil");
    if (d.scheme != NAT_SL) return fail(ctx, NAT_ERR_UNSUPPORTED, "compressible populations use SL streaming");
    if (!(d.gamma > 1.0)) return fail(ctx, NAT_ERR_ARG, "gamma must be > 1");
    if (d.prandtl < 0) return fail(ctx, NAT_ERR_ARG, "prandtl must be >= 0");
    d.backend = NAT_BACKEND_GENERAL;
  } else if (d.stencil == NAT_D2Q25H) {
    return fail(ctx, NAT_ERR_UNSUPPORTED, "the D2Q25H stencil is used with compressible populations only");
  }
  ctx->D = d.dim;
  ctx->Q = ctx->st.Q;
  ctx->p = d.order;
  ctx->fp32 = d.precision == NAT_FP32_SHIFTED;

  for (int a = 0; a < 3; a++) {
    if (a >= d.dim) {
      d.ncells[a] = 1;
      d.periodic[a] = 1;
    }
    if (d.ncells[a] < 1) return fail(ctx, NAT_ERR_ARG, "ncells must be >= 1");
    if (!d.periodic[a]) ctx->has_walls = true;
  }
  if (ctx->has_walls && d.scheme == NAT_SL) {
    // velocity-neq bounce-back walls of SL streaming (SemiLagrangian.cpp:358-385) are built in
    // the departure walk of the general backend; SEDG applies them as a wall flux
    // (VelocityNeqBounceBack.cpp:46-135) in its stage kernel
    if (d.backend == NAT_BACKEND_STRUCTURED)
      return fail(ctx, NAT_ERR_UNSUPPORTED, "the structured backend needs periodic boundaries on every axis");
    d.backend = NAT_BACKEND_GENERAL;
  }
  const int nz = d.dim == 3 ? d.ncells[2] + 1 : 1;
  const size_t nv = (size_t)(d.ncells[0] + 1) * (d.ncells[1] + 1) * nz;
  ctx->verts.assign(3 * nv, 0.0);
  for (size_t i = 0; i < nv; i++)
    for (int a = 0; a < d.dim; a++) ctx->verts[3 * i + a] = d.vertices[i * d.dim + a];
  d.vertices = nullptr;
  if (d.scheme != NAT_SL) d.cartesian = 0;  // DG uses the multilinear map

  // time step and relaxation time (CFDSolverUtilities.cpp:92-99, AuxiliaryCollisionFunctions.h:63-68)
  ctx->h_min = min_vertex_distance(ctx);
  ctx->dt = d.cfl * ctx->h_min / (ctx->st.max_speed * d.order * d.order);
  ctx->tau = d.viscosity / (ctx->dt * ctx->st.cs2) + 0.5;

  NAT_HIP(hipStreamCreate(&ctx->stream));
  NAT_HIP(hipStreamCreateWithFlags(&ctx->cap_stream, hipStreamNonBlocking));
  TRY(alloc(ctx, &ctx->d_status, 1));
  TRY(alloc(ctx, &ctx->d_sums, 8));
  NAT_HIP(hipMemset(ctx->d_status, 0, sizeof(int)));
  NAT_HIP(hipHostMalloc((void**)&ctx->h_status, sizeof(int), hipHostMallocDefault));
  *ctx->h_status = 0;
  NAT_HIP(hipEventCreate(&ctx->ev[0]));
  NAT_HIP(hipEventCreate(&ctx->ev[1]));
  TRY(nat_sums_scratch_setup(ctx));  // this device's address of the fixed sums scratch
  NAT_HIP(hipDeviceGetAttribute(&ctx->n_cu, hipDeviceAttributeMultiprocessorCount, ctx->device));
  TRY(alloc(ctx, &ctx->d_vert, ctx->verts.size()));
  NAT_HIP(hipMemcpy(ctx->d_vert, ctx->verts.data(), ctx->verts.size() * sizeof(double), hipMemcpyHostToDevice));
  double gll[MAXP + 1], gw[MAXP + 1];
  gll_points(d.order, gll, gw);
  TRY(alloc(ctx, &ctx->d_gll, MAXP + 1));
  NAT_HIP(hipMemcpy(ctx->d_gll, gll, sizeof(gll), hipMemcpyHostToDevice));

  if (d.scheme == NAT_SL) {
    for (int a = 0; a < 3; a++) {
      ctx->nfe[a] = a < d.dim ? d.ncells[a] * d.order + 1 : 1;
      ctx->nu[a] = a < d.dim ? d.ncells[a] * d.order : 1;
    }
    ctx->n_dofs = (int64_t)ctx->nfe[0] * ctx->nfe[1] * ctx->nfe[2];
    bool want_struct = d.backend != NAT_BACKEND_GENERAL && d.cartesian && is_uniform_cartesian(ctx) &&
                       structured_instantiated(d.dim, ctx->Q, d.order);
    if (d.backend == NAT_BACKEND_STRUCTURED && !want_struct)
      return fail(ctx, NAT_ERR_UNSUPPORTED, "structured backend needs a uniform Cartesian periodic mesh and a supported (stencil, order)");
    ctx->backend = want_struct ? NAT_BACKEND_STRUCTURED : NAT_BACKEND_GENERAL;
    if (ctx->backend == NAT_BACKEND_STRUCTURED) {
      bool ok = false;
      StructTables t;
      TRY(build_structured_tables(ctx, t, &ok));
      if (ok) commit_structured_tables(ctx, t);
      if (!ok) {
        if (d.backend == NAT_BACKEND_STRUCTURED)
          return fail(ctx, NAT_ERR_UNSUPPORTED, "operator is not a separable one-cell-layer shift; use the general backend");
        ctx->backend = NAT_BACKEND_GENERAL;
      }
    }
    ctx->n_store = ctx->backend == NAT_BACKEND_STRUCTURED ? (int64_t)ctx->nu[0] * ctx->nu[1] * ctx->nu[2] : ctx->n_dofs;
    // support points and owner cells of the DoFs, formed on the device
    TRY(alloc(ctx, &ctx->d_coords, 3 * ctx->n_dofs));
    TRY(alloc(ctx, &ctx->d_owner, ctx->n_dofs));
    TRY(nat_launch_support_points(ctx));
  } else {
    const int np = d.order + 1;
    const int64_t ncell = (int64_t)d.ncells[0] * d.ncells[1] * (d.dim == 3 ? d.ncells[2] : 1);
    ctx->n_dofs = ncell * (d.dim == 3 ? np * np * np : np * np);
    ctx->n_store = ctx->n_dofs;
    ctx->backend = NAT_BACKEND_GENERAL;
  }
  ctx->stride = (ctx->n_store + 127) / 128 * 128;
  const size_t es = ctx->fp32 ? sizeof(float) : sizeof(double);
  for (int b = 0; b < 2; b++) {
    hipError_t e = hipMalloc(&ctx->f[b], es * ctx->stride * ctx->Q);
    if (e != hipSuccess) return fail(ctx, NAT_ERR_NOMEM, "device allocation of populations failed");
    NAT_HIP(hipMemset(ctx->f[b], 0, es * ctx->stride * ctx->Q));
  }
  if (d.compressible) {
    for (int b = 0; b < 2; b++) {
      hipError_t e = hipMalloc(&ctx->g[b], es * ctx->stride * ctx->Q);
      if (e != hipSuccess) return fail(ctx, NAT_ERR_NOMEM, "device allocation of g populations failed");
      NAT_HIP(hipMemset(ctx->g[b], 0, es * ctx->stride * ctx->Q));
    }
    TRY(alloc(ctx, &ctx->T, ctx->n_store));
    TRY(alloc(ctx, &ctx->mask, ctx->n_store));
    NAT_HIP(hipMemset(ctx->T, 0, sizeof(double) * ctx->n_store));
    NAT_HIP(hipMemset(ctx->mask, 0, sizeof(double) * ctx->n_store));
    TRY(nat_comp_setup(ctx));
  }
  {
    hipError_t e = hipMalloc(&ctx->rho, es * ctx->n_store);
    if (e == hipSuccess) e = hipMalloc(&ctx->u, es * ctx->n_store * ctx->D);
    if (e != hipSuccess) return fail(ctx, NAT_ERR_NOMEM, "device allocation of macroscopic fields failed");
    NAT_HIP(hipMemset(ctx->rho, 0, es * ctx->n_store));
    NAT_HIP(hipMemset(ctx->u, 0, es * ctx->n_store * ctx->D));
  }
  if (d.scheme == NAT_SL) {
    if (ctx->backend == NAT_BACKEND_GENERAL) TRY(assemble_general(ctx));
  } else {
    TRY(nat_sedg_setup(ctx));
  }
  ctx->assembled = true;
  return 0;
}

int nat_node_coords(nat_ctx* ctx, double* xyz) {
  if (ctx->desc.scheme == NAT_SL) {  // the support points the assembly walks from
    NAT_HIP(hipMemcpyAsync(xyz, ctx->d_coords, 3 * ctx->n_dofs * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    NAT_HIP(hipStreamSynchronize(ctx->stream));
    return 0;
  }
  std::vector<double> x;
  dg_points(ctx, x);
  std::memcpy(xyz, x.data(), x.size() * sizeof(double));
  return 0;
}

int nat_set_vertices(nat_ctx* ctx, const double* vertices) {
  if (!vertices) return fail(ctx, NAT_ERR_ARG, "vertices required");
  nat_drop_graphs(ctx);
  const nat_desc& d = ctx->desc;
  const int nz = d.dim == 3 ? d.ncells[2] + 1 : 1;
  const size_t nv = (size_t)(d.ncells[0] + 1) * (d.ncells[1] + 1) * nz;
  const std::vector<double> old = ctx->verts;
  const double old_hmin = ctx->h_min;
  for (size_t i = 0; i < nv; i++)
    for (int a = 0; a < d.dim; a++) ctx->verts[3 * i + a] = vertices[i * d.dim + a];
  if (d.cartesian && !is_axis_aligned(ctx)) {
    ctx->verts = old;
    return fail(ctx, NAT_ERR_ARG, "a context created with cartesian = 1 needs axis-aligned cells");
  }
  if (d.scheme == NAT_SL && ctx->backend == NAT_BACKEND_STRUCTURED && !is_uniform_cartesian(ctx)) {
    ctx->verts = old;
    return fail(ctx, NAT_ERR_UNSUPPORTED, "the structured backend needs a uniform Cartesian mesh");
  }
  // Put the previous mesh back (host and device vertices, support points) after a failed
  // rebuild, so that the operator in use stays consistent with the geometry.
  auto restore = [&](int code, const std::string& msg) {
    ctx->verts = old;
    ctx->h_min = old_hmin;
    if (hipMemcpy(ctx->d_vert, ctx->verts.data(), ctx->verts.size() * sizeof(double), hipMemcpyHostToDevice) !=
            hipSuccess ||
        (d.scheme == NAT_SL && nat_launch_support_points(ctx) != 0))
      ctx->assembled = false;
    ctx->err = msg;
    return code;
  };
  NAT_HIP(hipMemcpy(ctx->d_vert, ctx->verts.data(), ctx->verts.size() * sizeof(double), hipMemcpyHostToDevice));
  ctx->h_min = min_vertex_distance(ctx);
  if (d.scheme != NAT_SL) return 0;  // SEDG evaluates the geometry from the vertices in every stage
  {
    const int rc = nat_launch_support_points(ctx);
    if (rc) return restore(rc, ctx->err);
  }
  if (ctx->backend == NAT_BACKEND_STRUCTURED) {
    bool ok = false;
    StructTables t;
    const int rc = build_structured_tables(ctx, t, &ok);
    if (rc) return restore(rc, ctx->err);
    if (!ok) return restore(NAT_ERR_UNSUPPORTED, "moved mesh gives a non-separable operator for the structured backend");
    commit_structured_tables(ctx, t);
    return 0;
  }
  const int rc = assemble_general(ctx);
  if (rc) {
    const std::string msg = ctx->err;
    restore(rc, msg);
    // the departure records were rewritten in place: rebuild them for the previous mesh
    if (ctx->assembled && assemble_general(ctx) != 0) ctx->assembled = false;
    ctx->err = msg;
    return rc;
  }
  return 0;
}

int nat_set_stabilizer(nat_ctx* ctx, int on) {
  if (ctx->stabilizer != (on != 0)) nat_drop_graphs(ctx);  // the flag is baked into captured steps
  ctx->stabilizer = on != 0;
  return 0;
}

int nat_set_dt(nat_ctx* ctx, double dt) {
  nat_drop_graphs(ctx);
  const double old_dt = ctx->dt, old_tau = ctx->tau;
  if (dt > 0) {
    ctx->dt = dt;
    ctx->tau = ctx->desc.viscosity / (ctx->dt * ctx->st.cs2) + 0.5;
  }
  // a rejected time step leaves dt, tau and the operator as they were
  auto restore = [&](int code, const std::string& msg) {
    ctx->dt = old_dt;
    ctx->tau = old_tau;
    ctx->err = msg;
    return code;
  };
  if (ctx->desc.scheme != NAT_SL) {
    const int rc = nat_sedg_setup(ctx);
    return rc ? restore(rc, ctx->err) : 0;
  }
  if (ctx->backend == NAT_BACKEND_STRUCTURED) {
    bool ok = false;
    StructTables t;
    const int rc = build_structured_tables(ctx, t, &ok);
    if (rc) return restore(rc, ctx->err);
    if (!ok) return restore(NAT_ERR_UNSUPPORTED, "time step gives a non-separable operator for the structured backend");
    commit_structured_tables(ctx, t);
    return 0;
  }
  const int rc = assemble_general(ctx);
  if (rc) {
    const std::string msg = ctx->err;
    restore(rc, msg);
    // the departure records were rewritten in place: rebuild them for the previous dt
    if (assemble_general(ctx) != 0) ctx->assembled = false;
    ctx->err = msg;
    return rc;
  }
  return 0;
}

// FP32 shift of the g populations per unit weight: rho0 (2 C_v - D), the g equilibrium at rest with T = 1
static double g_shift(const nat_ctx* ctx) {
  const double rho0 = ctx->desc.rho0 > 0 ? ctx->desc.rho0 : 1.0;
  return rho0 * (2.0 / (ctx->desc.gamma - 1.0) - ctx->D);
}

static int upload_pop(nat_ctx* ctx, void* dst, const double* f, bool is_g = false) {
  const int Q = ctx->Q;
  const int64_t n = ctx->n_dofs;
  if (ctx->fp32) {
    const double rho0 = is_g ? g_shift(ctx) : (ctx->desc.rho0 > 0 ? ctx->desc.rho0 : 1.0);
    std::vector<float> h((size_t)ctx->stride * Q, 0.0f);
    for (int q = 0; q < Q; q++)
      for (int64_t i = 0; i < n; i++) h[(size_t)q * ctx->stride + store_index(ctx, i)] = (float)(f[q * n + i] - ctx->st.w[q] * rho0);
    NAT_HIP(hipMemcpy(dst, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
  } else {
    std::vector<double> h((size_t)ctx->stride * Q, 0.0);
    for (int q = 0; q < Q; q++)
      for (int64_t i = 0; i < n; i++) h[(size_t)q * ctx->stride + store_index(ctx, i)] = f[q * n + i];
    NAT_HIP(hipMemcpy(dst, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  return 0;
}

static int download_pop(nat_ctx* ctx, const void* src, double* f, bool is_g = false) {
  const int Q = ctx->Q;
  const int64_t n = ctx->n_dofs;
  NAT_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->fp32) {
    const double rho0 = is_g ? g_shift(ctx) : (ctx->desc.rho0 > 0 ? ctx->desc.rho0 : 1.0);
    std::vector<float> h((size_t)ctx->stride * Q);
    NAT_HIP(hipMemcpy(h.data(), src, h.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (int q = 0; q < Q; q++)
      for (int64_t i = 0; i < n; i++) f[q * n + i] = (double)h[(size_t)q * ctx->stride + store_index(ctx, i)] + ctx->st.w[q] * rho0;
  } else {
    std::vector<double> h((size_t)ctx->stride * Q);
    NAT_HIP(hipMemcpy(h.data(), src, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int q = 0; q < Q; q++)
      for (int64_t i = 0; i < n; i++) f[q * n + i] = h[(size_t)q * ctx->stride + store_index(ctx, i)];
  }
  return 0;
}

int nat_upload_f(nat_ctx* ctx, const double* f) {
  NAT_HIP(hipStreamSynchronize(ctx->stream));
  return upload_pop(ctx, ctx->f[ctx->cur], f);
}
int nat_download_f(nat_ctx* ctx, double* f) { return download_pop(ctx, ctx->f[ctx->cur], f); }
int nat_upload_g(nat_ctx* ctx, const double* g) {
  if (!ctx->desc.compressible) return fail(ctx, NAT_ERR_UNSUPPORTED, "no g populations in this context");
  NAT_HIP(hipStreamSynchronize(ctx->stream));
  return upload_pop(ctx, ctx->g[ctx->cur], g, true);
}
int nat_download_g(nat_ctx* ctx, double* g) {
  if (!ctx->desc.compressible) return fail(ctx, NAT_ERR_UNSUPPORTED, "no g populations in this context");
  return download_pop(ctx, ctx->g[ctx->cur], g, true);
}

// CompressibleCFDSolver::calcQuarticEquilibrium in its loop form
// (CompressibleCFDSolver.h:790-848), g = f_eq T (2 C_v - D) (:930-943)
static int init_compressible(nat_ctx* ctx, const double* rho, const double* u, const double* T) {
  const int Q = ctx->Q, D = ctx->D;
  const int64_t n = ctx->n_dofs;
  const double s = ctx->st.s;
  const double cs2 = ctx->st.cs2 / (s * s);
  const double C_v = 1. / (ctx->desc.gamma - 1.0);
  std::vector<double> f((size_t)Q * n), g((size_t)Q * n);
  auto eye = [](int a, int b) { return a == b ? 1.0 : 0.0; };
  for (int64_t ii = 0; ii < n; ii++) {
    double vel[3] = {0, 0, 0};
    for (int j = 0; j < D; j++) vel[j] = u[j * n + ii] / s;
    const double density = rho[ii], temperature = T[ii];
    double uu_term = 0.0;
    for (int j = 0; j < D; j++) uu_term += -(vel[j] * vel[j]) / (2.0 * cs2);
    for (int i = 0; i < Q; i++) {
      double e[3] = {0, 0, 0};
      for (int j = 0; j < D; j++) e[j] = ctx->st.dir[i][j] / s;
      const double w = ctx->st.w[i];
      const double T1 = cs2 * (temperature - 1);
      double ue = 0.0;
      for (int j = 0; j < D; j++) ue += (vel[j] * e[j]) / cs2;
      double feq = w * density * (1 + ue * (1 + 0.5 * ue) + uu_term);
      for (int a = 0; a < D; a++)
        for (int b = 0; b < D; b++) {
          feq += density * w / (2.0 * cs2) *
                 ((temperature - 1) * eye(a, b) * e[a] * e[b] - cs2 * eye(a, b) * (temperature - 1));
          for (int c = 0; c < D; c++) {
            feq += w * density / (6. * cs2 * cs2 * cs2) *
                   (vel[a] * vel[b] * vel[c] + T1 * (eye(a, b) * vel[c] + eye(b, c) * vel[a] + eye(a, c) * vel[b])) *
                   (e[a] * e[b] * e[c] - cs2 * (e[c] * eye(a, b) + e[b] * eye(a, c) + e[a] * eye(b, c)));
            for (int d = 0; d < D; d++) {
              const double p4 = e[a] * e[b] * e[c] * e[d];
              const double p2 = e[a] * e[b] * eye(c, d) + e[a] * e[c] * eye(b, d) + e[a] * e[d] * eye(b, c) +
                                e[b] * e[c] * eye(a, d) + e[b] * e[d] * eye(a, c) + e[c] * e[d] * eye(a, b);
              const double p0 = eye(a, b) * eye(c, d) + eye(a, c) * eye(b, d) + eye(a, d) * eye(b, c);
              const double u4 = vel[a] * vel[b] * vel[c] * vel[d];
              const double u2 = vel[a] * vel[b] * eye(c, d) + vel[a] * vel[c] * eye(b, d) + vel[a] * vel[d] * eye(b, c) +
                                vel[b] * vel[c] * eye(a, d) + vel[b] * vel[d] * eye(a, c) + vel[c] * vel[d] * eye(a, b);
              feq += w * density / (24. * cs2 * cs2 * cs2 * cs2) * (p4 - cs2 * p2 + cs2 * cs2 * p0) *
                     (u4 + T1 * (u2 + T1 * p0));
            }
          }
        }
      f[(size_t)i * n + ii] = feq;
      g[(size_t)i * n + ii] = feq * temperature * (2.0 * C_v - D);
    }
  }
  ctx->time = 0;
  ctx->collided = false;
  TRY(nat_upload_f(ctx, f.data()));
  return nat_upload_g(ctx, g.data());
}

int nat_init_equilibrium(nat_ctx* ctx, const double* rho, const double* u, const double* T) {
  if (ctx->desc.compressible) {
    if (!T) return fail(ctx, NAT_ERR_ARG, "compressible initialisation needs the temperature field");
    return init_compressible(ctx, rho, u, T);
  }
  // BGKStandard::getEquilibriumDistribution with the scaled stencil (BGKStandard.cpp:23-41)
  const int Q = ctx->Q, D = ctx->D;
  const int64_t n = ctx->n_dofs;
  std::vector<double> f((size_t)Q * n);
  const double cs2 = ctx->st.cs2;
  for (int64_t i = 0; i < n; i++) {
    double uu = 0;
    for (int d = 0; d < D; d++) uu += u[d * n + i] * u[d * n + i];
    for (int q = 0; q < Q; q++) {
      const double pre = ctx->st.w[q] * rho[i];
      const double usq = -(uu) / (2 * cs2);
      if (q == 0) {
        f[(size_t)q * n + i] = pre * (1 + usq);
        continue;
      }
      double ue = 0;
      for (int d = 0; d < D; d++) ue += u[d * n + i] * ctx->st.dir[q][d];
      const double m = ue / cs2;
      f[(size_t)q * n + i] = pre * (1 + m * (1 + 0.5 * m) + usq);
    }
  }
  ctx->time = 0;
  ctx->collided = false;
  return nat_upload_f(ctx, f.data());
}

int nat_collide(nat_ctx* ctx) {
  if (ctx->desc.compressible)
    TRY(nat_launch_collide_comp(ctx));
  else
    TRY(nat_launch_collide(ctx, 0));
  ctx->collided = true;
  return check_status(ctx);
}

int nat_collide_init(nat_ctx* ctx, const double* u) {
  if (ctx->desc.compressible)
    return fail(ctx, NAT_ERR_UNSUPPORTED, "the initialisation collision is built for incompressible contexts");
  if (u) {
    // the velocity field the collision reads: physical units, DoF order -> storage order
    const int64_t n = ctx->n_dofs, ns = ctx->n_store;
    const int D = ctx->D;
    NAT_HIP(hipStreamSynchronize(ctx->stream));
    if (ctx->fp32) {
      std::vector<float> h((size_t)ns * D);
      for (int d = 0; d < D; d++)
        for (int64_t i = 0; i < n; i++) h[(size_t)d * ns + store_index(ctx, i)] = (float)u[(size_t)d * n + i];
      NAT_HIP(hipMemcpy(ctx->u, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
    } else {
      std::vector<double> h((size_t)ns * D);
      for (int d = 0; d < D; d++)
        for (int64_t i = 0; i < n; i++) h[(size_t)d * ns + store_index(ctx, i)] = u[(size_t)d * n + i];
      NAT_HIP(hipMemcpy(ctx->u, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    }
  }
  TRY(nat_launch_collide(ctx, 1));
  ctx->collided = true;
  return check_status(ctx);
}

int nat_stabilize(nat_ctx* ctx) {
  TRY(nat_launch_stabilize(ctx));
  return check_status(ctx);
}

static int step_launch(nat_ctx* ctx, int nsteps, bool macro_last = true) {
  if (!ctx->assembled) return fail(ctx, NAT_ERR_STATE, "operator not assembled");
  for (int s = 0; s < nsteps; s++) {
    const int wm = (macro_last && s == nsteps - 1) ? STEP_MACRO : 0;
    if (ctx->desc.compressible) {
      TRY(nat_launch_sl_comp(ctx, wm));
    } else if (ctx->desc.scheme == NAT_SL) {
      if (ctx->backend == NAT_BACKEND_STRUCTURED)
        TRY(nat_launch_sl_struct(ctx, wm));
      else
        TRY(nat_launch_sl_general(ctx, wm));
    } else {
      TRY(nat_sedg_step(ctx, 1, wm));
    }
    ctx->time += ctx->dt;
  }
  return 0;
}

// Long runs of steps go to the device as replays of a captured HIP graph of
// kGraphSteps fused steps: one submission per chunk instead of one per kernel, so
// dependent launches follow each other without host launch latency in between. An
// even chunk returns the ping-pong buffers to where it started; one graph is kept
// per starting buffer. Anything baked into the kernel arguments (dt, tables,
// stabilizer) drops the graphs when it changes.
static constexpr int kGraphSteps = 32;

// Captures nsteps fused steps (the last one writing the fields when macro_last) into a graph
// without running them; time and the ping-pong state are left as they were. The launches are
// recorded on the context's capture stream (see nat_ctx::cap_stream).
extern "C++" int nat_internal_capture(nat_ctx* ctx, int nsteps, bool macro_last, hipGraph_t* out) {
  *out = nullptr;
  const double t = ctx->time;
  const int par = ctx->cur;
  hipStream_t run = ctx->stream;
  ctx->stream = ctx->cap_stream;
  hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
  int rc = 0;
  if (e == hipSuccess) {
    rc = step_launch(ctx, nsteps, macro_last);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(ctx->stream, &g);
    if (rc || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
    } else {
      *out = g;
    }
  }
  ctx->stream = run;
  ctx->time = t;
  ctx->cur = par;
  if (rc) return rc;
  if (e != hipSuccess) return fail(ctx, NAT_ERR_HIP, std::string("step graph capture: ") + hipGetErrorString(e));
  return 0;
}

static int step_enqueue(nat_ctx* ctx, int nsteps) {
  int left = nsteps;
  // the last step writes the macroscopic fields and always runs as a plain launch
  while (ctx->use_graphs && left - 1 >= kGraphSteps) {
    const int par = ctx->cur;
    if (!ctx->gexec[par]) {
      hipGraph_t g = nullptr;
      TRY(nat_internal_capture(ctx, kGraphSteps, false, &g));
      ctx->graph[par] = g;
      NAT_HIP(hipGraphInstantiate(&ctx->gexec[par], g, nullptr, nullptr, 0));
    }
    NAT_HIP(hipGraphLaunch(ctx->gexec[par], ctx->stream));
    for (int st = 0; st < kGraphSteps; st++) ctx->time += ctx->dt;  // as step by step
    left -= kGraphSteps;
  }
  return step_launch(ctx, left);
}

int nat_step(nat_ctx* ctx, int nsteps) {
  TRY(step_enqueue(ctx, nsteps));
  return check_status(ctx);
}

int nat_sl_assemble(nat_ctx* ctx, double dt) {
  if (ctx->desc.scheme != NAT_SL) return fail(ctx, NAT_ERR_ARG, "nat_sl_assemble: the context uses the SEDG scheme");
  return nat_set_dt(ctx, dt);
}

int nat_sl_bgk_step(nat_ctx* ctx, int nsteps) {
  if (ctx->desc.scheme != NAT_SL || ctx->desc.compressible)
    return fail(ctx, NAT_ERR_ARG, "nat_sl_bgk_step: the context is not an incompressible SL context");
  return nat_step(ctx, nsteps);
}

int nat_sedg_rk5_step(nat_ctx* ctx, int nsteps) {
  if (ctx->desc.scheme == NAT_SL) return fail(ctx, NAT_ERR_ARG, "nat_sedg_rk5_step: the context uses the SL scheme");
  return nat_step(ctx, nsteps);
}

int nat_compressible_step(nat_ctx* ctx, int nsteps) {
  if (!ctx->desc.compressible) return fail(ctx, NAT_ERR_ARG, "nat_compressible_step: the context is incompressible");
  return nat_step(ctx, nsteps);
}

int nat_step_timed(nat_ctx* ctx, int nsteps, double* elapsed_ms) {
  NAT_HIP(hipEventRecord(ctx->ev[0], ctx->stream));
  TRY(step_enqueue(ctx, nsteps));
  NAT_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
  // check_status waits for the stream, which includes the second event
  const int rc = check_status(ctx);
  float ms = 0;
  NAT_HIP(hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]));
  *elapsed_ms = ms;
  return rc;
}

int nat_stream(nat_ctx* ctx) {
  if (!ctx->assembled) return fail(ctx, NAT_ERR_STATE, "operator not assembled");
  if (ctx->desc.compressible)
    TRY(nat_launch_sl_comp(ctx, STEP_STREAM_ONLY));
  else if (ctx->desc.scheme != NAT_SL)
    TRY(nat_sedg_step(ctx, 0, 0));
  else if (ctx->backend == NAT_BACKEND_STRUCTURED)
    TRY(nat_launch_sl_struct(ctx, STEP_STREAM_ONLY));
  else
    TRY(nat_launch_sl_general(ctx, STEP_STREAM_ONLY));
  ctx->time += ctx->dt;
  return check_status(ctx);
}

int nat_get_knudsen(nat_ctx* ctx, double* kn) {
  if (!ctx->desc.compressible) return fail(ctx, NAT_ERR_UNSUPPORTED, "no shock sensor in an incompressible context");
  NAT_HIP(hipStreamSynchronize(ctx->stream));
  NAT_HIP(hipMemcpy(kn, ctx->mask, ctx->n_store * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

int nat_get_macro(nat_ctx* ctx, double* rho, double* u, double* T) {
  NAT_HIP(hipStreamSynchronize(ctx->stream));
  const int64_t n = ctx->n_dofs, ns = ctx->n_store;
  const int D = ctx->D;
  if (T) {
    if (!ctx->desc.compressible) return fail(ctx, NAT_ERR_UNSUPPORTED, "no temperature in an incompressible context");
    NAT_HIP(hipMemcpy(T, ctx->T, ns * sizeof(double), hipMemcpyDeviceToHost));
  }
  if (ctx->fp32) {
    std::vector<float> r(ns), v((size_t)ns * D);
    NAT_HIP(hipMemcpy(r.data(), ctx->rho, ns * sizeof(float), hipMemcpyDeviceToHost));
    NAT_HIP(hipMemcpy(v.data(), ctx->u, ns * D * sizeof(float), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; i++) {
      const int64_t s = store_index(ctx, i);
      rho[i] = r[s];
      for (int d = 0; d < D; d++) u[d * n + i] = v[(size_t)d * ns + s];
    }
  } else {
    std::vector<double> r(ns), v((size_t)ns * D);
    NAT_HIP(hipMemcpy(r.data(), ctx->rho, ns * sizeof(double), hipMemcpyDeviceToHost));
    NAT_HIP(hipMemcpy(v.data(), ctx->u, ns * D * sizeof(double), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; i++) {
      const int64_t s = store_index(ctx, i);
      rho[i] = r[s];
      for (int d = 0; d < D; d++) u[d * n + i] = v[(size_t)d * ns + s];
    }
  }
  return 0;
}

static int sums_result(nat_ctx* ctx, double* mass, double* momentum, double* ekin) {
  double h[5];
  NAT_HIP(hipMemcpyAsync(h, ctx->d_sums, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  NAT_HIP(hipStreamSynchronize(ctx->stream));
  if (mass) *mass = h[0];
  if (momentum)
    for (int d = 0; d < ctx->D; d++) momentum[d] = h[1 + d];
  if (ekin) *ekin = 0.5 * h[4];
  return 0;
}

static int sums_layout(const nat_ctx* ctx) {
  return ctx->desc.scheme != NAT_SL ? 2 : (ctx->backend == NAT_BACKEND_STRUCTURED ? 1 : 0);
}

int nat_sums(nat_ctx* ctx, double* mass, double* momentum, double* ekin) {
  TRY(nat_launch_sums(ctx, sums_layout(ctx)));
  return sums_result(ctx, mass, momentum, ekin);
}

int nat_sums_noalloc(nat_ctx* ctx, double* mass, double* momentum, double* ekin) {
  TRY(nat_launch_sums_noalloc(ctx, sums_layout(ctx)));
  return sums_result(ctx, mass, momentum, ekin);
}

int nat_struct_tile(nat_ctx* ctx, int* tile) {
  if (ctx->desc.scheme != NAT_SL || ctx->backend != NAT_BACKEND_STRUCTURED || ctx->desc.compressible)
    return fail(ctx, NAT_ERR_ARG, "nat_struct_tile: not a structured SL context");
  return nat_launch_sl_struct_tile(ctx, tile);
}

int nat_pack_key(nat_ctx* ctx, int64_t* key) {
  if (!key) return fail(ctx, NAT_ERR_ARG, "nat_pack_key: null key");
  if (nat_internal_pack_key(ctx, key))
    return fail(ctx, NAT_ERR_UNSUPPORTED,
                "packed steps serve structured SL contexts (uniform periodic boxes) at p = 2-4");
  return 0;
}

int nat_step_packed(nat_ctx* const* ctxs, int n, int nsteps) {
  if (!ctxs || n < 1 || !ctxs[0]) return NAT_ERR_ARG;
  nat_ctx* ctx = ctxs[0];
  if (n > NAT_PACK_MAX) return fail(ctx, NAT_ERR_ARG, "nat_step_packed: more than 16 contexts");
  if (nsteps < 0) return fail(ctx, NAT_ERR_ARG, "nat_step_packed: negative step count");
  for (int i = 0; i < n; i++) {
    if (!ctxs[i]) return fail(ctx, NAT_ERR_ARG, "nat_step_packed: null context");
    if (ctxs[i]->device != ctx->device) return fail(ctx, NAT_ERR_ARG, "nat_step_packed: contexts on different GPUs");
    for (int j = 0; j < i; j++)
      if (ctxs[j] == ctxs[i]) return fail(ctx, NAT_ERR_ARG, "nat_step_packed: a context appears twice");
  }
  if (nsteps == 0) return 0;
  // the launches go to the first context's stream, after the others' pending work; their
  // later work waits for the launches
  for (int i = 1; i < n; i++) {
    NAT_HIP(hipEventRecord(ctxs[i]->ev[0], ctxs[i]->stream));
    NAT_HIP(hipStreamWaitEvent(ctx->stream, ctxs[i]->ev[0], 0));
  }
  const int rc = nat_internal_pack_step(ctxs, n, nsteps, ctx->stream);
  if (rc) {
    if (ctx->err.empty()) ctx->err = "nat_step_packed: packing failed";
    return rc;
  }
  NAT_HIP(hipEventRecord(ctx->ev[1], ctx->stream));
  for (int i = 1; i < n; i++) NAT_HIP(hipStreamWaitEvent(ctxs[i]->stream, ctx->ev[1], 0));
  // each member's status word (density below the collision limit, ...)
  for (int i = 0; i < n; i++) {
    const int st = check_status(ctxs[i]);
    if (st) {
      if (i > 0) ctx->err = "context " + std::to_string(i) + ": " + ctxs[i]->err;
      return st;
    }
  }
  return 0;
}

int nat_set_general_form(nat_ctx* ctx, int form) {
  if (!(ctx->desc.scheme == NAT_SL && !ctx->desc.compressible && ctx->backend == NAT_BACKEND_GENERAL))
    return fail(ctx, NAT_ERR_ARG, "nat_set_general_form: a general-backend SL context is needed");
  const bool two = ctx->fp32 && ctx->D == 3;
  if (form < -1 || form > (two ? 1 : 0)) return fail(ctx, NAT_ERR_ARG, "nat_set_general_form: form out of range");
  if (form != ctx->gen_pref) nat_drop_graphs(ctx);  // captured steps hold the other form's launches
  ctx->gen_pref = form;
  return 0;
}

int nat_autotune(nat_ctx* ctx, int nsteps, int* tile) {
  nat_drop_graphs(ctx);  // captured steps hold the previous tile's launches
  int t[3] = {0, 0, 0};
  const bool general = ctx->desc.scheme == NAT_SL && !ctx->desc.compressible && ctx->backend == NAT_BACKEND_GENERAL;
  const int rc = general ? nat_internal_autotune_general(ctx, nsteps, t, nullptr)
                         : nat_internal_autotune(ctx, nsteps, t, nullptr);
  if (rc) return rc;
  if (tile)
    for (int i = 0; i < 3; i++) tile[i] = t[i];
  // the timing launches leave the state untouched; a status bit they set (a context autotuned
  // before its populations were initialised) is not an error of the solve
  NAT_HIP(hipMemsetAsync(ctx->d_status, 0, sizeof(int), ctx->stream));
  NAT_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

int nat_sync(nat_ctx* ctx) {
  NAT_HIP(hipStreamSynchronize(ctx->stream));
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Internal entry points for the request server (nat_serve.cpp); not exported.
// ---------------------------------------------------------------------------
int nat_internal_enqueue(nat_ctx* ctx, int nsteps) { return step_launch(ctx, nsteps); }
int nat_internal_status(nat_ctx* ctx) { return check_status(ctx); }

// Device bytes a context for this descriptor allocates (populations, macroscopic
// fields, departure records or SEDG work vectors; tables are ignored).
int64_t nat_internal_footprint(const nat_desc* d) {
  using namespace nat;
  const int D = d->dim, p = d->order;
  const int Q = d->stencil == NAT_D2Q9 ? 9 : d->stencil == NAT_D3Q19 ? 19 : d->stencil == NAT_D3Q27 ? 27 : 25;
  const int np = p + 1;
  int64_t ncell = 1, nfe = 1, nu = 1;
  for (int a = 0; a < D; a++) {
    ncell *= d->ncells[a];
    nfe *= (int64_t)d->ncells[a] * p + 1;
    nu *= (int64_t)d->ncells[a] * p;
  }
  const int64_t es = d->precision == NAT_FP32_SHIFTED ? 4 : 8;
  int64_t n_store, recs = 0;
  if (d->scheme != NAT_SL) {
    n_store = ncell * (D == 3 ? np * np * np : np * np);
    recs = Q * ((n_store + 127) / 128 * 128) * es;  // RK5 df
  } else {
    bool walls = false;
    for (int a = 0; a < D; a++) walls = walls || !d->periodic[a];
    const bool structured = !d->compressible && !walls && d->cartesian && d->backend != NAT_BACKEND_GENERAL &&
                            structured_instantiated(D, Q, p);
    n_store = structured ? nu : nfe;
    if (!structured) {
      recs = (int64_t)(Q - 1) * nfe * (4 + 8 * D);
      if (es == 4 && !d->compressible) recs += (int64_t)(Q - 1) * nfe * (4 + 4 * D);
      // walls: node index per DoF, the wall entries (at most every node: mask + Q values)
      // and, during assembly, a mask per DoF and one value per record
      if (walls) recs += 4 * nfe + nfe * (4 + 8 * Q) + 4 * nfe + (int64_t)(Q - 1) * nfe * 8;
    }
  }
  const int64_t stride = (n_store + 127) / 128 * 128;
  int64_t bytes = 2 * Q * stride * es * (d->compressible ? 2 : 1);
  bytes += (1 + D) * n_store * es + (d->compressible ? 2 * 8 * n_store : 0);
  bytes += recs;
  return bytes + (1 << 20);
}
