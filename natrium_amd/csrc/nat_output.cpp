// Output and restart formats of the solver, on the host, from fields downloaded
// through the C ABI:
//   - the results-table line of SolverStats (solver/SolverStats.h:102-175) with the
//     quantities of PhysicalProperties (solver/PhysicalProperties.cpp:29-420);
//   - the VTU/PVTU solution files of CFDSolver::output (solver/CFDSolver.cpp:1001-1069);
//   - checkpoints: the status file of Checkpoint::write/load (solver/Checkpoint.cpp:74-140)
//     and a data file of FP64 populations in DoF order.
#include <dirent.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include <zlib.h>

#include "nat_internal.hpp"

using namespace nat;

namespace {

int fail(nat_ctx* ctx, int code, const std::string& msg) {
  ctx->err = msg;
  return code;
}

// Cell geometry and the local-to-global DoF map of the context's space (FE_Q lattice for
// SL, cell-major DG nodes for SEDG), GLL nodes and weights, 1-D differentiation matrix.
struct Space {
  const nat_ctx* c;
  int D, p, np, nl;
  int N[3];
  bool cartesian;
  double x[MAXP + 1], w[MAXP + 1];
  double dm[MAXP + 1][MAXP + 1];  // dm[i][j] = l_j'(x_i)

  explicit Space(const nat_ctx* ctx) : c(ctx) {
    D = ctx->D;
    p = ctx->p;
    np = p + 1;
    nl = D == 3 ? np * np * np : np * np;
    for (int a = 0; a < 3; a++) N[a] = a < D ? ctx->desc.ncells[a] : 1;
    cartesian = ctx->desc.scheme == NAT_SL && ctx->desc.cartesian;
    gll_points(p, x, w);
    for (int i = 0; i <= p; i++)
      for (int j = 0; j <= p; j++) {
        // derivative of the Lagrange polynomial j at node i
        double s = 0;
        if (i == j) {
          for (int k = 0; k <= p; k++)
            if (k != j) s += 1.0 / (x[j] - x[k]);
        } else {
          double num = 1, den = 1;
          for (int k = 0; k <= p; k++) {
            if (k != j) den *= x[j] - x[k];
            if (k != j && k != i) num *= x[i] - x[k];
          }
          s = num / den;
        }
        dm[i][j] = s;
      }
  }
  int64_t n_cells() const { return (int64_t)N[0] * N[1] * N[2]; }
  void cell_xyz(int64_t cell, int cc[3]) const {
    cc[0] = (int)(cell % N[0]);
    cc[1] = (int)((cell / N[0]) % N[1]);
    cc[2] = (int)(cell / ((int64_t)N[0] * N[1]));
  }
  const double* vert(int i, int j, int k) const {
    return &c->verts[3 * ((size_t)i + (size_t)(N[0] + 1) * ((size_t)j + (size_t)(N[1] + 1) * k))];
  }
  const double* cell_vertex(const int cc[3], int v) const {
    return vert(cc[0] + (v & 1), cc[1] + ((v >> 1) & 1), D == 3 ? cc[2] + ((v >> 2) & 1) : 0);
  }
  // global DoF of local node (m0, m1, m2) of a cell
  int64_t dof(int64_t cell, const int cc[3], int m0, int m1, int m2) const {
    if (c->desc.scheme != NAT_SL) return cell * nl + m0 + np * (m1 + np * m2);
    const int64_t i = (int64_t)cc[0] * p + m0, j = (int64_t)cc[1] * p + m1, k = (int64_t)cc[2] * p + m2;
    return i + (int64_t)c->nfe[0] * (j + (int64_t)c->nfe[1] * k);
  }
  // map and Jacobian J[d][b] = d x_d / d xi_b at unit point xi
  void map(const int cc[3], const double xi[3], double X[3], double J[3][3]) const {
    const int nv = D == 3 ? 8 : 4;
    for (int a = 0; a < 3; a++) {
      X[a] = 0;
      for (int b = 0; b < 3; b++) J[a][b] = (a == b && a >= D) ? 1.0 : 0.0;
    }
    if (cartesian) {
      const double* lo = cell_vertex(cc, 0);
      const double* hi = cell_vertex(cc, nv - 1);
      for (int a = 0; a < D; a++) {
        X[a] = lo[a] + xi[a] * (hi[a] - lo[a]);
        J[a][a] = hi[a] - lo[a];
      }
      return;
    }
    for (int v = 0; v < nv; v++) {
      const double* V = cell_vertex(cc, v);
      double Nv = 1;
      for (int a = 0; a < D; a++) Nv *= ((v >> a) & 1) ? xi[a] : 1 - xi[a];
      for (int d = 0; d < D; d++) X[d] += Nv * V[d];
      for (int b = 0; b < D; b++) {
        double dN = 1;
        for (int a = 0; a < D; a++) {
          const int bit = (v >> a) & 1;
          dN *= (a == b) ? (bit ? 1.0 : -1.0) : (bit ? xi[a] : 1 - xi[a]);
        }
        for (int d = 0; d < D; d++) J[d][b] += dN * V[d];
      }
    }
  }
};

double det3(const double J[3][3]) {
  return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
         J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}

// inverse of a 3x3 matrix (2D problems carry a unit third row/column)
void inv3(const double J[3][3], double I[3][3]) {
  const double d = det3(J);
  I[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / d;
  I[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / d;
  I[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / d;
  I[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / d;
  I[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / d;
  I[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / d;
  I[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / d;
  I[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / d;
  I[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / d;
}

bool make_dir(const std::string& dir) {
  struct stat st;
  if (stat(dir.c_str(), &st) == 0) return S_ISDIR(st.st_mode);
  return mkdir(dir.c_str(), 0755) == 0;
}

}  // namespace

extern "C" {

int nat_stats(nat_ctx* ctx, int64_t iteration, nat_stats_t* out) {
  if (!out) return fail(ctx, NAT_ERR_ARG, "null output");
  const int D = ctx->D, Q = ctx->Q;
  const int64_t n = ctx->n_dofs;
  std::vector<double> rho(n), u((size_t)D * n), f((size_t)Q * n);
  int rc = nat_get_macro(ctx, rho.data(), u.data(), nullptr);
  if (rc) return rc;
  if ((rc = nat_download_f(ctx, f.data()))) return rc;
  double mass = 0, mom[3] = {0, 0, 0}, ekin = 0;
  if ((rc = nat_sums(ctx, &mass, mom, &ekin))) return rc;

  nat_stats_t s{};
  s.iteration = iteration;
  s.time = ctx->time;
  s.mass = mass;
  s.kin_e = ekin;
  // getMaxVelocityNorm: the largest component-wise maximum norm
  double rmin = 1e300, rmax = -1e300;
  for (int64_t i = 0; i < n; i++) {
    rmin = std::min(rmin, rho[i]);
    rmax = std::max(rmax, rho[i]);
    for (int d = 0; d < D; d++) s.max_u = std::max(s.max_u, std::fabs(u[(size_t)d * n + i]));
  }
  // maximalPressure: (rho - 1) c_s^2 with the scaled speed of sound
  s.max_p = (rmax - 1.0) * ctx->st.cs2;
  s.min_p = (rmin - 1.0) * ctx->st.cs2;

  // quadrature quantities on the GLL nodes of every cell
  Space sp(ctx);
  const int np = sp.np, nz = D == 3 ? np : 1;
  double ens = 0, ens_sq = 0, ent = 0, ux_sum = 0, area = 0;
  std::vector<double> uloc((size_t)D * sp.nl);
  std::vector<int64_t> gl(sp.nl);
  for (int64_t cell = 0; cell < sp.n_cells(); cell++) {
    int cc[3];
    sp.cell_xyz(cell, cc);
    for (int m2 = 0; m2 < nz; m2++)
      for (int m1 = 0; m1 < np; m1++)
        for (int m0 = 0; m0 < np; m0++) {
          const int l = m0 + np * (m1 + np * m2);
          gl[l] = sp.dof(cell, cc, m0, m1, m2);
          for (int d = 0; d < D; d++) uloc[(size_t)d * sp.nl + l] = u[(size_t)d * n + gl[l]];
        }
    double cell_jxw = 0, cell_ent = 0, cell_ux = 0;
    for (int l = 0; l < sp.nl; l++) {
      const int64_t g = gl[l];
      cell_ux += u[g];
      for (int q = 0; q < Q; q++) {
        const double fq = f[(size_t)q * n + g];
        cell_ent += fq * std::log(fq / ctx->st.w[q]);
      }
    }
    for (int q2 = 0; q2 < nz; q2++)
      for (int q1 = 0; q1 < np; q1++)
        for (int q0 = 0; q0 < np; q0++) {
          const int qi[3] = {q0, q1, q2};
          const double xi[3] = {sp.x[q0], sp.x[q1], D == 3 ? sp.x[q2] : 0.0};
          double X[3], J[3][3], Ji[3][3];
          sp.map(cc, xi, X, J);
          inv3(J, Ji);
          const double jxw = det3(J) * sp.w[q0] * sp.w[q1] * (D == 3 ? sp.w[q2] : 1.0);
          cell_jxw += jxw;
          // reference-space derivatives of every velocity component at this node
          double g[3][3] = {};  // g[c][k] = d u_c / d x_k
          for (int c = 0; c < D; c++) {
            double dxi[3] = {0, 0, 0};
            for (int b = 0; b < D; b++)
              for (int m = 0; m < np; m++) {
                int mi[3] = {q0, q1, q2};
                mi[b] = m;
                const int l = mi[0] + np * (mi[1] + np * (D == 3 ? mi[2] : 0));
                dxi[b] += sp.dm[qi[b]][m] * uloc[(size_t)c * sp.nl + l];
              }
            for (int k = 0; k < D; k++)
              for (int b = 0; b < D; b++) g[c][k] += Ji[b][k] * dxi[b];
          }
          double e2;
          if (D == 2) {
            const double w = g[1][0] - g[0][1];  // vorticity dv/dx - du/dy
            e2 = w * w;
          } else {
            e2 = 0;  // Frobenius norm of the velocity gradient
            for (int c = 0; c < 3; c++)
              for (int k = 0; k < 3; k++) e2 += g[c][k] * g[c][k];
          }
          ens += e2 * jxw;
          ens_sq += e2 * e2 * jxw;
        }
    // entropy and mean(ux) sum over quadrature points and cell DoFs independently
    // (PhysicalProperties.cpp:323-400)
    ent -= cell_ent * cell_jxw;
    ux_sum += cell_ux * cell_jxw;
    area += cell_jxw * sp.nl;
  }
  s.enstrophy = ens;
  s.enstrophy_sq = ens_sq;
  s.entropy = ent;
  s.mean_ux = ux_sum / area;

  // residuals at every 100th iteration against the fields of the previous such call, once
  // 100 iterations have passed since the first (SolverStats::calculateResiduals, called by the
  // stop check every 100 iterations, CFDSolver.cpp:927-931); a second call at the same
  // iteration keeps them
  if (iteration % 100 == 0 && !(ctx->res_have && iteration == ctx->res_last)) {
    if (!ctx->res_rho.empty() && iteration - ctx->res_first >= 100) {
      double rr = 0, ru = 0;
      for (int64_t i = 0; i < n; i++) {
        rr = std::max(rr, std::fabs(ctx->res_rho[i] - rho[i]));
        double s2 = 0;
        for (int d = 0; d < D; d++) {
          const double du = ctx->res_u[(size_t)d * n + i] - u[(size_t)d * n + i];
          s2 += du * du;
        }
        ru = std::max(ru, s2);
      }
      ctx->res_rho_val = rr;
      ctx->res_u_val = std::sqrt(ru);
    }
    if (ctx->res_rho.empty()) ctx->res_first = iteration;
    ctx->res_rho = rho;
    ctx->res_u = u;
    ctx->res_have = true;
    ctx->res_last = iteration;
  }
  s.res_rho = ctx->res_rho_val;
  s.res_u = ctx->res_u_val;
  *out = s;
  return 0;
}

int nat_write_stats(nat_ctx* ctx, const char* path, int64_t iteration, int truncate) {
  if (!path) return fail(ctx, NAT_ERR_ARG, "null path");
  nat_stats_t s;
  int rc = nat_stats(ctx, iteration, &s);
  if (rc) return rc;
  std::ofstream out(path, truncate ? std::ios::out : (std::ios::out | std::ios::app));
  if (!out) return fail(ctx, NAT_ERR_ARG, std::string("cannot open ") + path);
  if (truncate)
    out << "#  i      t      max |u_numeric|   mass  kinE  enstrophy   entropy   maxP     minP    "
           "#residuum(rho)   residuum(u)    mean(ux)   enstr^2  \n";
  out << s.iteration << " " << s.time << " " << s.max_u << " " << s.mass << " " << s.kin_e << " " << s.enstrophy
      << " " << s.entropy << " " << s.max_p << " " << s.min_p << " " << s.res_rho << " " << s.res_u << " "
      << s.mean_ux << " " << s.enstrophy_sq << "\n";
  return out ? 0 : fail(ctx, NAT_ERR_ARG, std::string("write failed: ") + path);
}

// base64 (RFC 4648) of a byte string
static std::string base64(const unsigned char* p, size_t n) {
  static const char* tab = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
  std::string out;
  out.reserve((n + 2) / 3 * 4);
  for (size_t i = 0; i < n; i += 3) {
    const uint32_t b0 = p[i], b1 = i + 1 < n ? p[i + 1] : 0, b2 = i + 2 < n ? p[i + 2] : 0;
    const uint32_t v = (b0 << 16) | (b1 << 8) | b2;
    out += tab[(v >> 18) & 63];
    out += tab[(v >> 12) & 63];
    out += i + 1 < n ? tab[(v >> 6) & 63] : '=';
    out += i + 2 < n ? tab[v & 63] : '=';
  }
  return out;
}

// One VTK zlib block: base64 of the header (blocks = 1, block size, last block size,
// compressed size, UInt32 each), then base64 of the compressed bytes.
static bool zlib_base64(const void* data, size_t bytes, std::string& out) {
  uLongf clen = compressBound((uLong)bytes);
  std::vector<unsigned char> buf(clen);
  if (compress2(buf.data(), &clen, (const Bytef*)data, (uLong)bytes, Z_BEST_COMPRESSION) != Z_OK) return false;
  const uint32_t hdr[4] = {1u, (uint32_t)bytes, (uint32_t)bytes, (uint32_t)clen};
  out = base64((const unsigned char*)hdr, sizeof(hdr)) + base64(buf.data(), clen);
  return true;
}

// Patches as DataOut::build_patches(n) lays them out: per cell (n+1)^D points on a uniform
// unit-cell grid (n = p + 1, or 1 for p = 1), values interpolated from the cell's DoFs,
// positions through the cell map; n^D linear sub-cells per patch. Point data rho, ux, uy
// (uz) [, T], cell data subdomain = 0.
int nat_write_vtu(nat_ctx* ctx, const char* dir, int64_t iteration) {
  if (!dir) return fail(ctx, NAT_ERR_ARG, "null directory");
  if (!make_dir(dir)) return fail(ctx, NAT_ERR_ARG, std::string("cannot create ") + dir);
  const int D = ctx->D;
  const int64_t n = ctx->n_dofs;
  std::vector<double> rho(n), u((size_t)D * n), T;
  if (ctx->desc.compressible) T.resize(n);
  int rc = nat_get_macro(ctx, rho.data(), u.data(), T.empty() ? nullptr : T.data());
  if (rc) return rc;
  Space sp(ctx);
  const int sub = sp.p == 1 ? 1 : sp.p + 1;
  const int pp = sub + 1, ppz = D == 3 ? pp : 1;
  const int64_t pts_per = (int64_t)pp * pp * ppz, cells_per = (int64_t)sub * sub * (D == 3 ? sub : 1);
  const int64_t ncell = sp.n_cells();
  // Lagrange values of every node at the uniform patch coordinates
  std::vector<double> L((size_t)pp * sp.np);
  for (int a = 0; a < pp; a++) {
    const double t = (double)a / sub;
    for (int j = 0; j < sp.np; j++) {
      double v = 1;
      for (int k = 0; k < sp.np; k++)
        if (k != j) v *= (t - sp.x[k]) / (sp.x[j] - sp.x[k]);
      L[(size_t)a * sp.np + j] = v;
    }
  }
  const int nfields = 1 + D + (T.empty() ? 0 : 1);
  std::vector<std::vector<double>> val(nfields, std::vector<double>((size_t)ncell * pts_per));
  std::vector<double> xyz((size_t)ncell * pts_per * 3);
  const int npz = D == 3 ? sp.np : 1;
  for (int64_t cell = 0; cell < ncell; cell++) {
    int cc[3];
    sp.cell_xyz(cell, cc);
    for (int c2 = 0; c2 < ppz; c2++)
      for (int c1 = 0; c1 < pp; c1++)
        for (int c0 = 0; c0 < pp; c0++) {
          const int64_t pi = cell * pts_per + c0 + pp * (c1 + (int64_t)pp * c2);
          const double xi[3] = {(double)c0 / sub, (double)c1 / sub, D == 3 ? (double)c2 / sub : 0.0};
          double X[3], J[3][3];
          sp.map(cc, xi, X, J);
          for (int d = 0; d < 3; d++) xyz[3 * pi + d] = d < D ? X[d] : 0.0;
          for (int fi = 0; fi < nfields; fi++) {
            double v = 0;
            for (int m2 = 0; m2 < npz; m2++)
              for (int m1 = 0; m1 < sp.np; m1++)
                for (int m0 = 0; m0 < sp.np; m0++) {
                  const double w = L[(size_t)c0 * sp.np + m0] * L[(size_t)c1 * sp.np + m1] *
                                   (D == 3 ? L[(size_t)c2 * sp.np + m2] : 1.0);
                  const int64_t g = sp.dof(cell, cc, m0, m1, m2);
                  const double fv = fi == 0 ? rho[g] : (fi <= D ? u[(size_t)(fi - 1) * n + g] : T[g]);
                  v += w * fv;
                }
            val[fi][pi] = v;
          }
        }
  }
  std::ostringstream name;
  name << "t_0." << iteration << ".vtu";
  const std::string path = std::string(dir) + "/" + name.str();
  std::ofstream o(path, std::ios::binary);
  if (!o) return fail(ctx, NAT_ERR_ARG, "cannot open " + path);
  const int64_t npts = ncell * pts_per, ncl = ncell * cells_per;
  const char* names[5] = {"rho", "ux", "uy", "uz", "T"};
  // The arrays go out as deal.II's write_vtu writes them with zlib: format="binary", each
  // array one zlib block, base64 of the 4 x UInt32 block header, then base64 of the block.
  std::string err;
  auto arr = [&](const char* type, const char* attrs, const void* data, size_t bytes) {
    std::string enc;
    if (!zlib_base64(data, bytes, enc)) err = "zlib compression failed";
    o << "<DataArray type=\"" << type << "\" " << attrs << " format=\"binary\">\n" << enc << "\n</DataArray>\n";
  };
  o << "<?xml version=\"1.0\"?>\n<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\" "
       "compressor=\"vtkZLibDataCompressor\">\n"
    << "<UnstructuredGrid>\n<Piece NumberOfPoints=\"" << npts << "\" NumberOfCells=\"" << ncl << "\">\n<Points>\n";
  arr("Float64", "NumberOfComponents=\"3\"", xyz.data(), xyz.size() * sizeof(double));
  o << "</Points>\n<Cells>\n";
  std::vector<int32_t> conn;
  conn.reserve((size_t)ncl * (D == 3 ? 8 : 4));
  for (int64_t cell = 0; cell < ncell; cell++)
    for (int s2 = 0; s2 < (D == 3 ? sub : 1); s2++)
      for (int s1 = 0; s1 < sub; s1++)
        for (int s0 = 0; s0 < sub; s0++) {
          auto P = [&](int a, int b, int c) {
            return (int32_t)(cell * pts_per + (s0 + a) + pp * ((s1 + b) + (int64_t)pp * (s2 + c)));
          };
          for (int32_t v : {P(0, 0, 0), P(1, 0, 0), P(1, 1, 0), P(0, 1, 0)}) conn.push_back(v);
          if (D == 3)
            for (int32_t v : {P(0, 0, 1), P(1, 0, 1), P(1, 1, 1), P(0, 1, 1)}) conn.push_back(v);
        }
  if (npts > INT32_MAX) return fail(ctx, NAT_ERR_UNSUPPORTED, "VTU output: more than 2^31 patch points");
  arr("Int32", "Name=\"connectivity\"", conn.data(), conn.size() * sizeof(int32_t));
  const int nvc = D == 3 ? 8 : 4;
  std::vector<int32_t> offs((size_t)ncl);
  for (int64_t i = 0; i < ncl; i++) offs[(size_t)i] = (int32_t)((i + 1) * nvc);
  arr("Int32", "Name=\"offsets\"", offs.data(), offs.size() * sizeof(int32_t));
  std::vector<uint8_t> types((size_t)ncl, (uint8_t)(D == 3 ? 12 : 9));
  arr("UInt8", "Name=\"types\"", types.data(), types.size());
  o << "</Cells>\n<PointData Scalars=\"rho\">\n";
  for (int fi = 0; fi < nfields; fi++) {
    const std::string nm = std::string("Name=\"") + (fi == 0 ? names[0] : (fi <= D ? names[fi] : names[4])) + "\"";
    arr("Float64", nm.c_str(), val[fi].data(), val[fi].size() * sizeof(double));
  }
  o << "</PointData>\n<CellData>\n";
  std::vector<double> subdomain((size_t)ncl, 0.0);
  arr("Float64", "Name=\"subdomain\"", subdomain.data(), subdomain.size() * sizeof(double));
  o << "</CellData>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n";
  if (!err.empty()) return fail(ctx, NAT_ERR_ARG, err);
  if (!o) return fail(ctx, NAT_ERR_ARG, "write failed: " + path);
  // master record listing the single piece
  std::ostringstream pname;
  pname << std::string(dir) << "/t_0." << iteration << ".pvtu";
  std::ofstream pv(pname.str());
  pv << "<?xml version=\"1.0\"?>\n<VTKFile type=\"PUnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\" "
        "compressor=\"vtkZLibDataCompressor\">\n"
     << "<PUnstructuredGrid GhostLevel=\"0\">\n<PPointData Scalars=\"rho\">\n";
  for (int fi = 0; fi < nfields; fi++)
    pv << "<PDataArray type=\"Float64\" Name=\"" << (fi == 0 ? names[0] : (fi <= D ? names[fi] : names[4]))
       << "\"/>\n";
  pv << "</PPointData>\n<PCellData>\n<PDataArray type=\"Float64\" Name=\"subdomain\"/>\n</PCellData>\n"
     << "<PPoints>\n<PDataArray type=\"Float64\" NumberOfComponents=\"3\"/>\n</PPoints>\n"
     << "<Piece Source=\"" << name.str() << "\"/>\n</PUnstructuredGrid>\n</VTKFile>\n";
  return pv ? 0 : fail(ctx, NAT_ERR_ARG, "write failed: " + pname.str());
}

// Data file: magic "NATCKPT1", int32 dim, Q, order, precision flag, int64 n_dofs, then
// Q x n_dofs FP64 populations in DoF order.
static int write_pop(nat_ctx* ctx, const std::string& stat, const std::string& data, const std::vector<double>& f,
                     int64_t iteration) {
  {
    std::ofstream s(stat);
    if (!s) return fail(ctx, NAT_ERR_ARG, "cannot open " + stat);
    s.precision(17);
    s << iteration << "\n" << ctx->time << "\n" << ctx->st.s << "\n" << ctx->p << "\n";
    if (!s) return fail(ctx, NAT_ERR_ARG, "write failed: " + stat);
  }
  std::ofstream d(data, std::ios::binary);
  if (!d) return fail(ctx, NAT_ERR_ARG, "cannot open " + data);
  const int32_t hdr[4] = {ctx->D, ctx->Q, ctx->p, ctx->fp32 ? 1 : 0};
  d.write("NATCKPT1", 8);
  d.write((const char*)hdr, sizeof(hdr));
  d.write((const char*)&ctx->n_dofs, sizeof(int64_t));
  d.write((const char*)f.data(), (std::streamsize)(f.size() * sizeof(double)));
  return d ? 0 : fail(ctx, NAT_ERR_ARG, "write failed: " + data);
}

int nat_checkpoint_write(nat_ctx* ctx, const char* dir, int64_t iteration) {
  if (!dir) return fail(ctx, NAT_ERR_ARG, "null directory");
  if (!make_dir(dir)) return fail(ctx, NAT_ERR_ARG, std::string("cannot create ") + dir);
  std::vector<double> f((size_t)ctx->Q * ctx->n_dofs);
  int rc = nat_download_f(ctx, f.data());
  if (rc) return rc;
  const std::string base = std::string(dir) + "/";
  const std::string it = std::to_string(iteration);
  if ((rc = write_pop(ctx, base + "checkpoint_" + it + ".stat", base + "checkpoint_" + it + ".data", f, iteration)))
    return rc;
  if (ctx->desc.compressible) {
    if ((rc = nat_download_g(ctx, f.data()))) return rc;
    if ((rc = write_pop(ctx, base + "checkpointG_" + it + ".stat", base + "checkpointG_" + it + ".data", f,
                        iteration)))
      return rc;
  }
  return 0;
}

static int read_pop(nat_ctx* ctx, const std::string& stat, const std::string& data, std::vector<double>& f,
                    int64_t* iteration, double* time) {
  std::ifstream s(stat);
  if (!s) return fail(ctx, NAT_ERR_ARG, "cannot open " + stat);
  int64_t it = 0;
  double t = 0, scaling = 0;
  int order = 0;
  s >> it >> t >> scaling >> order;
  if (!s) return fail(ctx, NAT_ERR_ARG, "malformed status file " + stat);
  if (order != ctx->p)
    return fail(ctx, NAT_ERR_UNSUPPORTED, "checkpoint of another FE order: interpolation between orders is not built");
  if (std::fabs(scaling - ctx->st.s) > 1e-12 * std::fabs(scaling))
    return fail(ctx, NAT_ERR_UNSUPPORTED, "checkpoint of another stencil scaling: rescaling is not built");
  std::ifstream d(data, std::ios::binary);
  if (!d) return fail(ctx, NAT_ERR_ARG, "cannot open " + data);
  char magic[8];
  int32_t hdr[4];
  int64_t n = 0;
  d.read(magic, 8);
  d.read((char*)hdr, sizeof(hdr));
  d.read((char*)&n, sizeof(n));
  if (!d || std::memcmp(magic, "NATCKPT1", 8) != 0) return fail(ctx, NAT_ERR_ARG, "not a checkpoint data file: " + data);
  if (hdr[0] != ctx->D || hdr[1] != ctx->Q || hdr[2] != ctx->p || n != ctx->n_dofs)
    return fail(ctx, NAT_ERR_ARG, "checkpoint does not match this context (dim, Q, order or DoF count)");
  f.resize((size_t)ctx->Q * ctx->n_dofs);
  d.read((char*)f.data(), (std::streamsize)(f.size() * sizeof(double)));
  if (!d) return fail(ctx, NAT_ERR_ARG, "truncated checkpoint data file: " + data);
  *iteration = it;
  *time = t;
  return 0;
}

int nat_checkpoint_read(nat_ctx* ctx, const char* dir, int64_t iteration, int64_t* iteration_out) {
  if (!dir) return fail(ctx, NAT_ERR_ARG, "null directory");
  int64_t it = iteration;
  if (iteration == 1) {
    // the most recent checkpoint_<i>.stat in the directory (Checkpoint.cpp:21-45)
    bool found = false;
    DIR* dd = opendir(dir);
    if (!dd) return fail(ctx, NAT_ERR_ARG, std::string("cannot open ") + dir);
    while (dirent* e = readdir(dd)) {
      const std::string nm = e->d_name;
      if (nm.size() > 16 && nm.compare(0, 11, "checkpoint_") == 0 && nm.compare(nm.size() - 5, 5, ".stat") == 0) {
        const std::string num = nm.substr(11, nm.size() - 16);
        if (num.empty() || num.find_first_not_of("0123456789") != std::string::npos) continue;
        const int64_t i = std::stoll(num);
        if (i > it) it = i;
        found = true;
      }
    }
    closedir(dd);
    if (!found) return fail(ctx, NAT_ERR_ARG, std::string("no checkpoint in ") + dir);
  }
  const std::string base = std::string(dir) + "/";
  const std::string its = std::to_string(it);
  std::vector<double> f;
  int64_t iter = 0;
  double t = 0;
  int rc = read_pop(ctx, base + "checkpoint_" + its + ".stat", base + "checkpoint_" + its + ".data", f, &iter, &t);
  if (rc) return rc;
  std::vector<double> g;
  if (ctx->desc.compressible) {
    int64_t ig = 0;
    double tg = 0;
    rc = read_pop(ctx, base + "checkpointG_" + its + ".stat", base + "checkpointG_" + its + ".data", g, &ig, &tg);
    if (rc) return rc;
  }
  if ((rc = nat_upload_f(ctx, f.data()))) return rc;
  if (!g.empty() && (rc = nat_upload_g(ctx, g.data()))) return rc;
  ctx->time = t;
  if (iteration_out) *iteration_out = iter;
  return 0;
}

}  // extern "C"
