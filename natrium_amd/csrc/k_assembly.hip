// FP64 operator assembly on the GPU: departure points x_i - dt c_alpha, the
// face walk to the containing cell, and the unit coordinates there.
//
// Follows the reference semi-Lagrangian assembly:
//   advection/SemiLagrangian.cpp:196-245  departure point per (DoF, direction), start cell
//   advection/SemiLagrangian.cpp:246-443  path walk: life-time limit 50, periodic wrap, interior faces
//   advection/SemiLagrangian.cpp:358-385  velocity-neq bounce-back: reverse the path at a wall
//   boundaries/VelocityNeqBounceBack.cpp:137-195 value added per wall hit
//   advection/SemiLagrangian.cpp:525-699  faceCrossedFirst: unit-cell clamps (1e-10), x<y<z tie order
//   advection/SemiLagrangian.cpp:702-757  faceCrossedFirstWithoutMapping (inverse-map failure)
//   advection/SemiLagrangianTools.cpp:39-52 unit point of the departure point, clamped
//   boundaries/PeriodicBoundary.cpp:261-307 periodic coordinate offset
#include "nat_internal.hpp"

namespace nat {

__device__ __forceinline__ void clamp01(double& v) {
  if (fabs(v) < 1e-10) v = 0;
  if (fabs(v - 1) < 1e-10) v = 1;
}

__device__ __forceinline__ const double* vtx(const MeshDev& m, int i, int j, int k) {
  return m.vert + 3 * ((size_t)i + (size_t)(m.N[0] + 1) * ((size_t)j + (size_t)(m.N[1] + 1) * k));
}

template <int D>
__device__ __forceinline__ void cell_xyz(const MeshDev& m, int cell, int c[3]) {
  c[0] = cell % m.N[0];
  c[1] = (cell / m.N[0]) % m.N[1];
  c[2] = D == 3 ? cell / (m.N[0] * m.N[1]) : 0;
}

template <int D>
__device__ __forceinline__ const double* cell_vertex(const MeshDev& m, const int c[3], int v) {
  return vtx(m, c[0] + (v & 1), c[1] + ((v >> 1) & 1), D == 3 ? c[2] + ((v >> 2) & 1) : 0);
}

template <int D>
__device__ void map_fwd(const MeshDev& m, const int c[3], const double xi[3], double x[3]) {
  const int nv = D == 3 ? 8 : 4;
  if (m.cartesian) {
    const double* a = cell_vertex<D>(m, c, 0);
    const double* b = cell_vertex<D>(m, c, nv - 1);
    for (int d = 0; d < 3; d++) x[d] = d < D ? a[d] + xi[d] * (b[d] - a[d]) : 0.0;
    return;
  }
  x[0] = x[1] = x[2] = 0;
  for (int v = 0; v < nv; v++) {
    double Nv = 1;
    for (int a = 0; a < D; a++) Nv *= ((v >> a) & 1) ? xi[a] : (1 - xi[a]);
    const double* X = cell_vertex<D>(m, c, v);
    for (int d = 0; d < D; d++) x[d] += Nv * X[d];
  }
}

template <int D>
__device__ void jac(const MeshDev& m, const int c[3], const double xi[3], double J[3][3]) {
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) J[a][b] = (a == b && a >= D) ? 1.0 : 0.0;
  const int nv = D == 3 ? 8 : 4;
  for (int v = 0; v < nv; v++) {
    const double* X = cell_vertex<D>(m, c, v);
    for (int b = 0; b < D; b++) {
      double dN = 1;
      for (int a = 0; a < D; a++) {
        const int bit = (v >> a) & 1;
        dN *= (a == b) ? (bit ? 1.0 : -1.0) : (bit ? xi[a] : 1 - xi[a]);
      }
      for (int d = 0; d < D; d++) J[d][b] += dN * X[d];
    }
  }
}

// inverse map; false when the Newton iteration of the multilinear map fails
template <int D>
__device__ bool map_inv(const MeshDev& m, const int c[3], const double x[3], double xi[3]) {
  const int nv = D == 3 ? 8 : 4;
  if (m.cartesian) {
    const double* a = cell_vertex<D>(m, c, 0);
    const double* b = cell_vertex<D>(m, c, nv - 1);
    for (int d = 0; d < 3; d++) xi[d] = d < D ? (x[d] - a[d]) / (b[d] - a[d]) : 0.0;
    return true;
  }
  for (int d = 0; d < 3; d++) xi[d] = d < D ? 0.5 : 0.0;
  double prev = 1e300;
  for (int it = 0; it < 100; it++) {
    double y[3], J[3][3], r[3], s[3];
    map_fwd<D>(m, c, xi, y);
    for (int d = 0; d < 3; d++) r[d] = d < D ? y[d] - x[d] : 0.0;
    jac<D>(m, c, xi, J);
    if (D == 2) {
      const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      if (det == 0) return false;
      s[0] = (r[0] * J[1][1] - J[0][1] * r[1]) / det;
      s[1] = (J[0][0] * r[1] - r[0] * J[1][0]) / det;
    } else {
      const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
      const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
      const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
      const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
      if (det == 0) return false;
      s[0] = (c00 * r[0] + (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * r[1] + (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * r[2]) / det;
      s[1] = (c01 * r[0] + (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * r[1] + (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * r[2]) / det;
      s[2] = (c02 * r[0] + (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * r[1] + (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * r[2]) / det;
    }
    double mx = 0;
    for (int d = 0; d < D; d++) {
      xi[d] -= s[d];
      mx = fmax(mx, fabs(s[d]));
    }
    if (mx < 1e-15) return true;
    // converged to the rounding level of the residual: the step no longer shrinks
    if (mx < 1e-12 && mx > 0.25 * prev) return true;
    prev = mx;
    if (fabs(xi[0]) > 1e3) return false;
  }
  return true;
}

// neighbour across face f; -1 at a non-periodic boundary
template <int D>
__device__ int neighbor(const MeshDev& m, const int c[3], int f, bool* wrapped, int nc[3]) {
  const int a = f >> 1;
  nc[0] = c[0];
  nc[1] = c[1];
  nc[2] = c[2];
  nc[a] += (f & 1) ? 1 : -1;
  *wrapped = false;
  if (nc[a] < 0 || nc[a] >= m.N[a]) {
    if (!m.periodic[a]) return -1;
    nc[a] = nc[a] < 0 ? m.N[a] - 1 : 0;
    *wrapped = true;
  }
  return nc[0] + m.N[0] * (nc[1] + m.N[1] * nc[2]);
}

template <int D>
__device__ void periodic_shift(const MeshDev& m, const int c[3], int f, double p[3]) {
  const int a = f >> 1;
  int lo[3] = {c[0], c[1], c[2]}, hi[3] = {c[0], c[1], c[2]};
  lo[a] = 0;
  hi[a] = m.N[a];
  const double* vlo = vtx(m, lo[0], lo[1], D == 3 ? lo[2] : 0);
  const double* vhi = vtx(m, hi[0], hi[1], D == 3 ? hi[2] : 0);
  for (int d = 0; d < D; d++) {
    const double off = vlo[d] - vhi[d];
    p[d] = (f & 1) ? p[d] + off : p[d] - off;
  }
}

template <int D>
__device__ int crossed_without_mapping(const MeshDev& m, const int c[3], const double pin[3], const double pout[3],
                                       double pb[3]) {
  double lmin = 1;
  int imin = 0;
  for (int f = 0; f < 2 * D; f++) {
    const int a = f >> 1, side = f & 1;
    int fv[4], k = 0;
    for (int v = 0; v < (D == 3 ? 8 : 4); v++)
      if (((v >> a) & 1) == side) fv[k++] = v;
    const double* v0 = cell_vertex<D>(m, c, fv[0]);
    const double* v1 = cell_vertex<D>(m, c, fv[1]);
    double n[3];
    if (D == 2) {
      n[0] = v1[1] - v0[1];
      n[1] = -(v1[0] - v0[0]);
      n[2] = 0;
    } else {
      const double* v2 = cell_vertex<D>(m, c, fv[2]);
      const double d1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
      const double d2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
      n[0] = d1[1] * d2[2] - d1[2] * d2[1];
      n[1] = d1[2] * d2[0] - d1[0] * d2[2];
      n[2] = d1[0] * d2[1] - d1[1] * d2[0];
    }
    double num = 0, den = 0;
    for (int d = 0; d < D; d++) {
      num += (pin[d] - v0[d]) * n[d];
      den += (pin[d] - pout[d]) * n[d];
    }
    const double lam = num / den;
    if (lam < 0) continue;
    if (lam < lmin) {
      lmin = lam;
      imin = f;
    }
  }
  if (lmin == 1) return -1;
  for (int d = 0; d < 3; d++) pb[d] = pin[d] + (pout[d] - pin[d]) * lmin;
  return imin;
}

template <int D>
__device__ int crossed_first(const MeshDev& m, const int c[3], const double pin[3], const double pout[3], double pb[3]) {
  double pi[3], po[3];
  if (!map_inv<D>(m, c, pin, pi) || !map_inv<D>(m, c, pout, po)) return crossed_without_mapping<D>(m, c, pin, pout, pb);
  for (int i = 0; i < D; i++) {
    clamp01(pi[i]);
    clamp01(po[i]);
  }
  int face = -1;
  double lambda = 100;
  for (int a = 0; a < D; a++) {
    double lam;
    if (po[a] < 0)
      lam = (0 - pi[a]) / (po[a] - pi[a]);
    else if (po[a] > 1)
      lam = (1 - pi[a]) / (po[a] - pi[a]);
    else
      continue;
    if (lam < lambda) {
      lambda = lam;
      face = 2 * a + (po[a] > 1 ? 1 : 0);
    }
  }
  if (face < 0) return -1;
  double h[3] = {0, 0, 0};
  for (int i = 0; i < D; i++) {
    h[i] = pi[i] + (po[i] - pi[i]) * lambda;
    clamp01(h[i]);
  }
  map_fwd<D>(m, c, h, pb);
  return face;
}

// Walk from the owning cell to the cell holding x - dt c. Returns the cell index,
// -1 when the life-time limit is hit (identity row), -3 when the final unit point
// is outside its cell. At a wall face the path turns into the opposite direction
// at the hit point and continues for the remaining distance in the same cell;
// *nhit counts the wall hits (an odd count reads the opposite direction) and *bias
// accumulates 2 w (e . u_wall) / cs2 (rho = 1, destination direction e) per hit.
template <int D>
__device__ int locate(const MeshDev& m, int owner, const double x[3], const double disp[3], const double e[3],
                      double w, double cs2, double xi[3], int* nhit, double* bias) {
  double cur[3] = {x[0], x[1], x[2]};
  double dep[3] = {x[0] + disp[0], x[1] + disp[1], x[2] + disp[2]};
  int c[3];
  cell_xyz<D>(m, owner, c);
  int cell = owner;
  double sgn = 1.0;  // current direction = sgn * e
  for (int life = 1;; life++) {
    if (life > 50) return -1;
    double pb[3];
    const int face = crossed_first<D>(m, c, cur, dep, pb);
    if (face < 0) break;
    bool wrapped;
    int nc[3];
    const int nb = neighbor<D>(m, c, face, &wrapped, nc);
    if (nb < 0) {
      sgn = -sgn;
      double vel = 0, dist = 0;
      for (int d = 0; d < D; d++) {
        const double ed = sgn * e[d];
        vel += ed * ed;
        dist += (dep[d] - pb[d]) * (dep[d] - pb[d]);
      }
      vel = sqrt(vel);
      dist = sqrt(dist);
      for (int d = 0; d < D; d++) {
        cur[d] = pb[d];
        dep[d] = pb[d] - (sgn * e[d]) * dist / vel;
      }
      double exu = 0;
      for (int d = 0; d < D; d++) exu += m.wall_u[face][d] * e[d];
      *bias = *bias + 2 * w * 1.0 * exu / cs2;
      (*nhit)++;
      continue;
    }
    if (wrapped) {
      periodic_shift<D>(m, c, face, pb);
      periodic_shift<D>(m, c, face, dep);
    }
    for (int d = 0; d < 3; d++) cur[d] = pb[d];
    c[0] = nc[0];
    c[1] = nc[1];
    c[2] = nc[2];
    cell = nb;
  }
  if (!map_inv<D>(m, c, dep, xi)) return -3;
  for (int d = 0; d < D; d++) {
    clamp01(xi[d]);
    if (xi[d] < 0 || xi[d] > 1) return -3;
  }
  if (D == 2) xi[2] = 0;
  return cell;
}

// One thread per (streamed direction, DoF). dirw holds per streamed direction the
// scaled velocity (3) and the weight; with walls, flipmask[i] collects bit alpha
// for a reversed source direction and bit 0 for any hit, bias_out the added value.
template <int D>
__global__ __launch_bounds__(128) void k_assemble_general(MeshDev m, int64_t n, int Qm1, const double* __restrict__ coords,
                                   const int* __restrict__ owner, const double* __restrict__ disp /* Qm1 x 3 */,
                                   const double* __restrict__ dirw /* Qm1 x 4 */, double cs2,
                                   int* __restrict__ cell_out, double* __restrict__ xi_out,
                                   unsigned* __restrict__ flipmask, double* __restrict__ bias_out,
                                   int* __restrict__ status) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * Qm1) return;
  const int a = (int)(t / n);
  const int64_t i = t - (int64_t)a * n;
  const double x[3] = {coords[3 * i], coords[3 * i + 1], coords[3 * i + 2]};
  const double dp[3] = {disp[3 * a], disp[3 * a + 1], disp[3 * a + 2]};
  const double e[3] = {dirw[4 * a], dirw[4 * a + 1], dirw[4 * a + 2]};
  double xi[3] = {0, 0, 0};
  int nhit = 0;
  double bias = 0;
  const int c = locate<D>(m, owner[i], x, dp, e, dirw[4 * a + 3], cs2, xi, &nhit, &bias);
  cell_out[t] = c;
  for (int d = 0; d < D; d++) xi_out[t * D + d] = xi[d];
  if (flipmask) {
    bias_out[t] = bias;
    // an identity row (life-time limit) keeps direction alpha; its hits still add
    const bool flip = (nhit & 1) && c >= 0;
    const unsigned bits = (flip ? (1u << (a + 1)) : 0u) | (nhit ? 1u : 0u);
    if (bits) atomicOr(&flipmask[i], bits);
  } else if (nhit) {
    atomicOr(status, 2);
  }
  if (c == -3) atomicOr(status, 4);
  if (c == -1) atomicOr(status, 8);
}

// Compact wall table: for the k-th wall node list[k], its flip mask and the
// per-direction added values (Q per entry, direction 0 unused).
__global__ void k_wall_compact(int64_t n, int Q, const int* __restrict__ list, int64_t nwall,
                               const unsigned* __restrict__ flipmask, const double* __restrict__ bias_rec,
                               unsigned* __restrict__ wall_flip, double* __restrict__ wall_bias) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nwall) return;
  const int64_t i = list[k];
  wall_flip[k] = flipmask[i];
  wall_bias[k * Q] = 0;
  for (int a = 1; a < Q; a++) wall_bias[k * Q + a] = bias_rec[(int64_t)(a - 1) * n + i];
}

// Lagrange basis on GLL nodes: value of basis j at xi (product form).
__device__ __forceinline__ double lagrange(const double* nodes, int np, int j, double xi) {
  double num = 1, den = 1;
  for (int m = 0; m < np; m++)
    if (m != j) {
      num *= xi - nodes[m];
      den *= nodes[j] - nodes[m];
    }
  return num / den;
}

// Structured tables: for axis a and signed displacement d (one per shift id),
// class j of a node (position j in its cell) -> source cell offset and p+1 weights.
// One thread per (axis, shift, class); the representative node sits in the
// middle cell of the axis, so no periodic wrap is involved.
__global__ void k_assemble_axis(int D, const int* N, const double* vert_axis /* [3][MAXC+1] */, int maxc,
                                const double* gll, int p, const double* disp /* [3][MAXSHIFT] */,
                                const int* nshift, double* w_out /* [3][MAXSHIFT][MAXP][MAXP+1] */,
                                int* off_out /* [3][MAXSHIFT][MAXP] */, int* status) {
  const int t = threadIdx.x;
  const int a = t / (MAXSHIFT * MAXP);
  const int s = (t / MAXP) % MAXSHIFT;
  const int j = t % MAXP;
  if (a >= D || s >= nshift[a] || j >= p) return;
  const double* X = vert_axis + a * (maxc + 1);
  const int n = N[a];
  const int c0 = n / 2;
  // node position from its owning cell (the lower cell for a vertex node)
  int own = (j == 0) ? c0 - 1 : c0;
  if (own < 0) own = 0;
  const double xi_node = (j == 0 && own == c0 - 1) ? 1.0 : gll[j];
  const double x = X[own] + xi_node * (X[own + 1] - X[own]);
  const double dep = x + disp[a * MAXSHIFT + s];
  int cell = own;
  double xi = 0;
  for (int it = 0; it < 64; it++) {
    xi = (dep - X[cell]) / (X[cell + 1] - X[cell]);
    clamp01(xi);
    if (xi < 0) {
      cell--;
    } else if (xi > 1) {
      cell++;
    } else {
      break;
    }
    if (cell < 0 || cell >= n) {
      atomicOr(status, 16);
      return;
    }
  }
  const int o = cell - c0;
  off_out[(a * MAXSHIFT + s) * MAXP + j] = o;
  for (int m = 0; m <= p; m++) w_out[((a * MAXSHIFT + s) * MAXP + j) * (MAXP + 1) + m] = lagrange(gll, p + 1, m, xi);
}

}  // namespace nat

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
namespace nat {

// FE_Q support points of the lattice DoFs (x fastest; SemiLagrangian.cpp:214-229) through
// the cell map. Each DoF is owned by the lowest cell that contains it: a DoF at index
// k > 0 of an axis belongs to cell (k - 1) / p. The map is x = sum_v N_v X_v with
// N_v = prod (xi or 1 - xi), or x = a + xi (b - a) on Cartesian cells. Contraction is off
// in this kernel: every product and sum is rounded on its own, as a CPU without fused
// multiply-add forms them, so the points match a CPU evaluation of the same formula bit
// for bit.
template <int D>
__global__ __launch_bounds__(256) void k_support_points(MeshDev m, int p, GllTab g, int nfe0, int nfe1, int64_t n,
                                                        double* __restrict__ xyz, int* __restrict__ owner) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int idx[3] = {(int)(i % nfe0), (int)((i / nfe0) % nfe1), D == 3 ? (int)(i / ((int64_t)nfe0 * nfe1)) : 0};
  int cc[3] = {0, 0, 0};
  double xi[3] = {0, 0, 0};
#pragma unroll
  for (int a = 0; a < D; a++) {
    cc[a] = idx[a] == 0 ? 0 : (idx[a] - 1) / p;
    xi[a] = g.x[idx[a] - cc[a] * p];
  }
  double x[3] = {0, 0, 0};
  constexpr int nv = D == 3 ? 8 : 4;
  if (m.cartesian) {
    const double* a0 = cell_vertex<D>(m, cc, 0);
    const double* b0 = cell_vertex<D>(m, cc, nv - 1);
#pragma unroll
    for (int d = 0; d < D; d++) x[d] = a0[d] + xi[d] * (b0[d] - a0[d]);
  } else {
#pragma unroll
    for (int v = 0; v < nv; v++) {
      double Nv = 1;
#pragma unroll
      for (int a = 0; a < D; a++) Nv *= ((v >> a) & 1) ? xi[a] : (1 - xi[a]);
      const double* X = cell_vertex<D>(m, cc, v);
#pragma unroll
      for (int d = 0; d < D; d++) x[d] += Nv * X[d];
    }
  }
  xyz[3 * i] = x[0];
  xyz[3 * i + 1] = x[1];
  xyz[3 * i + 2] = x[2];
  owner[i] = cc[0] + m.N[0] * (cc[1] + m.N[1] * cc[2]);
}

}  // namespace nat

int nat_launch_support_points(nat_ctx* ctx) {
  nat::MeshDev m;
  m.D = ctx->D;
  for (int a = 0; a < 3; a++) {
    m.N[a] = a < ctx->D ? ctx->desc.ncells[a] : 1;
    m.periodic[a] = a < ctx->D ? ctx->desc.periodic[a] : 1;
  }
  m.cartesian = ctx->desc.cartesian;
  m.vert = ctx->d_vert;
  nat::GllTab g;
  double w[nat::MAXP + 1];
  nat::gll_points(ctx->p, g.x, w);
  const int bs = 256;
  const unsigned grid = (unsigned)((ctx->n_dofs + bs - 1) / bs);
  if (ctx->D == 3)
    hipLaunchKernelGGL(nat::k_support_points<3>, dim3(grid), dim3(bs), 0, ctx->stream, m, ctx->p, g, ctx->nfe[0],
                       ctx->nfe[1], ctx->n_dofs, ctx->d_coords, ctx->d_owner);
  else
    hipLaunchKernelGGL(nat::k_support_points<2>, dim3(grid), dim3(bs), 0, ctx->stream, m, ctx->p, g, ctx->nfe[0],
                       ctx->nfe[1], ctx->n_dofs, ctx->d_coords, ctx->d_owner);
  NAT_HIP(hipGetLastError());
  return 0;
}

int nat_launch_assemble_general(nat_ctx* ctx, const double* d_disp, const double* d_dirw, int* d_cell, double* d_xi,
                                unsigned* d_flipmask, double* d_bias) {
  nat::MeshDev m;
  m.D = ctx->D;
  for (int a = 0; a < 3; a++) {
    m.N[a] = a < ctx->D ? ctx->desc.ncells[a] : 1;
    m.periodic[a] = a < ctx->D ? ctx->desc.periodic[a] : 1;
  }
  for (int f = 0; f < 6; f++)
    for (int d = 0; d < 3; d++) m.wall_u[f][d] = d < ctx->D ? ctx->desc.wall_velocity[f][d] : 0.0;
  m.cartesian = ctx->desc.cartesian;
  m.vert = ctx->d_vert;
  const int64_t total = ctx->n_dofs * (ctx->Q - 1);
  const int bs = 128;
  if (ctx->D == 3)
    hipLaunchKernelGGL(nat::k_assemble_general<3>, dim3((unsigned)((total + bs - 1) / bs)), dim3(bs), 0, ctx->stream,
                       m, ctx->n_dofs, ctx->Q - 1, ctx->d_coords, ctx->d_owner, d_disp, d_dirw, ctx->st.cs2, d_cell,
                       d_xi, d_flipmask, d_bias, ctx->d_status);
  else
    hipLaunchKernelGGL(nat::k_assemble_general<2>, dim3((unsigned)((total + bs - 1) / bs)), dim3(bs), 0, ctx->stream,
                       m, ctx->n_dofs, ctx->Q - 1, ctx->d_coords, ctx->d_owner, d_disp, d_dirw, ctx->st.cs2, d_cell,
                       d_xi, d_flipmask, d_bias, ctx->d_status);
  NAT_HIP(hipGetLastError());
  return 0;
}

int nat_launch_wall_compact(nat_ctx* ctx, const int* d_list, const unsigned* d_flipmask, const double* d_bias) {
  if (ctx->n_wall == 0) return 0;
  const int bs = 256;
  hipLaunchKernelGGL(nat::k_wall_compact, dim3((unsigned)((ctx->n_wall + bs - 1) / bs)), dim3(bs), 0, ctx->stream,
                     ctx->n_dofs, ctx->Q, d_list, ctx->n_wall, d_flipmask, d_bias, ctx->wall_flip, ctx->wall_bias);
  NAT_HIP(hipGetLastError());
  return 0;
}

int nat_launch_assemble_axis(nat_ctx* ctx, const int* d_N, const double* d_vaxis, int maxc, const double* d_disp,
                             const int* d_nshift, double* d_w, int* d_off) {
  hipLaunchKernelGGL(nat::k_assemble_axis, dim3(1), dim3(3 * nat::MAXSHIFT * nat::MAXP), 0, ctx->stream, ctx->D, d_N,
                     d_vaxis, maxc, ctx->d_gll, ctx->p, d_disp, d_nshift, d_w, d_off, ctx->d_status);
  NAT_HIP(hipGetLastError());
  return 0;
}

namespace nat {
// FP32 departure records: code = cell * (p+1)^D + anchor (the GLL node nearest
// to the departure point, per axis, x fastest) and dxi = xi - x_anchor per axis.
// The small offset keeps FP32 relative precision where the weights need it.
// Compact records for FP32 storage: the departure cell and nearest GLL node packed in
// one int32, and the unit-coordinate offset from that node in fixed point (2^-30
// units: |offset| <= 1/2, resolution 9.3e-10), from which the kernel rebuilds FP64
// weights.
__global__ void k_records_f32(const int* __restrict__ cell, const double* __restrict__ xi, int64_t nrec, int D,
                              int p, GllTab g, int* __restrict__ rec) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrec) return;
  const int c = cell[r];
  int lin = 0, mult = 1;
  for (int d = 0; d < D; d++) {
    const double x = xi[r * D + d];
    int best = 0;
    for (int m = 1; m <= p; m++)
      if (fabs(x - g.x[m]) < fabs(x - g.x[best])) best = m;
    rec[r * (D + 1) + 1 + d] = (int)llrint((x - g.x[best]) * 1073741824.0);
    lin += best * mult;
    mult *= p + 1;
  }
  rec[r * (D + 1)] = c < 0 ? -1 : c * mult + lin;
}
}  // namespace nat

int nat_launch_records_f32(nat_ctx* ctx, int64_t nrec) {
  nat::GllTab g;
  double w[nat::MAXP + 1];
  nat::gll_points(ctx->p, g.x, w);
  const int bs = 256;
  hipLaunchKernelGGL(nat::k_records_f32, dim3((unsigned)((nrec + bs - 1) / bs)), dim3(bs), 0, ctx->stream,
                     ctx->dep_cell, ctx->dep_xi64, nrec, ctx->D, ctx->p, g, ctx->dep_rec32);
  NAT_HIP(hipGetLastError());
  return 0;
}
