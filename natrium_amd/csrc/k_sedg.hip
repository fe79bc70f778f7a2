// SEDG (Min-Lee) streaming operator inside the low-storage RK5 integrator,
// matrix-free.
//
// The reference assembles, per direction alpha, a block sparse matrix
// (advection/SEDGMinLee.cpp:321-536): on GLL collocation nodes the mass matrix is
// diagonal, M_i = JxW_i, and the row of node i is
//   volume:  -(e . D)_ij / M_i = -e . grad(phi_j)(x_i)
//   faces:   for every cell face through node i with e.n < 0 (upwind)
//            +JxW_f (e.n) / M_i on the diagonal and -JxW_f (e.n) / M_i on the
//            neighbour's coincident node; central flux: +1/2 of it on both.
// With the multilinear map (the DG space always uses MappingQ1) this is, per node,
// with G_b = sum_k e_k (J^-1)_bk:
//   (L f)_i = -sum_b G_b (d f / d xi_b)(xi_i)
//             + sum_{faces a} s_a G_a / w_a(i) * (f_i - f_nb)   [s_a G_a < 0]
// where s_a = -1 / +1 on the lower / upper face and w_a(i) is the 1-D GLL weight of
// the node along the face normal (JxW_f / JxW_i = w_tangential / w_i). Nothing is
// stored per node: J^-1 is rebuilt from the cell's vertices in FP64, and the
// derivative along each axis is the 1-D GLL differentiation matrix applied in LDS.
//
// RK5 (timeintegration/RungeKutta5LowStorage.cpp:56-105, coefficients
// RungeKutta5LowStorage.h:29-46), per stage s and entry:
//   af = dt * (L f); d = (s == 0) ? af : a_s df + af; d *= b_s; f += d; df = d * (1 / b_s)
// (the vectors' operator/= multiplies by the reciprocal: deal.II Vector and
// TrilinosWrappers::MPI::Vector)
// Each stage reads f from one buffer and writes the other (neighbour values at
// faces come from other blocks). The last stage of a step can collide in
// registers (CFDSolver::run order: stream, then collide).
//
// Walls (velocity-neq bounce-back on the faces of a non-periodic axis, assembled by
// VelocityNeqBounceBack::assembleBoundary, VelocityNeqBounceBack.cpp:46-135): the other
// side of a wall face is the opposite direction at the same node, and the wall velocity
// enters as the system vector b of the RK5 right-hand side, with rho = 1:
//   upwind (s_a G_a < 0): s_a G_a / w_a(i) * (f_i - f_opp,i) - 2 c w (e.u_wall) / cs2
//   central:  1/2 c (f_i + f_opp,i) - c w (e.u_wall) / cs2,   c = s_a G_a / w_a(i).
//
// FP32 storage keeps f~ = f - w rho0 in FP32. L annihilates constants for the
// upwind flux; for the central flux L(w rho0) = w rho0 sum_a s_a G_a / w_a(i) is
// added back explicitly. All arithmetic is FP64.
#include "nat_collide.hpp"
#include "nat_geom.hpp"

namespace nat {

// Carpenter-Kennedy 2N-storage RK5 coefficients (RungeKutta5LowStorage.h:29-46)
static const double kRkA[5] = {0.0, -567301805773.0 / 1357537059087.0, -2404267990393.0 / 2016746695238.0,
                               -3550918686646.0 / 2091501179385.0, -1275806237668.0 / 842570457699.0};
static const double kRkB[5] = {1432997174477.0 / 9575080441755.0, 5161836677717.0 / 13612068292357.0,
                               1720146321549.0 / 2090206949498.0, 3134564353537.0 / 4481467310338.0,
                               2277821191437.0 / 14882151754819.0};

template <typename T>
struct SedgArgs {
  const T* fin;
  T* fout;
  T* df;
  int64_t stride;
  int64_t ncell;
  MeshDev m;
  double scaling;
  double dt;
  double a, b;       // RK5 coefficients of this stage
  double binv;       // 1 / b (the division of df /= b is a multiplication by 1/b)
  double rho0;       // FP32 shift (0 for FP64)
  int first;         // first stage: df not read
  int store_df;      // 0 on the last stage (df is dead afterwards)
  int central;       // central flux
  double gll[MAXP + 1];
  double inv_gw[MAXP + 1];           // 1 / GLL weight
  double dl[MAXP + 1][MAXP + 1];     // dl[i][j] = phi_j'(x_i)
};

template <int D, int P>
struct SedgShape {
  static constexpr int NP = P + 1;
  static constexpr int NL = (D == 3) ? NP * NP * NP : NP * NP;
  // cells per block: ~256 threads; a 64-node cell (3D, p = 3) runs alone as one wave, whose
  // per-direction barrier then waits for no other wave (C3 SEDG +4% FP32, +1% FP64)
  static constexpr int CPB = (NL == 64) ? 1 : ((256 / NL) > 0 ? 256 / NL : 1);
  static constexpr int NT = CPB * NL;
};

__device__ __forceinline__ void sedg_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// WALLS: the mesh has a non-periodic axis (its own instantiation: the wall terms cost
// registers that the periodic kernels keep for occupancy)
template <int D, int Q, int P, typename T, bool COL, bool WALLS>
__global__ __launch_bounds__((SedgShape<D, P>::NT)) void k_sedg_stage(SedgArgs<T> s, CollideArgs<T> ca,
                                                                       int write_macro) {
  using L = Lattice<Q>;
  using S = SedgShape<D, P>;
  constexpr int NP = S::NP, NL = S::NL, NT = S::NT;
  __shared__ double buf[2][NT];
  __shared__ double dl[NP][NP];
  const int tid = threadIdx.x;
  for (int i = tid; i < NP * NP; i += NT) dl[i / NP][i % NP] = s.dl[i / NP][i % NP];

  const int cl = tid / NL, l = tid % NL;
  const int64_t cell = (int64_t)blockIdx.x * S::CPB + cl;
  const bool active = cell < s.ncell;
  const int64_t node = active ? cell * NL + l : 0;
  const int mi[3] = {l % NP, (l / NP) % NP, (D == 3) ? l / (NP * NP) : 0};
  const int lstr[3] = {1, NP, NP * NP};

  // geometry of this node
  double Ji[3][3];
  int64_t nbnode[3] = {0, 0, 0};
  double fac[3] = {0, 0, 0};  // s_a / w_a(i) for the face through this node along axis a, else 0
  int wallf[3] = {-1, -1, -1};  // box face 2 a + side when that face is a wall
  {
    const int N0 = s.m.N[0], N1 = s.m.N[1];
    const int c[3] = {(int)(cell % N0), (int)((cell / N0) % N1), (D == 3) ? (int)(cell / ((int64_t)N0 * N1)) : 0};
    const double xi[3] = {s.gll[mi[0]], s.gll[mi[1]], (D == 3) ? s.gll[mi[2]] : 0.0};
    double J[3][3];
    if (active) {
      cell_jac<D>(s.m, c, xi, J);
    } else {
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) J[a][b] = (a == b) ? 1.0 : 0.0;
    }
    inv_det(D, J, Ji);
#pragma unroll
    for (int a = 0; a < D; a++) {
      if (mi[a] == 0 || mi[a] == P) {
        const int up = (mi[a] == P) ? 1 : 0;
        int cn[3] = {c[0], c[1], c[2]};
        const int n = s.m.N[a];
        const bool edge = up ? (c[a] + 1 == n) : (c[a] == 0);
        if (WALLS && edge && !s.m.periodic[a]) {
          wallf[a] = 2 * a + up;
          nbnode[a] = node;  // read in the opposite direction
        } else {
          cn[a] = up ? (c[a] + 1 == n ? 0 : c[a] + 1) : (c[a] == 0 ? n - 1 : c[a] - 1);
          const int64_t nc = cn[0] + (int64_t)N0 * (cn[1] + (int64_t)N1 * cn[2]);
          nbnode[a] = active ? nc * NL + l + (up ? -P : P) * lstr[a] : 0;
        }
        fac[a] = (up ? 1.0 : -1.0) * s.inv_gw[mi[a]];
      }
    }
  }
  __syncthreads();

  const double dt = s.dt, ra = s.a, rb = s.b, sc = s.scaling;
  T fs[COL ? Q : 1];
  fs[0] = active ? s.fin[node] : T(0);
  if (!COL && active) s.fout[node] = fs[0];

  T vnext = active ? s.fin[1 * s.stride + node] : T(0);
  static_for<1, Q>([&](auto A) {
    // The stage arithmetic states its fused multiply-adds (fma) and forbids the compiler
    // any other contraction: a contraction left to the compiler's choice can differ
    // between the stream-only and the collide-after instantiations of this kernel, and
    // the two must give the same populations and fields bit for bit.
#pragma clang fp contract(off)
    constexpr int al = decltype(A)::value;
    constexpr int e0 = L::e[al][0], e1 = L::e[al][1], e2 = (D == 3) ? L::e[al][2] : 0;
    // per-direction addresses are formed here, not hoisted into long-lived registers
    const int64_t dofs = (int64_t)al * s.stride;
    int64_t nd = node;
    asm volatile("" : "+v"(nd));
    const T* __restrict__ fin = s.fin + dofs;
    const double v = (double)vnext;
    buf[al & 1][tid] = v;
    T dfo = T(0);
    if (!s.first && active) dfo = s.df[dofs + nd];
    // G_b = sum_k e_k (J^-1)_bk, scaled velocities
    double G[3];
#pragma unroll
    for (int b = 0; b < D; b++) {
      double g = 0;
      if (e0 != 0) g += e0 > 0 ? Ji[b][0] : -Ji[b][0];
      if (e1 != 0) g += e1 > 0 ? Ji[b][1] : -Ji[b][1];
      if (e2 != 0) g += e2 > 0 ? Ji[b][2] : -Ji[b][2];
      G[b] = g * sc;
    }
    // Face neighbours this direction reads. With FP64 storage they are loaded before the
    // barrier, so that their latency overlaps the LDS exchange and the volume term
    // (C3 SEDG FP64 0.88 -> 1.14 GLUPS). With FP32 storage they are loaded right after
    // the barrier, so that the volume term covers their latency: the stage then holds
    // 110 registers (4 waves per SIMD) and C3 FP32 takes 621 us per step, against 707
    // with the loads after the volume term (131 registers, 3 waves) and 758 before the
    // barrier.
    constexpr bool early = sizeof(T) == 8;
    T fnb[3] = {T(0), T(0), T(0)};
    constexpr int op = opposite<Q>(al);
    const T* __restrict__ fop = s.fin + (int64_t)op * s.stride;
    auto load_faces = [&]() {
#pragma unroll
      for (int a = 0; a < D; a++) {
        const double c = fac[a] * G[a];
        if (active && (s.central ? fac[a] != 0.0 : c < 0.0)) fnb[a] = (!WALLS || wallf[a] < 0 ? fin : fop)[nbnode[a]];
      }
    };
    if constexpr (early) load_faces();
    // the next direction's own value last: loads complete in order, so this one stays
    // in flight while the face values above are waited for
    if constexpr (al + 1 < Q) vnext = active ? s.fin[(al + 1) * s.stride + node] : T(0);
    sedg_barrier();
    if constexpr (!early) load_faces();
    double lf = 0;
#pragma unroll
    for (int b = 0; b < D; b++) {
      const int base = tid - mi[b] * lstr[b];
      double der = 0;
#pragma unroll
      for (int j = 0; j < NP; j++) der = fma(dl[mi[b]][j], buf[al & 1][base + j * lstr[b]], der);
      lf = fma(-G[b], der, lf);
    }
#pragma unroll
    for (int a = 0; a < D; a++) {
      const double c = fac[a] * G[a];
      // wall velocity term: c w (e.u_wall) / cs2 with e and cs2 scaled, cs2 = s^2 / 3
      double cwu = 0.0;
      if (WALLS && wallf[a] >= 0) {
        const double* uw = s.m.wall_u[wallf[a]];
        double eu = 0.0;
        if (e0 != 0) eu += e0 * uw[0];
        if (e1 != 0) eu += e1 * uw[1];
        if (e2 != 0) eu += e2 * uw[2];
        cwu = c * L::w[al] * (3.0 * eu / sc);
      }
      if (s.central) {
        if (fac[a] != 0.0) {
          const double fn = (double)fnb[a];
          lf += 0.5 * c * (v + fn) + c * L::w[al] * s.rho0;
          if constexpr (WALLS) lf -= cwu;
        }
      } else if (c < 0.0) {
        const double fn = (double)fnb[a];
        lf = fma(c, v - fn, lf);
        if constexpr (WALLS) lf -= 2.0 * cwu;
      }
    }
    const double af = lf * dt;
    double d = s.first ? af : fma((double)dfo, ra, af);
    d *= rb;
    const double fnew = v + d;
    d *= s.binv;
    if (active) {
      if (s.store_df) s.df[dofs + nd] = (T)d;
      if constexpr (COL)
        fs[al] = (T)fnew;
      else
        s.fout[dofs + nd] = (T)fnew;
    }
  });
  if constexpr (COL) {
    if (active) bgk_collide_store_pairs<Q, T>(fs, ca, node, write_macro != 0, s.fout, s.stride);
  }
}

template <int D, int Q, int P, typename T>
static int launch_stage(nat_ctx* ctx, const SedgArgs<T>& s, const CollideArgs<T>& ca, int collide, int wm) {
  using S = SedgShape<D, P>;
  const unsigned grid = (unsigned)((s.ncell + S::CPB - 1) / S::CPB);
  if (ctx->has_walls) {
    if (collide)
      hipLaunchKernelGGL((k_sedg_stage<D, Q, P, T, true, true>), dim3(grid), dim3(S::NT), 0, ctx->stream, s, ca, wm);
    else
      hipLaunchKernelGGL((k_sedg_stage<D, Q, P, T, false, true>), dim3(grid), dim3(S::NT), 0, ctx->stream, s, ca, 0);
  } else {
    if (collide)
      hipLaunchKernelGGL((k_sedg_stage<D, Q, P, T, true, false>), dim3(grid), dim3(S::NT), 0, ctx->stream, s, ca, wm);
    else
      hipLaunchKernelGGL((k_sedg_stage<D, Q, P, T, false, false>), dim3(grid), dim3(S::NT), 0, ctx->stream, s, ca, 0);
  }
  NAT_HIP(hipGetLastError());
  return 0;
}

template <typename T>
static int dispatch_stage(nat_ctx* ctx, const SedgArgs<T>& s, const CollideArgs<T>& ca, int collide, int wm) {
  const int Q = ctx->Q, P = ctx->p;
#define NAT_SE(DD, QQ, PP) \
  if (Q == QQ && P == PP) return launch_stage<DD, QQ, PP, T>(ctx, s, ca, collide, wm);
  NAT_SE(2, 9, 1)
  NAT_SE(2, 9, 2)
  NAT_SE(2, 9, 3)
  NAT_SE(2, 9, 4)
  NAT_SE(2, 9, 5)
  NAT_SE(2, 9, 6)
  NAT_SE(3, 19, 1)
  NAT_SE(3, 19, 2)
  NAT_SE(3, 19, 3)
  NAT_SE(3, 19, 4)
  NAT_SE(3, 27, 1)
  NAT_SE(3, 27, 2)
  NAT_SE(3, 27, 3)
  NAT_SE(3, 27, 4)
#undef NAT_SE
  ctx->err = "SEDG kernel not instantiated for this (stencil, order)";
  return NAT_ERR_UNSUPPORTED;
}

// 1-D GLL differentiation matrix dl[i][j] = phi_j'(x_i) (product rule on the
// Lagrange product form, as FE_Q evaluates its shape gradients)
static void diff_matrix(int p, const double* x, double dl[MAXP + 1][MAXP + 1]) {
  for (int i = 0; i <= MAXP; i++)
    for (int j = 0; j <= MAXP; j++) dl[i][j] = 0.0;
  for (int j = 0; j <= p; j++) {
    double den = 1;
    for (int m = 0; m <= p; m++)
      if (m != j) den *= (x[j] - x[m]);
    for (int i = 0; i <= p; i++) {
      double sum = 0;
      for (int k = 0; k <= p; k++) {
        if (k == j) continue;
        double t = 1;
        for (int m = 0; m <= p; m++)
          if (m != j && m != k) t *= (x[i] - x[m]);
        sum += t;
      }
      dl[i][j] = sum / den;
    }
  }
}

template <typename T>
static int sedg_step_typed(nat_ctx* ctx, int collide_after, int write_macro) {
  SedgArgs<T> s;
  s.df = (T*)ctx->sedg_df;
  s.stride = ctx->stride;
  s.m.D = ctx->D;
  for (int a = 0; a < 3; a++) {
    s.m.N[a] = a < ctx->D ? ctx->desc.ncells[a] : 1;
    s.m.periodic[a] = a < ctx->D ? (ctx->desc.periodic[a] ? 1 : 0) : 1;
  }
  for (int f = 0; f < 6; f++)
    for (int d = 0; d < 3; d++) s.m.wall_u[f][d] = d < ctx->D ? ctx->desc.wall_velocity[f][d] : 0.0;
  s.m.cartesian = 0;
  s.m.vert = ctx->d_vert;
  s.ncell = (int64_t)s.m.N[0] * s.m.N[1] * s.m.N[2];
  s.scaling = ctx->st.s;
  s.dt = ctx->dt;
  s.rho0 = ctx->fp32 ? (ctx->desc.rho0 > 0 ? ctx->desc.rho0 : 1.0) : 0.0;
  s.central = ctx->desc.scheme == NAT_SEDG_CENTRAL ? 1 : 0;
  double w[MAXP + 1];
  gll_points(ctx->p, s.gll, w);
  for (int i = 0; i <= MAXP; i++) s.inv_gw[i] = i <= ctx->p ? 1.0 / w[i] : 0.0;
  diff_matrix(ctx->p, s.gll, s.dl);
  CollideArgs<T> ca;
  ca.omega = (T)(1.0 / ctx->tau);
  ca.omega64 = 1.0 / ctx->tau;
  ca.rho0 = (T)s.rho0;
  ca.scaling = ctx->st.s;
  ca.rho = (T*)ctx->rho;
  ca.u = (T*)ctx->u;
  ca.mstride = ctx->n_store;
  ca.status = ctx->d_status;
  ca.stab = ctx->stab_on();
  for (int st = 0; st < 5; st++) {
    s.fin = (const T*)ctx->f[ctx->cur];
    s.fout = (T*)ctx->f[1 - ctx->cur];
    s.a = kRkA[st];
    s.b = kRkB[st];
    s.binv = 1.0 / kRkB[st];
    s.first = st == 0 ? 1 : 0;
    s.store_df = st < 4 ? 1 : 0;
    const int col = (st == 4) ? collide_after : 0;
    const int rc = dispatch_stage<T>(ctx, s, ca, col, col ? write_macro : 0);
    if (rc) return rc;
    ctx->cur = 1 - ctx->cur;
  }
  return 0;
}

}  // namespace nat

int nat_sedg_setup(nat_ctx* ctx) {
  using namespace nat;
  if (ctx->D == 3 && ctx->p > 4) {
    ctx->err = "SEDG kernels are instantiated for p <= 4 in 3D";
    return NAT_ERR_UNSUPPORTED;
  }
  if (ctx->D == 2 && ctx->p > 6) {
    ctx->err = "SEDG kernels are instantiated for p <= 6 in 2D";
    return NAT_ERR_UNSUPPORTED;
  }
  if (!ctx->sedg_df) {
    const size_t es = ctx->fp32 ? sizeof(float) : sizeof(double);
    NAT_HIP(hipMalloc(&ctx->sedg_df, es * ctx->stride * ctx->Q));
    NAT_HIP(hipMemset(ctx->sedg_df, 0, es * ctx->stride * ctx->Q));
  }
  return 0;
}

int nat_sedg_step(nat_ctx* ctx, int collide_after, int write_macro) {
  return ctx->fp32 ? nat::sedg_step_typed<float>(ctx, collide_after, write_macro)
                   : nat::sedg_step_typed<double>(ctx, collide_after, write_macro);
}
