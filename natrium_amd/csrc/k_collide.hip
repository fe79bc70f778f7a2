// Collision-only pass (CFDSolver::collide, CFDSolver.cpp:807-843) and FP64
// conservation sums (PhysicalProperties::mass / kineticEnergy,
// PhysicalProperties.cpp:29-130 and 270-310).
#include "nat_collide.hpp"
#include "nat_geom.hpp"

namespace nat {

template <int Q, typename T>
__global__ __launch_bounds__(256) void k_collide(T* f, int64_t stride, int64_t n, CollideArgs<T> ca) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T fs[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) fs[q] = f[(int64_t)q * stride + i];
  bgk_collide<Q, T>(fs, ca, i, true);
#pragma unroll
  for (int q = 0; q < Q; q++) f[(int64_t)q * stride + i] = fs[q];
}

// Layouts of the macroscopic fields for the quadrature sums.
enum { LAYOUT_FEQ = 0, LAYOUT_UNIQUE = 1, LAYOUT_DG = 2 };

struct SumArgs {
  MeshDev m;
  int p;
  int layout;
  int nn[3];        // FE_Q lattice nodes (LAYOUT_FEQ) or unique nodes (LAYOUT_UNIQUE) per axis
  int64_t mstride;  // stride of u components
  double gll[MAXP + 1];
  double gllw[MAXP + 1];
};

// one thread per (cell, local node); block partial sums -> FP64 atomics
template <typename T>
__global__ __launch_bounds__(256) void k_sums(SumArgs s, const T* __restrict__ rho, const T* __restrict__ u,
                                              double* __restrict__ out /* mass, mom[3], 2*ekin */) {
  const int np = s.p + 1;
  const int nl = s.m.D == 3 ? np * np * np : np * np;
  const int64_t ncell = (int64_t)s.m.N[0] * s.m.N[1] * (s.m.D == 3 ? s.m.N[2] : 1);
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[5] = {0, 0, 0, 0, 0};
  if (t < ncell * nl) {
    const int64_t cell = t / nl;
    const int l = (int)(t - cell * nl);
    int c[3] = {(int)(cell % s.m.N[0]), (int)((cell / s.m.N[0]) % s.m.N[1]),
                s.m.D == 3 ? (int)(cell / ((int64_t)s.m.N[0] * s.m.N[1])) : 0};
    int mi[3] = {l % np, (l / np) % np, s.m.D == 3 ? l / (np * np) : 0};
    double xi[3] = {s.gll[mi[0]], s.gll[mi[1]], s.m.D == 3 ? s.gll[mi[2]] : 0.0};
    double J[3][3];
    if (s.m.D == 3)
      cell_jac<3>(s.m, c, xi, J);
    else
      cell_jac<2>(s.m, c, xi, J);
    double det = s.m.D == 2 ? J[0][0] * J[1][1] - J[0][1] * J[1][0]
                            : J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                                  J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                                  J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    const double jxw = det * s.gllw[mi[0]] * s.gllw[mi[1]] * (s.m.D == 3 ? s.gllw[mi[2]] : 1.0);
    int64_t i;
    if (s.layout == LAYOUT_DG) {
      i = t;
    } else {
      int g[3] = {c[0] * s.p + mi[0], c[1] * s.p + mi[1], s.m.D == 3 ? c[2] * s.p + mi[2] : 0};
      if (s.layout == LAYOUT_UNIQUE)
        for (int a = 0; a < s.m.D; a++)
          if (g[a] >= s.nn[a]) g[a] -= s.nn[a];
      i = g[0] + (int64_t)s.nn[0] * (g[1] + (int64_t)s.nn[1] * g[2]);
    }
    const double r = (double)rho[i];
    double uu = 0;
    for (int d = 0; d < s.m.D; d++) {
      const double ud = (double)u[d * s.mstride + i];
      v[1 + d] = r * ud * jxw;
      uu += ud * ud;
    }
    v[0] = r * jxw;
    v[4] = r * uu * jxw;
  }
  __shared__ double red[5][256];
  for (int k = 0; k < 5; k++) red[k][threadIdx.x] = v[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 5; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 5; k++) atomicAdd(out + k, red[k][0]);
}

}  // namespace nat

int nat_launch_collide(nat_ctx* ctx) {
  using namespace nat;
  const int bs = 256;
  const unsigned grid = (unsigned)((ctx->n_store + bs - 1) / bs);
  if (ctx->fp32) {
    CollideArgs<float> ca;
    ca.omega = (float)(1.0 / ctx->tau);
    ca.omega64 = 1.0 / ctx->tau;
    ca.rho0 = (float)(ctx->desc.rho0 > 0 ? ctx->desc.rho0 : 1.0);
    ca.scaling = ctx->st.s;
    ca.rho = (float*)ctx->rho;
    ca.u = (float*)ctx->u;
    ca.mstride = ctx->n_store;
    ca.status = ctx->d_status;
    float* f = (float*)ctx->f[ctx->cur];
    if (ctx->Q == 9) hipLaunchKernelGGL((k_collide<9, float>), dim3(grid), dim3(bs), 0, ctx->stream, f, ctx->stride, ctx->n_store, ca);
    else if (ctx->Q == 19) hipLaunchKernelGGL((k_collide<19, float>), dim3(grid), dim3(bs), 0, ctx->stream, f, ctx->stride, ctx->n_store, ca);
    else if (ctx->Q == 27) hipLaunchKernelGGL((k_collide<27, float>), dim3(grid), dim3(bs), 0, ctx->stream, f, ctx->stride, ctx->n_store, ca);
    else { ctx->err = "collision not available for this stencil"; return NAT_ERR_UNSUPPORTED; }
  } else {
    CollideArgs<double> ca;
    ca.omega = 1.0 / ctx->tau;
    ca.omega64 = 1.0 / ctx->tau;
    ca.rho0 = 0.0;
    ca.scaling = ctx->st.s;
    ca.rho = (double*)ctx->rho;
    ca.u = (double*)ctx->u;
    ca.mstride = ctx->n_store;
    ca.status = ctx->d_status;
    double* f = (double*)ctx->f[ctx->cur];
    if (ctx->Q == 9) hipLaunchKernelGGL((k_collide<9, double>), dim3(grid), dim3(bs), 0, ctx->stream, f, ctx->stride, ctx->n_store, ca);
    else if (ctx->Q == 19) hipLaunchKernelGGL((k_collide<19, double>), dim3(grid), dim3(bs), 0, ctx->stream, f, ctx->stride, ctx->n_store, ca);
    else if (ctx->Q == 27) hipLaunchKernelGGL((k_collide<27, double>), dim3(grid), dim3(bs), 0, ctx->stream, f, ctx->stride, ctx->n_store, ca);
    else { ctx->err = "collision not available for this stencil"; return NAT_ERR_UNSUPPORTED; }
  }
  NAT_HIP(hipGetLastError());
  return 0;
}

int nat_launch_sums(nat_ctx* ctx, int layout) {
  using namespace nat;
  SumArgs s;
  s.m.D = ctx->D;
  for (int a = 0; a < 3; a++) {
    s.m.N[a] = a < ctx->D ? ctx->desc.ncells[a] : 1;
    s.m.periodic[a] = 1;
    s.nn[a] = layout == LAYOUT_UNIQUE ? ctx->nu[a] : ctx->nfe[a];
  }
  s.m.cartesian = (layout == LAYOUT_DG) ? 0 : ctx->desc.cartesian;
  s.m.vert = ctx->d_vert;
  s.p = ctx->p;
  s.layout = layout;
  s.mstride = ctx->n_store;
  gll_points(ctx->p, s.gll, s.gllw);
  const int np = ctx->p + 1;
  const int64_t total = (int64_t)s.m.N[0] * s.m.N[1] * s.m.N[2] * (ctx->D == 3 ? np * np * np : np * np);
  NAT_HIP(hipMemsetAsync(ctx->d_sums, 0, 5 * sizeof(double), ctx->stream));
  const int bs = 256;
  const unsigned grid = (unsigned)((total + bs - 1) / bs);
  if (ctx->fp32)
    hipLaunchKernelGGL((k_sums<float>), dim3(grid), dim3(bs), 0, ctx->stream, s, (const float*)ctx->rho,
                       (const float*)ctx->u, ctx->d_sums);
  else
    hipLaunchKernelGGL((k_sums<double>), dim3(grid), dim3(bs), 0, ctx->stream, s, (const double*)ctx->rho,
                       (const double*)ctx->u, ctx->d_sums);
  NAT_HIP(hipGetLastError());
  return 0;
}
