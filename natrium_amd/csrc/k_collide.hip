// Collision-only pass (CFDSolver::collide, CFDSolver.cpp:807-843) and FP64
// conservation sums (PhysicalProperties::mass / kineticEnergy,
// PhysicalProperties.cpp:29-130 and 270-310).
#include <mutex>
#include <string>

#include "nat_collide.hpp"
#include "nat_fastdiv.hpp"
#include "nat_geom.hpp"

namespace nat {

// One thread per node; the same FP64 pair-form collision as the fused step kernels,
// so stream-only followed by this pass equals a fused step bit for bit. INIT: the
// reference's initialisation collision (velocity read from the stored field).
template <int Q, typename T, bool INIT>
__global__ __launch_bounds__(256) void k_collide(T* f, int64_t stride, int64_t n, CollideArgs<T> ca) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  T fs[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) fs[q] = f[(int64_t)q * stride + i];
  bgk_collide_store_pairs<Q, T, INIT>(fs, ca, i, true, f, stride);
}

// The pseudo-entropic stabilizer as a pass of its own (PseudoEntropicStabilizer::apply,
// dataprocessors/PseudoEntropicStabilizer.cpp:271-311): every node's populations go
// through the w-orthogonal projection P onto w times the Hermite polynomials of order
// <= 2, (P f)_i = w_i (rho + e_i.j / cs2 + H2_i : (Pi - rho cs2 I) / (2 cs2^2)). P keeps
// w rho0, so the shifted FP32 populations are projected as they are. D2Q9 / D3Q19 only.
template <int Q, typename T>
__global__ __launch_bounds__(256) void k_stabilize(T* f, int64_t stride, int64_t n) {
  using L = Lattice<Q>;
  constexpr int D = L::D;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double fd[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) fd[q] = (double)f[(int64_t)q * stride + i];
  double m0 = 0, j[3] = {0, 0, 0};
#pragma unroll
  for (int q = 0; q < Q; q++) {
    m0 += fd[q];
#pragma unroll
    for (int d = 0; d < D; d++) {
      if (L::e[q][d] > 0) j[d] += fd[q];
      if (L::e[q][d] < 0) j[d] -= fd[q];
    }
  }
  // second-order coefficients 9/2 (Pi - m0 I / 3), as the collision's non-equilibrium
  // projection at u = 0 and omega = 0
  const double zero[3] = {0, 0, 0};
  const StabPi<Q> m = stab_neq<Q>(stab_moments<Q>(fd), m0, 0.0, zero, 0.0);
#pragma unroll
  for (int q = 0; q < Q; q++) {
    double ej = 0;
#pragma unroll
    for (int d = 0; d < D; d++) {
      if (L::e[q][d] > 0) ej += j[d];
      if (L::e[q][d] < 0) ej -= j[d];
    }
    f[(int64_t)q * stride + i] = (T)(L::w[q] * (m0 + 3.0 * ej) + stab_term<Q>(m, q));
  }
}

// Layouts of the macroscopic fields for the quadrature sums.
enum { LAYOUT_FEQ = 0, LAYOUT_UNIQUE = 1, LAYOUT_DG = 2 };

struct SumArgs {
  MeshDev m;
  int p;
  int layout;
  FastDivU div_nl, div_n0, div_n1;  // (cell, node) index -> cell, node; cell -> coordinates
  int nn[3];        // FE_Q lattice nodes (LAYOUT_FEQ) or unique nodes (LAYOUT_UNIQUE) per axis
  int64_t mstride;  // stride of u components
  double gll[MAXP + 1];
  double gllw[MAXP + 1];
};

// Sum of a double over the 64 lanes of a wave by a fixed butterfly (every lane ends
// with the same total, formed in the same order on every run).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

constexpr int SUM_BS = 256;  // threads per block of both passes

// Pass 1: one thread per (cell, local node); each block writes its five partial sums
// (mass, momentum, 2 E_kin) to partial[5 * block]. Wave sums by shuffles, then the
// block's waves added in wave order: no atomics, so the result does not depend on the
// order in which blocks run.
template <typename T>
__global__ __launch_bounds__(SUM_BS) void k_sums(SumArgs s, const T* __restrict__ rho, const T* __restrict__ u,
                                                 double* __restrict__ partial) {
  const int np = s.p + 1;
  const int nl = s.m.D == 3 ? np * np * np : np * np;
  const int64_t ncell = (int64_t)s.m.N[0] * s.m.N[1] * (s.m.D == 3 ? s.m.N[2] : 1);
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double v[5] = {0, 0, 0, 0, 0};
  if (t < ncell * nl) {
    // 32-bit index arithmetic by multiply-high divisions (the host keeps ncell * nl < 2^31):
    // 64-bit divisions by run-time values cost more than the sum itself
    const uint32_t tt = (uint32_t)t;
    const uint32_t cell = s.div_nl.div(tt);
    const int l = (int)(tt - cell * (uint32_t)nl);
    const uint32_t q0 = s.div_n0.div(cell), q1 = s.div_n1.div(q0);
    int c[3] = {(int)(cell - q0 * (uint32_t)s.m.N[0]), (int)(q0 - q1 * (uint32_t)s.m.N[1]),
                s.m.D == 3 ? (int)q1 : 0};
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
  __shared__ double red[SUM_BS / 64][5];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const double w = wave_sum(v[k]);
    if (lane == 0) red[wv][k] = w;
  }
  __syncthreads();
  if (threadIdx.x < 5) {
    double a = red[0][threadIdx.x];
#pragma unroll
    for (int w = 1; w < SUM_BS / 64; w++) a += red[w][threadIdx.x];
    partial[(int64_t)blockIdx.x * 5 + threadIdx.x] = a;
  }
}

// Pass 2: one block adds the nb partial sums: thread t takes blocks t, t + 256, ... in
// order, then the same wave and block reduction as pass 1.
__global__ __launch_bounds__(SUM_BS) void k_sums_final(const double* __restrict__ partial, int64_t nb,
                                                       double* __restrict__ out) {
  double v[5] = {0, 0, 0, 0, 0};
  for (int64_t b = threadIdx.x; b < nb; b += SUM_BS)
#pragma unroll
    for (int k = 0; k < 5; k++) v[k] += partial[b * 5 + k];
  __shared__ double red[SUM_BS / 64][5];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 5; k++) {
    const double w = wave_sum(v[k]);
    if (lane == 0) red[wv][k] = w;
  }
  __syncthreads();
  if (threadIdx.x < 5) {
    double a = red[0][threadIdx.x];
#pragma unroll
    for (int w = 1; w < SUM_BS / 64; w++) a += red[w][threadIdx.x];
    out[threadIdx.x] = a;
  }
}

}  // namespace nat

template <typename T>
static int launch_collide_typed(nat_ctx* ctx, bool init) {
  using namespace nat;
  const int bs = 256;
  const unsigned grid = (unsigned)((ctx->n_store + bs - 1) / bs);
  CollideArgs<T> ca;
  ca.omega = (T)(1.0 / ctx->tau);
  ca.omega64 = 1.0 / ctx->tau;
  ca.rho0 = (T)(ctx->fp32 ? (ctx->desc.rho0 > 0 ? ctx->desc.rho0 : 1.0) : 0.0);
  ca.scaling = ctx->st.s;
  ca.rho = (T*)ctx->rho;
  ca.u = (T*)ctx->u;
  ca.mstride = ctx->n_store;
  ca.status = ctx->d_status;
  ca.stab = 0;
  T* f = (T*)ctx->f[ctx->cur];
#define NAT_KC(QQ)                                                                                              \
  if (ctx->Q == QQ) {                                                                                          \
    if (init)                                                                                                  \
      hipLaunchKernelGGL((k_collide<QQ, T, true>), dim3(grid), dim3(bs), 0, ctx->stream, f, ctx->stride,       \
                         ctx->n_store, ca);                                                                    \
    else                                                                                                       \
      hipLaunchKernelGGL((k_collide<QQ, T, false>), dim3(grid), dim3(bs), 0, ctx->stream, f, ctx->stride,      \
                         ctx->n_store, ca);                                                                    \
    NAT_HIP(hipGetLastError());                                                                                \
    return 0;                                                                                                  \
  }
  NAT_KC(9)
  NAT_KC(19)
  NAT_KC(27)
#undef NAT_KC
  ctx->err = "collision not available for this stencil";
  return NAT_ERR_UNSUPPORTED;
}

int nat_launch_collide(nat_ctx* ctx, int init) {
  return ctx->fp32 ? launch_collide_typed<float>(ctx, init != 0) : launch_collide_typed<double>(ctx, init != 0);
}

int nat_launch_stabilize(nat_ctx* ctx) {
  using namespace nat;
  if (ctx->Q != 9 && ctx->Q != 19) return 0;  // the reference leaves other stencils alone
  const int bs = 256;
  const unsigned grid = (unsigned)((ctx->n_store + bs - 1) / bs);
  void* f = ctx->f[ctx->cur];
  if (ctx->fp32) {
    if (ctx->Q == 9)
      hipLaunchKernelGGL((k_stabilize<9, float>), dim3(grid), dim3(bs), 0, ctx->stream, (float*)f, ctx->stride, ctx->n_store);
    else
      hipLaunchKernelGGL((k_stabilize<19, float>), dim3(grid), dim3(bs), 0, ctx->stream, (float*)f, ctx->stride, ctx->n_store);
  } else {
    if (ctx->Q == 9)
      hipLaunchKernelGGL((k_stabilize<9, double>), dim3(grid), dim3(bs), 0, ctx->stream, (double*)f, ctx->stride, ctx->n_store);
    else
      hipLaunchKernelGGL((k_stabilize<19, double>), dim3(grid), dim3(bs), 0, ctx->stream, (double*)f, ctx->stride, ctx->n_store);
  }
  NAT_HIP(hipGetLastError());
  return 0;
}

// Quadrature-sum arguments of a context and the number of pass-1 blocks they need.
static nat::SumArgs sum_args(const nat_ctx* ctx, int layout, int64_t* grid) {
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
  const int nl = ctx->D == 3 ? np * np * np : np * np;
  const int64_t total = (int64_t)s.m.N[0] * s.m.N[1] * s.m.N[2] * nl;
  s.div_nl.set((uint32_t)nl);
  s.div_n0.set((uint32_t)s.m.N[0]);
  s.div_n1.set((uint32_t)s.m.N[1]);
  *grid = (total + SUM_BS - 1) / SUM_BS;
  return s;
}

// Both passes on the context's stream; partial holds 5 doubles per pass-1 block.
static int launch_sums_into(nat_ctx* ctx, const nat::SumArgs& s, int64_t grid, double* partial) {
  using namespace nat;
  if (grid * SUM_BS > INT32_MAX) {
    ctx->err = "conservation sums: more than 2^31 (cell, node) pairs";
    return NAT_ERR_UNSUPPORTED;
  }
  if (ctx->fp32)
    hipLaunchKernelGGL((k_sums<float>), dim3((unsigned)grid), dim3(SUM_BS), 0, ctx->stream, s,
                       (const float*)ctx->rho, (const float*)ctx->u, partial);
  else
    hipLaunchKernelGGL((k_sums<double>), dim3((unsigned)grid), dim3(SUM_BS), 0, ctx->stream, s,
                       (const double*)ctx->rho, (const double*)ctx->u, partial);
  NAT_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_sums_final, dim3(1), dim3(SUM_BS), 0, ctx->stream, (const double*)partial, grid,
                     ctx->d_sums);
  NAT_HIP(hipGetLastError());
  return 0;
}

int nat_launch_sums(nat_ctx* ctx, int layout) {
  int64_t grid = 0;
  const nat::SumArgs s = sum_args(ctx, layout, &grid);
  if (grid > ctx->n_partials) {
    if (ctx->d_partials) hipFree(ctx->d_partials);
    ctx->d_partials = nullptr;
    ctx->n_partials = 0;
    NAT_HIP(hipMalloc((void**)&ctx->d_partials, (size_t)grid * 5 * sizeof(double)));
    ctx->n_partials = grid;
  }
  return launch_sums_into(ctx, s, grid, ctx->d_partials);
}

// ---------------------------------------------------------------------------
// Sums without allocation: the pass-1 partial sums go to a fixed device array instead of
// the context's lazily allocated buffer. Each GPU holds its own copy of the array at its
// own address, looked up once per device with that device current (nat_sums_scratch_setup,
// called from nat_create); host code uses only the looked-up address, never the array's
// name. Every pass-1 block overwrites its own 5 entries, so nothing has to start at zero;
// calls that overlap on one GPU (two streams at once) share the array and race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[NAT_SUMS_SCRATCH_BLOCKS * 5];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int nat_sums_scratch_setup(nat_ctx* ctx) {
  int dev = 0;
  NAT_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    ctx->err = "sums scratch: device index beyond the supported range";
    return NAT_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    NAT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = (double*)p;
  }
  ctx->device = dev;
  return 0;
}

int nat_launch_sums_noalloc(nat_ctx* ctx, int layout) {
  int64_t grid = 0;
  const nat::SumArgs s = sum_args(ctx, layout, &grid);
  if (grid > NAT_SUMS_SCRATCH_BLOCKS) {
    ctx->err = "nat_sums_noalloc: " + std::to_string(grid * nat::SUM_BS) +
               " (cell, node) pairs exceed the fixed scratch of " +
               std::to_string((int64_t)NAT_SUMS_SCRATCH_BLOCKS * nat::SUM_BS) + "; use nat_sums";
    return NAT_ERR_ARG;
  }
  double* part = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (ctx->device >= 0 && ctx->device < kMaxDevices) part = g_scratch_ptr[ctx->device];
  }
  if (!part) {
    ctx->err = "nat_sums_noalloc: the device's scratch address was not looked up at context creation";
    return NAT_ERR_STATE;
  }
  return launch_sums_into(ctx, s, grid, part);
}
