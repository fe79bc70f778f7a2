// Fused semi-Lagrangian stream + BGK collide for arbitrary (non-affine) lattice
// hex meshes.
//
// The assembled reference operator stores, per (DoF, direction), the shape values
// of the departure cell at the departure point, dropping |w| < 1e-10
// (advection/SemiLagrangian.cpp:457-503, SemiLagrangianTools.cpp:15-64). Here each
// (direction, DoF) keeps only its departure cell and unit coordinates (assembled in
// FP64 on the GPU); the (p+1)^D tensor-product Lagrange weights are re-evaluated in
// registers every step with the same cut-off. That is 4 + D*sizeof(T) bytes per row
// instead of (p+1)^D stored weights plus column indices.
#include "nat_collide.hpp"

namespace nat {

template <typename T, typename XT>
struct GeneralArgs {
  const T* fin;
  T* fout;
  int64_t stride;
  int64_t n;          // FE_Q DoFs
  int nfe[3];         // FE_Q lattice nodes per axis
  int N[3];           // cells per axis
  const int* cell;    // FP64: [(alpha-1) * n + i] departure cell (-1: identity row)
  const XT* xi;       // FP64: unit coordinates; FP32: offsets from the anchor node, D per record
  const int* code;    // FP32: cell * (p+1)^D + anchor (-1: identity row)
  T gll[MAXP + 1];
  T inv_den[MAXP + 1];
  T dgl[MAXP + 1][MAXP + 1];  // x_a - x_k
};

template <int P, typename T>
__device__ __forceinline__ void lagrange_all(const T (&x)[MAXP + 1], const T (&inv)[MAXP + 1], T xi, T (&v)[P + 1]) {
#pragma unroll
  for (int j = 0; j <= P; j++) {
    T s = inv[j];
#pragma unroll
    for (int m = 0; m <= P; m++)
      if (m != j) s *= (xi - x[m]);
    v[j] = s;
  }
}

template <int D, int Q, int P, typename T, typename XT>
__global__ __launch_bounds__(256) void k_sl_general(GeneralArgs<T, XT> a, CollideArgs<T> ca, int write_macro) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  T fs[Q];
  fs[0] = a.fin[i];
  const int n0 = a.nfe[0], n1 = a.nfe[1];
#pragma unroll
  for (int al = 1; al < Q; al++) {
    const int64_t r = (int64_t)(al - 1) * a.n + i;
    const T* src = a.fin + (int64_t)al * a.stride;
    if constexpr (sizeof(T) == 8) {
      const int c = a.cell[r];
      if (c < 0) {  // tracker life-time limit: identity row
        fs[al] = src[i];
        continue;
      }
      T lx[P + 1], ly[P + 1], lz[P + 1];
      lagrange_all<P, T>(a.gll, a.inv_den, (T)a.xi[r * D + 0], lx);
      lagrange_all<P, T>(a.gll, a.inv_den, (T)a.xi[r * D + 1], ly);
      if constexpr (D == 3) lagrange_all<P, T>(a.gll, a.inv_den, (T)a.xi[r * D + 2], lz);
      const int cx = c % a.N[0];
      const int cy = (c / a.N[0]) % a.N[1];
      const int cz = (D == 3) ? c / (a.N[0] * a.N[1]) : 0;
      const int64_t base = (int64_t)cx * P + (int64_t)n0 * ((int64_t)cy * P + (int64_t)n1 * ((int64_t)cz * P));
      T acc = 0;
      if constexpr (D == 2) {
#pragma unroll
        for (int my = 0; my <= P; my++)
#pragma unroll
          for (int mx = 0; mx <= P; mx++) {
            const T w = lx[mx] * ly[my];
            if (!(fabs(w) < T(1e-10))) acc += w * src[base + mx + (int64_t)n0 * my];
          }
      } else {
#pragma unroll
        for (int mz = 0; mz <= P; mz++)
#pragma unroll
          for (int my = 0; my <= P; my++)
#pragma unroll
            for (int mx = 0; mx <= P; mx++) {
              const T w = lx[mx] * ly[my] * lz[mz];
              if (!(fabs(w) < T(1e-10))) acc += w * src[base + mx + (int64_t)n0 * (my + (int64_t)n1 * mz)];
            }
      }
      fs[al] = acc;
    } else {
      // FP32: weights from the anchor offsets, delta form around the anchor node
      // (its own term is w_a (f_a - f_a) = 0, so it needs no special case)
      const int code = a.code[r];
      if (code < 0) {
        fs[al] = src[i];
        continue;
      }
      constexpr int NP = P + 1;
      constexpr int NL = (D == 3) ? NP * NP * NP : NP * NP;
      const int c = code / NL;
      const int lin = code - c * NL;
      const int ax = lin % NP, ay = (lin / NP) % NP, az = (D == 3) ? lin / (NP * NP) : 0;
      T lx[NP], ly[NP], lz[NP];
      auto basis = [&](int anc, T d, T(&v)[NP]) {
#pragma unroll
        for (int j = 0; j <= P; j++) {
          T t = a.inv_den[j];
#pragma unroll
          for (int k = 0; k <= P; k++)
            if (k != j) t *= (a.dgl[anc][k] + d);
          v[j] = t;
        }
      };
      basis(ax, (T)a.xi[r * D + 0], lx);
      basis(ay, (T)a.xi[r * D + 1], ly);
      if constexpr (D == 3) basis(az, (T)a.xi[r * D + 2], lz);
      const int cx = c % a.N[0];
      const int cy = (c / a.N[0]) % a.N[1];
      const int cz = (D == 3) ? c / (a.N[0] * a.N[1]) : 0;
      const int64_t base = (int64_t)cx * P + (int64_t)n0 * ((int64_t)cy * P + (int64_t)n1 * ((int64_t)cz * P));
      const T fa = src[base + ax + (int64_t)n0 * (ay + (int64_t)n1 * az)];
      T acc = 0;
      if constexpr (D == 2) {
#pragma unroll
        for (int my = 0; my <= P; my++)
#pragma unroll
          for (int mx = 0; mx <= P; mx++) {
            const T w = lx[mx] * ly[my];
            if (!(fabs(w) < T(1e-10))) acc += w * (src[base + mx + (int64_t)n0 * my] - fa);
          }
      } else {
#pragma unroll
        for (int mz = 0; mz <= P; mz++)
#pragma unroll
          for (int my = 0; my <= P; my++)
#pragma unroll
            for (int mx = 0; mx <= P; mx++) {
              const T w = lx[mx] * ly[my] * lz[mz];
              if (!(fabs(w) < T(1e-10))) acc += w * (src[base + mx + (int64_t)n0 * (my + (int64_t)n1 * mz)] - fa);
            }
      }
      fs[al] = fa + acc;
    }
  }
  bgk_collide<Q, T>(fs, ca, i, write_macro != 0);
#pragma unroll
  for (int q = 0; q < Q; q++) a.fout[(int64_t)q * a.stride + i] = fs[q];
}

template <int D, int Q, int P, typename T, typename XT>
static int launch_general(nat_ctx* ctx, const GeneralArgs<T, XT>& a, const CollideArgs<T>& ca, int wm) {
  const int bs = 256;
  hipLaunchKernelGGL((k_sl_general<D, Q, P, T, XT>), dim3((unsigned)((a.n + bs - 1) / bs)), dim3(bs), 0, ctx->stream,
                     a, ca, wm);
  NAT_HIP(hipGetLastError());
  return 0;
}

template <typename T, typename XT>
static int dispatch_general(nat_ctx* ctx, const GeneralArgs<T, XT>& a, const CollideArgs<T>& ca, int wm) {
  const int Q = ctx->Q, P = ctx->p;
#define NAT_G(DD, QQ, PP) \
  if (Q == QQ && P == PP) return launch_general<DD, QQ, PP, T, XT>(ctx, a, ca, wm);
  NAT_G(2, 9, 1)
  NAT_G(2, 9, 2)
  NAT_G(2, 9, 3)
  NAT_G(2, 9, 4)
  NAT_G(2, 9, 5)
  NAT_G(2, 9, 6)
  NAT_G(3, 19, 1)
  NAT_G(3, 19, 2)
  NAT_G(3, 19, 3)
  NAT_G(3, 19, 4)
  NAT_G(3, 19, 6)
  NAT_G(3, 27, 1)
  NAT_G(3, 27, 2)
  NAT_G(3, 27, 3)
  NAT_G(3, 27, 4)
  NAT_G(3, 27, 6)
#undef NAT_G
  ctx->err = "general SL kernel not instantiated for this (stencil, order)";
  return NAT_ERR_UNSUPPORTED;
}

template <typename T, typename XT>
static void fill_general(nat_ctx* ctx, GeneralArgs<T, XT>& a, const XT* xi) {
  const int src = ctx->cur, dst = 1 - ctx->cur;
  a.fin = (const T*)ctx->f[src];
  a.fout = (T*)ctx->f[dst];
  a.stride = ctx->stride;
  a.n = ctx->n_dofs;
  for (int d = 0; d < 3; d++) {
    a.nfe[d] = ctx->nfe[d];
    a.N[d] = d < ctx->D ? ctx->desc.ncells[d] : 1;
  }
  a.cell = ctx->dep_cell;
  a.code = ctx->dep_code32;
  a.xi = xi;
  double x[MAXP + 1], w[MAXP + 1];
  gll_points(ctx->p, x, w);
  for (int j = 0; j <= MAXP; j++) {
    a.gll[j] = j <= ctx->p ? (T)x[j] : T(0);
    double d = 1;
    for (int m = 0; m <= ctx->p; m++)
      if (m != j && j <= ctx->p) d *= (x[j] - x[m]);
    a.inv_den[j] = j <= ctx->p ? (T)(1.0 / d) : T(0);
  }
  for (int j = 0; j <= MAXP; j++)
    for (int k = 0; k <= MAXP; k++) a.dgl[j][k] = (j <= ctx->p && k <= ctx->p) ? (T)(x[j] - x[k]) : T(0);
}

}  // namespace nat

int nat_launch_sl_general(nat_ctx* ctx, int write_macro) {
  using namespace nat;
  const int dst = 1 - ctx->cur;
  int rc;
  if (ctx->fp32) {
    GeneralArgs<float, float> a;
    fill_general<float, float>(ctx, a, ctx->dep_xi32);
    CollideArgs<float> ca;
    ca.omega = (float)(1.0 / ctx->tau);
    ca.omega64 = 1.0 / ctx->tau;
    ca.rho0 = (float)(ctx->desc.rho0 > 0 ? ctx->desc.rho0 : 1.0);
    ca.scaling = ctx->st.s;
    ca.rho = (float*)ctx->rho;
    ca.u = (float*)ctx->u;
    ca.mstride = ctx->n_store;
    ca.status = ctx->d_status;
    rc = dispatch_general<float, float>(ctx, a, ca, write_macro);
  } else {
    GeneralArgs<double, double> a;
    fill_general<double, double>(ctx, a, ctx->dep_xi64);
    CollideArgs<double> ca;
    ca.omega = 1.0 / ctx->tau;
    ca.omega64 = 1.0 / ctx->tau;
    ca.rho0 = 0.0;
    ca.scaling = ctx->st.s;
    ca.rho = (double*)ctx->rho;
    ca.u = (double*)ctx->u;
    ca.mstride = ctx->n_store;
    ca.status = ctx->d_status;
    rc = dispatch_general<double, double>(ctx, a, ca, write_macro);
  }
  if (rc == 0) ctx->cur = dst;
  return rc;
}
