// Fused semi-Lagrangian stream + BGK collide on affine periodic Cartesian meshes.
//
// On an axis-aligned uniform mesh the reference operator A_alpha
// (advection/SemiLagrangian.cpp:145-522) is a tensor product of 1-D operators:
// the departure point of node k along axis a lies in cell floor(k/p) + o[j] at a
// unit coordinate that only depends on the node's position j = k mod p in its cell
// and on the velocity component. Every row of A_alpha is therefore
//   f_alpha(x_k) = sum_{m,n,l} Wx[jx][m] Wy[jy][n] Wz[jz][l] f_alpha(src nodes)
// and is applied here by sum factorisation: per direction the source window of
// a tile is staged in LDS once and interpolated axis by axis (x, then y, then z),
// so each population is read from HBM once per step and written once.
//
// Storage: unique periodic nodes (N_a p per axis, x fastest); the FE_Q DoFs on the
// upper periodic faces are copies of the lower ones and are produced at download.
// Precision T = double, or float on shifted populations f - w rho0.
#include "nat_collide.hpp"

namespace nat {

template <typename T>
struct StructArgs {
  const T* fin;
  T* fout;
  int64_t stride;
  int n[3];
  const T* w;      // [3][MAXSHIFT][MAXP][MAXP+1]
};

// Source-cell pattern of a velocity component (checked at assembly time):
//   component > 0 (departure towards -x): the vertex node (j = 0) reads the cell
//   below, nodes j >= 1 read their own cell; component < 0: every node reads its
//   own cell. So the source window of a cell-aligned tile [X0, X0+T) is
//   [X0-P, X0+T] (component > 0), [X0, X0+T] (< 0) or [X0, X0+T) (= 0).
template <int C, int T_, int P>
struct Win {
  static constexpr int lo = C > 0 ? -P : 0;
  static constexpr int n = C > 0 ? T_ + P + 1 : (C < 0 ? T_ + 1 : T_);
};

constexpr int pow2ceil(int v) {
  int r = 1;
  while (r < v) r <<= 1;
  return r;
}

template <int D, int Q, int P, typename T, int TX, int TY, int TZ>
__global__ __launch_bounds__(TX* TY* TZ) void k_sl_struct(StructArgs<T> a, CollideArgs<T> ca, int write_macro) {
  using L = Lattice<Q>;
  static_assert(TX % P == 0 && TY % P == 0 && (D == 2 || TZ % P == 0), "tiles must be cell aligned");
  constexpr int NT = TX * TY * TZ;
  constexpr int WX = TX + P + 1, WY = TY + P + 1, WZ = (D == 3) ? TZ + P + 1 : 1;
  constexpr int RW = pow2ceil(WX);  // lanes per staged row
  static_assert(NT % TX == 0, "block must tile the x-pass rows");
  __shared__ T win[WX * WY * WZ];
  __shared__ T b1[TX * WY * WZ];
  __shared__ T b2[(D == 3) ? TX * TY * WZ : 1];
  __shared__ T wt_s[3 * 2 * P * (P + 1)];  // [axis][shift 0:+c 1:-c][class][m]

  const int tid = threadIdx.x;
  for (int t = tid; t < 3 * 2 * P * (P + 1); t += NT) {
    const int ax = t / (2 * P * (P + 1)), sh = (t / (P * (P + 1))) % 2, j = (t / (P + 1)) % P, m = t % (P + 1);
    wt_s[t] = a.w[((ax * MAXSHIFT + sh) * MAXP + j) * (MAXP + 1) + m];
  }
  __syncthreads();
  const int tid0 = tid;
  const int X00 = blockIdx.x * TX, Y00 = blockIdx.y * TY, Z00 = blockIdx.z * TZ;
  const int nx = a.n[0], ny = a.n[1], nz = a.n[2];
  const bool active = (X00 + tid % TX < nx) && (Y00 + (tid / TX) % TY < ny) && (Z00 + tid / (TX * TY) < nz);
  const int64_t node =
      (int64_t)(X00 + tid % TX) + (int64_t)nx * ((Y00 + (tid / TX) % TY) + (int64_t)ny * (Z00 + tid / (TX * TY)));

  T fs[Q];
  fs[0] = active ? a.fin[node] : T(0);

  static_for<1, Q>([&](auto A) {
    constexpr int al = decltype(A)::value;
    constexpr int ex = L::e[al][0], ey = L::e[al][1], ez = (D == 3) ? L::e[al][2] : 0;
    using Wx = Win<ex, TX, P>;
    using Wy = Win<ey, TY, P>;
    using Wz = Win<ez, (D == 3 ? TZ : 1), (D == 3 ? P : 0)>;
    constexpr int wnx = Wx::n, wny = Wy::n, wnz = (D == 3) ? Wz::n : 1;
    const T* __restrict__ src = a.fin + (int64_t)al * a.stride;
    // opaque copies: stop the compiler from sharing index arithmetic across
    // directions (it would keep every direction's addresses live at once)
    int tid = tid0, X0 = X00, Y0 = Y00, Z0 = Z00;
    asm volatile("" : "+v"(tid), "+s"(X0), "+s"(Y0), "+s"(Z0));
    const int ox = tid % TX, oy = (tid / TX) % TY, oz = tid / (TX * TY);
    const int jx = ox % P, jy = oy % P, jz = oz % P;
    const int wlx = X0 + Wx::lo, wly = Y0 + Wy::lo, wlz = Z0 + ((D == 3) ? Wz::lo : 0);

    // stage the window with periodic wrap; RW (power of two) slots per row
    {
      constexpr int rows = wny * wnz;
      for (int t = tid; t < rows * RW; t += NT) {
        const int ix = t % RW, r = t / RW;
        if (ix >= wnx) continue;
        const int iy = r % wny, iz = r / wny;
        // windows start at most P nodes below 0 and P <= n, so +n makes the
        // argument non-negative; the modulo also covers tiles wider than the axis
        const int gx = (wlx + ix + nx) % nx;
        const int gy = (wly + iy + ny) % ny;
        const int gz = (D == 3) ? (wlz + iz + nz) % nz : 0;
        win[ix + WX * (iy + WY * iz)] = src[(int64_t)gx + (int64_t)nx * (gy + (int64_t)ny * gz)];
      }
    }
    __syncthreads();

    // x pass: b1[ox][iy][iz] (win itself when ex == 0)
    if constexpr (ex != 0) {
      const T* wt = wt_s + ((0 * 2 + (ex > 0 ? 0 : 1)) * P + jx) * (P + 1);
      T wv[P + 1];
#pragma unroll
      for (int m = 0; m <= P; m++) wv[m] = wt[m];
      // source cell start inside the window
      const int s0 = (ox / P) * P + ((ex > 0 && jx != 0) ? P : 0);
      constexpr int rows = wny * wnz;
      for (int r = tid / TX; r < rows; r += NT / TX) {
        const int iy = r % wny, iz = r / wny;
        const T* wr = win + WX * (iy + WY * iz) + s0;
        T acc = 0;
#pragma unroll
        for (int m = 0; m <= P; m++) acc += wv[m] * wr[m];
        b1[ox + TX * (iy + WY * iz)] = acc;
      }
      __syncthreads();
    }
    const T* B1 = (ex != 0) ? b1 : win;
    constexpr int b1x = (ex != 0) ? TX : WX;

    T v;
    if constexpr (D == 2) {
      if constexpr (ey != 0) {
        const T* wt = wt_s + ((1 * 2 + (ey > 0 ? 0 : 1)) * P + jy) * (P + 1);
        const int s0 = (oy / P) * P + ((ey > 0 && jy != 0) ? P : 0);
        T acc = 0;
#pragma unroll
        for (int n = 0; n <= P; n++) acc += wt[n] * B1[ox + b1x * (s0 + n)];
        v = acc;
      } else {
        v = B1[ox + b1x * oy];
      }
    } else {
      const T* B2 = B1;
      constexpr int b2x = (ey != 0) ? TX : b1x;
      constexpr int b2y = (ey != 0) ? TY : WY;
      if constexpr (ey != 0) {
        const T* wt = wt_s + ((1 * 2 + (ey > 0 ? 0 : 1)) * P + jy) * (P + 1);
        T wv[P + 1];
#pragma unroll
        for (int n = 0; n <= P; n++) wv[n] = wt[n];
        const int s0 = (oy / P) * P + ((ey > 0 && jy != 0) ? P : 0);
#pragma unroll
        for (int iz = oz; iz < wnz; iz += TZ) {
          const T* br = B1 + ox + b1x * (s0 + WY * iz);
          T acc = 0;
#pragma unroll
          for (int n = 0; n <= P; n++) acc += wv[n] * br[b1x * n];
          b2[ox + TX * (oy + TY * iz)] = acc;
        }
        __syncthreads();
        B2 = b2;
      }
      if constexpr (ez != 0) {
        const T* wt = wt_s + ((2 * 2 + (ez > 0 ? 0 : 1)) * P + jz) * (P + 1);
        const int s0 = (oz / P) * P + ((ez > 0 && jz != 0) ? P : 0);
        T acc = 0;
#pragma unroll
        for (int l = 0; l <= P; l++) acc += wt[l] * B2[ox + b2x * (oy + b2y * (s0 + l))];
        v = acc;
      } else {
        v = B2[ox + b2x * (oy + b2y * oz)];
      }
    }
    // materialise this direction's value now (otherwise its operands stay live
    // until the collision)
    asm volatile("" : "+v"(v));
    fs[al] = v;
    __syncthreads();
    asm volatile("" ::: "memory");
  });

  if (!active) return;
  bgk_collide<Q, T>(fs, ca, node, write_macro != 0);
#pragma unroll
  for (int i = 0; i < Q; i++) a.fout[(int64_t)i * a.stride + node] = fs[i];
}

template <int D, int Q, int P, typename T, int TX, int TY, int TZ>
static int launch_struct(nat_ctx* ctx, const StructArgs<T>& a, const CollideArgs<T>& ca, int write_macro) {
  dim3 grid((a.n[0] + TX - 1) / TX, (a.n[1] + TY - 1) / TY, D == 3 ? (a.n[2] + TZ - 1) / TZ : 1);
  hipLaunchKernelGGL((k_sl_struct<D, Q, P, T, TX, TY, TZ>), grid, dim3(TX * TY * TZ), 0, ctx->stream, a, ca,
                     write_macro);
  NAT_HIP(hipGetLastError());
  return 0;
}

template <typename T>
static int dispatch_struct(nat_ctx* ctx, const StructArgs<T>& a, const CollideArgs<T>& ca, int wm) {
  const int Q = ctx->Q, P = ctx->p;
#define NAT_S3(QQ, PP, TX, TY, TZ) \
  if (Q == QQ && P == PP) return launch_struct<3, QQ, PP, T, TX, TY, TZ>(ctx, a, ca, wm);
#define NAT_S2(QQ, PP, TX, TY) \
  if (Q == QQ && P == PP) return launch_struct<2, QQ, PP, T, TX, TY, 1>(ctx, a, ca, wm);
  if (ctx->D == 3) {
    NAT_S3(27, 1, 16, 8, 4)
    NAT_S3(27, 2, 16, 8, 4)
    NAT_S3(27, 3, 12, 6, 6)
    NAT_S3(27, 4, 16, 8, 4)
    NAT_S3(27, 6, 12, 6, 6)
    NAT_S3(19, 1, 16, 8, 4)
    NAT_S3(19, 2, 16, 8, 4)
    NAT_S3(19, 3, 12, 6, 6)
    NAT_S3(19, 4, 16, 8, 4)
    NAT_S3(19, 6, 12, 6, 6)
  } else {
    NAT_S2(9, 1, 32, 16)
    NAT_S2(9, 2, 32, 16)
    NAT_S2(9, 3, 24, 18)
    NAT_S2(9, 4, 32, 16)
    NAT_S2(9, 6, 24, 18)
  }
#undef NAT_S3
#undef NAT_S2
  ctx->err = "structured SL kernel not instantiated for this (stencil, order)";
  return NAT_ERR_UNSUPPORTED;
}

}  // namespace nat

int nat_launch_sl_struct(nat_ctx* ctx, int write_macro) {
  using namespace nat;
  const int src = ctx->cur, dst = 1 - ctx->cur;
  if (ctx->fp32) {
    StructArgs<float> a;
    a.fin = (const float*)ctx->f[src];
    a.fout = (float*)ctx->f[dst];
    a.stride = ctx->stride;
    for (int i = 0; i < 3; i++) a.n[i] = ctx->nu[i];
    a.w = (const float*)ctx->ax_w;
    CollideArgs<float> ca;
    ca.omega = (float)(1.0 / ctx->tau);
    ca.rho0 = (float)(ctx->desc.rho0 > 0 ? ctx->desc.rho0 : 1.0);
    ca.scaling = ctx->st.s;
    ca.rho = (float*)ctx->rho;
    ca.u = (float*)ctx->u;
    ca.mstride = ctx->n_store;
    ca.status = ctx->d_status;
    int rc = dispatch_struct<float>(ctx, a, ca, write_macro);
    if (rc == 0) ctx->cur = dst;
    return rc;
  }
  StructArgs<double> a;
  a.fin = (const double*)ctx->f[src];
  a.fout = (double*)ctx->f[dst];
  a.stride = ctx->stride;
  for (int i = 0; i < 3; i++) a.n[i] = ctx->nu[i];
  a.w = (const double*)ctx->ax_w;
  CollideArgs<double> ca;
  ca.omega = 1.0 / ctx->tau;
  ca.rho0 = 0.0;
  ca.scaling = ctx->st.s;
  ca.rho = (double*)ctx->rho;
  ca.u = (double*)ctx->u;
  ca.mstride = ctx->n_store;
  ca.status = ctx->d_status;
  int rc = dispatch_struct<double>(ctx, a, ca, write_macro);
  if (rc == 0) ctx->cur = dst;
  return rc;
}
