// Fused semi-Lagrangian stream + BGK collide for arbitrary (non-affine) lattice
// hex meshes.
//
// The assembled reference operator stores, per (DoF, direction), the shape values
// of the departure cell at the departure point, dropping |w| < 1e-10
// (advection/SemiLagrangian.cpp:457-503, SemiLagrangianTools.cpp:15-64). Here each
// (direction, DoF) keeps only its departure point (assembled in FP64 on the GPU);
// the (p+1)^D tensor-product Lagrange weights are re-evaluated in FP64 registers
// every step with the same cut-off. Records: FP64 storage keeps the cell and FP64
// unit coordinates (4 + 8D bytes); FP32 storage keeps the cell and nearest GLL node
// packed in one int32 and the offsets from that node in 2^-30 fixed point (4 + 4D
// bytes, one interleaved record), instead of (p+1)^D stored weights plus column indices
// per row.
#include "nat_collide.hpp"
#include "nat_fastdiv.hpp"

#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>


namespace nat {

// Workgroup index regrouped so that each XCD (workgroup i is dispatched to XCD i % 8)
// runs a contiguous range of the grid: the gathers of a block read rows of the blocks
// one lattice plane away, which then sit in the same L2 (C3: FP32 1.60 -> 1.78 GLUPS,
// FP64 0.84 -> 0.92).
__device__ __forceinline__ int64_t xcd_block_index() {
  const int64_t b = blockIdx.x, per = gridDim.x / 8;
  return b < per * 8 ? (b % 8) * per + b / 8 : b;
}

template <typename T>
struct GeneralArgs {
  const T* fin;
  T* fout;
  int64_t stride;
  int64_t n;          // FE_Q DoFs
  int nfe[3];         // FE_Q lattice nodes per axis
  int N[3];           // cells per axis
  FastDivU divN0, divN1;  // division by N[0], N[1]
  int64_t rows;       // lattice rows along x (n / nfe[0]); 0: threads take DoFs in order
  FastDivU divM;      // division by nfe[0] - 2, the x-interior DoFs of a row
  FastDivU divG;      // division by 64 nfe[0], the DoFs of a group of 64 rows
  const int* cell;    // FP64 storage: [(alpha-1) * n + i] departure cell (-1: identity row)
  const double* xi;   //   unit coordinates, D per record
  const int* rec;     // FP32 storage: D + 1 int32 per record: cell * (p+1)^D + anchor node
                      //   (-1: identity row), then the offsets from the anchor (2^-30 units)
  int dir_bytes;      // bytes of one direction's array (buffer-load range)
  const int* wall_idx;        // walls: wall entry per DoF or -1 (null: no walls)
  const unsigned* wall_flip;  //   per entry: bit alpha set when alpha reads the opposite direction
  const double* wall_bias;    //   per entry: Q values added after interpolation
  double gll[MAXP + 1];
  double inv_den[MAXP + 1];
  double dgl[MAXP + 1][MAXP + 1];  // x_a - x_k
};

// Launch shape. The direction loop is bound by the latency of its gathers, so the
// kernel wants many waves per SIMD; holding the Q streamed populations in registers
// across that loop forces either few waves or scratch spills. Where Q populations of
// a block fit in LDS at >= 4 waves per SIMD (FP32 storage; FP64 up to Q = 19), each
// thread parks its populations in its own LDS column (no barrier: no thread reads
// another's column), the direction loop stays rolled, and the collision reloads the
// column. FP32: 6 waves (80 registers); FP64: the 4 waves its LDS allows. FP64 D3Q27
// (216 B per DoF) keeps the populations in registers at 6 waves. On 24^3 perturbed
// cells, p = 3, this shape gives D3Q19 FP32 1.25 -> 1.58 and D3Q27 FP32 0.86 -> 1.25
// GLUPS over the unrolled register form.
//
// Within one direction the rows of `zu` z-planes of the source block are loaded
// together: one plane at 80 registers, as many planes as keep the raw rows within 64
// registers at 4 waves (C3 FP64, two planes: 358 -> 342 us per step). The DEEP form of
// the FP32 3D kernel runs at 4 waves with those planes in flight and the next
// direction's record prefetched; it is launched when the grid fits one round of it,
// where each wave's dependent loads, not the number of waves, set the time (D3Q27 p = 3
// on 16^3 perturbed cells, 460 blocks: 219 us per step in the 6-wave form, 118 in the
// DEEP one). A grid that fills the 6-wave form (C3: 1520 blocks) keeps that form (179 us
// against 194 for the 4-wave one).
template <int D, int Q, int P, typename T, bool DEEP>
struct GeneralShape {
  static constexpr bool lds = sizeof(T) == 4 || Q * 8 <= 160;
  static constexpr int waves = DEEP ? 4 : (sizeof(T) == 4 ? 6 : (lds ? 4 : 6));
  // DoFs per block: 128 with FP64 storage (C3 FP64 +6% over 256), 256 with FP32
  // (128 measured 18% slower on D3Q27)
  static constexpr int bs = sizeof(T) == 8 ? 128 : 256;
  static constexpr int nz = D == 3 ? P + 1 : 1;
  static constexpr int plane_regs = (P + 1) * (P + 1) * (int)(sizeof(T) / 4);
  static constexpr int zfit = 64 / plane_regs < 1 ? 1 : (64 / plane_regs > nz ? nz : 64 / plane_regs);
  static constexpr int zu = (waves == 4 && lds) ? zfit : 1;
  // the next direction's record loaded before this direction's gathers (FP32 at 6 waves:
  // spills and loses 14%)
  static constexpr bool prefetch = sizeof(T) == 8 || DEEP;
};

// Lagrange values at x_a + d (anchor a), in FP64: prod_{k != j} (x_a - x_k + d) / den_j
template <int P>
__device__ __forceinline__ void basis_at(const double (&dgl)[MAXP + 1][MAXP + 1], const double (&inv)[MAXP + 1],
                                         int a, double d, double (&v)[P + 1]) {
#pragma unroll
  for (int j = 0; j <= P; j++) {
    double t = inv[j];
#pragma unroll
    for (int k = 0; k <= P; k++)
      if (k != j) t *= (dgl[a][k] + d);
    v[j] = t;
  }
}

template <int P>
__device__ __forceinline__ void basis_xi(const double (&x)[MAXP + 1], const double (&inv)[MAXP + 1], double xi,
                                         double (&v)[P + 1]) {
#pragma unroll
  for (int j = 0; j <= P; j++) {
    double t = inv[j];
#pragma unroll
    for (int m = 0; m <= P; m++)
      if (m != j) t *= (xi - x[m]);
    v[j] = t;
  }
}

// p+1 consecutive lattice values (one x row of a source cell) through raw buffer
// loads: byte offsets against a per-direction resource, 16-byte loads where the row
// length allows (buffer loads need dword alignment only).
template <int P, typename T>
__device__ __forceinline__ void load_row(__amdgpu_buffer_rsrc_t r, int off, double (&v)[P + 1]) {
  if constexpr (sizeof(T) == 4 && (P + 1) % 4 == 0) {
#pragma unroll
    for (int q = 0; q < (P + 1) / 4; q++) {
      const auto x = __builtin_amdgcn_raw_buffer_load_b128(r, off + 16 * q, 0, 0);
#pragma unroll
      for (int k = 0; k < 4; k++) v[4 * q + k] = (double)__builtin_bit_cast(float, (unsigned)x[k]);
    }
  } else if constexpr (sizeof(T) == 4 && (P + 1) % 2 == 0) {
#pragma unroll
    for (int q = 0; q < (P + 1) / 2; q++) {
      const auto x = __builtin_amdgcn_raw_buffer_load_b64(r, off + 8 * q, 0, 0);
      v[2 * q] = (double)__builtin_bit_cast(float, (unsigned)x[0]);
      v[2 * q + 1] = (double)__builtin_bit_cast(float, (unsigned)x[1]);
    }
  } else if constexpr (sizeof(T) == 8 && (P + 1) % 2 == 0) {
#pragma unroll
    for (int q = 0; q < (P + 1) / 2; q++) {
      const auto x = __builtin_amdgcn_raw_buffer_load_b128(r, off + 16 * q, 0, 0);
      v[2 * q] = __builtin_bit_cast(double, ((unsigned long)(unsigned)x[1] << 32) | (unsigned)x[0]);
      v[2 * q + 1] = __builtin_bit_cast(double, ((unsigned long)(unsigned)x[3] << 32) | (unsigned)x[2]);
    }
  } else {
#pragma unroll
    for (int m = 0; m <= P; m++) {
      if constexpr (sizeof(T) == 4)
        v[m] = (double)__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off + 4 * m, 0, 0));
      else
        v[m] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off + 8 * m, 0, 0));
    }
  }
}

// Weights, cut-off and arithmetic in FP64 for both storage precisions (FP32 weights
// rounded the same way at every step bias the FP32-shifted result).
//
// Velocity-neq bounce-back (SemiLagrangian.cpp:358-385, VelocityNeqBounceBack.cpp:
// 137-195): a DoF whose path met a wall reads the opposite direction's populations
// at the reflected departure point and adds 2 w (e . u_wall) / cs2 per hit. Such
// rows take the term-by-term path with per-lane source addresses. With FP32
// storage the shift w rho0 passes through, since opposite directions share w.
//
// The departure record of one (direction, DoF): FP64 storage the cell and FP64 unit
// coordinates, FP32 storage the packed cell/anchor code and fixed-point offsets.
template <int D, typename T>
struct DepRec {
  int c;
  double xi[D];
};
template <int D>
struct DepRec<D, float> {
  int c;
  int off[D];
};

template <int D, typename T>
__device__ __forceinline__ DepRec<D, T> load_rec(const GeneralArgs<T>& a, int64_t r) {
  DepRec<D, T> rec;
  if constexpr (sizeof(T) == 8) {
    rec.c = a.cell[r];
#pragma unroll
    for (int d = 0; d < D; d++) rec.xi[d] = a.xi[r * D + d];
  } else {
    if constexpr (D == 3) {  // one 16-byte load per record
      const int4 v = reinterpret_cast<const int4*>(a.rec)[r];
      rec.c = v.x;
      rec.off[0] = v.y;
      rec.off[1] = v.z;
      rec.off[2] = v.w;
    } else {
      rec.c = a.rec[r * 3];
      rec.off[0] = a.rec[r * 3 + 1];
      rec.off[1] = a.rec[r * 3 + 2];
    }
  }
  return rec;
}

// 1-D Lagrange weights of a departure record in FP64 and the lattice index of its cell's
// first node (records with rec.c >= 0 only).
template <int D, int P, typename T>
__device__ __forceinline__ int64_t dep_weights(const GeneralArgs<T>& a, const DepRec<D, T>& rec, double (&lx)[P + 1],
                                               double (&ly)[P + 1], double (&lz)[P + 1]) {
  constexpr int NP = P + 1;
  constexpr int NL = (D == 3) ? NP * NP * NP : NP * NP;
  int c;
  if constexpr (sizeof(T) == 8) {
    c = rec.c;
    basis_xi<P>(a.gll, a.inv_den, rec.xi[0], lx);
    basis_xi<P>(a.gll, a.inv_den, rec.xi[1], ly);
    if constexpr (D == 3) basis_xi<P>(a.gll, a.inv_den, rec.xi[2], lz);
  } else {
    const int code = rec.c;
    c = code / NL;
    const int lin = code - c * NL;
    const int ax = lin % NP, ay = (lin / NP) % NP, az = (D == 3) ? lin / (NP * NP) : 0;
    constexpr double q = 1.0 / 1073741824.0;
    basis_at<P>(a.dgl, a.inv_den, ax, (double)rec.off[0] * q, lx);
    basis_at<P>(a.dgl, a.inv_den, ay, (double)rec.off[1] * q, ly);
    if constexpr (D == 3) basis_at<P>(a.dgl, a.inv_den, az, (double)rec.off[2] * q, lz);
  }
  const uint32_t q0 = a.divN0.div((uint32_t)c), q1 = a.divN1.div(q0);
  const int cx = c - (int)q0 * a.N[0];
  const int cy = (int)q0 - (int)q1 * a.N[1];
  const int cz = (D == 3) ? (int)q1 : 0;
  return (int64_t)cx * P + (int64_t)a.nfe[0] * ((int64_t)cy * P + (int64_t)a.nfe[1] * ((int64_t)cz * P));
}

// Buffer resource over one direction's population array (wave-uniform base).
template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t dir_rsrc(const GeneralArgs<T>& a, const T* arr) {
  const uint64_t pb = (uint64_t)arr;
  const uint32_t phi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(pb >> 32));
  const uint32_t plo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)pb);
  return __builtin_amdgcn_make_buffer_rsrc((void*)(((uint64_t)phi << 32) | (uint64_t)plo), (short)0,
                                           __builtin_amdgcn_readfirstlane(a.dir_bytes), 0x00020000);
}

// The row of the assembled operator under the reference cut-off: every product weight
// (same factors, same order) is formed and the terms with |w| < 1e-10 dropped, the source
// rows read through buffer loads.
template <int D, int P, typename T>
__device__ __forceinline__ double kept_sum(__amdgpu_buffer_rsrc_t rs, int b0, int n0, int n1,
                                           const double (&lx)[P + 1], const double (&ly)[P + 1],
                                           const double (&lz)[P + 1]) {
  constexpr int NP = P + 1;
  constexpr int NZ = (D == 3) ? NP : 1;
  double kept = 0;
#pragma unroll 1
  for (int mz = 0; mz < NZ; mz++) {
#pragma unroll 1
    for (int my = 0; my <= P; my++) {
      double v[NP];
      load_row<P, T>(rs, b0 + (int)sizeof(T) * (n0 * (my + n1 * mz)), v);
#pragma unroll
      for (int mx = 0; mx <= P; mx++) {
        const double w = (D == 3) ? lx[mx] * ly[my] * lz[mz] : lx[mx] * ly[my];
        kept += fabs(w) < 1e-10 ? 0.0 : w * v[mx];
      }
    }
  }
  return kept;
}

// Returns the streamed population of direction al at DoF i from its departure record.
//
// Reference cut-off: product weights |w| < 1e-10 are dropped. When the product of the
// smallest 1-D magnitudes is >= 1e-10 no product can fall under it and the sum factorises
// row by row (rows through 16-, 8- or 4-byte buffer loads). Otherwise (a departure point
// on a node line of some axis, as for DoFs on the box faces, which almost every 64-DoF wave
// holds) the lane sums the kept terms again, one product at a time, with the source rows
// re-read one at a time (a compact loop: the whole wave runs it when one lane needs it, so
// it must not raise the register count of the direction loop). Lanes reflected at a wall
// (another direction's array) gather term by term.
template <int D, int Q, int P, typename T, int ZU>
__device__ __forceinline__ T gen_direction(const GeneralArgs<T>& a, int64_t i, int al, int widx, unsigned wmask,
                                           const DepRec<D, T>& rec) {
  constexpr int NP = P + 1;
  const int n0 = a.nfe[0], n1 = a.nfe[1];
  const T* own = a.fin + (int64_t)al * a.stride;
  const bool flip = (wmask >> al) & 1u;
  const T* src = flip ? a.fin + (int64_t)opposite<Q>(al) * a.stride : own;
  const double bias = widx >= 0 ? a.wall_bias[(int64_t)widx * Q + al] : 0.0;
  if (rec.c < 0) return (T)((double)own[i] + bias);  // tracker life-time limit: identity row
  double lx[NP], ly[NP], lz[NP];
  const int64_t base = dep_weights<D, P, T>(a, rec, lx, ly, lz);
  double acc = 0;
  double mn = 1;
  {
    double mx_ = fabs(lx[0]), my_ = fabs(ly[0]), mz_ = (D == 3) ? fabs(lz[0]) : 1.0;
#pragma unroll
    for (int m = 1; m <= P; m++) {
      mx_ = fmin(mx_, fabs(lx[m]));
      my_ = fmin(my_, fabs(ly[m]));
      if constexpr (D == 3) mz_ = fmin(mz_, fabs(lz[m]));
    }
    mn = mx_ * my_ * mz_;
  }
  constexpr int NZ = (D == 3) ? NP : 1;
  if (!flip) {
    const __amdgpu_buffer_rsrc_t rs = dir_rsrc<T>(a, own);
    const int b0 = (int)base * (int)sizeof(T);
#pragma unroll 1
    for (int mz0 = 0; mz0 < NZ; mz0 += ZU) {
#pragma unroll
      for (int dz = 0; dz < ZU; dz++) {
        const int mz = mz0 + dz;
        if (NZ % ZU != 0 && mz >= NZ) break;
        double accz = 0;
#pragma unroll
        for (int my = 0; my <= P; my++) {
          double v[NP];
          load_row<P, T>(rs, b0 + (int)sizeof(T) * (n0 * (my + n1 * mz)), v);
          double row = lx[0] * v[0];
#pragma unroll
          for (int mx = 1; mx <= P; mx++) row += lx[mx] * v[mx];
          accz += ly[my] * row;
        }
        acc += ((D == 3) ? lz[mz] : 1.0) * accz;
      }
    }
    if (!(mn >= 1e-10)) acc = kept_sum<D, P, T>(rs, b0, n0, n1, lx, ly, lz);
  } else if constexpr (D == 2) {
#pragma unroll
    for (int my = 0; my <= P; my++)
#pragma unroll
      for (int mx = 0; mx <= P; mx++) {
        const double w = lx[mx] * ly[my];
        if (!(fabs(w) < 1e-10)) acc += w * (double)src[base + mx + (int64_t)n0 * my];
      }
  } else {
#pragma unroll 1
    for (int mz = 0; mz <= P; mz++)
#pragma unroll
      for (int my = 0; my <= P; my++)
#pragma unroll
        for (int mx = 0; mx <= P; mx++) {
          const double w = lx[mx] * ly[my] * lz[mz];
          if (!(fabs(w) < 1e-10)) acc += w * (double)src[base + mx + (int64_t)n0 * (my + (int64_t)n1 * mz)];
        }
  }
  return (T)(acc + bias);
}

// DoF of thread g. Departure points of DoFs on the x faces of the box (x = 0 and x = L,
// unmoved vertices) sit on node lines for the directions along those faces, and such rows
// take the kernel's term-by-term cut-off path for the whole wave; in x-fastest order almost
// every 64-DoF wave holds one or two of them. Threads therefore take the lattice rows in
// groups of 64: first the x-interior DoFs of the group's rows (rows in order, DoFs in order:
// nfe[0] - 2 whole waves), then the group's 128 x-face DoFs (two waves), so that interior
// waves stay on the factorised path and the face waves sit beside their neighbours in the
// grid instead of forming a tail. Loads of an interior wave's own DoFs stay contiguous
// within each row.
template <typename T>
__device__ __forceinline__ int64_t dof_of(const GeneralArgs<T>& a, int64_t g) {
  if (a.rows == 0) return g;
  const int64_t nfe0 = a.nfe[0];
  const uint32_t m = (uint32_t)nfe0 - 2u;
  const uint32_t k = a.divG.div((uint32_t)g);
  const int64_t r0 = (int64_t)k * 64;
  const int64_t R = a.rows - r0 < 64 ? a.rows - r0 : 64;
  const uint32_t gg = (uint32_t)(g - r0 * nfe0);
  const uint32_t ni = (uint32_t)R * m;
  if (gg < ni) {
    const uint32_t q = a.divM.div(gg);
    return (r0 + q) * nfe0 + 1 + (int64_t)(gg - q * m);
  }
  const uint32_t h = gg - ni;
  return (r0 + (h >> 1)) * nfe0 + ((h & 1) ? nfe0 - 1 : 0);
}

template <int D, int Q, int P, typename T, bool DEEP>
__global__ __launch_bounds__((GeneralShape<D, Q, P, T, DEEP>::bs)) __attribute__((amdgpu_waves_per_eu(GeneralShape<D, Q, P, T, DEEP>::waves))) void k_sl_general(
    GeneralArgs<T> a, CollideArgs<T> ca, int write_macro) {
  using S = GeneralShape<D, Q, P, T, DEEP>;
  const int64_t g = xcd_block_index() * blockDim.x + threadIdx.x;
  if (g >= a.n) return;
  // the grouped order where it measured faster: the LDS forms at six waves (FP32) and four
  // (FP64); the DEEP form and the FP64 D3Q27 register form keep the lattice order
  const int64_t i = (S::lds && !DEEP) ? dof_of(a, g) : g;
  int widx = -1;
  unsigned wmask = 0;
  if (a.wall_idx) {
    widx = a.wall_idx[i];
    if (widx >= 0) wmask = a.wall_flip[widx];
  }
  T fs[Q];
  if constexpr (S::lds) {
    fs[0] = a.fin[i];
    __shared__ T col[Q - 1][S::bs];
    const int t = threadIdx.x;
    if constexpr (S::prefetch) {
      // the next direction's record is loaded before this direction's gathers, so its
      // latency overlaps them (C3 FP64 370 -> 358 us per step; FP32 at 6 waves, at its
      // 80-register cap, spilled more and lost 14%)
      DepRec<D, T> rec = load_rec<D, T>(a, i);
#pragma unroll 1
      for (int al = 1; al < Q; al++) {
        DepRec<D, T> next = rec;
        if (al + 1 < Q) next = load_rec<D, T>(a, (int64_t)al * a.n + i);
        col[al - 1][t] = gen_direction<D, Q, P, T, S::zu>(a, i, al, widx, wmask, rec);
        rec = next;
      }
    } else {
#pragma unroll 1
      for (int al = 1; al < Q; al++)
        col[al - 1][t] =
            gen_direction<D, Q, P, T, S::zu>(a, i, al, widx, wmask, load_rec<D, T>(a, (int64_t)(al - 1) * a.n + i));
    }
    // keep the compiler from forwarding the stores into registers held across the loop
    asm volatile("" ::: "memory");
#pragma unroll
    for (int al = 1; al < Q; al++) fs[al] = col[al - 1][t];
  } else {
#pragma unroll
    for (int al = 1; al < Q; al++)
      fs[al] = gen_direction<D, Q, P, T, S::zu>(a, i, al, widx, wmask, load_rec<D, T>(a, (int64_t)(al - 1) * a.n + i));
    fs[0] = a.fin[i];  // not held across the unrolled loop (FP64 D3Q27 p = 3: 181 -> 143 us)
  }
  if (write_macro & STEP_STREAM_ONLY) {  // the SL stream with its wall values (CFDSolver::stream)
#pragma unroll
    for (int q = 0; q < Q; q++) a.fout[(int64_t)q * a.stride + i] = fs[q];
    return;
  }
  bgk_collide_store_pairs<Q, T>(fs, ca, i, (write_macro & STEP_MACRO) != 0, a.fout, a.stride);
}

template <int D, int Q, int P, typename T, bool DEEP>
static int launch_general_shape(nat_ctx* ctx, const GeneralArgs<T>& a, const CollideArgs<T>& ca, int wm) {
  constexpr int bs = GeneralShape<D, Q, P, T, DEEP>::bs;
  hipLaunchKernelGGL((k_sl_general<D, Q, P, T, DEEP>), dim3((unsigned)((a.n + bs - 1) / bs)), dim3(bs), 0,
                     ctx->stream, a, ca, wm);
  NAT_HIP(hipGetLastError());
  return 0;
}

// The DEEP form (FP32 3D only) when the autotuner chose it or, by default, when its grid fits
// one round: 4 blocks of 256 DoFs per CU at 4 waves per SIMD.
template <int D, int Q, int P, typename T>
static bool deep_form(const nat_ctx* ctx, int64_t n) {
  if constexpr (sizeof(T) == 4 && D == 3) {
    constexpr int bs = GeneralShape<D, Q, P, T, true>::bs;
    const int64_t blocks = (n + bs - 1) / bs;
    return ctx->gen_pref >= 0 ? ctx->gen_pref == 1 : (ctx->n_cu > 0 && blocks <= 4 * (int64_t)ctx->n_cu);
  } else {
    return false;
  }
}

template <int D, int Q, int P, typename T>
static int launch_general(nat_ctx* ctx, const GeneralArgs<T>& a, const CollideArgs<T>& ca, int wm) {
  if constexpr (sizeof(T) == 4 && D == 3)
    if (deep_form<D, Q, P, T>(ctx, a.n)) return launch_general_shape<D, Q, P, T, true>(ctx, a, ca, wm);
  return launch_general_shape<D, Q, P, T, false>(ctx, a, ca, wm);
}

// Calls f(D, Q, P) (as integral constants) for the context's instantiation.
template <typename F>
static int dispatch_dqp(nat_ctx* ctx, F&& f) {
  const int Q = ctx->Q, P = ctx->p;
#define NAT_G(DD, QQ, PP) \
  if (Q == QQ && P == PP)   \
    return f(std::integral_constant<int, DD>{}, std::integral_constant<int, QQ>{}, std::integral_constant<int, PP>{});
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

template <typename T>
static int dispatch_general(nat_ctx* ctx, const GeneralArgs<T>& a, const CollideArgs<T>& ca, int wm) {
  return dispatch_dqp(ctx, [&](auto d, auto q, auto p) {
    return launch_general<decltype(d)::value, decltype(q)::value, decltype(p)::value, T>(ctx, a, ca, wm);
  });
}

template <typename T>
static void fill_general(nat_ctx* ctx, GeneralArgs<T>& a) {
  const int src = ctx->cur, dst = 1 - ctx->cur;
  a.fin = (const T*)ctx->f[src];
  a.fout = (T*)ctx->f[dst];
  a.stride = ctx->stride;
  a.n = ctx->n_dofs;
  for (int d = 0; d < 3; d++) {
    a.nfe[d] = ctx->nfe[d];
    a.N[d] = d < ctx->D ? ctx->desc.ncells[d] : 1;
  }
  a.divN0.set((uint32_t)a.N[0]);
  a.divN1.set((uint32_t)a.N[1]);
  a.rows = 0;
  if (a.nfe[0] >= 3 && a.n % a.nfe[0] == 0) {
    a.rows = a.n / a.nfe[0];
    a.divM.set((uint32_t)(a.nfe[0] - 2));
    a.divG.set((uint32_t)(64 * a.nfe[0]));
  }
  a.cell = ctx->dep_cell;
  a.xi = ctx->dep_xi64;
  a.rec = ctx->dep_rec32;
  a.dir_bytes = (int)(ctx->stride * (int64_t)sizeof(T));
  a.wall_idx = ctx->has_walls ? ctx->wall_idx : nullptr;
  a.wall_flip = ctx->wall_flip;
  a.wall_bias = ctx->wall_bias;
  double x[MAXP + 1], w[MAXP + 1];
  gll_points(ctx->p, x, w);
  for (int j = 0; j <= MAXP; j++) {
    a.gll[j] = j <= ctx->p ? x[j] : 0.0;
    double d = 1;
    for (int m = 0; m <= ctx->p; m++)
      if (m != j && j <= ctx->p) d *= (x[j] - x[m]);
    a.inv_den[j] = j <= ctx->p ? 1.0 / d : 0.0;
  }
  for (int j = 0; j <= MAXP; j++)
    for (int k = 0; k <= MAXP; k++) a.dgl[j][k] = (j <= ctx->p && k <= ctx->p) ? (x[j] - x[k]) : 0.0;
}

}  // namespace nat

namespace {

// One fused step from f[cur] into f[1 - cur], without advancing cur.
int general_launch(nat_ctx* ctx, int write_macro) {
  using namespace nat;
  if (ctx->stride * (int64_t)(ctx->fp32 ? 4 : 8) > INT32_MAX) {
    ctx->err = "general backend: one direction's array exceeds 2 GiB";
    return NAT_ERR_UNSUPPORTED;
  }
  if (ctx->fp32) {
    GeneralArgs<float> a;
    fill_general<float>(ctx, a);
    CollideArgs<float> ca;
    ca.omega = (float)(1.0 / ctx->tau);
    ca.omega64 = 1.0 / ctx->tau;
    ca.rho0 = (float)(ctx->desc.rho0 > 0 ? ctx->desc.rho0 : 1.0);
    ca.scaling = ctx->st.s;
    ca.rho = (float*)ctx->rho;
    ca.u = (float*)ctx->u;
    ca.mstride = ctx->n_store;
    ca.status = ctx->d_status;
    ca.stab = ctx->stab_on();
    return dispatch_general<float>(ctx, a, ca, write_macro);
  }
  GeneralArgs<double> a;
  fill_general<double>(ctx, a);
  CollideArgs<double> ca;
  ca.omega = 1.0 / ctx->tau;
  ca.omega64 = 1.0 / ctx->tau;
  ca.rho0 = 0.0;
  ca.scaling = ctx->st.s;
  ca.rho = (double*)ctx->rho;
  ca.u = (double*)ctx->u;
  ca.mstride = ctx->n_store;
  ca.status = ctx->d_status;
  ca.stab = ctx->stab_on();
  return dispatch_general<double>(ctx, a, ca, write_macro);
}

struct GenKey {
  int Q, p, walls;
  int64_t n;
  bool operator<(const GenKey& o) const {
    return std::tie(Q, p, walls, n) < std::tie(o.Q, o.p, o.walls, o.n);
  }
};
std::mutex g_gen_mutex;
std::map<GenKey, int> g_gen_cache;

}  // namespace

int nat_launch_sl_general(nat_ctx* ctx, int write_macro) {
  const int rc = general_launch(ctx, write_macro);
  if (rc == 0) ctx->cur = 1 - ctx->cur;
  return rc;
}

// Form of the general SL kernel picked by measurement. FP32 3D contexts have two (the
// six-wave form and the DEEP form): each is timed for nsteps steps on the context's own mesh,
// writing the idle ping-pong buffer (populations, fields and time stay as they were), and the
// faster is kept for this context and, process-wide, for later contexts of the same
// instantiation and DoF count. Both forms run the same arithmetic in the same order, so
// results do not depend on the choice. tile receives the chosen form's DoFs per block, waves
// per SIMD it is built for, and z-planes of source rows loaded together.
int nat_internal_autotune_general(nat_ctx* ctx, int nsteps, int* tile, double* best_us) {
  using namespace nat;
  if (!ctx->assembled) return NAT_ERR_STATE;
  if (best_us) *best_us = 0;
  if (ctx->fp32 && ctx->D == 3) {
    const GenKey key{ctx->Q, ctx->p, ctx->has_walls ? 1 : 0, ctx->n_dofs};
    int cached = -1;
    {
      std::lock_guard<std::mutex> lk(g_gen_mutex);
      auto it = g_gen_cache.find(key);
      if (it != g_gen_cache.end()) cached = it->second;
    }
    if (cached >= 0) {
      ctx->gen_pref = cached;
    } else {
      if (nsteps < 1) nsteps = 10;
      hipEvent_t e0 = nullptr, e1 = nullptr;
      NAT_HIP(hipEventCreate(&e0));
      if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        ctx->err = "hipEventCreate failed";
        return NAT_ERR_HIP;
      }
      int best = -1, rc = 0;
      float best_ms = 0;
      for (int pref = 0; pref < 2 && !rc; pref++) {
        ctx->gen_pref = pref;
        for (int s = 0; s < 2 && !rc; s++) rc = general_launch(ctx, 0);
        if (rc) break;
        if (hipEventRecord(e0, ctx->stream) != hipSuccess) {
          ctx->err = "autotune timing events failed";
          rc = NAT_ERR_HIP;
          break;
        }
        for (int s = 0; s < nsteps && !rc; s++) rc = general_launch(ctx, 0);
        if (rc) break;
        float ms = 0;
        if (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
          ctx->err = "autotune timing events failed";
          rc = NAT_ERR_HIP;
          break;
        }
        if (best < 0 || ms < best_ms) {
          best = pref;
          best_ms = ms;
        }
      }
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      if (rc) {
        ctx->gen_pref = -1;
        return rc;
      }
      ctx->gen_pref = best;
      if (best_us) *best_us = 1e3 * best_ms / nsteps;
      std::lock_guard<std::mutex> lk(g_gen_mutex);
      g_gen_cache[key] = best;
    }
  }
  return dispatch_dqp(ctx, [&](auto d, auto q, auto p) {
    constexpr int D = decltype(d)::value, Q = decltype(q)::value, P = decltype(p)::value;
    auto put = [&](auto shape) {
      using S = decltype(shape);
      tile[0] = S::bs;
      tile[1] = S::waves;
      tile[2] = S::zu;
    };
    if (ctx->fp32) {
      if (deep_form<D, Q, P, float>(ctx, ctx->n_dofs))
        put(GeneralShape<D, Q, P, float, true>{});
      else
        put(GeneralShape<D, Q, P, float, false>{});
    } else {
      put(GeneralShape<D, Q, P, double, false>{});
    }
    return 0;
  });
}
