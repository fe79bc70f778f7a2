// Fused semi-Lagrangian stream + BGK collide on affine periodic Cartesian meshes.
//
// On an axis-aligned uniform mesh the reference operator A_alpha
// (advection/SemiLagrangian.cpp:145-522) is a tensor product of 1-D operators:
// the departure point of node k along axis a lies in cell floor(k/p) + o[j] at a
// unit coordinate that only depends on the node's position j = k mod p in its cell
// and on the sign of the velocity component. Every row of A_alpha is therefore
//   f_alpha(x_k) = sum_{m,n,l} Wx[jx][m] Wy[jy][n] Wz[jz][l] f_alpha(src nodes)
// and is applied by sum factorisation, one 1-D pass per non-zero velocity
// component, in the order z, y, x.
//
// Work shapes. A block owns a tile of TX*TY*TZ output nodes, one thread per node
// for the collision. In a 1-D pass a work item owns one source cell of one line:
// it reads the cell's p+1 values once and produces the p outputs that depart from
// that cell (for a positive component the nodes 1..p of the cell, the last one
// being the next cell's vertex; otherwise the nodes 0..p-1). The first pass of a
// direction reads its source window straight from global memory (L2-resident
// halos); later passes read the previous pass's output from LDS; the last pass
// writes the tile's values, which each thread picks up for its own node.
//
// Scheduling. Directions are processed in groups of B, ordered by their number of
// passes so that a group needs as few barriers as its longest member: all first
// passes of a group, one barrier, all second passes, one barrier, ... The global
// loads of the next group's first passes are issued right after the current
// group's first passes, so they are in flight during the remaining passes
// (barriers here wait for LDS traffic only).
//
// Arithmetic. Interpolation weights, pass accumulations, LDS intermediates and
// the collision run in FP64 for both storage precisions (non-packed FP32 and FP64
// FMAs have the same rate on CDNA4, and FP64 intermediates need no conversions
// between passes). With FP32 storage the populations are shifted,
// f~ = f - w rho0, so only perturbations are rounded to FP32.
//
// Storage: unique periodic nodes (N_a p per axis, x fastest); the FE_Q DoFs on the
// upper periodic faces are copies of the lower ones and are produced at download.
#include <map>
#include <mutex>

#include "nat_collide.hpp"

// directions per group (FP32 / FP64 storage) and the D3Q27 p=4 tile
#ifndef NAT_STRUCT_B32
#define NAT_STRUCT_B32 3
#endif
#ifndef NAT_STRUCT_B64
#define NAT_STRUCT_B64 3
#endif
#ifndef NAT_PREFETCH32
#define NAT_PREFETCH32 1  // groups of first-pass loads in flight (FP32 storage)
#endif
#ifndef NAT_PREFETCH64
#define NAT_PREFETCH64 1
#endif
#ifndef NAT_TILE_27_4
#define NAT_TILE_27_4 16, 8, 8
#endif


namespace nat {

template <typename T>
struct StructArgs {
  const T* fin;
  T* fout;
  int64_t stride;
  int n[3];
  int tiles[3];       // tiles per axis
  int xcd_remap;      // 1: group spatially adjacent tiles on the same XCD
  const double* w;    // [3][MAXSHIFT][MAXP][MAXP+1] FP64 weights
  int dir_bytes;      // bytes of one direction's array (buffer-load range)
};

// Source window of a tile along one axis for a velocity component c (pattern
// checked at assembly): c > 0 -> nodes [X0-P, X0+T] (the vertex node reads the
// cell below), c < 0 -> [X0, X0+T], c = 0 -> [X0, X0+T).
constexpr int win_n(int c, int t, int p) { return c > 0 ? t + p + 1 : (c < 0 ? t + 1 : t); }
constexpr int win_lo(int c, int p) { return c > 0 ? -p : 0; }

// Compile-time geometry of one 1-D pass. A pass reads valid input coordinates
// (window or tile extents per axis) and writes the valid outputs of each source
// cell into a compact layout whose extent along the pass axis is the tile's. For a
// positive component the first and last source cells of a line are partial: the
// first only feeds the tile's first vertex, the last misses its final output.
struct PG {
  int D, P, T[3], c[3], K;
  int np, A, C;
  bool first, last;
  int ext[3];              // valid input extents
  int pin[3], inoff[3];    // physical input extents and offsets (K > 0)
  int pout[3], outoff;     // physical output extents, offset of valid outputs along A
  int NC, O1, O2, E1, E2, items;
  int sin[3], sout[3];     // physical strides
  constexpr PG(int D_, int P_, int tx, int ty, int tz, int c0, int c1, int c2, int K_)
      : D(D_), P(P_), T{tx, ty, tz}, c{c0, c1, c2}, K(K_), np(0), A(-1), C(0), first(K_ == 0), last(false),
        ext{}, pin{}, inoff{}, pout{}, outoff(0), NC(0), O1(0), O2(0), E1(0), E2(0), items(0), sin{}, sout{} {
    np = (c[0] != 0) + (c[1] != 0) + (c[2] != 0);
    A = axis_of(K);
    C = A >= 0 ? c[A] : 0;
    last = (K == np - 1);
    for (int a = 0; a < 3; a++) ext[a] = processed(a, K) ? T[a] : win_n(c[a], T[a], P);
    for (int a = 0; a < 3; a++) {
      pin[a] = ext[a];
      inoff[a] = 0;
      pout[a] = (a == A) ? T[a] : ext[a];
    }
    outoff = 0;
    if (A >= 0) NC = C > 0 ? T[A] / P + 1 : T[A] / P;
    O1 = (A == 0) ? 1 : 0;
    O2 = (A == 2) ? 1 : 2;
    E1 = ext[O1];
    E2 = ext[O2];
    items = E1 * E2 * NC;
    sin[0] = 1;
    sin[1] = pin[0];
    sin[2] = pin[0] * pin[1];
    // the last pass's output rows get one element of padding: its items run along y
    // with lanes one row apart, and an odd pitch spreads them over the LDS banks
    sout[0] = 1;
    sout[1] = pout[0] + (last ? 1 : 0);
    sout[2] = sout[1] * pout[1];
  }
  // pass order z, y, x over the non-zero components
  constexpr int axis_of(int k) const {
    int n = 0;
    for (int a = 2; a >= 0; a--)
      if (c[a] != 0) {
        if (n == k) return a;
        n++;
      }
    return -1;
  }
  constexpr bool processed(int a, int k) const {
    for (int j = 0; j < k; j++)
      if (axis_of(j) == a) return true;
    return false;
  }
  constexpr int out_size() const { return sout[2] * pout[2]; }
};

template <int D, int Q, int P, int TX, int TY, int TZ, int AL, int K>
struct PassGeo {
  static constexpr PG g{D, P, TX, TY, TZ, Lattice<Q>::e[AL][0], Lattice<Q>::e[AL][1],
                        (D == 3) ? Lattice<Q>::e[AL][2] : 0, K};
};

// LDS sizes: largest output of pass 0 (b1), pass 1 (b2) and of a last pass (b3).
template <int D, int Q, int P, int TX, int TY, int TZ>
struct BufSizes {
  int b1 = 1, b2 = 1, b3 = 1;
  constexpr BufSizes() {
    for (int al = 1; al < Q; al++) {
      const int c2 = (D == 3) ? Lattice<Q>::e[al][2] : 0;
      for (int k = 0; k < 3; k++) {
        const PG g(D, P, TX, TY, TZ, Lattice<Q>::e[al][0], Lattice<Q>::e[al][1], c2, k);
        if (k >= g.np) break;
        const int sz = g.out_size();
        if (g.last) {
          if (sz > b3) b3 = sz;
        } else if (k == 0) {
          if (sz > b1) b1 = sz;
        } else {
          if (sz > b2) b2 = sz;
        }
      }
    }
  }
};

// Processing order: directions grouped B at a time, groups never mixing
// directions with different numbers of passes.
template <int D, int Q, int B>
struct Schedule {
  int ngroups = 0;
  int dir[32][4] = {};
  int npass[32] = {};
  constexpr Schedule() {
    for (int cls = 1; cls <= 3; cls++) {
      int slot = B;
      for (int al = 1; al < Q; al++) {
        const int np = (Lattice<Q>::e[al][0] != 0) + (Lattice<Q>::e[al][1] != 0) +
                       ((D == 3) ? (Lattice<Q>::e[al][2] != 0) : 0);
        if (np != cls) continue;
        if (slot == B) {
          ngroups++;
          slot = 0;
          npass[ngroups - 1] = cls;
        }
        dir[ngroups - 1][slot++] = al;
      }
    }
  }
};
template <int D, int Q, int B>
inline constexpr Schedule<D, Q, B> kSchedule{};

// Exact division by a constant for n < N with one 24-bit multiply and a shift.
template <unsigned DV, unsigned N>
struct FastDiv {
  static constexpr unsigned shift() {
    for (unsigned s = 8; s < 24; s++) {
      const unsigned long long m = ((1ull << s) + DV - 1) / DV;
      if (m >= (1u << 24)) break;
      bool ok = true;
      for (unsigned n = 0; n < N && ok; n++) ok = ((unsigned long long)n * m >> s) == n / DV;
      if (ok) return s;
    }
    return 0;
  }
  static constexpr unsigned S = shift();
  static constexpr unsigned M = (unsigned)(((1ull << S) + DV - 1) / DV);
  static_assert((DV & (DV - 1)) == 0 || S > 0, "no 24-bit reciprocal for this divisor");
  __device__ static __forceinline__ unsigned div(unsigned n) {
    if constexpr ((DV & (DV - 1)) == 0)
      return n / DV;
    else
      return __umul24(n, M) >> S;
  }
};

// Compile-time tag of a direction group: value = output parity, grp = group index.
template <int G_>
struct GroupTag {
  static constexpr int value = G_ & 1;
  static constexpr int grp = G_;
};

// Barrier for LDS hand-over only: outstanding global loads stay in flight.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

template <int D, int Q, int P, int TX, int TY, int TZ, int B>
struct StructCfg {
  static constexpr int NT = TX * TY * TZ;
  static constexpr BufSizes<D, Q, P, TX, TY, TZ> BS{};
  static constexpr int B1 = BS.b1, B2 = BS.b2, B3 = BS.b3;
  // register slots for a first pass
  static constexpr int first_items() {
    int m = 1;
    for (int al = 1; al < Q; al++) {
      const PG g(D, P, TX, TY, TZ, Lattice<Q>::e[al][0], Lattice<Q>::e[al][1], (D == 3) ? Lattice<Q>::e[al][2] : 0, 0);
      if (g.items > m) m = g.items;
    }
    return m;
  }
  static constexpr int KI = (first_items() + NT - 1) / NT;
};

// SYM: the host found the 1-D tables identical on all axes and mirror symmetric,
// W(-, j, m) = W(+, (p - j) mod p, p - m) (GLL nodes are symmetric), so the p x (p+1)
// weights of the positive shift serve every pass and stay in scalar registers.
template <int D, int Q, int P, typename T, int TX, int TY, int TZ, int B, int WPE, bool SYM>
__global__ __launch_bounds__(TX* TY* TZ) __attribute__((amdgpu_waves_per_eu(WPE, WPE))) void k_sl_struct(
    StructArgs<T> a, const double* __restrict__ wtab, CollideArgs<T> ca, int write_macro) {
  int bid = blockIdx.x;
#include "k_sl_struct_body.inc"
}


// Packed steps: one launch advances several independent solves of the same instantiation
// (stencil, order, storage, tile) by one step each. Solve r owns the workgroups
// [block0[r], block0[r + 1]); each of its tiles runs exactly as in its own launch, so its
// populations do not depend on the other members of the pack.
template <typename T>
struct PackArgs {  // kernel argument block: sizeof <= 4 KiB (static_assert below)
  int n;
  int block0[NAT_PACK_MAX + 1];
  int write_macro[NAT_PACK_MAX];
  StructArgs<T> a[NAT_PACK_MAX];
  CollideArgs<T> ca[NAT_PACK_MAX];
};

static_assert(sizeof(PackArgs<double>) <= 4096, "packed step arguments exceed the kernel argument block");

template <int D, int Q, int P, typename T, int TX, int TY, int TZ, int B, int WPE, bool SYM>
__global__ __launch_bounds__(TX* TY* TZ) __attribute__((amdgpu_waves_per_eu(WPE, WPE))) void k_sl_struct_packed(
    PackArgs<T> pk) {
  int r = 0;
  while (r + 1 < pk.n && pk.block0[r + 1] <= (int)blockIdx.x) r++;  // uniform: scalar loads and branches
  const StructArgs<T>& a = pk.a[r];
  const double* __restrict__ wtab = a.w;
  const CollideArgs<T>& ca = pk.ca[r];
  const int write_macro = pk.write_macro[r];
  int bid = (int)blockIdx.x - pk.block0[r];
#include "k_sl_struct_body.inc"
}

// How a dispatch ends: report the tile, launch one solve, or launch a packed step.
struct StructMode {
  int* query = nullptr;        // != null: report the tile (TX, TY, TZ) and whether it is the small one
  const void* pack = nullptr;  // != null: host PackArgs<T> of a packed step (passed by value)
  int pack_blocks = 0;
  hipStream_t pack_stream = nullptr;
};

template <int D, int Q, int P, typename T, int TX, int TY, int TZ, bool SYM_ONLY = false>
static int launch_struct(nat_ctx* ctx, StructArgs<T> a, const CollideArgs<T>& ca, int write_macro,
                         const StructMode& md, int small = 0) {
  // FP64 at p = 6: groups of 2 directions (3 hold 3 x 7 FP64 first-pass values per item
  // slot and spill at the 128-register budget)
  constexpr int B = sizeof(T) == 4 ? NAT_STRUCT_B32 : (P >= 6 ? 2 : NAT_STRUCT_B64);
  constexpr int WPE = 4;  // 16 waves per CU: one 1024-thread tile, up to 128 VGPRs
  if (md.query) {
    md.query[0] = TX;
    md.query[1] = TY;
    md.query[2] = TZ;
    md.query[3] = small;
    md.query[4] = (SYM_ONLY || ctx->ax_sym) ? 1 : 0;
    return 0;
  }
  if (md.pack) {
    // instantiated where the packed form keeps the one-solve kernel's registers (no SGPR or
    // VGPR spills): symmetric tables, p = 2-4; pack_key admits only these
    if constexpr (P >= 2 && P <= 4) {
      if (SYM_ONLY || ctx->ax_sym) {
        const PackArgs<T>& pk = *(const PackArgs<T>*)md.pack;
        hipLaunchKernelGGL((k_sl_struct_packed<D, Q, P, T, TX, TY, TZ, B, WPE, true>), dim3(md.pack_blocks),
                           dim3(TX * TY * TZ), 0, md.pack_stream, pk);
        NAT_HIP(hipGetLastError());
        return 0;
      }
    }
    ctx->err = "packed structured step: not built for this (order, tables)";
    return NAT_ERR_UNSUPPORTED;
  }
  a.tiles[0] = (a.n[0] + TX - 1) / TX;
  a.tiles[1] = (a.n[1] + TY - 1) / TY;
  a.tiles[2] = D == 3 ? (a.n[2] + TZ - 1) / TZ : 1;
  const int nb = a.tiles[0] * a.tiles[1] * a.tiles[2];
  a.xcd_remap = (nb % 8 == 0) ? 1 : 0;
  if (SYM_ONLY || ctx->ax_sym)
    hipLaunchKernelGGL((k_sl_struct<D, Q, P, T, TX, TY, TZ, B, WPE, true>), dim3(nb), dim3(TX * TY * TZ), 0,
                       ctx->stream, a, a.w, ca, write_macro);
  else if constexpr (!SYM_ONLY)
    hipLaunchKernelGGL((k_sl_struct<D, Q, P, T, TX, TY, TZ, B, WPE, false>), dim3(nb), dim3(TX * TY * TZ), 0,
                       ctx->stream, a, a.w, ca, write_macro);
  NAT_HIP(hipGetLastError());
  return 0;
}

// Small meshes: when the tiles below would give fewer workgroups than the GPU has CUs, a
// tile of ~256 nodes spreads the step over more CUs (a 16^3-cell p = 2 box has 32 tiles of
// 16x8x8 but 128 of 8x8x4): 8^3-cell boxes at p = 2-4 step 20-40% faster. Instantiated for
// the periodic boxes whose 1-D tables are symmetric (all uniform boxes) at p = 2, 3, 4; at
// p = 6 a small tile is one cell plus its halo and measured slower. The autotuner
// (nat_autotune) may fix the choice per mesh instead (ctx->tile_pref).
template <int D, int Q, int P, typename T, int TX, int TY, int TZ, int SX, int SY, int SZ>
static int launch_struct_sized(nat_ctx* ctx, const StructArgs<T>& a, const CollideArgs<T>& ca, int wm,
                               const StructMode& md) {
  const int64_t nb = (int64_t)((a.n[0] + TX - 1) / TX) * ((a.n[1] + TY - 1) / TY) * (D == 3 ? (a.n[2] + TZ - 1) / TZ : 1);
  const bool small = ctx->ax_sym && (ctx->tile_pref >= 0 ? ctx->tile_pref == 1 : nb < 256);
  if (small) return launch_struct<D, Q, P, T, SX, SY, SZ, true>(ctx, a, ca, wm, md, 1);
  return launch_struct<D, Q, P, T, TX, TY, TZ>(ctx, a, ca, wm, md, 0);
}

template <typename T>
static int dispatch_struct(nat_ctx* ctx, const StructArgs<T>& a, const CollideArgs<T>& ca, int wm, const StructMode& md) {
  const int Q = ctx->Q, P = ctx->p;
#define NAT_S3(QQ, PP, TX, TY, TZ) \
  if (Q == QQ && P == PP) return launch_struct<3, QQ, PP, T, TX, TY, TZ>(ctx, a, ca, wm, md);
#define NAT_S3S(QQ, PP, TX, TY, TZ, SX, SY, SZ) \
  if (Q == QQ && P == PP) return launch_struct_sized<3, QQ, PP, T, TX, TY, TZ, SX, SY, SZ>(ctx, a, ca, wm, md);
#define NAT_S3X(...) NAT_S3S(__VA_ARGS__)
#define NAT_S2(QQ, PP, TX, TY) \
  if (Q == QQ && P == PP) return launch_struct<2, QQ, PP, T, TX, TY, 1>(ctx, a, ca, wm, md);
#define NAT_S2S(QQ, PP, TX, TY, SX, SY) \
  if (Q == QQ && P == PP) return launch_struct_sized<2, QQ, PP, T, TX, TY, 1, SX, SY, 1>(ctx, a, ca, wm, md);
  if (ctx->D == 3) {
    NAT_S3(27, 1, 16, 8, 8)
    NAT_S3S(27, 2, 16, 8, 8, 8, 8, 4)
    NAT_S3S(27, 3, 12, 12, 6, 6, 6, 6)
    NAT_S3X(27, 4, NAT_TILE_27_4, 8, 8, 4)
    NAT_S3(27, 6, 12, 12, 6)
    NAT_S3(19, 1, 16, 8, 8)
    NAT_S3S(19, 2, 16, 8, 8, 8, 8, 4)
    NAT_S3S(19, 3, 12, 12, 6, 6, 6, 6)
    NAT_S3S(19, 4, 16, 8, 8, 8, 8, 4)
    NAT_S3(19, 6, 12, 12, 6)
  } else {
    NAT_S2(9, 1, 32, 16)
    NAT_S2S(9, 2, 32, 16, 16, 16)
    NAT_S2S(9, 3, 24, 24, 12, 12)
    NAT_S2S(9, 4, 32, 16, 16, 16)
    NAT_S2(9, 6, 24, 24)
  }
#undef NAT_S3
#undef NAT_S3S
#undef NAT_S3X
#undef NAT_S2
#undef NAT_S2S
  ctx->err = "structured SL kernel not instantiated for this (stencil, order)";
  return NAT_ERR_UNSUPPORTED;
}

// kernel arguments of one step of ctx reading buffer src
template <typename T>
static int make_args(nat_ctx* ctx, int src, StructArgs<T>& a, CollideArgs<T>& ca) {
  if (!ctx->ax_w || !ctx->ax_off) {
    ctx->err = "structured backend: no 1-D operator tables (the last assembly failed)";
    return NAT_ERR_STATE;
  }
  a = StructArgs<T>{};
  a.fin = (const T*)ctx->f[src];
  a.fout = (T*)ctx->f[1 - src];
  a.stride = ctx->stride;
  for (int i = 0; i < 3; i++) a.n[i] = ctx->nu[i];
  a.w = (const double*)ctx->ax_w;
  if ((int64_t)ctx->stride * (int64_t)sizeof(T) > INT32_MAX) {
    ctx->err = "structured backend: one direction's array exceeds 2 GiB";
    return NAT_ERR_UNSUPPORTED;
  }
  a.dir_bytes = (int)(ctx->stride * (int64_t)sizeof(T));
  ca = CollideArgs<T>{};
  ca.omega = (T)(1.0 / ctx->tau);
  ca.omega64 = 1.0 / ctx->tau;
  ca.rho0 = (T)(ctx->fp32 ? (ctx->desc.rho0 > 0 ? ctx->desc.rho0 : 1.0) : 0.0);
  ca.scaling = ctx->st.s;
  ca.rho = (T*)ctx->rho;
  ca.u = (T*)ctx->u;
  ca.mstride = ctx->n_store;
  ca.status = ctx->d_status;
  ca.stab = ctx->stab_on();
  return 0;
}

template <typename T>
static int launch_typed(nat_ctx* ctx, int write_macro, int* query = nullptr) {
  StructArgs<T> a;
  CollideArgs<T> ca;
  if (int rc = make_args<T>(ctx, ctx->cur, a, ca)) return rc;
  StructMode md;
  md.query = query;
  const int rc = dispatch_struct<T>(ctx, a, ca, write_macro, md);
  if (rc == 0 && !query) ctx->cur = 1 - ctx->cur;
  return rc;
}

// ---- packed steps

// Packing key of a structured context: members of one pack share the kernel
// instantiation (dimension, stencil, order, storage, tile, weight symmetry).
static int pack_key(nat_ctx* ctx, int64_t* key) {
  if (ctx->desc.scheme != NAT_SL || ctx->desc.compressible || ctx->backend != NAT_BACKEND_STRUCTURED ||
      !ctx->assembled || !ctx->ax_sym || ctx->p < 2 || ctx->p > 4)
    return NAT_ERR_ARG;
  int q[5] = {0, 0, 0, 0, 0};
  const int rc = ctx->fp32 ? launch_typed<float>(ctx, 0, q) : launch_typed<double>(ctx, 0, q);
  if (rc) return rc;
  *key = ((((((int64_t)ctx->D * 64 + ctx->Q) * 16 + ctx->p) * 2 + (ctx->fp32 ? 1 : 0)) * 64 + q[0]) * 64 + q[1]) * 64 +
         q[2];
  *key = *key * 4 + q[4] * 2 + q[3];
  return 0;
}

template <typename T>
static int pack_step_typed(nat_ctx* const* ctxs, int n, int nsteps, hipStream_t stream) {
  nat_ctx* ctx = ctxs[0];  // errors of the pack as a whole are reported on the first member
  int q[5] = {0, 0, 0, 0, 0};
  if (int rc = launch_typed<T>(ctx, 0, q)) return rc;
  const int TX = q[0], TY = q[1], TZ = q[2];
  PackArgs<T> tab[3];  // [0] even steps, [1] odd steps, [2] the last step (kernel arguments)
  for (int t = 0; t < 3; t++) {
    tab[t] = PackArgs<T>{};
    tab[t].n = n;
  }
  int blocks = 0;
  for (int i = 0; i < n; i++) {
    nat_ctx* c = ctxs[i];
    for (int t = 0; t < 3; t++) {
      const int par = (t == 2) ? ((nsteps - 1) & 1) : t;
      StructArgs<T> a;
      CollideArgs<T> ca;
      if (int rc = make_args<T>(c, c->cur ^ par, a, ca)) {
        if (c != ctx) ctx->err = c->err;
        return rc;
      }
      a.tiles[0] = (a.n[0] + TX - 1) / TX;
      a.tiles[1] = (a.n[1] + TY - 1) / TY;
      a.tiles[2] = c->D == 3 ? (a.n[2] + TZ - 1) / TZ : 1;
      const int nb = a.tiles[0] * a.tiles[1] * a.tiles[2];
      a.xcd_remap = (nb % 8 == 0) ? 1 : 0;
      tab[t].a[i] = a;
      tab[t].ca[i] = ca;
      tab[t].write_macro[i] = (t == 2) ? STEP_MACRO : 0;
      tab[t].block0[i] = blocks;
      if (t == 2) blocks += nb;
    }
  }
  for (int t = 0; t < 3; t++) tab[t].block0[n] = blocks;
  StructMode md;
  md.pack_blocks = blocks;
  md.pack_stream = stream;
  StructArgs<T> a0;
  CollideArgs<T> c0;
  if (int rc = make_args<T>(ctx, ctx->cur, a0, c0)) return rc;
  for (int s = 0; s < nsteps; s++) {
    md.pack = &tab[s == nsteps - 1 ? 2 : (s & 1)];
    if (int rc = dispatch_struct<T>(ctx, a0, c0, 0, md)) return rc;
  }
  for (int i = 0; i < n; i++) {
    ctxs[i]->cur ^= (nsteps & 1);
    for (int st = 0; st < nsteps; st++) ctxs[i]->time += ctxs[i]->dt;  // as step by step
  }
  return 0;
}

}  // namespace nat

int nat_launch_sl_struct(nat_ctx* ctx, int write_macro) {
  return ctx->fp32 ? nat::launch_typed<float>(ctx, write_macro) : nat::launch_typed<double>(ctx, write_macro);
}

int nat_launch_sl_struct_tile(nat_ctx* ctx, int* tile) {
  int q[5] = {0, 0, 0, 0, 0};
  const int rc = ctx->fp32 ? nat::launch_typed<float>(ctx, 0, q) : nat::launch_typed<double>(ctx, 0, q);
  for (int i = 0; i < 3; i++) tile[i] = q[i];
  return rc;
}

int nat_internal_pack_key(nat_ctx* ctx, int64_t* key) { return nat::pack_key(ctx, key); }

// nsteps fused steps of every context in ctxs (1 <= n <= NAT_PACK_MAX, equal pack keys, each
// context idle) as packed launches on stream; the per-solve arguments travel as kernel
// arguments (at most NAT_PACK_MAX solves fit the 4 KiB argument block).
int nat_internal_pack_step(nat_ctx* const* ctxs, int n, int nsteps, hipStream_t stream) {
  if (n < 1 || n > NAT_PACK_MAX || nsteps < 1) return NAT_ERR_ARG;
  int64_t k0 = 0;
  if (int rc = nat::pack_key(ctxs[0], &k0)) return rc;
  for (int i = 1; i < n; i++) {
    int64_t k = 0;
    if (nat::pack_key(ctxs[i], &k) || k != k0) {
      ctxs[0]->err = "packed step: members of different kernel instantiations";
      return NAT_ERR_ARG;
    }
  }
  return ctxs[0]->fp32 ? nat::pack_step_typed<float>(ctxs, n, nsteps, stream)
                       : nat::pack_step_typed<double>(ctxs, n, nsteps, stream);
}

// ---- autotuner

namespace {

struct TuneKey {
  int D, Q, p, fp32, sym, n0, n1, n2;
  bool operator<(const TuneKey& o) const {
    const int a[8] = {D, Q, p, fp32, sym, n0, n1, n2}, b[8] = {o.D, o.Q, o.p, o.fp32, o.sym, o.n0, o.n1, o.n2};
    for (int i = 0; i < 8; i++)
      if (a[i] != b[i]) return a[i] < b[i];
    return false;
  }
};
std::mutex g_tune_mutex;
std::map<TuneKey, int> g_tune_cache;  // winning tile preference per (instantiation, mesh)

// device time of nsteps launches of the step kernel with the context's current tile choice;
// each reads the current buffer and writes the other one without advancing the state
template <typename T>
int time_tile(nat_ctx* ctx, int nsteps, hipEvent_t e0, hipEvent_t e1, float* ms) {
  nat::StructArgs<T> a;
  nat::CollideArgs<T> ca;
  if (int rc = nat::make_args<T>(ctx, ctx->cur, a, ca)) return rc;
  ca.rho = nullptr;
  ca.u = nullptr;
  nat::StructMode md;
  for (int s = 0; s < 2; s++)
    if (int rc = nat::dispatch_struct<T>(ctx, a, ca, 0, md)) return rc;
  NAT_HIP(hipEventRecord(e0, ctx->stream));
  for (int s = 0; s < nsteps; s++)
    if (int rc = nat::dispatch_struct<T>(ctx, a, ca, 0, md)) return rc;
  NAT_HIP(hipEventRecord(e1, ctx->stream));
  NAT_HIP(hipEventSynchronize(e1));
  NAT_HIP(hipEventElapsedTime(ms, e0, e1));
  return 0;
}

}  // namespace

// Tile shape of the structured step kernel picked by measurement: every tile instantiated for
// the context's (stencil, order) is timed for nsteps steps on the context's own mesh (the
// launches write the idle ping-pong buffer; populations, fields and time stay as they were),
// and the fastest is kept for this context and, process-wide, for every later context of the
// same instantiation and mesh size, which then skips the timing. Results do not depend on the
// tile: each node's interpolation and collision run the same arithmetic in any tile.
int nat_internal_autotune(nat_ctx* ctx, int nsteps, int* tile, double* best_us) {
  if (ctx->desc.scheme != NAT_SL || ctx->desc.compressible || ctx->backend != NAT_BACKEND_STRUCTURED) {
    ctx->err = "nat_autotune: the autotuner serves the structured and general SL backends";
    return NAT_ERR_ARG;
  }
  if (!ctx->assembled) return NAT_ERR_STATE;
  const TuneKey key{ctx->D, ctx->Q, ctx->p, ctx->fp32 ? 1 : 0, ctx->ax_sym ? 1 : 0, ctx->nu[0], ctx->nu[1], ctx->nu[2]};
  if (best_us) *best_us = 0;
  {
    std::lock_guard<std::mutex> lk(g_tune_mutex);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      ctx->tile_pref = it->second;
      return nat_launch_sl_struct_tile(ctx, tile);
    }
  }
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
    ctx->tile_pref = pref;
    int q[5] = {0, 0, 0, 0, 0};
    rc = ctx->fp32 ? nat::launch_typed<float>(ctx, 0, q) : nat::launch_typed<double>(ctx, 0, q);
    if (rc) break;
    if (pref == 1 && q[3] == 0) break;  // no small tile for this instantiation
    float ms = 0;
    rc = ctx->fp32 ? time_tile<float>(ctx, nsteps, e0, e1, &ms) : time_tile<double>(ctx, nsteps, e0, e1, &ms);
    if (!rc && (best < 0 || ms < best_ms)) {
      best = pref;
      best_ms = ms;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) {
    ctx->tile_pref = -1;
    return rc;
  }
  ctx->tile_pref = best;
  if (best_us) *best_us = 1e3 * best_ms / nsteps;
  {
    std::lock_guard<std::mutex> lk(g_tune_mutex);
    g_tune_cache[key] = best;
  }
  return nat_launch_sl_struct_tile(ctx, tile);
}
