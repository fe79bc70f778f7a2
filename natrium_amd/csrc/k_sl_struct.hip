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
  using Cfg = StructCfg<D, Q, P, TX, TY, TZ, B>;
  static_assert(TX % P == 0 && TY % P == 0 && (D == 2 || TZ % P == 0), "tiles must be cell aligned");
  constexpr int NT = Cfg::NT, KI = Cfg::KI;
  constexpr int TW = (TX > TY ? (TX > TZ ? TX : TZ) : (TY > TZ ? TY : TZ)) + P + 1;
  __shared__ double b1[B][Cfg::B1];
  __shared__ double b2[B][Cfg::B2];
  __shared__ T b3[2][B][Cfg::B3];
  __shared__ int gt[3][TW];  // byte offsets of the wrapped window nodes [X0-P, X0+T] per axis

  const unsigned tid = threadIdx.x;
  // tile index, optionally regrouped so that consecutive tiles (which share halos)
  // run on the same XCD: hardware dispatches workgroup i to XCD i % 8
  int bid = blockIdx.x;
  const int nb = a.tiles[0] * a.tiles[1] * a.tiles[2];
  if (a.xcd_remap) bid = (bid % 8) * (nb / 8) + bid / 8;
  const int org[3] = {(bid % a.tiles[0]) * TX, ((bid / a.tiles[0]) % a.tiles[1]) * TY,
                      (bid / (a.tiles[0] * a.tiles[1])) * TZ};
  const int nx = a.n[0], ny = a.n[1], nz = a.n[2];
  const unsigned ox0 = tid % TX, oy0 = (tid / TX) % TY, oz0 = tid / (TX * TY);
  const bool active = (org[0] + (int)ox0 < nx) && (org[1] + (int)oy0 < ny) && (org[2] + (int)oz0 < nz);
  const int node = (org[0] + ox0) + nx * ((org[1] + oy0) + ny * (org[2] + oz0));
  T fs[Q];
  fs[0] = active ? a.fin[node] : T(0);
  const int astr[3] = {1, nx, nx * ny};
  for (int i = tid; i < 3 * TW; i += NT) {
    const int ax = i / TW, k = i % TW;
    const int n = a.n[ax];
    gt[ax][k] = (ax < D) ? (((org[ax] - P + k) % n + n) % n) * astr[ax] * (int)sizeof(T) : 0;
  }

  double W0[SYM ? P : 1][P + 1];
  if constexpr (SYM) {
#pragma unroll
    for (int j = 0; j < P; j++)
#pragma unroll
      for (int m = 0; m <= P; m++) W0[j][m] = wtab[j * (MAXP + 1) + m];
  }

  // first-pass inputs of the next PF groups (register slot sets, rotating by group)
  constexpr int PF = (sizeof(T) == 4) ? NAT_PREFETCH32 : NAT_PREFETCH64;
  T xin[PF][B][KI][P + 1];
  int xo0[PF][B][KI];         // output base of the first-pass item, computed with its loads
  unsigned xsc[PF][B][KI];    // its source cell along the pass axis
  // the offset tables are LDS: wait for LDS only, so the rest population's global load
  // stays in flight (a full __syncthreads would wait for it here)
  lds_barrier();

  // raw buffer loads: 32-bit byte offsets against a per-direction resource (range checked)
  auto load_t = [](__amdgpu_buffer_rsrc_t r, int off) -> T {
    if constexpr (sizeof(T) == 4)
      return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
    else
      return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
  };

  // The directions of a group start their items at different waves, so that a
  // pass with fewer items than threads does not leave the same waves idle for all
  // directions of the group.
  auto rot = [](unsigned t, int b) -> unsigned {
    const unsigned off = ((unsigned)(b * NT / B) / 64u) * 64u;
    const unsigned r = t + off;
    return r >= (unsigned)NT ? r - NT : r;
  };

  // ---- global loads of a direction's first pass into register slot b
  auto load_first = [&](auto ALc, auto Bc, auto SETc) {
    constexpr int al = decltype(ALc)::value, b = decltype(Bc)::value, ps = decltype(SETc)::value;
    if constexpr (al > 0) {
      constexpr PG G = PassGeo<D, Q, P, TX, TY, TZ, al, 0>::g;
      constexpr int KK = (G.items + NT - 1) / NT;
      static_assert(KK <= KI, "first-pass register slots");
      // the base is uniform; readfirstlane lets the compiler keep the resource in SGPRs
      // (readfirstlane returns int: both halves go through uint32_t, no sign extension)
      const uint64_t pb = (uint64_t)(a.fin + (int64_t)al * a.stride);
      const uint32_t phi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(pb >> 32));
      const uint32_t plo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)pb);
      const uint64_t pu = ((uint64_t)phi << 32) | (uint64_t)plo;
      const __amdgpu_buffer_rsrc_t src = __builtin_amdgcn_make_buffer_rsrc(
          (void*)pu, (short)0, __builtin_amdgcn_readfirstlane(a.dir_bytes), 0x00020000);
      const unsigned t0 = rot(tid, b);
#pragma unroll
      for (int k = 0; k < KK; k++) {
        const unsigned it = t0 + k * NT;
        if (it < (unsigned)G.items) {
          const unsigned r = FastDiv<G.E1, KK * NT>::div(it);
          const unsigned i1 = it - r * G.E1;
          const unsigned sc = FastDiv<G.E2, (KK * NT) / G.E1 + 1>::div(r);
          const unsigned i2 = r - sc * G.E2;
          const int base = gt[G.O1][i1 + win_lo(G.c[G.O1], P) + P] + gt[G.O2][i2 + win_lo(G.c[G.O2], P) + P];
          const int* ga = &gt[G.A][sc * P + win_lo(G.C, P) + P];
#pragma unroll
          for (int m = 0; m <= P; m++) xin[ps][b][k][m] = load_t(src, base + ga[m]);
          constexpr int sh = (G.C > 0) ? 1 - P : 0;
          xo0[ps][b][k] = (int)(i1 * G.sout[G.O1] + i2 * G.sout[G.O2]) + ((int)sc * P + sh) * G.sout[G.A];
          xsc[ps][b][k] = sc;
        }
      }
    }
  };

  // ---- one 1-D pass of direction al (slot b, output parity par)
  auto run_pass = [&](auto ALc, auto Kc, auto Bc, auto PARc) {
#pragma clang fp contract(off)  // the one fma below is explicit: all instantiations round alike
    constexpr int al = decltype(ALc)::value, kp = decltype(Kc)::value;
    constexpr int b = decltype(Bc)::value, par = decltype(PARc)::value;
    constexpr int ps = decltype(PARc)::grp % PF;  // register slot set of this group
    constexpr PG G = PassGeo<D, Q, P, TX, TY, TZ, al, kp>::g;
    constexpr int KK = (G.items + NT - 1) / NT;
    // this pass's weight rows [class][m], uniform addresses: scalar loads
    const double* __restrict__ wt = wtab + (G.A * MAXSHIFT + (G.C > 0 ? 0 : 1)) * MAXP * (MAXP + 1);
    const unsigned t0 = rot(tid, b);
#pragma unroll
    for (int k = 0; k < KK; k++) {
      const unsigned it = t0 + k * NT;
      if (it < (unsigned)G.items) {
        double v[P + 1];
        int o0;
        unsigned sc;
        constexpr int sh = (G.C > 0) ? 1 - P : 0;
        if constexpr (G.first) {
#pragma unroll
          for (int m = 0; m <= P; m++) v[m] = (double)xin[ps][b][k][m];
          o0 = xo0[ps][b][k];
          sc = xsc[ps][b][k];
        } else {
          const unsigned r = FastDiv<G.E1, KK * NT>::div(it);
          const unsigned i1 = it - r * G.E1;
          sc = FastDiv<G.E2, (KK * NT) / G.E1 + 1>::div(r);
          const unsigned i2 = r - sc * G.E2;
          const double* in = (kp == 1) ? b1[b] : b2[b];
          const unsigned i0 = (i1 + G.inoff[G.O1]) * G.sin[G.O1] + (i2 + G.inoff[G.O2]) * G.sin[G.O2] +
                              sc * (P * G.sin[G.A]);
#pragma unroll
          for (int m = 0; m <= P; m++) v[m] = in[i0 + m * G.sin[G.A]];
          o0 = (int)(i1 * G.sout[G.O1] + i2 * G.sout[G.O2]) + ((int)sc * P + sh) * G.sout[G.A];
        }
        // the P accumulations are independent chains: interleave them (m outer)
        double acc[P];
#pragma unroll
        for (int m = 0; m <= P; m++) {
#pragma unroll
          for (int kk = 0; kk < P; kk++) {
            double w;
            if constexpr (SYM)
              w = (G.C > 0) ? W0[(kk + 1) % P][m] : W0[(P - kk) % P][P - m];
            else
              w = wt[((kk + (G.C > 0 ? 1 : 0)) % P) * (MAXP + 1) + m];
            acc[kk] = (m == 0) ? w * v[0] : fma(w, v[m], acc[kk]);
          }
        }
        // output kk is the node sc*P + kk (+1-P for a positive component) of the tile
        const bool lo_ok = G.C < 0 || sc != 0;                      // outputs kk < P-1
        const bool hi_ok = G.C < 0 || sc != (unsigned)(G.NC - 1);   // output kk = P-1
#pragma unroll
        for (int kk = 0; kk < P; kk++) {
          if (kk < P - 1 ? lo_ok : hi_ok) {
            if constexpr (G.last)
              b3[par][b][o0 + kk * G.sout[G.A]] = (T)acc[kk];
            else
              ((kp == 0) ? b1[b] : b2[b])[o0 + kk * G.sout[G.A]] = acc[kk];
          }
        }
      }
    }
  };

  constexpr int NG = kSchedule<D, Q, B>.ngroups;
  static_for<0, (PF < NG ? PF : NG)>([&](auto Gc) {
    constexpr int g = decltype(Gc)::value;
    static_for<0, B>([&](auto Bc) {
      constexpr int b = decltype(Bc)::value;
      load_first(std::integral_constant<int, kSchedule<D, Q, B>.dir[g][b]>{}, Bc, std::integral_constant<int, g % PF>{});
    });
  });

  static_for<0, NG>([&](auto Gc) {
    constexpr int g = decltype(Gc)::value;
    constexpr int np = kSchedule<D, Q, B>.npass[g];
    using PAR = GroupTag<g>;
    // first passes (register inputs)
    static_for<0, B>([&](auto Bc) {
      constexpr int al = kSchedule<D, Q, B>.dir[g][decltype(Bc)::value];
      if constexpr (al > 0) run_pass(std::integral_constant<int, al>{}, std::integral_constant<int, 0>{}, Bc, PAR{});
    });
    // the inputs of group g + PF go into the slot set just consumed; they are in flight
    // during the remaining passes of this group and all of the next PF - 1 groups
    if constexpr (g + PF < NG) {
      static_for<0, B>([&](auto Bc) {
        constexpr int b = decltype(Bc)::value;
        load_first(std::integral_constant<int, kSchedule<D, Q, B>.dir[g + PF][b]>{}, Bc,
                   std::integral_constant<int, g % PF>{});
      });
    }
    lds_barrier();
    static_for<1, np>([&](auto Kc) {
      static_for<0, B>([&](auto Bc) {
        constexpr int al = kSchedule<D, Q, B>.dir[g][decltype(Bc)::value];
        if constexpr (al > 0) run_pass(std::integral_constant<int, al>{}, Kc, Bc, PAR{});
      });
      lds_barrier();
    });
    // pick up this group's values for the own node
    static_for<0, B>([&](auto Bc) {
      constexpr int b = decltype(Bc)::value;
      constexpr int al = kSchedule<D, Q, B>.dir[g][b];
      if constexpr (al > 0) {
        constexpr PG G = PassGeo<D, Q, P, TX, TY, TZ, al, kSchedule<D, Q, B>.npass[g] - 1>::g;
        constexpr unsigned off = (G.A == 0 ? G.outoff : 0) + G.sout[1] * (G.A == 1 ? G.outoff : 0) +
                                 G.sout[2] * (G.A == 2 ? G.outoff : 0);
        T val = b3[g & 1][b][off + ox0 + G.sout[1] * oy0 + G.sout[2] * oz0];
        asm volatile("" : "+v"(val));
        fs[al] = val;
      }
    });
  });

  if (!active) return;
  if (write_macro & STEP_STREAM_ONLY) {  // the SL stream alone (CFDSolver::stream)
#pragma unroll
    for (int q = 0; q < Q; q++) a.fout[(int64_t)q * a.stride + node] = fs[q];
    return;
  }
  bgk_collide_store_pairs<Q, T>(fs, ca, node, (write_macro & STEP_MACRO) != 0, a.fout, a.stride);
}

// query != null: report the tile (TX, TY, TZ) instead of launching
template <int D, int Q, int P, typename T, int TX, int TY, int TZ, bool SYM_ONLY = false>
static int launch_struct(nat_ctx* ctx, StructArgs<T> a, const CollideArgs<T>& ca, int write_macro,
                         int* query) {
  if (query) {
    query[0] = TX;
    query[1] = TY;
    query[2] = TZ;
    return 0;
  }
  a.tiles[0] = (a.n[0] + TX - 1) / TX;
  a.tiles[1] = (a.n[1] + TY - 1) / TY;
  a.tiles[2] = D == 3 ? (a.n[2] + TZ - 1) / TZ : 1;
  const int nb = a.tiles[0] * a.tiles[1] * a.tiles[2];
  a.xcd_remap = (nb % 8 == 0) ? 1 : 0;
  // FP64 at p = 6: groups of 2 directions (3 hold 3 x 7 FP64 first-pass values per item
  // slot and spill at the 128-register budget)
  constexpr int B = sizeof(T) == 4 ? NAT_STRUCT_B32 : (P >= 6 ? 2 : NAT_STRUCT_B64);
  constexpr int WPE = 4;  // 16 waves per CU: one 1024-thread tile, up to 128 VGPRs
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
// p = 6 a small tile is one cell plus its halo and measured slower.
template <int D, int Q, int P, typename T, int TX, int TY, int TZ, int SX, int SY, int SZ>
static int launch_struct_sized(nat_ctx* ctx, const StructArgs<T>& a, const CollideArgs<T>& ca, int wm, int* query) {
  const int64_t nb = (int64_t)((a.n[0] + TX - 1) / TX) * ((a.n[1] + TY - 1) / TY) * (D == 3 ? (a.n[2] + TZ - 1) / TZ : 1);
  if (ctx->ax_sym && nb < 256) return launch_struct<D, Q, P, T, SX, SY, SZ, true>(ctx, a, ca, wm, query);
  return launch_struct<D, Q, P, T, TX, TY, TZ>(ctx, a, ca, wm, query);
}

template <typename T>
static int dispatch_struct(nat_ctx* ctx, const StructArgs<T>& a, const CollideArgs<T>& ca, int wm, int* query) {
  const int Q = ctx->Q, P = ctx->p;
#define NAT_S3(QQ, PP, TX, TY, TZ) \
  if (Q == QQ && P == PP) return launch_struct<3, QQ, PP, T, TX, TY, TZ>(ctx, a, ca, wm, query);
#define NAT_S3S(QQ, PP, TX, TY, TZ, SX, SY, SZ) \
  if (Q == QQ && P == PP) return launch_struct_sized<3, QQ, PP, T, TX, TY, TZ, SX, SY, SZ>(ctx, a, ca, wm, query);
#define NAT_S3X(...) NAT_S3S(__VA_ARGS__)
#define NAT_S2(QQ, PP, TX, TY) \
  if (Q == QQ && P == PP) return launch_struct<2, QQ, PP, T, TX, TY, 1>(ctx, a, ca, wm, query);
#define NAT_S2S(QQ, PP, TX, TY, SX, SY) \
  if (Q == QQ && P == PP) return launch_struct_sized<2, QQ, PP, T, TX, TY, 1, SX, SY, 1>(ctx, a, ca, wm, query);
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

template <typename T>
static int launch_typed(nat_ctx* ctx, int write_macro, int* query = nullptr) {
  if (!ctx->ax_w || !ctx->ax_off) {
    ctx->err = "structured backend: no 1-D operator tables (the last assembly failed)";
    return NAT_ERR_STATE;
  }
  const int src = ctx->cur, dst = 1 - ctx->cur;
  StructArgs<T> a;
  a.fin = (const T*)ctx->f[src];
  a.fout = (T*)ctx->f[dst];
  a.stride = ctx->stride;
  for (int i = 0; i < 3; i++) a.n[i] = ctx->nu[i];
  a.w = (const double*)ctx->ax_w;
  if ((int64_t)ctx->stride * (int64_t)sizeof(T) > INT32_MAX) {
    ctx->err = "structured backend: one direction's array exceeds 2 GiB";
    return NAT_ERR_UNSUPPORTED;
  }
  a.dir_bytes = (int)(ctx->stride * (int64_t)sizeof(T));
  CollideArgs<T> ca;
  ca.omega = (T)(1.0 / ctx->tau);
  ca.omega64 = 1.0 / ctx->tau;
  ca.rho0 = (T)(ctx->fp32 ? (ctx->desc.rho0 > 0 ? ctx->desc.rho0 : 1.0) : 0.0);
  ca.scaling = ctx->st.s;
  ca.rho = (T*)ctx->rho;
  ca.u = (T*)ctx->u;
  ca.mstride = ctx->n_store;
  ca.status = ctx->d_status;
  ca.stab = ctx->stab_on();
  int rc = dispatch_struct<T>(ctx, a, ca, write_macro, query);
  if (rc == 0 && !query) ctx->cur = dst;
  return rc;
}

}  // namespace nat

int nat_launch_sl_struct(nat_ctx* ctx, int write_macro) {
  return ctx->fp32 ? nat::launch_typed<float>(ctx, write_macro) : nat::launch_typed<double>(ctx, write_macro);
}

int nat_launch_sl_struct_tile(nat_ctx* ctx, int* tile) {
  return ctx->fp32 ? nat::launch_typed<float>(ctx, 0, tile) : nat::launch_typed<double>(ctx, 0, tile);
}
