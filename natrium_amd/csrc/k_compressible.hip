// Compressible populations (f, g) with the quartic Hermite equilibrium and FP64
// temperature: SL streaming of both sets with one set of weights, fused with the
// compressible BGK collision.
//
// Collision semantics (per node, FP64), following
// collision_advanced/CollisionOperator.h:113-224, CollisionSchemes.h:43-118 and
// AuxiliaryCollisionFunctions.h:290-307,421-431:
//   rho = sum f, u = sum e f / rho (unscaled e)
//   T   = 1/2 sum_i [ |e_i - u|^2 f_i / cs2 + g_i ] / (rho C_v),  C_v = 1/(gamma-1)
//   f_eq = quartic Hermite polynomial (Equilibria.h:120-267), g_eq = f_eq T (2 C_v - D)
//   tau_v = (tau - 1/2) S(T) / (T rho) + 1/2, S = (T/0.85)^0.7 with Sutherland's law
//   optional Prandtl number: f* = w Q:(e e e - 3 cs2 e I)/(6 cs2^3), g* = w F.e / T
//   f -= omega_v (f - f_eq) - (omega_v - omega_p) f*;  g likewise with g_eq, g*
//   Knudsen estimate sum |f - f_eq| / w / Q written as the shock-sensor field.
// The per-direction constants (unscaled e, w, |e|^2, the independent H3 and H4
// components) come from a small FP64 table read with uniform addresses.
//
// FP32 storage keeps f~ = f - w rho0 and g~ = g - w rho0 (2 C_v - D), the
// equilibria at rest with T = 1; the collision un-shifts in FP64.
#include "nat_collide.hpp"
#include "nat_fastdiv.hpp"

namespace nat {

constexpr int CTAB = 16;  // doubles per direction in the constant table
enum { CT_E0 = 0, CT_E1 = 1, CT_W = 2, CT_EE = 3, CT_H3 = 4, CT_H4 = 8, CT_IW = 13 };

template <typename T>
struct CompArgs {
  const T* fin;
  const T* gin;
  T* fout;
  T* gout;
  int64_t stride;
  int64_t n;
  int nfe[3];
  int N[3];
  FastDivU divN0;  // division by N[0]
  const int* cell;    // [(alpha-1) * n + i] departure cell (-1: identity row)
  const double* xi;   // unit coordinates, 2 per record
  const int* wall_idx;        // walls: wall entry per DoF or -1 (null: no walls)
  const unsigned* wall_flip;  //   per entry: bit alpha set when alpha reads the opposite direction
  const double* wall_bias;    //   per entry: Q values added to f after interpolation
  int opp[MAXQ];              // opposite direction index
  double gll[MAXP + 1];
  double inv_den[MAXP + 1];
  double cs2;         // unscaled c_s^2
  double ics2;        // 1 / cs2
  double k3, k4;      // 1 / (6 cs2^3), 1 / (24 cs2^4)
  double cv;          // C_v = 1 / (gamma - 1)
  double tau;
  double gamma;
  double prandtl;
  int prandtl_set;
  int sutherland;
  double rho0;        // FP32 shift of f (0 for FP64)
  double gshift;      // FP32 shift factor of g: w rho0 (2 C_v - D)
  double scaling;
  T* rho;
  T* u;
  double* temp;
  double* mask;
  int64_t mstride;
  int* status;
  int stream_only;  // CompressibleCFDSolver::stream alone: store the streamed f and g, no collision
};

// Quartic equilibrium of direction q for the node scalars (2D). The node factors
// (powers of 1/cs2 and rho) are folded into the Hermite coefficients once per node,
// so a direction costs multiply-adds only (no divisions).
struct QuarticNode {
  double rho, us0, us1, uu_term, k2a, k2b;
  double b_xxx, b_xxy, b_xyy, b_yyy, b_xxxx, b_yyyy, b_xxxy, b_xyyy, b_xxyy;
  template <typename T>
  __device__ void init(double r, double x, double y, double T_, const CompArgs<T>& a) {
    const double cs2 = a.cs2, ics2 = a.ics2;
    rho = r;
    us0 = x * ics2;
    us1 = y * ics2;
    uu_term = -(x * x) * (0.5 * ics2) - (y * y) * (0.5 * ics2);
    const double Tm1 = T_ - 1;
    // rho w (Tm1 |e|^2 - 2 cs2 Tm1) / (2 cs2)
    k2a = r * Tm1 * (0.5 * ics2);
    k2b = r * Tm1;
    const double T1 = cs2 * (T_ - 1);
    const double r3 = r * a.k3, r4 = r * a.k4;
    b_xxx = r3 * (x * x * x + T1 * (x + x + x));
    b_xxy = 3.0 * r3 * (x * x * y + T1 * y);
    b_xyy = 3.0 * r3 * (x * y * y + T1 * x);
    b_yyy = r3 * (y * y * y + T1 * (y + y + y));
    b_xxxx = r4 * (x * x * x * x + T1 * x * x * 6.0 + T1 * T1 * 3.0);
    b_yyyy = r4 * (y * y * y * y + T1 * y * y * 6.0 + T1 * T1 * 3.0);
    b_xxxy = 4.0 * r4 * (x * x * x * y + T1 * (x * y * 3.0));
    b_xyyy = 4.0 * r4 * (x * y * y * y + T1 * (x * y * 3.0));
    b_xxyy = 6.0 * r4 * (x * x * y * y + T1 * (x * x + y * y) + T1 * T1);
  }
  __device__ double feq(const double* __restrict__ t) const {
    const double ue = us0 * t[CT_E0] + us1 * t[CT_E1];
    double f = rho * (1 + ue * (1 + 0.5 * ue) + uu_term);
    f += k2a * t[CT_EE] - k2b;
    f += b_xxx * t[CT_H3 + 0] + b_xxy * t[CT_H3 + 1] + b_xyy * t[CT_H3 + 2] + b_yyy * t[CT_H3 + 3];
    f += b_xxxx * t[CT_H4 + 0] + b_yyyy * t[CT_H4 + 1] + b_xxyy * t[CT_H4 + 2] + b_xyyy * t[CT_H4 + 3] +
         b_xxxy * t[CT_H4 + 4];
    return t[CT_W] * f;
  }
};

// Node state of the compressible collision: moments, temperature, equilibrium
// coefficients and relaxation rates.
struct CompNode {
  double rho, u0, u1, temp, itemp, gfac, om_v, om_diff;
  QuarticNode qn;
};

// Node state from complete sums: rho, u, and tf = sum |e - u|^2 f, tg = sum g.
template <typename T>
__device__ __forceinline__ CompNode comp_node_sums(double rho, double u0, double u1, double tf, double tg,
                                                   const CompArgs<T>& a) {
  constexpr int D = 2;
  const double C_v = a.cv;
  CompNode nd;
  const double temp = (tf * a.ics2 + tg) * 0.5 / (rho * C_v);
  nd.rho = rho;
  nd.u0 = u0;
  nd.u1 = u1;
  nd.temp = temp;
  nd.itemp = 1.0 / temp;
  nd.qn.init(rho, u0, u1, temp, a);
  nd.gfac = temp * (2.0 * C_v - D);
  double S = 1.0;
  if (a.sutherland) S = pow(temp / 0.85, 0.7);
  const double visc_tau = (a.tau - 0.5) * S / (temp * rho) + 0.5;
  const double prandtl_tau = (visc_tau - 0.5) / (a.prandtl_set ? a.prandtl : 1.0) + 0.5;
  nd.om_v = 1. / visc_tau;
  nd.om_diff = nd.om_v - 1. / prandtl_tau;
  return nd;
}

// Sums over directions of the non-equilibrium parts: the Knudsen estimate and, with
// a Prandtl number (PR), the third moments Q of f_neq and first moments F of g_neq
// about u.
struct CompNeq {
  double kn = 0;
  double Qt[2][2][2] = {{{0, 0}, {0, 0}}, {{0, 0}, {0, 0}}};
  double Fg[2] = {0, 0};
};

// Adds direction t's terms; returns its f_eq.
template <bool PR, typename T>
__device__ __forceinline__ double comp_neq_add(CompNeq& s, const CompNode& nd, const CompArgs<T>& a,
                                               const double* __restrict__ t, double fq, double gq) {
  const double fe = nd.qn.feq(t);
  const double fneq = fq - fe;
  s.kn += fabs(fneq) * t[CT_IW];
  if (PR && a.prandtl_set) {
    const double gneq = gq - fe * nd.gfac;
    const double d[2] = {t[CT_E0] - nd.u0, t[CT_E1] - nd.u1};
    for (int x = 0; x < 2; x++)
      for (int y = 0; y < 2; y++)
        for (int z = 0; z < 2; z++) s.Qt[x][y][z] += (d[x] * d[y] * d[z]) * fneq;
    s.Fg[0] += d[0] * gneq;
    s.Fg[1] += d[1] * gneq;
  }
  return fe;
}

// Relaxed f and g of direction t, given its f_eq and the complete sums s.
template <bool PR, typename T>
__device__ __forceinline__ void comp_relax(const CompNode& nd, const CompNeq& s, const CompArgs<T>& a,
                                           const double* __restrict__ t, double fe, double fq, double gq, double& fn,
                                           double& gn) {
  const double cs2 = a.cs2;
  const double ge = fe * nd.gfac;
  double fstar = 0, gstar = 0;
  if (PR && a.prandtl_set) {
    // w Q:(e e e - 3 cs2 e I) / (6 cs2^3) with 1 / (6 cs2^3) = k3
    const double e[2] = {t[CT_E0], t[CT_E1]};
    double acc = 0;
    for (int x = 0; x < 2; x++)
      for (int y = 0; y < 2; y++)
        for (int z = 0; z < 2; z++) acc += s.Qt[x][y][z] * (e[x] * e[y] * e[z] - 3 * cs2 * e[z] * (x == y ? 1.0 : 0.0));
    fstar = t[CT_W] * acc * a.k3;
    gstar = t[CT_W] * (s.Fg[0] * e[0] + s.Fg[1] * e[1]) * nd.itemp;
  }
  fn = fq - (nd.om_v * (fq - fe) - nd.om_diff * fstar);
  gn = gq - (nd.om_v * (gq - ge) - nd.om_diff * gstar);
}

template <typename T>
__device__ __forceinline__ void comp_write_macro(const CompArgs<T>& a, const CompNode& nd, double kn_sum, int Q,
                                                 int64_t node) {
  a.rho[node] = (T)nd.rho;
  a.u[node] = (T)(nd.u0 * a.scaling);
  a.u[a.mstride + node] = (T)(nd.u1 * a.scaling);
  a.temp[node] = nd.temp;
  a.mask[node] = kn_sum / Q;
}

template <int P>
__device__ __forceinline__ void lag_all(const double (&x)[MAXP + 1], const double (&inv)[MAXP + 1], double xi,
                                        double (&v)[P + 1]) {
#pragma unroll
  for (int j = 0; j <= P; j++) {
    double s = inv[j];
#pragma unroll
    for (int m = 0; m <= P; m++)
      if (m != j) s *= (xi - x[m]);
    v[j] = s;
  }
}

// Streamed f and g of direction al at DoF i (al uniform across the wave), with the
// FP32 shift added back. One set of FP64 weights serves f and g.
template <int Q, int P, typename T>
__device__ __forceinline__ void comp_stream_dir(const CompArgs<T>& a, const double* __restrict__ tab, int64_t i,
                                                int al, int widx, unsigned wmask, double& fo, double& go) {
  const int n0 = a.nfe[0];
  const int64_t r = (int64_t)(al - 1) * a.n + i;
  const int sd = ((wmask >> al) & 1u) ? a.opp[al] : al;
  const T* fsrc = a.fin + (int64_t)sd * a.stride;
  const T* gsrc = a.gin + (int64_t)sd * a.stride;
  const double bias = widx >= 0 ? a.wall_bias[(int64_t)widx * Q + al] : 0.0;
  const int c = a.cell[r];
  double af, ag;
  if (c < 0) {
    af = (double)a.fin[(int64_t)al * a.stride + i];
    ag = (double)a.gin[(int64_t)al * a.stride + i];
  } else {
    double lx[P + 1], ly[P + 1];
    lag_all<P>(a.gll, a.inv_den, a.xi[2 * r + 0], lx);
    lag_all<P>(a.gll, a.inv_den, a.xi[2 * r + 1], ly);
    const int cy = (int)a.divN0.div((uint32_t)c), cx = c - cy * a.N[0];
    const int64_t base = (int64_t)cx * P + (int64_t)n0 * ((int64_t)cy * P);
    af = 0;
    ag = 0;
    double mnx = fabs(lx[0]), mny = fabs(ly[0]);
#pragma unroll
    for (int m = 1; m <= P; m++) {
      mnx = fmin(mnx, fabs(lx[m]));
      mny = fmin(mny, fabs(ly[m]));
    }
    if (mnx * mny >= 1e-10) {
      // no product weight reaches the cut-off: every term is taken, with the loads of
      // the whole stencil issued together (same terms in the same order)
      double vf[P + 1][P + 1], vg[P + 1][P + 1];
#pragma unroll
      for (int my = 0; my <= P; my++)
#pragma unroll
        for (int mx = 0; mx <= P; mx++) {
          const int64_t k = base + mx + (int64_t)n0 * my;
          vf[my][mx] = (double)fsrc[k];
          vg[my][mx] = (double)gsrc[k];
        }
#pragma unroll
      for (int my = 0; my <= P; my++)
#pragma unroll
        for (int mx = 0; mx <= P; mx++) {
          const double w = lx[mx] * ly[my];
          af += w * vf[my][mx];
          ag += w * vg[my][mx];
        }
    } else {
#pragma unroll
      for (int my = 0; my <= P; my++)
#pragma unroll
        for (int mx = 0; mx <= P; mx++) {
          const double w = lx[mx] * ly[my];
          if (!(fabs(w) < 1e-10)) {
            const int64_t k = base + mx + (int64_t)n0 * my;
            af += w * (double)fsrc[k];
            ag += w * (double)gsrc[k];
          }
        }
    }
  }
  af = af + bias;
  fo = af + tab[al * CTAB + CT_W] * a.rho0;
  go = ag + tab[al * CTAB + CT_W] * a.gshift;
}

// Fused SL stream of f and g + compressible collision. The work of a block is spread
// over its waves: 64 DoFs per block, CW waves each streaming every CW-th direction of
// those DoFs (one thread per DoF would give a 66k-DoF grid only four waves per CU to hide
// its gather latency). The streamed values meet in LDS. The collision is spread the
// same way: each wave sums the moments over its own directions, the partial sums meet in
// LDS and are added in wave order (so a run is reproducible), every wave forms the node
// state from them and relaxes its own directions. Without a Prandtl number the
// equilibrium of a direction serves both its Knudsen term and its relaxation; with one
// (PR) the third moments must be complete first, through one more LDS reduction.
// Waves per block: 8, or 4 with a Prandtl number, whose 11-value reduction buffer would
// otherwise leave room for one block per CU (C4 with Pr = 0.71: 0.97 -> 1.60 GLUPS FP32;
// 8 waves measured best without it).
template <bool PR>
constexpr int comp_waves() {
  return PR ? 4 : 8;
}
// ST = false: no streaming, each DoF's own populations are collided in place (the
// initial collision of CFDSolver::run; fin and fout may be the same arrays, since every
// value of the block is read into LDS before the barrier and only this block touches
// its 64 DoFs).
template <int Q, int P, typename T, bool PR, bool ST = true>
__global__ __launch_bounds__(comp_waves<PR>() * 64) void k_sl_comp_wide(CompArgs<T> a,
                                                                        const double* __restrict__ tab,
                                                                        int write_macro) {
  constexpr int CW = comp_waves<PR>();
  constexpr int NV = PR ? 11 : 5;  // kn, Q (8), F (2); moments: rho, j (3) + tf, tg (2)
  __shared__ double sf[Q][64], sg[Q][64];
  __shared__ double red[CW][NV][64];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int64_t i = (int64_t)blockIdx.x * 64 + lane;
  const bool live = i < a.n;
  int widx = -1;
  unsigned wmask = 0;
  if (ST && live && a.wall_idx) {
    widx = a.wall_idx[i];
    if (widx >= 0) wmask = a.wall_flip[widx];
  }
  // stream only (uniform over the grid): each wave stores the directions it streamed
  const bool so = ST && a.stream_only;
  if (live) {
    if (w == 0) {
      if (so) {
        a.fout[i] = a.fin[i];
        a.gout[i] = a.gin[i];
      } else {
        sf[0][lane] = (double)a.fin[i] + tab[CT_W] * a.rho0;
        sg[0][lane] = (double)a.gin[i] + tab[CT_W] * a.gshift;
      }
    }
    for (int al = 1 + w; al < Q; al += CW) {
      double fo, go;
      if constexpr (ST) {
        comp_stream_dir<Q, P, T>(a, tab, i, al, widx, wmask, fo, go);
      } else {
        fo = (double)a.fin[(int64_t)al * a.stride + i] + tab[al * CTAB + CT_W] * a.rho0;
        go = (double)a.gin[(int64_t)al * a.stride + i] + tab[al * CTAB + CT_W] * a.gshift;
      }
      if (so) {
        a.fout[(int64_t)al * a.stride + i] = (T)(fo - tab[al * CTAB + CT_W] * a.rho0);
        a.gout[(int64_t)al * a.stride + i] = (T)(go - tab[al * CTAB + CT_W] * a.gshift);
      } else {
        sf[al][lane] = fo;
        sg[al][lane] = go;
      }
    }
  }
  if (so) return;
  __syncthreads();
  const auto fl = [&](int q) { return sf[q][lane]; };
  const auto gl = [&](int q) { return sg[q][lane]; };
  // moments: each wave sums its own directions; partial sums meet in LDS, added in wave order
  double m0 = 0, m1 = 0, m2 = 0;
  if (live)
    for (int q = w; q < Q; q += CW) {
      const double f = fl(q);
      m0 += f;
      m1 += tab[q * CTAB + CT_E0] * f;
      m2 += tab[q * CTAB + CT_E1] * f;
    }
  red[w][0][lane] = m0;
  red[w][1][lane] = m1;
  red[w][2][lane] = m2;
  __syncthreads();
  double rho = red[0][0][lane], j0 = red[0][1][lane], j1 = red[0][2][lane];
  for (int v = 1; v < CW; v++) {
    rho += red[v][0][lane];
    j0 += red[v][1][lane];
    j1 += red[v][2][lane];
  }
  const double u0 = j0 * (1.0 / rho), u1 = j1 * (1.0 / rho);
  double tf = 0, tg = 0;
  if (live)
    for (int q = w; q < Q; q += CW) {
      const double d0 = tab[q * CTAB + CT_E0] - u0, d1 = tab[q * CTAB + CT_E1] - u1;
      tf += (d0 * d0 + d1 * d1) * fl(q);
      tg += gl(q);
    }
  red[w][3][lane] = tf;
  red[w][4][lane] = tg;
  __syncthreads();
  tf = red[0][3][lane];
  tg = red[0][4][lane];
  for (int v = 1; v < CW; v++) {
    tf += red[v][3][lane];
    tg += red[v][4][lane];
  }
  CompNode nd = {};
  if (live) nd = comp_node_sums<T>(rho, u0, u1, tf, tg, a);
  if (w == 0 && live && !(nd.rho >= 1e-10)) atomicOr(a.status, 1);
  auto store = [&](int q, double fn, double gn) {
    a.fout[(int64_t)q * a.stride + i] = (T)(fn - tab[q * CTAB + CT_W] * a.rho0);
    a.gout[(int64_t)q * a.stride + i] = (T)(gn - tab[q * CTAB + CT_W] * a.gshift);
  };
  CompNeq s;
  if constexpr (PR) {
    if (live)
      for (int q = w; q < Q; q += CW) comp_neq_add<true, T>(s, nd, a, tab + q * CTAB, fl(q), gl(q));
    __syncthreads();  // every wave has read the moment sums
    red[w][0][lane] = s.kn;
#pragma unroll
    for (int k = 0; k < 8; k++) red[w][1 + k][lane] = s.Qt[k >> 2][(k >> 1) & 1][k & 1];
    red[w][9][lane] = s.Fg[0];
    red[w][10][lane] = s.Fg[1];
    __syncthreads();
    CompNeq t;
    t.kn = red[0][0][lane];
#pragma unroll
    for (int k = 0; k < 8; k++) t.Qt[k >> 2][(k >> 1) & 1][k & 1] = red[0][1 + k][lane];
    t.Fg[0] = red[0][9][lane];
    t.Fg[1] = red[0][10][lane];
    for (int v = 1; v < CW; v++) {
      t.kn += red[v][0][lane];
#pragma unroll
      for (int k = 0; k < 8; k++) t.Qt[k >> 2][(k >> 1) & 1][k & 1] += red[v][1 + k][lane];
      t.Fg[0] += red[v][9][lane];
      t.Fg[1] += red[v][10][lane];
    }
    if (live)
      for (int q = w; q < Q; q += CW) {
        const double* tq = tab + q * CTAB;
        double fn, gn;
        comp_relax<true, T>(nd, t, a, tq, nd.qn.feq(tq), fl(q), gl(q), fn, gn);
        store(q, fn, gn);
      }
    s.kn = t.kn;
  } else {
    if (live)
      for (int q = w; q < Q; q += CW) {
        const double* tq = tab + q * CTAB;
        const double fe = comp_neq_add<false, T>(s, nd, a, tq, fl(q), gl(q));
        double fn, gn;
        comp_relax<false, T>(nd, s, a, tq, fe, fl(q), gl(q), fn, gn);
        store(q, fn, gn);
      }
    if (write_macro) {  // uniform over the block
      red[w][0][lane] = s.kn;
      __syncthreads();
      s.kn = red[0][0][lane];
      for (int v = 1; v < CW; v++) s.kn += red[v][0][lane];
    }
  }
  if (write_macro && w == 0 && live) comp_write_macro<T>(a, nd, s.kn, Q, i);
}

template <typename T>
static void fill_comp(nat_ctx* ctx, CompArgs<T>& a, int src, int dst) {
  a.fin = (const T*)ctx->f[src];
  a.gin = (const T*)ctx->g[src];
  a.fout = (T*)ctx->f[dst];
  a.gout = (T*)ctx->g[dst];
  a.stride = ctx->stride;
  a.n = ctx->n_dofs;
  for (int d = 0; d < 3; d++) {
    a.nfe[d] = ctx->nfe[d];
    a.N[d] = d < ctx->D ? ctx->desc.ncells[d] : 1;
  }
  a.divN0.set((uint32_t)a.N[0]);
  a.cell = ctx->dep_cell;
  a.xi = ctx->dep_xi64;
  a.wall_idx = ctx->has_walls ? ctx->wall_idx : nullptr;
  a.wall_flip = ctx->wall_flip;
  a.wall_bias = ctx->wall_bias;
  for (int q = 0; q < MAXQ; q++) a.opp[q] = q;
  for (int q = 0; q < ctx->Q; q++)
    for (int r = 0; r < ctx->Q; r++)
      if (ctx->st.dir[r][0] == -ctx->st.dir[q][0] && ctx->st.dir[r][1] == -ctx->st.dir[q][1] &&
          ctx->st.dir[r][2] == -ctx->st.dir[q][2])
        a.opp[q] = r;
  double x[MAXP + 1], w[MAXP + 1];
  gll_points(ctx->p, x, w);
  for (int j = 0; j <= MAXP; j++) {
    a.gll[j] = j <= ctx->p ? x[j] : 0.0;
    double d = 1;
    for (int m = 0; m <= ctx->p; m++)
      if (m != j && j <= ctx->p) d *= (x[j] - x[m]);
    a.inv_den[j] = j <= ctx->p ? 1.0 / d : 0.0;
  }
  const double s = ctx->st.s;
  a.cs2 = ctx->st.cs2 / (s * s);
  a.ics2 = 1.0 / a.cs2;
  a.k3 = 1.0 / (6.0 * a.cs2 * a.cs2 * a.cs2);
  a.k4 = 1.0 / (24.0 * a.cs2 * a.cs2 * a.cs2 * a.cs2);
  a.cv = 1.0 / (ctx->desc.gamma - 1.0);
  a.tau = ctx->tau;
  a.gamma = ctx->desc.gamma;
  a.prandtl = ctx->desc.prandtl;
  a.prandtl_set = ctx->desc.prandtl > 0 ? 1 : 0;
  a.sutherland = ctx->desc.sutherland;
  const double C_v = 1. / (ctx->desc.gamma - 1.0);
  const double rho0 = ctx->fp32 ? (ctx->desc.rho0 > 0 ? ctx->desc.rho0 : 1.0) : 0.0;
  a.rho0 = rho0;
  a.gshift = rho0 * (2.0 * C_v - ctx->D);
  a.scaling = s;
  a.rho = (T*)ctx->rho;
  a.u = (T*)ctx->u;
  a.temp = ctx->T;
  a.mask = ctx->mask;
  a.mstride = ctx->n_store;
  a.status = ctx->d_status;
  a.stream_only = 0;
}

template <typename T>
static int step_comp(nat_ctx* ctx, int mode) {
  CompArgs<T> a;
  const int src = ctx->cur, dst = 1 - ctx->cur;
  fill_comp<T>(ctx, a, src, dst);
  a.stream_only = (mode & STEP_STREAM_ONLY) ? 1 : 0;
  const int write_macro = (mode & STEP_MACRO) ? 1 : 0;
  const unsigned grid = (unsigned)((a.n + 63) / 64);
  const double* tab = ctx->comp_tab;
  const dim3 blk(comp_waves<true>() * 64), blk0(comp_waves<false>() * 64);
#define NAT_CW(PP)                                                                                              \
  case PP:                                                                                                     \
    if (a.prandtl_set)                                                                                         \
      hipLaunchKernelGGL((k_sl_comp_wide<25, PP, T, true>), dim3(grid), blk, 0, ctx->stream, a, tab, write_macro); \
    else                                                                                                       \
      hipLaunchKernelGGL((k_sl_comp_wide<25, PP, T, false>), dim3(grid), blk0, 0, ctx->stream, a, tab,           \
                         write_macro);                                                                         \
    break;
  switch (ctx->p) {
    NAT_CW(1)
    NAT_CW(2)
    NAT_CW(3)
    NAT_CW(4)
    default:
      ctx->err = "compressible SL kernel instantiated for p <= 4";
      return NAT_ERR_UNSUPPORTED;
  }
#undef NAT_CW
  NAT_HIP(hipGetLastError());
  ctx->cur = dst;
  return 0;
}

template <typename T>
static int collide_comp(nat_ctx* ctx) {
  CompArgs<T> a;
  fill_comp<T>(ctx, a, ctx->cur, ctx->cur);
  const unsigned grid = (unsigned)((a.n + 63) / 64);
  const double* tab = ctx->comp_tab;
  if (a.prandtl_set)
    hipLaunchKernelGGL((k_sl_comp_wide<25, 1, T, true, false>), dim3(grid), dim3(comp_waves<true>() * 64), 0,
                       ctx->stream, a, tab, 1);
  else
    hipLaunchKernelGGL((k_sl_comp_wide<25, 1, T, false, false>), dim3(grid), dim3(comp_waves<false>() * 64), 0,
                       ctx->stream, a, tab, 1);
  NAT_HIP(hipGetLastError());
  return 0;
}

}  // namespace nat

// Constant table: unscaled e, w, |e|^2 and the independent H3 / H4 components
// (AuxiliaryCollisionFunctions.h:519-566) per direction.
int nat_comp_setup(nat_ctx* ctx) {
  using namespace nat;
  if (ctx->D != 2 || ctx->Q != 25) {
    ctx->err = "compressible populations are implemented for the D2Q25H stencil";
    return NAT_ERR_UNSUPPORTED;
  }
  const double s = ctx->st.s;
  const double cs2 = ctx->st.cs2 / (s * s);
  std::vector<double> t((size_t)ctx->Q * CTAB, 0.0);
  for (int q = 0; q < ctx->Q; q++) {
    const double e[2] = {ctx->st.dir[q][0] / s, ctx->st.dir[q][1] / s};
    double* r = &t[(size_t)q * CTAB];
    r[CT_E0] = e[0];
    r[CT_E1] = e[1];
    r[CT_W] = ctx->st.w[q];
    r[CT_EE] = e[0] * e[0] + e[1] * e[1];
    auto eye = [](int a, int b) { return a == b ? 1.0 : 0.0; };
    auto H3 = [&](int a, int b, int c) {
      return e[a] * e[b] * e[c] - cs2 * (e[a] * eye(b, c) + e[b] * eye(a, c) + e[c] * eye(a, b));
    };
    auto H4 = [&](int a, int b, int c, int d) {
      const double p4 = e[a] * e[b] * e[c] * e[d];
      const double p2 = e[a] * e[b] * eye(c, d) + e[a] * e[c] * eye(b, d) + e[a] * e[d] * eye(b, c) +
                        e[b] * e[c] * eye(a, d) + e[b] * e[d] * eye(a, c) + e[c] * e[d] * eye(a, b);
      const double p0 = eye(a, b) * eye(c, d) + eye(a, c) * eye(b, d) + eye(a, d) * eye(b, c);
      return p4 - cs2 * p2 + cs2 * cs2 * p0;
    };
    r[CT_H3 + 0] = H3(0, 0, 0);
    r[CT_H3 + 1] = H3(0, 0, 1);
    r[CT_H3 + 2] = H3(0, 1, 1);
    r[CT_H3 + 3] = H3(1, 1, 1);
    r[CT_H4 + 0] = H4(0, 0, 0, 0);
    r[CT_H4 + 1] = H4(1, 1, 1, 1);
    r[CT_H4 + 2] = H4(0, 0, 1, 1);
    r[CT_H4 + 3] = H4(0, 1, 1, 1);
    r[CT_H4 + 4] = H4(0, 0, 0, 1);
    r[CT_IW] = 1.0 / ctx->st.w[q];
  }
  if (!ctx->comp_tab) NAT_HIP(hipMalloc(&ctx->comp_tab, t.size() * sizeof(double)));
  NAT_HIP(hipMemcpy(ctx->comp_tab, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

int nat_launch_sl_comp(nat_ctx* ctx, int write_macro) {
  return ctx->fp32 ? nat::step_comp<float>(ctx, write_macro) : nat::step_comp<double>(ctx, write_macro);
}

int nat_launch_collide_comp(nat_ctx* ctx) {
  return ctx->fp32 ? nat::collide_comp<float>(ctx) : nat::collide_comp<double>(ctx);
}
