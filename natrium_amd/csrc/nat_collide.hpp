// Pointwise BGK collision in registers, shared by the fused step kernels.
//
// Semantics of collision_advanced/CollisionOperator.h:52-102 with
// BGKCollision::relax (CollisionSchemes.h:28-41) and BGKEquilibrium
// (Equilibria.h:66-83): rho = sum f, u = sum e f / rho (unscaled e), relax
// f -= (f - f_eq)/tau, write rho and u*scaling.
//
// Populations may be stored shifted, f~ = f - w rho0 (rho0 = 0: unshifted FP64).
// Because sum_i w_i = 1 and sum_i w_i e_i = 0, rho = rho0 + sum f~ and
// rho u = sum e f~ exactly, and f~_eq = w (drho + rho (ue (1 + ue/2) - u.u/(2 cs2))),
// so BGK relaxes f~ without ever forming the O(1) background.
#pragma once
#include "nat_internal.hpp"

namespace nat {

template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (I < N) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, N>(f);
  }
}

// Pseudo-entropic stabilizer (dataprocessors/PseudoEntropicStabilizer.cpp:271-311), applied
// after the collision of each loop step (solver/CFDSolver.cpp:877-892): the post-collision
// populations are projected onto w times the Hermite polynomials of order <= 2, the
// w-orthogonal projection P with P f_eq = f_eq. Since P keeps density and momentum,
// P (f - omega (f - f_eq)) = f_eq + (1 - omega) P f_neq with
// (P f_neq)_i = w_i (e_i e_i - cs2 I) : Pi_neq / (2 cs2^2), Pi_neq = sum_j e_j e_j f_neq_j
// = Pi(f) - rho (u u + cs2 I). In the shifted variables Pi(f~) = Pi(f) - rho0 cs2 I, so
// Pi_neq = Pi(f~) - rho u u - drho cs2 I. Unscaled units: cs2 = 1/3.
template <int Q>
struct StabPi {
  double xx = 0, yy = 0, zz = 0, xy = 0, xz = 0, yz = 0;
};

// second moments of the populations (e_i integer)
template <int Q>
__device__ __forceinline__ StabPi<Q> stab_moments(const double (&f)[Q]) {
  using L = Lattice<Q>;
  StabPi<Q> m;
#pragma unroll
  for (int i = 0; i < Q; i++) {
    const int ex = L::e[i][0], ey = L::e[i][1], ez = L::e[i][2];
    if (ex) m.xx += f[i];
    if (ey) m.yy += f[i];
    if (ez) m.zz += f[i];
    if (ex * ey > 0) m.xy += f[i];
    if (ex * ey < 0) m.xy -= f[i];
    if (ex * ez > 0) m.xz += f[i];
    if (ex * ez < 0) m.xz -= f[i];
    if (ey * ez > 0) m.yz += f[i];
    if (ey * ez < 0) m.yz -= f[i];
  }
  return m;
}

// Pi_neq from the shifted moments, scaled by (1 - omega) 9/2 so that the correction of
// direction i is w_i (sum_ab e_a e_b P_ab - tr P / 3)
template <int Q>
__device__ __forceinline__ StabPi<Q> stab_neq(StabPi<Q> m, double drho, double rho, const double (&u)[3],
                                              double om) {
#pragma clang fp contract(off)
  const double k = (1.0 - om) * 4.5;
  const double c = drho / 3.0;
  m.xx = k * (m.xx - rho * u[0] * u[0] - c);
  m.yy = k * (m.yy - rho * u[1] * u[1] - c);
  m.zz = k * (m.zz - rho * u[2] * u[2] - (Lattice<Q>::D == 3 ? c : 0.0));
  m.xy = k * (m.xy - rho * u[0] * u[1]);
  m.xz = k * (m.xz - rho * u[0] * u[2]);
  m.yz = k * (m.yz - rho * u[1] * u[2]);
  return m;
}

template <int Q>
__device__ __forceinline__ double stab_term(const StabPi<Q>& m, int i) {
#pragma clang fp contract(off)
  using L = Lattice<Q>;
  const int ex = L::e[i][0], ey = L::e[i][1], ez = L::e[i][2];
  double v = 0;
  if (ex) v += m.xx;
  if (ey) v += m.yy;
  if (ez) v += m.zz;
  if (ex * ey > 0) v += 2 * m.xy;
  if (ex * ey < 0) v -= 2 * m.xy;
  if (ex * ez > 0) v += 2 * m.xz;
  if (ex * ez < 0) v -= 2 * m.xz;
  if (ey * ez > 0) v += 2 * m.yz;
  if (ey * ez < 0) v -= 2 * m.yz;
  const double tr = (Lattice<Q>::D == 3) ? (m.xx + m.yy + m.zz) : (m.xx + m.yy);
  return L::w[i] * (v - tr / 3.0);
}

// Opposite direction of i in the integer velocity table.
template <int Q>
constexpr int opposite(int i) {
  for (int j = 0; j < Q; j++)
    if (Lattice<Q>::e[j][0] == -Lattice<Q>::e[i][0] && Lattice<Q>::e[j][1] == -Lattice<Q>::e[i][1] &&
        Lattice<Q>::e[j][2] == -Lattice<Q>::e[i][2])
      return j;
  return -1;
}

// FP64 BGK collision of populations held in storage precision, every relaxed population
// stored as soon as it is formed. The equilibrium of each opposite pair (i, i') comes from
// shared even and odd parts: f_eq(i/i') = w (E +- O), E = drho + rho (ue^2/2 - 3u.u/2),
// O = rho ue (the direction-by-direction form of Equilibria.h:66-83 up to FP64 rounding).
// Every step kernel and the collision-only pass use this one function, so their results
// agree bit for bit. Stores go through a uniform direction base and a 32-bit byte offset.
//
// INIT: the collision of the reference's iterative initialisation
// (collision_advanced/CollisionOperator.h:79-91, inInitializationProcedure): the velocity
// is read from the stored field (u / scaling) instead of computed, and only rho is written.
template <int Q, typename T, bool INIT = false>
__device__ __forceinline__ void bgk_collide_store_pairs(const T (&f)[Q], const CollideArgs<T>& ca, int64_t node,
                                                        bool write_macro, T* __restrict__ fout, int64_t stride) {
  // Fused multiply-adds are written out and no other contraction is allowed: every kernel
  // that inlines this function rounds the same way, whatever the compiler decides elsewhere.
#pragma clang fp contract(off)
  using L = Lattice<Q>;
  constexpr int D = L::D;
  double fd[Q];
#pragma unroll
  for (int i = 0; i < Q; i++) fd[i] = (double)f[i];
  double drho = 0;
#pragma unroll
  for (int i = 0; i < Q; i++) drho += fd[i];
  double j[3] = {0, 0, 0};
#pragma unroll
  for (int i = 0; i < Q; i++) {
#pragma unroll
    for (int d = 0; d < D; d++) {
      if (L::e[i][d] > 0) j[d] += fd[i];
      if (L::e[i][d] < 0) j[d] -= fd[i];
    }
  }
  const double rho = (double)ca.rho0 + drho;
  if (!(rho >= 1e-10)) atomicOr(ca.status, 1);
  const double inv = 1.0 / rho;
  double u[3], u3[3];
#pragma unroll
  for (int d = 0; d < D; d++) {
    if constexpr (INIT)
      u[d] = (double)ca.u[d * ca.mstride + node] / ca.scaling;
    else
      u[d] = j[d] * inv;
    u3[d] = 3.0 * u[d];
  }
  double uu = 0;
#pragma unroll
  for (int d = 0; d < D; d++) uu = fma(u[d], u[d], uu);
  const double base = fma(-rho, 1.5 * uu, drho);  // even part without the ue^2 term
  const double hr = 0.5 * rho;
  const double om = ca.omega64;
  const uint32_t nb = (uint32_t)node * (uint32_t)sizeof(T);
  // raw buffer stores against a uniform direction base with the sc1 cache policy
  // (device-scope write-through): the step's output reaches memory while the kernel
  // runs instead of in the write-back at the kernel boundary (measured 1.8% on the
  // 16^3 p=4 D3Q27 step in FP32, 3% in FP64). Offsets stay below 2^31: the launcher
  // checks one direction's array against INT32_MAX bytes.
  auto store = [&](int i, double v) {
    const uint64_t pb = (uint64_t)(fout + (int64_t)i * stride);
    const uint32_t phi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(pb >> 32));
    const uint32_t plo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)pb);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(((uint64_t)phi << 32) | (uint64_t)plo), (short)0, 0x7fffffff, 0x00020000);
    if constexpr (sizeof(T) == 4)
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, (float)v), rs, (int)nb, 0, 16);
    else
      __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(unsigned __attribute__((ext_vector_type(2))), v), rs,
                                            (int)nb, 0, 16);
  };
  if ((Q == 9 || Q == 19) && ca.stab) {
    // stabilized: f_eq + (1 - omega) P f_neq; the correction is even in e
    double uu3[3] = {u[0], u[1], D == 3 ? u[2] : 0.0};
    const StabPi<Q> pn = stab_neq<Q>(stab_moments<Q>(fd), drho, rho, uu3, om);
    store(0, fma(L::w[0], base, stab_term<Q>(pn, 0)));
    static_for<1, Q>([&](auto I) {
      constexpr int i = decltype(I)::value;
      constexpr int o = opposite<Q>(i);
      if constexpr (o > i) {
        double ue = 0;
#pragma unroll
        for (int d = 0; d < D; d++) {
          if (L::e[i][d] > 0) ue += u3[d];
          if (L::e[i][d] < 0) ue -= u3[d];
        }
        const double E = fma(hr, ue * ue, base);
        const double O = rho * ue;
        const double wE = fma(L::w[i], E, stab_term<Q>(pn, i)), wO = L::w[i] * O;
        store(i, wE + wO);
        store(o, wE - wO);
      }
    });
  } else {
    store(0, fma(-om, fd[0] - L::w[0] * base, fd[0]));
    static_for<1, Q>([&](auto I) {
      constexpr int i = decltype(I)::value;
      constexpr int o = opposite<Q>(i);
      if constexpr (o > i) {
        double ue = 0;
#pragma unroll
        for (int d = 0; d < D; d++) {
          if (L::e[i][d] > 0) ue += u3[d];
          if (L::e[i][d] < 0) ue -= u3[d];
        }
        const double E = fma(hr, ue * ue, base);
        const double O = rho * ue;
        const double wE = L::w[i] * E, wO = L::w[i] * O;
        store(i, fma(-om, fd[i] - (wE + wO), fd[i]));
        store(o, fma(-om, fd[o] - (wE - wO), fd[o]));
      }
    });
  }
  if (write_macro && ca.rho) {
    ca.rho[node] = (T)rho;
    if constexpr (!INIT) {
#pragma unroll
      for (int d = 0; d < D; d++) ca.u[d * ca.mstride + node] = (T)(u[d] * ca.scaling);
    }
  }
}

}  // namespace nat
