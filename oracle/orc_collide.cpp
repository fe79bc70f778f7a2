// Oracle collision operators. TEST INFRASTRUCTURE ONLY (see orc_collide.hpp).
#include "orc_collide.hpp"

#include <cmath>

namespace orc {

namespace {

struct GenData {
  int D, Q;
  double e[32][3];  // unscaled directions (AuxiliaryCollisionFunctions.h:161-165)
  double w[32];
  double cs2;       // unscaled (AuxiliaryCollisionFunctions.h:172)
  double tau;       // AuxiliaryCollisionFunctions.h:63-68,177
  double scaling;
};

GenData make_gendata(const Stencil& st, const CollisionParams& cp) {
  GenData g;
  g.D = st.D;
  g.Q = st.Q;
  g.scaling = cp.scaling;
  for (int i = 0; i < st.D; i++)
    for (int j = 0; j < st.Q; j++) g.e[j][i] = st.dir[j][i] / cp.scaling;
  for (int j = 0; j < st.Q; j++) {
    if (st.D == 2) g.e[j][2] = 0;
    g.w[j] = st.w[j];
  }
  g.cs2 = st.cs2 / (cp.scaling * cp.scaling);
  g.tau = (cp.viscosity) / (cp.dt * st.cs2) + 0.5;
  return g;
}

inline double eye(int a, int b) { return a == b ? 1.0 : 0.0; }

void velocity_generic(const GenData& g, const double* f, double rho, double* u) {
  for (int j = 0; j < g.D; j++) {
    u[j] = 0.0;
    for (int i = 0; i < g.Q; i++) u[j] += g.e[i][j] * f[i];
    u[j] = u[j] * 1.0 / rho;
  }
}

void velocity(const GenData& g, const double* f, double rho, double* u) {
  if (g.D == 2 && g.Q == 9) {
    // AuxiliaryCollisionFunctions.h:258-268
    u[0] = 1.0 / rho * (f[1] + f[5] + f[8] - f[3] - f[6] - f[7]);
    u[1] = 1.0 / rho * (f[2] + f[5] + f[6] - f[4] - f[7] - f[8]);
  } else if (g.D == 3 && g.Q == 19) {
    // AuxiliaryCollisionFunctions.h:270-287
    u[0] = 1.0 / rho * (f[1] - f[3] + f[7] - f[8] - f[9] + f[10] + f[11] + f[12] - f[13] - f[14]);
    u[1] = 1.0 / rho * (-f[5] + f[6] - f[11] + f[12] + f[13] - f[14] - f[15] + f[16] + f[17] - f[18]);
    u[2] = 1.0 / rho * (f[2] - f[4] + f[7] + f[8] - f[9] - f[10] + f[15] + f[16] - f[17] - f[18]);
  } else {
    velocity_generic(g, f, rho, u);
  }
}

void feq_bgk(const GenData& g, double rho, const double* u, double* feq) {
  if (g.D == 2 && g.Q == 9) {
    // Equilibria.h:33-63
    double prefactor = 1. / g.cs2;
    double scalar_product = u[0] * u[0] + u[1] * u[1];
    double uSquareTerm = -scalar_product / (2 * g.cs2);
    double weighting = 4. / 9. * rho;
    feq[0] = weighting * (1 + uSquareTerm);
    weighting = 1. / 9. * rho;
    double m = prefactor * (u[0]);
    feq[1] = weighting * (1 + m * (1 + 0.5 * m) + uSquareTerm);
    feq[3] = weighting * (1 - m * (1 - 0.5 * m) + uSquareTerm);
    m = prefactor * (u[1]);
    feq[2] = weighting * (1 + m * (1 + 0.5 * m) + uSquareTerm);
    feq[4] = weighting * (1 - m * (1 - 0.5 * m) + uSquareTerm);
    weighting = 1. / 36. * rho;
    m = prefactor * (u[0] + u[1]);
    feq[5] = weighting * (1 + m * (1 + 0.5 * m) + uSquareTerm);
    feq[7] = weighting * (1 - m * (1 - 0.5 * m) + uSquareTerm);
    m = prefactor * (-u[0] + u[1]);
    feq[6] = weighting * (1 + m * (1 + 0.5 * m) + uSquareTerm);
    feq[8] = weighting * (1 - m * (1 - 0.5 * m) + uSquareTerm);
    return;
  }
  // Equilibria.h:66-83
  double uu_term = 0.0;
  for (int j = 0; j < g.D; j++) uu_term += -(u[j] * u[j]) / (2.0 * g.cs2);
  for (int i = 0; i < g.Q; i++) {
    double ue = 0.0;
    for (int j = 0; j < g.D; j++) ue += (u[j] * g.e[i][j]) / g.cs2;
    feq[i] = g.w[i] * rho * (1 + ue * (1 + 0.5 * (ue)) + uu_term);
  }
}

struct HTensors {
  double H3[32][3][3][3];
  double H4[32][3][3][3][3];
};

void make_H(const GenData& g, HTensors& h) {
  // AuxiliaryCollisionFunctions.h:519-566
  const int D = g.D;
  for (int i = 0; i < g.Q; i++) {
    const double* e = g.e[i];
    for (int a = 0; a < D; a++)
      for (int b = 0; b < D; b++)
        for (int c = 0; c < D; c++) {
          h.H3[i][a][b][c] = e[a] * e[b] * e[c] - g.cs2 * (e[a] * eye(b, c) + e[b] * eye(a, c) + e[c] * eye(a, b));
          for (int d = 0; d < D; d++) {
            const double power4 = e[a] * e[b] * e[c] * e[d];
            const double power2 = e[a] * e[b] * eye(c, d) + e[a] * e[c] * eye(b, d) + e[a] * e[d] * eye(b, c) +
                                  e[b] * e[c] * eye(a, d) + e[b] * e[d] * eye(a, c) + e[c] * e[d] * eye(a, b);
            const double power0 = eye(a, b) * eye(c, d) + eye(a, c) * eye(b, d) + eye(a, d) * eye(b, c);
            h.H4[i][a][b][c][d] = (power4 - g.cs2 * power2 + g.cs2 * g.cs2 * power0);
          }
        }
  }
}

void feq_quartic(const GenData& g, const HTensors& h, double rho, const double* u, double T, double* feq) {
  // Equilibria.h:119-267 (QuarticEquilibrium::polynomial)
  const int D = g.D;
  const double cs2 = g.cs2;
  double uu_term = 0.0;
  for (int j = 0; j < D; j++) uu_term += -(u[j] * u[j]) / (2.0 * cs2);
  const double T1 = cs2 * (T - 1);
  const double a_xxx = u[0] * u[0] * u[0] + T1 * (u[0] + u[0] + u[0]);
  const double a_xxy = u[0] * u[0] * u[1] + T1 * (u[1]);
  const double a_xyy = u[0] * u[1] * u[1] + T1 * (u[0]);
  const double a_yyy = u[1] * u[1] * u[1] + T1 * (u[1] + u[1] + u[1]);
  const double a_xxxx = u[0] * u[0] * u[0] * u[0] + T1 * u[0] * u[0] * 6.0 + T1 * T1 * 3.0;
  const double a_yyyy = u[1] * u[1] * u[1] * u[1] + T1 * u[1] * u[1] * 6.0 + T1 * T1 * 3.0;
  const double a_xxxy = u[0] * u[0] * u[0] * u[1] + T1 * (u[0] * u[1] * 3.0);
  const double a_xyyy = u[0] * u[1] * u[1] * u[1] + T1 * (u[0] * u[1] * 3.0);
  const double a_xxyy = u[0] * u[0] * u[1] * u[1] + T1 * (u[0] * u[0] + u[1] * u[1]) + T1 * T1;
  double a_zzz = 0, a_xxz = 0, a_xzz = 0, a_yzz = 0, a_yyz = 0, a_xyz = 0;
  double a_zzzz = 0, a_xzzz = 0, a_xxzz = 0, a_xxxz = 0, a_yzzz = 0, a_yyzz = 0, a_yyyz = 0, a_xxyz = 0, a_xyyz = 0,
         a_xyzz = 0;
  if (D == 3) {
    a_zzz = u[2] * u[2] * u[2] + T1 * (u[2] + u[2] + u[2]);
    a_xxz = u[0] * u[0] * u[2] + T1 * (u[2]);
    a_xzz = u[0] * u[2] * u[2] + T1 * (u[0]);
    a_yzz = u[1] * u[2] * u[2] + T1 * (u[1]);
    a_yyz = u[1] * u[1] * u[2] + T1 * (u[2]);
    a_xyz = u[0] * u[1] * u[2];
    a_zzzz = u[2] * u[2] * u[2] * u[2] + T1 * u[2] * u[2] * 6.0 + T1 * T1 * 3.0;
    a_xxxz = u[0] * u[0] * u[0] * u[2] + T1 * (u[0] * u[2] * 3.0);
    a_yyyz = u[1] * u[1] * u[1] * u[2] + T1 * (u[1] * u[2] * 3.0);
    a_xzzz = u[0] * u[2] * u[2] * u[2] + T1 * (u[0] * u[2] * 3.0);
    a_yzzz = u[1] * u[2] * u[2] * u[2] + T1 * (u[1] * u[2] * 3.0);
    a_xxzz = u[0] * u[0] * u[2] * u[2] + T1 * (u[0] * u[0] + u[2] * u[2]) + T1 * T1;
    a_yyzz = u[1] * u[1] * u[2] * u[2] + T1 * (u[1] * u[1] + u[2] * u[2]) + T1 * T1;
    a_xyzz = u[0] * u[1] * u[2] * u[2] + T1 * (u[0] * u[1]);
    a_xyyz = u[0] * u[1] * u[1] * u[2] + T1 * (u[0] * u[2]);
    a_xxyz = u[0] * u[0] * u[1] * u[2] + T1 * (u[1] * u[2]);
  }
  for (int i = 0; i < g.Q; i++) {
    const double* e = g.e[i];
    const double w = g.w[i];
    double ue = 0.0;
    for (int j = 0; j < D; j++) ue += (u[j] * e[j]) / cs2;
    feq[i] = w * rho * (1 + ue * (1 + 0.5 * (ue)) + uu_term);
    for (int alp = 0; alp < D; alp++)
      for (int bet = 0; bet < D; bet++)
        feq[i] += rho * w / (2.0 * cs2) *
                  ((T - 1) * eye(alp, bet) * e[alp] * e[bet] - cs2 * eye(alp, bet) * (T - 1));
    const auto& H3 = h.H3[i];
    const auto& H4 = h.H4[i];
    feq[i] += w * rho / (6. * cs2 * cs2 * cs2) *
              (a_xxx * H3[0][0][0] + 3 * (a_xxy * H3[0][0][1] + a_xyy * H3[0][1][1]) + a_yyy * H3[1][1][1]);
    if (D == 3) {
      feq[i] += w * rho / (6. * cs2 * cs2 * cs2) *
                (a_zzz * H3[2][2][2] +
                 3 * (a_xxz * H3[0][0][2] + a_xzz * H3[0][2][2] + a_yzz * H3[1][2][2] + a_yyz * H3[1][1][2]) +
                 6.0 * a_xyz * H3[0][1][2]);
    }
    feq[i] += w * rho / (24. * cs2 * cs2 * cs2 * cs2) *
              (H4[0][0][0][0] * a_xxxx + H4[1][1][1][1] * a_yyyy + 6.0 * H4[0][0][1][1] * a_xxyy +
               4.0 * H4[0][1][1][1] * a_xyyy + 4.0 * H4[0][0][0][1] * a_xxxy);
    if (D == 3) {
      feq[i] += w * rho / (24. * cs2 * cs2 * cs2 * cs2) *
                (H4[2][2][2][2] * a_zzzz +
                 4.0 * (H4[0][2][2][2] * a_xzzz + H4[1][2][2][2] * a_yzzz + H4[0][0][0][2] * a_xxxz +
                        H4[1][1][1][2] * a_yyyz) +
                 6.0 * (H4[0][0][2][2] * a_xxzz + H4[1][1][2][2] * a_yyzz) +
                 12.0 * (H4[0][0][1][2] * a_xxyz + H4[0][1][1][2] * a_xyyz + H4[0][1][2][2] * a_xyzz));
    }
  }
}

}  // namespace

int collide_bgk(const Stencil& st, const CollisionParams& cp, int64_t n, double* const* f, double* rho,
                double* const* u, bool init) {
  const GenData g = make_gendata(st, cp);
  int err = 0;
#pragma omp parallel for schedule(static) reduction(| : err)
  for (int64_t ii = 0; ii < n; ii++) {
    double fl[32], feq[32], vel[3] = {0, 0, 0};
    for (int p = 0; p < g.Q; ++p) fl[p] = f[p][ii];
    double density = 0.0;
    for (int p = 0; p < g.Q; ++p) density += fl[p];
    if (density < 1e-10) err |= 1;
    rho[ii] = density;
    velocity(g, fl, density, vel);
    if (!init)
      for (int j = 0; j < g.D; ++j) u[j][ii] = vel[j] * g.scaling;
    else
      for (int j = 0; j < g.D; ++j) vel[j] = u[j][ii] / g.scaling;
    feq_bgk(g, density, vel, feq);
    for (int p = 0; p < g.Q; ++p) fl[p] -= 1. / g.tau * (fl[p] - feq[p]);
    for (int p = 0; p < g.Q; ++p) f[p][ii] = fl[p];
  }
  return err;
}

int collide_compressible(const Stencil& st, const CollisionParams& cp, int64_t n, double* const* f, double* const* g_,
                         double* rho, double* const* u, double* T, double* mask) {
  const GenData g = make_gendata(st, cp);
  HTensors* h = new HTensors;
  make_H(g, *h);
  const double gamma = cp.gamma;
  const double C_v = 1. / (gamma - 1.0);
  int err = 0;
#pragma omp parallel for schedule(static) reduction(| : err)
  for (int64_t ii = 0; ii < n; ii++) {
    double fl[32], gl[32], feq[32], geq[32], vel[3] = {0, 0, 0};
    for (int p = 0; p < g.Q; ++p) fl[p] = f[p][ii];
    for (int p = 0; p < g.Q; ++p) gl[p] = g_[p][ii];
    double density = 0.0;
    for (int p = 0; p < g.Q; ++p) density += fl[p];
    if (density < 1e-10) err |= 1;
    rho[ii] = density;
    velocity(g, fl, density, vel);
    // calculateTemperature (AuxiliaryCollisionFunctions.h:289-307)
    double temperature = 0.0;
    for (int i = 0; i < g.Q; i++) {
      double sum = 0.0;
      for (int a = 0; a < g.D; a++) sum += (g.e[i][a] - vel[a]) * (g.e[i][a] - vel[a]);
      temperature += sum * fl[i] / g.cs2 + gl[i];
    }
    temperature = temperature * 0.5 / (density * C_v);
    T[ii] = temperature;
    for (int j = 0; j < g.D; ++j) u[j][ii] = vel[j] * g.scaling;
    // relaxWithG (CollisionSchemes.h:43-118)
    feq_quartic(g, *h, density, vel, temperature, feq);
    for (int i = 0; i < g.Q; i++) geq[i] = feq[i] * (temperature) * (2.0 * C_v - g.D);
    double fStar[32] = {0}, gStar[32] = {0}, fNeq[32], gNeq[32];
    for (int p = 0; p < g.Q; ++p) {
      fNeq[p] = fl[p] - feq[p];
      gNeq[p] = gl[p] - geq[p];
    }
    if (cp.prandtl_set) {
      double Qt[3][3][3] = {};
      double Fg[3] = {0, 0, 0};
      for (int i = 0; i < g.Q; i++)
        for (int a = 0; a < g.D; a++)
          for (int b = 0; b < g.D; b++)
            for (int c = 0; c < g.D; c++)
              Qt[a][b][c] += ((g.e[i][a] - vel[a]) * (g.e[i][b] - vel[b]) * (g.e[i][c] - vel[c])) * fNeq[i];
      for (int i = 0; i < g.Q; i++)
        for (int a = 0; a < g.D; a++) Fg[a] += (g.e[i][a] - vel[a]) * gNeq[i];
      const double cs6 = 6.0 * g.cs2 * g.cs2 * g.cs2;
      for (int a = 0; a < g.D; a++)
        for (int b = 0; b < g.D; b++)
          for (int c = 0; c < g.D; c++)
            for (int i = 0; i < g.Q; i++)
              fStar[i] += g.w[i] *
                          (Qt[a][b][c] * (g.e[i][a] * g.e[i][b] * g.e[i][c] - 3 * g.cs2 * g.e[i][c] * eye(a, b))) /
                          cs6;
      for (int a = 0; a < g.D; a++)
        for (int i = 0; i < g.Q; i++) gStar[i] += g.w[i] * (Fg[a] * g.e[i][a]) / temperature;
    }
    double sutherland_factor = 1.0;
    if (cp.sutherland_set) sutherland_factor = std::pow(temperature / 0.85, 0.7);
    const double visc_tau = (g.tau - 0.5) * sutherland_factor / (temperature * density) + 0.5;
    double kn = 0.0;  // calculateKnudsenNumberEstimate: only feeds the shock mask
    for (int i = 0; i < g.Q; i++) kn += std::fabs(fl[i] - feq[i]) / g.w[i];
    kn = kn / g.Q;
    mask[ii] = kn;
    const double prandtl_tau = (visc_tau - 0.5) / cp.prandtl + 0.5;
    const double visc_omega = 1. / visc_tau;
    const double prandtl_omega = 1. / prandtl_tau;
    const double prandtl_diff = visc_omega - prandtl_omega;
    for (int p = 0; p < g.Q; ++p) {
      fl[p] -= visc_omega * fNeq[p] - prandtl_diff * fStar[p];
      gl[p] -= visc_omega * gNeq[p] - prandtl_diff * gStar[p];
    }
    for (int p = 0; p < g.Q; ++p) {
      f[p][ii] = fl[p];
      g_[p][ii] = gl[p];
    }
  }
  delete h;
  return err;
}

void init_equilibrium(const Stencil& st, int64_t n, const double* rho, const double* const* u, double* const* f) {
  // BGKStandard::getEquilibriumDistribution (collision/BGKStandard.cpp:23-41), scaled stencil
#pragma omp parallel for schedule(static)
  for (int64_t ii = 0; ii < n; ii++) {
    double uu = 0;
    for (int d = 0; d < st.D; d++) uu += u[d][ii] * u[d][ii];
    for (int i = 0; i < st.Q; i++) {
      double prefactor = st.w[i] * rho[ii];
      double uSquareTerm = -(uu) / (2 * st.cs2);
      if (i == 0) {
        f[i][ii] = prefactor * (1 + uSquareTerm);
        continue;
      }
      double ue = 0;
      for (int d = 0; d < st.D; d++) ue += u[d][ii] * st.dir[i][d];
      double mixedTerm = (ue) / st.cs2;
      f[i][ii] = prefactor * (1 + mixedTerm * (1 + 0.5 * (mixedTerm)) + uSquareTerm);
    }
  }
}

void init_compressible(const Stencil& st, const CollisionParams& cp, int64_t n, const double* rho,
                       const double* const* u, const double* T, double* const* f, double* const* g) {
  // CompressibleCFDSolver.h:790-848 (calcQuarticEquilibrium) and :930-943
  const GenData gd = make_gendata(st, cp);
  const int D = st.D;
  const double cs2 = gd.cs2;
  const double C_v = 1. / (cp.gamma - 1.0);
#pragma omp parallel for schedule(static)
  for (int64_t ii = 0; ii < n; ii++) {
    double vel[3] = {0, 0, 0};
    for (int j = 0; j < D; j++) vel[j] = u[j][ii] / cp.scaling;
    const double density = rho[ii], temperature = T[ii];
    double uu_term = 0.0;
    for (int j = 0; j < D; j++) uu_term += -(vel[j] * vel[j]) / (2.0 * cs2);
    for (int i = 0; i < st.Q; i++) {
      const double* e = gd.e[i];
      const double w = gd.w[i];
      double T1 = cs2 * (temperature - 1);
      double ue = 0.0;
      for (int j = 0; j < D; j++) ue += (vel[j] * e[j]) / cs2;
      double feq = w * density * (1 + ue * (1 + 0.5 * (ue)) + uu_term);
      for (int alp = 0; alp < D; alp++)
        for (int bet = 0; bet < D; bet++) {
          feq += density * w / (2.0 * cs2) *
                 ((temperature - 1) * eye(alp, bet) * e[alp] * e[bet] - cs2 * eye(alp, bet) * (temperature - 1));
          for (int gam = 0; gam < D; gam++) {
            feq += w * density / (6. * cs2 * cs2 * cs2) *
                   (vel[alp] * vel[bet] * vel[gam] +
                    T1 * (eye(alp, bet) * vel[gam] + eye(bet, gam) * vel[alp] + eye(alp, gam) * vel[bet])) *
                   (e[alp] * e[bet] * e[gam] -
                    cs2 * (e[gam] * eye(alp, bet) + e[bet] * eye(alp, gam) + e[alp] * eye(bet, gam)));
            for (int det = 0; det < D; det++) {
              double power4 = e[alp] * e[bet] * e[gam] * e[det];
              double power2 = e[alp] * e[bet] * eye(gam, det) + e[alp] * e[gam] * eye(bet, det) +
                              e[alp] * e[det] * eye(bet, gam) + e[bet] * e[gam] * eye(alp, det) +
                              e[bet] * e[det] * eye(alp, gam) + e[gam] * e[det] * eye(alp, bet);
              double power0 = eye(alp, bet) * eye(gam, det) + eye(alp, gam) * eye(bet, det) + eye(alp, det) * eye(bet, gam);
              double u4 = vel[alp] * vel[bet] * vel[gam] * vel[det];
              double u2 = vel[alp] * vel[bet] * eye(gam, det) + vel[alp] * vel[gam] * eye(bet, det) +
                          vel[alp] * vel[det] * eye(bet, gam) + vel[bet] * vel[gam] * eye(alp, det) +
                          vel[bet] * vel[det] * eye(alp, gam) + vel[gam] * vel[det] * eye(alp, bet);
              double multieye = eye(alp, bet) * eye(gam, det) + eye(alp, gam) * eye(bet, det) + eye(alp, det) * eye(bet, gam);
              feq += w * density / (24. * cs2 * cs2 * cs2 * cs2) * (power4 - cs2 * power2 + cs2 * cs2 * power0) *
                     (u4 + T1 * (u2 + T1 * multieye));
            }
          }
        }
      f[i][ii] = feq;
      g[i][ii] = feq * (temperature) * (2.0 * C_v - D);
    }
  }
}

void quartic_feq(const Stencil& st, double rho, const double u[3], double T, double* feq) {
  CollisionParams cp;
  cp.scaling = st.scaling;
  cp.viscosity = 1;
  cp.dt = 1;
  const GenData g = make_gendata(st, cp);
  HTensors* h = new HTensors;
  make_H(g, *h);
  feq_quartic(g, *h, rho, u, T, feq);
  delete h;
}

void stabilizer_matrix(const Stencil& st, std::vector<double>& P) {
  const int Q = st.Q, D = st.D;
  const double s = st.scaling, cs2 = 1.0 / 3.0;
  // Hermite polynomials of order <= 2 at the unscaled velocities
  std::vector<std::vector<double>> H;
  std::vector<double> h(Q);
  for (int i = 0; i < Q; i++) h[i] = 1.0;
  H.push_back(h);
  for (int a = 0; a < D; a++) {
    for (int i = 0; i < Q; i++) h[i] = st.dir[i][a] / s;
    H.push_back(h);
  }
  for (int a = 0; a < D; a++)
    for (int b = a; b < D; b++) {
      for (int i = 0; i < Q; i++) h[i] = (st.dir[i][a] / s) * (st.dir[i][b] / s) - (a == b ? cs2 : 0.0);
      H.push_back(h);
    }
  const int K = (int)H.size();
  // the basis is w-orthogonal on D2Q9 and D3Q19, so G is diagonal
  std::vector<double> g(K, 0.0);
  for (int k = 0; k < K; k++)
    for (int i = 0; i < Q; i++) g[k] += st.w[i] * H[k][i] * H[k][i];
  P.assign((size_t)Q * Q, 0.0);
  for (int i = 0; i < Q; i++)
    for (int j = 0; j < Q; j++) {
      double v = 0;
      for (int k = 0; k < K; k++) v += H[k][i] * H[k][j] / g[k];
      P[(size_t)i * Q + j] = st.w[i] * v;
    }
}

void apply_stabilizer(const Stencil& st, const std::vector<double>& P, int64_t n, double* const* f) {
  const int Q = st.Q;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n; i++) {
    double in[32], out[32];
    for (int q = 0; q < Q; q++) in[q] = f[q][i];
    for (int r = 0; r < Q; r++) {
      double v = 0;
      for (int q = 0; q < Q; q++) v += P[(size_t)r * Q + q] * in[q];
      out[r] = v;
    }
    for (int q = 0; q < Q; q++) f[q][i] = out[q];
  }
}

}  // namespace orc
