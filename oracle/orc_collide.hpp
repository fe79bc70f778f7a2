// Oracle collision operators. TEST INFRASTRUCTURE ONLY.
//
// Reference (src/library/natrium/):
//   collision_advanced/CollisionOperator.h:26-111   incompressible collideAll
//   collision_advanced/CollisionOperator.h:113-224  compressible collideAll (f, g)
//   collision_advanced/CollisionSchemes.h:28-41     BGKCollision::relax
//   collision_advanced/CollisionSchemes.h:43-118    BGKCollision::relaxWithG
//   collision_advanced/Equilibria.h:34-83           BGKEquilibrium (D2Q9 specialised + generic)
//   collision_advanced/Equilibria.h:120-267         QuarticEquilibrium::polynomial
//   collision_advanced/AuxiliaryCollisionFunctions.h:46-68,106-202,232-307,421-431,475-481,520-566
//   collision/BGKStandard.cpp:23-41                 legacy equilibrium used for initial f
//   solver/CompressibleCFDSolver.h:790-848,850-1004 compressible initial f, g
//   dataprocessors/PseudoEntropicStabilizer.cpp:271-289 stabilizer after each loop collision
#pragma once
#include <cstdint>
#include <vector>

#include "orc_core.hpp"

namespace orc {

struct CollisionParams {
  double scaling = 1, viscosity = 0, dt = 0, gamma = 1.4;
  bool prandtl_set = false, sutherland_set = false;
  double prandtl = 1.0;
};

// f: Q arrays of n; rho: n; u: D arrays of n (written as u*scaling).
// Returns 0, or 1 if a density fell below 1e-10 (the reference throws).
// init: the initialisation collision (CollisionOperator.h:79-91), u is read (u / scaling
// replaces the computed velocity) and not written.
int collide_bgk(const Stencil& st, const CollisionParams& cp, int64_t n, double* const* f, double* rho,
                double* const* u, bool init = false);

// Compressible BGK with quartic equilibrium on f and g; writes rho, u*scaling, T, mask.
int collide_compressible(const Stencil& st, const CollisionParams& cp, int64_t n, double* const* f, double* const* g,
                         double* rho, double* const* u, double* T, double* mask);

// f = legacy BGKStandard equilibrium (scaled stencil) of (rho, u physical).
void init_equilibrium(const Stencil& st, int64_t n, const double* rho, const double* const* u, double* const* f);

// compressible initial f (quartic, loop form) and g = f T (2 C_v - D).
void init_compressible(const Stencil& st, const CollisionParams& cp, int64_t n, const double* rho,
                       const double* const* u, const double* T, double* const* f, double* const* g);

// quartic equilibrium (polynomial form) for one node, e unscaled, cs2 unscaled
void quartic_feq(const Stencil& st, double rho, const double u[3], double T, double* feq);

// Pseudo-entropic stabilizer of D2Q9 / D3Q19 (PseudoEntropicStabilizer::apply): a fixed
// linear map f <- P f per node, P the projection onto the populations spanned by
// w_i H_k(e_i) for the Hermite polynomials of order <= 2 (1, e_a, e_a e_b - cs2 delta_ab),
// orthogonal in the w-weighted inner product. It keeps density, momentum and the second
// moments and removes the higher (ghost) moments. Built here from the unscaled stencil:
// P = diag(w) H G^-1 H^T with G = H^T diag(w) H. Q x Q row-major.
void stabilizer_matrix(const Stencil& st, std::vector<double>& P);
void apply_stabilizer(const Stencil& st, const std::vector<double>& P, int64_t n, double* const* f);

}  // namespace orc
