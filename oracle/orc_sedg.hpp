// Oracle: SEDG (Min-Lee) operator on FE_DGQ(p) with GLL collocation, and the
// low-storage 5-stage Runge-Kutta integrator. TEST INFRASTRUCTURE ONLY.
//
// Reference (src/library/natrium/):
//   advection/SEDGMinLee.cpp:92-184   reassemble (cell loop, alpha loop)
//   advection/SEDGMinLee.cpp:321-358  mass / derivative matrices (collocated at GLL points)
//   advection/SEDGMinLee.cpp:361-412  face loop incl. periodic faces
//   advection/SEDGMinLee.cpp:414-458  -(e . D) M^-1 cell block
//   advection/SEDGMinLee.cpp:461-536  upwind / central face terms
//   advection/SEDGMinLee.cpp:361-412, boundaries/VelocityNeqBounceBack.cpp:46-135
//                                      velocity-neq bounce-back wall faces
//   timeintegration/RungeKutta5LowStorage.h:29-46, .cpp:56-105
#pragma once
#include <vector>

#include "orc_core.hpp"
#include "orc_sl.hpp"

namespace orc {

struct DGSpace {
  Mesh mesh;
  int p = 1;
  int nl = 1;  // nodes per cell
  std::vector<double> gll, gllw;
  Lagrange1D lag;
  std::vector<std::array<double, 3>> x;  // node coordinates, cell-major
  void init(const Mesh& m, int order);
  int64_t n_dofs() const { return (int64_t)mesh.n_cells() * nl; }
};

// per-direction CSR operator L_alpha (alpha = 1..Q-1): block[alpha-1]
void assemble_sedg(const DGSpace& V, const Stencil& st, bool central_flux, SLOperator& op);

// One RK5 low-storage step of df/dt = L f over all streamed blocks.
// f: Q-1 arrays (directions 1..Q-1) of n. work: 2*(Q-1)*n scratch.
void rk5_step(const SLOperator& op, double dt, double* const* f, std::vector<double>& work);

extern const double RK5_A[5];
extern const double RK5_B[5];

}  // namespace orc
