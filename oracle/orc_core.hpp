// Oracle core: stencils, Gauss-Lobatto nodes, Lagrange basis, lattice hex mesh
// with Cartesian and trilinear (Q1) mappings.
//
// TEST INFRASTRUCTURE ONLY. This is an FP64 CPU restatement of the NATriuM
// reference algorithm, used as the checker for the HIP library and as the timed
// CPU baseline. Nothing in the product (natrium_amd/) links or calls it.
//
// Reference files followed (paths relative to src/library/natrium/):
//   stencils/D2Q9.cpp:40-61, stencils/D3Q19.cpp:37-60, stencils/D3Q27.cpp:37-110,
//   stencils/D2Q25H.cpp:42-82 (weights/directions), *.h getMaxParticleVelocityMagnitude
//   advection/AdvectionTools.cpp:48-93 (FE_Q / FE_DGQ on QGaussLobatto(p+1) points)
//   advection/AdvectionOperator.cpp:59-62 (MappingCartesian if Cartesian and not DG, else Q1)
//   boundaries/PeriodicBoundary.cpp:184-310 (periodic coordinate transform)
#pragma once
#include <array>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

enum StencilId { ST_D2Q9 = 0, ST_D3Q19 = 1, ST_D3Q27 = 2, ST_D2Q25H = 3 };

struct Stencil {
  int id = 0, D = 0, Q = 0;
  double scaling = 1.0;
  double cs2 = 0.0;                        // scaled: scaling^2/3
  std::vector<std::array<double, 3>> dir;  // scaled directions (z=0 in 2D)
  std::vector<double> w;
  std::vector<int> opp;
  double max_speed = 0.0;                  // sqrt(2)*scaling (reference convention)
};

Stencil make_stencil(int id, double scaling);

// Gauss-Lobatto points/weights on [0,1], n = p+1 points, ascending.
void gll_01(int p, std::vector<double>& x, std::vector<double>& w);

// 1-D Lagrange basis on given nodes: values and derivatives at xi.
struct Lagrange1D {
  int n = 0;
  std::vector<double> nodes, inv_den;
  void init(const std::vector<double>& x);
  double value(int j, double xi) const;
  double deriv(int j, double xi) const;
};

// Logically structured hex/quad mesh (lattice of cells), arbitrary vertex positions.
// Cells are numbered lexicographically (x fastest); vertex v of a cell is
// v = vx + 2 vy + 4 vz (deal.II ordering). Faces: 0:-x 1:+x 2:-y 3:+y 4:-z 5:+z.
struct Mesh {
  int D = 2;
  int N[3] = {1, 1, 1};
  bool periodic[3] = {true, true, true};
  // velocity of the velocity-neq bounce-back wall on box face 2*axis+side, used on
  // non-periodic axes (boundary ids of a colorized hyper-rectangle)
  double wall_u[6][3] = {};
  bool cartesian = true;  // MappingCartesian if true, else MappingQ1
  std::vector<std::array<double, 3>> vert;  // (N0+1)(N1+1)(N2+1) vertices

  int nv(int a) const { return N[a] + 1; }
  int vidx(int i, int j, int k) const { return i + nv(0) * (j + nv(1) * k); }
  int n_cells() const { return N[0] * N[1] * (D == 3 ? N[2] : 1); }
  int cell_index(const int c[3]) const { return c[0] + N[0] * (c[1] + N[1] * (D == 3 ? c[2] : 0)); }
  void cell_coords(int cell, int c[3]) const {
    c[0] = cell % N[0];
    c[1] = (cell / N[0]) % N[1];
    c[2] = D == 3 ? cell / (N[0] * N[1]) : 0;
  }
  int n_cell_vertices() const { return D == 3 ? 8 : 4; }
  const std::array<double, 3>& cell_vertex(int cell, int v) const {
    int c[3];
    cell_coords(cell, c);
    return vert[vidx(c[0] + (v & 1), c[1] + ((v >> 1) & 1), D == 3 ? c[2] + ((v >> 2) & 1) : 0)];
  }

  // forward map unit -> real
  void map(int cell, const double xi[3], double x[3]) const;
  // Jacobian dx/dxi at xi (row a = d x_a / d xi_b)
  void jacobian(int cell, const double xi[3], double J[3][3]) const;
  // inverse map; returns false if the iteration does not converge
  bool inverse_map(int cell, const double x[3], double xi[3]) const;
  double min_vertex_distance() const;
};

Mesh make_box_mesh(int D, const int N[3], const double lo[3], const double hi[3], bool cartesian);

inline void clamp_unit(double& v) {
  if (std::fabs(v) < 1e-10) v = 0;
  if (std::fabs(v - 1) < 1e-10) v = 1;
}

}  // namespace orc
