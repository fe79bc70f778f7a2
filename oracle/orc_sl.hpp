// Oracle: FE_Q lattice space and semi-Lagrangian operator assembly.
// TEST INFRASTRUCTURE ONLY.
//
// Reference: advection/SemiLagrangian.cpp:145-522 (fillSparseObject),
//            advection/SemiLagrangian.cpp:525-757 (faceCrossedFirst[WithoutMapping]),
//            advection/SemiLagrangianTools.cpp:15-64 (shapeFunctionValue),
//            boundaries/PeriodicBoundary.cpp:184-310 (coordinatesAcrossPeriodicBoundary),
//            advection/SemiLagrangian.cpp:358-385 (velocity-neq bounce-back: reverse the path),
//            boundaries/VelocityNeqBounceBack.cpp:137-195 (value added per wall hit),
//            boundaries/SemiLagrangianBoundaryHandler.cpp:14-158 (hit list; g untouched).
#pragma once
#include <cstdint>
#include <vector>

#include "orc_core.hpp"

namespace orc {

// Continuous FE_Q(p) space on a lattice mesh. DoFs are the lattice nodes
// (N_a p + 1 per axis, lexicographic, x fastest); nodes on opposite periodic
// faces are separate DoFs (no periodicity constraints), as in the reference.
struct FEQSpace {
  Mesh mesh;
  int p = 1;
  int n[3] = {1, 1, 1};
  std::vector<double> gll, gllw;
  Lagrange1D lag;
  std::vector<std::array<double, 3>> x;  // support point of every DoF
  std::vector<int> owner;                // first cell (iteration order) holding the DoF

  void init(const Mesh& m, int order);
  int64_t n_dofs() const { return (int64_t)n[0] * n[1] * n[2]; }
  int64_t dof(int i, int j, int k) const { return i + (int64_t)n[0] * (j + (int64_t)n[1] * k); }
  // global DoF of local node (m0,m1,m2) of a cell
  int64_t cell_dof(int cell, int m0, int m1, int m2) const;
  int n_local() const { return mesh.D == 3 ? (p + 1) * (p + 1) * (p + 1) : (p + 1) * (p + 1); }
  // tensor-product shape values of a cell's local DoFs (lexicographic) at unit point xi
  void shape_values(const double xi[3], double* out) const;
};

// SL operator: one CSR matrix per streamed destination direction alpha=1..Q-1.
// Columns are DoFs; `src` holds the source direction of each entry when a wall
// couples alpha with its opposite (block (alpha, opp alpha) of the reference's
// block matrix), and is empty when every entry reads direction alpha.
struct CSR {
  std::vector<int64_t> rowptr;
  std::vector<int64_t> col;
  std::vector<double> val;
  std::vector<int8_t> src;
};
struct SLOperator {
  int Q = 0;
  int64_t n = 0;
  std::vector<CSR> block;  // block[alpha-1]
  int64_t n_identity_rows = 0;  // rows written by the tracker life-time limit
  // Velocity-neq bounce-back: per destination direction, the values added to the
  // streamed f after the product (empty when no path reaches a wall), and the
  // hit-list counts of the boundary handler (hits, distinct cells holding hits).
  std::vector<std::vector<double>> bvec;
  int64_t n_hits = 0, n_hit_cells = 0;
  // SEDG with velocity-neq bounce-back walls (VelocityNeqBounceBack::assembleBoundary): the
  // diagonal of the block (alpha, opposite alpha) per destination block (empty without walls),
  // the block index of the opposite direction, and bvec above as the system vector b
  std::vector<std::vector<double>> opp_diag;
  std::vector<int> opp_block;
};

struct WallHit {
  int face;  // box face 2*axis+side
  int cell;  // cell in which the hit lies
};

// Returns the crossed face id or -1 (point inside). Throws on inverse-map failure
// when without_mapping is false (caller then retries without mapping).
int face_crossed_first(const Mesh& mesh, int cell, const double p_in[3], const double p_out[3], double p_b[3],
                       double* lambda, bool without_mapping);

// Neighbour across face f; returns -1 at a non-periodic boundary. *wrapped is set
// when the face is a periodic boundary face.
int neighbor_cell(const Mesh& mesh, int cell, int f, bool* wrapped);

// PeriodicBoundary::coordinatesAcrossPeriodicBoundary for a cell at a periodic face f.
void across_periodic(const Mesh& mesh, int cell, int f, const double p[3], double out[3]);

// Locate the departure point of DoF `dof` for direction alpha (walk from its
// owning cell). Returns the final cell (or -1 when the life-time limit hit) and
// the unit coordinates there; *src_dir receives the direction the path ends in
// (alpha, or its opposite after an odd number of wall reflections) and `hits`
// the walls met on the way.
int locate_departure(const FEQSpace& V, const Stencil& st, double dt, int64_t dof, int alpha, double xi[3],
                     int* src_dir = nullptr, std::vector<WallHit>* hits = nullptr);

void assemble_sl(const FEQSpace& V, const Stencil& st, double dt, SLOperator& op);

// y = A_alpha x for one block whose entries all read x (no wall coupling)
void csr_vmult(const CSR& A, const double* x, double* y, int64_t n);

// Streaming of all directions: y_alpha = sum_entries val x_{src}[col] (+ bvec_alpha
// when with_boundary), x and y laid out direction-major (Q x n); y_0 is not touched.
void sl_apply(const SLOperator& op, const double* x, double* y, bool with_boundary);

}  // namespace orc
