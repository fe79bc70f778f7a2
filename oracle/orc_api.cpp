// Oracle C API: a CFDSolver-like object over the FP64 restatement.
// TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
// bench.py's cpu_baseline leg as the checker / CPU baseline, never by the product.
//
// Loop order follows solver/CFDSolver.cpp:877-902 (run: collide(); loop{stream(); collide();})
// and solver/CFDSolver.cpp:659-754 (stream: f_tmp = f; f_stream = A f_tmp_stream; t += dt),
// solver/CompressibleCFDSolver.h:1006-1051 (stream f, stream g, collide).
// dt: utilities/CFDSolverUtilities.cpp:92-99.
// Sums: solver/PhysicalProperties.cpp:29-130 (kinetic energy), :270-310 (mass).
#include <omp.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "orc_collide.hpp"
#include "orc_core.hpp"
#include "orc_sedg.hpp"
#include "orc_sl.hpp"

using namespace orc;

struct orc_solver {
  Stencil st;
  Mesh mesh;
  int p = 1;
  int scheme = 0;  // 0 SL, 1 SEDG upwind, 2 SEDG central
  bool compressible = false;
  CollisionParams cp;
  double cfl = 0.4, dt = 0, time = 0;
  FEQSpace feq;
  DGSpace dg;
  SLOperator op;
  int64_t n = 0;
  std::vector<double> f, g, rho, u, T, mask, work, ftmp;
  bool stabilizer = false;  // PseudoEntropicStabilizer data processor
  std::vector<double> stab_P;
  std::string err;
  std::vector<double*> fptr(std::vector<double>& v, int Q) {
    std::vector<double*> r(Q);
    for (int i = 0; i < Q; i++) r[i] = v.data() + (size_t)i * n;
    return r;
  }
};

extern "C" {

orc_solver* orc_create(int dim, int stencil_id, int p, const int* ncells, const double* vertices, int cartesian,
                       const int* periodic, double scaling, double viscosity, double cfl, int scheme, int compressible,
                       double gamma) {
  auto* s = new orc_solver;
  try {
    s->st = make_stencil(stencil_id, scaling);
    if (s->st.D != dim) throw std::runtime_error("stencil dimension mismatch");
    Mesh& m = s->mesh;
    m.D = dim;
    for (int a = 0; a < 3; a++) {
      m.N[a] = a < dim ? ncells[a] : 1;
      m.periodic[a] = a < dim ? (periodic ? periodic[a] != 0 : true) : true;
    }
    m.cartesian = cartesian != 0;
    const int nz = dim == 3 ? m.N[2] + 1 : 1;
    const size_t nv = (size_t)(m.N[0] + 1) * (m.N[1] + 1) * nz;
    m.vert.resize(nv);
    for (size_t i = 0; i < nv; i++)
      for (int d = 0; d < 3; d++) m.vert[i][d] = d < dim ? vertices[i * dim + d] : 0.0;
    s->p = p;
    s->scheme = scheme;
    s->compressible = compressible != 0;
    s->cfl = cfl;
    s->cp.scaling = scaling;
    s->cp.viscosity = viscosity;
    s->cp.gamma = gamma;
    if (scheme == 0) {
      s->feq.init(m, p);
      s->n = s->feq.n_dofs();
    } else {
      m.cartesian = false;  // DG always uses the Q1 mapping (AdvectionOperator.cpp:59-62)
      s->dg.init(m, p);
      s->n = s->dg.n_dofs();
    }
    // CFDSolverUtilities::calculateTimestep
    double dx = m.min_vertex_distance();
    double umax = s->st.max_speed;
    s->dt = cfl * dx / (umax * p * p);
    s->cp.dt = s->dt;
    const int Q = s->st.Q;
    s->f.assign((size_t)Q * s->n, 0.0);
    if (s->compressible) s->g.assign((size_t)Q * s->n, 0.0);
    s->rho.assign(s->n, 0.0);
    s->u.assign((size_t)dim * s->n, 0.0);
    s->T.assign(s->n, 1.0);
    s->mask.assign(s->n, 0.0);
  } catch (std::exception& e) {
    s->err = e.what();
  }
  return s;
}

// Compressible options of the reference configuration (SolverConfiguration
// "Prandtl number" and "Sutherland law"): prandtl <= 0 leaves it unset.
void orc_set_compressible_options(orc_solver* s, double prandtl, int sutherland) {
  s->cp.prandtl_set = prandtl > 0;
  s->cp.prandtl = prandtl > 0 ? prandtl : 1.0;
  s->cp.sutherland_set = sutherland != 0;
}

const char* orc_error(orc_solver* s) { return s->err.c_str(); }
void orc_destroy(orc_solver* s) { delete s; }
int64_t orc_n_dofs(orc_solver* s) { return s->n; }
double orc_dt(orc_solver* s) { return s->dt; }
double orc_time(orc_solver* s) { return s->time; }
double orc_tau(orc_solver* s) { return s->cp.viscosity / (s->dt * s->st.cs2) + 0.5; }
int orc_Q(orc_solver* s) { return s->st.Q; }
void orc_set_threads(int n) { omp_set_num_threads(n); }
int orc_get_threads() { return omp_get_max_threads(); }

void orc_node_coords(orc_solver* s, double* xyz) {
  for (int64_t i = 0; i < s->n; i++)
    for (int d = 0; d < 3; d++) xyz[i * 3 + d] = s->scheme == 0 ? s->feq.x[i][d] : s->dg.x[i][d];
}

// setDeltaT + reassemble (SemiLagrangian.h:175-179, SemiLagrangian.cpp:76-90)
int orc_assemble(orc_solver* s, double dt) {
  try {
    if (dt > 0) {
      s->dt = dt;
      s->cp.dt = dt;
    }
    if (s->scheme == 0)
      assemble_sl(s->feq, s->st, s->dt, s->op);
    else
      assemble_sedg(s->dg, s->st, s->scheme == 2, s->op);
  } catch (std::exception& e) {
    s->err = e.what();
    return 1;
  }
  return 0;
}

int64_t orc_op_nnz(orc_solver* s, int alpha) { return s->op.block[alpha - 1].rowptr[s->n]; }
int64_t orc_op_identity_rows(orc_solver* s) { return s->op.n_identity_rows; }

// Velocity of the velocity-neq bounce-back wall on each box face (6 x 3, face
// 2*axis+side), used on non-periodic axes; set before orc_assemble.
void orc_set_walls(orc_solver* s, const double* wall_u) {
  for (int f = 0; f < 6; f++)
    for (int d = 0; d < 3; d++) s->mesh.wall_u[f][d] = wall_u[3 * f + d];
  if (s->scheme == 0)
    s->feq.mesh = s->mesh;
  else
    s->dg.mesh = s->mesh;
}

// Boundary hit list counts of the last assembly (SemiLagrangianBoundaryHandler::n_hits, n_cells)
void orc_op_hits(orc_solver* s, int64_t* n_hits, int64_t* n_cells) {
  *n_hits = s->op.n_hits;
  *n_cells = s->op.n_hit_cells;
}

// Source direction of every entry of block alpha (returns 0 and writes nothing when
// all entries read direction alpha), and the boundary values added after the product.
int orc_op_src(orc_solver* s, int alpha, int8_t* src) {
  const CSR& A = s->op.block[alpha - 1];
  if (A.src.empty()) return 0;
  std::memcpy(src, A.src.data(), A.src.size());
  return 1;
}
int orc_op_bvec(orc_solver* s, int alpha, double* b) {
  const std::vector<double>& v = s->op.bvec[alpha - 1];
  if (v.empty()) return 0;
  std::memcpy(b, v.data(), sizeof(double) * v.size());
  return 1;
}
// SEDG walls: diagonal of the block (alpha, opposite alpha) (returns 0 without walls)
int orc_op_opp(orc_solver* s, int alpha, double* d) {
  if (s->op.opp_diag.empty()) return 0;
  const std::vector<double>& v = s->op.opp_diag[alpha - 1];
  std::memcpy(d, v.data(), sizeof(double) * v.size());
  return 1;
}
void orc_op_get(orc_solver* s, int alpha, int64_t* rowptr, int64_t* col, double* val) {
  const CSR& A = s->op.block[alpha - 1];
  std::memcpy(rowptr, A.rowptr.data(), sizeof(int64_t) * (s->n + 1));
  std::memcpy(col, A.col.data(), sizeof(int64_t) * A.col.size());
  std::memcpy(val, A.val.data(), sizeof(double) * A.val.size());
}

// departure cell and unit coordinates of (dof, alpha)
int orc_locate(orc_solver* s, int64_t dof, int alpha, double* xi) {
  try {
    return locate_departure(s->feq, s->st, s->dt, dof, alpha, xi);
  } catch (std::exception& e) {
    s->err = e.what();
    return -2;
  }
}

// initial f from (rho, u[, T]) at the DoFs: CFDSolver.cpp:1104-1129 or CompressibleCFDSolver.h:850-946
void orc_init(orc_solver* s, const double* rho, const double* u, const double* T) {
  const int Q = s->st.Q, D = s->st.D;
  std::memcpy(s->rho.data(), rho, sizeof(double) * s->n);
  std::memcpy(s->u.data(), u, sizeof(double) * s->n * D);
  std::vector<const double*> up(D);
  for (int d = 0; d < D; d++) up[d] = s->u.data() + (size_t)d * s->n;
  auto fp = s->fptr(s->f, Q);
  if (s->compressible) {
    std::memcpy(s->T.data(), T, sizeof(double) * s->n);
    auto gp = s->fptr(s->g, Q);
    init_compressible(s->st, s->cp, s->n, s->rho.data(), up.data(), s->T.data(), fp.data(), gp.data());
  } else {
    init_equilibrium(s->st, s->n, s->rho.data(), up.data(), fp.data());
  }
  s->time = 0;
}

int orc_collide(orc_solver* s) {
  const int Q = s->st.Q, D = s->st.D;
  auto fp = s->fptr(s->f, Q);
  std::vector<double*> up(D);
  for (int d = 0; d < D; d++) up[d] = s->u.data() + (size_t)d * s->n;
  int e;
  if (s->compressible) {
    auto gp = s->fptr(s->g, Q);
    e = collide_compressible(s->st, s->cp, s->n, fp.data(), gp.data(), s->rho.data(), up.data(), s->T.data(),
                             s->mask.data());
  } else {
    e = collide_bgk(s->st, s->cp, s->n, fp.data(), s->rho.data(), up.data());
  }
  if (e) s->err = "Densities too small (< 1e-10) for collisions. Decrease time step size.";
  return e;
}

// selectCollision(..., init = true) (CollisionOperator.h:79-91): the stored velocity
// (replaced by u first when u is not null) drives the relaxation; incompressible only.
int orc_collide_init(orc_solver* s, const double* u) {
  if (s->compressible) {
    s->err = "initialisation collision: incompressible only";
    return -1;
  }
  const int Q = s->st.Q, D = s->st.D;
  if (u) std::memcpy(s->u.data(), u, sizeof(double) * s->n * D);
  auto fp = s->fptr(s->f, Q);
  std::vector<double*> up(D);
  for (int d = 0; d < D; d++) up[d] = s->u.data() + (size_t)d * s->n;
  const int e = collide_bgk(s->st, s->cp, s->n, fp.data(), s->rho.data(), up.data(), true);
  if (e) s->err = "Densities too small (< 1e-10) for collisions. Decrease time step size.";
  return e;
}

// One SL / SEDG streaming of a population vector. For SL, with_boundary adds the
// velocity-neq bounce-back values (applyBoundaryConditions after the product,
// CFDSolver.cpp:669-721); the g populations of the compressible solver get the
// product only, since applyBoundaryConditionsToG acts on velocity-equilibrium
// boundaries alone (CompressibleCFDSolver.h:279-314,
// SemiLagrangianBoundaryHandler.cpp:111-158).
static void stream_vec(orc_solver* s, std::vector<double>& v, bool with_boundary) {
  const int Q = s->st.Q;
  const int64_t n = s->n;
  if (s->scheme == 0) {
    s->ftmp = v;  // DistributionFunctions f_tmp(m_f)
    sl_apply(s->op, s->ftmp.data(), v.data(), with_boundary);
  } else {
    std::vector<double*> fp(Q - 1);
    for (int a = 1; a < Q; a++) fp[a - 1] = v.data() + (size_t)a * n;
    rk5_step(s->op, s->dt, fp.data(), s->work);
  }
}

int orc_stream(orc_solver* s) {
  stream_vec(s, s->f, true);
  if (s->compressible) stream_vec(s, s->g, false);
  s->time += s->dt;
  return 0;
}

// Append the pseudo-entropic stabilizer as a data processor (CFDSolver::appendDataProcessor):
// applied after the collision of every loop iteration, for D2Q9 and D3Q19 only
// (PseudoEntropicStabilizer.cpp:271-289 does nothing for other stencils).
void orc_set_stabilizer(orc_solver* s, int on) {
  s->stabilizer = on != 0;
  if (s->stabilizer) stabilizer_matrix(s->st, s->stab_P);
}

// n full steps (stream + collide + data processors, CFDSolver.cpp:877-892); the caller runs
// the initial collide() once.
int orc_steps(orc_solver* s, int nsteps) {
  for (int i = 0; i < nsteps; i++) {
    orc_stream(s);
    if (orc_collide(s)) return 1;
    if (s->stabilizer && (s->st.Q == 9 || s->st.Q == 19)) {
      auto fp = s->fptr(s->f, s->st.Q);
      apply_stabilizer(s->st, s->stab_P, s->n, fp.data());
    }
  }
  return 0;
}

void orc_stabilizer_matrix(int id, double scaling, double* P) {
  Stencil st = make_stencil(id, scaling);
  std::vector<double> m;
  stabilizer_matrix(st, m);
  std::memcpy(P, m.data(), sizeof(double) * m.size());
}

void orc_get_f(orc_solver* s, double* f) { std::memcpy(f, s->f.data(), sizeof(double) * s->f.size()); }
void orc_set_f(orc_solver* s, const double* f) { std::memcpy(s->f.data(), f, sizeof(double) * s->f.size()); }
void orc_get_g(orc_solver* s, double* g) { std::memcpy(g, s->g.data(), sizeof(double) * s->g.size()); }
void orc_set_g(orc_solver* s, const double* g) { std::memcpy(s->g.data(), g, sizeof(double) * s->g.size()); }
void orc_get_macro(orc_solver* s, double* rho, double* u, double* T, double* mask) {
  std::memcpy(rho, s->rho.data(), sizeof(double) * s->n);
  std::memcpy(u, s->u.data(), sizeof(double) * s->u.size());
  if (T) std::memcpy(T, s->T.data(), sizeof(double) * s->n);
  if (mask) std::memcpy(mask, s->mask.data(), sizeof(double) * s->n);
}

// mass = sum_cells sum_q rho_q JxW_q ; E_kin = 1/2 sum rho |u|^2 JxW (GLL collocation)
void orc_sums(orc_solver* s, double* mass, double* ekin) {
  const Mesh& m = s->mesh;
  const int D = m.D, np = s->p + 1;
  std::vector<double> gll, gllw;
  gll_01(s->p, gll, gllw);
  const int nl = D == 3 ? np * np * np : np * np;
  double msum = 0, esum = 0;
  for (int c = 0; c < m.n_cells(); c++) {
    int cc[3];
    m.cell_coords(c, cc);
    for (int l = 0; l < nl; l++) {
      int mi[3] = {l % np, (l / np) % np, D == 3 ? l / (np * np) : 0};
      double xi[3] = {gll[mi[0]], gll[mi[1]], D == 3 ? gll[mi[2]] : 0.0};
      double J[3][3];
      Mesh mm = m;
      if (s->scheme != 0) mm.cartesian = false;
      mm.jacobian(c, xi, J);
      double det = D == 2 ? J[0][0] * J[1][1] - J[0][1] * J[1][0]
                          : J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                                J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                                J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
      double jxw = det * gllw[mi[0]] * gllw[mi[1]] * (D == 3 ? gllw[mi[2]] : 1.0);
      int64_t i = s->scheme == 0 ? s->feq.cell_dof(c, mi[0], mi[1], mi[2]) : (int64_t)c * nl + l;
      double uu = 0;
      for (int d = 0; d < D; d++) uu += s->u[(size_t)d * s->n + i] * s->u[(size_t)d * s->n + i];
      msum += s->rho[i] * jxw;
      esum += s->rho[i] * uu * jxw;
    }
  }
  *mass = msum;
  *ekin = 0.5 * esum;
}

// faceCrossedFirst on a given mesh (test hook); returns face id, fills p_b and lambda
int orc_face_crossed_first(orc_solver* s, int cell, const double* pin, const double* pout, double* pb, double* lambda,
                           int without_mapping) {
  double a[3] = {pin[0], pin[1], s->mesh.D == 3 ? pin[2] : 0.0};
  double b[3] = {pout[0], pout[1], s->mesh.D == 3 ? pout[2] : 0.0};
  try {
    return face_crossed_first(s->mesh, cell, a, b, pb, lambda, without_mapping != 0);
  } catch (std::exception& e) {
    s->err = e.what();
    return -3;
  }
}

// Stencil and basis tables (test hooks)
void orc_stencil(int id, double scaling, double* dirs, double* w, int* opp, double* cs2) {
  Stencil st = make_stencil(id, scaling);
  for (int i = 0; i < st.Q; i++) {
    for (int d = 0; d < 3; d++) dirs[i * 3 + d] = st.dir[i][d];
    w[i] = st.w[i];
    opp[i] = st.opp[i];
  }
  *cs2 = st.cs2;
}
void orc_gll(int p, double* x, double* w) {
  std::vector<double> xs, ws;
  gll_01(p, xs, ws);
  for (int i = 0; i <= p; i++) {
    x[i] = xs[i];
    w[i] = ws[i];
  }
}
void orc_lagrange(int p, int j, int n, const double* xi, double* val, double* der) {
  std::vector<double> xs, ws;
  gll_01(p, xs, ws);
  Lagrange1D L;
  L.init(xs);
  for (int i = 0; i < n; i++) {
    val[i] = L.value(j, xi[i]);
    der[i] = L.deriv(j, xi[i]);
  }
}
void orc_quartic_feq(int id, double scaling, double rho, const double* u, double T, double* feq) {
  Stencil st = make_stencil(id, scaling);
  double uu[3] = {u[0], u[1], st.D == 3 ? u[2] : 0.0};
  quartic_feq(st, rho, uu, T, feq);
}
// one RK5 step of the scalar ODE f' = lambda f (RungeKutta5LowStorage_test.cpp:25-59)
double orc_rk5_scalar(double lambda, double f0, double dt, int nsteps) {
  SLOperator op;
  op.Q = 2;
  op.n = 1;
  op.block.resize(1);
  op.block[0].rowptr = {0, 1};
  op.block[0].col = {0};
  op.block[0].val = {lambda};
  std::vector<double> w;
  double f = f0;
  double* fp[1] = {&f};
  for (int i = 0; i < nsteps; i++) rk5_step(op, dt, fp, w);
  return f;
}

}  // extern "C"
