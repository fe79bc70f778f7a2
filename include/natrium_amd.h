/* natrium_amd: MI355X (gfx950) time-stepping hot path of a high-order
 * semi-Lagrangian / SEDG lattice Boltzmann solver, behind a plain C ABI.
 *
 * Every entry point returns 0 on success or a NAT_ERR_* code; the message of the
 * last failure on a context is available from nat_last_error(). No exception and
 * no C++ type crosses this boundary. Populations and operators stay in device
 * memory between calls; each context owns one HIP stream and calls on different
 * contexts may run from different threads. Calls on one context must not overlap.
 * nat_step captures a HIP graph of its steps the first time it runs more than 32 steps
 * from a given ping-pong state; the HIP runtime fails a capture, and the other call, when
 * another thread meanwhile issues work that synchronises with the legacy default stream
 * (creating or initialising a context, copies), so threads that do that concurrently
 * serialise it against nat_step calls of that length (the request server does).
 *
 * Array conventions (host pointers unless the name says device):
 *   DoF numbering   SL:   lattice nodes (N_a p + 1 per axis), x fastest; nodes on
 *                         opposite periodic faces are distinct DoFs, as in the
 *                         reference's unconstrained FE_Q space.
 *                   SEDG: cell-major, cell c (x fastest) then local GLL node
 *                         (x fastest), (p+1)^D nodes per cell.
 *   populations     f[alpha * n_dofs + i], alpha = 0..Q-1 (structure of arrays).
 *   velocities      u[d * n_dofs + i], in physical units (u_lattice * scaling).
 *   vertices        (N_0+1)(N_1+1)[(N_2+1)] points, x fastest, dim doubles each.
 */
#ifndef NATRIUM_AMD_H
#define NATRIUM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* stencils (stencils/D2Q9.cpp, D3Q19.cpp, D3Q27.cpp, D2Q25H.cpp of the reference) */
enum { NAT_D2Q9 = 0, NAT_D3Q19 = 1, NAT_D3Q27 = 2, NAT_D2Q25H = 3 };
/* advection schemes (solver/SolverConfiguration.cpp:45 "Advection scheme") */
enum { NAT_SL = 0, NAT_SEDG = 1, NAT_SEDG_CENTRAL = 2 };
/* arithmetic of the bulk populations */
enum { NAT_FP64 = 0, NAT_FP32_SHIFTED = 1 };
/* equilibria (collision_advanced/Equilibria.h) */
enum { NAT_EQ_BGK = 0, NAT_EQ_QUARTIC = 1 };
/* SL backend selection */
enum { NAT_BACKEND_AUTO = 0, NAT_BACKEND_STRUCTURED = 1, NAT_BACKEND_GENERAL = 2 };

enum {
  NAT_OK = 0,
  NAT_ERR_ARG = 1,          /* invalid argument / descriptor */
  NAT_ERR_UNSUPPORTED = 2,  /* combination not supported by this build */
  NAT_ERR_HIP = 3,          /* HIP runtime failure */
  NAT_ERR_DENSITY = 4,      /* density < 1e-10 in collision (DensityZeroException) */
  NAT_ERR_STATE = 5,        /* call out of order (e.g. stepping before set_dt) */
  NAT_ERR_NOMEM = 6         /* device allocation failed */
};

typedef struct nat_ctx nat_ctx;

typedef struct nat_desc {
  int dim;               /* 2 or 3 */
  int stencil;           /* NAT_D2Q9 ... */
  int order;             /* FE order p >= 1 */
  int ncells[3];         /* cells per axis */
  const double* vertices;/* lattice vertices, see conventions above */
  int cartesian;         /* 1: axis-aligned cells (MappingCartesian); 0: trilinear map */
  int periodic[3];       /* periodic flag per axis; both faces of a non-periodic axis are
                            velocity-neq bounce-back walls (SL scheme only) */
  double scaling;        /* stencil scaling s (|c_axis| = s) */
  double viscosity;      /* kinematic viscosity nu */
  double cfl;            /* CFL number; dt = CFL h_min / (sqrt(2) s p^2) */
  int scheme;            /* NAT_SL, NAT_SEDG or NAT_SEDG_CENTRAL */
  int precision;         /* NAT_FP64 or NAT_FP32_SHIFTED */
  int equilibrium;       /* NAT_EQ_BGK or NAT_EQ_QUARTIC */
  int compressible;      /* 1: f and g populations, quartic equilibrium, FP64 temperature */
  double gamma;          /* heat capacity ratio (compressible) */
  double rho0;           /* reference density of the FP32 shift f - w rho0 (0 -> 1.0) */
  int backend;           /* NAT_BACKEND_* (SL only) */
  double prandtl;        /* compressible: Prandtl number (0 = not set, unit Prandtl) */
  int sutherland;        /* compressible: 1 = Sutherland viscosity law (T/0.85)^0.7 */
  double wall_velocity[6][3]; /* velocity of the wall on box face 2*axis+side (side 0: lower),
                                 constant in time; read on non-periodic axes. Replaces
                                 VelocityNeqBounceBack(boundary id, velocity)
                                 (boundaries/VelocityNeqBounceBack.cpp:15-38) with the boundary
                                 ids of a colorized box */
} nat_desc;

/* Context lifetime. Replaces the CFDSolver<dim> constructor's set-up of the
 * stencil, DoF handler and distribution vectors (solver/CFDSolver.cpp:77-255). */
int nat_create(const nat_desc* desc, nat_ctx** out);
void nat_destroy(nat_ctx* ctx);
const char* nat_last_error(const nat_ctx* ctx);

int64_t nat_n_dofs(const nat_ctx* ctx);
int nat_n_stored_nodes(const nat_ctx* ctx, int64_t* n); /* nodes held on the device */
int nat_Q(const nat_ctx* ctx);
double nat_dt(const nat_ctx* ctx);     /* CFDSolverUtilities::calculateTimestep (CFDSolverUtilities.cpp:92-99) */
double nat_tau(const nat_ctx* ctx);    /* calculateTauFromNu (AuxiliaryCollisionFunctions.h:63-68) */
double nat_time(const nat_ctx* ctx);
int nat_backend(const nat_ctx* ctx);   /* NAT_BACKEND_STRUCTURED or NAT_BACKEND_GENERAL for SL */
int nat_node_coords(nat_ctx* ctx, double* xyz /* n_dofs x 3 */);

/* AdvectionOperator::setDeltaT + reassemble (SemiLagrangian.h:175-179,
 * SemiLagrangian.cpp:76-90; SEDGMinLee.cpp:92-184). Departure points and basis
 * weights are located on the GPU in FP64. dt <= 0 keeps the CFL time step. */
int nat_set_dt(nat_ctx* ctx, double dt);

/* Moved mesh, same lattice topology: the transformed triangulation followed by
 * AdvectionOperator::reassemble (SemiLagrangian.cpp:76-90, the departure walk of
 * SemiLagrangian.cpp:145-757 through MappingQ1 in FP64 on the GPU). The populations stay
 * attached to their DoFs and dt is kept. A cartesian = 1 context needs axis-aligned cells;
 * the structured backend needs a uniform Cartesian mesh. SEDG contexts read the new
 * geometry in their next stage. */
int nat_set_vertices(nat_ctx* ctx, const double* vertices);

/* CFDSolver::appendDataProcessor(PseudoEntropicStabilizer) (solver/CFDSolver.h:69-373,
 * dataprocessors/PseudoEntropicStabilizer.cpp:271-311): on = 1 applies the stabilizer after the
 * collision of every nat_step iteration (not after nat_collide, as run() applies data
 * processors after its loop collide only). Like the reference it acts on D2Q9 and D3Q19 and
 * leaves other stencils unchanged. */
int nat_set_stabilizer(nat_ctx* ctx, int on);

/* CFDSolver::initializeDistributions (CFDSolver.cpp:1104-1129; compressible:
 * CompressibleCFDSolver.h:850-946): f = equilibrium(rho, u[, T]); T may be NULL. */
int nat_init_equilibrium(nat_ctx* ctx, const double* rho, const double* u, const double* T);

/* DistributionFunctions copies (solver/DistributionFunctions.h:47-300): FP64 host arrays. */
int nat_upload_f(nat_ctx* ctx, const double* f);
int nat_download_f(nat_ctx* ctx, double* f);
int nat_upload_g(nat_ctx* ctx, const double* g);
int nat_download_g(nat_ctx* ctx, double* g);

/* CFDSolver::collide (CFDSolver.cpp:807-843 -> selectCollision ->
 * CollisionOperator::collideAll, CollisionOperator.h:26-224) with init == false. Writes rho,
 * u (and T). The collision arithmetic is that of the fused steps: nat_stream followed by
 * nat_collide gives the populations of nat_step(ctx, 1) bit for bit (without a stabilizer),
 * except in FP32 compressible contexts, whose fused step keeps the streamed f and g in FP64
 * registers where nat_stream stores them in FP32. */
int nat_collide(nat_ctx* ctx);

/* selectCollision(..., init == true) (CollisionSelection.h:61-67 -> collideAll,
 * CollisionOperator.h:79-91), the collision of the iterative initialisation
 * (CFDSolver.cpp:1269-1311): rho is computed and written, the velocity is read from the
 * stored field instead of computed, and u is not written. u (dim x n_dofs, physical units)
 * replaces the stored field first; NULL keeps it. Incompressible contexts only. */
int nat_collide_init(nat_ctx* ctx, const double* u);

/* CFDSolver::stream (CFDSolver.cpp:659-754) alone: for SL, f = A f_old with the wall values
 * of applyBoundaryConditions (SemiLagrangian.h:150-161; compressible contexts stream g as
 * well, CompressibleCFDSolver.h:181-314); for SEDG, one low-storage RK5 step
 * (RungeKutta5LowStorage.cpp:56-105). f_0 is unchanged; time advances by dt. Lets a caller
 * interpose filter() between stream() and collide() (CFDSolver.cpp:884-891). */
int nat_stream(nat_ctx* ctx);

/* PseudoEntropicStabilizer::apply (dataprocessors/PseudoEntropicStabilizer.cpp:271-311) as a
 * pass of its own, for a caller that runs the loop body as nat_stream, nat_collide and the
 * data processors. D2Q9 and D3Q19 only; other stencils are left unchanged, as in the
 * reference. (nat_set_stabilizer fuses the same projection into nat_step.) */
int nat_stabilize(nat_ctx* ctx);

/* nsteps iterations of the CFDSolver::run loop body (CFDSolver.cpp:884-888):
 * stream(); collide(); with stream and collide fused on the device. The caller
 * runs the initial collide() of run() once with nat_collide(). */
int nat_step(nat_ctx* ctx, int nsteps);

/* The same operations with one entry point per path, for callers that bind a single
 * path; each returns NAT_ERR_ARG when the context is of another kind.
 *   nat_sl_assemble: SemiLagrangian::setDeltaT + reassemble (SemiLagrangian.h:175-179,
 *     SemiLagrangian.cpp:76-90,145-757) of an SL context (incompressible or compressible);
 *     dt <= 0 keeps the CFL time step. As nat_set_dt.
 *   nat_sl_bgk_step: nsteps of the run() loop body with the SL stream and the BGK collision
 *     (CFDSolver.cpp:669-721,807-843,884-888) of an incompressible SL context. As nat_step.
 *   nat_sedg_rk5_step: nsteps of the loop body whose stream is one low-storage RK5 step of
 *     the SEDG operator (RungeKutta5LowStorage.cpp:56-105, SEDGMinLee.cpp:321-536). As nat_step.
 *   nat_compressible_step: nsteps of CompressibleCFDSolver::run's loop body
 *     (CompressibleCFDSolver.h:181-314,739-1040): stream f and g with the SL operator, then
 *     the compressible collision. As nat_step. */
int nat_sl_assemble(nat_ctx* ctx, double dt);
int nat_sl_bgk_step(nat_ctx* ctx, int nsteps);
int nat_sedg_rk5_step(nat_ctx* ctx, int nsteps);
int nat_compressible_step(nat_ctx* ctx, int nsteps);

/* Measurement helper with no counterpart in the reference (its solver prints
 * "million DoF updates per second" from clock(), CFDSolver.cpp:975-982): runs
 * nat_step(ctx, nsteps) between two HIP events recorded on the context's stream
 * and returns the elapsed device time in milliseconds. */
int nat_step_timed(nat_ctx* ctx, int nsteps, double* elapsed_ms);

/* macroscopic fields written by the last collision (FP64 host arrays; T may be NULL,
 * and must be NULL for incompressible contexts) */
int nat_get_macro(nat_ctx* ctx, double* rho, double* u, double* T);

/* Knudsen-number estimate sum_i |f_i - f_eq,i| / w_i / Q of the last compressible
 * collision, the field behind the reference's shock mask (calculateKnudsenNumberEstimate,
 * AuxiliaryCollisionFunctions.h:474-481; maskShockSensor, CollisionSchemes.h:90-100). */
int nat_get_knudsen(nat_ctx* ctx, double* kn);

/* PhysicalProperties::mass / kineticEnergy (PhysicalProperties.cpp:29-130, 270-310),
 * plus total momentum, reduced on the device in FP64 from the last collision's fields. */
int nat_sums(nat_ctx* ctx, double* mass, double* momentum /* dim */, double* ekin);

/* nat_sums without any device allocation. No counterpart in the reference: it is for callers
 * that cannot allocate (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc). The partial sums go to a fixed per-GPU scratch of NAT_SUMS_SCRATCH_BLOCKS x 5
 * doubles (2.5 MiB) held until the process ends, instead of the context's lazily allocated
 * buffer; meshes with more than 16,777,216 (cell, node) pairs return NAT_ERR_ARG. The result
 * is bit-identical to nat_sums. Calls that overlap on one GPU (two streams or threads at
 * once) share the scratch and race: serialise them. */
int nat_sums_noalloc(nat_ctx* ctx, double* mass, double* momentum /* dim */, double* ekin);

int nat_sync(nat_ctx* ctx);

/* Launch shape of the structured SL step kernel, no counterpart in the reference: the node
 * tile (x, y, z extents; z = 1 in 2D) one workgroup of the fused stream + collide kernel
 * owns on this mesh. */
int nat_struct_tile(nat_ctx* ctx, int* tile /* 3 */);

/* Packed stepping, no counterpart in the reference (which runs one solve per process): for
 * callers that advance many small solves, e.g. a request server. nat_step_packed runs nsteps
 * fused stream + collide steps of each of the n contexts (1 <= n <= 16) as one kernel launch
 * per step that advances all of them, on the first context's stream (after the pending work
 * of every member; their later work waits for it). Members must be distinct, on one GPU, and
 * share a packing key: structured SL contexts (uniform periodic boxes) at p = 2-4 with the same
 * dimension, stencil, order, storage precision and tile. nat_pack_key returns that key, or
 * NAT_ERR_UNSUPPORTED for a context that cannot be packed. Every member's populations and
 * fields are bit-identical to those of nat_step(member, nsteps); an error of member i is
 * reported on the first context with the prefix "context i". */
int nat_pack_key(nat_ctx* ctx, int64_t* key);
int nat_step_packed(nat_ctx* const* ctxs, int n, int nsteps);

/* Launch-shape autotuner, no counterpart in the reference: times every launch shape the SL
 * step kernel is built with for this (stencil, order) on the context's own mesh, nsteps steps
 * each (nsteps <= 0: 10), and keeps the fastest for this context and, process-wide, for later
 * contexts of the same kernel instantiation and mesh size (which then skip the timing). The
 * timing launches write only the idle population buffer: populations, fields and time are
 * unchanged, and results do not depend on the shape. SL contexts (not compressible):
 *   structured backend: the tiles; tile receives the chosen (x, y, z) extents;
 *   general backend: with FP32 storage in 3D the six-wave form and the DEEP form (4 waves,
 *     more source rows and the next departure record in flight; by default chosen when its
 *     grid fits one round of the device); tile receives the chosen form's (DoFs per block,
 *     waves per SIMD it is built for, z-planes of source rows loaded together). */
int nat_autotune(nat_ctx* ctx, int nsteps, int* tile /* 3, may be NULL */);

/* Pins the general SL backend's kernel form, no counterpart in the reference: form -1 the
 * default rule (DEEP when its grid fits one round of the device), 0 the six-wave form, 1 the
 * DEEP form (FP32 storage in 3D; other contexts have one form and accept only -1 and 0). Both
 * forms give bit-identical results; the choice only moves the step time. NAT_ERR_ARG for
 * another value or backend. */
int nat_set_general_form(nat_ctx* ctx, int form);

/* ---- Output and restart (host side, from fields downloaded through this API) ---- */

/* One row of the reference's results table: SolverStats::update/printNewLine
 * (solver/SolverStats.h:102-175) with PhysicalProperties (solver/PhysicalProperties.cpp:29-420):
 * max |u| is the largest component-wise maximum, enstrophy is int (dv/dx - du/dy)^2 in 2D and
 * int |grad u|_F^2 in 3D on the GLL quadrature (enstrophy_sq the integral of its square),
 * entropy is -sum_cells (sum_q JxW_q) sum_i sum_j f_j ln(f_j / w_j), pressures are
 * (rho - 1) c_s^2, mean(ux) the volume-weighted mean of the cell DoF averages. Residuals are
 * formed at iterations divisible by 100 against the fields of the previous such call, once 100
 * iterations have passed since the first (SolverStats::calculateResiduals, which the
 * reference's stop check calls every 100 iterations, CFDSolver.cpp:927-931); otherwise, and on
 * a second call at the same iteration, the last values are repeated (1.0 before the first
 * residual, CFDSolver.cpp:398-399). */
typedef struct nat_stats_t {
  int64_t iteration;
  double time, max_u, mass, kin_e, enstrophy, entropy, max_p, min_p, res_rho, res_u, mean_ux, enstrophy_sq;
} nat_stats_t;
int nat_stats(nat_ctx* ctx, int64_t iteration, nat_stats_t* out);
/* Appends the row to the table file (columns as results_table.txt); truncate = 1 starts a
 * new file with the header line. */
int nat_write_stats(nat_ctx* ctx, const char* path, int64_t iteration, int truncate);

/* CFDSolver::output solution files (solver/CFDSolver.cpp:1001-1069): dir/t_0.<it>.vtu with one
 * patch per cell sampled like DataOut::build_patches(p + 1) (p = 1: one sub-cell), point data
 * rho, ux, uy[, uz][, T] and cell data subdomain, plus the dir/t_0.<it>.pvtu record. */
int nat_write_vtu(nat_ctx* ctx, const char* dir, int64_t iteration);

/* Checkpoint::write / load (solver/Checkpoint.cpp:21-140): dir/checkpoint_<it>.stat holds the
 * iteration, time, stencil scaling and FE order, one per line, as the reference writes it;
 * dir/checkpoint_<it>.data holds the FP64 populations in DoF order (the reference stores a
 * p4est/deal.II serialization there). Compressible contexts add checkpointG_<it>.* for g.
 * Reading with iteration == 1 picks the most recent checkpoint, as in the reference; it sets
 * the context time and returns the stored iteration. Another order or scaling returns
 * NAT_ERR_UNSUPPORTED (the reference interpolates and rescales). */
int nat_checkpoint_write(nat_ctx* ctx, const char* dir, int64_t iteration);
int nat_checkpoint_read(nat_ctx* ctx, const char* dir, int64_t iteration, int64_t* iteration_out);

/* ------------------------------------------------------------------------
 * Request server. No counterpart in the reference (one solve per process):
 * continuous batching of independent solves on one GPU. Requests are admitted
 * in arrival order while their device footprint fits the HBM budget and an
 * active slot is free, and built on setup threads. Active requests advance in
 * chunks of fused steps: alone on their own stream as replays of a captured HIP
 * graph, or, when several idle requests share a packing key, together in packed
 * launches (one launch per step for all of them), so requests overlap on the
 * device and a retired request's memory admits the next one. A request's results
 * do not depend on the other requests in flight. All functions are thread-safe.
 * ------------------------------------------------------------------------ */
enum { NAT_PROBLEM_TGV2D = 0, NAT_PROBLEM_TGV3D = 1, NAT_PROBLEM_SHOCK_TUBE = 2 };

typedef struct nat_request {
  int problem;     /* NAT_PROBLEM_*: TGV on [0,2pi]^D, or a density step (1 | 0.8) on a
                      periodic 1 x 1/4 strip */
  int stencil;     /* NAT_D2Q9 (2D problems), NAT_D3Q19 or NAT_D3Q27 (TGV3D) */
  int refinement;  /* 2^refinement cells along x */
  int order;       /* FE order p */
  double reynolds; /* Re = U L / nu with the benchmark's length (2 pi in TGV2D, 1 otherwise) */
  double mach;     /* Ma = U / c_s, U = 1 */
  double cfl;      /* 0 -> 0.4 */
  int nsteps;      /* fused stream+collide steps after the initial collision */
  int precision;   /* NAT_FP64 or NAT_FP32_SHIFTED */
} nat_request;

typedef struct nat_result {
  int status;              /* NAT_OK or the NAT_ERR_* the solve ended with */
  int64_t n_dofs;
  double queue_ms;         /* submit -> admitted */
  double run_ms;           /* admitted -> retired (set-up, steps, read-back) */
  double latency_ms;       /* submit -> retired */
  double mass, ekin;       /* FP64 conservation sums after the last step */
  int64_t footprint_bytes; /* device bytes reserved for the request */
  char error[256];         /* message of a failed request, empty otherwise */
} nat_result;

typedef struct nat_server_opts {
  int64_t hbm_budget_bytes; /* <= 0: 80% of the free device memory at creation */
  int max_active;           /* concurrently resident requests (<= 0: 8) */
  int chunk_steps;          /* steps per scheduling chunk (<= 0: 50) */
  int use_graphs;           /* 1: replay one captured HIP graph per chunk */
  int pack;                 /* 1: idle requests with equal nat_pack_key advance together in
                               packed launches (nat_step_packed), up to 16 per launch */
  int autotune;             /* 1: pick each structured request's tile with nat_autotune */
  int setup_threads;        /* threads building admitted requests (mesh, operator, initial
                               state) beside the stepping (<= 0: 4) */
} nat_server_opts;

typedef struct nat_server_stats_t {
  int64_t queued, active, used_bytes;
  int64_t node_updates;  /* DoF updates of completed requests */
  double busy_ms;        /* first admission -> last retirement */
  int64_t packed_chunks; /* chunks issued as packed launches of several requests */
} nat_server_stats_t;

typedef struct nat_server nat_server;

int nat_server_create(const nat_server_opts* opts, nat_server** out);
int nat_server_submit(nat_server* s, const nat_request* r, int64_t* id);
/* blocks until request id is retired; copies its velocity field (dim x n_dofs,
 * physical units) into u when u_len > 0; the id is released afterwards */
int nat_server_wait(nat_server* s, int64_t id, nat_result* res, double* u, int64_t u_len);
int nat_server_stats(nat_server* s, nat_server_stats_t* st);
/* finishes every queued request, then stops the scheduler */
void nat_server_destroy(nat_server* s);

#ifdef __cplusplus
}
#endif

#endif /* NATRIUM_AMD_H */
