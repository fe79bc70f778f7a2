// C++ host layer of natrium_amd: the reference's solver-side interfaces, with the
// populations and operators resident on one MI355X behind the C ABI of
// include/natrium_amd.h. Header-only; link against libnatrium_amd.so.
//
// Mirrors, without deal.II or Trilinos types:
//   - the exceptions of utilities/NATriuMException.h:22-75 (DensityZeroException,
//     AuxiliaryCollisionFunctions.h:28; SemiLagrangianException, SemiLagrangian.h:43;
//     CFDSolverException, CFDSolver.h:46; ...), thrown where the C ABI returns an error;
//   - DistributionFunctions (solver/DistributionFunctions.h:47-300): Q host vectors, at(i);
//   - SemiLagrangian<dim> (advection/SemiLagrangian.h:145-179): setupDoFs, setDeltaT,
//     reassemble, stream(f_old, f, t) -> dt with f_old = f; f = A f_old; boundaries;
//   - SEDGMinLee<dim> + RungeKutta5LowStorage::step(f, A, b, t, dt) -> t + dt
//     (timeintegration/RungeKutta5LowStorage.cpp:56-105);
//   - selectCollision(..., init) (collision_advanced/CollisionSelection.h:61-67,116-122) on
//     host vectors, including the init == true form (CollisionOperator.h:79-91);
//   - CFDSolver<dim> (solver/CFDSolver.h:69-373) with run() in the loop order of
//     CFDSolver.cpp:877-892 (collide once; then output, stream, filter, collide, data
//     processors until a stop condition of CFDSolver.cpp:905-945 holds), and
//     CompressibleCFDSolver<dim>::run (CompressibleCFDSolver.h:1006-1040).
//
// Two kinds of call: the host-vector forms (stream(f_old, f, t), selectCollision) copy
// populations over PCIe on every call, exactly like a caller that keeps its own
// DistributionFunctions; CFDSolver keeps them on the device and, between two outputs,
// hands whole runs of loop iterations to nat_step (stream and collide fused). Without a
// stabilizer that gives the populations of the one-operation-at-a-time loop bit for bit
// (FP32-shifted compressible contexts excepted: their fused step keeps the streamed values
// in FP64); with the stabilizer the fused projection rounds differently, and both forms
// match the reference's loop to the FP64 parity bound.
#pragma once

#include <cmath>
#include <cstdint>
#include <functional>
#include <limits>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "natrium_amd.h"

namespace natrium_amd {

// ---------------------------------------------------------------- exceptions
class NATriuMException : public std::exception {
 public:
  explicit NATriuMException(std::string msg, int code = NAT_ERR_ARG) : msg_(std::move(msg)), code_(code) {}
  const char* what() const noexcept override { return msg_.c_str(); }
  int code() const noexcept { return code_; }

 private:
  std::string msg_;
  int code_;
};
class NotImplementedException : public NATriuMException {
  using NATriuMException::NATriuMException;
};
class CollisionException : public NATriuMException {
  using NATriuMException::NATriuMException;
};
class DensityZeroException : public CollisionException {
  using CollisionException::CollisionException;
};
class AdvectionSolverException : public NATriuMException {
  using NATriuMException::NATriuMException;
};
class SemiLagrangianException : public AdvectionSolverException {
  using AdvectionSolverException::AdvectionSolverException;
};
class TimeIntegratorException : public NATriuMException {
  using NATriuMException::NATriuMException;
};
class CFDSolverException : public NATriuMException {
  using NATriuMException::NATriuMException;
};
class DistributionFunctionsException : public NATriuMException {
  using NATriuMException::NATriuMException;
};

// Error codes of the C ABI to the exception a reference caller expects at that seam.
enum class Seam { Collision, Advection, TimeIntegration, Solver };

[[noreturn]] inline void rethrow(int rc, const std::string& msg, Seam where) {
  if (rc == NAT_ERR_DENSITY) throw DensityZeroException(msg, rc);
  if (rc == NAT_ERR_UNSUPPORTED) throw NotImplementedException(msg, rc);
  switch (where) {
    case Seam::Collision: throw CollisionException(msg, rc);
    case Seam::Advection: throw SemiLagrangianException(msg, rc);
    case Seam::TimeIntegration: throw TimeIntegratorException(msg, rc);
    default: throw CFDSolverException(msg, rc);
  }
}

// ---------------------------------------------------------------- device context
// Owns one nat_ctx (populations, operator, stream on the device).
class GpuContext {
 public:
  explicit GpuContext(const nat_desc& d) {
    const int rc = nat_create(&d, &ctx_);
    if (rc) {
      const std::string m = ctx_ ? nat_last_error(ctx_) : "nat_create failed";
      if (ctx_) nat_destroy(ctx_);
      ctx_ = nullptr;
      rethrow(rc, m, Seam::Solver);
    }
    desc_ = d;
    desc_.vertices = nullptr;  // the caller's array need not outlive the call
  }
  ~GpuContext() {
    if (ctx_) nat_destroy(ctx_);
  }
  GpuContext(const GpuContext&) = delete;
  GpuContext& operator=(const GpuContext&) = delete;

  nat_ctx* get() const { return ctx_; }
  const nat_desc& desc() const { return desc_; }
  size_t n_dofs() const { return (size_t)nat_n_dofs(ctx_); }
  size_t Q() const { return (size_t)nat_Q(ctx_); }
  int dim() const { return desc_.dim; }
  bool compressible() const { return desc_.compressible != 0; }
  void check(int rc, Seam where) const {
    if (rc) rethrow(rc, nat_last_error(ctx_), where);
  }

 private:
  nat_ctx* ctx_ = nullptr;
  nat_desc desc_{};
};

// ---------------------------------------------------------------- populations on the host
// Q vectors of n FP64 values, f_0 first, each direction contiguous (the layout of the
// reference's f_0 vector plus the fStream block vector, DistributionFunctions.h:47-300).
class DistributionFunctions {
 public:
  DistributionFunctions() = default;
  DistributionFunctions(size_t Q, size_t n) { reinit(Q, n); }
  void reinit(size_t Q, size_t n) {
    Q_ = Q;
    n_ = n;
    data_.assign(Q * n, 0.0);
  }
  size_t getQ() const { return Q_; }
  size_t size() const { return n_; }
  double* at(size_t i) {
    if (i >= Q_) throw DistributionFunctionsException("direction index out of range");
    return data_.data() + i * n_;
  }
  const double* at(size_t i) const {
    if (i >= Q_) throw DistributionFunctionsException("direction index out of range");
    return data_.data() + i * n_;
  }
  double* data() { return data_.data(); }
  const double* data() const { return data_.data(); }
  DistributionFunctions& operator=(const DistributionFunctions&) = default;

 private:
  size_t Q_ = 0, n_ = 0;
  std::vector<double> data_;
};

// ---------------------------------------------------------------- advection operators
// AdvectionOperator<dim> for the semi-Lagrangian scheme (SemiLagrangian.h:145-179).
template <size_t dim>
class SemiLagrangian {
 public:
  explicit SemiLagrangian(std::shared_ptr<GpuContext> c) : c_(std::move(c)) {
    if (c_->desc().scheme != NAT_SL) throw SemiLagrangianException("SemiLagrangian needs an SL context");
    if ((size_t)c_->dim() != dim) throw SemiLagrangianException("dimension of the context differs");
  }
  // DoFs are the lattice nodes of the context: nothing to distribute; the support points
  // are available from getSupportPoints().
  void setupDoFs() {}
  // SemiLagrangian::setDeltaT: the new time step, then the operator is assembled again on
  // the device (departure walk in FP64). A step the structured tables cannot take leaves
  // the previous operator in use and throws.
  void setDeltaT(double dt) { c_->check(nat_sl_assemble(c_->get(), dt), Seam::Advection); }
  double getDeltaT() const { return nat_dt(c_->get()); }
  // reassemble() with the current time step and mesh
  void reassemble() { c_->check(nat_set_dt(c_->get(), getDeltaT()), Seam::Advection); }
  // reassemble() after a mesh transformation: new vertex coordinates, same lattice
  void reassemble(const std::vector<double>& vertices) {
    c_->check(nat_set_vertices(c_->get(), vertices.data()), Seam::Advection);
  }
  std::vector<double> getSupportPoints() const {
    std::vector<double> xyz(3 * c_->n_dofs());
    c_->check(nat_node_coords(c_->get(), xyz.data()), Seam::Advection);
    return xyz;
  }
  // stream(f_old, f, t): f_old = f; f = A f_old with the wall values; returns dt. The
  // host vectors go to the device and back.
  double stream(DistributionFunctions& f_old, DistributionFunctions& f, double t) {
    if (&f_old == &f) throw SemiLagrangianException("stream: f_old and f must differ");
    (void)t;  // the wall velocities are constant in time
    f_old = f;
    c_->check(nat_upload_f(c_->get(), f.data()), Seam::Advection);
    c_->check(nat_stream(c_->get()), Seam::Advection);
    c_->check(nat_download_f(c_->get(), f.data()), Seam::Advection);
    return getDeltaT();
  }
  // the same on the device-resident populations of the context
  double stream() {
    c_->check(nat_stream(c_->get()), Seam::Advection);
    return getDeltaT();
  }
  GpuContext& context() { return *c_; }

 private:
  std::shared_ptr<GpuContext> c_;
};

// SEDG operator (SEDGMinLee.cpp:92-536), applied matrix-free on the device; it is the
// "systemMatrix" argument of RungeKutta5LowStorage::step.
template <size_t dim>
class SEDGMinLee {
 public:
  explicit SEDGMinLee(std::shared_ptr<GpuContext> c) : c_(std::move(c)) {
    if (c_->desc().scheme == NAT_SL) throw AdvectionSolverException("SEDGMinLee needs an SEDG context");
    if ((size_t)c_->dim() != dim) throw AdvectionSolverException("dimension of the context differs");
  }
  void setupDoFs() {}
  void reassemble() {}  // the operator is rebuilt from the vertices in every stage
  void reassemble(const std::vector<double>& vertices) {
    c_->check(nat_set_vertices(c_->get(), vertices.data()), Seam::Advection);
  }
  double getDeltaT() const { return nat_dt(c_->get()); }
  void setDeltaT(double dt) { c_->check(nat_set_dt(c_->get(), dt), Seam::Advection); }
  GpuContext& context() const { return *c_; }

 private:
  std::shared_ptr<GpuContext> c_;
};

// The system vector b of the reference's step: the SEDG boundary contributions. Periodic
// meshes (the only SEDG meshes built) have none.
struct SystemVector {
  bool empty() const { return true; }
};

// RungeKutta5LowStorage (Carpenter-Kennedy, RungeKutta5LowStorage.{h:29-46,cpp:56-105}).
class RungeKutta5LowStorage {
 public:
  explicit RungeKutta5LowStorage(double timeStepSize) : dt_(timeStepSize) {}
  double getTimeStepSize() const { return dt_; }
  void setTimeStepSize(double dt) { dt_ = dt; }
  // One step of f' = A f + b on the host vectors f: dt == 0 keeps the step size, another
  // value replaces it (and the operator's time step). Returns t + dt.
  template <size_t dim>
  double step(DistributionFunctions& f, const SEDGMinLee<dim>& A, const SystemVector& b, double t, double dt) {
    if (!b.empty()) throw TimeIntegratorException("SEDG boundary vectors are not supported");
    GpuContext& c = A.context();
    if (dt != 0.0 && dt != dt_) dt_ = dt;
    if (dt == 0.0) dt = dt_;
    if (nat_dt(c.get()) != dt_) c.check(nat_set_dt(c.get(), dt_), Seam::TimeIntegration);
    c.check(nat_upload_f(c.get(), f.data()), Seam::TimeIntegration);
    c.check(nat_stream(c.get()), Seam::TimeIntegration);
    c.check(nat_download_f(c.get(), f.data()), Seam::TimeIntegration);
    return t + dt;
  }

 private:
  double dt_;
};

// selectCollision(..., init) on host vectors (CollisionSelection.h:61-67): in-place
// update of f; writes rho and u * scaling; with init == true reads u instead of writing it.
inline void selectCollision(GpuContext& c, DistributionFunctions& f, std::vector<double>& rho, std::vector<double>& u,
                            bool init) {
  const size_t n = c.n_dofs();
  if (f.getQ() != c.Q() || f.size() != n) throw CollisionException("populations do not match the context");
  rho.resize(n);
  if (u.size() != (size_t)c.dim() * n) {
    if (init) throw CollisionException("init collision needs the velocity field");
    u.resize((size_t)c.dim() * n);
  }
  c.check(nat_upload_f(c.get(), f.data()), Seam::Collision);
  if (init)
    c.check(nat_collide_init(c.get(), u.data()), Seam::Collision);
  else
    c.check(nat_collide(c.get()), Seam::Collision);
  c.check(nat_download_f(c.get(), f.data()), Seam::Collision);
  std::vector<double> uo((size_t)c.dim() * n);
  c.check(nat_get_macro(c.get(), rho.data(), uo.data(), nullptr), Seam::Collision);
  if (!init) u.swap(uo);
}

// ---------------------------------------------------------------- solver
// The SolverConfiguration entries this path reads (solver/SolverConfiguration.cpp).
struct SolverConfiguration {
  size_t numberOfTimeSteps = 1000000000;  // "Number of time steps"
  double simulationEndTime = 1e10;        // "Simulation end time"
  double convergenceThreshold = 1e-10;    // "Convergence threshold"
  size_t outputTableInterval = 0;         // "Output table interval" (0: no table)
  std::string tableFile;                  // results table path
  size_t outputSolutionInterval = 0;      // "Output solution interval" (0: no VTU)
  std::string outputDirectory;            // VTU files
  size_t outputCheckpointInterval = 0;    // "Output checkpoint interval"
  bool filtering = false;                 // "Filtering" (spectral filter: not built)
  bool fuseSteps = true;                  // hand runs of iterations to nat_step
};

// CFDSolver<dim> (CFDSolver.h:69-373): the loop of run() on device-resident populations.
template <size_t dim>
class CFDSolver {
 public:
  CFDSolver(const nat_desc& d, SolverConfiguration cfg)
      : c_(std::make_shared<GpuContext>(d)), cfg_(std::move(cfg)) {
    if ((size_t)d.dim != dim) throw CFDSolverException("dimension of the descriptor differs");
    if (cfg_.filtering) throw NotImplementedException("the spectral filter is not part of this build");
    m_time = 0.0;
  }
  virtual ~CFDSolver() = default;

  // initializeDistributions (CFDSolver.cpp:1104-1129): f = f_eq(rho, u[, T])
  void initializeDistributions(const std::vector<double>& rho, const std::vector<double>& u,
                               const std::vector<double>* T = nullptr) {
    c_->check(nat_init_equilibrium(c_->get(), rho.data(), u.data(), T ? T->data() : nullptr), Seam::Solver);
  }
  // CFDSolver::stream (CFDSolver.cpp:659-754): one SL stream or one RK5 step; t += dt
  void stream() {
    c_->check(nat_stream(c_->get()), Seam::Advection);
    m_time = nat_time(c_->get());
  }
  void filter() {}  // filtering is rejected at construction
  // CFDSolver::collide (CFDSolver.cpp:807-843) with init == false
  void collide() { c_->check(nat_collide(c_->get()), Seam::Collision); }
  // appendDataProcessor(PseudoEntropicStabilizer)
  void appendStabilizer() { stabilizer_ = true; }
  void applyDataProcessors() {
    if (stabilizer_) c_->check(nat_stabilize(c_->get()), Seam::Solver);
  }
  // run() (CFDSolver.cpp:877-892)
  virtual void run() {
    m_i = m_iterationStart;
    collide();
    while (true) {
      if (stopConditionMet()) break;
      output(m_i, false);
      m_i++;
      const size_t k = fusedRun();
      if (k > 0) {
        // iterations m_i .. m_i + k - 1 have no output and no stop check in between
        stepFused(k);
        m_i += k - 1;
        continue;
      }
      loopBody();
    }
    output(m_i, true);
  }
  // stopConditionMet (CFDSolver.cpp:905-945): iterations, end time, and every 100
  // iterations the velocity residual against the fields of the previous check (1.0 until
  // two checks have been made, CFDSolver.cpp:398-399)
  bool stopConditionMet() {
    if (m_i >= cfg_.numberOfTimeSteps) return true;
    if (m_time >= cfg_.simulationEndTime) return true;
    if (m_i % 100 == 0) {
      nat_stats_t st{};
      c_->check(nat_stats(c_->get(), (int64_t)m_i, &st), Seam::Solver);
      if (st.res_u < cfg_.convergenceThreshold) return true;
    }
    return false;
  }
  // output (CFDSolver.cpp:947-1100): results table, solution files and checkpoints
  void output(size_t it, bool is_final) {
    if (cfg_.outputTableInterval && (it % cfg_.outputTableInterval == 0 || is_final) && !cfg_.tableFile.empty()) {
      c_->check(nat_write_stats(c_->get(), cfg_.tableFile.c_str(), (int64_t)it, tableStarted_ ? 0 : 1), Seam::Solver);
      tableStarted_ = true;
    }
    if (cfg_.outputSolutionInterval && (it % cfg_.outputSolutionInterval == 0 || is_final) &&
        !cfg_.outputDirectory.empty())
      c_->check(nat_write_vtu(c_->get(), cfg_.outputDirectory.c_str(), (int64_t)it), Seam::Solver);
    if (cfg_.outputCheckpointInterval && it % cfg_.outputCheckpointInterval == 0 && !cfg_.outputDirectory.empty() &&
        it > m_iterationStart)
      c_->check(nat_checkpoint_write(c_->get(), cfg_.outputDirectory.c_str(), (int64_t)it), Seam::Solver);
  }

  DistributionFunctions getF() const {
    DistributionFunctions f(c_->Q(), c_->n_dofs());
    c_->check(nat_download_f(c_->get(), f.data()), Seam::Solver);
    return f;
  }
  void setF(const DistributionFunctions& f) { c_->check(nat_upload_f(c_->get(), f.data()), Seam::Solver); }
  void getMacro(std::vector<double>& rho, std::vector<double>& u) const {
    rho.resize(c_->n_dofs());
    u.resize((size_t)dim * c_->n_dofs());
    c_->check(nat_get_macro(c_->get(), rho.data(), u.data(), nullptr), Seam::Solver);
  }
  std::vector<double> getSupportPoints() const {
    std::vector<double> xyz(3 * c_->n_dofs());
    c_->check(nat_node_coords(c_->get(), xyz.data()), Seam::Solver);
    return xyz;
  }
  double getTime() const { return m_time; }
  size_t getIteration() const { return m_i; }
  void setIterationStart(size_t i) { m_iterationStart = i; }
  double getTimeStepSize() const { return nat_dt(c_->get()); }
  double getTau() const { return nat_tau(c_->get()); }
  std::shared_ptr<GpuContext> context() const { return c_; }
  SolverConfiguration& configuration() { return cfg_; }

 protected:
  virtual void loopBody() {
    stream();
    filter();
    collide();
    applyDataProcessors();
  }
  // nat_step fuses stream and collide (and the stabilizer, when set) per iteration
  void stepFused(size_t k) {
    c_->check(nat_set_stabilizer(c_->get(), stabilizer_ ? 1 : 0), Seam::Solver);
    c_->check(nat_step(c_->get(), (int)k), Seam::Solver);
    m_time = nat_time(c_->get());
  }
  // iterations from m_i on that can run fused: up to (excluding) the next iteration with
  // an output, a residual check or the iteration limit; 0 when fusion is off
  size_t fusedRun() const {
    if (!cfg_.fuseSteps) return 0;
    size_t end = cfg_.numberOfTimeSteps;
    auto next_multiple = [&](size_t iv) { return iv ? ((m_i + iv - 1) / iv) * iv : std::numeric_limits<size_t>::max(); };
    end = std::min(end, next_multiple(100));
    end = std::min(end, next_multiple(cfg_.outputTableInterval));
    end = std::min(end, next_multiple(cfg_.outputSolutionInterval));
    end = std::min(end, next_multiple(cfg_.outputCheckpointInterval));
    if (cfg_.simulationEndTime < 1e10) {
      // stop on the first iteration whose start time reaches the end time
      const double dt = nat_dt(c_->get());
      size_t k = 0;
      double t = m_time;
      while (m_i + k < end && t < cfg_.simulationEndTime) {
        t += dt;
        k++;
      }
      end = std::min(end, m_i + k);
    }
    // iterations m_i .. end: the loop checks the stop condition before end + 1
    if (end < m_i) return 0;
    const size_t k = end - m_i + 1;
    return k > (size_t)std::numeric_limits<int>::max() ? (size_t)std::numeric_limits<int>::max() : k;
  }

  std::shared_ptr<GpuContext> c_;
  SolverConfiguration cfg_;
  size_t m_i = 0, m_iterationStart = 0;
  double m_time = 0.0;
  bool stabilizer_ = false;
  bool tableStarted_ = false;
};

// CompressibleCFDSolver<dim>::run (CompressibleCFDSolver.h:1006-1040): stream() moves f and g
// with the one SL operator (gStream), then the compressible collision.
template <size_t dim>
class CompressibleCFDSolver : public CFDSolver<dim> {
 public:
  CompressibleCFDSolver(const nat_desc& d, SolverConfiguration cfg) : CFDSolver<dim>(d, std::move(cfg)) {
    if (!d.compressible) throw CFDSolverException("CompressibleCFDSolver needs a compressible descriptor");
  }
  void gStream() {}  // nat_stream moves g with f
  DistributionFunctions getG() const {
    DistributionFunctions g(this->c_->Q(), this->c_->n_dofs());
    this->c_->check(nat_download_g(this->c_->get(), g.data()), Seam::Solver);
    return g;
  }
  void getTemperature(std::vector<double>& T) const {
    std::vector<double> rho(this->c_->n_dofs()), u((size_t)dim * this->c_->n_dofs());
    T.resize(this->c_->n_dofs());
    this->c_->check(nat_get_macro(this->c_->get(), rho.data(), u.data(), T.data()), Seam::Solver);
  }

 protected:
  void loopBody() override {
    this->stream();
    gStream();
    this->filter();
    this->collide();
    this->applyDataProcessors();
  }
};

}  // namespace natrium_amd
