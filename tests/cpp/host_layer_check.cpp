// Exercises the C++ host layer (natrium_amd/host/natrium_gpu.hpp) on the GPU and writes what
// it computed to raw FP64 files, which tests/test_host_layer_gpu.py compares with the oracle.
//
//   host_layer_check <out_dir>
//
// Cases (all on the 2D Taylor-Green vortex, D2Q9, 8 x 8 cells, p = 4, FP64, Re 10, Ma 0.05):
//   run_loop.f64   CFDSolver<2>::run() for 20 iterations, one operation at a time
//   run_fused.f64  the same with runs of iterations handed to nat_step
//   sl_stream.f64  SemiLagrangian<2>::stream(f_old, f, t) on host vectors after the initial
//                  equilibrium, then selectCollision(init = false)
//   sl_fold.f64    f_old of that stream call (the populations before it)
//   sedg_rk5.f64   SEDG (upwind) on 4 x 4 cells, p = 3: three RungeKutta5LowStorage::step calls
// and prints one line per check of exception behaviour.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../natrium_amd/host/natrium_gpu.hpp"

using namespace natrium_amd;

static void save(const std::string& path, const double* p, size_t n) {
  FILE* fp = std::fopen(path.c_str(), "wb");
  if (!fp || std::fwrite(p, sizeof(double), n, fp) != n) {
    std::fprintf(stderr, "cannot write %s\n", path.c_str());
    std::exit(2);
  }
  std::fclose(fp);
}

static std::vector<double> box(int N, double L) {
  std::vector<double> v;
  for (int j = 0; j <= N; j++)
    for (int i = 0; i <= N; i++) {
      v.push_back(L * i / N);
      v.push_back(L * j / N);
    }
  return v;
}

static nat_desc tgv_desc(const std::vector<double>& v, int N, int p, int scheme) {
  const double Ma = 0.05, cs = 1.0 / Ma;
  nat_desc d{};
  d.dim = 2;
  d.stencil = NAT_D2Q9;
  d.order = p;
  d.ncells[0] = d.ncells[1] = N;
  d.ncells[2] = 1;
  d.vertices = v.data();
  d.cartesian = 1;
  d.periodic[0] = d.periodic[1] = d.periodic[2] = 1;
  d.scaling = std::sqrt(3.0) * cs;
  d.viscosity = 2 * M_PI / 10.0;  // Re = 10 with U = 1, L = 2 pi
  d.cfl = 0.4;
  d.scheme = scheme;
  d.precision = NAT_FP64;
  d.equilibrium = NAT_EQ_BGK;
  d.backend = NAT_BACKEND_AUTO;
  return d;
}

// TaylorGreenVortex2D (TaylorGreenVortex2D.cpp:36-58) at the support points
static void tgv_fields(const std::vector<double>& xyz, std::vector<double>& rho, std::vector<double>& u) {
  const size_t n = xyz.size() / 3;
  const double cs = 1.0 / 0.05;
  rho.assign(n, 1.0);
  u.assign(2 * n, 0.0);
  for (size_t i = 0; i < n; i++) {
    const double X = xyz[3 * i], Y = xyz[3 * i + 1];
    u[i] = std::sin(X) * std::cos(Y);
    u[n + i] = -std::cos(X) * std::sin(Y);
    rho[i] = 1.0 + 0.25 * (std::cos(2 * X) + std::cos(2 * Y)) / (cs * cs);
  }
}

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: host_layer_check <out_dir>\n");
    return 2;
  }
  const std::string out = argv[1];
  const int N = 8, p = 4;
  const std::vector<double> v = box(N, 2 * M_PI);
  std::vector<double> rho, u;

  // CFDSolver::run, one operation at a time and fused
  for (int fused = 0; fused < 2; fused++) {
    SolverConfiguration cfg;
    cfg.numberOfTimeSteps = 20;
    cfg.fuseSteps = fused != 0;
    CFDSolver<2> s(tgv_desc(v, N, p, NAT_SL), cfg);
    tgv_fields(s.getSupportPoints(), rho, u);
    s.initializeDistributions(rho, u);
    s.run();
    const DistributionFunctions f = s.getF();
    save(out + (fused ? "/run_fused.f64" : "/run_loop.f64"), f.data(), f.getQ() * f.size());
    std::printf("run fused=%d iterations=%zu time=%.17g\n", fused, s.getIteration(), s.getTime());
  }

  // SemiLagrangian::stream on host vectors, then selectCollision
  {
    auto c = std::make_shared<GpuContext>(tgv_desc(v, N, p, NAT_SL));
    SemiLagrangian<2> sl(c);
    sl.setupDoFs();
    sl.setDeltaT(0.0);  // keep the CFL step, assemble
    tgv_fields(sl.getSupportPoints(), rho, u);
    if (nat_init_equilibrium(c->get(), rho.data(), u.data(), nullptr)) return 3;
    DistributionFunctions f(c->Q(), c->n_dofs()), f_old;
    if (nat_download_f(c->get(), f.data())) return 3;
    const double dt = sl.stream(f_old, f, 0.0);
    save(out + "/sl_fold.f64", f_old.data(), f_old.getQ() * f_old.size());
    selectCollision(*c, f, rho, u, false);
    save(out + "/sl_stream.f64", f.data(), f.getQ() * f.size());
    std::printf("stream dt=%.17g\n", dt);
    try {
      sl.stream(f, f, 0.0);
      std::printf("same-vector stream: no exception\n");
    } catch (const SemiLagrangianException&) {
      std::printf("same-vector stream: SemiLagrangianException\n");
    }
    // vanishing density: the reference's DensityZeroException
    DistributionFunctions z(c->Q(), c->n_dofs());
    try {
      selectCollision(*c, z, rho, u, false);
      std::printf("zero density: no exception\n");
    } catch (const DensityZeroException& e) {
      std::printf("zero density: DensityZeroException (%s)\n", e.what());
    }
  }

  // RK5 steps of the SEDG operator on host vectors
  {
    const int Ns = 4, ps = 3;
    const std::vector<double> vs = box(Ns, 2 * M_PI);
    nat_desc d = tgv_desc(vs, Ns, ps, NAT_SEDG);
    d.cartesian = 0;  // the Q1 map, as the SEDG parity tests run it
    auto c = std::make_shared<GpuContext>(d);
    SEDGMinLee<2> A(c);
    std::vector<double> xyz(3 * c->n_dofs());
    if (nat_node_coords(c->get(), xyz.data())) return 3;
    tgv_fields(xyz, rho, u);
    if (nat_init_equilibrium(c->get(), rho.data(), u.data(), nullptr)) return 3;
    DistributionFunctions f(c->Q(), c->n_dofs());
    if (nat_download_f(c->get(), f.data())) return 3;
    RungeKutta5LowStorage rk(A.getDeltaT());
    double t = 0.0;
    for (int i = 0; i < 3; i++) t = rk.step(f, A, SystemVector{}, t, 0.0);
    save(out + "/sedg_rk5.f64", f.data(), f.getQ() * f.size());
    std::printf("rk5 t=%.17g dt=%.17g\n", t, rk.getTimeStepSize());
  }

  // a descriptor the build does not support: compressible populations with SEDG streaming
  {
    nat_desc d = tgv_desc(v, N, 2, NAT_SEDG);
    d.stencil = NAT_D2Q25H;
    d.compressible = 1;
    d.equilibrium = NAT_EQ_QUARTIC;
    d.gamma = 1.4;
    try {
      CompressibleCFDSolver<2> s(d, SolverConfiguration{});
      std::printf("compressible sedg: no exception\n");
    } catch (const NotImplementedException&) {
      std::printf("compressible sedg: NotImplementedException\n");
    } catch (const NATriuMException& e) {
      std::printf("compressible sedg: NATriuMException code %d\n", e.code());
    }
  }
  std::printf("host_layer_check done\n");
  return 0;
}
