// natrium_amd internals shared by the C-ABI layer and the HIP kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/natrium_amd.h"

namespace nat {

constexpr int MAXQ = 32;
constexpr int MAXP = 8;
constexpr int MAXSHIFT = 4;  // distinct non-zero velocity components per axis

// ---------------------------------------------------------------------------
// Stencils. Integer velocity sets (D2Q9, D3Q19, D3Q27) are compile-time tables
// so the collision and moment sums unroll with exact +-1/0 coefficients.
// Ordering and weights follow stencils/D2Q9.cpp:40-61, D3Q19.cpp:37-60,
// D3Q27.cpp:37-110 of the reference.
// ---------------------------------------------------------------------------
template <int Q>
struct Lattice;

template <>
struct Lattice<9> {
  static constexpr int D = 2;
  static constexpr int e[9][3] = {{0, 0, 0}, {1, 0, 0},  {0, 1, 0},  {-1, 0, 0}, {0, -1, 0},
                                  {1, 1, 0}, {-1, 1, 0}, {-1, -1, 0}, {1, -1, 0}};
  static constexpr double w[9] = {4. / 9., 1. / 9., 1. / 9., 1. / 9., 1. / 9., 1. / 36., 1. / 36., 1. / 36., 1. / 36.};
};

template <>
struct Lattice<19> {
  static constexpr int D = 3;
  static constexpr int e[19][3] = {{0, 0, 0},  {1, 0, 0},  {0, 0, 1},   {-1, 0, 0}, {0, 0, -1}, {0, -1, 0}, {0, 1, 0},
                                   {1, 0, 1},  {-1, 0, 1}, {-1, 0, -1}, {1, 0, -1}, {1, -1, 0}, {1, 1, 0},  {-1, 1, 0},
                                   {-1, -1, 0}, {0, -1, 1}, {0, 1, 1},   {0, 1, -1}, {0, -1, -1}};
  static constexpr double w[19] = {1. / 3.,  1. / 18., 1. / 18., 1. / 18., 1. / 18., 1. / 18., 1. / 18.,
                                   1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.,
                                   1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.};
};

template <>
struct Lattice<27> {
  static constexpr int D = 3;
  static constexpr int e[27][3] = {{0, 0, 0},   {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},   {0, -1, 0},   {0, 0, 1},
                                   {0, 0, -1},  {0, 1, 1},   {0, -1, -1}, {0, 1, -1}, {0, -1, 1},   {1, 0, 1},
                                   {-1, 0, -1}, {1, 0, -1},  {-1, 0, 1}, {1, 1, 0},   {-1, -1, 0},  {1, -1, 0},
                                   {-1, 1, 0},  {1, 1, 1},   {-1, -1, -1}, {1, 1, -1}, {-1, -1, 1}, {1, -1, 1},
                                   {-1, 1, -1}, {1, -1, -1}, {-1, 1, 1}};
  static constexpr double w[27] = {8. / 27.,  2. / 27.,  2. / 27.,  2. / 27.,  2. / 27.,  2. / 27.,  2. / 27.,
                                   1. / 54.,  1. / 54.,  1. / 54.,  1. / 54.,  1. / 54.,  1. / 54.,  1. / 54.,
                                   1. / 54.,  1. / 54.,  1. / 54.,  1. / 54.,  1. / 54.,  1. / 216., 1. / 216.,
                                   1. / 216., 1. / 216., 1. / 216., 1. / 216., 1. / 216., 1. / 216.};
};

// Host-side description of any supported stencil (incl. D2Q25H).
struct StencilHost {
  int id = 0, D = 0, Q = 0;
  double s = 1.0;
  double cs2 = 0.0;              // scaled c_s^2 = s^2/3
  double dir[MAXQ][3] = {};      // scaled velocities
  double w[MAXQ] = {};
  double max_speed = 0.0;        // sqrt(2) s: reference dt convention
};
bool make_stencil(int id, double s, StencilHost& st);

// Gauss-Lobatto points and weights on [0,1].
void gll_points(int p, double* x, double* w);

// ---------------------------------------------------------------------------
// Kernel argument blocks
// ---------------------------------------------------------------------------
struct MeshDev {
  int D;
  int N[3];
  int periodic[3];
  int cartesian;
  const double* vert;  // (N0+1)(N1+1)(N2+1) x 3 (z = 0 in 2D)
  double wall_u[6][3] = {};  // velocity-neq bounce-back wall velocity per box face (non-periodic axes)
};

// Mode bits of the fused step kernels' launch argument.
enum {
  STEP_MACRO = 1,       // write rho and u (the last step of a batch)
  STEP_STREAM_ONLY = 2  // CFDSolver::stream without the collision: store the streamed values
};

struct GllTab {
  double x[MAXP + 1];
};

// Collision constants shared by all step kernels (unscaled lattice units).
template <typename T>
struct CollideArgs {
  T omega;     // 1/tau
  double omega64;  // 1/tau for kernels that collide in FP64
  T rho0;      // FP32 shift reference density (0 for unshifted FP64)
  double scaling;
  T* rho;      // optional outputs (may be null)
  T* u;        // D arrays, physical units
  int64_t mstride;
  int* status;  // bit 0: density < 1e-10 seen
  int stab = 0;  // pseudo-entropic stabilizer after the collision (D2Q9 / D3Q19)
};

}  // namespace nat

// Pass-1 blocks (256 (cell, node) pairs each) that the fixed scratch of nat_sums_noalloc
// holds: 65536 x 5 doubles = 2.5 MiB per GPU, up to 16,777,216 pairs.
#define NAT_SUMS_SCRATCH_BLOCKS 65536

// Solves one packed structured launch can advance together.
#define NAT_PACK_MAX 16

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
struct nat_ctx {
  nat_desc desc{};
  std::vector<double> verts;  // 3 per vertex
  nat::StencilHost st;
  int D = 2, Q = 9, p = 1;
  int64_t n_dofs = 0;      // reference DoF count (FE_Q lattice with periodic duplicates, or DG)
  int64_t n_store = 0;     // nodes held per direction on the device
  int64_t stride = 0;      // elements between direction arrays (padded)
  int nfe[3] = {1, 1, 1};  // FE_Q lattice nodes per axis
  int nu[3] = {1, 1, 1};   // unique periodic nodes per axis (structured backend)
  double dt = 0, tau = 0, time = 0, h_min = 0;
  bool fp32 = false;
  bool assembled = false;
  bool collided = false;
  int backend = NAT_BACKEND_GENERAL;
  hipStream_t stream = nullptr;
  // step graphs are captured here, not on `stream`: the context stream is a blocking stream,
  // and a capture on it would be broken by any legacy-stream operation another thread makes
  // meanwhile (a second context being set up)
  hipStream_t cap_stream = nullptr;
  int device = -1;  // HIP device current at nat_create (the context's allocations live there)
  int n_cu = 0;     // compute units of that device

  // device state
  void* f[2] = {nullptr, nullptr};
  void* g[2] = {nullptr, nullptr};
  int cur = 0;
  void* rho = nullptr;  // macro fields in compute precision (compressible T is FP64)
  void* u = nullptr;
  double* T = nullptr;
  double* mask = nullptr;
  double* d_vert = nullptr;
  double* d_gll = nullptr;
  double* d_coords = nullptr;  // SL: FE_Q support points (n_dofs x 3)
  int* d_owner = nullptr;      // SL: owning cell per FE_Q DoF
  int* d_status = nullptr;
  int* h_status = nullptr;     // pinned host copy of the status word (one D2H copy per call, no staging)
  double* d_sums = nullptr;
  double* d_partials = nullptr;  // per-block partial conservation sums (two-pass reduction)
  int64_t n_partials = 0;
  hipEvent_t ev[2] = {nullptr, nullptr};  // timing events of nat_step_timed, created once

  // SL general backend: departure records per (alpha-1, node)
  int* dep_cell = nullptr;
  double* dep_xi64 = nullptr;  // D per record
  // FP32 storage: per record D + 1 int32, interleaved so that one load brings a record: the
  // code cell * (p+1)^D + anchor node, then the fixed-point offsets from the anchor (2^-30 units)
  int* dep_rec32 = nullptr;

  // velocity-neq bounce-back walls (general SL backend): per node the index of its wall
  // entry or -1; per wall entry a mask of the directions read from the opposite
  // direction (bit alpha) and the value added per direction after interpolation
  int* wall_idx = nullptr;
  uint32_t* wall_flip = nullptr;
  double* wall_bias = nullptr;  // n_wall x Q
  int64_t n_wall = 0;
  bool has_walls = false;

  // captured step chunks, one per starting ping-pong buffer
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  bool use_graphs = true;

  // residuals of the results table: fields of the previous 100th-iteration call
  std::vector<double> res_rho, res_u;
  int64_t res_first = 0, res_last = 0;
  bool res_have = false;
  double res_rho_val = 1.0, res_u_val = 1.0;  // the solver's start values (CFDSolver.cpp:398-399)

  // pseudo-entropic stabilizer as a data processor after each loop step's collision
  // (D2Q9 and D3Q19 only, as in the reference)
  bool stabilizer = false;
  int stab_on() const { return stabilizer && (Q == 9 || Q == 19) ? 1 : 0; }

  // SL structured backend: 1-D tables
  void* ax_w = nullptr;   // [3][MAXSHIFT][MAXP][MAXP+1] weights (T)
  int* ax_off = nullptr;  // [3][MAXSHIFT][MAXP] source cell offsets
  bool ax_sym = false;    // tables equal on all axes and mirror symmetric in the shift sign
  int tile_pref = -1;     // structured tile: -1 size heuristic, 0 large tile, 1 small tile (autotuner)
  int gen_pref = -1;      // general SL FP32 3D kernel: -1 grid-fill rule, 0 six-wave form, 1 DEEP form (autotuner)
  int sid[nat::MAXQ][3] = {};
  int off_lo[3][nat::MAXSHIFT] = {}, off_hi[3][nat::MAXSHIFT] = {};
  int nshift[3] = {0, 0, 0};

  // SEDG backend
  void* sedg_df = nullptr;
  double* sedg_geo = nullptr;

  // compressible: per-direction constant table (unscaled e, w, H3/H4 components)
  double* comp_tab = nullptr;

  std::string err;
};

// error helpers
#define NAT_HIP(call)                                                           \
  do {                                                                          \
    hipError_t _e = (call);                                                     \
    if (_e != hipSuccess) {                                                     \
      ctx->err = std::string(#call) + ": " + hipGetErrorString(_e);             \
      return NAT_ERR_HIP;                                                       \
    }                                                                           \
  } while (0)
