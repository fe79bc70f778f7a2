// Request server: continuous batching of independent solves on one GPU.
//
// No counterpart in the reference, which runs one solve per process. Requests
// (problem, mesh refinement, order, stencil, Reynolds and Mach number, step count,
// precision) are queued by nat_server_submit from any thread. A scheduler thread
//   - admits queued requests in arrival order while their device footprint fits
//     the HBM budget and an active slot is free, building the context (mesh,
//     operator, equilibrium initialisation, initial collision) on its own stream;
//   - advances every active request by chunks of fused steps, each chunk a
//     replay of a HIP graph captured once per request (or plain launches for the
//     remainder), so the work of different requests overlaps on the device;
//   - retires a request when its last chunk completes: status check, FP64
//     conservation sums, velocity field copied to the host, context freed, and
//     the freed memory immediately admits the next request.
// Each request owns its context and stream, so its results do not depend on
// which other requests share the device.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "nat_internal.hpp"

int nat_internal_enqueue(nat_ctx* ctx, int nsteps);
int nat_internal_status(nat_ctx* ctx);
int64_t nat_internal_footprint(const nat_desc* d);

namespace {

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Req {
  int64_t id = 0;
  nat_request r{};
  nat_desc desc{};
  std::vector<double> verts;
  int64_t bytes = 0;
  double t_submit = 0, t_start = 0, t_end = 0;
  int state = 0;  // 0 queued, 1 active, 2 done
  int rc = 0;
  std::string err;
  nat_ctx* ctx = nullptr;
  int remaining = 0;
  bool in_flight = false;
  hipEvent_t ev = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // results
  int64_t n = 0;
  double mass = 0, ekin = 0;
  std::vector<double> u;
};

}  // namespace

struct nat_server {
  nat_server_opts o{};
  std::mutex m;
  std::condition_variable cv_work, cv_done;
  std::deque<Req*> queued;
  std::vector<Req*> active;
  std::unordered_map<int64_t, std::unique_ptr<Req>> all;
  int64_t next_id = 1;
  int64_t used = 0;
  bool stop = false;
  std::thread th;
  // totals
  int64_t node_updates = 0;
  double busy_t0 = -1, busy_t1 = 0;
};

namespace {

// Mesh and descriptor of a request (problem set of the serving workload).
void build_problem(Req& q) {
  const nat_request& r = q.r;
  nat_desc& d = q.desc;
  d = nat_desc{};
  const double U = 1.0;
  const double cs = U / r.mach;
  d.scaling = std::sqrt(3.0) * cs;
  d.order = r.order;
  d.stencil = r.stencil;
  d.dim = r.stencil == NAT_D2Q9 ? 2 : 3;
  d.cartesian = 1;
  d.cfl = r.cfl > 0 ? r.cfl : 0.4;
  d.scheme = NAT_SL;
  d.precision = r.precision;
  d.equilibrium = NAT_EQ_BGK;
  d.rho0 = 1.0;
  d.backend = NAT_BACKEND_AUTO;
  const int N = 1 << r.refinement;
  double L[3] = {2 * M_PI, 2 * M_PI, 2 * M_PI};
  int n[3] = {N, N, N};
  if (r.problem == NAT_PROBLEM_TGV2D) {
    d.viscosity = U * 2 * M_PI / r.reynolds;  // TaylorGreenVortex2D: L = 2 pi
  } else if (r.problem == NAT_PROBLEM_TGV3D) {
    d.viscosity = U * 1.0 / r.reynolds;  // TaylorGreenVortex3D: characteristic length 1
  } else {  // shock tube: unit-length strip, 4:1 cells, walls at rest at both ends
            // (SodShockTube.cpp:109-128: velocity-neq bounce-back on x, periodic across)
    L[0] = 1.0;
    L[1] = 0.25;
    L[2] = 0.25;
    n[1] = n[2] = N / 4 > 0 ? N / 4 : 1;
    d.viscosity = U * 1.0 / r.reynolds;
  }
  for (int a = 0; a < 3; a++) {
    d.ncells[a] = a < d.dim ? n[a] : 1;
    d.periodic[a] = (r.problem == NAT_PROBLEM_SHOCK_TUBE && a == 0) ? 0 : 1;
  }
  const int nz = d.dim == 3 ? d.ncells[2] + 1 : 1;
  q.verts.clear();
  for (int k = 0; k < nz; k++)
    for (int j = 0; j <= d.ncells[1]; j++)
      for (int i = 0; i <= d.ncells[0]; i++) {
        q.verts.push_back(L[0] * i / d.ncells[0]);
        q.verts.push_back(L[1] * j / d.ncells[1]);
        if (d.dim == 3) q.verts.push_back(L[2] * k / d.ncells[2]);
      }
  d.vertices = q.verts.data();
}

// Initial (rho, u) at the DoFs: Taylor-Green fields (TaylorGreenVortex2D.cpp:36-58,
// TaylorGreenVortex3D.cpp:38-59) or a density step at rest.
void initial_fields(const Req& q, const std::vector<double>& xyz, std::vector<double>& rho, std::vector<double>& u) {
  const int64_t n = (int64_t)xyz.size() / 3;
  const int D = q.desc.dim;
  const double cs = 1.0 / q.r.mach;
  rho.assign(n, 1.0);
  u.assign((size_t)D * n, 0.0);
  for (int64_t i = 0; i < n; i++) {
    const double X = xyz[3 * i], Y = xyz[3 * i + 1], Z = xyz[3 * i + 2];
    if (q.r.problem == NAT_PROBLEM_TGV2D) {
      u[i] = std::sin(X) * std::cos(Y);
      u[n + i] = -std::cos(X) * std::sin(Y);
      rho[i] = 1.0 + 0.25 * (std::cos(2 * X) + std::cos(2 * Y)) / (cs * cs);
    } else if (q.r.problem == NAT_PROBLEM_TGV3D) {
      u[i] = std::sin(X) * std::cos(Y) * std::cos(Z);
      u[n + i] = -std::cos(X) * std::sin(Y) * std::cos(Z);
      rho[i] = 1.0 + (1.0 / 16.0) * (std::cos(2 * X) + std::cos(2 * Y)) * (std::cos(2 * Z) + 2) / (cs * cs);
    } else {
      rho[i] = (X < 0.5) ? 1.0 : 0.8;
    }
  }
}

int setup(nat_server* s, Req& q) {
  int rc = nat_create(&q.desc, &q.ctx);
  if (rc) {
    q.err = q.ctx ? nat_last_error(q.ctx) : "allocation failed";
    return rc;
  }
  nat_ctx* c = q.ctx;
  q.n = c->n_dofs;
  std::vector<double> xyz((size_t)q.n * 3), rho, u;
  if ((rc = nat_node_coords(c, xyz.data()))) return rc;
  initial_fields(q, xyz, rho, u);
  if ((rc = nat_init_equilibrium(c, rho.data(), u.data(), nullptr))) return rc;
  if ((rc = nat_collide(c))) return rc;
  if (hipEventCreateWithFlags(&q.ev, hipEventDisableTiming) != hipSuccess) return NAT_ERR_HIP;
  q.remaining = q.r.nsteps;
  (void)s;
  return 0;
}

// enqueue the next chunk of steps of an active request on its stream
int issue(nat_server* s, Req& q) {
  nat_ctx* c = q.ctx;
  const int chunk = s->o.chunk_steps > 0 ? s->o.chunk_steps : 50;
  int k = q.remaining < chunk ? q.remaining : chunk;
  if (s->o.use_graphs && k == chunk && chunk % 2 == 0) {
    if (!q.exec) {
      // capture one chunk: an even number of ping-pong steps returns to the same buffers
      const double t = c->time;
      const int cur = c->cur;
      if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return NAT_ERR_HIP;
      int rc = nat_internal_enqueue(c, chunk);
      hipGraph_t g = nullptr;
      if (hipStreamEndCapture(c->stream, &g) != hipSuccess || rc) return rc ? rc : NAT_ERR_HIP;
      q.graph = g;
      if (hipGraphInstantiate(&q.exec, g, nullptr, nullptr, 0) != hipSuccess) return NAT_ERR_HIP;
      c->time = t;
      c->cur = cur;
    }
    if (hipGraphLaunch(q.exec, c->stream) != hipSuccess) return NAT_ERR_HIP;
    c->time += chunk * c->dt;
  } else {
    int rc = nat_internal_enqueue(c, k);
    if (rc) return rc;
  }
  q.remaining -= k;
  if (hipEventRecord(q.ev, c->stream) != hipSuccess) return NAT_ERR_HIP;
  q.in_flight = true;
  return 0;
}

void finish(nat_server* s, Req& q) {
  nat_ctx* c = q.ctx;
  if (c) {
    if (!q.rc) q.rc = nat_internal_status(c);
    if (!q.rc) {
      double mom[3];
      q.rc = nat_sums(c, &q.mass, mom, &q.ekin);
    }
    if (!q.rc) {
      q.u.assign((size_t)q.n * q.desc.dim, 0.0);
      std::vector<double> rho(q.n);
      q.rc = nat_get_macro(c, rho.data(), q.u.data(), nullptr);
    }
    if (q.rc && q.err.empty()) q.err = nat_last_error(c);
    if (q.exec) (void)hipGraphExecDestroy(q.exec);
    if (q.graph) (void)hipGraphDestroy(q.graph);
    if (q.ev) (void)hipEventDestroy(q.ev);
    nat_destroy(c);
    q.ctx = nullptr;
  }
  std::lock_guard<std::mutex> lk(s->m);
  q.t_end = now_ms();
  q.state = 2;
  s->used -= q.bytes;
  if (!q.rc) s->node_updates += (int64_t)q.r.nsteps * q.n;
  s->busy_t1 = q.t_end;
  s->cv_done.notify_all();
}

void run(nat_server* s) {
  const int max_active = s->o.max_active > 0 ? s->o.max_active : 8;
  for (;;) {
    std::vector<Req*> admit;
    {
      std::unique_lock<std::mutex> lk(s->m);
      for (;;) {
        while (!s->queued.empty() && (int)(s->active.size() + admit.size()) < max_active &&
               (s->used + s->queued.front()->bytes <= s->o.hbm_budget_bytes || (s->active.empty() && admit.empty()))) {
          Req* q = s->queued.front();
          s->queued.pop_front();
          s->used += q->bytes;
          q->state = 1;
          q->t_start = now_ms();
          if (s->busy_t0 < 0) s->busy_t0 = q->t_start;
          admit.push_back(q);
        }
        if (!admit.empty() || !s->active.empty()) break;
        if (s->stop) return;
        s->cv_work.wait(lk);
      }
    }
    for (Req* q : admit) {
      q->rc = setup(s, *q);
      if (q->rc) {
        if (q->err.empty() && q->ctx) q->err = nat_last_error(q->ctx);
        finish(s, *q);
      } else {
        std::lock_guard<std::mutex> lk(s->m);
        s->active.push_back(q);
      }
    }
    // keep every active request busy
    for (Req* q : s->active)
      if (!q->in_flight && q->remaining > 0) {
        q->rc = issue(s, *q);
        if (q->rc) q->remaining = 0;
      }
    // retire completed requests
    bool progressed = false;
    for (size_t i = 0; i < s->active.size();) {
      Req* q = s->active[i];
      if (q->in_flight && hipEventQuery(q->ev) == hipSuccess) {
        q->in_flight = false;
        progressed = true;
      }
      if (!q->in_flight && (q->remaining == 0 || q->rc)) {
        finish(s, *q);
        std::lock_guard<std::mutex> lk(s->m);
        s->active[i] = s->active.back();
        s->active.pop_back();
        progressed = true;
        continue;
      }
      i++;
    }
    if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
}

}  // namespace

extern "C" {

int nat_server_create(const nat_server_opts* opts, nat_server** out) {
  *out = nullptr;
  auto* s = new (std::nothrow) nat_server;
  if (!s) return NAT_ERR_NOMEM;
  if (opts) s->o = *opts;
  if (s->o.hbm_budget_bytes <= 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
      delete s;
      return NAT_ERR_HIP;
    }
    s->o.hbm_budget_bytes = (int64_t)(0.8 * (double)fr);
  }
  s->th = std::thread(run, s);
  *out = s;
  return 0;
}

int nat_server_submit(nat_server* s, const nat_request* r, int64_t* id) {
  if (!s || !r || !id) return NAT_ERR_ARG;
  if (r->problem < NAT_PROBLEM_TGV2D || r->problem > NAT_PROBLEM_SHOCK_TUBE) return NAT_ERR_ARG;
  if (r->refinement < 1 || r->refinement > 9 || r->order < 1 || r->order > 8 || r->nsteps < 0) return NAT_ERR_ARG;
  if (!(r->reynolds > 0) || !(r->mach > 0)) return NAT_ERR_ARG;
  const bool d3 = r->stencil == NAT_D3Q19 || r->stencil == NAT_D3Q27;
  if (r->stencil != NAT_D2Q9 && !d3) return NAT_ERR_UNSUPPORTED;
  if ((r->problem == NAT_PROBLEM_TGV3D) != d3) return NAT_ERR_ARG;
  auto q = std::make_unique<Req>();
  q->r = *r;
  build_problem(*q);
  q->bytes = nat_internal_footprint(&q->desc);
  std::lock_guard<std::mutex> lk(s->m);
  q->id = s->next_id++;
  q->t_submit = now_ms();
  *id = q->id;
  s->queued.push_back(q.get());
  s->all[q->id] = std::move(q);
  s->cv_work.notify_one();
  return 0;
}

int nat_server_wait(nat_server* s, int64_t id, nat_result* res, double* u, int64_t u_len) {
  std::unique_lock<std::mutex> lk(s->m);
  auto it = s->all.find(id);
  if (it == s->all.end()) return NAT_ERR_ARG;
  Req* q = it->second.get();
  s->cv_done.wait(lk, [&] { return q->state == 2; });
  if (res) {
    res->status = q->rc;
    res->n_dofs = q->n;
    res->queue_ms = q->t_start - q->t_submit;
    res->run_ms = q->t_end - q->t_start;
    res->latency_ms = q->t_end - q->t_submit;
    res->mass = q->mass;
    res->ekin = q->ekin;
    res->footprint_bytes = q->bytes;
  }
  if (u && u_len > 0) {
    const int64_t m = (int64_t)q->u.size() < u_len ? (int64_t)q->u.size() : u_len;
    for (int64_t i = 0; i < m; i++) u[i] = q->u[i];
  }
  const int rc = q->rc;
  s->all.erase(it);
  return rc;
}

int nat_server_stats(nat_server* s, nat_server_stats_t* st) {
  std::lock_guard<std::mutex> lk(s->m);
  st->queued = (int64_t)s->queued.size();
  st->active = (int64_t)s->active.size();
  st->used_bytes = s->used;
  st->node_updates = s->node_updates;
  st->busy_ms = s->busy_t0 < 0 ? 0.0 : s->busy_t1 - s->busy_t0;
  return 0;
}

void nat_server_destroy(nat_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->m);
    s->stop = true;
    s->cv_work.notify_all();
  }
  if (s->th.joinable()) s->th.join();
  delete s;
}

}  // extern "C"
