// Request server: continuous batching of independent solves on one GPU.
//
// No counterpart in the reference, which runs one solve per process. Requests
// (problem, mesh refinement, order, stencil, Reynolds and Mach number, step count,
// precision) are queued by nat_server_submit from any thread. A scheduler thread
//   - admits queued requests in arrival order while their device footprint fits
//     the HBM budget and an active slot is free, and hands them to setup threads,
//     which build the context (mesh, operator, equilibrium initialisation, initial
//     collision, and for structured meshes the autotuned tile) on its own stream
//     while the scheduler keeps the device busy with the others;
//   - advances every active request by chunks of fused steps. Idle requests that
//     share a packing key (structured SL, same kernel instantiation) advance
//     together: one launch per step covers the tiles of all of them (up to 16), so
//     small solves fill the GPU instead of leaving most CUs idle. Others run alone
//     on their own stream, each chunk a replay of a HIP graph captured once per
//     request (or plain launches for the remainder);
//   - retires a request when its last chunk completes: status check, FP64
//     conservation sums, velocity field copied to the host, context freed, and
//     the freed memory immediately admits the next request.
// A packed launch runs each member's tiles exactly as its own launch would, and
// each request owns its context, so its results do not depend on which other
// requests share the device.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdio>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "nat_internal.hpp"

int nat_internal_enqueue(nat_ctx* ctx, int nsteps);
int nat_internal_capture(nat_ctx* ctx, int nsteps, bool macro_last, hipGraph_t* out);
int nat_internal_status(nat_ctx* ctx);
int64_t nat_internal_footprint(const nat_desc* d);
int nat_internal_pack_key(nat_ctx* ctx, int64_t* key);
int nat_internal_pack_step(nat_ctx* const* ctxs, int n, int nsteps, hipStream_t stream);

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
  bool packable = false;
  int64_t key = 0;  // packing key (nat_pack_key) when packable
  // results
  int64_t n = 0;
  double mass = 0, ekin = 0;
  std::vector<double> u;
};

}  // namespace

// a stream for packed chunks; busy until its last chunk's event has completed
struct PackSlot {
  hipStream_t stream = nullptr;
  hipEvent_t ev = nullptr;
  bool busy = false;
};

struct nat_server {
  nat_server_opts o{};
  std::mutex m;
  std::condition_variable cv_work, cv_done, cv_setup;
  std::deque<Req*> queued;
  std::deque<Req*> to_setup;  // admitted, waiting for a setup thread
  std::vector<Req*> ready;    // built, waiting to join the active set
  int n_setup = 0;            // admitted requests not yet active (to_setup, being built, ready)
  std::vector<Req*> active;
  std::vector<std::thread> setup_th;
  std::vector<PackSlot> slots;
  // HIP invalidates a stream capture when another thread meanwhile makes a call that
  // synchronises with the legacy default stream (set-up copies, launches on blocking streams),
  // and fails that call too: set-ups hold this shared, a capture holds it exclusively and is
  // skipped (plain launches for that chunk) while a set-up runs
  std::shared_mutex cap_mu;
  std::unordered_map<int64_t, std::unique_ptr<Req>> all;
  int64_t next_id = 1;
  int64_t used = 0;
  bool stop = false;
  std::thread th;
  // totals
  int64_t node_updates = 0;
  int64_t packed_chunks = 0;
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

void release(Req& q);
void mark_done_locked(nat_server* s, Req& q);

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
  if (c->desc.scheme == NAT_SL && c->backend == NAT_BACKEND_STRUCTURED && !c->desc.compressible) {
    if (s->o.autotune && (rc = nat_autotune(c, 10, nullptr))) return rc;
    q.packable = s->o.pack && nat_internal_pack_key(c, &q.key) == 0;
  }
  if (hipEventCreateWithFlags(&q.ev, hipEventDisableTiming) != hipSuccess) return NAT_ERR_HIP;
  q.remaining = q.r.nsteps;
  return 0;
}

// setup threads: build admitted requests while the scheduler steps the active ones
void setup_worker(nat_server* s) {
  for (;;) {
    Req* q = nullptr;
    {
      std::unique_lock<std::mutex> lk(s->m);
      // once stopping, admission only drains the queue: with both empty nothing is left to build
      s->cv_setup.wait(lk, [&] { return !s->to_setup.empty() || (s->stop && s->queued.empty()); });
      if (s->to_setup.empty()) return;
      q = s->to_setup.front();
      s->to_setup.pop_front();
    }
    {
      std::shared_lock<std::shared_mutex> cap(s->cap_mu);
      q->rc = setup(s, *q);
      if (q->rc) {
        if (q->err.empty() && q->ctx) q->err = nat_last_error(q->ctx);
        release(*q);
      }
    }
    if (q->rc) {
      {
        std::lock_guard<std::mutex> lk(s->m);
        s->n_setup--;
        mark_done_locked(s, *q);
      }
      s->cv_done.notify_all();
    } else {
      std::lock_guard<std::mutex> lk(s->m);
      s->ready.push_back(q);
    }
    s->cv_work.notify_one();
  }
}

// enqueue the next chunk of steps of an active request (on the context's non-blocking
// stream, see issue())
int issue_chunk(nat_server* s, Req& q) {
  nat_ctx* c = q.ctx;
  const int chunk = s->o.chunk_steps > 0 ? s->o.chunk_steps : 50;
  int k = q.remaining < chunk ? q.remaining : chunk;
  bool graph = s->o.use_graphs && k == chunk && chunk % 2 == 0;
  if (graph && !q.exec) {
    // capture one chunk (an even number of ping-pong steps returns to the same buffers),
    // unless a set-up is running (then this chunk goes as plain launches)
    std::unique_lock<std::shared_mutex> cap(s->cap_mu, std::try_to_lock);
    if (cap.owns_lock()) {
      hipGraph_t g = nullptr;
      if (int rc = nat_internal_capture(c, chunk, true, &g)) return rc;
      q.graph = g;
      if (hipGraphInstantiate(&q.exec, g, nullptr, nullptr, 0) != hipSuccess) return NAT_ERR_HIP;
    } else {
      graph = false;
    }
  }
  if (graph) {
    if (hipGraphLaunch(q.exec, c->stream) != hipSuccess) return NAT_ERR_HIP;
    for (int st = 0; st < chunk; st++) c->time += c->dt;  // as step by step
  } else {
    int rc = nat_internal_enqueue(c, k);
    if (rc) return rc;
  }
  q.remaining -= k;
  if (hipEventRecord(q.ev, c->stream) != hipSuccess) return NAT_ERR_HIP;
  q.in_flight = true;
  return 0;
}

// Chunks run on the context's non-blocking stream, so the legacy-stream copies of other
// requests' set-ups do not wait for them; everything else (set-up, retirement) uses the
// context's own stream once the chunk's event has completed.
int issue(nat_server* s, Req& q) {
  nat_ctx* c = q.ctx;
  hipStream_t own = c->stream;
  c->stream = c->cap_stream;
  const int rc = issue_chunk(s, q);
  c->stream = own;
  return rc;
}

// enqueue one packed chunk of the idle requests qs (equal keys, 2..NAT_PACK_MAX) on a free
// pack slot; false when no slot is free or the pack cannot be launched (the caller then
// issues them alone)
bool issue_pack(nat_server* s, const std::vector<Req*>& qs) {
  PackSlot* slot = nullptr;
  for (PackSlot& sl : s->slots)
    if (!sl.busy) {
      slot = &sl;
      break;
    }
  if (!slot) return false;
  const int chunk = s->o.chunk_steps > 0 ? s->o.chunk_steps : 50;
  int k = chunk;
  std::vector<nat_ctx*> cs;
  for (Req* q : qs) {
    k = q->remaining < k ? q->remaining : k;
    cs.push_back(q->ctx);
  }
  if (k < 1) return false;
  if (nat_internal_pack_step(cs.data(), (int)cs.size(), k, slot->stream)) return false;
  for (Req* q : qs) {
    q->remaining -= k;
    (void)hipEventRecord(q->ev, slot->stream);
    q->in_flight = true;
  }
  (void)hipEventRecord(slot->ev, slot->stream);
  slot->busy = true;
  std::lock_guard<std::mutex> lk(s->m);
  s->packed_chunks++;
  return true;
}

// GPU-side end of a request: results read back, graph, event and context released.
void release(Req& q) {
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
}

// Book-keeping of an ended request, with s->m held by the caller in the same critical
// section that takes the request out of its list, so that a waiter woken by state 2 sees
// the server's counts without it.
void mark_done_locked(nat_server* s, Req& q) {
  q.t_end = now_ms();
  q.state = 2;
  s->used -= q.bytes;
  if (!q.rc) s->node_updates += (int64_t)q.r.nsteps * q.n;
  s->busy_t1 = q.t_end;
}

void run(nat_server* s) {
  const int max_active = s->o.max_active > 0 ? s->o.max_active : 8;
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(s->m);
      for (;;) {
        bool admitted = false;
        while (!s->queued.empty() && (int)s->active.size() + s->n_setup < max_active &&
               (s->used + s->queued.front()->bytes <= s->o.hbm_budget_bytes ||
                (s->active.empty() && s->n_setup == 0))) {
          Req* q = s->queued.front();
          s->queued.pop_front();
          s->used += q->bytes;
          q->state = 1;
          q->t_start = now_ms();
          if (s->busy_t0 < 0) s->busy_t0 = q->t_start;
          s->to_setup.push_back(q);
          s->n_setup++;
          admitted = true;
        }
        if (admitted) s->cv_setup.notify_all();
        for (Req* q : s->ready) s->active.push_back(q);
        s->n_setup -= (int)s->ready.size();
        s->ready.clear();
        if (!s->active.empty()) break;
        if (s->stop && s->queued.empty() && s->n_setup == 0) return;
        s->cv_work.wait_for(lk, std::chrono::milliseconds(s->n_setup ? 1 : 50));
      }
    }
    // free pack slots whose chunk has completed
    for (PackSlot& sl : s->slots)
      if (sl.busy && hipEventQuery(sl.ev) == hipSuccess) sl.busy = false;
    // keep every active request busy: packable idle requests by key, together
    std::vector<std::pair<int64_t, std::vector<Req*>>> groups;
    for (Req* q : s->active) {
      if (q->in_flight || q->remaining <= 0 || q->rc) continue;
      if (q->packable) {
        auto g = groups.begin();
        while (g != groups.end() && g->first != q->key) ++g;
        if (g == groups.end()) {
          groups.emplace_back(q->key, std::vector<Req*>());
          g = groups.end() - 1;
        }
        if (g->second.size() < (size_t)NAT_PACK_MAX) {
          g->second.push_back(q);
          continue;
        }
      }
      q->rc = issue(s, *q);
      if (q->rc) q->remaining = 0;
    }
    for (auto& g : groups) {
      if (g.second.size() >= 2 && issue_pack(s, g.second)) continue;
      for (Req* q : g.second) {
        q->rc = issue(s, *q);
        if (q->rc) q->remaining = 0;
      }
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
        release(*q);
        {
          std::lock_guard<std::mutex> lk(s->m);
          s->active[i] = s->active.back();
          s->active.pop_back();
          mark_done_locked(s, *q);
        }
        s->cv_done.notify_all();
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
  for (int i = 0; i < 4; i++) {  // pack streams (four hardware queues per process)
    PackSlot sl;
    if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess) {
      for (PackSlot& x : s->slots) {
        (void)hipStreamDestroy(x.stream);
        (void)hipEventDestroy(x.ev);
      }
      if (sl.stream) (void)hipStreamDestroy(sl.stream);
      delete s;
      return NAT_ERR_HIP;
    }
    s->slots.push_back(sl);
  }
  const int nthreads = s->o.setup_threads > 0 ? s->o.setup_threads : 4;
  for (int i = 0; i < nthreads; i++) s->setup_th.emplace_back(setup_worker, s);
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
    std::snprintf(res->error, sizeof(res->error), "%s", q->rc ? q->err.c_str() : "");
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
  st->packed_chunks = s->packed_chunks;
  return 0;
}

void nat_server_destroy(nat_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->m);
    s->stop = true;
    s->cv_work.notify_all();
    s->cv_setup.notify_all();
  }
  // the scheduler returns once every queued request has been built and retired
  if (s->th.joinable()) s->th.join();
  {
    std::lock_guard<std::mutex> lk(s->m);
    s->cv_setup.notify_all();
  }
  for (std::thread& t : s->setup_th)
    if (t.joinable()) t.join();
  for (PackSlot& sl : s->slots) {
    (void)hipStreamSynchronize(sl.stream);
    (void)hipStreamDestroy(sl.stream);
    (void)hipEventDestroy(sl.ev);
  }
  delete s;
}

}  // extern "C"
