// serve.hip -- the serving loop: solves arrive one by one (each on one of the resident meshes,
// with its own sources and optionally its own speed field) and are solved together.  No
// counterpart in the reference (one solve per process).
//
// A worker thread owns `launches` launch sequences (one HIP stream each).  Whenever one of them
// is idle and requests are waiting, it admits waiting requests in arrival order while the
// device-memory budget for in-flight request state allows (at least one, so a request larger
// than the budget still runs alone), up to max_solves_per_launch, and starts one multi-solve
// launch for them (multi.h): their work shares the GPU with the launches already running.  When
// a launch completes (its end event has fired) the results are copied out and the budget freed.
// Requests that arrive while launches are running are admitted into the next free launch at
// once, so a long solve does not hold back the ones behind it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "multi.h"

namespace eik {

struct Request {
  int64_t id = 0;
  int32_t mesh = 0;
  std::vector<int32_t> seeds;
  const float* factors = nullptr;  // caller-owned, one per element, or null
  float* T_out = nullptr;          // caller-owned, nv values, or null
  int64_t bytes = 0;
  int32_t status = 0;  // 0 queued, 1 running, 2 done, < 0 failed (-code)
  int32_t launch = -1;
  double t_submit = 0, t_start = 0, t_end = 0;
  std::string error;
};

struct Slots {  // per launch: time, seen and factor arrays for max_solves requests
  float* T = nullptr;
  float* seen = nullptr;
  float* mult = nullptr;   // a request's per-element factors as uploaded
  float* rmult = nullptr;  // the same in record order (what the visits read)
};

struct Server {
  std::vector<Mesh*> meshes;
  eik_server_opts_t opts;
  int device = 0;
  int64_t max_nv = 0, max_nseen = 0, max_ne = 0, max_nrec = 0, max_blocks = 0;
  std::vector<Launch> L;
  std::vector<Slots> slots;
  std::vector<std::vector<Request*>> inflight;
  std::deque<Request*> pending;
  std::unordered_map<int64_t, Request*> reqs;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::thread worker;
  bool stop = false;
  int64_t used = 0;
  int64_t next_id = 1;
  std::atomic<int64_t> launches_started{0};
  std::chrono::steady_clock::time_point t0;
  std::string fatal;

  double now_ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
};

static int64_t request_bytes(const Server& s, int32_t mesh, bool factors) {
  const Mesh* m = s.meshes[mesh];
  // T, out, seen, factors per element and per record
  return (int64_t)sizeof(float) *
         (2 * m->nv + m->dev->nseen + (factors ? m->ne + m->layout.total_elems : 0));
}

static void server_free(Server* s) {
  for (size_t k = 0; k < s->L.size(); ++k) {
    launch_free(s->L[k]);
    Slots& z = s->slots[k];
    if (z.T) (void)hipFree(z.T);
    if (z.seen) (void)hipFree(z.seen);
    if (z.mult) (void)hipFree(z.mult);
    if (z.rmult) (void)hipFree(z.rmult);
  }
  for (auto& kv : s->reqs) delete kv.second;
  delete s;
}

// Starts one launch on launch sequence k for the given requests (called by the worker).
static int start_launch(Server* s, int k, std::vector<Request*>& rs) {
  Launch& L = s->L[k];
  Slots& z = s->slots[k];
  int64_t off = 0, bbase = 0, nseeds = 0;
  const int64_t cap = L.cap_seeds;
  for (size_t j = 0; j < rs.size(); ++j) {
    Request* r = rs[j];
    Mesh* m = s->meshes[r->mesh];
    float* rmult = nullptr;
    if (r->factors) {
      float* mult = z.mult + (int64_t)j * s->max_ne;
      rmult = z.rmult + (int64_t)j * s->max_nrec;
      HIP_TRY(hipMemcpyAsync(mult, r->factors, sizeof(float) * m->ne, hipMemcpyHostToDevice, L.stream));
      record_factors(*m, mult, rmult, L.stream);
    }
    SolveDesc& d = L.hsolves[j];
    solve_desc(d, m, z.T + (int64_t)j * s->max_nv, z.seen + (int64_t)j * s->max_nseen, rmult);
    d.out_off = off;
    d.bbase = (int32_t)bbase;
    off += m->nv;
    bbase += d.nblocks;
    for (int32_t v : r->seeds) {
      L.h_seed[nseeds] = (int32_t)j;
      L.h_seed[cap + nseeds] = v;
      const float zero = 0.f;
      std::memcpy(&L.h_seed[2 * cap + nseeds], &zero, 4);
      ++nseeds;
    }
  }
  L.h_seed[3 * cap] = (int32_t)nseeds;
  eik_solve_opts_t o;
  eik_solve_opts_default(&o);
  o.max_inner = s->opts.max_inner;
  MultiArgs a[2];
  int rc = launch_args(L, (int32_t)rs.size(), bbase, o, a);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(L.ev0, L.stream));
  rc = launch_issue(L, (int32_t)rs.size(), bbase, off, a, true);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(L.ev1, L.stream));
  return EIK_OK;
}

static void finish_launch(Server* s, int k) {
  Launch& L = s->L[k];
  int64_t off = 0;
  const double t = s->now_ms();
  for (size_t j = 0; j < s->inflight[k].size(); ++j) {
    Request* r = s->inflight[k][j];
    const Mesh* m = s->meshes[r->mesh];
    if (r->T_out) std::memcpy(r->T_out, L.h_out + off, sizeof(float) * m->nv);
    off += m->nv;
    r->t_end = t;
  }
  std::lock_guard<std::mutex> g(s->mu);
  for (Request* r : s->inflight[k]) {
    r->status = 2;
    s->used -= r->bytes;
  }
  s->inflight[k].clear();
  s->cv_done.notify_all();
}

static void fail_launch(Server* s, int k, std::vector<Request*>& rs, int rc) {
  std::lock_guard<std::mutex> g(s->mu);
  const std::string msg = eik_last_error();
  for (Request* r : rs) {
    r->status = -rc;
    r->error = msg;
    s->used -= r->bytes;
  }
  rs.clear();
  s->cv_done.notify_all();
}

static void worker_main(Server* s) {
  if (hipSetDevice(s->device) != hipSuccess) {
    std::lock_guard<std::mutex> g(s->mu);
    s->fatal = "cannot select the device in the serving thread";
    s->cv_done.notify_all();
    return;
  }
  const int nl = (int)s->L.size();
  for (;;) {
    bool progressed = false;
    // completions
    for (int k = 0; k < nl; ++k) {
      if (s->inflight[k].empty()) continue;
      hipError_t e = hipEventQuery(s->L[k].ev1);
      if (e == hipSuccess) {
        int rc = EIK_OK;
        if (hipStreamSynchronize(s->L[k].stream) != hipSuccess) {
          fail_launch(s, k, s->inflight[k], fail(EIK_ERR_HIP, "serving launch failed"));
        } else if ((rc = launch_check_end(s->L[k])) != EIK_OK) {
          fail_launch(s, k, s->inflight[k], rc);
        } else {
          finish_launch(s, k);
        }
        progressed = true;
      } else if (e != hipErrorNotReady) {
        fail_launch(s, k, s->inflight[k],
                    fail(EIK_ERR_HIP, std::string("serving launch: ") + hipGetErrorString(e)));
        progressed = true;
      }
    }
    // admissions
    for (int k = 0; k < nl; ++k) {
      if (!s->inflight[k].empty()) continue;
      std::vector<Request*> rs;
      {
        std::lock_guard<std::mutex> g(s->mu);
        while (!s->pending.empty() && (int32_t)rs.size() < s->opts.max_solves_per_launch) {
          Request* r = s->pending.front();
          const bool fits = s->used + r->bytes <= s->opts.budget_bytes;
          if (!fits && !(rs.empty() && s->used == 0)) break;
          s->pending.pop_front();
          s->used += r->bytes;
          r->status = 1;
          r->launch = k;
          rs.push_back(r);
        }
      }
      if (rs.empty()) continue;
      const double t = s->now_ms();
      for (Request* r : rs) r->t_start = t;
      int rc = start_launch(s, k, rs);
      if (rc) {
        fail_launch(s, k, rs, rc);
      } else {
        s->inflight[k] = rs;
        ++s->launches_started;
      }
      progressed = true;
    }
    if (progressed) continue;
    std::unique_lock<std::mutex> lk(s->mu);
    bool busy = false;
    for (int k = 0; k < nl; ++k) busy |= !s->inflight[k].empty();
    if (s->stop && s->pending.empty() && !busy) return;
    if (busy) {
      // a launch is running: poll its event again shortly (new requests also wake us)
      s->cv_work.wait_for(lk, std::chrono::microseconds(20));
    } else {
      s->cv_work.wait_for(lk, std::chrono::milliseconds(5),
                          [&] { return s->stop || !s->pending.empty(); });
    }
  }
}

static int server_create(std::vector<Mesh*> ms, const eik_server_opts_t& o, Server** out) {
  if (ms.empty()) return fail(EIK_ERR_ARG, "no resident mesh");
  if (o.max_solves_per_launch <= 0 || o.launches <= 0 || o.budget_bytes <= 0 || o.max_seeds <= 0)
    return fail(EIK_ERR_ARG, "server options must be positive");
  for (Mesh* m : ms)
    if (m->nblocks <= 0) return fail(EIK_ERR_STATE, "every resident mesh must be partitioned");
  int rc = build_layouts_parallel(ms);
  if (rc) return rc;
  Server* s = new Server();
  s->meshes = ms;
  s->opts = o;
  auto bail = [&](int code) {
    server_free(s);
    return code;
  };
  if (hipGetDevice(&s->device) != hipSuccess) return bail(fail(EIK_ERR_HIP, "no HIP device"));
  for (Mesh* m : ms) {
    rc = upload_layout(*m);
    if (!rc) rc = ensure_split(*m);
    if (rc) return bail(rc);
    s->max_nv = std::max<int64_t>(s->max_nv, m->nv);
    s->max_nseen = std::max<int64_t>(s->max_nseen, m->dev->nseen);
    s->max_ne = std::max<int64_t>(s->max_ne, m->ne);
    s->max_nrec = std::max<int64_t>(s->max_nrec, m->layout.total_elems);
    s->max_blocks = std::max<int64_t>(s->max_blocks, m->dev->nblocks);
  }
  const int R = o.max_solves_per_launch;
  s->L.resize(o.launches);
  s->slots.resize(o.launches);
  s->inflight.resize(o.launches);
  for (int k = 0; k < o.launches; ++k) {
    rc = launch_alloc(s->L[k], ms, R, (int64_t)R * s->max_blocks, (int64_t)R * s->max_nv,
                      (int64_t)R * o.max_seeds);
    if (rc) return bail(rc);
    Slots& z = s->slots[k];
    if (hipMalloc((void**)&z.T, sizeof(float) * R * s->max_nv) != hipSuccess ||
        hipMalloc((void**)&z.seen, sizeof(float) * R * s->max_nseen) != hipSuccess ||
        hipMalloc((void**)&z.mult, sizeof(float) * R * s->max_ne) != hipSuccess ||
        hipMalloc((void**)&z.rmult, sizeof(float) * R * s->max_nrec) != hipSuccess)
      return bail(fail(EIK_ERR_NOMEM, "serving slot allocation failed"));
  }
  s->t0 = std::chrono::steady_clock::now();
  s->worker = std::thread(worker_main, s);
  *out = s;
  return EIK_OK;
}

static int server_submit(Server* s, int32_t mesh, const int32_t* seeds, int32_t nseeds,
                         const float* factors, float* T_out, int64_t* id) {
  if (mesh < 0 || mesh >= (int32_t)s->meshes.size()) return fail(EIK_ERR_ARG, "mesh index out of range");
  if (!seeds || nseeds <= 0) return fail(EIK_ERR_ARG, "at least one seed vertex is required");
  if (nseeds > s->opts.max_seeds) return fail(EIK_ERR_CAPACITY, "more seeds than max_seeds");
  const Mesh* m = s->meshes[mesh];
  for (int32_t i = 0; i < nseeds; ++i)
    if (seeds[i] < 0 || seeds[i] >= m->nv) return fail(EIK_ERR_ARG, "seed vertex out of range");
  Request* r = new Request();
  r->mesh = mesh;
  r->seeds.assign(seeds, seeds + nseeds);
  r->factors = factors;
  r->T_out = T_out;
  r->bytes = request_bytes(*s, mesh, factors != nullptr);
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (!s->fatal.empty()) {
      delete r;
      return fail(EIK_ERR_HIP, s->fatal);
    }
    r->id = s->next_id++;
    r->t_submit = s->now_ms();
    s->reqs[r->id] = r;
    s->pending.push_back(r);
  }
  s->cv_work.notify_one();
  if (id) *id = r->id;
  return EIK_OK;
}

static void fill_info(const Request* r, eik_request_info_t* info) {
  if (!info) return;
  info->status = r->status;
  info->launch = r->launch;
  info->submit_ms = r->t_submit;
  info->start_ms = r->t_start;
  info->end_ms = r->t_end;
}

}  // namespace eik

using namespace eik;

extern "C" {

void eik_server_opts_default(eik_server_opts_t* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->max_solves_per_launch = 32;
  o->launches = 4;
  o->budget_bytes = (int64_t)4 << 30;
  o->max_seeds = 64;
  o->max_inner = 1;
}

int eik_server_create(const eik_mesh_t* meshes, int32_t n, const eik_server_opts_t* opts,
                      eik_server_t* out) {
  if (!meshes || n <= 0 || !out) return fail(EIK_ERR_ARG, "null or empty argument");
  std::vector<Mesh*> ms(n);
  for (int32_t i = 0; i < n; ++i) {
    ms[i] = reinterpret_cast<Mesh*>(meshes[i]);
    if (!ms[i]) return fail(EIK_ERR_ARG, "null mesh");
  }
  eik_server_opts_t o;
  eik_server_opts_default(&o);
  if (opts) o = *opts;
  Server* s = nullptr;
  int rc = server_create(ms, o, &s);
  if (rc) return rc;
  *out = reinterpret_cast<eik_server_t>(s);
  return EIK_OK;
}

int eik_server_submit(eik_server_t server, int32_t mesh, const int32_t* seeds, int32_t nseeds,
                      const float* factors, float* T_out, int64_t* request_id) {
  Server* s = reinterpret_cast<Server*>(server);
  if (!s) return fail(EIK_ERR_ARG, "null server");
  return server_submit(s, mesh, seeds, nseeds, factors, T_out, request_id);
}

int eik_server_poll(eik_server_t server, int64_t request_id, eik_request_info_t* info) {
  Server* s = reinterpret_cast<Server*>(server);
  if (!s) return fail(EIK_ERR_ARG, "null server");
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->reqs.find(request_id);
  if (it == s->reqs.end()) return fail(EIK_ERR_ARG, "unknown request");
  fill_info(it->second, info);
  return EIK_OK;
}

int eik_server_wait(eik_server_t server, int64_t request_id, eik_request_info_t* info) {
  Server* s = reinterpret_cast<Server*>(server);
  if (!s) return fail(EIK_ERR_ARG, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(request_id);
  if (it == s->reqs.end()) return fail(EIK_ERR_ARG, "unknown request");
  Request* r = it->second;
  s->cv_done.wait(lk, [&] { return r->status == 2 || r->status < 0 || !s->fatal.empty(); });
  fill_info(r, info);
  const int32_t st = r->status;
  const std::string err = !s->fatal.empty() ? s->fatal : r->error;
  s->reqs.erase(it);
  delete r;
  if (st < 0) return fail(-st, err);
  if (st != 2) return fail(EIK_ERR_HIP, err);
  return EIK_OK;
}

int eik_server_destroy(eik_server_t server) {
  Server* s = reinterpret_cast<Server*>(server);
  if (!s) return EIK_OK;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  server_free(s);
  return EIK_OK;
}

int eik_server_launches(eik_server_t server, int64_t* started) {
  Server* s = reinterpret_cast<Server*>(server);
  if (!s || !started) return fail(EIK_ERR_ARG, "null argument");
  *started = s->launches_started.load();
  return EIK_OK;
}

}  // extern "C"
