// serve_host_model.cpp -- the serving loop's host code (csrc/serve.hip, compiled for the host
// only) against a model device, for ThreadSanitizer and stall checks on the CPU.
//
// The model device has one thread per stream that runs the stream's operations in order, each
// after a short random delay; events and stream synchronisation wait on what those threads have
// completed, with the happens-before edges of a real device (a completed event publishes every
// write of the operations before it).  The multi-solve launch layer of batch.hip is replaced by
// operations with the same effects on the launch's pinned buffers: a "visit" writes per-request
// results (solve index + vertex index + 1 into the gathered output) and the counters.  Copies
// from pageable caller memory are done at enqueue time (the source may be reused once the call
// returns, as with HIP).
//
// usage: serve_host_model <requests> <submit threads> <launches> <solves per launch> <seed>
// Exit status 0: every request completed with the expected values and nothing stalled.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "multi.h"

// ------------------------------------------------------------------------------------------
// model device
// ------------------------------------------------------------------------------------------
namespace model {

struct Stream {
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::pair<int, std::function<void()>>> ops;  // (delay us, effect)
  uint64_t enqueued = 0, completed = 0;
  bool stop = false;
  std::thread th;
  std::mt19937 rng;
};

struct Event {
  Stream* s = nullptr;
  uint64_t seq = 0;
};

std::mutex g_mu;
std::vector<std::unique_ptr<Stream>> g_streams;
std::atomic<int> g_max_delay_us{400};

void run(Stream* s) {
  for (;;) {
    std::pair<int, std::function<void()>> op;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv.wait(lk, [&] { return s->stop || !s->ops.empty(); });
      if (s->ops.empty()) return;
      op = std::move(s->ops.front());
      s->ops.pop_front();
    }
    if (op.first > 0) std::this_thread::sleep_for(std::chrono::microseconds(op.first));
    op.second();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      ++s->completed;
    }
    s->cv.notify_all();
  }
}

Stream* new_stream() {
  std::lock_guard<std::mutex> g(g_mu);
  g_streams.emplace_back(new Stream());
  Stream* s = g_streams.back().get();
  s->rng.seed((unsigned)g_streams.size() * 7919u);
  s->th = std::thread(run, s);
  return s;
}

void enqueue(Stream* s, std::function<void()> fn, bool slow = false) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    const int d = slow ? (int)(s->rng() % (unsigned)g_max_delay_us.load()) : 0;
    s->ops.emplace_back(d, std::move(fn));
    ++s->enqueued;
  }
  s->cv.notify_all();
}

void sync(Stream* s) {
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv.wait(lk, [&] { return s->completed >= s->enqueued; });
}

void shutdown() {
  for (auto& s : g_streams) {
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stop = true;
    }
    s->cv.notify_all();
    s->th.join();
  }
  g_streams.clear();
}

}  // namespace model

static model::Stream* S(hipStream_t s) { return reinterpret_cast<model::Stream*>(s); }

// ------------------------------------------------------------------------------------------
// the HIP runtime calls the serving loop makes
// ------------------------------------------------------------------------------------------
extern "C" {
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDevice(int* d) {
  *d = 0;
  return hipSuccess;
}
const char* hipGetErrorString(hipError_t) { return "model device error"; }
hipError_t hipMalloc(void** p, size_t n) {
  *p = std::calloc(n ? n : 1, 1);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p) {
  std::free(p);
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s) {
  auto* ev = reinterpret_cast<model::Event*>(e);
  std::lock_guard<std::mutex> lk(S(s)->mu);
  ev->s = S(s);
  ev->seq = S(s)->enqueued;
  return hipSuccess;
}
hipError_t hipEventQuery(hipEvent_t e) {
  auto* ev = reinterpret_cast<model::Event*>(e);
  std::lock_guard<std::mutex> lk(ev->s->mu);
  return ev->s->completed >= ev->seq ? hipSuccess : hipErrorNotReady;
}
hipError_t hipStreamSynchronize(hipStream_t s) {
  model::sync(S(s));
  return hipSuccess;
}
// device memory is host memory here; a copy from caller memory (pageable) is taken at once, any
// other copy runs in stream order
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t n, hipMemcpyKind kind, hipStream_t s) {
  if (kind == hipMemcpyHostToDevice) {
    std::vector<char> staged((const char*)src, (const char*)src + n);
    model::enqueue(S(s), [dst, staged]() { std::memcpy(dst, staged.data(), staged.size()); });
  } else {
    model::enqueue(S(s), [dst, src, n]() { std::memcpy(dst, src, n); });
  }
  return hipSuccess;
}
}  // extern "C"

// ------------------------------------------------------------------------------------------
// the library pieces the serving loop uses (errors, options, the launch layer)
// ------------------------------------------------------------------------------------------
namespace eik {
static thread_local std::string g_err;
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
int build_layouts_parallel(const std::vector<Mesh*>&) { return EIK_OK; }
int ensure_split(Mesh&) { return EIK_OK; }
int upload_layout(Mesh& m) { return m.dev ? EIK_OK : fail(EIK_ERR_STATE, "model mesh has no device state"); }

int launch_alloc(Launch& L, const std::vector<Mesh*>& kinds, int32_t cap_solves, int64_t cap_blocks,
                 int64_t cap_nv, int64_t cap_seeds) {
  L = Launch();
  for (Mesh* m : kinds) L.grp[m->dim == 3 ? 0 : 1].present = true;
  L.cap_solves = cap_solves;
  L.cap_blocks = cap_blocks;
  L.cap_nv = cap_nv;
  L.cap_seeds = std::max<int64_t>(cap_seeds, 1);
  L.stream = reinterpret_cast<hipStream_t>(model::new_stream());
  L.ev0 = reinterpret_cast<hipEvent_t>(new model::Event());
  L.ev1 = reinterpret_cast<hipEvent_t>(new model::Event());
  L.dsolves = new SolveDesc[cap_solves];
  L.hsolves = new SolveDesc[cap_solves];
  L.out = new float[cap_nv];
  L.h_out = new float[cap_nv];
  L.h_seed = new int32_t[3 * L.cap_seeds + 1];
  L.counters = new unsigned long long[kNumCounters];
  L.h_cnt = new unsigned long long[kNumCounters];
  return EIK_OK;
}

void launch_free(Launch& L) {
  if (L.stream) model::sync(S(L.stream));
  delete reinterpret_cast<model::Event*>(L.ev0);
  delete reinterpret_cast<model::Event*>(L.ev1);
  delete[] L.dsolves;
  delete[] L.hsolves;
  delete[] L.out;
  delete[] L.h_out;
  delete[] L.h_seed;
  delete[] L.counters;
  delete[] L.h_cnt;
  L = Launch();
}

void solve_desc(SolveDesc& s, Mesh* m, float* T, float* seen, const float* rmult) {
  std::memset(&s, 0, sizeof(s));
  s.T = T;
  s.seen = seen;
  s.rmult = rmult;
  s.nv = m->nv;
  s.nseen = m->dev->nseen;
  s.nblocks = m->dev->nblocks;
  s.dim = m->dim;
}

void record_factors(const Mesh& m, const float* mult, float* rmult, hipStream_t s) {
  const int64_t n = m.layout.total_elems, ne = m.ne;
  model::enqueue(S(s), [=]() {
    for (int64_t q = 0; q < n; ++q) rmult[q] = mult[q % ne];
  });
}

int launch_args(Launch& L, int32_t nsolves, int64_t total_blocks, const eik_solve_opts_t&,
                MultiArgs a[2]) {
  if (nsolves <= 0 || nsolves > L.cap_solves || total_blocks > L.cap_blocks)
    return fail(EIK_ERR_CAPACITY, "launch beyond its capacity");
  std::memset(a, 0, 2 * sizeof(MultiArgs));
  return EIK_OK;
}

// The launch's effects on the buffers the serving loop reads: descriptors in, then a "visit"
// (after a random delay) that writes every solve's gathered times as solve + vertex + 1 and
// checks the factors it was given, then the copies to the pinned output and counters.
int launch_issue(Launch& L, int32_t nsolves, int64_t, int64_t total_nv, const MultiArgs*,
                 bool upload_desc, bool output) {
  model::Stream* s = S(L.stream);
  if (upload_desc) {
    SolveDesc* d = L.dsolves;
    const SolveDesc* h = L.hsolves;
    model::enqueue(s, [d, h, nsolves]() { std::memcpy(d, h, sizeof(SolveDesc) * nsolves); });
  }
  SolveDesc* d = L.dsolves;
  float* out = L.out;
  unsigned long long* cnt = L.counters;
  model::enqueue(
      s,
      [d, out, cnt, nsolves, total_nv]() {
        int64_t off = 0;
        for (int32_t j = 0; j < nsolves; ++j) {
          float f = 1.f;
          if (d[j].rmult) f = d[j].rmult[0];
          for (int64_t v = 0; v < d[j].nv && off + v < total_nv; ++v)
            out[off + v] = (float)(j + v + 1) * f;
          off += d[j].nv;
        }
        for (int k = 0; k < kNumCounters; ++k) cnt[k] = 0;
      },
      true);
  unsigned long long* hc = L.h_cnt;
  model::enqueue(s, [hc, cnt]() { std::memcpy(hc, cnt, sizeof(unsigned long long) * kNumCounters); });
  return output ? launch_output(L, nsolves, total_nv) : EIK_OK;
}

// The gathered times to the pinned output buffer (the gather itself is part of the "visit").
int launch_output(Launch& L, int32_t, int64_t total_nv) {
  float* hout = L.h_out;
  float* out = L.out;
  model::enqueue(S(L.stream), [hout, out, total_nv]() { std::memcpy(hout, out, sizeof(float) * total_nv); });
  return EIK_OK;
}

int launch_check_end(const Launch& L) { return L.h_cnt[kCntAbandoned] ? fail(EIK_ERR_TIMEOUT, "abandoned") : EIK_OK; }
}  // namespace eik

extern "C" {
const char* eik_last_error(void) { return eik::g_err.c_str(); }
void eik_solve_opts_default(eik_solve_opts_t* o) { std::memset(o, 0, sizeof(*o)); }
}

// ------------------------------------------------------------------------------------------
// driver: Poisson-like submissions from several threads, waits from others
// ------------------------------------------------------------------------------------------
int main(int argc, char** argv) {
  const int nreq = argc > 1 ? std::atoi(argv[1]) : 400;
  const int nsub = argc > 2 ? std::atoi(argv[2]) : 3;
  const int launches = argc > 3 ? std::atoi(argv[3]) : 4;
  const int per_launch = argc > 4 ? std::atoi(argv[4]) : 8;
  const unsigned seed = argc > 5 ? (unsigned)std::atoi(argv[5]) : 1u;
  // three resident meshes of different sizes (a triangle mesh and two tet meshes)
  std::vector<eik::Mesh*> meshes;
  const int64_t nvs[3] = {500, 300, 900};
  for (int i = 0; i < 3; ++i) {
    eik::Mesh* m = new eik::Mesh();
    m->dim = i == 0 ? 2 : 3;
    m->nv = nvs[i];
    m->ne = 2 * nvs[i];
    m->nblocks = (int32_t)(nvs[i] / 64 + 1);
    m->layout.total_elems = 3 * nvs[i];
    m->dev = new eik::DeviceState();
    m->dev->nseen = nvs[i] + 100;
    m->dev->nblocks = m->nblocks;
    meshes.push_back(m);
  }
  eik_server_opts_t o;
  eik_server_opts_default(&o);
  o.launches = launches;
  o.max_solves_per_launch = per_launch;
  o.budget_bytes = 200000;  // a few requests at a time: admission holds the rest back
  eik_server_t srv = nullptr;
  std::vector<eik_mesh_t> hs;
  for (eik::Mesh* m : meshes) hs.push_back(reinterpret_cast<eik_mesh_t>(m));
  if (eik_server_create(hs.data(), 3, &o, &srv) != EIK_OK) {
    std::fprintf(stderr, "create: %s\n", eik_last_error());
    return 3;
  }
  struct Req {
    int mesh;
    std::vector<int32_t> seeds;
    std::vector<float> factors, out;
    int64_t id = -1;
  };
  std::vector<Req> reqs(nreq);
  std::mt19937 rng(seed);
  for (Req& r : reqs) {
    r.mesh = (int)(rng() % 3);
    r.seeds.assign(1 + rng() % 8, (int32_t)(rng() % nvs[r.mesh]));
    if (rng() % 2) r.factors.assign(meshes[r.mesh]->ne, 2.f);
    r.out.assign(nvs[r.mesh], -1.f);
  }
  std::atomic<int> submitted{0}, failures{0};
  std::vector<std::atomic<int64_t>> ids(nreq);
  for (auto& x : ids) x = -1;
  // submitters: request i goes from thread i % nsub after an exponential gap
  std::vector<std::thread> th;
  for (int t = 0; t < nsub; ++t)
    th.emplace_back([&, t]() {
      std::mt19937 r2(seed * 31 + t);
      std::exponential_distribution<double> gap(1.0 / 150.0);  // mean 150 us per thread
      for (int i = t; i < nreq; i += nsub) {
        std::this_thread::sleep_for(std::chrono::microseconds((int)gap(r2)));
        Req& q = reqs[i];
        int64_t id = -1;
        if (eik_server_submit(srv, q.mesh, q.seeds.data(), (int32_t)q.seeds.size(),
                              q.factors.empty() ? nullptr : q.factors.data(), q.out.data(), &id) != EIK_OK) {
          std::fprintf(stderr, "submit: %s\n", eik_last_error());
          ++failures;
        }
        ids[i] = id;
        ++submitted;
      }
    });
  // waiters: two threads wait on alternate requests as soon as they are submitted
  std::atomic<int> done{0};
  for (int w = 0; w < 2; ++w)
    th.emplace_back([&, w]() {
      for (int i = w; i < nreq; i += 2) {
        while (ids[i].load() < 0) std::this_thread::sleep_for(std::chrono::microseconds(50));
        eik_request_info_t info;
        if (eik_server_wait(srv, ids[i].load(), &info) != EIK_OK || info.status != 2) {
          std::fprintf(stderr, "wait %d: %s\n", i, eik_last_error());
          ++failures;
          continue;
        }
        // every value of the request's own slot: (its position in the launch) + v + 1, times
        // its factor; the position is unknown here, but consecutive values differ by the factor
        const Req& q = reqs[i];
        const float f = q.factors.empty() ? 1.f : 2.f;
        for (size_t v = 1; v < q.out.size(); ++v)
          if (q.out[v] - q.out[v - 1] != f) {
            std::fprintf(stderr, "request %d: wrong values at %zu\n", i, v);
            ++failures;
            break;
          }
        ++done;
      }
    });
  // watchdog: the whole stream must finish in bounded time
  std::atomic<bool> finished{false};
  std::thread dog([&]() {
    for (int k = 0; k < 600 && !finished; ++k) std::this_thread::sleep_for(std::chrono::milliseconds(100));
    if (!finished) {
      std::fprintf(stderr, "STALL: %d of %d requests done after 60 s\n", done.load(), nreq);
      std::_Exit(4);
    }
  });
  for (auto& t : th) t.join();
  finished = true;
  dog.join();
  int64_t nl = 0;
  eik_server_launches(srv, &nl);
  eik_server_destroy(srv);
  model::shutdown();
  std::printf("requests %d, launches %lld, failures %d\n", nreq, (long long)nl, failures.load());
  for (eik::Mesh* m : meshes) {
    delete m->dev;
    delete m;
  }
  return failures.load() ? 2 : 0;
}
