// solver.hip -- device state, kernels and the solve entry points of the C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <string>
#include <vector>

#include "eik_internal.h"
#include "fim_kernels.h"

namespace eik {

#define HIP_TRY(expr)                                                                    \
  do {                                                                                   \
    hipError_t _e = (expr);                                                              \
    if (_e != hipSuccess)                                                                \
      return fail(EIK_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));      \
  } while (0)

// counters[0..4]: visits, vertex updates, inner sweeps, local solves, algorithmic bytes
constexpr int kNumCounters = 5;

struct DeviceState {
  int device = -1;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t evA = nullptr, evB = nullptr;  // around each block-visit launch
  int dim = 3;
  int32_t nblocks = 0;
  int64_t nv = 0;
  BlockDesc* blocks = nullptr;
  void* elems = nullptr;
  int32_t* halo = nullptr;
  int32_t* iperm = nullptr;
  int32_t* sup_start = nullptr;
  int32_t* sup = nullptr;
  int32_t* dep_start = nullptr;
  int32_t* dep = nullptr;
  int32_t* owner = nullptr;
  float* T = nullptr;
  float* out = nullptr;
  int32_t* chg = nullptr;
  int32_t* self = nullptr;
  int32_t* list[2] = {nullptr, nullptr};
  int32_t* count = nullptr;  // 2 ints
  unsigned long long* counters = nullptr;
  int32_t* seeds = nullptr;
  float* seedvals = nullptr;
  int64_t seed_cap = 0;
  // async scheduler
  int32_t* queue = nullptr;
  uint32_t* qflag = nullptr;
  int32_t* qctl = nullptr;  // head, tail, pending, done
  int32_t qcap = 0;
  size_t lds_bytes = 0;
  int32_t queue_cap = 0;
  bool has_result = false;
};

static int dev_free(DeviceState* d) {
  if (!d) return EIK_OK;
  void* ptrs[] = {d->blocks, d->elems,  d->halo,  d->iperm,   d->sup_start, d->sup,
                  d->dep_start, d->dep, d->owner, d->T,      d->out,      d->chg,
                  d->self,   d->list[0], d->list[1], d->count, d->counters, d->seeds,
                  d->seedvals, d->queue, d->qflag, d->qctl};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (d->ev0) (void)hipEventDestroy(d->ev0);
  if (d->ev1) (void)hipEventDestroy(d->ev1);
  if (d->evA) (void)hipEventDestroy(d->evA);
  if (d->evB) (void)hipEventDestroy(d->evB);
  if (d->stream) (void)hipStreamDestroy(d->stream);
  delete d;
  return EIK_OK;
}

void device_release(Mesh& m) {
  dev_free(m.dev);
  m.dev = nullptr;
}

static int ensure_stream(Mesh& m) {
  if (!m.dev) m.dev = new DeviceState();
  DeviceState* d = m.dev;
  if (!d->stream) {
    int ndev = 0;
    HIP_TRY(hipGetDeviceCount(&ndev));
    if (ndev <= 0) return fail(EIK_ERR_HIP, "no HIP device");
    HIP_TRY(hipGetDevice(&d->device));
    HIP_TRY(hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&d->ev0));
    HIP_TRY(hipEventCreate(&d->ev1));
    HIP_TRY(hipEventCreate(&d->evA));
    HIP_TRY(hipEventCreate(&d->evB));
  }
  return EIK_OK;
}

template <typename T>
static int upload(T** dst, const std::vector<T>& src, hipStream_t s) {
  size_t n = std::max<size_t>(src.size(), 1);
  HIP_TRY(hipMalloc((void**)dst, n * sizeof(T)));
  if (!src.empty())
    HIP_TRY(hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, s));
  return EIK_OK;
}

// ------------------------------------------------------------------------------------------
// Kernels
// ------------------------------------------------------------------------------------------
__global__ void k_fill(float* p, int64_t n, float v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// seeds: caller ids -> new ids; T = value; the owner block of each seed is marked changed
// (boundary and self) at stamp 0 so the first activation pass schedules it and its dependents.
__global__ void k_seed(const int32_t* seeds, const float* vals, int64_t n, const int32_t* iperm,
                       const int32_t* owner, float* T, int32_t* chg, int32_t* self) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int32_t v = iperm[seeds[i]];
  atomicMin(reinterpret_cast<unsigned*>(&T[v]), __float_as_uint(vals ? vals[i] : 0.f));
  int32_t b = owner[v];
  chg[b] = 0;
  self[b] = 0;
}

// Next active list: block c is active at iteration it+1 if it did not converge in iteration it
// or a block it depends on (a supplier of its halo) changed boundary values in iteration it.
// Replaces the reference's host loop over blockCon and m_BlockNeighbor
// (meshFIM3dEikonal.cu:445-488, 518-527; meshFIM2d_eikonal.cu:730-799).
__global__ void k_activate(int32_t nb, const int32_t* __restrict__ sup_start,
                           const int32_t* __restrict__ sup, const int32_t* __restrict__ chg,
                           const int32_t* __restrict__ self, int32_t it, int32_t* __restrict__ list,
                           int32_t* __restrict__ count) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  bool act = false;
  if (c < nb) {
    act = self[c] == it;
    for (int s = sup_start[c]; !act && s < sup_start[c + 1]; ++s) act = chg[sup[s]] == it;
  }
  uint64_t bal = __ballot(act);
  int n = __popcll(bal);
  int base = 0;
  if (__lane_id() == 0 && n) base = atomicAdd(count, n);
  base = __shfl(base, 0, 64);
  if (act) list[base + __popcll(bal & lane_mask_lt())] = c;
}

struct SyncArgs {
  VisitArgs v;
  const int32_t* list;
  int32_t it;
  int32_t* chg;
  int32_t* self;
  unsigned long long* counters;
  int32_t rec_bytes;
};

__device__ void block_counters(unsigned long long* counters, const BlockDesc& d, int sweeps,
                               uint64_t solves, int rec_bytes, unsigned long long* sRed) {
  // one atomic per counter per visit
  uint64_t s = solves;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (__lane_id() == 0) atomicAdd(sRed, (unsigned long long)s);
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(&counters[0], 1ull);
    atomicAdd(&counters[1], (unsigned long long)d.nown);
    atomicAdd(&counters[2], (unsigned long long)sweeps);
    atomicAdd(&counters[3], *sRed);
    unsigned long long bytes = (unsigned long long)rec_bytes * d.nelem + 8ull * d.nhalo +
                               8ull * d.nown + sizeof(BlockDesc);
    atomicAdd(&counters[4], bytes);
  }
}

template <int DIM>
__global__ __launch_bounds__(kBlockThreads) void k_visit_sync(SyncArgs a) {
  extern __shared__ float lds[];
  __shared__ int sFlag[4];
  __shared__ int sQn;
  __shared__ unsigned long long sRed;
  const int b = a.list[blockIdx.x];
  const BlockDesc d = a.v.blocks[b];
  float* sT = lds;
  float* sOld = sT + (d.nown + d.nhalo);
  int* sChg = reinterpret_cast<int*>(sOld + d.nown);
  int* sQ = sChg + (d.nown + d.nhalo);
  if (threadIdx.x == 0) sRed = 0;
  uint64_t solves = 0;
  VisitResult r = visit_block<DIM>(a.v, b, sT, sOld, sChg, sQ, sFlag, &sQn, &solves);
  if (threadIdx.x == 0) {
    if (r.bnd_changed) a.chg[b] = a.it;
    if (!r.converged) a.self[b] = a.it;
  }
  block_counters(a.counters, d, r.sweeps, solves, a.rec_bytes, &sRed);
}

// T_out[caller id] = T[new id]
__global__ void k_gather(const float* T, const int32_t* iperm, int64_t nv, float* out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nv) out[i] = T[iperm[i]];
}

// Local-solver test kernel: one thread per element, every corner.
template <int DIM>
__global__ void k_debug_local(const float* T, const float* L, int64_t n, float* out, float large) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  constexpr int K = DIM == 3 ? 4 : 3, NL = DIM == 3 ? 6 : 3;
  float t[K], l[NL];
  for (int k = 0; k < K; ++k) t[k] = T[K * i + k];
  for (int k = 0; k < NL; ++k) l[k] = L[NL * i + k];
  for (int k = 0; k < K; ++k) {
    float v = ElemTraits<DIM>::solve(k, t, l, large);
    out[K * i + k] = __builtin_fminf(t[k], v);
  }
}

// ------------------------------------------------------------------------------------------
// Host orchestration
// ------------------------------------------------------------------------------------------
static size_t visit_lds_bytes(const Layout& L, int32_t queue_cap) {
  return sizeof(float) * ((size_t)L.max_local * 2 + L.max_own) + sizeof(int32_t) * queue_cap;
}

static int upload_layout(Mesh& m) {
  int rc = ensure_stream(m);
  if (rc) return rc;
  if (!m.layout.valid) {
    rc = build_layout(m);
    if (rc) return rc;
  }
  if (!m.device_dirty && m.dev->blocks) return EIK_OK;
  DeviceState* old = m.dev;
  // keep the stream and events, drop the buffers
  DeviceState* d = new DeviceState();
  d->device = old->device;
  d->stream = old->stream;
  d->ev0 = old->ev0;
  d->ev1 = old->ev1;
  d->evA = old->evA;
  d->evB = old->evB;
  old->stream = nullptr;
  old->ev0 = old->ev1 = old->evA = old->evB = nullptr;
  dev_free(old);
  m.dev = d;
  const Layout& L = m.layout;
  hipStream_t s = d->stream;
  d->dim = m.dim;
  d->nblocks = L.nblocks;
  d->nv = m.nv;
  // queue: one chunk of up to 2048 work items (8 KB), sized so LDS stays <= 64 KB
  int32_t qcap = 2048;
  size_t lds = visit_lds_bytes(L, qcap);
  if (lds > 65536)
    return fail(EIK_ERR_CAPACITY, "block too large for LDS staging (" + std::to_string(L.max_local) +
                                      " local vertices); use fewer vertices per block");
  d->lds_bytes = lds;
  d->queue_cap = qcap;
  if ((rc = upload(&d->blocks, L.blocks, s))) return rc;
  if (m.dim == 3) {
    if ((rc = upload(reinterpret_cast<TetRec**>(&d->elems), L.tets, s))) return rc;
  } else {
    if ((rc = upload(reinterpret_cast<TriRec**>(&d->elems), L.tris, s))) return rc;
  }
  if ((rc = upload(&d->halo, L.halo, s))) return rc;
  if ((rc = upload(&d->iperm, L.iperm, s))) return rc;
  if ((rc = upload(&d->sup_start, L.sup_start, s))) return rc;
  if ((rc = upload(&d->sup, L.sup, s))) return rc;
  if ((rc = upload(&d->dep_start, L.dep_start, s))) return rc;
  if ((rc = upload(&d->dep, L.dep, s))) return rc;
  if ((rc = upload(&d->owner, L.owner, s))) return rc;
  HIP_TRY(hipMalloc((void**)&d->T, sizeof(float) * m.nv));
  HIP_TRY(hipMalloc((void**)&d->out, sizeof(float) * m.nv));
  HIP_TRY(hipMalloc((void**)&d->chg, sizeof(int32_t) * L.nblocks));
  HIP_TRY(hipMalloc((void**)&d->self, sizeof(int32_t) * L.nblocks));
  HIP_TRY(hipMalloc((void**)&d->list[0], sizeof(int32_t) * L.nblocks));
  HIP_TRY(hipMalloc((void**)&d->list[1], sizeof(int32_t) * L.nblocks));
  HIP_TRY(hipMalloc((void**)&d->count, sizeof(int32_t) * 4));
  HIP_TRY(hipMalloc((void**)&d->counters, sizeof(unsigned long long) * kNumCounters));
  HIP_TRY(hipStreamSynchronize(s));
  m.device_dirty = false;
  return EIK_OK;
}

static inline unsigned grid1d(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

static int solve_sync(Mesh& m, const eik_solve_opts_t& o, eik_solve_stats_t& st) {
  DeviceState* d = m.dev;
  hipStream_t s = d->stream;
  const int32_t nb = d->nblocks;
  const float large = m.dim == 3 ? EIK_LARGENUM_3D : EIK_LARGENUM_2D;
  int32_t h_count = 0;
  int64_t it = 0;
  // activation for iteration 1 from the seed stamps (stamp 0)
  HIP_TRY(hipMemsetAsync(d->count, 0, sizeof(int32_t), s));
  hipLaunchKernelGGL(k_activate, dim3(grid1d(nb, 256)), dim3(256), 0, s, nb, d->sup_start, d->sup,
                     d->chg, d->self, 0, d->list[0], d->count);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(&h_count, d->count, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  SyncArgs a;
  a.v.blocks = d->blocks;
  a.v.elems = d->elems;
  a.v.halo = d->halo;
  a.v.T = d->T;
  a.v.large = large;
  a.v.eps = o.eps;
  a.v.max_inner = o.max_inner > 0 ? o.max_inner : 64;
  a.v.queue_cap = d->queue_cap;
  a.chg = d->chg;
  a.self = d->self;
  a.counters = d->counters;
  a.rec_bytes = m.dim == 3 ? (int)sizeof(TetRec) : (int)sizeof(TriRec);
  m.history.clear();
  int cur = 0;
  // values only decrease, so the loop ends; the cap bounds it on pathological inputs
  const int64_t cap = o.max_iterations > 0 ? o.max_iterations : 1000000;
  while (h_count > 0) {
    if (it >= cap) break;
    ++it;
    a.list = d->list[cur];
    a.it = (int32_t)it;
    HIP_TRY(hipEventRecord(d->evA, s));
    if (m.dim == 3)
      hipLaunchKernelGGL(k_visit_sync<3>, dim3(h_count), dim3(kBlockThreads), d->lds_bytes, s, a);
    else
      hipLaunchKernelGGL(k_visit_sync<2>, dim3(h_count), dim3(kBlockThreads), d->lds_bytes, s, a);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(d->evB, s));
    HIP_TRY(hipMemsetAsync(d->count, 0, sizeof(int32_t), s));
    hipLaunchKernelGGL(k_activate, dim3(grid1d(nb, 256)), dim3(256), 0, s, nb, d->sup_start,
                       d->sup, d->chg, d->self, (int32_t)it, d->list[cur ^ 1], d->count);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(&h_count, d->count, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    if (o.record_history) {
      std::vector<float> snap(m.nv);
      hipLaunchKernelGGL(k_gather, dim3(grid1d(m.nv, 256)), dim3(256), 0, s, d->T, d->iperm, m.nv,
                         d->out);
      HIP_TRY(hipMemcpyAsync(snap.data(), d->out, sizeof(float) * m.nv, hipMemcpyDeviceToHost, s));
      HIP_TRY(hipStreamSynchronize(s));
      m.history.push_back(std::move(snap));
    } else {
      HIP_TRY(hipStreamSynchronize(s));
    }
    float vms = 0.f;
    HIP_TRY(hipEventElapsedTime(&vms, d->evA, d->evB));
    st.visit_ms += vms;
    st.visit_launches += 1;
    cur ^= 1;
  }
  st.outer_iterations = it;
  return EIK_OK;
}

}  // namespace eik

using namespace eik;

extern "C" {

void eik_solve_opts_default(eik_solve_opts_t* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->scheduler = EIK_SCHED_SYNC;
  o->max_iterations = 0;
  o->max_inner = 64;
  o->eps = 0.f;
  o->record_history = 0;
  o->workgroups = 0;
}

int eik_mesh_stream(eik_mesh_t h, void** out) {
  Mesh* m = reinterpret_cast<Mesh*>(h);
  if (!m || !out) return fail(EIK_ERR_ARG, "null argument");
  int rc = ensure_stream(*m);
  if (rc) return rc;
  *out = (void*)m->dev->stream;
  return EIK_OK;
}

int eik_mesh_upload(eik_mesh_t h) {
  Mesh* m = reinterpret_cast<Mesh*>(h);
  if (!m) return fail(EIK_ERR_ARG, "null mesh");
  if (m->nblocks <= 0) return fail(EIK_ERR_STATE, "mesh is not partitioned");
  return upload_layout(*m);
}

int eik_solve(eik_mesh_t h, const int32_t* seeds, int64_t nseeds, const float* seed_vals,
              const eik_solve_opts_t* opts, float* T_out, eik_solve_stats_t* stats) {
  Mesh* m = reinterpret_cast<Mesh*>(h);
  if (!m) return fail(EIK_ERR_ARG, "null mesh");
  if (!seeds || nseeds <= 0) return fail(EIK_ERR_ARG, "at least one seed vertex is required");
  for (int64_t i = 0; i < nseeds; ++i) {
    if (seeds[i] < 0 || seeds[i] >= m->nv) return fail(EIK_ERR_ARG, "seed vertex out of range");
    if (seed_vals && !(seed_vals[i] >= 0.f && seed_vals[i] < 1e30f))
      return fail(EIK_ERR_ARG, "seed values must be finite and >= 0");
  }
  if (m->nblocks <= 0) return fail(EIK_ERR_STATE, "mesh is not partitioned");
  eik_solve_opts_t o;
  eik_solve_opts_default(&o);
  if (opts) o = *opts;
  if (o.scheduler != EIK_SCHED_SYNC && o.scheduler != EIK_SCHED_ASYNC)
    return fail(EIK_ERR_ARG, "unknown scheduler");
  if (o.scheduler == EIK_SCHED_ASYNC && o.record_history)
    return fail(EIK_ERR_ARG, "record_history needs the sync scheduler");
  int rc = upload_layout(*m);
  if (rc) return rc;
  DeviceState* d = m->dev;
  hipStream_t s = d->stream;
  auto t0 = std::chrono::steady_clock::now();
  const float large = m->dim == 3 ? EIK_LARGENUM_3D : EIK_LARGENUM_2D;
  if (nseeds > d->seed_cap) {
    if (d->seeds) HIP_TRY(hipFree(d->seeds));
    if (d->seedvals) HIP_TRY(hipFree(d->seedvals));
    d->seeds = nullptr;
    d->seedvals = nullptr;
    HIP_TRY(hipMalloc((void**)&d->seeds, sizeof(int32_t) * nseeds));
    HIP_TRY(hipMalloc((void**)&d->seedvals, sizeof(float) * nseeds));
    d->seed_cap = nseeds;
  }
  HIP_TRY(hipEventRecord(d->ev0, s));
  HIP_TRY(hipMemcpyAsync(d->seeds, seeds, sizeof(int32_t) * nseeds, hipMemcpyHostToDevice, s));
  if (seed_vals)
    HIP_TRY(hipMemcpyAsync(d->seedvals, seed_vals, sizeof(float) * nseeds, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_fill, dim3(grid1d(m->nv, 256)), dim3(256), 0, s, d->T, m->nv, large);
  hipLaunchKernelGGL(k_fill_i32, dim3(grid1d(d->nblocks, 256)), dim3(256), 0, s, d->chg,
                     (int64_t)d->nblocks, -1);
  hipLaunchKernelGGL(k_fill_i32, dim3(grid1d(d->nblocks, 256)), dim3(256), 0, s, d->self,
                     (int64_t)d->nblocks, -1);
  HIP_TRY(hipMemsetAsync(d->counters, 0, sizeof(unsigned long long) * kNumCounters, s));
  hipLaunchKernelGGL(k_seed, dim3(grid1d(nseeds, 256)), dim3(256), 0, s, d->seeds,
                     seed_vals ? d->seedvals : nullptr, nseeds, d->iperm, d->owner, d->T, d->chg,
                     d->self);
  HIP_TRY(hipGetLastError());
  eik_solve_stats_t st;
  std::memset(&st, 0, sizeof(st));
  rc = solve_sync(*m, o, st);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(d->ev1, s));
  unsigned long long cnt[kNumCounters];
  HIP_TRY(hipMemcpyAsync(cnt, d->counters, sizeof(cnt), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  auto t1 = std::chrono::steady_clock::now();
  float ms = 0.f;
  HIP_TRY(hipEventElapsedTime(&ms, d->ev0, d->ev1));
  st.kernel_ms = ms;
  st.wall_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  st.block_visits = (int64_t)cnt[0];
  st.vertex_updates = (int64_t)cnt[1];
  st.inner_sweeps = (int64_t)cnt[2];
  st.local_solves = (int64_t)cnt[3];
  st.algo_bytes = (int64_t)cnt[4];
  st.history_len = (int64_t)m->history.size();
  d->has_result = true;
  if (T_out) {
    hipLaunchKernelGGL(k_gather, dim3(grid1d(m->nv, 256)), dim3(256), 0, s, d->T, d->iperm, m->nv,
                       d->out);
    HIP_TRY(hipMemcpyAsync(T_out, d->out, sizeof(float) * m->nv, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
  if (stats) *stats = st;
  return EIK_OK;
}

int eik_get_result(eik_mesh_t h, float* T_out) {
  Mesh* m = reinterpret_cast<Mesh*>(h);
  if (!m || !T_out) return fail(EIK_ERR_ARG, "null argument");
  if (!m->dev || !m->dev->has_result) return fail(EIK_ERR_STATE, "no solve has run");
  DeviceState* d = m->dev;
  hipLaunchKernelGGL(k_gather, dim3(grid1d(m->nv, 256)), dim3(256), 0, d->stream, d->T, d->iperm,
                     m->nv, d->out);
  HIP_TRY(hipMemcpyAsync(T_out, d->out, sizeof(float) * m->nv, hipMemcpyDeviceToHost, d->stream));
  HIP_TRY(hipStreamSynchronize(d->stream));
  return EIK_OK;
}

int eik_get_history(eik_mesh_t h, int64_t i, float* T_out) {
  Mesh* m = reinterpret_cast<Mesh*>(h);
  if (!m || !T_out) return fail(EIK_ERR_ARG, "null argument");
  if (i < 0 || i >= (int64_t)m->history.size()) return fail(EIK_ERR_ARG, "history index out of range");
  std::memcpy(T_out, m->history[i].data(), sizeof(float) * m->nv);
  return EIK_OK;
}

int eik_debug_local_solve(int32_t dim, const float* T, const float* L, int64_t n, float* out) {
  if ((dim != 2 && dim != 3) || !T || !L || !out || n <= 0) return fail(EIK_ERR_ARG, "bad argument");
  const int K = dim == 3 ? 4 : 3, NL = dim == 3 ? 6 : 3;
  float *dT = nullptr, *dL = nullptr, *dO = nullptr;
  HIP_TRY(hipMalloc((void**)&dT, sizeof(float) * K * n));
  HIP_TRY(hipMalloc((void**)&dL, sizeof(float) * NL * n));
  HIP_TRY(hipMalloc((void**)&dO, sizeof(float) * K * n));
  HIP_TRY(hipMemcpy(dT, T, sizeof(float) * K * n, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dL, L, sizeof(float) * NL * n, hipMemcpyHostToDevice));
  if (dim == 3)
    hipLaunchKernelGGL(k_debug_local<3>, dim3(grid1d(n, 256)), dim3(256), 0, 0, dT, dL, n, dO,
                       EIK_LARGENUM_3D);
  else
    hipLaunchKernelGGL(k_debug_local<2>, dim3(grid1d(n, 256)), dim3(256), 0, 0, dT, dL, n, dO,
                       EIK_LARGENUM_2D);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpy(out, dO, sizeof(float) * K * n, hipMemcpyDeviceToHost));
  HIP_TRY(hipFree(dT));
  HIP_TRY(hipFree(dL));
  HIP_TRY(hipFree(dO));
  return EIK_OK;
}

}  // extern "C"
