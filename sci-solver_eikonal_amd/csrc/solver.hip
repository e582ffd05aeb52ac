// solver.hip -- device state, kernels and the solve entry points of the C ABI.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "device.h"

namespace eik {

static int dev_free(DeviceState* d) {
  if (!d) return EIK_OK;
  for (float* p : d->hist) (void)hipFree(p);
  void* ptrs[] = {d->esel, d->elen, d->rec_elem, d->vmask, d->vh_start, d->vh_pos, d->seen, d->blocks, d->elems,  d->halo,  d->iperm,   d->sup_start, d->sup,
                  d->dep_start, d->dep, d->owner, d->T,      d->out,      d->chg,
                  d->self,   d->list[0], d->list[1], d->count, d->counters, d->sync_acc, d->seeds,
                  d->seedvals, d->queue, d->qflag, d->bstate, d->qctl};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (d->h_seed) (void)hipHostFree(d->h_seed);
  if (d->h_cnt) (void)hipHostFree(d->h_cnt);
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

int ensure_stream(Mesh& m) {
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

// ------------------------------------------------------------------------------------------
// Kernels
// ------------------------------------------------------------------------------------------
__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// The per-solve resets in one launch: segment k sets n[k] 32-bit words at p[k] to v[k]
// (a launch per array cost about 5 us each, most of a small solve's set-up).
struct FillSegs {
  static constexpr int kMax = 10;
  int32_t* p[kMax];
  int64_t n[kMax];
  int32_t v[kMax];
  int count;
  void add(void* ptr, int64_t words, int32_t value) {
    p[count] = reinterpret_cast<int32_t*>(ptr);
    n[count] = words;
    v[count] = value;
    ++count;
  }
};

__global__ void k_fill_segs(FillSegs f) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int k = 0; k < f.count; ++k) {
    int32_t* p = f.p[k];
    const int32_t v = f.v[k];
    for (int64_t i = i0; i < f.n[k]; i += stride) p[i] = v;
  }
}

// seeds: caller ids -> new ids; T = value; the owner block of each seed is marked changed
// (boundary and self) at stamp 0 so the first activation pass schedules it and its dependents.
// A seed outside [0, nv) is skipped (the entry point that takes seeds in device memory cannot
// check them on the host).  chg / self may be null (the persistent scheduler needs no stamps).
__global__ void k_seed(const int32_t* seeds, const float* vals, int64_t n, const int32_t* iperm,
                       const int32_t* owner, float* T, int32_t* chg, int32_t* self, int64_t nv) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || seeds[i] < 0 || seeds[i] >= nv) return;
  int32_t v = iperm[seeds[i]];
  atomicMin(reinterpret_cast<unsigned*>(&T[v]), __float_as_uint(vals ? vals[i] : 0.f));
  int32_t b = owner[v];
  if (chg) chg[b] = 0;
  if (self) self[b] = 0;
}

// Next active list: block c is active at iteration it+1 if it did not converge in iteration it
// or a block it depends on (a supplier of its halo) changed boundary values in iteration it.
// Replaces the reference's host loop over blockCon and m_BlockNeighbor
// (meshFIM3dEikonal.cu:445-488, 518-527; meshFIM2d_eikonal.cu:730-799).
// The workgroups from act_wgs on copy T to snap instead (the per-iteration history, when snap
// is set and the iteration just ended visited a block: *count_cur > 0), which saves a launch
// per iteration.
__global__ void k_activate(int32_t nb, const int32_t* __restrict__ sup_start,
                           const int32_t* __restrict__ sup, const int32_t* __restrict__ chg,
                           const int32_t* __restrict__ self, int32_t it, int32_t* __restrict__ list,
                           int32_t* __restrict__ count, int32_t act_wgs, const float* __restrict__ T,
                           int64_t nv, const int32_t* __restrict__ count_cur,
                           float* __restrict__ snap) {
  if ((int32_t)blockIdx.x >= act_wgs) {
    if (!snap || *count_cur <= 0) return;
    const int64_t stride = (int64_t)(gridDim.x - act_wgs) * blockDim.x;
    for (int64_t i = (int64_t)(blockIdx.x - act_wgs) * blockDim.x + threadIdx.x; i < nv; i += stride)
      snap[i] = T[i];
    return;
  }
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  bool act = false;
  if (c < nb) {
    // suppliers 16 at a time, all their loads in flight at once (a block of the 64^3 cube has
    // 13 on average): two memory round trips per batch instead of two per supplier
    constexpr int kB = 16;
    const int s1 = sup_start[c + 1];
    act = self[c] == it;
    for (int s = sup_start[c]; !act && s < s1; s += kB) {
      int32_t g[kB], v[kB];
#pragma unroll
      for (int j = 0; j < kB; ++j) g[j] = s + j < s1 ? sup[s + j] : -1;
#pragma unroll
      for (int j = 0; j < kB; ++j) v[j] = g[j] >= 0 ? chg[g[j]] : -2;  // stamps are >= -1
#pragma unroll
      for (int j = 0; j < kB; ++j) act |= v[j] == it;
    }
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
  size_t rec_lds_bytes;
  const int32_t* list;
  const int32_t* count;  // active blocks of this iteration (the grid may be larger)
  int32_t* count_next;   // cleared here: the activation after this iteration fills it
  int32_t* iters;        // outer iterations that visited at least one block
  int32_t it;
  int32_t* chg;
  int32_t* self;
  unsigned long long* acc;  // kSyncSlots slot lines of counters
  int32_t elem_bytes;  // algorithmic bytes per element (visit_algo_bytes)
};

// Adds one visit's counters to the workgroup's slot line (kSyncSlots lines, summed by
// k_sum_slots after the solve).
__device__ void block_counters(unsigned long long* acc, const BlockDesc& d, int sweeps,
                               uint64_t solves, int elem_bytes, unsigned long long busy_ticks,
                               const VisitResult& r, unsigned long long* sRed) {
  uint64_t s = solves;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (__lane_id() == 0) atomicAdd(sRed, (unsigned long long)s);
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long* c = acc + (blockIdx.x % kSyncSlots) * kSyncSlotWords;
    atomicAdd(&c[0], 1ull);
    atomicAdd(&c[1], (unsigned long long)d.nown);
    atomicAdd(&c[2], (unsigned long long)sweeps);
    atomicAdd(&c[3], *sRed);
    atomicAdd(&c[4], visit_algo_bytes(elem_bytes, d));
    atomicAdd(&c[kCntAlgoLoaded], visit_algo_bytes(elem_bytes, d));
    atomicAdd(&c[5], busy_ticks);
    atomicAdd(&c[7], r.t_load);
    atomicAdd(&c[8], r.t_sweep);
    atomicAdd(&c[9], r.t_wb);
  }
}

// counters[k] += the sum of word k over the kSyncSlots slot lines
__global__ void k_sum_slots(const unsigned long long* acc, unsigned long long* counters) {
  const int k = threadIdx.x;
  if (k >= kNumCounters) return;
  unsigned long long s = 0;
  for (int i = 0; i < kSyncSlots; ++i) s += acc[i * kSyncSlotWords + k];
  counters[k] += s;
}

// One outer iteration: workgroup i visits the i-th active block.  The host enqueues several
// iterations without waiting, so the grid covers every block and the workgroups beyond this
// iteration's active count leave at once (an iteration after convergence is empty).
template <int DIM>
__global__ __launch_bounds__(kBlockThreads) void k_visit_sync(SyncArgs a) {
  extern __shared__ uint4 lds4[];
  __shared__ int sFlag[4];
  __shared__ int sQn[2];  // work-item counters of alternate chunks
  __shared__ unsigned long long sRed;
  const int n = *a.count;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *a.count_next = 0;
    if (n > 0) *a.iters = a.it;
  }
  if ((int)blockIdx.x >= n) return;
  const int b = a.list[blockIdx.x];
  const BlockDesc d = a.v.blocks[b];
  VisitLds l = carve_lds<DIM>(lds4, a.rec_lds_bytes, a.v.queue_cap, d);
  if (threadIdx.x == 0) sRed = 0;
  uint64_t solves = 0;
  const unsigned long long t0 = wall_clock64();
  VisitResult r = visit_block<DIM, false>(a.v, b, l.rec, l.T, l.old, l.chg, l.hidx, l.vm, l.q, l.mark,
                                          sFlag, sQn, &solves);
  if (threadIdx.x == 0) {
    if (r.bnd_changed) a.chg[b] = a.it;
    if (!r.converged) a.self[b] = a.it;
  }
  block_counters(a.acc, d, r.sweeps, solves, a.elem_bytes, wall_clock64() - t0, r, &sRed);
}

struct AsyncArgs {
  VisitArgs v;
  size_t rec_lds_bytes;
  int32_t* queue;     // qcap slots of kSlotStride ints: 0 = empty, else block + 1 or kDoneMark
  uint32_t* qflag;    // per block: 1 while queued
  uint32_t* bstate;   // per block: entries taken since its current visit began (0 = idle)
  QueueCtl* ctl;
  uint32_t qmask;     // qcap - 1
  int32_t bmask;      // entries carry the block index only
  const int32_t* dep_start;
  const int32_t* dep;
  unsigned long long* counters;
  int32_t elem_bytes;  // algorithmic bytes per element (visit_algo_bytes)
  int32_t nblocks;
  int64_t max_visits;  // runaway guard: visits of one workgroup after which it queues nothing
  unsigned long long wait_limit;  // wall-clock ticks a queue wait may last (q_take)
};

// initial queue: the owner block of every seed and that block's dependents, one workgroup per
// seed (grid-stride beyond the grid) and one lane per block to queue (lane 0 the owner), so the
// pushes' device-scope round trips overlap instead of following each other
__global__ void k_async_seed(AsyncArgs a, const int32_t* seeds, int64_t n, const int32_t* iperm,
                             const int32_t* owner, int64_t nv) {
  for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
    if (seeds[i] < 0 || seeds[i] >= nv) continue;
    const int32_t b = owner[iperm[seeds[i]]];
    const int32_t s0 = a.dep_start[b], nd = a.dep_start[b + 1] - s0;
    for (int j = threadIdx.x; j <= nd; j += blockDim.x) q_push(a, j == 0 ? b : a.dep[s0 + j - 1]);
  }
}

static dim3 seed_grid(int64_t nseeds) {
  return dim3((unsigned)std::min<int64_t>(nseeds, 1 << 16));
}

template <int DIM>
__global__ __launch_bounds__(kBlockThreads) void k_visit_async(AsyncArgs a) {
  extern __shared__ uint4 lds4[];
  __shared__ int sFlag[4];
  __shared__ int sQn[2];  // work-item counters of alternate chunks
  __shared__ int sBlock;
  __shared__ unsigned long long sRed;
  __shared__ unsigned long long sAcc[kNumCounters];
  __shared__ int sAgain;
  __shared__ int sNext;  // block taken while the previous visit queued its dependents, or -2
  __shared__ unsigned sPend;  // pending count returned by this workgroup's last push
  if (threadIdx.x < kNumCounters) sAcc[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    sNext = -2;
    sPend = 0;
  }
  for (;;) {
    if (threadIdx.x == 0) {
      int nb = sNext;
      if (nb == -2) {
        const unsigned long long w0 = wall_clock64();
        nb = q_take(a);
        sAcc[6] += wall_clock64() - w0;
        if (nb == kAbandoned) sAcc[kCntAbandoned] += 1;
      }
      sBlock = nb;
      sNext = -2;
    }
    __syncthreads();
    const int b = sBlock;
    if (b < 0) break;
    bool rerun = false;      // the previous visit of this workgroup was of block b
    bool preloaded = false;  // ... and its times for this visit are in LDS already (rerun_load)
    do {
      if (threadIdx.x == 0) sRed = 0;
      const unsigned long long t0 = wall_clock64();
      const BlockDesc d = a.v.blocks[b];
      VisitLds l = carve_lds<DIM>(lds4, a.rec_lds_bytes, a.v.queue_cap, d);
      uint64_t solves = 0;
      // latency-bound regime (see kLatencyPending); sPend is only written after the barrier
      // that follows this read
      const bool latency = sPend < kLatencyPending * gridDim.x;
      const int inner = latency ? latency_sweeps<DIM>(a.v.max_inner) : 0;
      VisitResult r = visit_block<DIM, true>(a.v, b, l.rec, l.T, l.old, l.chg, l.hidx, l.vm, l.q, l.mark,
                                             sFlag, sQn, &solves, rerun, inner, preloaded);
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) solves += __shfl_down(solves, off, 64);
      // sRed is read by thread 0 after the next barrier
      if (__lane_id() == 0) atomicAdd(&sRed, (unsigned long long)solves);
      const unsigned long long tq = wall_clock64();
      // visits beyond the bound queue nothing more (runaway guard; converging solves stay far
      // below it)
      const bool cap_ok = (int64_t)sAcc[0] < a.max_visits;
      // items: the dependents whose halo holds a lowered boundary vertex, then the block itself
      // when it hit the sweep cap.  The first 64 dependents are flagged (qflag) while thread 0
      // releases the block: whether the block runs again decides if this visit's pending count
      // is dropped in the same atomic that queues the items, and whether the next block is
      // taken now.  In the latency-bound regime an unconverged block is not released: this
      // workgroup runs it again.
      const int32_t s0 = a.dep_start[b];
      const int ndep = r.bnd_changed && cap_ok ? a.dep_start[b + 1] - s0 : 0;
      bool want0 = false;
      int32_t c0 = b;
      if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        if (lane < ndep) {
          const uint64_t mask = a.v.vmask ? ((uint64_t)(uint32_t)l.mark[0] |
                                             ((uint64_t)(uint32_t)l.mark[1] << 32))
                                          : 0ull;
          want0 = a.v.vmask ? ((mask >> lane) & 1ull) != 0ull : l.mark[lane] != 0;
          c0 = a.dep[s0 + lane];
          if (want0) want0 = atomicExch(&a.qflag[c0], 1u) == 0u;
        }
        if (lane == 0)
          sAgain = (latency && !r.converged && cap_ok) || (release_or_rerun(a, b) && cap_ok);
      }
      __syncthreads();
      const int again = sAgain;
      if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        const uint64_t mask = a.v.vmask ? ((uint64_t)(uint32_t)l.mark[0] |
                                           ((uint64_t)(uint32_t)l.mark[1] << 32))
                                        : 0ull;
        const int nitems = ndep + 1;
        for (int j0 = 0; j0 < nitems; j0 += 64) {
          const int j = j0 + lane;
          bool want = false;
          int32_t c = b;
          if (j0 == 0 && j < ndep) {
            want = want0;
            c = c0;
          } else {
            if (j < ndep) {
              want = a.v.vmask ? ((mask >> j) & 1ull) != 0ull : l.mark[j] != 0;
              c = a.dep[s0 + j];
            } else if (j == ndep) {
              want = !r.converged && !again && cap_ok;  // a re-run continues the block anyway
            }
            if (want) want = atomicExch(&a.qflag[c], 1u) == 0u;
          }
          const uint64_t bal = __ballot(want);
          const unsigned n = (unsigned)__popcll(bal);
          const unsigned dec = (j0 + 64 >= nitems && !again) ? 1u : 0u;
          if (n == 0u && dec == 0u) continue;
          uint32_t base = 0;
          if (lane == 0) {
            // tail += n, pending += n - dec, in one 64-bit add (pending >= 1 while this visit
            // counts, so the low word cannot borrow)
            const unsigned long long inc = ((unsigned long long)n << 32) + n - dec;
            const unsigned long long old = __hip_atomic_fetch_add(
                &a.ctl->tp, inc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            base = (uint32_t)(old >> 32);
            sPend = (unsigned)(old & 0xffffffffull);
            if (n == 0u && (old & 0xffffffffull) == 1ull) q_finish(a);
          }
          base = __shfl(base, 0, 64);
          if (want)
            __hip_atomic_store(q_slot(a, base + (uint32_t)__popcll(bal & lane_mask_lt())), c + 1,
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) {
          tally_visit(sAcc, d, r, sRed, a.elem_bytes, tq - t0, rerun, rerun && a.v.rec_lds);
          sAcc[10] += wall_clock64() - tq;
        }
        if (again) {
          // the re-run starts with clear dependent marks (they have all been read above)
          if (a.v.vmask) {
            if (lane < 2) l.mark[lane] = 0;
          } else if (a.v.vh_start) {
            for (int i = lane; i < a.dep_start[b + 1] - s0; i += 64) l.mark[i] = 0;
          }
        }
      } else if (again) {
        // the other waves prepare the re-run while wave 0 queues the dependents
        rerun_load<true, kBlockThreads>(a.v, d, r.sweeps, l.T, l.old, l.chg, l.hidx, sFlag, 64);
      } else if (threadIdx.x == 64) {
        // the next block is taken while wave 0 queues the dependents: this visit is over (no
        // re-run), so the workgroup holds no work that the wait for a ticket could depend on
        const unsigned long long w0 = wall_clock64();
        const int32_t nx = q_take(a);
        sNext = nx;
        sAcc[6] += wall_clock64() - w0;
        if (nx == kAbandoned) sAcc[kCntAbandoned] += 1;
      }
      __syncthreads();
      rerun = true;
      preloaded = again;
    } while (sAgain);
  }
  if (threadIdx.x < kNumCounters && sAcc[threadIdx.x])
    atomicAdd(&a.counters[threadIdx.x], sAcc[threadIdx.x]);
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
  if constexpr (DIM == 3) {
    const float t0 = T[4 * i], t1 = T[4 * i + 1], t2 = T[4 * i + 2], t3 = T[4 * i + 3];
    const float* l = L + 6 * i;
    const float L0 = l[0], L1 = l[1], L2 = l[2], L3 = l[3], L4 = l[4], L5 = l[5];
    out[4 * i + 0] = __builtin_fminf(t0, tet_rotation(0, t0, t1, t2, t3, L0, L1, L2, L3, L4, L5, large));
    out[4 * i + 1] = __builtin_fminf(t1, tet_rotation(1, t0, t1, t2, t3, L0, L1, L2, L3, L4, L5, large));
    out[4 * i + 2] = __builtin_fminf(t2, tet_rotation(2, t0, t1, t2, t3, L0, L1, L2, L3, L4, L5, large));
    out[4 * i + 3] = __builtin_fminf(t3, tet_rotation(3, t0, t1, t2, t3, L0, L1, L2, L3, L4, L5, large));
  } else {
    const float t0 = T[3 * i], t1 = T[3 * i + 1], t2 = T[3 * i + 2];
    const float L0 = L[3 * i], L1 = L[3 * i + 1], L2 = L[3 * i + 2];
    out[3 * i + 0] = __builtin_fminf(t0, tri_rotation(0, t0, t1, t2, L0, L1, L2, large));
    out[3 * i + 1] = __builtin_fminf(t1, tri_rotation(1, t0, t1, t2, L0, L1, L2, large));
    out[3 * i + 2] = __builtin_fminf(t2, tri_rotation(2, t0, t1, t2, L0, L1, L2, large));
  }
}

// ------------------------------------------------------------------------------------------
// Host orchestration
// ------------------------------------------------------------------------------------------
size_t visit_lds_bytes(const Layout& L, int dim, int32_t queue_cap, size_t rec_bytes) {
  // records; times, stamps and halo indices of the local vertices; owned times at the start and
  // (triangles) the owned vertices' dependent masks
  return rec_bytes + sizeof(float) * ((size_t)L.max_local * 3 + L.max_own) +
         (dim == 2 ? sizeof(uint64_t) * ((size_t)L.max_own + 1) : 0) +
         sizeof(uint16_t) * (((size_t)queue_cap + 1) & ~(size_t)1) +
         sizeof(int32_t) * std::max<size_t>(L.max_dep, 2);
}

int upload_layout(Mesh& m) {
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
  // work queue: all (element, corner) pairs of the largest block if that fits, else chunks of
  // 2048; element records are staged in LDS when the whole visit fits the per-workgroup LDS
  int max_lds = 65536;
  {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, d->device) ==
            hipSuccess && v > 0)
      max_lds = v - 1024;  // static shared variables of the kernels
  }
  const int K = m.dim == 3 ? 4 : 3;
  const size_t rb = m.dim == 3 ? sizeof(TetRec) : sizeof(TriRec);
  int32_t qcap = std::min(8192, std::max(256, K * L.max_elem));
  size_t rec_bytes = ((size_t)L.max_elem * rb + 15) & ~(size_t)15;
  size_t lds = visit_lds_bytes(L, m.dim, qcap, rec_bytes);
  d->rec_lds = 1;
  if (lds > (size_t)max_lds) {
    d->rec_lds = 0;
    rec_bytes = 0;
    qcap = std::min(qcap, 2048);
    lds = visit_lds_bytes(L, m.dim, qcap, 0);
  }
  if (lds > (size_t)max_lds)
    return fail(EIK_ERR_CAPACITY, "block too large for LDS staging (" + std::to_string(L.max_local) +
                                      " local vertices); use fewer vertices per block");
  if (lds > 65536) {
    // dynamic LDS beyond 64 KB (MI355X has 160 KB per CU) is opted into per kernel
    const void* ks[] = {(const void*)k_visit_sync<3>, (const void*)k_visit_sync<2>,
                        (const void*)k_visit_async<3>, (const void*)k_visit_async<2>};
    for (const void* k : ks)
      HIP_TRY(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  }
  d->lds_bytes = lds;
  d->lds_bytes_direct = visit_lds_bytes(L, m.dim, qcap, 0);
  d->rec_lds_bytes = rec_bytes;
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
  if ((rc = upload(&d->vh_start, L.vh_start, s))) return rc;
  if ((rc = upload(&d->vh_pos, L.vh_pos, s))) return rc;
  if (!L.vmask.empty())
    if ((rc = upload(&d->vmask, L.vmask, s))) return rc;
  if ((rc = upload(&d->rec_elem, L.rec_elem, s))) return rc;
  HIP_TRY(hipMalloc((void**)&d->T, sizeof(float) * m.nv));
  HIP_TRY(hipMalloc((void**)&d->out, sizeof(float) * m.nv));
  d->nseen = m.nv + (int64_t)L.halo.size();
  HIP_TRY(hipMalloc((void**)&d->seen, sizeof(float) * d->nseen));
  HIP_TRY(hipMalloc((void**)&d->chg, sizeof(int32_t) * L.nblocks));
  HIP_TRY(hipMalloc((void**)&d->self, sizeof(int32_t) * L.nblocks));
  HIP_TRY(hipMalloc((void**)&d->list[0], sizeof(int32_t) * L.nblocks));
  HIP_TRY(hipMalloc((void**)&d->list[1], sizeof(int32_t) * L.nblocks));
  HIP_TRY(hipMalloc((void**)&d->count, sizeof(int32_t) * 4));
  HIP_TRY(hipMalloc((void**)&d->counters, sizeof(unsigned long long) * kNumCounters));
  HIP_TRY(hipMalloc((void**)&d->sync_acc, sizeof(unsigned long long) * kSyncSlots * kSyncSlotWords));
  HIP_TRY(hipStreamSynchronize(s));
  m.device_dirty = false;
  return EIK_OK;
}

// Outer iterations are enqueued kSyncChunk at a time and the host reads the next iteration's
// active count once per chunk (the reference copies block flags back and forth and synchronises
// several times per iteration, meshFIM3dEikonal.cu:431-531).  Iterations enqueued after
// convergence find no active block and cost one near-empty launch each.
constexpr int kSyncChunk = 8;

static int solve_sync(Mesh& m, const eik_solve_opts_t& o, eik_solve_stats_t& st) {
  DeviceState* d = m.dev;
  hipStream_t s = d->stream;
  const int32_t nb = d->nblocks;
  const float large = m.dim == 3 ? EIK_LARGENUM_3D : EIK_LARGENUM_2D;
  // d->count (cleared with the solve's other arrays): [0], [1] the active counts of even and odd
  // iterations, [2] the last iteration that visited a block
  int32_t* cnt = d->count;
  // activation for iteration 1 from the seed stamps (stamp 0)
  const int32_t act_wgs = (int32_t)grid1d(nb, 256);
  hipLaunchKernelGGL(k_activate, dim3(act_wgs), dim3(256), 0, s, nb, d->sup_start, d->sup,
                     d->chg, d->self, 0, d->list[1], cnt + 1, act_wgs, nullptr, (int64_t)0,
                     nullptr, nullptr);
  HIP_TRY(hipGetLastError());
  SyncArgs a;
  a.v.blocks = d->blocks;
  a.v.elems = d->elems;
  a.v.halo = d->halo;
  a.v.T = d->T;
  a.v.large = large;
  a.v.eps = o.eps;
  a.v.max_inner = default_inner_sweeps(m.dim, o.max_inner);
  a.v.queue_cap = d->queue_cap;
  a.v.seen = d->seen;
  // Every block of an iteration is visited at once, so this kernel is throughput-bound: records
  // read from memory (smaller workgroups, more per CU) were never slower than staged ones and
  // up to 21% faster (96^3 cube 7.53 -> 5.91 ms, 64^3 2.42 -> 2.06, level-7 icospheres 8%);
  // AUTO reads them from memory.
  const bool stage = d->rec_lds && o.stage_records == EIK_STAGE_LDS;
  const size_t lds = stage ? d->lds_bytes : d->lds_bytes_direct;
  a.v.rec_lds = stage ? 1 : 0;
  a.v.vh_start = nullptr;  // the per-iteration activation works on block stamps
  a.v.vh_pos = nullptr;
  a.v.vmask = nullptr;
  a.v.rmult = nullptr;
  a.v.lens = nullptr;
  a.v.dep_start = d->dep_start;
  a.rec_lds_bytes = stage ? d->rec_lds_bytes : 0;
  a.chg = d->chg;
  a.self = d->self;
  a.acc = d->sync_acc;
  a.elem_bytes = m.dim == 3 ? kAlgoElemBytes<3> : kAlgoElemBytes<2>;
  a.iters = cnt + 2;
  // values only decrease, so the loop ends; the cap bounds it on pathological inputs
  const int64_t cap = o.max_iterations > 0 ? o.max_iterations : 1000000;
  int32_t* h = reinterpret_cast<int32_t*>(d->h_cnt);  // pinned, free until the counters return
  int64_t it = 0;
  HIP_TRY(hipEventRecord(d->evA, s));
  while (it < cap) {
    const int64_t n = std::min<int64_t>(kSyncChunk, cap - it);
    for (int64_t k = 0; k < n; ++k) {
      ++it;
      const int p = (int)(it & 1);
      a.list = d->list[p];
      a.count = cnt + p;
      a.count_next = cnt + (p ^ 1);
      a.it = (int32_t)it;
      if (m.dim == 3)
        hipLaunchKernelGGL(k_visit_sync<3>, dim3(nb), dim3(kBlockThreads), lds, s, a);
      else
        hipLaunchKernelGGL(k_visit_sync<2>, dim3(nb), dim3(kBlockThreads), lds, s, a);
      HIP_TRY(hipGetLastError());
      float* snap = nullptr;
      unsigned snap_wgs = 0;
      if (o.record_history) {
        const size_t slot = (size_t)(it - 1);
        while (d->hist.size() <= slot / kHistChunk) {
          float* c = nullptr;
          HIP_TRY(hipMalloc((void**)&c, sizeof(float) * kHistChunk * (size_t)m.nv));
          d->hist.push_back(c);
        }
        snap = d->hist[slot / kHistChunk] + (slot % kHistChunk) * (size_t)m.nv;
        snap_wgs = (unsigned)std::min<int64_t>(grid1d(m.nv, 256), 1024);
      }
      hipLaunchKernelGGL(k_activate, dim3((unsigned)act_wgs + snap_wgs), dim3(256), 0, s, nb,
                         d->sup_start, d->sup, d->chg, d->self, (int32_t)it, d->list[p ^ 1],
                         cnt + (p ^ 1), act_wgs, d->T, m.nv, cnt + p, snap);
      HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(d->evB, s));
    HIP_TRY(hipMemcpyAsync(h, cnt, sizeof(int32_t) * 3, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (h[(it + 1) & 1] == 0) break;  // the next iteration has no active block
  }
  hipLaunchKernelGGL(k_sum_slots, dim3(1), dim3(64), 0, s, d->sync_acc, d->counters);
  HIP_TRY(hipGetLastError());
  float vms = 0.f;
  HIP_TRY(hipEventElapsedTime(&vms, d->evA, d->evB));
  st.visit_ms = vms;
  st.outer_iterations = h[2];
  st.visit_launches = h[2];
  if (o.record_history) d->hist_len = h[2];
  return EIK_OK;
}

// Workgroups of the single-solve persistent kernel.  A workgroup for every queued block mostly
// visits blocks whose suppliers are still changing, which only makes more visits, while too few
// workgroups leave the fronts' blocks waiting; the grid is sized to the fronts.  Triangle
// meshes: the fronts are curves, about sqrt(blocks x sources) blocks, and triangle visits
// publish lowered times during their sweeps; three workgroups per front block, with three
// sweeps per visit, against every workgroup that fits with four: level-7 anisotropic icosphere
// with 4 sources 1.47 -> 1.23 ms (13.1M -> 5.3M local solves), 16 and 64 sources 5% faster,
// isotropic with 1 source 1.96 -> 1.72 ms, 257^2 square 1.27 -> 1.14 ms, 32^2 square 0.24 ->
// 0.21 ms, level-8 icosphere equal.  Tet meshes: surfaces, about blocks^(2/3) x sources^(1/3)
// blocks; five workgroups per front block measured best over cubes with one source: 32^3
// 0.59 -> 0.47 ms, 16^3 and 24^3 about 10% faster, from 48^3 on (the 64^3 cube included) every
// workgroup that fits (64^3: 1.17 ms at 768, 1.31 at 640, 1.59 at 512).  At least one per CU.
static int persistent_grid(int dim, size_t lds, int requested, int64_t nblocks, int64_t nseeds,
                           int* out) {
  int dev = 0, ncu = 0, per_cu = 0;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  if (dim == 3)
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_visit_async<3>, kBlockThreads, lds));
  else
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_visit_async<2>, kBlockThreads, lds));
  int resident = std::max(1, ncu * std::max(per_cu, 1));
  resident = std::min(resident, 64 * ncu);
  if (requested > 0) {
    *out = std::min(requested, resident);
    return EIK_OK;
  }
  const double nb = (double)std::max<int64_t>(nblocks, 1), ns = (double)std::max<int64_t>(nseeds, 1);
  const double front = dim == 2 ? 3.0 * std::sqrt(nb * ns) : 5.0 * std::cbrt(nb * nb * ns);
  *out = (int)std::min<double>(resident, std::max<double>(ncu, front));
  return EIK_OK;
}

// Inner sweeps per visit of the single-solve persistent kernel when the caller leaves the
// choice: triangle blocks, visited by the smaller grid above, measured best at 3 (icosphere
// with 4 sources at 256 workgroups: 1.24 vs 1.30 ms at 4); other schedulers and launches keep
// default_inner_sweeps.
static int32_t async_inner_sweeps(int dim, int32_t requested) {
  return requested > 0 ? requested : (dim == 2 ? 3 : default_inner_sweeps(dim, 0));
}

// The work queue of the single-solve persistent scheduler, allocated on first use.  Ring
// capacity: every block queued at once plus a waiting ticket per resident workgroup (at most 64
// per CU), twice over.
static int ensure_queue(DeviceState* d) {
  if (d->queue) return EIK_OK;
  int dev = 0, ncu = 0;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  d->max_workgroups = 64 * std::max(ncu, 1);
  uint32_t qcap = 1;
  while (qcap < 2u * ((uint32_t)d->nblocks + (uint32_t)d->max_workgroups) + 2u) qcap <<= 1;
  d->qcap = (int32_t)qcap;
  HIP_TRY(hipMalloc((void**)&d->queue, sizeof(int32_t) * kSlotStride * qcap));
  HIP_TRY(hipMalloc((void**)&d->qflag, sizeof(uint32_t) * d->nblocks));
  HIP_TRY(hipMalloc((void**)&d->bstate, sizeof(uint32_t) * d->nblocks));
  HIP_TRY(hipMalloc((void**)&d->qctl, sizeof(QueueCtl)));
  return EIK_OK;
}

// Per-solve device state of the persistent scheduler: the mesh's own buffers (eik_solve) or
// carved from the scratch array (eik_solve_noalloc).
struct AsyncBufs {
  float* T;
  float* seen;
  int32_t* queue;
  uint32_t* qflag;
  uint32_t* bstate;
  QueueCtl* ctl;
  unsigned long long* counters;
  uint32_t qcap;
};

static AsyncArgs async_args(const Mesh& m, const eik_solve_opts_t& o, const AsyncBufs& b,
                            bool stage) {
  const DeviceState* d = m.dev;
  AsyncArgs a;
  a.v.blocks = d->blocks;
  a.v.elems = d->elems;
  a.v.halo = d->halo;
  a.v.T = b.T;
  a.v.large = m.dim == 3 ? EIK_LARGENUM_3D : EIK_LARGENUM_2D;
  a.v.eps = o.eps;
  a.v.max_inner = async_inner_sweeps(m.dim, o.max_inner);
  a.v.queue_cap = d->queue_cap;
  a.v.seen = b.seen;
  a.v.rec_lds = stage ? 1 : 0;
  a.v.vh_start = d->vh_start;
  a.v.vh_pos = d->vh_pos;
  a.v.vmask = d->vmask;
  a.v.rmult = nullptr;
  a.v.lens = nullptr;
  a.v.dep_start = d->dep_start;
  a.rec_lds_bytes = stage ? d->rec_lds_bytes : 0;
  a.queue = b.queue;
  a.qflag = b.qflag;
  a.bstate = b.bstate;
  a.ctl = b.ctl;
  a.qmask = b.qcap - 1u;
  a.bmask = 0x7fffffff;
  a.dep_start = d->dep_start;
  a.dep = d->dep;
  a.counters = b.counters;
  a.elem_bytes = m.dim == 3 ? kAlgoElemBytes<3> : kAlgoElemBytes<2>;
  a.nblocks = d->nblocks;
  // runaway guard: a workgroup that has made this many visits queues no more work (its visits
  // still finish), so a solve ends even if it never converged; the bound is max_iterations
  // outer iterations' worth of visits of every block (1000 without a cap), far above what a
  // converging solve needs
  a.max_visits = (int64_t)(o.max_iterations > 0 ? o.max_iterations : 1000) * d->nblocks;
  a.wait_limit = wait_limit_ticks(d->device, o.wait_limit_us);
  return a;
}

static int solve_async(Mesh& m, const eik_solve_opts_t& o, int64_t nseeds, eik_solve_stats_t& st) {
  DeviceState* d = m.dev;
  hipStream_t s = d->stream;
  int rc0 = ensure_queue(d);
  if (rc0) return rc0;
  const bool stage = d->rec_lds && o.stage_records != EIK_STAGE_MEMORY;
  const size_t lds = stage ? d->lds_bytes : d->lds_bytes_direct;
  const AsyncBufs bufs = {d->T,     d->seen,   d->queue, d->qflag, d->bstate,
                          reinterpret_cast<QueueCtl*>(d->qctl), d->counters, (uint32_t)d->qcap};
  const AsyncArgs a = async_args(m, o, bufs, stage);
  // the queue, its flags, block states and control words were reset with the solve's other
  // arrays (eik_solve)
  hipLaunchKernelGGL(k_async_seed, seed_grid(nseeds), dim3(64), 0, s, a, d->seeds, nseeds,
                     d->iperm, d->owner, m.nv);
  HIP_TRY(hipGetLastError());
  int grid = 0;
  int rc = persistent_grid(m.dim, lds, o.workgroups, d->nblocks, nseeds, &grid);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(d->evA, s));
  if (m.dim == 3)
    hipLaunchKernelGGL(k_visit_async<3>, dim3(grid), dim3(kBlockThreads), lds, s, a);
  else
    hipLaunchKernelGGL(k_visit_async<2>, dim3(grid), dim3(kBlockThreads), lds, s, a);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(d->evB, s));
  HIP_TRY(hipStreamSynchronize(s));
  float vms = 0.f;
  HIP_TRY(hipEventElapsedTime(&vms, d->evA, d->evB));
  st.visit_ms = vms;
  st.visit_launches = 1;
  st.outer_iterations = 1;
  return EIK_OK;
}

}  // namespace eik

using namespace eik;

extern "C" {

void eik_solve_opts_default(eik_solve_opts_t* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->scheduler = EIK_SCHED_ASYNC;
  o->max_iterations = 0;
  o->max_inner = 0;  // per element kind (default_inner_sweeps)
  o->eps = 0.f;
  o->record_history = 0;
  o->workgroups = 0;
  o->wait_limit_us = 0;  // 10 s
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
  if (o.wait_limit_us < 0) return fail(EIK_ERR_ARG, "wait_limit_us must be >= 0");
  int rc = upload_layout(*m);
  if (rc) return rc;
  DeviceState* d = m->dev;
  hipStream_t s = d->stream;
  auto t0 = std::chrono::steady_clock::now();
  const float large = m->dim == 3 ? EIK_LARGENUM_3D : EIK_LARGENUM_2D;
  // the times are reset below: until this solve succeeds there is no result to read
  d->has_result = false;
  // the staging buffers may still feed a copy of a solve that ended in an error
  HIP_TRY(hipStreamSynchronize(s));
  if (nseeds > d->seed_cap) {
    if (d->seeds) HIP_TRY(hipFree(d->seeds));
    if (d->seedvals) HIP_TRY(hipFree(d->seedvals));
    if (d->h_seed) HIP_TRY(hipHostFree(d->h_seed));
    d->seeds = nullptr;
    d->seedvals = nullptr;
    d->h_seed = nullptr;
    d->seed_cap = 0;
    HIP_TRY(hipMalloc((void**)&d->seeds, sizeof(int32_t) * nseeds));
    HIP_TRY(hipMalloc((void**)&d->seedvals, sizeof(float) * nseeds));
    HIP_TRY(hipHostMalloc((void**)&d->h_seed, sizeof(int32_t) * 2 * nseeds));
    d->seed_cap = nseeds;
  }
  if (!d->h_cnt) HIP_TRY(hipHostMalloc((void**)&d->h_cnt, sizeof(unsigned long long) * kNumCounters));
  d->hist_len = 0;
  int32_t* h_vals = d->h_seed + d->seed_cap;
  std::memcpy(d->h_seed, seeds, sizeof(int32_t) * nseeds);
  if (seed_vals) std::memcpy(h_vals, seed_vals, sizeof(float) * nseeds);
  HIP_TRY(hipEventRecord(d->ev0, s));
  HIP_TRY(hipMemcpyAsync(d->seeds, d->h_seed, sizeof(int32_t) * nseeds, hipMemcpyHostToDevice, s));
  if (seed_vals)
    HIP_TRY(hipMemcpyAsync(d->seedvals, h_vals, sizeof(float) * nseeds, hipMemcpyHostToDevice, s));
  if (o.scheduler == EIK_SCHED_ASYNC && (rc = ensure_queue(d))) return rc;
  {
    FillSegs f;
    f.count = 0;
    f.add(d->T, m->nv, (int32_t)__builtin_bit_cast(uint32_t, large));
    f.add(d->seen, d->nseen, (int32_t)__builtin_bit_cast(uint32_t, EIK_INF));
    f.add(d->chg, d->nblocks, -1);
    f.add(d->self, d->nblocks, -1);
    f.add(d->counters, 2 * kNumCounters, 0);
    f.add(d->count, 4, 0);
    if (o.scheduler != EIK_SCHED_ASYNC)
      f.add(d->sync_acc, 2 * kSyncSlots * kSyncSlotWords, 0);
    if (o.scheduler == EIK_SCHED_ASYNC) {
      f.add(d->queue, (int64_t)d->qcap * kSlotStride, 0);
      f.add(d->qflag, d->nblocks, 0);
      f.add(d->bstate, d->nblocks, 0);
      f.add(d->qctl, (int64_t)(sizeof(QueueCtl) / 4), 0);
    }
    int64_t most = 0;
    for (int k = 0; k < f.count; ++k) most = std::max(most, f.n[k]);
    hipLaunchKernelGGL(k_fill_segs, dim3((unsigned)std::min<int64_t>(grid1d(most, 256), 2048)),
                       dim3(256), 0, s, f);
  }
  hipLaunchKernelGGL(k_seed, dim3(grid1d(nseeds, 256)), dim3(256), 0, s, d->seeds,
                     seed_vals ? d->seedvals : nullptr, nseeds, d->iperm, d->owner, d->T, d->chg,
                     d->self, m->nv);
  HIP_TRY(hipGetLastError());
  eik_solve_stats_t st;
  std::memset(&st, 0, sizeof(st));
  rc = o.scheduler == EIK_SCHED_ASYNC ? solve_async(*m, o, nseeds, st) : solve_sync(*m, o, st);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(d->ev1, s));
  HIP_TRY(hipMemcpyAsync(d->h_cnt, d->counters, sizeof(unsigned long long) * kNumCounters,
                         hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  const unsigned long long* cnt = d->h_cnt;
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
  st.algo_bytes_loaded = (int64_t)cnt[kCntAlgoLoaded];
  {
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, d->device) == hipSuccess &&
        khz > 0) {
      st.busy_ms = (double)cnt[5] / khz;
      st.wait_ms = (double)cnt[6] / khz;
      st.load_ms = (double)cnt[7] / khz;
      st.sweep_ms = (double)cnt[8] / khz;
      st.writeback_ms = (double)cnt[9] / khz;
      st.queue_ms = (double)cnt[10] / khz;
    }
  }
  st.reruns = (int64_t)cnt[11];
  st.abandoned_waits = (int64_t)cnt[kCntAbandoned];
  st.history_len = d->hist_len;
  if (st.abandoned_waits > 0) {
    if (stats) *stats = st;
    rc = check_queue_end(reinterpret_cast<const QueueCtl*>(d->qctl), s, cnt[kCntAbandoned]);
    if (rc) return rc;
  }
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
  if (!m->dev || !m->dev->has_result) return fail(EIK_ERR_STATE, "no completed solve (none has run, or the last one failed)");
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
  DeviceState* d = m->dev;
  if (!d || !d->has_result || i < 0 || i >= d->hist_len)
    return fail(EIK_ERR_ARG, "history index out of range");
  const float* snap = d->hist[(size_t)i / kHistChunk] + ((size_t)i % kHistChunk) * (size_t)m->nv;
  hipLaunchKernelGGL(k_gather, dim3(grid1d(m->nv, 256)), dim3(256), 0, d->stream, snap, d->iperm,
                     m->nv, d->out);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(T_out, d->out, sizeof(float) * m->nv, hipMemcpyDeviceToHost, d->stream));
  HIP_TRY(hipStreamSynchronize(d->stream));
  return EIK_OK;
}

int eik_debug_local_solve(int32_t dim, const float* T, const float* L, int64_t n, float* out) {
  if ((dim != 2 && dim != 3) || !T || !L || !out || n <= 0) return fail(EIK_ERR_ARG, "bad argument");
  const int K = dim == 3 ? 4 : 3, NL = dim == 3 ? 6 : 3;
  // the three buffers are freed on every return path
  struct Bufs {
    float *T = nullptr, *L = nullptr, *O = nullptr;
    ~Bufs() {
      for (float* p : {T, L, O})
        if (p) (void)hipFree(p);
    }
  } b;
  HIP_TRY(hipMalloc((void**)&b.T, sizeof(float) * K * n));
  HIP_TRY(hipMalloc((void**)&b.L, sizeof(float) * NL * n));
  HIP_TRY(hipMalloc((void**)&b.O, sizeof(float) * K * n));
  HIP_TRY(hipMemcpy(b.T, T, sizeof(float) * K * n, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(b.L, L, sizeof(float) * NL * n, hipMemcpyHostToDevice));
  if (dim == 3)
    hipLaunchKernelGGL(k_debug_local<3>, dim3(grid1d(n, 256)), dim3(256), 0, 0, b.T, b.L, n, b.O,
                       EIK_LARGENUM_3D);
  else
    hipLaunchKernelGGL(k_debug_local<2>, dim3(grid1d(n, 256)), dim3(256), 0, 0, b.T, b.L, n, b.O,
                       EIK_LARGENUM_2D);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpy(out, b.O, sizeof(float) * K * n, hipMemcpyDeviceToHost));
  return EIK_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// The solve that allocates nothing
// ------------------------------------------------------------------------------------------
namespace eik {

// Per-solve state of eik_solve_noalloc (times, seen, queue ring, flags, counters), one array per
// device for the life of the process.  Host code never uses the array's name as an address: the
// device address comes from hipGetSymbolAddress, per device, on the first call on that device.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static void* g_scratch[kMaxDevices];
static std::mutex g_scratch_mu;

static int scratch_address(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(EIK_ERR_HIP, "device index out of range");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(EIK_ERR_HIP, std::string("scratch address: ") + hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  *out = reinterpret_cast<unsigned char*>(g_scratch[dev]);
  return EIK_OK;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace eik

extern "C" {

int eik_solve_noalloc(eik_mesh_t h, const int32_t* seeds_dev, int64_t nseeds, float* T_out_dev,
                      void* stream) {
  Mesh* m = reinterpret_cast<Mesh*>(h);
  if (!m || !seeds_dev || !T_out_dev || nseeds <= 0) return fail(EIK_ERR_ARG, "bad argument");
  DeviceState* d = m->dev;
  if (m->nblocks <= 0 || !d || !d->blocks || m->device_dirty || !m->layout.valid)
    return fail(EIK_ERR_STATE, "the mesh layout must be uploaded first (eik_mesh_upload)");
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev != d->device)
    return fail(EIK_ERR_STATE, "the current device is not the one the mesh was uploaded to");
  int ncu = 0;
  HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  uint32_t qcap = 1;
  while (qcap < 2u * ((uint32_t)d->nblocks + 64u * (uint32_t)std::max(ncu, 1)) + 2u) qcap <<= 1;
  // [control block][counters][queued flags][block states][ring][T][seen], 256-byte aligned
  const size_t o_cnt = align256(sizeof(QueueCtl));
  const size_t o_flag = o_cnt + align256(sizeof(unsigned long long) * kNumCounters);
  const size_t o_state = o_flag + align256(sizeof(uint32_t) * (size_t)d->nblocks);
  const size_t o_ring = o_state + align256(sizeof(uint32_t) * (size_t)d->nblocks);
  const size_t o_T = o_ring + align256(sizeof(int32_t) * kSlotStride * (size_t)qcap);
  const size_t o_seen = o_T + align256(sizeof(float) * (size_t)m->nv);
  const size_t need = o_seen + align256(sizeof(float) * (size_t)d->nseen);
  if (need > kScratchBytes)
    return fail(EIK_ERR_CAPACITY, "this mesh needs " + std::to_string(need) +
                                      " bytes of solve state; the allocation-free solve has " +
                                      std::to_string(kScratchBytes) + " (use eik_solve)");
  unsigned char* base = nullptr;
  int rc = scratch_address(dev, &base);
  if (rc) return rc;
  AsyncBufs b;
  b.ctl = reinterpret_cast<QueueCtl*>(base);
  b.counters = reinterpret_cast<unsigned long long*>(base + o_cnt);
  b.qflag = reinterpret_cast<uint32_t*>(base + o_flag);
  b.bstate = reinterpret_cast<uint32_t*>(base + o_state);
  b.queue = reinterpret_cast<int32_t*>(base + o_ring);
  b.T = reinterpret_cast<float*>(base + o_T);
  b.seen = reinterpret_cast<float*>(base + o_seen);
  b.qcap = qcap;
  hipStream_t s = stream ? reinterpret_cast<hipStream_t>(stream) : d->stream;
  eik_solve_opts_t o;
  eik_solve_opts_default(&o);
  const bool stage = d->rec_lds != 0;
  const size_t lds = stage ? d->lds_bytes : d->lds_bytes_direct;
  const float large = m->dim == 3 ? EIK_LARGENUM_3D : EIK_LARGENUM_2D;
  // the array keeps the previous call's contents: everything the solve reads first is reset,
  // with a fill kernel (no memset nodes if the caller captures this into a graph)
  FillSegs f;
  f.count = 0;
  f.add(b.T, m->nv, (int32_t)__builtin_bit_cast(uint32_t, large));
  f.add(b.seen, d->nseen, (int32_t)__builtin_bit_cast(uint32_t, EIK_INF));
  f.add(b.counters, 2 * kNumCounters, 0);
  f.add(b.queue, (int64_t)qcap * kSlotStride, 0);
  f.add(b.qflag, d->nblocks, 0);
  f.add(b.bstate, d->nblocks, 0);
  f.add(b.ctl, (int64_t)(sizeof(QueueCtl) / 4), 0);
  int64_t most = 0;
  for (int k = 0; k < f.count; ++k) most = std::max(most, f.n[k]);
  hipLaunchKernelGGL(k_fill_segs, dim3((unsigned)std::min<int64_t>(grid1d(most, 256), 2048)),
                     dim3(256), 0, s, f);
  hipLaunchKernelGGL(k_seed, dim3(grid1d(nseeds, 256)), dim3(256), 0, s, seeds_dev, nullptr,
                     nseeds, d->iperm, d->owner, b.T, nullptr, nullptr, m->nv);
  const AsyncArgs a = async_args(*m, o, b, stage);
  hipLaunchKernelGGL(k_async_seed, seed_grid(nseeds), dim3(64), 0, s, a, seeds_dev, nseeds,
                     d->iperm, d->owner, m->nv);
  int grid = 0;
  if ((rc = persistent_grid(m->dim, lds, 0, d->nblocks, nseeds, &grid))) return rc;
  if (m->dim == 3)
    hipLaunchKernelGGL(k_visit_async<3>, dim3(grid), dim3(kBlockThreads), lds, s, a);
  else
    hipLaunchKernelGGL(k_visit_async<2>, dim3(grid), dim3(kBlockThreads), lds, s, a);
  hipLaunchKernelGGL(k_gather, dim3(grid1d(m->nv, 256)), dim3(256), 0, s, b.T, d->iperm, m->nv,
                     T_out_dev);
  HIP_TRY(hipGetLastError());
  return EIK_OK;
}

}  // extern "C"
