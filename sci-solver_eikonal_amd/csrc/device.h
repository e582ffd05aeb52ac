// device.h -- device state and the device-side pieces shared by the solve translation units:
// the per-mesh device state, the persistent work queue, LDS carving and visit tallies.
#pragma once

#include <hip/hip_runtime.h>

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

// counters[0..13]: visits, vertex updates, inner sweeps, local solves, algorithmic bytes,
// wall-clock ticks summed over workgroups: inside visits, waiting for work, then the parts of a
// visit: staging loads, sweeps, write-back, queueing of dependents; visits that followed the
// same workgroup's visit of the same block (re-runs); queue waits abandoned at the wait limit;
// algorithmic bytes without the element data of re-runs whose records were still staged
constexpr int kNumCounters = 14;
constexpr int kCntAbandoned = 12;
constexpr int kCntAlgoLoaded = 13;
// The one-launch-per-iteration kernel adds each visit's counters to one of kSyncSlots lines of
// 128 bytes (workgroup i to line i mod kSyncSlots); they are summed into counters after the
// solve.  Every visit adding to the same line serialised the peak iterations of the 64^3 cube
// (about 4,000 visits each) on that line's atomics.
constexpr int kSyncSlots = 128;
constexpr int kSyncSlotWords = 16;

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
  float* seen = nullptr;
  int64_t nseen = 0;
  int32_t* vh_start = nullptr;
  uint16_t* vh_pos = nullptr;
  uint64_t* vmask = nullptr;
  int32_t* rec_elem = nullptr;
  int32_t rec_lds = 0;
  size_t rec_lds_bytes = 0;
  int32_t* chg = nullptr;
  int32_t* self = nullptr;
  int32_t* list[2] = {nullptr, nullptr};
  int32_t* count = nullptr;  // active counts of even and odd iterations, last visited iteration
  unsigned long long* counters = nullptr;
  unsigned long long* sync_acc = nullptr;  // kSyncSlots * kSyncSlotWords
  int32_t* seeds = nullptr;
  float* seedvals = nullptr;
  int64_t seed_cap = 0;
  // pinned host staging, so the per-solve copies are plain DMA: the seeds then their values
  // (2 * seed_cap words) on the way in, the counters on the way out
  int32_t* h_seed = nullptr;
  unsigned long long* h_cnt = nullptr;
  // async scheduler
  int32_t* queue = nullptr;
  uint32_t* qflag = nullptr;
  uint32_t* bstate = nullptr;
  void* qctl = nullptr;  // QueueCtl
  int32_t qcap = 0;
  int32_t max_workgroups = 0;
  size_t lds_bytes = 0;         // a visit with its records staged (when rec_lds)
  size_t lds_bytes_direct = 0;  // a visit reading its records from memory
  int32_t queue_cap = 0;
  bool has_result = false;
  // split element records of the multi-solve visits that read them from memory (ensure_split)
  void* esel = nullptr;
  float* elen = nullptr;
  // per-outer-iteration snapshots of T (block-contiguous order) of the last sync solve with
  // record_history, kHistChunk snapshots per allocation
  std::vector<float*> hist;
  int64_t hist_len = 0;
};

constexpr int kHistChunk = 8;

template <typename T>
inline int upload(T** dst, const std::vector<T>& src, hipStream_t s) {
  size_t n = std::max<size_t>(src.size(), 1);
  HIP_TRY(hipMalloc((void**)dst, n * sizeof(T)));
  if (!src.empty())
    HIP_TRY(hipMemcpyAsync(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, s));
  return EIK_OK;
}

static inline unsigned grid1d(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }


// solver.hip
int ensure_stream(Mesh& m);
int upload_layout(Mesh& m);
size_t visit_lds_bytes(const Layout& L, int dim, int32_t queue_cap, size_t rec_bytes);
__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v);

// Carves the dynamic LDS of a visit: [records (rec_lds_bytes)] [sT] [sOld] [sChg] [sQ].
struct VisitLds {
  void* rec;
  float* T;
  float* old;
  int* chg;
  int* hidx;
  uint64_t* vm;
  uint16_t* q;
  int* mark;
};

template <int DIM>
__device__ __forceinline__ VisitLds carve_lds(void* base, size_t rec_bytes, int queue_cap,
                                              const BlockDesc& d) {
  VisitLds l;
  char* p = reinterpret_cast<char*>(base);
  l.rec = p;
  l.T = reinterpret_cast<float*>(p + rec_bytes);
  l.old = l.T + (d.nown + d.nhalo);
  l.chg = reinterpret_cast<int*>(l.old + d.nown);
  l.hidx = l.chg + (d.nown + d.nhalo);
  if constexpr (DIM == 2) {  // dependent masks of the owned vertices (triangle visits only)
    l.vm = reinterpret_cast<uint64_t*>((reinterpret_cast<uintptr_t>(l.hidx + d.nhalo) + 7) &
                                       ~static_cast<uintptr_t>(7));
    l.q = reinterpret_cast<uint16_t*>(l.vm + d.nown);
  } else {
    l.vm = nullptr;
    l.q = reinterpret_cast<uint16_t*>(l.hidx + d.nhalo);
  }
  l.mark = reinterpret_cast<int*>(l.q + ((queue_cap + 1) & ~1));
  return l;
}

// ------------------------------------------------------------------------------------------
// Persistent scheduler: one launch per solve.  Workgroups take blocks from a device ring buffer,
// visit them and queue the dependents of blocks whose boundary times decreased (and the block
// itself if it hit the inner-sweep cap).  A block sits in the queue at most once (qflag).
// `pending` counts queued plus in-flight blocks; the workgroup that brings it to zero ends the
// solve.  Replaces the reference's host loop (meshFIM3dEikonal.cu:411-542,
// meshFIM2d_eikonal.cu:683-816) and its five launches and blocking copies per iteration.
//
// Taking work is a ticket: one atomic add on `head` hands out the next slot index, and the
// workgroup then waits on that slot alone (slots are 64 bytes apart), so idle workgroups never
// poll a shared word.  The end of the solve sets the DONE bit of `head` with an atomic or; a
// ticket taken after that sees the bit, and the slots of tickets taken before it (all at or
// beyond the final tail, since nothing is queued any more) receive a DONE marker.
// ------------------------------------------------------------------------------------------
constexpr uint32_t kDoneBit = 0x80000000u;
constexpr int32_t kDoneMark = 0x7fffffff;
constexpr int kSlotStride = 16;  // ints between ring slots (64 bytes)
// A queue entry is (solve << bbits | g) + 1 (< kDoneMark): g indexes the per-block flags of all
// the blocks of a launch, `solve` the solve the block belongs to when one launch runs several.
// Each launch picks bbits (its arguments carry bbits and bmask = 2^bbits - 1): 31 for a single
// solve, just enough for the launch's blocks otherwise.

struct QueueCtl {
  uint32_t head;  // tickets handed out | kDoneBit
  uint32_t pad0[31];
  // (tail << 32) | pending: slots filled, and blocks queued or being visited; one word, so a
  // push raises both in one atomic and the count never drops to zero while work is queued
  unsigned long long tp;
  uint32_t pad1[30];
};

// Queue words and slots are device-scope atomics, coherent across XCDs.  No fences are needed:
// a visit's time write-backs are device-scope atomics that have returned (completed at the
// coherence point) before the workgroup barrier that precedes the push of its dependents, and
// a consumer issues its (device-scope) time loads only after it has read the slot.
template <class A>
__device__ __forceinline__ int32_t* q_slot(const A& a, uint32_t i) {
  return a.queue + (size_t)(i & a.qmask) * kSlotStride;
}

// Single-thread push (seeding).
template <class A>
__device__ __forceinline__ void q_push(const A& a, int32_t c) {
  if (atomicExch(&a.qflag[c & a.bmask], 1u) != 0u) return;  // already queued
  unsigned long long old = __hip_atomic_fetch_add(&a.ctl->tp, (1ull << 32) | 1ull,
                                                  __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(q_slot(a, (uint32_t)(old >> 32)), c + 1, __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// Block ownership.  bstate[b] counts the queue entries taken for block b since its current
// visit began (0 = idle).  The workgroup whose increment finds 0 visits the block; one that
// finds more drops its entry, and the visiting workgroup, which sees a count above 1 when it
// ends, visits the block again: two overlapping visits of one block would each converge against
// their own snapshot, and the merged result need not be converged.
template <class A>
__device__ __forceinline__ bool claim(const A& a, int32_t b) {
  return __hip_atomic_fetch_add(&a.bstate[b], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u;
}

// Ends a visit; returns true if the block must be visited again (entries were taken meanwhile).
template <class A>
__device__ __forceinline__ bool release_or_rerun(const A& a, int32_t b) {
  uint32_t expect = 1u;
  if (__hip_atomic_compare_exchange_strong(&a.bstate[b], &expect, 0u, __ATOMIC_RELAXED,
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    return false;
  // keep one count for the visit about to start; entries taken from now on add to it
  __hip_atomic_fetch_sub(&a.bstate[b], expect - 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

template <class A>
__device__ void q_finish(const A& a);

// q_take's result when its wait for a slot outlasted a.wait_limit (wall-clock ticks): the
// workgroup leaves the ticket and exits like at the end of the solve.  Every wait of a
// persistent kernel is thus bounded; the host, seeing the count, checks whether the queue still
// reached its end (then the result is complete) and otherwise reports the solve as unfinished.
constexpr int32_t kAbandoned = -3;

// Next block for this workgroup, claimed for a visit, -1 when the solve is finished, or
// kAbandoned.
template <class A>
__device__ __forceinline__ int32_t q_take(const A& a) {
  for (;;) {
    uint32_t r = __hip_atomic_fetch_add(&a.ctl->head, 1u, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
    if (r & kDoneBit) return -1;
    int32_t* slot = q_slot(a, r);
    int32_t v;
    int nap = 1;
    unsigned long long since = 0;
    while ((v = __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == 0) {
      const unsigned long long now = wall_clock64();
      if (since == 0) since = now;
      else if (now - since > a.wait_limit) return kAbandoned;
      __builtin_amdgcn_s_sleep(1);
      if (nap < 64) {
        for (int i = 0; i < nap; ++i) __builtin_amdgcn_s_sleep(1);
        nap <<= 1;
      } else {
        __builtin_amdgcn_s_sleep(63);
      }
    }
    __hip_atomic_store(slot, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (v == kDoneMark) return -1;
    const int32_t c = v - 1;
    const int32_t g = c & a.bmask;
    // Dequeued: a change made from now on must queue the block again.  The clear has to be
    // performed before this visit loads any time: a producer that writes times and then finds
    // the flag still set skips the push, so a clear still in flight while the loads run would
    // lose that update.  The claim is issued behind it and the wait (vmcnt covers stores on this
    // target) completes both.
    __hip_atomic_store(&a.qflag[g], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    const bool own = claim(a, g);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (own) return c;
    // the block is under visit, whose workgroup will visit it again: this entry is done
    unsigned long long old = __hip_atomic_fetch_add(&a.ctl->tp, ~0ull, __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
    if ((old & 0xffffffffull) == 1ull) q_finish(a);
  }
}

// Called by the workgroup whose decrement brought `pending` to zero.
template <class A>
__device__ void q_finish(const A& a) {
  uint32_t h = __hip_atomic_fetch_or(&a.ctl->head, kDoneBit, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT) & ~kDoneBit;
  uint32_t t = (uint32_t)(__hip_atomic_load(&a.ctl->tp, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT) >> 32);
  for (uint32_t i = t; (int32_t)(h - i) > 0; ++i)
    __hip_atomic_store(q_slot(a, i), kDoneMark, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wall-clock ticks of a queue-wait limit given in microseconds (0: 10 s).
inline unsigned long long wait_limit_ticks(int device, int32_t us) {
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) != hipSuccess || khz <= 0)
    khz = 100000;
  const unsigned long long u = us > 0 ? (unsigned long long)us : 10000000ull;
  return u * (unsigned long long)khz / 1000ull;
}

// After a launch whose workgroups abandoned queue waits: EIK_OK if the queue behind ctl still
// reached its end (every queued block was visited, so the result is complete), else
// EIK_ERR_TIMEOUT.  Synchronises stream s.
inline int check_queue_end(const QueueCtl* ctl, hipStream_t s, unsigned long long abandoned) {
  uint32_t head = 0;
  HIP_TRY(hipMemcpyAsync(&head, &ctl->head, sizeof(head), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (head & kDoneBit) return EIK_OK;
  return fail(EIK_ERR_TIMEOUT, std::to_string(abandoned) +
                                   " queue waits outlasted wait_limit_us and the solve did not "
                                   "finish (incomplete result)");
}

// Visit policy of the single-solve persistent kernel.  The pending count (blocks queued or under
// visit) returned by a workgroup's last push tells how busy the solve is.  Below
// kLatencyPending per resident workgroup the solve is bound by the latency of its chains of
// dependent visits rather than by throughput: a visit of a tet block runs one sweep, so the
// times it lowers reach its dependents sooner, and a block that has not converged is visited
// again at once by the same workgroup (records still staged, no queue round trip).  Above it, a
// visit runs up to max_inner sweeps and an unconverged block goes back to the queue tail, where
// it waits while its suppliers advance (re-running it at once would sweep ahead on stale halo
// values).  Triangle blocks keep max_inner sweeps in both regimes (measured faster).
// Multi-solve launches always use the queue: their many small solves keep it long, and
// immediate re-runs there only added visits.
constexpr unsigned kLatencyPending = 8;
template <int DIM>
__device__ __forceinline__ int latency_sweeps(int max_inner) {
  return DIM == 3 ? 1 : max_inner;
}

// Algorithmic HBM bytes of one block visit: the element data a visit needs (6 metric values and
// 4 local ids per tet = 32 B, 3 values and 3 ids plus padding per triangle = 20 B), the owned
// and halo times read, the owned times written, and one flag word.  The kernels' records are
// larger (corner heights for pruning) and a visit also touches `seen`, halo ids and its
// descriptor, while a re-run reads no records at all: the bytes actually moved are measured
// with hardware counters, and this figure is the common yardstick.
template <int DIM>
constexpr int kAlgoElemBytes = DIM == 3 ? 32 : 20;
__device__ __forceinline__ unsigned long long visit_algo_bytes(int elem_bytes, const BlockDesc& d) {
  return (unsigned long long)elem_bytes * d.nelem + 4ull * (d.nown + d.nhalo) + 4ull * d.nown + 4ull;
}

// Adds one visit to a workgroup's own tally (thread 0; flushed once when the workgroup exits,
// so the visits of a solve do not contend on shared counter words).  recs_staged: a re-run whose
// element records are still in LDS from the previous visit (it reads none from memory).
__device__ __forceinline__ void tally_visit(unsigned long long* acc, const BlockDesc& d,
                                            const VisitResult& r, unsigned long long solves,
                                            int elem_bytes, unsigned long long busy_ticks,
                                            bool rerun, bool recs_staged) {
  const unsigned long long ab = visit_algo_bytes(elem_bytes, d);
  acc[0] += 1;
  acc[1] += (unsigned long long)d.nown;
  acc[2] += (unsigned long long)r.sweeps;
  acc[3] += solves;
  acc[4] += ab;
  acc[kCntAlgoLoaded] += recs_staged ? ab - (unsigned long long)elem_bytes * d.nelem : ab;
  acc[11] += rerun ? 1ull : 0ull;
  acc[5] += busy_ticks;
  acc[7] += r.t_load;
  acc[8] += r.t_sweep;
  acc[9] += r.t_wb;
}

}  // namespace eik
