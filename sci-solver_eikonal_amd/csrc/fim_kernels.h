// fim_kernels.h -- block-visit kernels of the FIM solve (device code, included by solver.hip).
//
// A block visit (the reference's FIMCuda, src/core3d/cuda/CUDA_kernels.cu:383-489, and
// src/core2d/cuda/CUDA_kernels.cu:36-409) stages the block's owned and halo vertex times in LDS,
// sweeps the block's elements until no owned time decreases (the reference runs a fixed 2 or 7
// sweeps and tests |old - new| < EPS), and writes the owned times back.  One value per vertex
// replaces the reference's per-corner slot copies, so there is no slot reconciliation
// (CUDA_kernels.cu:460-474) and no halo scatter kernel (CopyOutBack, :604-668): an update is an
// LDS atomic minimum on the target vertex, and neighbour blocks read the written times directly.
//
// Work inside a sweep is compacted: a wave first tests its elements' corners against the exact
// pruning rules (target owned and updatable, some other corner known and smaller than the
// target by more than the corner's height, and, after the first sweep, some corner of the
// element changed in the previous sweep) and appends the surviving (element, corner) pairs to an
// LDS queue with per-corner ballots; the local solves then run over the queue with every lane
// busy.
//
// Element data is handled as scalars (records loaded as 16-byte vectors and unpacked): selecting
// among elements of a local array would be turned into an indexed access of a scratch copy.
#pragma once

#include <hip/hip_runtime.h>

#include "eik_internal.h"
#include "local_solver.h"

namespace eik {

constexpr int kBlockThreads = 256;
// elements a thread tests per pass of a sweep's selection phase
constexpr int kSelBatch = 3;

struct VisitArgs {
  const BlockDesc* blocks;
  const void* elems;
  const int32_t* halo;
  float* T;
  float* seen;        // per block-local vertex: the time its block last accounted for
  float large;
  float eps;
  int32_t max_inner;
  int32_t queue_cap;  // (element, corner) entries the LDS queue holds per chunk
  int32_t rec_lds;    // element records are staged in LDS (block fits)
  // optional: per vertex, the positions in its owner's dependent list of the blocks holding it
  // in their halo; with them a visit marks exactly the dependents its boundary changes affect
  const int32_t* vh_start;
  const uint16_t* vh_pos;
  const int32_t* dep_start;
  // optional, preferred: the same as one 64-bit mask per vertex (bit k = k-th dependent); the
  // visit then collects the dependents to queue in one 64-bit LDS word (sMark[0..1])
  const uint64_t* vmask;
  // optional: a factor per record (in record order) on its squared edge lengths, a request's
  // own speed field over a resident mesh; applied while the records are staged, or at every
  // use when they are read from memory
  const float* rmult;
  // split records (visits that read them from memory at every use, SPLIT = true): `elems` then
  // holds the 16-byte selection records (SelRec) and `lens` the squared edge lengths, 6 floats
  // per tet / 4 (3 and a pad) per triangle, in the same order
  const float* lens;
};

struct VisitResult {
  bool converged;
  bool bnd_changed;
  int sweeps;
  // wall-clock ticks (thread 0): staging loads, sweeps, write-back
  unsigned long long t_load, t_sweep, t_wb;
};

// One element unpacked to scalars.  Triangles leave v3 / L3..L5 / H3 unused.
struct Elem {
  int v0, v1, v2, v3;
  unsigned mask;
  float L0, L1, L2, L3, L4, L5;
  float H0, H1, H2, H3;
};

// Record sizes in 16-byte words: tets {v01, v23, L0, L1} {L2, L3, L4, L5} {H0, H1, H2, H3},
// triangles {v01, v2 | mask, L0, L1} {L2, H0, H1, H2}.
template <int DIM>
constexpr int kRecWords = DIM == 3 ? 3 : 2;

// What a sweep phase reads of a record: the corner selection needs the corners and heights,
// the local solve the corners and edge lengths.  A tet record's second 16-byte word holds only
// lengths and its third only heights, so each phase skips one of them (a third of the bytes
// when the records are read from memory); a triangle record is read whole.
enum RecPart { kRecAll, kRecSelect, kRecSolve };

// Split records (multi-solve visits that read element data from memory at every use): the
// selection phase reads a 16-byte record of corners and heights, the solve phase the corners
// again and the lengths from their own array, so neither phase streams the other's data.
// Heights are fp16, rounded down on the host (still lower bounds of every solve).
__device__ __forceinline__ float h2f(uint32_t bits) {
  const uint16_t u = (uint16_t)bits;
  _Float16 h;
  __builtin_memcpy(&h, &u, 2);
  return (float)h;
}

template <int DIM>
__device__ __forceinline__ Elem load_sel(const void* sel, int e) {
  const uint4 q = reinterpret_cast<const uint4*>(sel)[e];
  Elem x;
  x.v0 = q.x & 0xffffu;
  x.v1 = q.x >> 16;
  x.v2 = q.y & 0xffffu;
  x.L0 = x.L1 = x.L2 = x.L3 = x.L4 = x.L5 = 0.f;
  x.H0 = h2f(q.z & 0xffffu);
  x.H1 = h2f(q.z >> 16);
  x.H2 = h2f(q.w & 0xffffu);
  if constexpr (DIM == 3) {
    x.v3 = q.y >> 16;
    x.mask = 0xfu;
    x.H3 = h2f(q.w >> 16);
  } else {
    x.v3 = 0;
    x.mask = q.y >> 16;
    x.H3 = 0.f;
  }
  return x;
}

template <int DIM>
__device__ __forceinline__ void load_len(const float* lens, int e, Elem& x) {
  if constexpr (DIM == 3) {
    const float2* p = reinterpret_cast<const float2*>(lens + 6 * (int64_t)e);
    const float2 a = p[0], b = p[1], c = p[2];
    x.L0 = a.x;
    x.L1 = a.y;
    x.L2 = b.x;
    x.L3 = b.y;
    x.L4 = c.x;
    x.L5 = c.y;
  } else {
    const float4 q = reinterpret_cast<const float4*>(lens)[e];
    x.L0 = q.x;
    x.L1 = q.y;
    x.L2 = q.z;
  }
}

template <int DIM, RecPart PART = kRecAll>
__device__ __forceinline__ Elem load_elem(const void* recs, int e) {
  Elem x;
  const uint4* p = reinterpret_cast<const uint4*>(recs) + kRecWords<DIM> * (int64_t)e;
  const uint4 q0 = p[0];
  const uint4 q1 = (DIM == 3 && PART == kRecSelect) ? make_uint4(0u, 0u, 0u, 0u) : p[1];
  x.v0 = q0.x & 0xffffu;
  x.v1 = q0.x >> 16;
  x.v2 = q0.y & 0xffffu;
  x.L0 = __uint_as_float(q0.z);
  x.L1 = __uint_as_float(q0.w);
  x.L2 = __uint_as_float(q1.x);
  if constexpr (DIM == 3) {
    const uint4 q2 = PART == kRecSolve ? make_uint4(0u, 0u, 0u, 0u) : p[2];
    x.v3 = q0.y >> 16;
    x.mask = 0xfu;
    x.L3 = __uint_as_float(q1.y);
    x.L4 = __uint_as_float(q1.z);
    x.L5 = __uint_as_float(q1.w);
    x.H0 = __uint_as_float(q2.x);
    x.H1 = __uint_as_float(q2.y);
    x.H2 = __uint_as_float(q2.z);
    x.H3 = __uint_as_float(q2.w);
  } else {
    x.v3 = 0;
    x.mask = q0.y >> 16;
    x.L3 = x.L4 = x.L5 = 0.f;
    x.H0 = __uint_as_float(q1.y);
    x.H1 = __uint_as_float(q1.z);
    x.H2 = __uint_as_float(q1.w);
    x.H3 = 0.f;
  }
  return x;
}

// Time-array access.  A visit inside one launch of the persistent scheduler runs beside visits
// of neighbouring blocks on other XCDs (separate L2s), so its loads and stores of the time array
// are agent-scope atomics, which the hardware keeps coherent across XCDs without cache
// write-back or invalidation fences; the owned write-back is an atomic minimum, whose returned
// value tells whether the stored time dropped.  The one-launch-per-iteration scheduler visits
// each block once per launch and uses plain accesses (the kernel boundary between iterations
// makes every write-back visible to the next).
template <bool COHERENT>
__device__ __forceinline__ float load_T(const float* p) {
  if constexpr (COHERENT)
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else
    return *p;
}

// Returns true if the stored value decreased.
template <bool COHERENT>
__device__ __forceinline__ bool store_min_T(float* p, float v, float old) {
  if constexpr (COHERENT) {
    unsigned prev = __hip_atomic_fetch_min(reinterpret_cast<unsigned*>(p), __float_as_uint(v),
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return v < __uint_as_float(prev);
  } else {
    *p = v;
    return v < old;
  }
}

// Number of set bits of m in the lanes below this one.
__device__ __forceinline__ int mbcnt64(uint64_t m) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                        __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

__device__ __forceinline__ uint64_t lane_mask_lt() {
  unsigned lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// (element, corner) pairs of one element that may lower their target: target owned and
// updatable, smallest other corner known, and the target above that smallest time plus the
// corner's height (a lower bound of any solve through the element).
template <int DIM>
__device__ __forceinline__ unsigned select_corners(const Elem& x, float t0, float t1, float t2,
                                                   float t3, int nown, float large) {
  unsigned sel = 0;
  if constexpr (DIM == 3) {
    const float m01 = __builtin_fminf(t0, t1), m23 = __builtin_fminf(t2, t3);
    const float mn0 = __builtin_fminf(t1, m23), mn1 = __builtin_fminf(t0, m23);
    const float mn2 = __builtin_fminf(m01, t3), mn3 = __builtin_fminf(m01, t2);
    sel |= (x.v0 < nown && mn0 < large && t0 > mn0 + x.H0) ? 1u : 0u;
    sel |= (x.v1 < nown && mn1 < large && t1 > mn1 + x.H1) ? 2u : 0u;
    sel |= (x.v2 < nown && mn2 < large && t2 > mn2 + x.H2) ? 4u : 0u;
    sel |= (x.v3 < nown && mn3 < large && t3 > mn3 + x.H3) ? 8u : 0u;
  } else {
    const float mn0 = __builtin_fminf(t1, t2), mn1 = __builtin_fminf(t0, t2);
    const float mn2 = __builtin_fminf(t0, t1);
    sel |= (x.v0 < nown && mn0 < large && t0 > mn0 + x.H0) ? 1u : 0u;
    sel |= (x.v1 < nown && mn1 < large && t1 > mn1 + x.H1) ? 2u : 0u;
    sel |= (x.v2 < nown && mn2 < large && t2 > mn2 + x.H2) ? 4u : 0u;
    sel &= x.mask;
  }
  return sel;
}

// LDS layout of one visit (dynamic shared memory):
//   sRec[max_elem]  element records (only when rec_lds), 16-byte aligned
//   sT[nloc]        times of owned + halo vertices
//   sOld[nown]      owned times at the start of the visit
//   sChg[nloc]      sweep stamp of the last decrease of each local vertex (int); -1 = the
//                   vertex has not changed since the block's previous visit
//   sHidx[nhalo]    global ids of the halo vertices (a re-run reloads their times from here)
//   sVm[nown]       (triangles only) per owned vertex, the mask of dependents whose halo holds
//                   it (when the layout has masks), 8-byte aligned
//   sQ[queue_cap]   compacted (element << 2 | corner) work items (16-bit: a block holds at most
//                   16383 elements)
//   sMark[max_dep]  dependents whose halo holds a boundary vertex this visit lowered (int)
// plus static: per-sweep flags (4 ints) and the queue counter.
//
// The `seen` array (one value per block-local vertex, block after block) holds, per vertex, the
// time every element around it has been evaluated with; a vertex whose time is still the same
// cannot have changed any solve of the block, so the first sweep of a re-visit only evaluates
// elements around vertices that changed since.  At the end of a visit that is every vertex but
// the owned ones lowered in its last sweep (when that sweep was cut off by the sweep cap, the
// elements around them are still to be evaluated).  Visits of one block never overlap (the
// persistent kernels claim a block before visiting it, and the one-launch-per-iteration kernel
// lists each block once per launch), so only the visiting workgroup writes a block's entries.
//
// keep_recs: the workgroup's previous visit was of the same block, so its staged records are
// still in LDS and are not copied again.  inner > 0 caps this visit at `inner` sweeps instead
// of a.max_inner.  preloaded: the times, stamps and flags are already in LDS (rerun_load ran
// after the previous visit of the same block), so the visit starts with its sweeps.
template <int DIM, bool COHERENT, int BT = kBlockThreads, bool SPLIT = false>
__device__ VisitResult visit_block(const VisitArgs& a, int b, void* sRec, float* sT, float* sOld,
                                   int* sChg, int* sHidx, uint64_t* sVm, uint16_t* sQ, int* sMark,
                                   int* sFlag, int* sQn,
                                   uint64_t* nsolves, bool keep_recs = false, int inner = 0,
                                   bool preloaded = false) {
  const BlockDesc d = a.blocks[b];
  constexpr int RW = SPLIT ? 4 : 4 * kRecWords<DIM>;  // record size in 4-byte words
  const uint32_t* grec = reinterpret_cast<const uint32_t*>(a.elems) + (int64_t)RW * d.estart;
  const void* recs = !SPLIT && a.rec_lds ? (const void*)sRec : (const void*)grec;
  const float* glen = SPLIT ? a.lens + (DIM == 3 ? 6 : 4) * (int64_t)d.estart : nullptr;
  constexpr int K = DIM == 3 ? 4 : 3;
  const int tid = threadIdx.x;
  const int nown = d.nown;
  const float large = a.large;
  float* seen = a.seen + d.vstart + d.hstart;  // prefix sums of nown and nhalo
  const unsigned long long c0 = wall_clock64();
  const int nhalo = d.nhalo;
  if (!preloaded) {
    // this thread's first owned time and first halo index go out before the record copy and the
    // halo time right behind the first batch of record loads, so the three memory round trips of
    // the staging (descriptor, then records / owned times / halo indices, then halo times)
    // overlap instead of following each other
    float own0 = 0.f, sown0 = 0.f, shalo0 = 0.f, hv0 = 0.f;
    int32_t hidx0 = 0;
    uint64_t vm0 = 0;
    if (tid < nown) {
      own0 = load_T<COHERENT>(a.T + d.vstart + tid);
      sown0 = seen[tid];
      if (DIM == 2 && a.vmask) vm0 = a.vmask[d.vstart + tid];
    }
    if (tid < nhalo) {
      hidx0 = a.halo[d.hstart + tid];
      shalo0 = seen[nown + tid];
    }
    bool hv_issued = false;

    if (!SPLIT && a.rec_lds && !keep_recs) {
      // copy the records in batches of 8 loads per thread, all in flight before the stores, so
      // the copy costs one memory round trip per 8*256 16-byte words instead of one per word
      constexpr int RW4 = kRecWords<DIM>;
      const int nw = RW4 * d.nelem;  // 16-byte words
      const uint4* src = reinterpret_cast<const uint4*>(grec);
      uint4* dst = reinterpret_cast<uint4*>(sRec);
      for (int i0 = 0; i0 < nw; i0 += 8 * BT) {
        uint4 r[8];
  #pragma unroll
        for (int j = 0; j < 8; ++j) {
          int i = i0 + j * BT + tid;
          r[j] = i < nw ? src[i] : make_uint4(0, 0, 0, 0);
        }
        if (!hv_issued) {
          if (tid < nhalo) hv0 = load_T<COHERENT>(a.T + hidx0);
          hv_issued = true;
        }
        if (a.rmult) {
          // a factor f on the squared lengths scales the heights by sqrt(f)
  #pragma unroll
          for (int j = 0; j < 8; ++j) {
            int i = i0 + j * BT + tid;
            if (i < nw) {
              const int e = i / RW4, w = i - RW4 * e;
              const float f = a.rmult[d.estart + e];
              const float sf = __builtin_sqrtf(f);
              float fx, fy, fz, fw;  // factor per component of this word
              if (w == 0) {
                fx = fy = 1.f;
                fz = fw = f;
              } else if (DIM == 3 && w == 1) {
                fx = fy = fz = fw = f;
              } else if (DIM == 3) {
                fx = fy = fz = fw = sf;
              } else {
                fx = f;
                fy = fz = fw = sf;
              }
              r[j].x = fx == 1.f ? r[j].x : __float_as_uint(__uint_as_float(r[j].x) * fx);
              r[j].y = fy == 1.f ? r[j].y : __float_as_uint(__uint_as_float(r[j].y) * fy);
              r[j].z = __float_as_uint(__uint_as_float(r[j].z) * fz);
              r[j].w = __float_as_uint(__uint_as_float(r[j].w) * fw);
            }
          }
        }
  #pragma unroll
        for (int j = 0; j < 8; ++j) {
          int i = i0 + j * BT + tid;
          if (i < nw) dst[i] = r[j];
        }
      }
    }
    if (!hv_issued && tid < nhalo) hv0 = load_T<COHERENT>(a.T + hidx0);
    if (tid < nown) {
      sT[tid] = own0;
      sOld[tid] = own0;
      sChg[tid] = own0 < sown0 ? 0 : -1;
      if constexpr (DIM == 2) sVm[tid] = vm0;
    }
    if (tid < nhalo) {
      sT[nown + tid] = hv0;
      sChg[nown + tid] = hv0 < shalo0 ? 0 : -1;
      sHidx[tid] = hidx0;
    }
    for (int i = tid + BT; i < nown; i += BT) {
      float v = load_T<COHERENT>(a.T + d.vstart + i);
      sT[i] = v;
      sOld[i] = v;
      sChg[i] = v < seen[i] ? 0 : -1;
      if constexpr (DIM == 2) sVm[i] = a.vmask ? a.vmask[d.vstart + i] : 0ull;
    }
    for (int i = tid + BT; i < nhalo; i += BT) {
      const int32_t h = a.halo[d.hstart + i];
      float v = load_T<COHERENT>(a.T + h);
      sT[nown + i] = v;
      sChg[nown + i] = v < seen[nown + i] ? 0 : -1;
      sHidx[i] = h;
    }
    if (tid < 4) sFlag[tid] = 0;
    if (tid < 2) sQn[tid] = 0;
    if (a.vmask) {
      if (tid < 2) sMark[tid] = 0;
    } else if (a.vh_start) {
      const int ndep = a.dep_start[b + 1] - a.dep_start[b];
      for (int i = tid; i < ndep; i += BT) sMark[i] = 0;
    }
    __syncthreads();
  }
  const unsigned long long c1 = wall_clock64();

  // records read from memory with factors: the factors are applied at every use, as the
  // staging copy applies them (the same products, so both paths give the same values)
  const float* rm = (!SPLIT && a.rec_lds) || !a.rmult ? nullptr : a.rmult + d.estart;
  uint64_t solves = 0;
  bool converged = false;
  int sweep = 0;
  const int chunk = a.queue_cap / K;
  const int max_inner = inner > 0 ? inner : a.max_inner;
  // Triangle visits of the persistent kernels (several sweeps each) store every lowered time
  // as soon as it drops, so dependents already running see it: level-7 anisotropic icosphere
  // 1.80 -> 1.52 ms.  Tet visits publish at the write-back: with early stores the 64^3 cube
  // ran 6% more visits for no gain beyond noise and the 256-cube batch was 3% slower.
  constexpr bool kEarly = COHERENT && DIM == 2;
  // chunks alternate between the two work-item counters: the one a chunk fills was cleared by
  // thread 0 during the previous chunk, after every thread had read it for the chunk before,
  // so no barrier is needed to clear it (both are 0 when a visit starts)
  int qc = 0;
  for (; sweep < max_inner; ++sweep) {
    const int stamp = sweep + 1;  // decreases made in this sweep carry this stamp
    for (int e0 = 0; e0 < d.nelem; e0 += chunk, qc ^= 1) {
      const int e1 = min(d.nelem, e0 + chunk);
      // phase 1: select the (element, corner) pairs that can improve.  A thread takes kSelBatch
      // elements at a time and issues all their LDS reads (records, then corner times and change
      // stamps) before using any, so the batch costs one chain of LDS round trips instead of one
      // per element; the batch's items are then placed with one LDS atomic per wave.
      for (int eb = e0; eb < e1; eb += kSelBatch * BT) {
        unsigned sel[kSelBatch];
        Elem x[kSelBatch];
#pragma unroll
        for (int u = 0; u < kSelBatch; ++u) {
          const int e = eb + u * BT + tid;
          if constexpr (SPLIT)
            x[u] = load_sel<DIM>(recs, e < e1 ? e : e0);
          else
            x[u] = load_elem<DIM, kRecSelect>(recs, e < e1 ? e : e0);
        }
#pragma unroll
        for (int u = 0; u < kSelBatch; ++u) {
          const int e = eb + u * BT + tid;
          sel[u] = 0;
          if (e < e1) {
            if (rm) {
              const float sf = __builtin_sqrtf(rm[e]);
              x[u].H0 *= sf;
              x[u].H1 *= sf;
              x[u].H2 *= sf;
              x[u].H3 *= sf;
            }
            const float t0 = sT[x[u].v0], t1 = sT[x[u].v1], t2 = sT[x[u].v2];
            const float t3 = DIM == 3 ? sT[x[u].v3] : 0.f;
            int fresh = max(max(sChg[x[u].v0], sChg[x[u].v1]), sChg[x[u].v2]);
            if constexpr (DIM == 3) fresh = max(fresh, sChg[x[u].v3]);
            // an element whose corners did not change since its last evaluation cannot improve
            if (fresh >= sweep) sel[u] = select_corners<DIM>(x[u], t0, t1, t2, t3, nown, large);
          }
        }
        // compaction: one ballot per (element of the batch, corner); the wave's items are laid
        // out ballot by ballot, each lane's position within a ballot given by the count of lower
        // lanes (mbcnt)
        uint64_t mk[kSelBatch][4];
        int total = 0;
#pragma unroll
        for (int u = 0; u < kSelBatch; ++u)
#pragma unroll
          for (int c = 0; c < K; ++c) {
            mk[u][c] = __ballot((sel[u] >> c) & 1u);
            total += __popcll(mk[u][c]);
          }
        if (total > 0) {
          int base = 0;
          if (__lane_id() == 0) base = atomicAdd(&sQn[qc], total);
          base = __builtin_amdgcn_readfirstlane(base);
#pragma unroll
          for (int u = 0; u < kSelBatch; ++u) {
            const uint16_t item = (uint16_t)((eb + u * BT + tid) << 2);
#pragma unroll
            for (int c = 0; c < K; ++c) {
              if ((sel[u] >> c) & 1u) sQ[base + mbcnt64(mk[u][c])] = item | (uint16_t)c;
              base += __popcll(mk[u][c]);
            }
          }
        }
      }
      __syncthreads();
      const int nq = sQn[qc];
      if (tid == 0) sQn[qc ^ 1] = 0;
      // phase 2: local solves over the compacted queue
      for (int q = tid; q < nq; q += BT) {
        const int item = (int)sQ[q];
        const int e = item >> 2, k = item & 3;
        Elem x;
        if constexpr (SPLIT) {
          x = load_sel<DIM>(recs, e);
          load_len<DIM>(glen, e, x);
        } else {
          x = load_elem<DIM, kRecSolve>(recs, e);
        }
        if (rm) {
          const float f = rm[e];
          x.L0 *= f;
          x.L1 *= f;
          x.L2 *= f;
          x.L3 *= f;
          x.L4 *= f;
          x.L5 *= f;
        }
        const float t0 = sT[x.v0], t1 = sT[x.v1], t2 = sT[x.v2];
        float cand, tk;
        int vk;
        if constexpr (DIM == 3) {
          const float t3 = sT[x.v3];
          cand = tet_rotation(k, t0, t1, t2, t3, x.L0, x.L1, x.L2, x.L3, x.L4, x.L5, large);
          tk = k == 0 ? t0 : k == 1 ? t1 : k == 2 ? t2 : t3;
          vk = k == 0 ? x.v0 : k == 1 ? x.v1 : k == 2 ? x.v2 : x.v3;
        } else {
          cand = tri_rotation(k, t0, t1, t2, x.L0, x.L1, x.L2, large);
          tk = k == 0 ? t0 : k == 1 ? t1 : t2;
          vk = k == 0 ? x.v0 : k == 1 ? x.v1 : x.v2;
        }
        // arrival times are non-negative; the clamp keeps the unsigned-bits minimum ordered
        cand = __builtin_fmaxf(cand, 0.f);
        ++solves;
        if (cand < tk) {
          unsigned prev = atomicMin(reinterpret_cast<unsigned*>(&sT[vk]), __float_as_uint(cand));
          float pv = __uint_as_float(prev);
          // triangle blocks in persistent kernels publish the new time at once (no result
          // awaited); the write-back then only waits for these atomics to complete
          if constexpr (kEarly) {
            if (cand < pv)
              (void)__hip_atomic_fetch_min(reinterpret_cast<unsigned*>(a.T + d.vstart + vk),
                                           __float_as_uint(cand), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
          }
          // a decrease larger than eps (relative) marks the vertex changed for the next sweep;
          // smaller ones are kept but do not trigger re-evaluation
          if (cand < pv && pv - cand > a.eps * cand) {
            sChg[vk] = stamp;
            sFlag[sweep % 3] = 1;
          }
        }
      }
      __syncthreads();
    }
    // block-wide OR of the changes through a rotating set of three flags: flag s%3 is set by any
    // thread that lowers a time in sweep s, read after the barrier that ends the sweep's last
    // chunk, and cleared by thread 0 in sweep s+1 (as (s+1+2)%3)
    const int any = sFlag[sweep % 3];
    if (tid == 0) sFlag[(sweep + 2) % 3] = 0;
    if (!any) {
      converged = true;
      ++sweep;
      break;
    }
  }
  // the barrier that ended the last chunk also ends the sweeps (times and stamps are final)
  if (tid == 0) {  // a following re-run starts without a load phase
    sQn[0] = 0;
    sQn[1] = 0;
  }
  const unsigned long long c2 = wall_clock64();
  // write back the owned times that decreased, and record in `seen` what the sweeps accounted
  // for: everything but the owned vertices lowered in the last sweep of a visit cut off by the
  // sweep cap (stamp == sweep), whose old `seen` value makes them count as changed next time
  bool bchg = false;
  for (int i = tid; i < nown; i += BT) {
    float v = sT[i];
    if (converged || sChg[i] < sweep) seen[i] = v;
    if (v < sOld[i]) {
      uint64_t mk = 0;
      if constexpr (DIM == 2) mk = sVm[i];
      else if (a.vmask) mk = a.vmask[d.vstart + i];
      // published during the sweeps when kEarly (only the visiting workgroup writes a block's
      // owned times, so every drop here is a drop of the stored time)
      bool dec = true;
      if constexpr (!kEarly) dec = store_min_T<COHERENT>(a.T + d.vstart + i, v, sOld[i]);
      if (dec && i >= d.nbnd0) {
        bchg = true;
        if (a.vmask) {
          if ((uint32_t)mk) atomicOr(reinterpret_cast<unsigned*>(&sMark[0]), (uint32_t)mk);
          if ((uint32_t)(mk >> 32))
            atomicOr(reinterpret_cast<unsigned*>(&sMark[1]), (uint32_t)(mk >> 32));
        } else if (a.vh_start) {
          const int vg = d.vstart + i;
          for (int p = a.vh_start[vg]; p < a.vh_start[vg + 1]; ++p) sMark[a.vh_pos[p]] = 1;
        }
      }
    }
  }
  // halo times do not change during a visit, and the first sweep evaluated every element
  // around a changed one
  if (sweep > 0)
    for (int i = tid; i < d.nhalo; i += BT) seen[nown + i] = sT[nown + i];
  // With COHERENT, the time write-backs are device-scope atomics, coherent across XCDs without
  // cache maintenance: the tet path consumes each result above, and the kEarly path waits for
  // this thread's outstanding memory operations (vmcnt covers stores and atomics on this
  // target), so either way they have completed at the device coherence point before the barrier
  // below, and the caller queues dependents only after it.  (A release fence would also write
  // back the L2: several times slower.)
  if constexpr (kEarly) {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
  }
  *nsolves += solves;
  // sFlag[3] was cleared before the sweeps (load phase or rerun_load); callers put a barrier
  // between this visit's read of it and the next visit's clearing
  if (bchg) sFlag[3] = 1;
  __syncthreads();
  VisitResult res;
  res.converged = converged;
  res.bnd_changed = sFlag[3] != 0;
  res.sweeps = sweep;
  res.t_load = c1 - c0;
  res.t_sweep = c2 - c1;
  res.t_wb = wall_clock64() - c2;
  return res;
}

// LDS state for a re-run of the block the workgroup has just visited, built
// by threads [t0, BT) while the others queue the visit's dependents: the owned times are already
// in LDS (only the visiting workgroup writes them); an owned vertex counts as changed if the last
// sweep (stamp last_stamp) lowered it, a halo vertex if its time dropped since it was loaded.
// This is what a fresh load computes from `seen`, without reading it.  The caller clears the
// dependent marks after they have been read, and a barrier follows before the sweeps.
template <bool COHERENT, int BT>
__device__ __forceinline__ void rerun_load(const VisitArgs& a, const BlockDesc& d, int last_stamp,
                                           float* sT, float* sOld, int* sChg, const int* sHidx,
                                           int* sFlag, int t0) {
  const int n = BT - t0, t = (int)threadIdx.x - t0;
  if (t < 4) sFlag[t] = 0;
  for (int i = t; i < d.nhalo; i += n) {
    const float v = load_T<COHERENT>(a.T + sHidx[i]);
    sChg[d.nown + i] = v < sT[d.nown + i] ? 0 : -1;
    sT[d.nown + i] = v;
  }
  for (int i = t; i < d.nown; i += n) {
    sOld[i] = sT[i];
    sChg[i] = sChg[i] == last_stamp ? 0 : -1;
  }
}

}  // namespace eik
