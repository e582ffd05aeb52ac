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
// target, and, after the first sweep, some corner of the element changed in the previous sweep)
// and appends the surviving (element, corner) pairs to an LDS queue with a ballot/prefix; the
// local solves then run over the queue with every lane busy.
#pragma once

#include <hip/hip_runtime.h>

#include "eik_internal.h"
#include "local_solver.h"

namespace eik {

constexpr int kBlockThreads = 256;
constexpr int kWaves = kBlockThreads / 64;

struct VisitArgs {
  const BlockDesc* blocks;
  const void* elems;
  const int32_t* halo;
  float* T;
  float large;
  float eps;
  int32_t max_inner;
  int32_t queue_cap;  // (element, corner) entries the LDS queue holds per chunk
};

struct VisitResult {
  bool converged;
  bool bnd_changed;
  int sweeps;
};

template <int DIM>
struct ElemTraits;

template <>
struct ElemTraits<3> {
  using Rec = TetRec;
  static constexpr int K = 4;
  static constexpr int NL = 6;
  __device__ static float solve(int c, const float* t, const float* L, float large) {
    return tet_rotation(c, t, L, large);
  }
};

template <>
struct ElemTraits<2> {
  using Rec = TriRec;
  static constexpr int K = 3;
  static constexpr int NL = 3;
  __device__ static float solve(int c, const float* t, const float* L, float large) {
    return tri_rotation(c, t, L, large);
  }
};

__device__ __forceinline__ uint64_t lane_mask_lt() {
  unsigned lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// LDS layout of one visit (dynamic shared memory, 4-byte words):
//   sT[max_local]        times of owned + halo vertices
//   sOld[max_own]        owned times at the start of the visit
//   sChg[max_local]      sweep stamp of the last decrease of each local vertex (int)
//   sQ[queue_cap]        compacted (element << 2 | corner) work items (int)
// plus static: per-sweep flags, queue counter.
template <int DIM>
__device__ VisitResult visit_block(const VisitArgs& a, int b, float* sT, float* sOld, int* sChg,
                                   int* sQ, int* sFlag, int* sQn, uint64_t* nsolves) {
  using Tr = ElemTraits<DIM>;
  using Rec = typename Tr::Rec;
  const BlockDesc d = a.blocks[b];
  const Rec* recs = reinterpret_cast<const Rec*>(a.elems) + d.estart;
  const int tid = threadIdx.x;
  const int nloc = d.nown + d.nhalo;

  for (int i = tid; i < d.nown; i += kBlockThreads) {
    float v = a.T[d.vstart + i];
    sT[i] = v;
    sOld[i] = v;
  }
  for (int i = tid; i < d.nhalo; i += kBlockThreads) sT[d.nown + i] = a.T[a.halo[d.hstart + i]];
  for (int i = tid; i < nloc; i += kBlockThreads) sChg[i] = 0;  // "changed before sweep 1"
  if (tid < 3) sFlag[tid] = 0;
  __syncthreads();

  uint64_t solves = 0;
  bool converged = false;
  int sweep = 0;
  for (; sweep < a.max_inner; ++sweep) {
    const int stamp = sweep + 1;  // decreases made in this sweep carry this stamp
    bool changed = false;
    // process the elements in chunks that fit the queue
    for (int e0 = 0; e0 < d.nelem; e0 += a.queue_cap / Tr::K) {
      const int e1 = min(d.nelem, e0 + a.queue_cap / Tr::K);
      if (tid == 0) *sQn = 0;
      __syncthreads();
      // phase 1: select (element, corner) pairs that can improve
      for (int eb = e0; eb < e1; eb += kBlockThreads) {
        const int e = eb + tid;
        unsigned sel = 0;
        if (e < e1) {
          const Rec r = recs[e];
          float t[Tr::K];
          bool fresh = false;
#pragma unroll
          for (int k = 0; k < Tr::K; ++k) {
            t[k] = sT[r.v[k]];
            fresh |= sChg[r.v[k]] >= sweep;  // changed in the previous sweep (or initial)
          }
          if (fresh) {
#pragma unroll
            for (int k = 0; k < Tr::K; ++k) {
              bool upd = r.v[k] < d.nown;
              if constexpr (DIM == 2) upd = upd && ((r.mask >> k) & 1);
              float mn = EIK_INF;
#pragma unroll
              for (int j = 0; j < Tr::K; ++j)
                if (j != k) mn = __builtin_fminf(mn, t[j]);
              if (upd && mn < a.large && t[k] > mn) sel |= 1u << k;
            }
          }
        }
        const int cnt = __builtin_popcount(sel);
        // wave prefix of counts
        int incl = cnt;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          int y = __shfl_up(incl, off, 64);
          if ((int)__lane_id() >= off) incl += y;
        }
        const int total = __shfl(incl, 63, 64);
        int base = 0;
        if (__lane_id() == 63 && total > 0) base = atomicAdd(sQn, total);
        base = __shfl(base, 63, 64);
        int pos = base + incl - cnt;
        while (sel) {
          int k = __builtin_ctz(sel);
          sel &= sel - 1;
          sQ[pos++] = (e << 2) | k;
        }
      }
      __syncthreads();
      const int nq = *sQn;
      // phase 2: local solves over the compacted queue
      for (int q = tid; q < nq; q += kBlockThreads) {
        const int item = sQ[q];
        const int e = item >> 2, k = item & 3;
        const Rec r = recs[e];
        float t[Tr::K];
#pragma unroll
        for (int j = 0; j < Tr::K; ++j) t[j] = sT[r.v[j]];
        // arrival times are non-negative; the clamp keeps the unsigned-bits minimum ordered
        float cand = __builtin_fmaxf(Tr::solve(k, t, r.L, a.large), 0.f);
        ++solves;
        if (cand < t[k]) {
          unsigned prev = atomicMin(reinterpret_cast<unsigned*>(&sT[r.v[k]]), __float_as_uint(cand));
          float pv = __uint_as_float(prev);
          if (cand < pv) {
            sChg[r.v[k]] = stamp;
            if (pv - cand > a.eps) changed = true;
          }
        }
      }
      __syncthreads();
    }
    // block-wide OR of `changed` through a rotating set of three flags
    if (changed) sFlag[sweep % 3] = 1;
    __syncthreads();
    const int any = sFlag[sweep % 3];
    if (tid == 0) sFlag[(sweep + 2) % 3] = 0;
    if (!any) {
      converged = true;
      ++sweep;
      break;
    }
  }
  __syncthreads();
  // write back the owned times that decreased
  bool bchg = false;
  for (int i = tid; i < d.nown; i += kBlockThreads) {
    float v = sT[i];
    if (v < sOld[i]) {
      a.T[d.vstart + i] = v;
      if (i >= d.nbnd0) bchg = true;
    }
  }
  *nsolves += solves;
  if (tid == 0) sFlag[3] = 0;
  __syncthreads();
  if (bchg) sFlag[3] = 1;
  __syncthreads();
  VisitResult res;
  res.converged = converged;
  res.bnd_changed = sFlag[3] != 0;
  res.sweeps = sweep;
  return res;
}

}  // namespace eik
