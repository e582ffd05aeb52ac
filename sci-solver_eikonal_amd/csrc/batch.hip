// batch.hip -- kernels of multi-solve launches (see multi.h), their host helpers, and batches:
// several meshes (or one resident mesh several times, each with its own speed field) solved by
// one launch sequence, captured into a HIP graph and replayed.  No counterpart in the reference,
// which runs one solve per process with a host loop per outer iteration
// (meshFIM3dEikonal.cu:411-542).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <thread>

#include "multi.h"

namespace eik {

// Per solve (blockIdx.y): T = LARGENUM, seen = +inf.
__global__ void k_multi_init(const SolveDesc* solves) {
  const SolveDesc& s = solves[blockIdx.y];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < s.nseen; i += stride) {
    if (i < s.nv) s.T[i] = s.large;
    s.seen[i] = EIK_INF;
  }
}

// Seeds: (solve, caller vertex, value) triples; the owner block of each seed and its dependents
// are queued, one workgroup per seed and one lane per block to queue (lane 0 the owner), so the
// pushes' device-scope round trips overlap.  *n is read from device memory so a captured graph
// serves any seed count.  The arguments of a kind without solves in the launch are all zero, so
// the descriptors come from `solves`, and a seed's own kind (present by construction) gives the
// queue.
__global__ void k_multi_seed(const SolveDesc* solves, MultiArgs a3, MultiArgs a2,
                             const int32_t* seed_solve, const int32_t* seed_vert,
                             const float* seed_val, const int32_t* n) {
  for (int64_t i = blockIdx.x; i < *n; i += gridDim.x) {
    const int32_t si = seed_solve[i];
    const SolveDesc& s = solves[si];
    const MultiArgs& a = s.dim == 3 ? a3 : a2;
    const int32_t v = s.iperm[seed_vert[i]];
    if (threadIdx.x == 0) atomicMin(reinterpret_cast<unsigned*>(&s.T[v]), __float_as_uint(seed_val[i]));
    const int32_t b = s.owner[v];
    const int32_t tag = si << a.bbits;
    const int32_t s0 = s.dep_start[b], nd = s.dep_start[b + 1] - s0;
    for (int j = threadIdx.x; j <= nd; j += blockDim.x)
      q_push(a, tag | (s.bbase + (j == 0 ? b : s.dep[s0 + j - 1])));
  }
}

// Per solve (blockIdx.y): out[out_off + caller id] = T[new id].
__global__ void k_multi_gather(const SolveDesc* solves, float* out) {
  const SolveDesc& s = solves[blockIdx.y];
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < s.nv; i += stride)
    out[s.out_off + i] = s.T[s.iperm[i]];
}

// rmult[q] = mult[rec_elem[q]]: per-element factors in record order.
__global__ void k_record_factors(const float* mult, const int32_t* rec_elem, int64_t n,
                                 float* rmult) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride)
    rmult[q] = mult[rec_elem[q]];
}

// *bad = 1 if a factor is not finite and > 0 (the caller clears *bad first).
__global__ void k_check_factors(const float* f, int64_t n, int32_t* bad) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int b = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    b |= !(f[i] > 0.f && f[i] < 1e30f);
  if (b) bad[0] = 1;
}

__global__ void k_fill_u64(unsigned long long* p, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

// One launch per element kind: a launch's solves are all tets or all triangles (the two kinds
// keep separate queues), so the kernel is compiled per kind and keeps the register budget of
// the single-solve kernel.  Two shapes: BT = kBlockThreads with the element records staged in
// LDS (the per-solve factors are applied while staging), and BT = kDirectThreads reading the
// records from memory at every use, whose small LDS footprint lets twice as many workgroups
// per CU hide each other's latencies (launches without factors: many independent solves keep
// them all busy).
constexpr int kDirectThreads = 128;

// WPE: waves per SIMD the register allocation is sized for (4: at most 128 VGPRs, 3: at most
// 168).  Both are compiled; a launch takes 3 when its LDS already limits the CU to that many
// waves (launch_args), so the larger budget costs no resident workgroup and saves the spills.
template <int DIM, int BT, int WPE>
__global__ __launch_bounds__(BT) __attribute__((amdgpu_waves_per_eu(WPE))) void k_visit_multi(
    MultiArgs a, const SolveDesc* __restrict__ solves) {
  // the shape reading records from memory uses the split records (selection part and lengths)
  constexpr bool kSplit = BT == kDirectThreads;
  extern __shared__ uint4 lds4[];
  __shared__ int sFlag[4];
  __shared__ int sQn[2];  // work-item counters of alternate chunks
  __shared__ int sEntry;
  __shared__ unsigned long long sRed;
  __shared__ unsigned long long sAcc[kNumCounters];
  __shared__ int sAgain;
  __shared__ int sNext;  // entry taken while the previous visit queued its dependents, or -2
  if (threadIdx.x < kNumCounters) sAcc[threadIdx.x] = 0;
  if (threadIdx.x == 0) sNext = -2;
  // No seed of this kind was queued (the seed kernel ran before this launch): nothing will ever
  // bring `pending` to zero, so the launch ends at once.  Every workgroup checks on entry rather
  // than one chosen workgroup: with several launches in flight, a given workgroup of this kernel
  // may not become resident until other kernels' workgroups leave, and those may be waiting on
  // their own launch in the same way.  pending is zero only before any seed or after the end,
  // and a repeated q_finish only marks slots no workgroup waits on (it is skipped once the end
  // is marked).
  if (threadIdx.x == 0 &&
      (__hip_atomic_load(&a.ctl->tp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 0xffffffffull) == 0ull &&
      (__hip_atomic_load(&a.ctl->head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & kDoneBit) == 0u)
    q_finish(a);
  for (;;) {
    if (threadIdx.x == 0) {
      int ne = sNext;
      if (ne == -2) {
        const unsigned long long w0 = wall_clock64();
        ne = q_take(a);
        sAcc[6] += wall_clock64() - w0;
        if (ne == kAbandoned) sAcc[kCntAbandoned] += 1;
      }
      sEntry = ne;
      sNext = -2;
    }
    __syncthreads();
    // uniform across the workgroup: kept in scalar registers, the descriptor read with scalar
    // loads
    const int entry = __builtin_amdgcn_readfirstlane(sEntry);
    if (entry < 0) break;
    const int si = entry >> a.bbits;
    const SolveDesc& sD = solves[si];
    const int g = entry & a.bmask;
    const int b = g - sD.bbase;
    VisitArgs v;
    v.blocks = sD.blocks;
    v.elems = kSplit ? sD.esel : sD.elems;
    v.halo = sD.halo;
    v.T = sD.T;
    v.seen = sD.seen;
    v.large = sD.large;
    v.eps = a.eps;
    v.max_inner = a.max_inner;
    v.queue_cap = a.queue_cap;
    v.rec_lds = BT == kBlockThreads;
    v.vh_start = nullptr;
    v.vh_pos = nullptr;
    v.dep_start = sD.dep_start;
    v.vmask = sD.vmask;
    v.rmult = sD.rmult;
    v.lens = sD.elen;
    bool rerun = false;  // the previous visit of this workgroup was of this block
    do {
      if (threadIdx.x == 0) sRed = 0;
      const unsigned long long t0 = wall_clock64();
      const BlockDesc d = v.blocks[b];
      VisitLds l = carve_lds<DIM>(lds4, a.rec_lds_bytes, a.queue_cap, d);
      uint64_t solves = 0;
      VisitResult r = visit_block<DIM, true, BT, kSplit>(v, b, l.rec, l.T, l.old, l.chg, l.hidx, l.vm, l.q,
                                                 l.mark, sFlag, sQn, &solves, rerun);
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) solves += __shfl_down(solves, off, 64);
      // sRed is read by thread 0 after the next barrier
      if (__lane_id() == 0) atomicAdd(&sRed, (unsigned long long)solves);
      const unsigned long long tq = wall_clock64();
      const bool cap_ok = (int64_t)sAcc[0] < a.max_visits;
      // the first 64 dependents are flagged while thread 0 releases the block (see
      // k_visit_async)
      const int32_t s0 = v.dep_start[b];
      const int ndep = r.bnd_changed && cap_ok ? v.dep_start[b + 1] - s0 : 0;
      const uint64_t mask = v.vmask ? ((uint64_t)(uint32_t)l.mark[0] |
                                       ((uint64_t)(uint32_t)l.mark[1] << 32))
                                    : ~0ull;
      bool want0 = false;
      int32_t c0 = g;
      if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        if (lane < ndep) {
          want0 = !v.vmask || ((mask >> lane) & 1ull) != 0ull;
          c0 = sD.bbase + sD.dep[s0 + lane];
          if (want0) want0 = atomicExch(&a.qflag[c0], 1u) == 0u;
        }
        if (lane == 0) sAgain = release_or_rerun(a, g) && cap_ok;
      }
      __syncthreads();
      const int again = sAgain;
      if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        const int32_t tag = si << a.bbits;
        const int nitems = ndep + 1;
        for (int j0 = 0; j0 < nitems; j0 += 64) {
          const int j = j0 + lane;
          bool want = false;
          int32_t c = g;
          if (j0 == 0 && j < ndep) {
            want = want0;
            c = c0;
          } else {
            if (j < ndep) {
              want = !v.vmask || ((mask >> (j & 63)) & 1ull) != 0ull;
              c = sD.bbase + sD.dep[s0 + j];
            } else if (j == ndep) {
              want = !r.converged && !again && cap_ok;
            }
            if (want) want = atomicExch(&a.qflag[c], 1u) == 0u;
          }
          const uint64_t bal = __ballot(want);
          const unsigned n = (unsigned)__popcll(bal);
          const unsigned dec = (j0 + 64 >= nitems && !again) ? 1u : 0u;
          if (n == 0u && dec == 0u) continue;
          uint32_t base = 0;
          if (lane == 0) {
            const unsigned long long inc = ((unsigned long long)n << 32) + n - dec;
            const unsigned long long old = __hip_atomic_fetch_add(
                &a.ctl->tp, inc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            base = (uint32_t)(old >> 32);
            if (n == 0u && (old & 0xffffffffull) == 1ull) q_finish(a);
          }
          base = __shfl(base, 0, 64);
          if (want)
            __hip_atomic_store(q_slot(a, base + (uint32_t)__popcll(bal & lane_mask_lt())),
                               (tag | c) + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) {
          tally_visit(sAcc, d, r, sRed, kAlgoElemBytes<DIM>, tq - t0, rerun, rerun && v.rec_lds);
          sAcc[10] += wall_clock64() - tq;
        }
      } else if (threadIdx.x == 64 && !again) {
        // next entry taken while wave 0 queues the dependents (no re-run: nothing held)
        const unsigned long long w0 = wall_clock64();
        const int32_t nx = q_take(a);
        sNext = nx;
        sAcc[6] += wall_clock64() - w0;
        if (nx == kAbandoned) sAcc[kCntAbandoned] += 1;
      }
      __syncthreads();
      rerun = true;
    } while (sAgain);
  }
  if (threadIdx.x < kNumCounters && sAcc[threadIdx.x])
    atomicAdd(&a.counters[threadIdx.x], sAcc[threadIdx.x]);
}

// ------------------------------------------------------------------------------------------
// Host helpers shared with the serving loop
// ------------------------------------------------------------------------------------------
template <class F>
static int parallel_over(size_t n, F&& fn) {
  // each layout build runs its own host threads too (at most 16): 8 x 16 threads at most
  const unsigned nt = std::max(1u, std::min<unsigned>(std::thread::hardware_concurrency(), 8u));
  std::vector<int> rcs(n, EIK_OK);
  std::vector<std::string> errs(n);
  std::atomic<size_t> next{0};
  auto work = [&]() {
    for (;;) {
      size_t i = next.fetch_add(1);
      if (i >= n) return;
      rcs[i] = fn(i);
      if (rcs[i]) errs[i] = eik_last_error();
    }
  };
  std::vector<std::thread> th;
  for (unsigned t = 0; t < std::min<unsigned>(nt, (unsigned)n); ++t) th.emplace_back(work);
  for (auto& t : th) t.join();
  for (size_t i = 0; i < n; ++i)
    if (rcs[i]) return fail(rcs[i], "mesh " + std::to_string(i) + ": " + errs[i]);
  return EIK_OK;
}

int ensure_split(Mesh& m) {
  DeviceState* d = m.dev;
  if (!d || !d->blocks) return fail(EIK_ERR_STATE, "mesh layout is not uploaded");
  if (d->esel) return EIK_OK;
  std::vector<uint32_t> sel;
  std::vector<float> len;
  split_records(m.layout, m.dim, sel, len);
  int rc = upload(reinterpret_cast<uint32_t**>(&d->esel), sel, d->stream);
  if (!rc) rc = upload(&d->elen, len, d->stream);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(d->stream));
  return EIK_OK;
}

// Builds the block layouts of the meshes that have none, several meshes at a time.
int build_layouts_parallel(const std::vector<Mesh*>& ms) {
  std::vector<Mesh*> todo;
  for (Mesh* m : ms)
    if (!m->layout.valid && std::find(todo.begin(), todo.end(), m) == todo.end()) todo.push_back(m);
  return parallel_over(todo.size(), [&](size_t i) { return build_layout(*todo[i]); });
}

// Partitions every mesh for a block size and builds its layout, several meshes at a time.
int partition_parallel(const std::vector<Mesh*>& ms, int32_t vpb) {
  std::vector<Mesh*> todo;
  for (Mesh* m : ms)
    if (std::find(todo.begin(), todo.end(), m) == todo.end()) todo.push_back(m);
  return parallel_over(todo.size(), [&](size_t i) {
    int rc = partition_by_size(*todo[i], vpb);
    return rc ? rc : build_layout(*todo[i]);
  });
}

// Resident workgroups of a launch shape and the register budget it runs with: the 168-VGPR
// build when it fits as many workgroups per CU as the 128-VGPR one (the LDS is the limit), else
// the 128-VGPR build.  C4 (256 cubes, 192 vertices per block, 6 workgroups per CU either way):
// 19.2 -> 18.2 ms; at 128 vertices per block the 128-VGPR build fits 8 per CU against 6 and is
// faster (19.7 against 20.4 ms; eight 96^3 cubes 23.0 against 26.7 ms).
template <int DIM, int BT>
static int group_grid(size_t lds, int requested, int* out, int* wpe) {
  int dev = 0, ncu = 0, per4 = 0, per3 = 0;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per4, k_visit_multi<DIM, BT, 4>, BT, lds));
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per3, k_visit_multi<DIM, BT, 3>, BT, lds));
  *wpe = per3 >= per4 ? 3 : 4;
  int resident = std::max(1, ncu * std::max(std::max(per3, per4), 1));
  resident = std::min(resident, 64 * ncu);
  *out = requested > 0 ? std::min(requested, resident) : resident;
  return EIK_OK;
}

template <int DIM, int BT>
static void launch_visits(const Group& g, size_t lds, hipStream_t s, const MultiArgs& a,
                          const SolveDesc* sd) {
  if (g.wpe == 3)
    hipLaunchKernelGGL((k_visit_multi<DIM, BT, 3>), dim3(g.grid), dim3(BT), lds, s, a, sd);
  else
    hipLaunchKernelGGL((k_visit_multi<DIM, BT, 4>), dim3(g.grid), dim3(BT), lds, s, a, sd);
}

void launch_free(Launch& L) {
  void* ptrs[] = {L.grp[0].queue, L.grp[0].qctl, L.grp[1].queue, L.grp[1].qctl, L.dsolves,
                  L.qflag,        L.bstate,      L.counters,     L.out,         L.d_sseed,
                  L.d_vseed,      L.d_nseed,     L.d_fseed};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (L.hsolves) (void)hipHostFree(L.hsolves);
  if (L.h_seed) (void)hipHostFree(L.h_seed);
  if (L.h_out) (void)hipHostFree(L.h_out);
  if (L.h_cnt) (void)hipHostFree(L.h_cnt);
  if (L.ev0) (void)hipEventDestroy(L.ev0);
  if (L.ev1) (void)hipEventDestroy(L.ev1);
  if (L.stream) (void)hipStreamDestroy(L.stream);
  L = Launch();
}

int launch_alloc(Launch& L, const std::vector<Mesh*>& kinds, int32_t cap_solves, int64_t cap_blocks,
                 int64_t cap_nv, int64_t cap_seeds) {
  int max_lds = 65536, ncu = 0;
  {
    int dev = 0, v = 0;
    HIP_TRY(hipGetDevice(&dev));
    L.device = dev;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
      max_lds = v - 1024;
    HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  }
  // per element kind: LDS for its largest block, with the records read from memory (the
  // default shape) and staged (EIK_STAGE_LDS, when the records fit)
  for (int gi = 0; gi < 2; ++gi) {
    Group& g = L.grp[gi];
    g = Group();
    g.dim = gi == 0 ? 3 : 2;
    int32_t max_local = 0, max_own = 0, max_dep = 2, qcap_l = 256;
    size_t rec_bytes = 0;
    for (Mesh* m : kinds) {
      if (m->dim != g.dim) continue;
      g.present = true;
      const Layout& Ly = m->layout;
      const int K = m->dim == 3 ? 4 : 3;
      const size_t rb = m->dim == 3 ? sizeof(TetRec) : sizeof(TriRec);
      max_local = std::max(max_local, Ly.max_local);
      max_own = std::max(max_own, Ly.max_own);
      max_dep = std::max(max_dep, Ly.max_dep);
      qcap_l = std::max(qcap_l, std::min(8192, K * Ly.max_elem));
      rec_bytes = std::max(rec_bytes, ((size_t)Ly.max_elem * rb + 15) & ~(size_t)15);
    }
    if (!g.present) continue;
    Layout probe;
    probe.max_local = max_local;
    probe.max_own = max_own;
    probe.max_dep = max_dep;
    const size_t staged = visit_lds_bytes(probe, g.dim, qcap_l, rec_bytes);
    const size_t direct = visit_lds_bytes(probe, g.dim, qcap_l, 0);
    if (direct > (size_t)max_lds)
      return fail(EIK_ERR_CAPACITY, "a block does not fit the LDS; use fewer vertices per block");
    g.lds[0] = staged <= (size_t)max_lds ? staged : 0;  // 0: the staged shape is unavailable
    g.lds[1] = direct;
    for (int sh = 0; sh < 2; ++sh) {
      if (g.lds[sh] <= 65536) continue;
      const void* k[2];
      if (sh == 0) {
        k[0] = g.dim == 3 ? (const void*)k_visit_multi<3, kBlockThreads, 4>
                          : (const void*)k_visit_multi<2, kBlockThreads, 4>;
        k[1] = g.dim == 3 ? (const void*)k_visit_multi<3, kBlockThreads, 3>
                          : (const void*)k_visit_multi<2, kBlockThreads, 3>;
      } else {
        k[0] = g.dim == 3 ? (const void*)k_visit_multi<3, kDirectThreads, 4>
                          : (const void*)k_visit_multi<2, kDirectThreads, 4>;
        k[1] = g.dim == 3 ? (const void*)k_visit_multi<3, kDirectThreads, 3>
                          : (const void*)k_visit_multi<2, kDirectThreads, 3>;
      }
      for (const void* f : k)
        HIP_TRY(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)g.lds[sh]));
    }
    g.rec_lds_bytes = rec_bytes;
    g.queue_cap = qcap_l;
    // ring capacity: every block queued at once plus a waiting ticket per resident workgroup
    // (at most 64 per CU), twice over
    uint32_t qcap = 1;
    while (qcap < 2u * ((uint32_t)cap_blocks + 64u * (uint32_t)std::max(ncu, 1)) + 2u) qcap <<= 1;
    g.qcap = qcap;
    HIP_TRY(hipMalloc((void**)&g.queue, sizeof(int32_t) * kSlotStride * qcap));
    HIP_TRY(hipMalloc((void**)&g.qctl, sizeof(QueueCtl)));
  }
  L.cap_solves = cap_solves;
  L.cap_blocks = cap_blocks;
  L.cap_nv = cap_nv;
  L.cap_seeds = std::max<int64_t>(cap_seeds, 1);
  HIP_TRY(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&L.ev0));
  HIP_TRY(hipEventCreate(&L.ev1));
  HIP_TRY(hipMalloc((void**)&L.dsolves, sizeof(SolveDesc) * cap_solves));
  HIP_TRY(hipHostMalloc((void**)&L.hsolves, sizeof(SolveDesc) * cap_solves));
  HIP_TRY(hipMalloc((void**)&L.qflag, sizeof(uint32_t) * cap_blocks));
  HIP_TRY(hipMalloc((void**)&L.bstate, sizeof(uint32_t) * cap_blocks));
  HIP_TRY(hipMalloc((void**)&L.counters, sizeof(unsigned long long) * kNumCounters));
  HIP_TRY(hipMalloc((void**)&L.out, sizeof(float) * cap_nv));
  HIP_TRY(hipMalloc((void**)&L.d_sseed, sizeof(int32_t) * L.cap_seeds));
  HIP_TRY(hipMalloc((void**)&L.d_vseed, sizeof(int32_t) * L.cap_seeds));
  HIP_TRY(hipMalloc((void**)&L.d_fseed, sizeof(float) * L.cap_seeds));
  HIP_TRY(hipMalloc((void**)&L.d_nseed, sizeof(int32_t)));
  HIP_TRY(hipHostMalloc((void**)&L.h_seed, sizeof(int32_t) * (3 * L.cap_seeds + 1)));
  HIP_TRY(hipHostMalloc((void**)&L.h_out, sizeof(float) * cap_nv));
  HIP_TRY(hipHostMalloc((void**)&L.h_cnt, sizeof(unsigned long long) * kNumCounters));
  return EIK_OK;
}

void record_factors(const Mesh& m, const float* mult, float* rmult, hipStream_t s) {
  const int64_t n = m.layout.total_elems;
  hipLaunchKernelGGL(k_record_factors, dim3((unsigned)std::min<int64_t>(grid1d(n, 256), 4096)),
                     dim3(256), 0, s, mult, (const int32_t*)m.dev->rec_elem, n, rmult);
}

void solve_desc(SolveDesc& s, Mesh* m, float* T, float* seen, const float* rmult) {
  DeviceState* d = m->dev;
  s.blocks = d->blocks;
  s.elems = d->elems;
  s.esel = d->esel;
  s.elen = d->elen;
  s.halo = d->halo;
  s.dep_start = d->dep_start;
  s.dep = d->dep;
  s.vmask = d->vmask;
  s.rmult = rmult;
  s.iperm = d->iperm;
  s.owner = d->owner;
  s.T = T;
  s.seen = seen;
  s.nv = m->nv;
  s.nseen = d->nseen;
  s.nblocks = d->nblocks;
  s.dim = m->dim;
  s.large = m->dim == 3 ? EIK_LARGENUM_3D : EIK_LARGENUM_2D;
}

int launch_args(Launch& L, int32_t nsolves, int64_t total_blocks, const eik_solve_opts_t& o,
                MultiArgs a[2]) {
  // queue entries: solve index above the launch-wide block index, both in 31 bits
  int32_t bbits = 1;
  while ((1ll << bbits) <= total_blocks) ++bbits;
  if (bbits > 30 || ((int64_t)nsolves << bbits) >= (1ll << 31) - 1)
    return fail(EIK_ERR_CAPACITY, "too many blocks and solves in one launch");
  for (int gi = 0; gi < 2; ++gi) {
    Group& g = L.grp[gi];
    std::memset(&a[gi], 0, sizeof(MultiArgs));
    if (!g.present) continue;
    // many independent solves keep every workgroup busy: records read from memory, smaller
    // workgroups, unless staging is asked for
    g.shape = o.stage_records == EIK_STAGE_LDS ? 0 : 1;
    if (g.shape == 0 && g.lds[0] == 0)
      return fail(EIK_ERR_CAPACITY, "the element records of a block do not fit the LDS for "
                                    "staging; use fewer vertices per block or read the records "
                                    "from memory (EIK_STAGE_AUTO)");
    int rc;
    if (g.shape == 0)
      rc = gi == 0 ? group_grid<3, kBlockThreads>(g.lds[0], o.workgroups, &g.grid, &g.wpe)
                   : group_grid<2, kBlockThreads>(g.lds[0], o.workgroups, &g.grid, &g.wpe);
    else
      rc = gi == 0 ? group_grid<3, kDirectThreads>(g.lds[1], o.workgroups, &g.grid, &g.wpe)
                   : group_grid<2, kDirectThreads>(g.lds[1], o.workgroups, &g.grid, &g.wpe);
    if (rc) return rc;
    MultiArgs& x = a[gi];
    x.solves = L.dsolves;
    x.queue = g.queue;
    x.qflag = L.qflag;
    x.bstate = L.bstate;
    x.ctl = g.qctl;
    x.qmask = g.qcap - 1u;
    x.bbits = bbits;
    x.bmask = (int32_t)((1u << bbits) - 1u);
    x.counters = L.counters;
    x.eps = o.eps;
    x.max_inner = default_inner_sweeps(g.dim, o.max_inner);
    x.queue_cap = g.queue_cap;
    x.rec_lds_bytes = g.shape == 0 ? g.rec_lds_bytes : 0;
    x.max_visits = (int64_t)(o.max_iterations > 0 ? o.max_iterations : 1000) * std::max<int64_t>(total_blocks, 1);
    x.wait_limit = wait_limit_ticks(L.device, o.wait_limit_us);
  }
  return EIK_OK;
}

int launch_check_end(const Launch& L) {
  const unsigned long long n = L.h_cnt[kCntAbandoned];
  if (n == 0) return EIK_OK;
  for (int gi = 0; gi < 2; ++gi)
    if (L.grp[gi].present) {
      int rc = check_queue_end(L.grp[gi].qctl, L.stream, n);
      if (rc) return rc;
    }
  return EIK_OK;
}

int launch_output(Launch& L, int32_t nsolves, int64_t total_nv) {
  hipLaunchKernelGGL(k_multi_gather, dim3(64, nsolves), dim3(256), 0, L.stream, L.dsolves, L.out);
  HIP_TRY(hipMemcpyAsync(L.h_out, L.out, sizeof(float) * total_nv, hipMemcpyDeviceToHost, L.stream));
  HIP_TRY(hipGetLastError());
  return EIK_OK;
}

int launch_issue(Launch& L, int32_t nsolves, int64_t total_blocks, int64_t total_nv,
                 const MultiArgs a[2], bool upload_desc, bool output) {
  hipStream_t s = L.stream;
  const int64_t cap = L.cap_seeds;
  if (upload_desc)
    HIP_TRY(hipMemcpyAsync(L.dsolves, L.hsolves, sizeof(SolveDesc) * nsolves, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(L.d_sseed, L.h_seed, sizeof(int32_t) * cap, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(L.d_vseed, L.h_seed + cap, sizeof(int32_t) * cap, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(L.d_fseed, L.h_seed + 2 * cap, sizeof(float) * cap, hipMemcpyHostToDevice, s));
  HIP_TRY(hipMemcpyAsync(L.d_nseed, L.h_seed + 3 * cap, sizeof(int32_t), hipMemcpyHostToDevice, s));
  for (int gi = 0; gi < 2; ++gi) {
    const Group& g = L.grp[gi];
    if (!g.present) continue;
    const int64_t qwords = (int64_t)g.qcap * kSlotStride;
    hipLaunchKernelGGL(k_fill_i32, dim3(grid1d(qwords, 256)), dim3(256), 0, s, g.queue, qwords, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, s, reinterpret_cast<int32_t*>(g.qctl),
                       (int64_t)(sizeof(QueueCtl) / 4), 0);
  }
  if (total_blocks > 0) {
    hipLaunchKernelGGL(k_fill_i32, dim3(grid1d(total_blocks, 256)), dim3(256), 0, s,
                       reinterpret_cast<int32_t*>(L.qflag), total_blocks, 0);
    hipLaunchKernelGGL(k_fill_i32, dim3(grid1d(total_blocks, 256)), dim3(256), 0, s,
                       reinterpret_cast<int32_t*>(L.bstate), total_blocks, 0);
  }
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, L.counters, (int64_t)kNumCounters);
  hipLaunchKernelGGL(k_multi_init, dim3(64, nsolves), dim3(256), 0, s, L.dsolves);
  hipLaunchKernelGGL(k_multi_seed, dim3((unsigned)std::min<int64_t>(cap, 1 << 16)), dim3(64), 0, s,
                     (const SolveDesc*)L.dsolves, a[0], a[1], L.d_sseed, L.d_vseed, L.d_fseed,
                     L.d_nseed);
  const SolveDesc* sd = L.dsolves;
  const Group& g3 = L.grp[0];
  const Group& g2 = L.grp[1];
  if (g3.present && g3.shape == 0) launch_visits<3, kBlockThreads>(g3, g3.lds[0], s, a[0], sd);
  if (g3.present && g3.shape == 1) launch_visits<3, kDirectThreads>(g3, g3.lds[1], s, a[0], sd);
  if (g2.present && g2.shape == 0) launch_visits<2, kBlockThreads>(g2, g2.lds[0], s, a[1], sd);
  if (g2.present && g2.shape == 1) launch_visits<2, kDirectThreads>(g2, g2.lds[1], s, a[1], sd);
  HIP_TRY(hipMemcpyAsync(L.h_cnt, L.counters, sizeof(unsigned long long) * kNumCounters,
                         hipMemcpyDeviceToHost, s));
  HIP_TRY(hipGetLastError());
  return output ? launch_output(L, nsolves, total_nv) : EIK_OK;
}

void launch_stats(const Launch& L, float kernel_ms, eik_solve_stats_t* out) {
  eik_solve_stats_t st;
  std::memset(&st, 0, sizeof(st));
  const unsigned long long* c = L.h_cnt;
  st.kernel_ms = kernel_ms;
  st.outer_iterations = 1;
  st.block_visits = (int64_t)c[0];
  st.vertex_updates = (int64_t)c[1];
  st.inner_sweeps = (int64_t)c[2];
  st.local_solves = (int64_t)c[3];
  st.algo_bytes = (int64_t)c[4];
  st.algo_bytes_loaded = (int64_t)c[kCntAlgoLoaded];
  st.visit_ms = kernel_ms;
  st.visit_launches = (L.grp[0].present ? 1 : 0) + (L.grp[1].present ? 1 : 0);
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, L.device) == hipSuccess && khz > 0) {
    st.busy_ms = (double)c[5] / khz;
    st.wait_ms = (double)c[6] / khz;
    st.load_ms = (double)c[7] / khz;
    st.sweep_ms = (double)c[8] / khz;
    st.writeback_ms = (double)c[9] / khz;
    st.queue_ms = (double)c[10] / khz;
  }
  st.reruns = (int64_t)c[11];
  st.abandoned_waits = (int64_t)c[kCntAbandoned];
  *out = st;
}

// ------------------------------------------------------------------------------------------
// Batches
// ------------------------------------------------------------------------------------------
struct Batch {
  std::vector<Mesh*> meshes;
  Launch L;
  float** rmult = nullptr;  // per solve: its factors in record order (owned), or null
  float** mult = nullptr;   // per solve: its factors as uploaded (element order, owned), or null
  int32_t* dbad = nullptr;  // factor check flag (device) and its copy (pinned host)
  int32_t* hbad = nullptr;
  int64_t total_blocks = 0;
  int64_t total_nv = 0;
  float* T = nullptr;     // per solve: its own times and seen values (a mesh may appear in
  float* seen = nullptr;  // several solves of the batch)
  int64_t seed_cap = 0;   // seeds the launch buffers hold
  bool use_graph = true;
  hipGraphExec_t gexec = nullptr;
  hipGraph_t graph = nullptr;
  MultiArgs graph_args[2];  // arguments, grids, shapes and register budgets the graph was
  int graph_grid[2] = {0, 0};  // captured with
  int graph_shape[2] = {-1, -1};
  int graph_wpe[2] = {0, 0};
  bool solved = false;  // a solve has run since the solve arrays were (re)allocated
};

static void batch_free_graph(Batch* b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
}

static void batch_free(Batch* b) {
  if (!b) return;
  batch_free_graph(b);
  launch_free(b->L);
  if (b->T) (void)hipFree(b->T);
  if (b->seen) (void)hipFree(b->seen);
  if (b->rmult) {
    for (size_t i = 0; i < b->meshes.size(); ++i)
      if (b->rmult[i]) (void)hipFree(b->rmult[i]);
    delete[] b->rmult;
  }
  if (b->mult) {
    for (size_t i = 0; i < b->meshes.size(); ++i)
      if (b->mult[i]) (void)hipFree(b->mult[i]);
    delete[] b->mult;
  }
  if (b->dbad) (void)hipFree(b->dbad);
  if (b->hbad) (void)hipHostFree(b->hbad);
  delete b;
}

static int batch_alloc(Batch* b, int64_t seed_cap) {
  // (re)allocates the launch resources for a seed capacity; the descriptors are re-uploaded
  batch_free_graph(b);
  launch_free(b->L);
  int rc = launch_alloc(b->L, b->meshes, (int32_t)b->meshes.size(), b->total_blocks, b->total_nv,
                        seed_cap);
  if (rc) return rc;
  b->seed_cap = b->L.cap_seeds;
  int64_t so = 0, off = 0, bbase = 0;
  for (size_t i = 0; i < b->meshes.size(); ++i) {
    Mesh* m = b->meshes[i];
    SolveDesc& s = b->L.hsolves[i];
    solve_desc(s, m, b->T + off, b->seen + so, b->rmult[i]);
    s.out_off = off;
    s.bbase = (int32_t)bbase;
    so += s.nseen;
    off += m->nv;
    bbase += s.nblocks;
  }
  HIP_TRY(hipMemcpy(b->L.dsolves, b->L.hsolves, sizeof(SolveDesc) * b->meshes.size(),
                    hipMemcpyHostToDevice));
  return EIK_OK;
}

static int batch_create(std::vector<Mesh*> ms, Batch** out) {
  if (ms.empty()) return fail(EIK_ERR_ARG, "empty batch");
  for (Mesh* m : ms)
    if (m->nblocks <= 0) return fail(EIK_ERR_STATE, "every mesh of a batch must be partitioned");
  int rc = build_layouts_parallel(ms);
  if (rc) return rc;
  Batch* b = new Batch();
  b->meshes = ms;
  b->rmult = new float*[ms.size()]();
  b->mult = new float*[ms.size()]();
  auto bail = [&](int code) {
    batch_free(b);
    return code;
  };
  // device layouts (each mesh keeps its own, also usable by single solves)
  for (Mesh* m : ms) {
    rc = upload_layout(*m);
    if (!rc) rc = ensure_split(*m);
    if (rc) return bail(rc);
  }
  int64_t total_seen = 0;
  for (Mesh* m : ms) {
    b->total_blocks += m->dev->nblocks;
    b->total_nv += m->nv;
    total_seen += m->dev->nseen;
  }
  if (hipMalloc((void**)&b->T, sizeof(float) * b->total_nv) != hipSuccess ||
      hipMalloc((void**)&b->seen, sizeof(float) * total_seen) != hipSuccess)
    return bail(fail(EIK_ERR_NOMEM, "batch allocation failed"));
  rc = batch_alloc(b, 1024);
  if (rc) return bail(rc);
  *out = b;
  return EIK_OK;
}

static bool same_args(const MultiArgs& x, const MultiArgs& y) {
  return std::memcmp(&x, &y, sizeof(MultiArgs)) == 0;
}

static int batch_solve(Batch* b, const int64_t* seed_off, const int32_t* seeds,
                       const float* seed_vals, const eik_solve_opts_t& o, float* T_out,
                       eik_solve_stats_t* stats) {
  const size_t ns = b->meshes.size();
  const int64_t nseeds = seed_off[ns];
  if (seed_off[0] != 0) return fail(EIK_ERR_ARG, "seed offsets must start at 0");
  for (size_t i = 0; i < ns; ++i) {
    if (seed_off[i + 1] < seed_off[i]) return fail(EIK_ERR_ARG, "seed offsets must not decrease");
    for (int64_t k = seed_off[i]; k < seed_off[i + 1]; ++k) {
      if (seeds[k] < 0 || seeds[k] >= b->meshes[i]->nv)
        return fail(EIK_ERR_ARG, "seed vertex out of range (solve " + std::to_string(i) + ")");
      if (seed_vals && !(seed_vals[k] >= 0.f && seed_vals[k] < 1e30f))
        return fail(EIK_ERR_ARG, "seed values must be finite and >= 0");
    }
  }
  int rc;
  b->solved = false;  // until this solve has finished
  if (nseeds > b->seed_cap && (rc = batch_alloc(b, nseeds))) return rc;
  Launch& L = b->L;
  const int64_t cap = L.cap_seeds;
  for (size_t i = 0; i < ns; ++i)
    for (int64_t k = seed_off[i]; k < seed_off[i + 1]; ++k) {
      L.h_seed[k] = (int32_t)i;
      L.h_seed[cap + k] = seeds[k];
      float v = seed_vals ? seed_vals[k] : 0.f;
      std::memcpy(&L.h_seed[2 * cap + k], &v, 4);
    }
  L.h_seed[3 * cap] = (int32_t)nseeds;
  MultiArgs a[2];
  rc = launch_args(L, (int32_t)ns, b->total_blocks, o, a);
  if (rc) return rc;
  hipStream_t s = L.stream;
  auto t0 = std::chrono::steady_clock::now();
  if (b->use_graph) {
    if (b->gexec && (!same_args(a[0], b->graph_args[0]) || !same_args(a[1], b->graph_args[1]) ||
                     b->graph_grid[0] != L.grp[0].grid || b->graph_grid[1] != L.grp[1].grid ||
                     b->graph_shape[0] != L.grp[0].shape || b->graph_shape[1] != L.grp[1].shape ||
                     b->graph_wpe[0] != L.grp[0].wpe || b->graph_wpe[1] != L.grp[1].wpe))
      batch_free_graph(b);
    if (!b->gexec) {
      HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      rc = launch_issue(L, (int32_t)ns, b->total_blocks, b->total_nv, a, false, false);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(s, &g);
      if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (e != hipSuccess) return fail(EIK_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
      b->graph = g;
      HIP_TRY(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
      for (int gi = 0; gi < 2; ++gi) {
        b->graph_args[gi] = a[gi];
        b->graph_grid[gi] = L.grp[gi].grid;
        b->graph_shape[gi] = L.grp[gi].shape;
        b->graph_wpe[gi] = L.grp[gi].wpe;
      }
    }
    HIP_TRY(hipEventRecord(L.ev0, s));
    HIP_TRY(hipGraphLaunch(b->gexec, s));
    HIP_TRY(hipEventRecord(L.ev1, s));
  } else {
    HIP_TRY(hipEventRecord(L.ev0, s));
    rc = launch_issue(L, (int32_t)ns, b->total_blocks, b->total_nv, a, false, false);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(L.ev1, s));
  }
  // the times leave the device only when asked for (eik_batch_get_result fetches one solve's)
  if (T_out && (rc = launch_output(L, (int32_t)ns, b->total_nv))) return rc;
  HIP_TRY(hipStreamSynchronize(s));
  auto t1 = std::chrono::steady_clock::now();
  if (stats) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, L.ev0, L.ev1));
    launch_stats(L, ms, stats);
    stats->wall_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  }
  rc = launch_check_end(L);
  if (rc) return rc;
  b->solved = true;
  if (T_out) std::memcpy(T_out, L.h_out, sizeof(float) * b->total_nv);
  return EIK_OK;
}

// Times of solve i of the last batched solve, in caller order: gathered on the device and
// copied (nv of that mesh's values).
static int batch_get_result(Batch* b, int32_t i, float* T_out) {
  if (i < 0 || i >= (int32_t)b->meshes.size()) return fail(EIK_ERR_ARG, "solve index out of range");
  if (!b->solved) return fail(EIK_ERR_STATE, "the batch has not been solved");
  Launch& L = b->L;
  const SolveDesc& d = L.hsolves[i];
  hipLaunchKernelGGL(k_multi_gather, dim3(64, 1), dim3(256), 0, L.stream, L.dsolves + i, L.out);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(T_out, L.out + d.out_off, sizeof(float) * d.nv, hipMemcpyDeviceToHost,
                         L.stream));
  HIP_TRY(hipStreamSynchronize(L.stream));
  return EIK_OK;
}

static int batch_set_factors(Batch* b, int32_t i, const float* f, int64_t count) {
  if (i < 0 || i >= (int32_t)b->meshes.size()) return fail(EIK_ERR_ARG, "solve index out of range");
  Mesh* m = b->meshes[i];
  if (f && count != m->ne) return fail(EIK_ERR_ARG, "one factor per element is required");
  hipStream_t s = b->L.stream;
  // the batch's stream is idle between calls (batch_solve ends with a synchronisation), so the
  // buffers below are not in use by a launch
  if (!f) {
    if (b->rmult[i]) HIP_TRY(hipFree(b->rmult[i]));
    if (b->mult[i]) HIP_TRY(hipFree(b->mult[i]));
    b->rmult[i] = nullptr;
    b->mult[i] = nullptr;
  } else {
    // both buffers stay with the solve slot, so a slot that takes a new speed field per solve
    // (serving) allocates nothing after its first
    if (!b->rmult[i])
      HIP_TRY(hipMalloc((void**)&b->rmult[i], sizeof(float) * std::max<int64_t>(m->layout.total_elems, 1)));
    if (!b->mult[i]) HIP_TRY(hipMalloc((void**)&b->mult[i], sizeof(float) * count));
    if (!b->dbad) HIP_TRY(hipMalloc((void**)&b->dbad, sizeof(int32_t)));
    if (!b->hbad) HIP_TRY(hipHostMalloc((void**)&b->hbad, sizeof(int32_t)));
    // the values are checked on the device, where they land anyway (a host pass over a large
    // field costs about a millisecond), before they replace the solve's current factors
    HIP_TRY(hipMemcpyAsync(b->mult[i], f, sizeof(float) * count, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, b->dbad, (int64_t)1, 0);
    hipLaunchKernelGGL(k_check_factors, dim3((unsigned)std::min<int64_t>(grid1d(count, 256), 4096)),
                       dim3(256), 0, s, b->mult[i], count, b->dbad);
    HIP_TRY(hipMemcpyAsync(b->hbad, b->dbad, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    // also: the caller's array has been read when the call returns
    HIP_TRY(hipStreamSynchronize(s));
    if (*b->hbad) return fail(EIK_ERR_ARG, "factors must be finite and > 0");
    // ordered before the next solve on the same stream
    record_factors(*m, b->mult[i], b->rmult[i], s);
    HIP_TRY(hipGetLastError());
  }
  if (b->L.hsolves[i].rmult != b->rmult[i]) {
    b->L.hsolves[i].rmult = b->rmult[i];
    HIP_TRY(hipMemcpy(&b->L.dsolves[i], &b->L.hsolves[i], sizeof(SolveDesc), hipMemcpyHostToDevice));
  }
  return EIK_OK;
}

// Times the batched solve of the meshes for every (block size, inner sweep cap) candidate:
// one warm-up and `reps` timed solves each (device time of the launch sequence), median kept.
// The meshes end partitioned with the fastest block size.
static int autotune(const std::vector<Mesh*>& ms, const int64_t* seed_off, const int32_t* seeds,
                    const int32_t* vpb, int32_t nvpb, const int32_t* inner, int32_t ninner,
                    int32_t reps, eik_tune_result_t* best, double* times) {
  double best_ms = 1e300;
  int32_t best_vpb = vpb[0], best_inner = inner[0];
  for (int32_t i = 0; i < nvpb; ++i) {
    int rc = partition_parallel(ms, vpb[i]);
    if (rc) return rc;
    Batch* b = nullptr;
    rc = batch_create(ms, &b);
    if (rc) return rc;
    for (int32_t j = 0; j < ninner; ++j) {
      eik_solve_opts_t o;
      eik_solve_opts_default(&o);
      o.max_inner = inner[j];
      std::vector<double> t;
      for (int32_t r = 0; r <= std::max(1, reps); ++r) {
        eik_solve_stats_t st;
        rc = batch_solve(b, seed_off, seeds, nullptr, o, nullptr, &st);
        if (rc) {
          batch_free(b);
          return rc;
        }
        if (r > 0) t.push_back(st.kernel_ms);
      }
      std::sort(t.begin(), t.end());
      const double med = t[t.size() / 2];
      if (times) times[i * ninner + j] = med;
      if (med < best_ms) {
        best_ms = med;
        best_vpb = vpb[i];
        best_inner = inner[j];
      }
    }
    batch_free(b);
  }
  int rc = partition_parallel(ms, best_vpb);
  if (rc) return rc;
  best->verts_per_block = best_vpb;
  best->max_inner = best_inner;
  best->ms = best_ms;
  return EIK_OK;
}

}  // namespace eik

using namespace eik;

extern "C" {

int eik_batch_create(const eik_mesh_t* meshes, int32_t n, eik_batch_t* out) {
  if (!meshes || n <= 0 || !out) return fail(EIK_ERR_ARG, "null or empty argument");
  std::vector<Mesh*> ms(n);
  for (int32_t i = 0; i < n; ++i) {
    ms[i] = reinterpret_cast<Mesh*>(meshes[i]);
    if (!ms[i]) return fail(EIK_ERR_ARG, "null mesh in batch");
  }
  Batch* b = nullptr;
  int rc = batch_create(ms, &b);
  if (rc) return rc;
  *out = reinterpret_cast<eik_batch_t>(b);
  return EIK_OK;
}

int eik_batch_destroy(eik_batch_t batch) {
  batch_free(reinterpret_cast<Batch*>(batch));
  return EIK_OK;
}

int eik_batch_info(eik_batch_t batch, int64_t* total_verts, int64_t* total_blocks) {
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (!b) return fail(EIK_ERR_ARG, "null batch");
  if (total_verts) *total_verts = b->total_nv;
  if (total_blocks) *total_blocks = b->total_blocks;
  return EIK_OK;
}

int eik_batch_set_graph(eik_batch_t batch, int32_t enable) {
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (!b) return fail(EIK_ERR_ARG, "null batch");
  b->use_graph = enable != 0;
  if (!b->use_graph) batch_free_graph(b);
  return EIK_OK;
}

int eik_batch_set_factors(eik_batch_t batch, int32_t i, const float* factors, int64_t count) {
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (!b) return fail(EIK_ERR_ARG, "null batch");
  return batch_set_factors(b, i, factors, count);
}

int eik_batch_get_result(eik_batch_t batch, int32_t i, float* T_out) {
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (!b || !T_out) return fail(EIK_ERR_ARG, "null argument");
  return batch_get_result(b, i, T_out);
}

int eik_batch_solve(eik_batch_t batch, const int64_t* seed_offsets, const int32_t* seeds,
                    const float* seed_vals, const eik_solve_opts_t* opts, float* T_out,
                    eik_solve_stats_t* stats) {
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (!b || !seed_offsets || !seeds) return fail(EIK_ERR_ARG, "null argument");
  eik_solve_opts_t o;
  eik_solve_opts_default(&o);
  if (opts) o = *opts;
  if (o.wait_limit_us < 0) return fail(EIK_ERR_ARG, "wait_limit_us must be >= 0");
  return batch_solve(b, seed_offsets, seeds, seed_vals, o, T_out, stats);
}

int eik_batch_autotune(const eik_mesh_t* meshes, int32_t n, const int64_t* seed_offsets,
                       const int32_t* seeds, const int32_t* verts_per_block, int32_t nvpb,
                       const int32_t* max_inner, int32_t ninner, int32_t reps,
                       eik_tune_result_t* best, double* times) {
  if (!meshes || n <= 0 || !seed_offsets || !seeds || !verts_per_block || nvpb <= 0 ||
      !max_inner || ninner <= 0 || !best)
    return fail(EIK_ERR_ARG, "null or empty argument");
  std::vector<Mesh*> ms(n);
  for (int32_t i = 0; i < n; ++i) {
    ms[i] = reinterpret_cast<Mesh*>(meshes[i]);
    if (!ms[i]) return fail(EIK_ERR_ARG, "null mesh");
  }
  for (int32_t i = 0; i < nvpb; ++i)
    if (verts_per_block[i] <= 0) return fail(EIK_ERR_ARG, "verts_per_block must be > 0");
  for (int32_t j = 0; j < ninner; ++j)
    if (max_inner[j] <= 0) return fail(EIK_ERR_ARG, "max_inner must be > 0");
  return autotune(ms, seed_offsets, seeds, verts_per_block, nvpb, max_inner, ninner, reps, best,
                  times);
}

}  // extern "C"
