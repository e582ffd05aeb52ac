// multi.h -- several solves in one persistent launch: the solve descriptors, the launch
// resources, and the host helpers shared by batches (batch.hip) and the serving loop
// (serve.hip).
//
// Every solve of a launch has a descriptor (its mesh's device layout, its own time arrays and,
// optionally, a per-element factor on the squared edge lengths: a request's own speed field over
// a resident mesh).  The blocks of all solves share one set of queue flags, solve i's blocks
// being numbered bbase_i + b, and a queue entry carries the solve index, so the workgroups of the
// launch pull blocks of any solve from one queue: small solves whose wavefront would leave most
// of the GPU idle fill it together.  Tets and triangles keep separate queues and kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "device.h"

namespace eik {

struct alignas(16) SolveDesc {
  const BlockDesc* blocks;
  const void* elems;
  const void* esel;          // split records: selection part (SelRec) and lengths
  const float* elen;
  const int32_t* halo;
  const int32_t* dep_start;
  const int32_t* dep;
  const uint64_t* vmask;     // null: every dependent is queued on a boundary change
  const float* rmult;        // per record (record order) factor on the squared edge lengths,
                             // or null
  const int32_t* iperm;      // caller id -> new id
  const int32_t* owner;      // new id -> block
  float* T;
  float* seen;
  int64_t nv;
  int64_t nseen;
  int64_t out_off;           // first value of this solve in the gathered output
  int32_t bbase;             // first launch-wide block index
  int32_t nblocks;
  int32_t dim;
  float large;
};

struct MultiArgs {
  const SolveDesc* solves;
  int32_t* queue;
  uint32_t* qflag;
  uint32_t* bstate;
  QueueCtl* ctl;
  uint32_t qmask;
  int32_t bbits;
  int32_t bmask;
  unsigned long long* counters;
  float eps;
  int32_t max_inner;
  int32_t queue_cap;
  size_t rec_lds_bytes;
  int64_t max_visits;
  unsigned long long wait_limit;  // wall-clock ticks a queue wait may last (q_take)
};

// The solves of one element kind in a launch: their own queue and launch shape.  Shape 0
// stages the element records in LDS (EIK_STAGE_LDS; lds[0] is 0 when the largest block's
// records do not fit), shape 1 reads them from memory with smaller workgroups, the per-solve
// factors applied at every use (batch.hip, k_visit_multi).
struct Group {
  int dim = 3;
  bool present = false;
  int32_t* queue = nullptr;
  QueueCtl* qctl = nullptr;
  uint32_t qcap = 0;
  size_t lds[2] = {0, 0};
  size_t rec_lds_bytes = 0;  // shape 0
  int32_t queue_cap = 0;
  int shape = 0;             // chosen per launch (launch_args)
  int grid = 0;
  int wpe = 4;               // register budget of the visit kernel: waves per SIMD (group_grid)
};

// Device and pinned host resources of one launch sequence, sized for a capacity: at most
// cap_solves solves, cap_blocks blocks and cap_nv vertices (all solves together), cap_seeds
// seeds.  The descriptors are written per launch from the pinned hsolves.
struct Launch {
  int device = -1;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  Group grp[2];  // tets, triangles
  int32_t cap_solves = 0;
  int64_t cap_blocks = 0, cap_nv = 0, cap_seeds = 0;
  SolveDesc* dsolves = nullptr;
  SolveDesc* hsolves = nullptr;  // pinned
  uint32_t* qflag = nullptr;
  uint32_t* bstate = nullptr;
  unsigned long long* counters = nullptr;
  float* out = nullptr;
  int32_t *d_sseed = nullptr, *d_vseed = nullptr, *d_nseed = nullptr;
  float* d_fseed = nullptr;
  int32_t* h_seed = nullptr;  // pinned: [solve x cap][vertex x cap][value bits x cap][count]
  float* h_out = nullptr;     // pinned: gathered times
  unsigned long long* h_cnt = nullptr;
};

// batch.hip
int build_layouts_parallel(const std::vector<Mesh*>& ms);
// Uploads the split element records of an uploaded mesh (split_records) if it has none: the
// multi-solve visits that read element data from memory use them.
int ensure_split(Mesh& m);
int partition_parallel(const std::vector<Mesh*>& ms, int32_t vpb);
// Shapes the per-kind groups for every mesh that may take part, allocates for the capacity.
int launch_alloc(Launch& L, const std::vector<Mesh*>& kinds, int32_t cap_solves,
                 int64_t cap_blocks, int64_t cap_nv, int64_t cap_seeds);
void launch_free(Launch& L);
// Fills the descriptor of a solve of mesh m with the given arrays (out_off, bbase by caller).
void solve_desc(SolveDesc& s, Mesh* m, float* T, float* seen, const float* rmult);
// Per-record factors of mesh m (rmult, layout.total_elems values) from per-element factors
// (mult, device), on stream s.
void record_factors(const Mesh& m, const float* mult, float* rmult, hipStream_t s);
// Kernel arguments for a launch of nsolves solves and total_blocks blocks.
int launch_args(Launch& L, int32_t nsolves, int64_t total_blocks, const eik_solve_opts_t& o,
                MultiArgs a[2]);
// Issues one launch sequence on L.stream: descriptors (from hsolves when upload_desc), seeds
// (from h_seed), clears, init, seeding, the per-kind visit kernels, the copy of the counters to
// the pinned buffer and, with output, the times (launch_output).
int launch_issue(Launch& L, int32_t nsolves, int64_t total_blocks, int64_t total_nv,
                 const MultiArgs a[2], bool upload_desc, bool output = true);
// Gathers the times of the launch's solves into caller order and copies them (total_nv values)
// to the pinned buffer, on L.stream.
int launch_output(Launch& L, int32_t nsolves, int64_t total_nv);
void launch_stats(const Launch& L, float kernel_ms, eik_solve_stats_t* st);
// After a finished launch: EIK_ERR_TIMEOUT if workgroups abandoned queue waits and a queue
// never reached its end (check_queue_end), else EIK_OK.
int launch_check_end(const Launch& L);

}  // namespace eik
