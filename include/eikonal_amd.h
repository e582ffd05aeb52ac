/*
 * eikonal_amd.h -- C ABI of the MI355X-native Fast Iterative Method (FIM) Eikonal solver.
 *
 * This is the drop-in boundary for the reference's solve path (SCIInstitute/SCI-Solver_Eikonal).
 * The reference exposes one C++ class, `Eikonal` (src/Eikonal.h:14-61), over two host drivers
 * (meshFIM3dEikonal, meshFIM2dEikonal) and CUDA kernels.  Every entry point below names the
 * reference interface it replaces.  The C++ facade in sci-solver_eikonal_amd/csrc/Eikonal.h keeps
 * the reference's class API on top of these functions.
 *
 * Conventions
 *   - Every function returns an int status: EIK_OK (0) or one of the EIK_ERR_* codes; the
 *     library never calls exit() (the facade maps codes back to the reference's exit(1)/exit(0)).
 *     eik_last_error() returns a message for the last failure on the calling thread.
 *   - Arrays are caller-owned host arrays (plain pointers and element counts); device state is
 *     owned by the mesh handle.  Handles are not thread-safe: one thread per handle at a time.
 *   - Vertex ids, element ids and seeds are 0-based in the caller's numbering; results are
 *     returned in the caller's vertex order.
 *   - Times are fp32.  Unreached vertices read EIK_LARGENUM_3D (tet meshes) or EIK_LARGENUM_2D
 *     (triangle meshes), the reference's LARGENUM values (tetmesh.h:11-13, TriMesh.h:11-13).
 *   - All device work of a handle is issued on the handle's own HIP stream (eik_mesh_stream).
 */
#ifndef EIKONAL_AMD_H
#define EIKONAL_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EIK_OK 0
#define EIK_ERR_ARG 1      /* invalid argument (null pointer, bad size, bad index, ...) */
#define EIK_ERR_IO 2       /* mesh file missing or unreadable */
#define EIK_ERR_STATE 3    /* call out of order (e.g. solve before partition) */
#define EIK_ERR_HIP 4      /* HIP runtime error (no device, launch failure, ...) */
#define EIK_ERR_CAPACITY 5 /* a block exceeds the kernel's local capacity */
#define EIK_ERR_NOMEM 6    /* host or device allocation failed */
#define EIK_ERR_TIMEOUT 7  /* async / batch / served solve: a workgroup waited longer than
                              wait_limit_us for queued work and the queue never reached its end
                              (the result is incomplete; the device is left in a usable state) */

#define EIK_LARGENUM_3D 1.0e9f
#define EIK_LARGENUM_2D 1.0e7f

typedef struct eik_mesh* eik_mesh_t;

/* ---------------------------------------------------------------------------------------- */
/* Errors                                                                                   */
/* ---------------------------------------------------------------------------------------- */
const char* eik_last_error(void);
const char* eik_version(void);

/* ---------------------------------------------------------------------------------------- */
/* Meshes                                                                                   */
/* ---------------------------------------------------------------------------------------- */

/* Tet mesh from arrays.  xyz: 3*nv floats; tets: 4*nt vertex ids, 0- or 1-based (the minimum
 * index decides, as TetMesh::init does: src/core3d/cuda/tetmesh.cu:9-53).
 * Replaces tetgenio + TetMesh::init (src/Eikonal.cu:111-127). */
int eik_mesh_create_tet(const float* xyz, int64_t nv, const int32_t* tets, int64_t nt,
                        eik_mesh_t* out);

/* Triangle mesh (planar or surface) from arrays; faces: 3*nf 0-based vertex ids.
 * Replaces TriMesh construction + meshFIM2dEikonal::SetMesh (meshFIM2d_eikonal.h:40-72). */
int eik_mesh_create_tri(const float* xyz, int64_t nv, const int32_t* faces, int64_t nf,
                        eik_mesh_t* out);

/* TetGen .node/.ele pair, `basename` without extension.
 * Replaces tetgenio::load_tetmesh (src/core3d/cuda/tetgen.cu:1993-2320); EIK_ERR_IO where the
 * reference exits with code 1 (src/Eikonal.cu:112-115). */
int eik_mesh_load_tetgen(const char* basename, eik_mesh_t* out);

/* ASCII or binary PLY triangle mesh.  Replaces TriMesh::read (src/core2d/cuda/TriMesh_io.cu:154);
 * EIK_ERR_IO where the reference exits with code 1 (src/Eikonal.cu:102-107). */
int eik_mesh_load_ply(const char* path, eik_mesh_t* out);

int eik_mesh_destroy(eik_mesh_t mesh);

/* Partition quality follows from the block fields (set once a partition exists, kept across speed
 * changes): element duplication = block_elems / (nelem + nvirtual), imbalance = max_block_verts *
 * nblocks / nv, halo vertices per block = block_halo / nblocks. */
typedef struct {
  int32_t dim;           /* 3 = tetrahedra, 2 = triangles */
  int32_t nblocks;       /* 0 before partitioning */
  int64_t nv;            /* vertices */
  int64_t nelem;         /* tets or faces */
  int64_t nvirtual;      /* 2D virtual faces for obtuse corners (TriMesh.h:191-313) */
  int64_t block_elems;   /* elements summed over blocks (halo copies counted per block) */
  int32_t max_block_verts;  /* owned vertices, largest block */
  int32_t max_block_local;  /* owned + halo vertices, largest block */
  int32_t max_block_elems;  /* elements, largest block */
  int32_t pad;
  int64_t block_halo;    /* halo vertices summed over blocks */
} eik_mesh_info_t;

int eik_mesh_info(eik_mesh_t mesh, eik_mesh_info_t* out);
int eik_mesh_get_vertices(eik_mesh_t mesh, float* xyz_out);       /* 3*nv */
int eik_mesh_get_elements(eik_mesh_t mesh, int32_t* elems_out);   /* 4*nt or 3*nf, 0-based */

/* HIP stream (hipStream_t) on which the handle's device work runs; creates it on first use. */
int eik_mesh_stream(eik_mesh_t mesh, void** stream_out);

/* ---------------------------------------------------------------------------------------- */
/* Speeds and metrics (Eikonal::initSpeedMtxMultipliers, src/Eikonal.cu:62-67)              */
/* ---------------------------------------------------------------------------------------- */

/* Reference semantics.  Tet meshes (TetMesh::init, tetmesh.cu:54-79): count == 6*nt sets a
 * symmetric tensor per tet stored [0 1 2; 1 3 4; 2 4 5]; count == nt sets M = v*I; any other
 * count (including 0) sets M = I.  Travel time along an edge e is sqrt(e^T M e), so a scalar v
 * means 1/speed^2.  Triangle meshes (TriMesh::InitializeAttributes, TriMesh_stats.cu:56-69):
 * count == nf gives speedInv = 1/v per face; any other count gives speedInv = 1. */
int eik_mesh_set_speed(eik_mesh_t mesh, const float* values, int64_t count);

/* Anisotropic triangle metric (no reference counterpart: the reference's 2D kernel is
 * isotropic only, src/core2d/cuda/CUDA_kernels.cu:129-356).  Three values per face
 * {m11, m12, m22}: a 2x2 SPD tensor in the face frame t1 = unit(v1 - v0),
 * t2 = unit(n x t1), n = unit((v1 - v0) x (v2 - v0)). */
int eik_mesh_set_metric_tri(eik_mesh_t mesh, const float* m3, int64_t nf);

/* ---------------------------------------------------------------------------------------- */
/* Synthetic meshes (the measured workloads)                                                 */
/* ---------------------------------------------------------------------------------------- */

/* Unit square as an n x n grid, each cell split along its (i,j)-(i+1,j+1) diagonal:
 * (n+1)^2 vertices, id j*(n+1)+i at (i/n, j/n, 0), 2n^2 triangles. */
int eik_gen_square(int32_t n, eik_mesh_t* out);

/* Unit cube as an n^3 grid of cells, each split into 6 Kuhn tets along (0,0,0)-(1,1,1):
 * (n+1)^3 vertices, id (k*(n+1)+j)*(n+1)+i, 6n^3 tets.  Interior vertices are jittered by
 * uniform(-jitter*h, +jitter*h) per coordinate (h = 1/n) from a splitmix64 stream seeded with
 * `seed`.  Fails with EIK_ERR_ARG if a tet would not keep a positive volume. */
int eik_gen_kuhn_cube(int32_t n, float jitter, uint64_t seed, eik_mesh_t* out);

/* Icosahedron subdivided `level` times (midpoint split), projected to the unit sphere:
 * 10*4^level + 2 vertices, 20*4^level faces. */
int eik_gen_icosphere(int32_t level, eik_mesh_t* out);

/* ---------------------------------------------------------------------------------------- */
/* Block partition (the reference's METIS / structured partitions)                          */
/* ---------------------------------------------------------------------------------------- */

/* Partition into ceil(nv / verts_per_block) blocks (tets; at least 2) or floor(nv /
 * verts_per_block) blocks (triangles; at least 2), the block counts of
 * meshFIM3dEikonal::GraphPartition_METIS2 (meshFIM3dEikonal.cu:153-237) and
 * meshFIM2dEikonal::GraphPartition_METIS2 (meshFIM2d_eikonal.cu:111-204).  METIS k-way is
 * replaced by recursive coordinate bisection with balanced block sizes.  *nblocks_out may be
 * NULL.  Every partition call also builds the block layout on the host (the reference's
 * PartitionTets / PartitionFaces, meshFIM3dEikonal.cu:568-648), so eik_mesh_info reports the
 * partition's quality at once. */
int eik_partition(eik_mesh_t mesh, int32_t verts_per_block, int32_t* nblocks_out);

/* Structured tiling of a grid whose vertex id is (k*W + i)*L + j:
 * label = (k/bd)*nbl*nbw + (i/bw)*nbl + j/bl with nbl = L/bl, nbw = W/bw
 * (meshFIM3dEikonal::GraphPartition_Square, meshFIM3dEikonal.cu:239-283; pass D = bd = 1 for
 * the 2D form, meshFIM2d_eikonal.cu:206-252). */
int eik_partition_structured(eik_mesh_t mesh, int32_t L, int32_t W, int32_t D, int32_t bl,
                             int32_t bw, int32_t bd, int32_t* nblocks_out);

/* Caller-provided labels (e.g. from an external METIS run); labels in [0, nblocks). */
int eik_partition_set_labels(eik_mesh_t mesh, const int32_t* labels, int32_t nblocks);

/* Current labels (nv entries). */
int eik_partition_labels(eik_mesh_t mesh, int32_t* labels_out);

/* ---------------------------------------------------------------------------------------- */
/* Solve (Eikonal::solveEikonal, src/Eikonal.cu:135-231; meshFIM3dEikonal::GenerateData,     */
/* meshFIM3dEikonal.cu:285-566; meshFIM2dEikonal::GenerateData, meshFIM2d_eikonal.cu:349-864) */
/* ---------------------------------------------------------------------------------------- */

#define EIK_SCHED_SYNC 0   /* one launch per outer iteration, host reads the active count */
#define EIK_SCHED_ASYNC 1  /* one persistent launch with a device work queue */

/* Element records of a block visit: copied to LDS once per visit, or read from memory (L1/L2)
 * at every use with smaller workgroups, more of them per CU.  AUTO stages the persistent single
 * solve (one solve's wavefront is latency-bound: fewer, faster visits win) and reads from
 * memory in the one-launch-per-iteration solve and in multi-solve launches (batches, serving:
 * throughput-bound, more resident workgroups win). */
#define EIK_STAGE_AUTO 0
#define EIK_STAGE_LDS 1
#define EIK_STAGE_MEMORY 2

typedef struct {
  int32_t scheduler;       /* EIK_SCHED_* (default EIK_SCHED_ASYNC) */
  int32_t max_iterations;  /* sync: outer iterations cap (reference maxIterations_, default 100
                              at src/Eikonal.cu:19); <= 0 means unbounded */
  int32_t max_inner;       /* inner sweeps per block visit before the block is re-queued;
                              <= 0 (the default) chooses per element kind: 2 for tets, 4 for
                              triangles, 3 for triangles in the async single solve (the
                              reference runs 2 in 3D, 7 in 2D).  The async scheduler runs one
                              sweep per tet visit while few blocks are pending and then re-runs
                              an unconverged block at once in the same workgroup (stats.reruns) */
  float eps;               /* relative: a decrease by more than eps * T marks a vertex changed
                              (its elements are re-solved) and keeps its block unconverged;
                              default 0, the exact fixed point (the reference stops at an
                              absolute 1e-5, CUDADefines.h:20) */
  int32_t record_history;  /* sync only: keep T after every outer iteration */
  int32_t workgroups;      /* async: persistent workgroups (0 = auto, sized to the fronts:
                              about 5 x (blocks^2 x seeds)^(1/3) for tets, 3 x sqrt(blocks x
                              seeds) for triangles, at least one per CU, at most all that fit) */
  int32_t stage_records;   /* EIK_STAGE_*: where a block visit reads its element records */
  int32_t wait_limit_us;   /* async / batch / served: longest wait of a persistent workgroup
                              for queued work, in microseconds (0 = 10 s, the default; < 0 is
                              invalid).  A workgroup that waits longer exits; if the queue then
                              never reaches its end the solve fails with EIK_ERR_TIMEOUT, so a
                              solve cannot hang the caller.  Normal waits last microseconds. */
} eik_solve_opts_t;

void eik_solve_opts_default(eik_solve_opts_t* opts);

typedef struct {
  double kernel_ms;        /* device time of the solve (HIP events on the handle's stream) */
  double wall_ms;          /* host time of the solve call (excludes mesh upload) */
  int64_t outer_iterations;/* sync: outer iterations; async: 1 */
  int64_t block_visits;    /* block visits (the reference's "total iteration number",
                              meshFIM3dEikonal.cu:428) */
  int64_t vertex_updates;  /* owned vertices summed over block visits */
  int64_t inner_sweeps;    /* inner sweeps summed over block visits */
  int64_t local_solves;    /* element-corner local solves evaluated */
  int64_t algo_bytes;      /* algorithmic HBM bytes summed over block visits: per visit 32 B per
                              tet (20 per triangle) + 4 B per owned and halo time read + 4 B per
                              owned time written + 4 */
  int64_t history_len;     /* iterations stored (record_history) */
  double visit_ms;         /* device time of the block-visit kernel launches (HIP events on the
                              handle's stream around each launch) */
  int64_t visit_launches;  /* block-visit kernel launches */
  double busy_ms;          /* time inside block visits, summed over workgroups */
  double wait_ms;          /* async: time workgroups waited for work, summed */
  double load_ms;          /* parts of busy_ms: staging loads, sweeps, write-back */
  double sweep_ms;
  double writeback_ms;
  double queue_ms;         /* async: queueing dependents and releasing blocks, summed */
  int64_t reruns;          /* async: visits that directly followed the same workgroup's visit of
                              the same block (no queue round trip; records still staged) */
  int64_t abandoned_waits; /* async / batch: queue waits that outlasted wait_limit_us (0 in a
                              healthy solve) */
  int64_t algo_bytes_loaded; /* algo_bytes without the element term of the re-runs whose element
                                records were still staged in LDS from the visit before (they
                                read none from memory); equal to algo_bytes when records are
                                read from memory at every use */
} eik_solve_stats_t;

/* Solve with `nseeds` source vertices.  seed_vals may be NULL (all sources at 0, the
 * reference's only mode: Eikonal.cu:145-158, 187-201); non-NULL values must be >= 0 and are
 * start times: the result is the fixed point with T <= start time at every source, so a source
 * reached earlier from another source takes the earlier time, and a vertex listed twice keeps
 * its smaller start time.
 * T_out (nv floats, may be NULL to leave the result on the device) receives the arrival times in
 * the caller's vertex order.  Builds and uploads the block layout on first use after a
 * partition or speed change. */
int eik_solve(eik_mesh_t mesh, const int32_t* seeds, int64_t nseeds, const float* seed_vals,
              const eik_solve_opts_t* opts, float* T_out, eik_solve_stats_t* stats);

/* The solve of eik_solve (persistent scheduler, default options, sources at 0) for callers that
 * cannot allocate device memory: code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc.  No counterpart in the reference (its solve allocates on every call,
 * meshFIM3dEikonal.cu:368-389).  Allocates nothing and does not synchronise: the solve's state
 * (times, the seen values, the work queue, its flags and counters) lives in a device array of the
 * library of 64 MiB per device, held until the process ends; a mesh whose state needs more fails
 * with EIK_ERR_CAPACITY (tet meshes up to about 3 M vertices fit at 64 vertices per block;
 * larger ones use eik_solve).
 * The mesh must be uploaded (eik_mesh_upload or a previous solve) on the current device.
 * seeds_dev: nseeds caller vertex ids in device memory (ids outside [0, nv) are skipped);
 * T_out_dev: nv floats of device memory that receive the times in the caller's vertex order when
 * the stream reaches the end of the work enqueued here.  stream: a hipStream_t, or NULL for the
 * mesh's stream.  Every call on a device shares that device's array: calls that overlap on one
 * device (different streams, or host threads) race and must be serialised by the caller.  The
 * call returns before the solve ends, so it reports no statistics and cannot report a queue
 * wait that outlasted the default wait limit (10 s; the result is then incomplete). */
int eik_solve_noalloc(eik_mesh_t mesh, const int32_t* seeds_dev, int64_t nseeds, float* T_out_dev,
                      void* stream);

/* Result of the last solve (nv floats, caller's vertex order). */
int eik_get_result(eik_mesh_t mesh, float* T_out);

/* With record_history: T after outer iteration `i` (0-based), as the reference's
 * getResultAtIteration (src/Eikonal.cu:42-44). */
int eik_get_history(eik_mesh_t mesh, int64_t i, float* T_out);

/* Build and upload the block layout now (otherwise done by the first solve). */
int eik_mesh_upload(eik_mesh_t mesh);

/* ---------------------------------------------------------------------------------------- */
/* Diagnostics                                                                               */
/* ---------------------------------------------------------------------------------------- */

/* Virtual faces of a triangle mesh: rows (obtuse vertex, a, b) and the originating face.
 * cap is the row capacity; *n_out receives the count. */
int eik_mesh_virtual_faces(eik_mesh_t mesh, int32_t* rows_out, int32_t* origin_out, int64_t cap,
                           int64_t* n_out);

/* Read-only view of the block layout the solve uploads (built on the host if needed; no device
 * work).  Pointers stay valid until the next partition or speed change of the mesh.
 * blocks: 8 int32 per block {vstart, nown, nbnd0, hstart, nhalo, estart, nelem, 0};
 * elems: nelems records of rec_bytes bytes: tets {uint16 v[4]; float L[6]} with L the squared
 * metric edge lengths {01,02,03,12,13,23}; triangles {uint16 v[3]; uint16 mask; float L[3]}
 * with L = {01,12,20}; v are block-local vertex indices (< nown: owned vertex vstart + v, else
 * halo[hstart + v - nown]).  Vertex ids in the layout are the block-contiguous numbering:
 * perm[new] = caller id, iperm[caller] = new.  sup: blocks owning a block's halo vertices;
 * dep: the transpose.  No counterpart in the reference (its layout is internal to GenerateData,
 * meshFIM3dEikonal.cu:680-923); for testing the layout against the oracle. */
typedef struct {
  int32_t dim;
  int32_t nblocks;
  int64_t nv;
  int64_t nelems;
  int64_t nhalo;
  int32_t rec_bytes;
  int32_t max_local;
  const int32_t* perm;
  const int32_t* iperm;
  const int32_t* blocks;
  const int32_t* halo;
  const void* elems;
  const int32_t* sup_start;
  const int32_t* sup;
  const int32_t* dep_start;
  const int32_t* dep;
} eik_layout_view_t;

int eik_mesh_layout(eik_mesh_t mesh, eik_layout_view_t* out);

/* Run the device local solver on n independent elements (tets: T 4n, L 6n squared edge
 * lengths {01,02,03,12,13,23}; triangles: T 3n, L 3n {01,12,20}); out receives, per corner,
 * min(T_corner, local solve from the other corners).  For testing the kernel arithmetic. */
int eik_debug_local_solve(int32_t dim, const float* T, const float* L, int64_t n, float* out);

/* ---------------------------------------------------------------------------------------- */
/* Batched solves (no reference counterpart: the reference solves one mesh per process,       */
/* Eikonal::solveEikonal, src/Eikonal.cu:135-231, with a host loop per outer iteration)       */
/* ---------------------------------------------------------------------------------------- */
typedef struct eik_batch* eik_batch_t;

/* Groups partitioned meshes (tets and triangles may be mixed; solves x blocks < 2^31) so that one
 * persistent launch solves all of them.  Layouts missing on the host are built in parallel
 * threads; each mesh's device layout is uploaded (and stays usable by eik_solve).  The meshes
 * stay owned by the caller and must outlive the batch; re-partitioning a mesh invalidates the
 * batch.  A mesh may appear several times: every solve has its own time arrays and, with
 * eik_batch_set_factors, its own speed field.  EIK_ERR_CAPACITY if a block's times and work
 * queue do not fit the LDS; blocks whose element records do not fit it as well are solved with
 * the records read from memory (stage_records EIK_STAGE_LDS then fails with EIK_ERR_CAPACITY). */
int eik_batch_create(const eik_mesh_t* meshes, int32_t n, eik_batch_t* out);
int eik_batch_destroy(eik_batch_t batch);
int eik_batch_info(eik_batch_t batch, int64_t* total_verts, int64_t* total_blocks);

/* Solves every mesh of the batch in one launch.  Seeds of mesh i are seeds[seed_offsets[i] ..
 * seed_offsets[i+1]) (caller vertex ids; seed_offsets has n+1 entries; a mesh may have none,
 * its times then stay LARGENUM); seed_vals as in eik_solve (NULL = 0).  T_out (may be NULL)
 * receives the times of mesh 0, then mesh 1, ... (sum of nv floats); with NULL the times stay
 * on the device (eik_batch_get_result copies one solve's).  opts: max_inner, eps, max_iterations, workgroups as in eik_solve
 * (the persistent scheduler always).  The launch sequence is captured into a HIP graph on the
 * first call and replayed (eik_batch_set_graph(batch, 0) runs it directly instead). */
int eik_batch_solve(eik_batch_t batch, const int64_t* seed_offsets, const int32_t* seeds,
                    const float* seed_vals, const eik_solve_opts_t* opts, float* T_out,
                    eik_solve_stats_t* stats);
int eik_batch_set_graph(eik_batch_t batch, int32_t enable);

/* Per-solve speed field over the mesh's own metric: factors[e] (one per element, > 0)
 * multiplies every squared edge length of element e, i.e. the metric of e (tets: the
 * reference's scalar per tet, 1/speed^2; triangles: 1/speed^2 = speedInv^2).  NULL clears it.
 * Lets several solves of one resident mesh carry their own speeds.  The factors are read before
 * the call returns and checked on the device: a value that is not finite and > 0 gives
 * EIK_ERR_ARG and leaves the solve's current factors.  The solve keeps its device buffers, so
 * replacing its factors for every solve allocates nothing after the first call. */
int eik_batch_set_factors(eik_batch_t batch, int32_t i, const float* factors, int64_t count);

/* Times of solve i of the last eik_batch_solve (nv floats of that mesh, caller's vertex order),
 * without copying the other solves' (no counterpart in the reference, which solves one mesh at a
 * time; the batch's analogue of Eikonal::getFinalResult, src/Eikonal.cu:38-40).  EIK_ERR_STATE
 * before a batch solve has finished. */
int eik_batch_get_result(eik_batch_t batch, int32_t i, float* T_out);

/* Autotuner for batched solves (no reference counterpart: the reference's block size is the
 * user option maxVertsPerBlock_, src/Eikonal.h:47, and its inner sweep count a constant,
 * src/core3d/include/CUDADefines.h).  For every verts_per_block candidate the meshes are
 * re-partitioned (several at a time) and batched, and for every max_inner candidate one
 * warm-up and `reps` timed batched solves run from the given seeds (offsets as in
 * eik_batch_solve); the median device time of each pair goes to times[i*ninner + j] (may be
 * NULL).  The meshes end partitioned with the fastest block size. */
typedef struct {
  int32_t verts_per_block;
  int32_t max_inner;
  double ms;  /* median batched-solve time of the winner */
} eik_tune_result_t;
int eik_batch_autotune(const eik_mesh_t* meshes, int32_t n, const int64_t* seed_offsets,
                       const int32_t* seeds, const int32_t* verts_per_block, int32_t nvpb,
                       const int32_t* max_inner, int32_t ninner, int32_t reps,
                       eik_tune_result_t* best, double* times);

/* ---------------------------------------------------------------------------------------- */
/* Serving (no reference counterpart): solves arriving one at a time over resident meshes     */
/* ---------------------------------------------------------------------------------------- */
typedef struct eik_server* eik_server_t;

typedef struct {
  int32_t max_solves_per_launch; /* requests solved together by one launch (default 32) */
  int32_t launches;              /* launch sequences in flight at once, one stream each (4) */
  int64_t budget_bytes;          /* device memory for in-flight request state (4 GiB); a request
                                    needs 4 * (2 nv + block-local vertices + elements if it has
                                    factors) bytes; requests wait while the budget is full */
  int32_t max_seeds;             /* seeds per request (64) */
  int32_t max_inner;             /* inner sweeps per block visit (1) */
} eik_server_opts_t;

typedef struct {
  int32_t status;   /* 0 waiting, 1 running, 2 done (negative: failed with that EIK_ERR code) */
  int32_t launch;   /* launch sequence that ran it */
  double submit_ms; /* host clock, ms since the server was created */
  double start_ms;  /* admitted into a launch */
  double end_ms;    /* launch observed complete, results copied */
} eik_request_info_t;

void eik_server_opts_default(eik_server_opts_t* opts);

/* Starts a serving loop (a worker thread and its launch resources) over partitioned meshes,
 * which stay resident in device memory and must outlive the server. */
int eik_server_create(const eik_mesh_t* meshes, int32_t n, const eik_server_opts_t* opts,
                      eik_server_t* out);

/* Queues one solve of mesh `mesh` from nseeds sources (caller vertex ids, value 0).  factors
 * (NULL or one per element, as eik_batch_set_factors) is the request's own speed field; T_out
 * (NULL or nv floats) receives the times.  Both stay owned by the caller and must stay valid
 * until eik_server_wait returns for the request.  Thread-safe. */
int eik_server_submit(eik_server_t server, int32_t mesh, const int32_t* seeds, int32_t nseeds,
                      const float* factors, float* T_out, int64_t* request_id);

/* State of a request (does not block). */
int eik_server_poll(eik_server_t server, int64_t request_id, eik_request_info_t* info);

/* Blocks until the request is done, fills info, and forgets the request. */
int eik_server_wait(eik_server_t server, int64_t request_id, eik_request_info_t* info);

/* Launches started so far. */
int eik_server_launches(eik_server_t server, int64_t* started);

/* Finishes the queued and running requests, stops the worker and frees the server. */
int eik_server_destroy(eik_server_t server);

#ifdef __cplusplus
}
#endif

#endif /* EIKONAL_AMD_H */
