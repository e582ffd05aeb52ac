"""Benchmark: 3D FIM on the 64^3 jittered Kuhn-tet cube (274,625 vertices, 1,572,864 tets),
per-tet log-normal(0, 0.5) speed, one source at the centre vertex, 64-vertex blocks, fp32.

One step = one complete solve of that workload with the mesh layout resident in HBM (seeding,
the persistent block-visit kernel to convergence; no host<->device copy of the mesh).
Metric: vertex-updates/s, a vertex-update being one owned vertex of a block processed in one
block visit (inner LDS sweeps are not counted), summed over the K timed solves.

Prints one JSON line.  `roofline` is the block-visit kernel's algorithmic bytes (per visit: 32 B
per tet of element data, 4 B per owned and halo time read, 4 B per owned time written, 4 B of
flag; counted on the device) divided by its duration measured with HIP events on the mesh's
stream; `traffic` is the kernel's measured HBM bytes per launch.
`cpu_baseline` is the oracle's blocked FIM in FP32 (oracle/fim_oracle.c) over the same block
layout on the host's cores (OpenMP; OMP_NUM_THREADS when set), run on one full solve of the same
mesh, with the same in FP64 and the serial vertex-level FP64 oracle on one core beside it.
"""
import argparse
import collections
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

# HBM bytes per launch of the block-visit kernels, measured with rocprofv3 (FETCH_SIZE and
# WRITE_SIZE passes, each scaled by the reading of a plain 256 MiB copy), are read from the
# profile summaries in profiles/: each names its workload, the bytes per launch and the hash of the
# kernel sources it was measured on.  A summary whose hash differs from the current sources is
# stale and gives no figure (traffic null).
PROFILES = {
    "fim3d_kuhn64_lognormal": "k_visit_async_fim3d_kuhn64_lognormal.md",
    "batch256_kuhn16to48_twolayer": "k_visit_multi_batch256_kuhn16to48.md",
    "ico7_aniso_4src": "k_visit_async_ico7_aniso.md",
}


def kernel_sources_hash():
    """sha256 over the device and ABI sources (sorted by name), the build a profile describes."""
    import glob
    import hashlib
    h = hashlib.sha256()
    files = sorted(glob.glob(os.path.join(ROOT, "sci-solver_eikonal_amd", "csrc", "*.hip")) +
                   glob.glob(os.path.join(ROOT, "sci-solver_eikonal_amd", "csrc", "*.h")) +
                   [os.path.join(ROOT, "include", "eikonal_amd.h")])
    for f in files:
        h.update(os.path.relpath(f, ROOT).encode())
        h.update(open(f, "rb").read())
    return h.hexdigest()[:16]


def measured_traffic(workload):
    """(bytes per launch, note) from the workload's profile summary; bytes is None when there is
    no summary or it was measured on other kernel sources."""
    import re
    name = PROFILES.get(workload)
    path = os.path.join(ROOT, "profiles", name) if name else None
    if not path or not os.path.exists(path):
        return None, "no profile"
    text = open(path).read()
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", text)
    hs = re.search(r"kernel_sources_hash:\s*([0-9a-f]+)", text)
    if not m or not hs:
        return None, "profile without a traffic figure"
    if hs.group(1) != kernel_sources_hash():
        return None, "profile measured on other kernel sources (%s)" % hs.group(1)
    return float(m.group(1)), "profiles/" + name


def progress(msg):
    """One line per finished step on stderr (stdout carries only the JSON line)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


def load_pkg():
    spec = importlib.util.spec_from_file_location(
        "eikonal_amd", os.path.join(ROOT, "sci-solver_eikonal_amd", "eikonal_amd.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def make_workload(E, n, vpb, seed):
    m = E.Mesh.kuhn_cube(n, jitter=0.2, seed=seed)
    nt = m.info()["nelem"]
    rng = np.random.default_rng(seed)
    speed = np.exp(rng.normal(0.0, 0.5, nt))
    mult = (1.0 / speed ** 2).astype(np.float32)
    m.set_speed(mult)
    m.partition(vpb)
    c = ((n // 2) * (n + 1) + n // 2) * (n + 1) + n // 2
    return m, mult, c


def aniso_error(m, M3, seeds, T):
    """The level-7 anisotropic icosphere against the serial FP64 oracle (the anisotropic triangle solve over the same faces
    and virtual faces); floor: a quarter of the shortest metric edge."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    xyz, f = m.vertices(), m.elements()
    M = M3.astype(np.float64)
    G = O.tri_gram(xyz, f, metric=M)
    vf, origin = O.tri_virtual_faces(xyz, f)
    Gv = O.tri_gram(xyz, vf, metric=M[origin], frame_faces=f[origin]) if len(vf) else None
    ref, _ = O.fim2d(xyz.astype(np.float64), f, seeds, G, 1, vfaces=vf, vparam=Gv)
    tau = 0.25 * float(np.sqrt(G[:, 0].min()))
    return float((np.abs(T.astype(np.float64) - ref) / np.maximum(ref, tau)).max())


def cpu_threads():
    """Host threads for the multi-core CPU baselines: OMP_NUM_THREADS when set (the machine's
    CPU share), else every core this process may use."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return max(1, len(os.sched_getaffinity(0)))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(m, mult, c, max_inner):
    """CPU baselines of the same solve, timed on the host, all over the product's own 64-vertex
    block layout or the same mesh:
      (i) the reported value: the blocked FIM in FP32 (the oracle's restatement of the
          reference's local solver, evaluated at the reference's arithmetic width), its active
          blocks visited concurrently on cpu_threads() threads (OpenMP);
      (ii) the same blocked FIM in FP64 on the same threads;
      (iii) the serial FP64 vertex-level oracle on one core (reference formulas, SMALLNUM
          regularisation as shipped).
    Also returns the serial regularised and unregularised oracle solutions for the error
    figures."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    lay = m.layout()
    nthr = cpu_threads()
    xyz = m.vertices().astype(np.float64)
    tets = m.elements()
    M6 = O.tet_metric(tets.shape[0], mult.astype(np.float64))
    t0 = time.perf_counter()
    T32, it32, visits32, updates32 = O.fim_blocked(lay, [c], max_inner=max_inner, threads=nthr,
                                                   counts=True, fp32=True)
    dt32 = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, it_b, visits_b, updates_b = O.fim_blocked(lay, [c], max_inner=max_inner, threads=nthr,
                                                 counts=True)
    dt_b = time.perf_counter() - t0
    t0 = time.perf_counter()
    T, st = O.fim3d(xyz, tets, [c], M6=M6)
    dt = time.perf_counter() - t0
    with O.exact():
        Te, _ = O.fim3d(xyz, tets, [c], M6=M6)
    tau = 0.25 * float(np.sqrt(mult.min())) / round(len(xyz) ** (1 / 3) - 1)
    err32 = float((np.abs(T32.astype(np.float64) - Te) / np.maximum(Te, tau)).max())
    return T, Te, {"value": updates32 / dt32, "unit": "vertex-updates/s", "cores": nthr,
                   "kind": "port", "nproc": os.cpu_count(), "cpu_model": cpu_model(),
                   "solve_ms": dt32 * 1e3, "outer_iterations": it32, "block_visits": visits32,
                   "max_rel_err_vs_exact_oracle": err32,
                   "sample": "one full solve of the same 64^3 workload by the blocked FIM in FP32 "
                             "(the oracle's restatement of the reference's local solver at the "
                             "reference's arithmetic width) over the same 64-vertex block layout "
                             "(outer iterations of active blocks, %d inner sweeps), the blocks of "
                             "each iteration visited on %d host threads (OpenMP)"
                             % (max_inner, nthr),
                   "fp64_blocked": {"value": updates_b / dt_b, "unit": "vertex-updates/s",
                                    "cores": nthr, "solve_ms": dt_b * 1e3,
                                    "outer_iterations": it_b, "block_visits": visits_b},
                   "serial": {"value": st["updates"] / dt, "unit": "vertex-updates/s",
                              "cores": 1, "solve_ms": dt * 1e3,
                              "sample": "one full solve by the serial FP64 vertex-level "
                                        "active-list oracle (a vertex-update = one vertex "
                                        "solve)"}}


def roofline_with_traffic(achieved, algo_bytes, visit_ms, workload, profiled_config):
    """The roofline object of a leg: algorithmic bytes over the visit kernel's time, and the
    measured traffic per launch when a current profile of this configuration exists."""
    traffic, src = measured_traffic(workload) if profiled_config else (None, "other config")
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "algo_bytes": algo_bytes, "traffic": traffic,
            "traffic_source": src,
            "traffic_frac": (traffic / (visit_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if traffic else None),
            "traffic_over_algorithmic": traffic / algo_bytes if traffic else None}


def history_leg(m, c, max_inner, reps=5):
    """The same solve as the reference's driver runs it: one launch per outer iteration
    (scheduler 0), every iteration's time field kept as a snapshot in device memory
    (record_history); against the persistent scheduler's solve with its result copied out, which
    is what the C++ facade runs when only the final field is read."""
    import torch
    def timed(**kw):
        ts = []
        for r in range(reps + 1):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, st = m.solve([c], out=True, max_inner=max_inner, **kw)
            dt = time.perf_counter() - t0
            if r:
                ts.append(dt * 1e3)
        return float(np.median(ts)), st
    t_async, _ = timed(scheduler=1)
    t_sync, st = timed(scheduler=0, record_history=1)
    t_plain, st0 = timed(scheduler=0)
    return {"async_ms": t_async, "sync_history_ms": t_sync, "sync_ms": t_plain,
            "outer_iterations": st["outer_iterations"], "history_len": st["history_len"],
            "sync_history_vs_async": t_sync / t_async,
            "note": "wall time per solve including the copy of the final field; the snapshots "
                    "stay in device memory until read"}


def make_batch_workload(E, count, seed):
    """256-cube batch: Kuhn cubes with n ~ U{16..48} cells per side (jitter 0.2 h), speed 1 below
    a random plane z = z0 ~ U(0, 1) and one U[0.2, 5] speed above it (per tet, by centroid),
    1-8 random sources per cube."""
    rng = np.random.default_rng(seed)
    meshes, seeds = [], []
    for i in range(count):
        n = int(rng.integers(16, 49))
        m = E.Mesh.kuhn_cube(n, jitter=0.2, seed=seed + 1 + i)
        xyz, t = m.vertices(), m.elements()
        zc = xyz[:, 2][t].mean(axis=1)
        z0, above = rng.uniform(0.0, 1.0), rng.uniform(0.2, 5.0)
        speed = np.where(zc < z0, 1.0, above)
        m.set_speed((1.0 / speed ** 2).astype(np.float32))
        seeds.append(rng.choice(m.nv, int(rng.integers(1, 9)), replace=False).astype(np.int32))
        meshes.append(m)
    return meshes, seeds


def batch_leg(E, count, seed, steps, tune_sample, fixed=None, plain_baseline=True):
    """Batched solves/s (one persistent launch per batch, graph replay) against back-to-back
    plain solves of the same meshes; block size and inner sweep cap autotuned on a sample
    (or fixed = (verts_per_block, max_inner))."""
    t0 = time.perf_counter()
    meshes, seeds = make_batch_workload(E, count, seed)
    t_gen = time.perf_counter() - t0
    k = min(tune_sample, count)
    if fixed is None:
        best, table = E.autotune(meshes[:k], seeds[:k], vpb=(32, 64, 128, 192), inner=(1, 2, 4),
                                 reps=2)
        rest = meshes[k:]
    else:
        best, table, k = {"verts_per_block": fixed[0], "max_inner": fixed[1]}, {}, 0
        rest = meshes
    t_tune = time.perf_counter() - t0 - t_gen
    progress("batch workload generated and tuned (%d/%d)" % (best["verts_per_block"],
                                                              best["max_inner"]))
    for m in rest:
        m.partition(best["verts_per_block"])
    b = E.Batch(meshes)
    info = b.info()
    opts = dict(max_inner=best["max_inner"])
    b.solve(seeds, out=False, **opts)  # warm-up (captures the graph)
    times = []
    for _ in range(steps):
        t1 = time.perf_counter()
        _, st = b.solve(seeds, out=False, **opts)
        times.append(time.perf_counter() - t1)
    Tb, st = b.solve(seeds, **opts)
    t_batch = float(np.median(times))
    b.close()
    kernel_ms = st["kernel_ms"]
    achieved = st["algo_bytes"] / (kernel_ms * 1e-3) / 1e9
    res = {
        "workload": "batch%d_kuhn16to48_twolayer" % count, "solves": count,
        "verts": int(info["total_verts"]), "blocks": int(info["total_blocks"]),
        "tets": int(sum(m.info()["nelem"] for m in meshes)),
        "autotuned": {"verts_per_block": best["verts_per_block"],
                      "max_inner": best["max_inner"], "sample_solves": k,
                      "candidates_ms": {"%d/%d" % kv: round(v, 3) for kv, v in table.items()}},
        "batch_ms": t_batch * 1e3, "batch_kernel_ms": kernel_ms,
        "solves_per_s": count / t_batch,
        "vertex_updates_per_s": st["vertex_updates"] / t_batch,
        "block_visits": st["block_visits"], "local_solves": st["local_solves"],
        "roofline": roofline_with_traffic(achieved, st["algo_bytes"], st["visit_ms"],
                                          "batch%d_kuhn16to48_twolayer" % count,
                                          (best["verts_per_block"], best["max_inner"]) == (192, 1)),
        "setup_s": {"generate": t_gen, "autotune": t_tune},
    }
    if not plain_baseline:
        return res

    def plain(popts):
        # back-to-back plain solves (one persistent launch and one host synchronisation each)
        for m, s in zip(meshes, seeds):  # first solves allocate each mesh's queue
            m.solve(s, out=False, scheduler=1, **popts)
        t1 = time.perf_counter()
        for m, s in zip(meshes, seeds):
            m.solve(s, out=False, scheduler=1, **popts)
        return time.perf_counter() - t1

    # plain solves at the batch's block size and sweep cap, and at the library defaults (64
    # vertices per block); the faster of the two is the baseline
    progress("batched solves timed")
    t_plain_tuned = plain(opts)
    progress("plain solves timed")
    worst = 0.0
    for i in range(0, count, max(1, count // 16)):
        alone, _ = meshes[i].solve(seeds[i], scheduler=1, **opts)
        worst = max(worst, float(np.max(np.abs(Tb[i] - alone) / np.maximum(alone, 1e-3))))
    t_plain_default = t_plain_tuned
    if best["verts_per_block"] != 64 or best["max_inner"] != 2:
        for m in meshes:
            m.partition(64)
        t_plain_default = plain({})
    t_plain = min(t_plain_tuned, t_plain_default)
    res.update({
        "plain_ms": t_plain * 1e3, "plain_solves_per_s": count / t_plain,
        "plain_ms_at_batch_config": t_plain_tuned * 1e3,
        "plain_ms_at_defaults": t_plain_default * 1e3,
        "speedup_vs_plain": t_plain / t_batch,
        "max_rel_diff_batched_vs_alone": worst,
    })
    return res


def face_metrics(nf, seed):
    """Per-face SPD 2x2 metric in the face's tangent frame: rotation theta ~ U[0, pi),
    eigenvalues 1 and kappa ~ U[1, 10], folded as (m11, m12, m22)."""
    rng = np.random.default_rng(seed)
    th = rng.uniform(0.0, np.pi, nf)
    k = rng.uniform(1.0, 10.0, nf)
    c, s = np.cos(th), np.sin(th)
    return np.stack([c * c + k * s * s, (1.0 - k) * c * s, s * s + k * c * c],
                    axis=1).astype(np.float32)


def square_leg(E, steps):
    """The 2D plain case: the unit square as a 32x32 grid split along the cell diagonals (1,089
    vertices, 2,048 triangles), speed 1, one source at vertex 0 = (0, 0), default blocks: ms per
    solve and the error against the analytic distance sqrt(x^2 + y^2)."""
    m = E.Mesh.square(32)
    m.partition(64)
    m.solve([0], out=False)
    times = []
    for _ in range(steps):
        t1 = time.perf_counter()
        T, st = m.solve([0])
        times.append(time.perf_counter() - t1)
    xyz = m.vertices().astype(np.float64)
    d = np.sqrt(xyz[:, 0] ** 2 + xyz[:, 1] ** 2)
    return {"workload": "square32_iso_corner", "verts": m.nv, "faces": m.info()["nelem"],
            "blocks": m.info()["nblocks"], "ms_per_solve": float(np.median(times)) * 1e3,
            "kernel_ms": st["kernel_ms"], "block_visits": st["block_visits"],
            "max_abs_err_vs_analytic": float(np.abs(T - d).max()),
            "rms_err_vs_analytic": float(np.sqrt(np.mean((T - d) ** 2)))}, (m, T)


def square_error(m, T):
    """The 2D plain case against the serial FP64 oracle (isotropic corner solve, speed 1)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    xyz, f = m.vertices().astype(np.float64), m.elements()
    vf, origin = O.tri_virtual_faces(m.vertices(), f)
    F = np.ones(f.shape[0])
    ref, _ = O.fim2d(xyz, f, [0], F, 0, vfaces=vf, vparam=F[origin] if len(vf) else None)
    return float((np.abs(T.astype(np.float64) - ref) / np.maximum(ref, 0.25 / 32)).max())


def aniso_leg(E, seed, steps):
    """Level-7 icosphere (163,842 vertices, 327,680 faces), a random SPD metric per face
    (condition number in [1, 10]), 4 random sources, 64-vertex blocks: ms per solve and
    vertex-updates/s; the error against the FP64 oracle is added by the CPU baseline section."""
    m = E.Mesh.icosphere(7)
    M3 = face_metrics(m.info()["nelem"], seed + 7)
    m.set_metric_tri(M3)
    m.partition(64)
    rng = np.random.default_rng(seed + 7)
    seeds = rng.choice(m.nv, 4, replace=False).astype(np.int32)
    m.solve(seeds, out=False)
    times, st = [], None
    for _ in range(steps):
        t1 = time.perf_counter()
        _, st = m.solve(seeds, out=False)
        times.append(time.perf_counter() - t1)
    T = m.result()
    t = float(np.median(times))
    achieved = st["algo_bytes"] / (st["visit_ms"] * 1e-3) / 1e9
    res = {"workload": "ico7_aniso_4src", "verts": m.nv, "faces": m.info()["nelem"],
           "ms_per_solve": t * 1e3, "kernel_ms": st["kernel_ms"], "visit_ms": st["visit_ms"],
           "vertex_updates_per_s": st["vertex_updates"] / t, "block_visits": st["block_visits"],
           "reruns": st["reruns"], "local_solves": st["local_solves"],
           "roofline": roofline_with_traffic(achieved, st["algo_bytes"], st["visit_ms"],
                                             "ico7_aniso_4src", True)}
    return res, (m, M3, seeds, T)


class PinnedPool:
    """float32 arrays in page-locked host memory allocated by the HIP runtime the library links
    (where a server receives request payloads), freed by close()."""

    def __init__(self):
        import ctypes
        self.c = ctypes
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipHostMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t,
                                           ctypes.c_uint]
        self.hip.hipHostFree.argtypes = [ctypes.c_void_p]
        self.ptrs = []

    def array(self, values):
        n = len(values)
        p = self.c.c_void_p()
        if self.hip.hipHostMalloc(self.c.byref(p), max(4 * n, 4), 0) != 0:
            raise MemoryError("hipHostMalloc failed")
        self.ptrs.append(p.value)
        a = np.ctypeslib.as_array((self.c.c_float * n).from_address(p.value))
        a[:] = values
        return a

    def close(self):
        for p in self.ptrs:
            self.hip.hipHostFree(p)
        self.ptrs = []


def serve_workload(E, nreq, rate, seed, fields_per_mesh=8, vpb=128, n_big=96, pool=None):
    """The served workload: three resident meshes (the level-7 icosphere, a 32^3 and an
    n_big^3 jittered Kuhn cube) partitioned at `vpb` vertices per block and uploaded once;
    `nreq` requests (mesh, 1-16 random sources, a log-normal(0, 0.5) speed field drawn from a
    pool of `fields_per_mesh` per mesh) arriving as a Poisson stream of mean `rate` per second
    (arrival times in s from the first).  With a PinnedPool the speed fields live in pinned host
    memory."""
    rng = np.random.default_rng(seed)
    meshes = [E.Mesh.icosphere(7), E.Mesh.kuhn_cube(32, jitter=0.2, seed=seed + 32),
              E.Mesh.kuhn_cube(n_big, jitter=0.2, seed=seed + n_big)]
    setup_ms = []  # admitting a mesh: partition, block layout and upload
    for m in meshes:
        t1 = time.perf_counter()
        m.partition(vpb)
        m.upload()
        setup_ms.append((time.perf_counter() - t1) * 1e3)
    fields = []
    for m in meshes:
        ne = m.info()["nelem"]
        fl = [(1.0 / np.exp(rng.normal(0.0, 0.5, ne)) ** 2).astype(np.float32)
              for _ in range(fields_per_mesh)]
        fields.append([pool.array(f) for f in fl] if pool is not None else fl)
    reqs = []
    for _ in range(nreq):
        k = int(rng.integers(0, 3))
        reqs.append((k, rng.choice(meshes[k].nv, int(rng.integers(1, 17)), replace=False)
                     .astype(np.int32), int(rng.integers(0, fields_per_mesh))))
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    arrivals -= arrivals[0]
    return meshes, fields, reqs, arrivals, setup_ms


def batch_serve(b, slot_mesh, reqs, arrivals, fields, keep=(), max_inner=1):
    """Serves requests (mesh k, sources, field index) arriving at `arrivals` (s) through one
    eik_batch whose solve slots belong to the meshes (slot_mesh[i] = k), one launch in flight at
    a time: whenever requests are waiting, a launch takes them in arrival order into the slots
    of their mesh (the slots bound the device memory in use; requests beyond a mesh's slots wait
    for the next launch), uploads each request's speed field into its slot and solves them
    together, and copies each request's times to the host (only its own: eik_batch_get_result);
    requests arriving meanwhile wait for the next launch.  Returns arrival-to-result latencies
    (s), the span from the first arrival to the last result (s), the launches, the results of
    the requests listed in `keep`, and the time spent uploading speed fields and in the launches
    with their result copies (s)."""
    nreq = len(reqs)
    slots = collections.defaultdict(list)
    for i, k in enumerate(slot_mesh):
        slots[k].append(i)
    lat = np.zeros(nreq)
    kept = {}
    keep = set(keep)
    waiting = collections.deque()
    nxt = done = launches = 0
    t_up = t_solve = 0.0
    t0 = time.perf_counter()
    while done < nreq:
        now = time.perf_counter() - t0
        while nxt < nreq and arrivals[nxt] <= now:
            waiting.append(nxt)
            nxt += 1
        if not waiting:
            time.sleep(max(0.0, min(0.0005, arrivals[nxt] - now)))
            continue
        used = collections.Counter()
        assign, rest = [], collections.deque()
        for r in waiting:
            k = reqs[r][0]
            if used[k] < len(slots[k]):
                assign.append((slots[k][used[k]], r))
                used[k] += 1
            else:
                rest.append(r)
        waiting = rest
        seeds = [np.zeros(0, np.int32) for _ in slot_mesh]
        t1 = time.perf_counter()
        for slot, r in assign:
            k, sd, fi = reqs[r]
            b.set_factors(slot, fields[k][fi])
            seeds[slot] = sd
        t2 = time.perf_counter()
        b.solve(seeds, out=False, max_inner=max_inner)
        for slot, r in assign:
            T = b.result(slot)
            lat[r] = time.perf_counter() - t0 - arrivals[r]
            if r in keep:
                kept[r] = T
        t_up += t2 - t1
        t_solve += time.perf_counter() - t2
        launches += 1
        done += len(assign)
    return lat, time.perf_counter() - t0, launches, kept, (t_up, t_solve)


def serve_batch_leg(E, nreq, rate, seed, slots=(16, 16, 8), vpb=192, max_inner=1,
                    plain_alt=(128, 2)):
    """The served workload through batch_serve (eik_batch, one launch in flight; the eik_server
    path with several launches in flight is off), against back-to-back launches of one request
    each (a one-slot batch per mesh, the request's speed field uploaded the same way; timed at
    the served block size and sweep cap and at `plain_alt`, the faster kept), and each of a
    sample of results in the mix against the same request solved alone."""
    pool = PinnedPool()
    meshes, fields, reqs, arrivals, setup_ms = serve_workload(E, nreq, rate, seed, vpb=vpb,
                                                              pool=pool)
    slot_mesh = [k for k in range(3) for _ in range(slots[k])]
    b = E.Batch([meshes[k] for k in slot_mesh])
    for k in range(3):  # warm-up: first launch of each kind, graph capture
        seeds = [np.zeros(0, np.int32) for _ in slot_mesh]
        seeds[slot_mesh.index(k)] = np.zeros(1, np.int32)
        b.solve(seeds, out=False, max_inner=max_inner)
    progress("serving meshes ready")
    sample = list(range(min(24, nreq)))
    lat, span, launches, kept, tt = batch_serve(b, slot_mesh, reqs, arrivals, fields, keep=sample,
                                                max_inner=max_inner)
    progress("poisson stream served (%d launches)" % launches)
    lat_s, span_s, launches_s, _, tt_s = batch_serve(b, slot_mesh, reqs, np.zeros(nreq), fields,
                                                     max_inner=max_inner)
    progress("saturated run served (%d launches)" % launches_s)
    b.close()
    nplain = min(nreq, 300)
    worst = 0.0
    t_plain = {}
    for cfg in ((vpb, max_inner), tuple(plain_alt)):
        if cfg in t_plain:
            continue
        if cfg[0] != vpb:
            for m in meshes:
                m.partition(cfg[0])
                m.upload()
        one = [E.Batch([m]) for m in meshes]
        for k in range(3):
            one[k].solve([np.zeros(1, np.int32)], out=False, max_inner=cfg[1])
        t1 = time.perf_counter()
        for k, sd, fi in reqs[:nplain]:
            one[k].set_factors(0, fields[k][fi])
            one[k].solve([sd], out=False, max_inner=cfg[1])
            one[k].result(0)
        t_plain[cfg] = time.perf_counter() - t1
        if cfg == (vpb, max_inner):
            for r in sample:
                k, sd, fi = reqs[r]
                one[k].set_factors(0, fields[k][fi])
                one[k].solve([sd], out=False, max_inner=max_inner)
                Ta = one[k].result(0)
                worst = max(worst, float(np.max(np.abs(kept[r] - Ta) / np.maximum(Ta, 1e-3))))
        for x in one:
            x.close()
    best_cfg = min(t_plain, key=t_plain.get)
    fields = None
    pool.close()
    progress("single-request launches timed")
    return {
        "workload": "serve_poisson%d_ico7_kuhn32_kuhn96" % int(rate), "requests": nreq,
        "rate_per_s": rate, "meshes": [m.info()["nv"] for m in meshes],
        "path": "eik_batch_solve from a host loop, one launch in flight; requests that arrive "
                "during a launch wait for the next (the eik_server path, several launches in "
                "flight, is off)",
        "slots": {"ico7": slots[0], "kuhn32": slots[1], "kuhn96": slots[2]},
        "verts_per_block": vpb, "max_inner": max_inner,
        "served_solves_per_s": nreq / span, "p50_ms": float(np.percentile(lat, 50)) * 1e3,
        "p99_ms": float(np.percentile(lat, 99)) * 1e3, "launches": launches,
        "saturated_solves_per_s": nreq / span_s,
        "saturated_p50_ms": float(np.percentile(lat_s, 50)) * 1e3,
        "saturated_p99_ms": float(np.percentile(lat_s, 99)) * 1e3,
        "saturated_launches": launches_s,
        "saturated_upload_s": tt_s[0], "saturated_launch_s": tt_s[1], "poisson_upload_s": tt[0],
        "poisson_launch_s": tt[1],
        "plain_solves_per_s": nplain / t_plain[best_cfg], "plain_sample": nplain,
        "plain_config": {"verts_per_block": best_cfg[0], "max_inner": best_cfg[1]},
        "plain_solves_per_s_by_config": {"%d/%d" % c: nplain / t for c, t in t_plain.items()},
        "saturated_speedup_vs_plain": (nreq / span_s) / (nplain / t_plain[best_cfg]),
        "max_rel_diff_alone_vs_mix": worst, "alone_vs_mix_sample": len(sample),
        "mesh_setup_ms": [round(t, 1) for t in setup_ms],
        "note": "latency from arrival to the request's times in host memory; each request's speed "
                "field is uploaded with it from pinned host memory, in the served runs and the "
                "single-request launches alike",
    }


def copy_calibration():
    import torch
    torch.cuda.set_device(0)
    n = 64 * 1024 * 1024
    a = torch.ones(n, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"launches": 3, "bytes_read": n * 4,
                                           "bytes_written": n * 4}}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--vpb", type=int, default=64)
    ap.add_argument("--seed", type=int, default=20240607)
    ap.add_argument("--scheduler", type=int, default=1)
    ap.add_argument("--workgroups", type=int, default=0)
    ap.add_argument("--max-inner", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--serve", action="store_true",
                    help="the served Poisson stream through eik_server, several launches in flight "
                         "(disabled: it has stalled without output in some runs and the cause is "
                         "not yet shown; the run exits with an error)")
    ap.add_argument("--no-serve", action="store_true", help="(the default; kept for old command lines)")
    ap.add_argument("--no-aniso", action="store_true", help="skip the level-7 anisotropic icosphere")
    ap.add_argument("--no-history", action="store_true",
                    help="skip the per-iteration history leg (one launch per outer iteration)")
    ap.add_argument("--no-square", action="store_true", help="skip the 2D 32x32 square")
    ap.add_argument("--serve-requests", type=int, default=2000,
                    help="requests of the served workload (through eik_batch)")
    ap.add_argument("--no-serve-batch", action="store_true",
                    help="skip the served workload through eik_batch (one launch in flight)")
    ap.add_argument("--serve-batch-only", action="store_true",
                    help="only the served workload through eik_batch; prints its JSON and exits")
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--batch-only", action="store_true",
                    help="only the batched leg, at --batch-vpb / --batch-inner when both are given "
                         "(no autotune) and without the plain baseline; prints its JSON and exits")
    ap.add_argument("--batch-vpb", type=int, default=0)
    ap.add_argument("--batch-inner", type=int, default=0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy 3 times (scales the FETCH_SIZE / "
                         "WRITE_SIZE counters of a profile run) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.serve:
        print("bench.py: --serve is disabled: the served Poisson stream has stalled without "
              "output in some runs and the cause is not yet shown", file=sys.stderr)
        sys.exit(2)
    if args.gpus != 1:
        print("bench.py: this solver runs on one GPU; --gpus %d is not supported (multi-GPU "
              "sharding is a design note only)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    assert torch.cuda.is_available()
    torch.cuda.set_device(0)
    E = load_pkg()
    if args.serve_batch_only:
        print(json.dumps(serve_batch_leg(E, args.serve_requests, args.serve_rate, args.seed)))
        return
    if args.batch_only:
        fixed = (args.batch_vpb, args.batch_inner) if args.batch_vpb and args.batch_inner else None
        print(json.dumps(batch_leg(E, args.batch_count, args.seed, args.batch_steps, 32, fixed,
                                   plain_baseline=False)))
        return
    m, mult, c = make_workload(E, args.n, args.vpb, args.seed)
    m.upload()
    info = m.info()
    for _ in range(args.warmup):
        m.solve([c], out=False, scheduler=args.scheduler, workgroups=args.workgroups,
                max_inner=args.max_inner)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stats = []
    for _ in range(args.steps):
        _, st = m.solve([c], out=False, scheduler=args.scheduler, workgroups=args.workgroups,
                        max_inner=args.max_inner)
        stats.append(st)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    T = m.result()
    vu = sum(s["vertex_updates"] for s in stats)
    value = vu / dt
    visit_ms = float(np.mean([s["visit_ms"] for s in stats]))
    launches = float(np.mean([s["visit_launches"] for s in stats]))
    algo_bytes = float(np.mean([s["algo_bytes"] for s in stats]))
    algo_loaded = float(np.mean([s["algo_bytes_loaded"] for s in stats]))
    achieved = algo_bytes / (visit_ms * 1e-3) / 1e9
    achieved_loaded = algo_loaded / (visit_ms * 1e-3) / 1e9
    traffic, traffic_src = (measured_traffic("fim3d_kuhn%d_lognormal" % args.n)
                            if (args.vpb, args.scheduler) == (64, 1) else (None, "other config"))
    ms_step = dt * 1e3 / args.steps
    out = {
        "metric": "3D FIM vertex-updates/s (64^3 Kuhn cube, log-normal speed)",
        "value": value,
        "unit": "vertex-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        # rates that do not depend on how many visits the scheduler makes: mesh vertices and
        # whole solves per second
        "mesh_vertices_per_s": info["nv"] / (ms_step * 1e-3),
        "solves_per_s": 1e3 / ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered Kuhn cube and log-normal speeds)",
        "config": {"workload": "fim3d_kuhn%d_lognormal" % args.n, "verts": info["nv"],
                   "tets": info["nelem"], "blocks": info["nblocks"],
                   "verts_per_block": args.vpb, "scheduler": args.scheduler,
                   "seed": args.seed, "parallelism": "single GPU"},
        # partition quality (from the partition call): element copies per element (halo
        # elements), halo vertices per block, largest block against the mean
        "partition": {"method": "rcb_gap", "element_copies": info["block_elems"] / info["nelem"],
                      "halo_per_block": info["block_halo"] / info["nblocks"],
                      "mean_block_verts": info["nv"] / info["nblocks"],
                      "max_block_verts": info["max_block_verts"],
                      "imbalance": info["max_block_verts"] * info["nblocks"] / info["nv"]},
        "solve": {"block_visits": stats[-1]["block_visits"],
                  "vertex_updates": stats[-1]["vertex_updates"],
                  "inner_sweeps": stats[-1]["inner_sweeps"],
                  "local_solves": stats[-1]["local_solves"],
                  "reruns": stats[-1]["reruns"],
                  # wall-clock of the visit kernel's workgroups, summed over workgroups (ms):
                  # inside visits (load, sweeps, write-back), waiting for work, queueing
                  "wg_ms": {k: round(stats[-1][k], 3) for k in ("busy_ms", "load_ms", "sweep_ms",
                                                                "writeback_ms", "wait_ms",
                                                                "queue_ms")},
                  "outer_iterations": stats[-1]["outer_iterations"],
                  "kernel_ms": float(np.mean([s["kernel_ms"] for s in stats])),
                  "visit_ms": visit_ms, "visit_launches": launches},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic,
                     "traffic_source": traffic_src,
                     # the measured HBM bytes per launch over the kernel's duration
                     "traffic_frac": (traffic / (visit_ms * 1e-3) / 1e9 / HBM_PEAK_GBS
                                      if traffic else None),
                     "traffic_over_algorithmic": traffic / algo_bytes if traffic else None,
                     "algo_bytes_per_solve": algo_bytes,
                     # the same yardstick without the element data of re-runs whose records
                     # stay staged in LDS (they read none)
                     "achieved_loaded": achieved_loaded,
                     "frac_loaded": achieved_loaded / HBM_PEAK_GBS,
                     "algo_bytes_loaded_per_solve": algo_loaded},
    }
    progress("single solves done (%.3f ms per solve)" % ms_step)
    if not args.no_history:
        out["history"] = history_leg(m, c, args.max_inner)
        progress("history leg done (%.3f ms with per-iteration snapshots)"
                 % out["history"]["sync_history_ms"])
    c1 = None
    if not args.no_square:
        out["square2d"], c1 = square_leg(E, 20)
        progress("2D square leg done (%.3f ms per solve)" % out["square2d"]["ms_per_solve"])
    c3 = None
    if not args.no_aniso:
        out["aniso2d"], c3 = aniso_leg(E, args.seed, 5)
        progress("anisotropic surface leg done (%.3f ms per solve)" % out["aniso2d"]["ms_per_solve"])
    if not args.no_batch:
        out["batched"] = batch_leg(E, args.batch_count, args.seed, args.batch_steps, 32)
        progress("batched leg done (%.1f ms per batch)" % out["batched"]["batch_ms"])
    out["served"] = {"skipped": "the served Poisson stream has stalled without output in some "
                                "runs (no GPU fault); it stays off until the cause is shown"}
    if not args.no_serve_batch:
        out["served_batch"] = serve_batch_leg(E, args.serve_requests, args.serve_rate, args.seed)
        progress("served workload through eik_batch done (p99 %.1f ms)" % out["served_batch"]["p99_ms"])
    if not args.no_cpu:
        ref, ref_exact, cb = cpu_baseline(m, mult, c, args.max_inner)
        progress("cpu baseline done")
        tau = 0.25 * float(np.sqrt(mult.min())) / args.n
        Td = T.astype(np.float64)
        cb["max_rel_err_gpu_vs_exact_oracle"] = float(
            (np.abs(Td - ref_exact) / np.maximum(ref_exact, tau)).max())
        cb["rms_err_gpu_vs_exact_oracle"] = float(np.sqrt(np.mean((Td - ref_exact) ** 2)))
        cb["max_abs_err_gpu_vs_exact_oracle"] = float(np.abs(Td - ref_exact).max())
        cb["rel_err_floor_tau"] = tau
        cb["max_rel_err_gpu_vs_regularised_oracle"] = float(
            (np.abs(Td - ref) / np.maximum(ref, tau)).max())
        out["cpu_baseline"] = cb
        if c1 is not None:
            out["square2d"]["max_rel_err_vs_oracle"] = square_error(*c1)
        if c3 is not None:
            out["aniso2d"]["max_rel_err_vs_oracle"] = aniso_error(*c3)
            progress("anisotropic surface checked against the oracle")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
