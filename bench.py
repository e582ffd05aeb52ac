"""Benchmark: 3D FIM on the 64^3 jittered Kuhn-tet cube (274,625 vertices, 1,572,864 tets),
per-tet log-normal(0, 0.5) speed, one source at the centre vertex, 64-vertex blocks, fp32.

One step = one complete solve of that workload with the mesh layout resident in HBM (seeding,
the persistent block-visit kernel to convergence; no host<->device copy of the mesh).
Metric: vertex-updates/s, a vertex-update being one owned vertex of a block processed in one
block visit (inner LDS sweeps are not counted), summed over the K timed solves.

Prints one JSON line.  `roofline` is the block-visit kernel's algorithmic bytes (element records
+ halo indices and values + owned values read and written + block descriptor, per visit, counted
on the device) divided by its duration measured with HIP events on the mesh's stream.
`cpu_baseline` is the serial FP64 oracle FIM (oracle/fim_oracle.c) on one host core, run on one
full solve of the same mesh.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

# HBM bytes per launch of the block-visit kernel, measured with rocprofv3 (FETCH_SIZE and
# WRITE_SIZE passes, each scaled by the reading of a plain 256 MiB copy): the figure and the
# commands are in profiles/k_visit_async_fim3d_kuhn64_lognormal.md.  Keyed by the configuration
# it was measured on; other configurations report null.
MEASURED_TRAFFIC = {
    ("fim3d_kuhn64_lognormal", 64, 1, 2): 2.497e9,
}


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


def cpu_baseline(m, mult, c):
    """Times the serial FP64 oracle (reference formulas, SMALLNUM regularisation as shipped) on
    one full solve; also returns the unregularised oracle solution for the error figure."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    xyz = m.vertices().astype(np.float64)
    tets = m.elements()
    M6 = O.tet_metric(tets.shape[0], mult.astype(np.float64))
    t0 = time.perf_counter()
    T, st = O.fim3d(xyz, tets, [c], M6=M6)
    dt = time.perf_counter() - t0
    with O.exact():
        Te, _ = O.fim3d(xyz, tets, [c], M6=M6)
    return T, Te, {"value": st["updates"] / dt, "unit": "vertex-updates/s", "cores": 1,
                   "kind": "port", "solve_ms": dt * 1e3,
                   "sample": "one full solve of the same 64^3 workload by the serial FP64 "
                             "oracle (vertex-level active-list FIM; a vertex-update = one "
                             "vertex solve)"}


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
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy 3 times (scales the FETCH_SIZE / "
                         "WRITE_SIZE counters of a profile run) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print("bench.py: this solver runs on one GPU; --gpus %d is not supported (multi-GPU "
              "sharding is a design note only)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch
    assert torch.cuda.is_available()
    torch.cuda.set_device(0)
    E = load_pkg()
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
    achieved = algo_bytes / (visit_ms * 1e-3) / 1e9
    out = {
        "metric": "3D FIM vertex-updates/s (64^3 Kuhn cube, log-normal speed)",
        "value": value,
        "unit": "vertex-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded jittered Kuhn cube and log-normal speeds)",
        "config": {"workload": "fim3d_kuhn%d_lognormal" % args.n, "verts": info["nv"],
                   "tets": info["nelem"], "blocks": info["nblocks"],
                   "verts_per_block": args.vpb, "scheduler": args.scheduler,
                   "seed": args.seed, "parallelism": "single GPU"},
        "solve": {"block_visits": stats[-1]["block_visits"],
                  "vertex_updates": stats[-1]["vertex_updates"],
                  "inner_sweeps": stats[-1]["inner_sweeps"],
                  "local_solves": stats[-1]["local_solves"],
                  "outer_iterations": stats[-1]["outer_iterations"],
                  "kernel_ms": float(np.mean([s["kernel_ms"] for s in stats])),
                  "visit_ms": visit_ms, "visit_launches": launches},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": MEASURED_TRAFFIC.get(
                         ("fim3d_kuhn%d_lognormal" % args.n, args.vpb, args.scheduler,
                          args.max_inner)),
                     "algo_bytes_per_solve": algo_bytes},
    }
    if not args.no_cpu:
        ref, ref_exact, cb = cpu_baseline(m, mult, c)
        tau = 0.25 * float(np.sqrt(mult.min())) / args.n
        Td = T.astype(np.float64)
        cb["max_rel_err_gpu_vs_exact_oracle"] = float(
            (np.abs(Td - ref_exact) / np.maximum(ref_exact, tau)).max())
        cb["max_rel_err_gpu_vs_regularised_oracle"] = float(
            (np.abs(Td - ref) / np.maximum(ref, tau)).max())
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
