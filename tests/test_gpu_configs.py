"""The measured configurations at their full sizes on the GPU, against the FP64 oracle
(1e-5 relative, floor as in the parity tests):

* the 64^3 jittered Kuhn cube (274,625 vertices, 1,572,864 tets), log-normal(0, 0.5) speed
  per tet, centre source, 64-vertex blocks: the bench workload;
* the level-7 icosphere (163,842 vertices, 327,680 faces), a random SPD 2x2 metric per face
  with condition number in [1, 10], 4 random sources;
* a 16-cube sample of the 256-cube batch (n ~ U{16..48} cells per side, two-layer speed, 1-8 sources)
  solved in one launch: every solve equal to the same mesh solved alone, three of them against
  the oracle."""
import numpy as np
import pytest

from conftest import center_vertex, rel_err
from test_gpu_parity import _face_metrics

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _say(*a):
    print(*a, flush=True)


def _tau_tet(xyz, tets, seeds, mult):
    """A quarter of the shortest metric edge of the tets around the seeds: the floor of the
    relative error (times near a source are compared in absolute terms)."""
    xyz = np.asarray(xyz, np.float64)
    near = np.isin(tets, np.atleast_1d(seeds)).any(axis=1)
    t = tets[near]
    f = np.sqrt(mult[near].astype(np.float64))
    best = np.inf
    for i in range(4):
        for j in range(i + 1, 4):
            best = min(best, float((f * np.linalg.norm(xyz[t[:, i]] - xyz[t[:, j]], axis=1)).min()))
    return 0.25 * best


def test_c2_full_size_against_exact_oracle(E, O):
    n, seed = 64, 20240607
    _say("64^3 Kuhn cube, log-normal speed, centre source")
    m = E.Mesh.kuhn_cube(n, jitter=0.2, seed=seed)
    nt = m.info()["nelem"]
    assert (m.nv, nt) == (274625, 1572864)
    rng = np.random.default_rng(seed)
    mult = (1.0 / np.exp(rng.normal(0.0, 0.5, nt)) ** 2).astype(np.float32)
    m.set_speed(mult)
    m.partition(64)
    c = center_vertex(n)
    T, st = m.solve([c])
    xyz, tets = m.vertices().astype(np.float64), m.elements()
    with O.exact():
        ref, _ = O.fim3d(xyz, tets, [c], M6=O.tet_metric(nt, mult.astype(np.float64)))
    err = rel_err(T, ref, _tau_tet(xyz, tets, [c], mult))
    _say("  blocks %d, visits %d, max rel err %.3e" % (m.info()["nblocks"], st["block_visits"], err))
    assert err <= TOL


def test_c3_full_size_anisotropic_icosphere(E, O):
    _say("icosphere level 7, anisotropic metric, 4 sources")
    m = E.Mesh.icosphere(7)
    xyz, f = m.vertices(), m.elements()
    assert (xyz.shape[0], f.shape[0]) == (163842, 327680)
    M3 = _face_metrics(f.shape[0], 1007).astype(np.float32)
    m.set_metric_tri(M3)
    m.partition(64)
    rng = np.random.default_rng(7)
    seeds = rng.choice(xyz.shape[0], 4, replace=False)
    T, st = m.solve(seeds)
    M = M3.astype(np.float64)
    G = O.tri_gram(xyz, f, metric=M)
    vf, origin = O.tri_virtual_faces(xyz, f)
    Gv = O.tri_gram(xyz, vf, metric=M[origin], frame_faces=f[origin]) if len(vf) else None
    ref, _ = O.fim2d(xyz.astype(np.float64), f, seeds, G, 1, vfaces=vf, vparam=Gv)
    err = rel_err(T, ref, 0.25 * np.sqrt(G[:, 0].min()))
    _say("  visits %d, max rel err %.3e" % (st["block_visits"], err))
    assert err <= TOL


def test_c4_batch_sample_alone_and_oracle(E, O):
    rng = np.random.default_rng(20240607)
    meshes, seeds, mults = [], [], []
    for i in range(16):
        n = int(rng.integers(16, 49))
        m = E.Mesh.kuhn_cube(n, jitter=0.2, seed=100 + i)
        xyz, t = m.vertices(), m.elements()
        zc = xyz[:, 2][t].mean(axis=1)
        z0, above = rng.uniform(0.0, 1.0), rng.uniform(0.2, 5.0)
        speed = np.where(zc < z0, 1.0, above)
        mult = (1.0 / speed ** 2).astype(np.float32)
        m.set_speed(mult)
        m.partition(128)
        meshes.append(m)
        mults.append(mult)
        seeds.append(rng.choice(m.nv, int(rng.integers(1, 9)), replace=False).astype(np.int32))
    _say("batch sample: 16 cubes, %d vertices" % sum(m.nv for m in meshes))
    b = E.Batch(meshes)
    Ts, st = b.solve(seeds, max_inner=1)
    _say("  one launch: visits %d, kernel %.2f ms" % (st["block_visits"], st["kernel_ms"]))
    for i, m in enumerate(meshes):
        alone, _ = m.solve(seeds[i])
        np.testing.assert_allclose(Ts[i], alone, rtol=2e-6, atol=1e-7)
    for i in (0, 7, 15):
        m = meshes[i]
        xyz, tets = m.vertices().astype(np.float64), m.elements()
        nt = tets.shape[0]
        with O.exact():
            ref, _ = O.fim3d(xyz, tets, list(seeds[i]), M6=O.tet_metric(nt, mults[i].astype(np.float64)))
        err = rel_err(Ts[i], ref, _tau_tet(xyz, tets, list(seeds[i]), mults[i]))
        _say("  cube %d (%d vertices): max rel err vs oracle %.3e" % (i, m.nv, err))
        assert err <= TOL
    b.close()


def test_large_cube_async_equals_sync_and_bounds(E):
    """128^3 Kuhn cube (2.1 M vertices, 12.6 M tets), log-normal speeds: the persistent scheduler
    passes through both visit regimes (throughput-bound while the front is wide, latency-bound
    at the start and the end) and must reach the same fixed point as the one-launch-per-outer-
    iteration scheduler; every time is at least the straight-line distance at the fastest speed
    (no oracle at this size)."""
    n = 128
    m = E.Mesh.kuhn_cube(n, jitter=0.2, seed=128)
    nt = m.info()["nelem"]
    rng = np.random.default_rng(128)
    mult = (1.0 / np.exp(rng.normal(0.0, 0.5, nt)) ** 2).astype(np.float32)
    m.set_speed(mult)
    m.partition(64)
    c = center_vertex(n)
    Ta, sa = m.solve([c], scheduler=1)
    Ts, ss = m.solve([c], scheduler=0)
    print("128^3: async %.2f ms, %d visits, %d re-runs; sync %d outer iterations"
          % (sa["kernel_ms"], sa["block_visits"], sa["reruns"], ss["outer_iterations"]), flush=True)
    assert 0 < sa["reruns"] < sa["block_visits"]
    assert sa["inner_sweeps"] > sa["block_visits"]  # the throughput regime ran
    rel = np.abs(Ta.astype(np.float64) - Ts) / np.maximum(Ts, 1e-3)
    assert rel.max() <= 1e-5, rel.max()
    xyz = m.vertices().astype(np.float64)
    lower = np.sqrt(float(mult.min())) * np.linalg.norm(xyz - xyz[c], axis=1)
    assert np.all(Ta >= lower * (1 - 1e-5) - 1e-6)
