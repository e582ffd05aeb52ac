"""Batched solves (one persistent launch for several meshes) on the GPU: every solve of a batch
equals the same mesh solved alone and the FP64 oracle (1e-5 relative, floor as in the parity
tests), for mixed tet / triangle batches, repeated solves through the captured graph, the
direct (no graph) path, and per-solve speed factors over one resident mesh."""
import numpy as np
import pytest

from conftest import center_vertex, golden, lognormal_mult, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _say(*a):
    print(*a, flush=True)


def _face_metrics(nf, seed):
    rng = np.random.default_rng(seed)
    th = rng.uniform(0.0, np.pi, nf)
    k = rng.uniform(1.0, 10.0, nf)
    c, s = np.cos(th), np.sin(th)
    return np.stack([c * c + k * s * s, (1.0 - k) * c * s, s * s + k * c * c],
                    axis=1).astype(np.float32)


def _mixed(E):
    meshes, refs = [], []
    for n, seed in ((8, 1), (12, 2), (16, 3)):
        m = E.Mesh.kuhn_cube(n, 0.2, seed=seed)
        mult = lognormal_mult(m.info()["nelem"], seed)
        m.set_speed(mult)
        m.partition(64)
        meshes.append(m)
        refs.append(("tet", mult))
    m = E.Mesh.load_tetgen(golden("sphere334"))
    m.partition(32)
    meshes.append(m)
    refs.append(("tet", None))
    m = E.Mesh.square(32)
    m.partition(64)
    meshes.append(m)
    refs.append(("tri", None))
    m = E.Mesh.icosphere(4)
    M3 = _face_metrics(m.info()["nelem"], 4)
    m.set_metric_tri(M3)
    m.partition(64)
    meshes.append(m)
    refs.append(("aniso", M3))
    return meshes, refs


def _oracle(O, m, kind, param, seeds):
    xyz, el = m.vertices().astype(np.float64), m.elements()
    if kind == "tet":
        M6 = None if param is None else O.tet_metric(el.shape[0], param.astype(np.float64))
        with O.exact():
            return O.fim3d(xyz, el, seeds, M6=M6)[0]
    if kind == "tri":
        vf, origin = O.tri_virtual_faces(xyz, el)
        F = np.ones(el.shape[0])
        return O.fim2d(xyz, el, seeds, F, 0, vfaces=vf, vparam=F[origin])[0]
    G = O.tri_gram(xyz, el, metric=param.astype(np.float64))
    return O.fim2d(xyz, el, seeds, G, 1)[0]


def _seeds(meshes, rng):
    return [rng.choice(m.nv, size=rng.integers(1, 4), replace=False) for m in meshes]


def test_batch_equals_single_solves_and_oracle(E, O):
    meshes, refs = _mixed(E)
    b = E.Batch(meshes)
    rng = np.random.default_rng(7)
    seeds = _seeds(meshes, rng)
    _say("batch of %d meshes, %s" % (len(meshes), b.info()))
    Ts, st = b.solve(seeds)
    _say("  visits %d, kernel %.3f ms" % (st["block_visits"], st["kernel_ms"]))
    for i, (m, (kind, param)) in enumerate(zip(meshes, refs)):
        alone, _ = m.solve(seeds[i], scheduler=1)
        np.testing.assert_allclose(Ts[i], alone, rtol=2e-6, atol=1e-7)
        ref = _oracle(O, m, kind, param, seeds[i])
        err = rel_err(Ts[i], ref, 1e-3 * np.max(ref[ref < 1e6]))
        _say("  solve %d (%s): max rel err vs oracle %.3e" % (i, kind, err))
        assert err <= TOL


def test_batch_graph_replay_and_direct_path(E):
    meshes, _ = _mixed(E)
    b = E.Batch(meshes)
    rng = np.random.default_rng(8)
    for rep in range(3):
        seeds = _seeds(meshes, rng)
        _say("replay %d" % rep)
        Tg, _ = b.solve(seeds, max_inner=2)
        alone = [m.solve(s, scheduler=1)[0] for m, s in zip(meshes, seeds)]
        for x, y in zip(Tg, alone):
            np.testing.assert_allclose(x, y, rtol=2e-6, atol=1e-7)
    b.set_graph(False)
    Td, _ = b.solve(seeds, max_inner=2)
    for x, y in zip(Td, alone):
        np.testing.assert_allclose(x, y, rtol=2e-6, atol=1e-7)
    # a mesh without seeds keeps LARGENUM
    empty = [s if i else np.zeros(0, np.int32) for i, s in enumerate(seeds)]
    b.set_graph(True)
    Te, _ = b.solve(empty)
    assert np.all(Te[0] == np.float32(1e9))
    np.testing.assert_allclose(Te[1], alone[1], rtol=2e-6, atol=1e-7)


def test_batch_factors_over_one_resident_mesh(E):
    """Three solves of one mesh, each with its own speed field: equal to solving copies of the
    mesh whose speeds were set the reference's way (scalar per tet = 1/speed^2)."""
    n = 12
    m = E.Mesh.kuhn_cube(n, 0.2, seed=5)
    m.partition(64)
    nt = m.info()["nelem"]
    b = E.Batch([m, m, m])
    fields = [lognormal_mult(nt, 50 + k) for k in range(3)]
    for k, f in enumerate(fields):
        b.set_factors(k, f)
    seeds = [[center_vertex(n)], [0], [0, m.nv - 1]]
    Ts, _ = b.solve(seeds)
    for k, f in enumerate(fields):
        c = E.Mesh.kuhn_cube(n, 0.2, seed=5)
        c.set_speed(f)
        c.partition(64)
        alone, _ = c.solve(seeds[k], scheduler=1)
        err = np.max(np.abs(Ts[k] - alone) / np.maximum(alone, 1e-3))
        _say("  field %d: max rel diff vs separately set speeds %.3e" % (k, err))
        assert err <= 2e-6
    b.set_factors(1, None)
    Tn, _ = b.solve(seeds)
    plain, _ = m.solve(seeds[1], scheduler=1)
    np.testing.assert_allclose(Tn[1], plain, rtol=2e-6, atol=1e-7)


def test_batch_factors_replaced_and_rejected(E):
    """A solve slot that takes a new speed field before every solve (the serving pattern) gives
    the same times each time a field comes back; factors that are not finite and > 0 are
    rejected (checked on the device) and leave the slot's current field in place."""
    n = 10
    m = E.Mesh.kuhn_cube(n, 0.2, seed=15)
    m.partition(64)
    nt = m.info()["nelem"]
    b = E.Batch([m, m])
    fa, fb = lognormal_mult(nt, 71), lognormal_mult(nt, 72)
    seeds = [[center_vertex(n)], [0]]
    b.set_factors(0, fa)
    Ta, _ = b.solve(seeds)
    b.set_factors(0, fb)
    Tb, _ = b.solve(seeds)
    assert np.max(np.abs(Ta[0] - Tb[0])) > 1e-3  # the field took effect
    b.set_factors(0, fa)
    Ta2, _ = b.solve(seeds)
    np.testing.assert_allclose(Ta2[0], Ta[0], rtol=2e-6, atol=1e-7)
    for bad_value in (-1.0, np.nan, np.inf, 0.0):
        bad = fb.copy()
        bad[nt // 2] = bad_value
        _say("rejected factor value %r" % bad_value)
        with pytest.raises(E.EikError) as ei:
            b.set_factors(0, bad)
        assert ei.value.code == E.EIK_ERR_ARG
    Tk, _ = b.solve(seeds)  # still field a
    np.testing.assert_allclose(Tk[0], Ta[0], rtol=2e-6, atol=1e-7)
    np.testing.assert_allclose(Tk[1], Ta[1], rtol=2e-6, atol=1e-7)


def test_batch_result_of_one_solve(E):
    """A solve whose times stay on the device (out=False) gives each solve's times on request,
    equal to what a solve with output returns; before any solve there is no result."""
    a = E.Mesh.kuhn_cube(9, 0.2, seed=21)
    a.partition(64)
    t = E.Mesh.icosphere(3)
    t.partition(64)
    b = E.Batch([a, t, a])
    with pytest.raises(E.EikError) as ei:
        b.result(0)
    assert ei.value.code == E.EIK_ERR_STATE
    seeds = [[0], [5, 60], np.zeros(0, np.int32)]
    Tall, _ = b.solve(seeds)
    b.solve([[a.nv - 1], [0], [3]], out=False)  # other times in between
    b.solve(seeds, out=False)
    for i in range(3):
        np.testing.assert_allclose(b.result(i), Tall[i], rtol=2e-6, atol=1e-7)
    assert np.all(b.result(2) == np.float32(1e9))
    with pytest.raises(E.EikError) as ei:
        b.result(3)
    assert ei.value.code == E.EIK_ERR_ARG
    # the direct launch path (no graph), and a solve with more seeds than the launch buffers held
    # (they are allocated again)
    b.set_graph(False)
    b.solve(seeds, out=False)
    for i in range(3):
        np.testing.assert_allclose(b.result(i), Tall[i], rtol=2e-6, atol=1e-7)
    b.set_graph(True)
    many = [np.tile(np.arange(a.nv, dtype=np.int32), 2)[:1100], [5], [0]]  # repeats allowed
    Tm, _ = b.solve(many)
    b.solve(many, out=False)
    for i in range(3):
        np.testing.assert_allclose(b.result(i), Tm[i], rtol=2e-6, atol=1e-7)
    assert np.all(b.result(0) == 0.0)  # every vertex is a source


def test_batch_kind_without_seeds(E):
    """A batch whose triangle solves have no seeds at all: their launch ends at once (nothing
    was queued) and their times stay LARGENUM; the tet solves are unaffected."""
    a = E.Mesh.kuhn_cube(8, 0.2, seed=1)
    a.partition(64)
    t = E.Mesh.square(16)
    t.partition(64)
    b = E.Batch([a, t])
    _say("tets seeded, triangles not")
    Ts, st = b.solve([[0], np.zeros(0, np.int32)])
    alone, _ = a.solve([0], scheduler=1)
    np.testing.assert_allclose(Ts[0], alone, rtol=2e-6, atol=1e-7)
    assert np.all(Ts[1] == np.float32(1e7))
    _say("nothing seeded")
    Tn, _ = b.solve([np.zeros(0, np.int32), np.zeros(0, np.int32)])
    assert np.all(Tn[0] == np.float32(1e9)) and np.all(Tn[1] == np.float32(1e7))


@pytest.mark.parametrize("stage", [1, 2])
def test_batch_record_shapes_with_factors(E, stage):
    """Both workgroup shapes of a multi-solve launch (element records staged in LDS, or read from
    memory with the per-solve factors applied at every use) give the same times as solving
    copies whose speeds or metrics were set directly, for tets and anisotropic triangles, with
    and without factors."""
    n = 10
    m = E.Mesh.kuhn_cube(n, 0.2, seed=9)
    m.partition(64)
    t = E.Mesh.icosphere(3)
    M3 = _face_metrics(t.info()["nelem"], 11)
    t.set_metric_tri(M3)
    t.partition(64)
    b = E.Batch([m, m, t, t])
    fm = lognormal_mult(m.info()["nelem"], 61)
    ft = lognormal_mult(t.info()["nelem"], 62)
    b.set_factors(1, fm)
    b.set_factors(3, ft)
    seeds = [[center_vertex(n)], [0, m.nv - 1], [0], [5, 77]]
    _say("stage_records=%d" % stage)
    Ts, _ = b.solve(seeds, stage_records=stage)
    cm = E.Mesh.kuhn_cube(n, 0.2, seed=9)
    cm.set_speed(fm)
    cm.partition(64)
    ct = E.Mesh.icosphere(3)
    ct.set_metric_tri((M3 * ft[:, None]).astype(np.float32))
    ct.partition(64)
    refs = [m.solve(seeds[0], scheduler=1)[0], cm.solve(seeds[1], scheduler=1)[0],
            t.solve(seeds[2], scheduler=1)[0], ct.solve(seeds[3], scheduler=1)[0]]
    for k in range(4):
        err = np.max(np.abs(Ts[k] - refs[k]) / np.maximum(refs[k], 1e-3))
        _say("  solve %d: max rel diff vs directly set speeds %.3e" % (k, err))
        assert err <= 2e-6


@pytest.mark.parametrize("vpb", [64, 1024])
def test_batch_block_sizes_both_register_budgets(E, vpb):
    """The batch kernel is built for two register budgets and a launch takes the larger one when
    its LDS already limits the CU to few workgroups: small blocks (64 vertices) run the 128-VGPR
    build, large ones (1,024 vertices: a few workgroups per CU) the 168-VGPR build.  For tets and
    triangles alike, every solve of the batch equals the same mesh solved alone.  The records of
    a 1,024-vertex tet block (about 6,800 elements, 326 KB) do not fit the LDS: asking to stage
    them fails cleanly, and the batch solves again afterwards."""
    a = E.Mesh.kuhn_cube(20, 0.2, seed=31)
    a.set_speed(lognormal_mult(a.info()["nelem"], 31))
    t = E.Mesh.icosphere(5)
    M3 = _face_metrics(t.info()["nelem"], 32)
    t.set_metric_tri(M3)
    for m in (a, t):
        m.partition(vpb)
    _say("%d vertices per block: %d and %d blocks" % (vpb, a.info()["nblocks"], t.info()["nblocks"]))
    b = E.Batch([a, t, a])
    seeds = [[center_vertex(20)], [0, 777], [0]]
    Ts, st = b.solve(seeds, max_inner=1)
    _say("  visits %d" % st["block_visits"])
    for i, m in enumerate((a, t, a)):
        alone, _ = m.solve(seeds[i], scheduler=1)
        err = np.max(np.abs(Ts[i] - alone) / np.maximum(alone, 1e-3))
        _say("  solve %d: max rel diff vs alone %.3e" % (i, err))
        assert err <= 2e-6
    if vpb == 1024:
        with pytest.raises(E.EikError) as ei:
            b.solve(seeds, max_inner=1, stage_records=1)
        assert ei.value.code == E.EIK_ERR_CAPACITY
        T2, _ = b.solve(seeds, max_inner=1)
        for x, y in zip(T2, Ts):
            np.testing.assert_allclose(x, y, rtol=2e-6, atol=1e-7)
