"""GPU edge cases of the solve path through the C ABI, each against the FP64 oracle (same
tolerance as the parity tests: max |T - T_oracle| / max(T_oracle, tau) <= 1e-5):

* blocks too large for LDS staging: element records are read from memory and each sweep works
  through the block's (element, corner) pairs in chunks of the work queue;
* a single block (no dependents, no halo);
* seed values other than zero, duplicate seeds (the smaller value holds), every vertex a seed.
"""
import numpy as np
import pytest

from conftest import center_vertex, lognormal_mult, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _say(*a):
    print(*a, flush=True)


def _tau_tet(xyz, tets, seeds, mult):
    xyz = np.asarray(xyz, np.float64)
    s = np.zeros(xyz.shape[0], bool)
    s[np.atleast_1d(seeds)] = True
    sel = s[tets].any(axis=1)
    t, f = tets[sel], np.sqrt(mult[sel].astype(np.float64))
    best = np.inf
    for i in range(4):
        for j in range(i + 1, 4):
            best = min(best, float((f * np.linalg.norm(xyz[t[:, i]] - xyz[t[:, j]],
                                                       axis=1)).min()))
    return 0.25 * best


def _oracle_3d(O, m, mult, seeds, seed_vals=None):
    xyz, tets = m.vertices().astype(np.float64), m.elements()
    with O.exact():
        ref, _ = O.fim3d(xyz, tets, seeds, M6=O.tet_metric(tets.shape[0], mult.astype(np.float64)),
                         seed_vals=seed_vals)
    return ref


@pytest.mark.parametrize("sched", [0, 1])
def test_tet_blocks_beyond_lds_read_records_from_memory(E, O, sched):
    n = 16
    _say("kuhn n=%d, 1200-vertex blocks (records beyond LDS), scheduler %d" % (n, sched))
    m = E.Mesh.kuhn_cube(n, 0.2, seed=31)
    nt = m.info()["nelem"]
    mult = lognormal_mult(nt, 31)
    m.set_speed(mult)
    m.partition(1200)
    lay = m.layout()
    max_elem = int(lay["blocks"][:, 6].max())
    # 48-byte records of the largest block exceed the 160 KiB of LDS a workgroup can hold,
    # and its (element, corner) pairs exceed one chunk of the work queue
    assert max_elem * 48 > 160 * 1024 and 4 * max_elem > 2048
    c = center_vertex(n)
    T, st = m.solve([c], scheduler=sched)
    ref = _oracle_3d(O, m, mult, [c])
    err = rel_err(T, ref, _tau_tet(m.vertices(), m.elements(), [c], mult))
    _say("  blocks %d, largest %d tets, visits %d, max rel err %.3e"
         % (lay["nblocks"], max_elem, st["block_visits"], err))
    assert err <= TOL


@pytest.mark.parametrize("sched", [0, 1])
def test_single_triangle_block_beyond_lds(E, O, sched):
    """One block holding the whole 50x50 square (5000 faces plus virtual faces): no
    dependents, no halo, records read from memory, the sweep in queue chunks."""
    _say("square 50 in one block, scheduler %d" % sched)
    m = E.Mesh.square(50)
    m.set_labels(np.zeros(m.nv, np.int32), 1)
    lay = m.layout()
    assert lay["nblocks"] == 1 and int(lay["blocks"][0, 6]) * 32 > 160 * 1024
    T, st = m.solve([0], scheduler=sched)
    xyz, f = m.vertices(), m.elements()
    vf, origin = O.tri_virtual_faces(xyz, f)
    F = np.ones(f.shape[0])
    ref, _ = O.fim2d(xyz.astype(np.float64), f, [0], F, 0, vfaces=vf, vparam=F[origin])
    err = rel_err(T, ref, 0.25 / 50)
    _say("  visits %d, max rel err %.3e" % (st["block_visits"], err))
    assert err <= TOL
    # one block, so the visits are the block's re-runs only
    assert st["block_visits"] >= 1


def test_nonzero_and_duplicate_seed_values(E, O):
    """Sources with start times: the solve is the minimum over sources of (start time + travel
    time); a seed given twice keeps its smaller value."""
    n = 12
    _say("kuhn n=%d, seeds with start times" % n)
    m = E.Mesh.kuhn_cube(n, 0.2, seed=41)
    nt = m.info()["nelem"]
    mult = lognormal_mult(nt, 41)
    m.set_speed(mult)
    m.partition(64)
    seeds = np.array([0, center_vertex(n), m.nv - 1, center_vertex(n)], np.int32)
    vals = np.array([0.0, 0.15, 0.4, 0.05])
    T, _ = m.solve(seeds, seed_vals=vals, scheduler=1)
    ref = _oracle_3d(O, m, mult, seeds, seed_vals=vals)
    assert T[center_vertex(n)] == np.float32(0.05)
    err = rel_err(T, ref, _tau_tet(m.vertices(), m.elements(), seeds, mult))
    _say("  max rel err %.3e" % err)
    assert err <= TOL
    # the time at a vertex is never below the earliest start time
    assert T.min() == 0.0


def test_every_vertex_a_seed(E):
    n = 8
    _say("kuhn n=%d, every vertex a seed" % n)
    m = E.Mesh.kuhn_cube(n, 0.2, seed=43)
    m.partition(64)
    rng = np.random.default_rng(43)
    vals = rng.uniform(0.0, 1e-3, m.nv).astype(np.float32)
    for sched in (0, 1):
        T, _ = m.solve(np.arange(m.nv, dtype=np.int32), seed_vals=vals, scheduler=sched)
        # start times differ by less than any edge's travel time (edges of at least 0.6 h at
        # speed 1), so no source lowers another's start time
        np.testing.assert_array_equal(T, vals)


def _fan(n):
    """A disc as a fan of n triangles around a centre vertex (id 0); rim ids 1..n."""
    th = 2.0 * np.pi * np.arange(n) / n
    xyz = np.zeros((n + 1, 3), np.float32)
    xyz[1:, 0], xyz[1:, 1] = np.cos(th), np.sin(th)
    faces = np.array([[0, 1 + i, 1 + (i + 1) % n] for i in range(n)], np.int32)
    return xyz, faces


def test_block_with_more_than_64_dependents(E, O):
    """Every face of a fan holds the centre vertex, so with 4-vertex blocks the centre's block is
    in every other block's halo: more than 64 dependents, which are then tracked as per-vertex
    lists instead of 64-bit masks (single solves under both schedulers, and a multi-solve
    launch, whose kernel queues every dependent of such a block)."""
    xyz, f = _fan(400)
    m = E.Mesh.tri(xyz, f)
    m.partition(4)
    lay = m.layout()
    ndep = np.diff(lay["dep_start"])
    _say("fan of 400 triangles, %d blocks, most dependents %d" % (lay["nblocks"], ndep.max()))
    assert ndep.max() > 64
    seeds = np.array([1, 200], np.int32)
    F = np.ones(f.shape[0])
    ref, _ = O.fim2d(xyz.astype(np.float64), f, seeds, F, 0)
    tau = 0.25 * 2.0 * np.sin(np.pi / 400)
    for sched in (0, 1):
        T, st = m.solve(seeds, scheduler=sched)
        err = rel_err(T, ref, tau)
        _say("  scheduler %d: visits %d, max rel err %.3e" % (sched, st["block_visits"], err))
        assert err <= TOL
    b = E.Batch([m, m])
    Ts, st = b.solve([seeds, seeds[::-1].copy()])
    b.close()
    for T in Ts:
        err = rel_err(T, ref, tau)
        _say("  multi-solve launch: max rel err %.3e" % err)
        assert err <= TOL


def test_wait_limit_bounds_every_queue_wait(E):
    """Every queue wait of the persistent kernels is bounded (wait_limit_us).  A limit of 1 us,
    set from the host, makes idle workgroups leave their waits and exit: the call returns either
    a complete result (the queue still reached its end) or EIK_ERR_TIMEOUT with the abandoned
    waits counted and no result kept; the next solve with the default limit (10 s) is exact
    again, since every launch resets its queue.  Single solves and a multi-solve launch."""
    import ctypes
    m = E.Mesh.kuhn_cube(20, jitter=0.2, seed=5)
    m.partition(64)
    seeds = np.array([0, m.nv // 2], np.int32)
    ref, st0 = m.solve(seeds)
    assert st0["abandoned_waits"] == 0
    tau = 0.25 / 20
    abandoned = 0
    for rep in range(3):
        o = E.default_opts(wait_limit_us=1)
        st = E.SolveStats()
        T = np.zeros(m.nv, np.float32)
        rc = E.lib().eik_solve(m.h, seeds.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                               seeds.size, None, ctypes.byref(o),
                               T.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.byref(st))
        _say("  single solve, 1 us limit, rep %d: rc %d, abandoned waits %d, visits %d"
             % (rep, rc, st.abandoned_waits, st.block_visits))
        assert rc in (E.EIK_OK, E.EIK_ERR_TIMEOUT)
        abandoned += st.abandoned_waits
        if rc == E.EIK_OK:
            assert rel_err(T, ref, tau) <= TOL
        else:
            assert st.abandoned_waits > 0
            with pytest.raises(E.EikError) as ei:
                m.result()
            assert ei.value.code == E.EIK_ERR_STATE
    assert abandoned > 0
    T, st = m.solve(seeds)
    assert st["abandoned_waits"] == 0 and rel_err(T, ref, tau) <= TOL
    with pytest.raises(E.EikError) as ei:
        m.solve(seeds, wait_limit_us=-1)
    assert ei.value.code == E.EIK_ERR_ARG

    b = E.Batch([m, m])
    try:
        try:
            Ts, st = b.solve([seeds, seeds[::-1].copy()], wait_limit_us=1)
            _say("  multi-solve launch, 1 us limit: complete, abandoned waits %d"
                 % st["abandoned_waits"])
            for T in Ts:
                assert rel_err(T, ref, tau) <= TOL
        except E.EikError as e:
            _say("  multi-solve launch, 1 us limit: %s" % e)
            assert e.code == E.EIK_ERR_TIMEOUT
        Ts, st = b.solve([seeds, seeds[::-1].copy()])
        assert st["abandoned_waits"] == 0
        for T in Ts:
            assert rel_err(T, ref, tau) <= TOL
    finally:
        b.close()
