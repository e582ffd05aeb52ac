"""eik_solve_noalloc, the solve that allocates nothing (its state in the library's per-device
scratch array): through the C ABI, against eik_solve on the same meshes and sources (1e-5
relative), on two calls in a row, the second on a smaller mesh (the array keeps the first call's
contents, which the second must clear), and a mesh too large for the array fails cleanly."""
import numpy as np
import pytest

from conftest import center_vertex, lognormal_mult, rel_err

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _mesh(E, n, seed):
    m = E.Mesh.kuhn_cube(n, 0.2, seed=seed)
    m.set_speed(lognormal_mult(m.info()["nelem"], seed))
    m.partition(64)
    m.upload()
    return m


class _Dev:
    """Device buffers through the HIP runtime the library itself links (libamdhip64, already
    loaded with it), so the test initialises no second runtime."""

    def __init__(self):
        import ctypes
        self.c = ctypes
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipFree.argtypes = [ctypes.c_void_p]
        self.h = h
        self.ptrs = []

    def alloc(self, nbytes):
        p = self.c.c_void_p()
        assert self.h.hipMalloc(self.c.byref(p), nbytes) == 0
        self.ptrs.append(p.value)
        return p.value

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = self.alloc(arr.nbytes)
        assert self.h.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, n, dtype):
        out = np.empty(n, dtype)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.ptrs:
            self.h.hipFree(p)
        self.ptrs = []


def _noalloc(E, m, seeds):
    dev = _Dev()
    try:
        sd = dev.put(np.asarray(seeds, np.int32))
        out = dev.put(np.full(m.nv, -1.0, np.float32))
        m.solve_noalloc(sd, len(seeds), out)
        return dev.get(out, m.nv, np.float32)  # synchronises the device first
    finally:
        dev.free()


def test_noalloc_matches_solve_two_calls_second_smaller(E):
    big = _mesh(E, 24, 5)
    small = _mesh(E, 12, 6)
    for m, n, seeds in ((big, 24, [center_vertex(24), 7]), (small, 12, [center_vertex(12)])):
        _say("allocation-free solve, %d^3 cube, seeds %s" % (n, seeds))
        T = _noalloc(E, m, seeds)
        ref, _ = m.solve(seeds)
        err = rel_err(T, ref, 1e-3 * float(ref.max()))
        _say("  max rel diff vs eik_solve %.3e" % err)
        assert err <= 1e-5
        assert np.all(T >= 0) and T[seeds[0]] == 0.0


def test_noalloc_matches_solve_on_triangles(E):
    """Triangle meshes take the same path with their own kernel (and a grid sized by the number
    of sources): an icosphere with a per-face speed field, one and three sources."""
    m = E.Mesh.icosphere(5)
    nf = m.info()["nelem"]
    m.set_speed(np.exp(np.random.default_rng(11).normal(0.0, 0.4, nf)).astype(np.float32))
    m.partition(64)
    m.upload()
    for seeds in ([0], [0, 1234, 9000]):
        _say("allocation-free solve, level-5 icosphere, seeds %s" % seeds)
        T = _noalloc(E, m, seeds)
        ref, _ = m.solve(seeds)
        err = rel_err(T, ref, 1e-3 * float(ref.max()))
        _say("  max rel diff vs eik_solve %.3e" % err)
        assert err <= 1e-5
        assert T[seeds[0]] == 0.0


def test_noalloc_twice_on_one_mesh_and_bad_seed_skipped(E):
    m = _mesh(E, 10, 7)
    c = center_vertex(10)
    a = _noalloc(E, m, [c])
    b = _noalloc(E, m, [c, -5, m.nv + 3])  # out-of-range ids are skipped
    np.testing.assert_allclose(a, b, rtol=2e-6, atol=1e-7)


def test_noalloc_rejects_meshes_beyond_its_array(E):
    m = E.Mesh.kuhn_cube(170, 0.0, seed=1)  # 5.0 M vertices: more state than the 64 MiB array
    m.partition(64)
    m.upload()
    dev = _Dev()
    try:
        sd = dev.put(np.zeros(1, np.int32))
        out = dev.alloc(4 * m.nv)
        with pytest.raises(E.EikError) as ei:
            m.solve_noalloc(sd, 1, out)
        assert ei.value.code == E.EIK_ERR_CAPACITY and "eik_solve" in str(ei.value)
    finally:
        dev.free()


def test_noalloc_needs_an_uploaded_mesh(E):
    m = E.Mesh.kuhn_cube(6, 0.2, seed=1)
    m.partition(64)
    dev = _Dev()
    try:
        sd = dev.put(np.zeros(1, np.int32))
        out = dev.alloc(4 * m.nv)
        with pytest.raises(E.EikError) as ei:
            m.solve_noalloc(sd, 1, out)
        assert ei.value.code == E.EIK_ERR_STATE
    finally:
        dev.free()
