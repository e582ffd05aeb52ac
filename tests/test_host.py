"""Host side of the library, CPU only: the C ABI loads and exports what the header declares,
mesh readers and generators, speed semantics, partitions, the block layout, and error codes.
No test here launches device code.

The layout is checked twice: structurally (every element sits in each block owning one of its
corners, local indices resolve to the right vertices, records carry the right squared metric
lengths) and by running the oracle's blocked FIM over the product's own layout, which must
reach the serial oracle's fixed point.
"""
import os

import numpy as np
import pytest

from conftest import center_vertex, golden, lognormal_mult

HAS_DEVICE = os.path.exists("/dev/kfd")


# ------------------------------------------------------------------------------------------
# the C ABI
# ------------------------------------------------------------------------------------------
def test_library_exports_every_header_symbol(E):
    L = E.lib()
    names = E.header_symbols()
    assert len(names) >= 25
    for n in names:
        getattr(L, n)  # AttributeError if the symbol is missing
    assert set(names) == set(E.SIGNATURES), set(names) ^ set(E.SIGNATURES)
    assert E.lib().eik_version().decode()


def test_solve_opts_defaults(E):
    o = E.default_opts()
    assert (o.scheduler, o.max_iterations, o.max_inner, o.eps, o.record_history,
            o.wait_limit_us) == (E.SCHED_ASYNC, 0, 0, 0.0, 0, 0)


def test_struct_mirrors_match_header(E, tmp_path):
    """The ctypes mirrors have the sizes and field offsets of the C structs, as a C compiler lays
    them out from include/eikonal_amd.h (a caller binding by layout sees the same)."""
    import ctypes
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    pairs = [("eik_mesh_info_t", E.MeshInfo), ("eik_layout_view_t", E.LayoutView),
             ("eik_solve_opts_t", E.SolveOpts), ("eik_solve_stats_t", E.SolveStats),
             ("eik_server_opts_t", E.ServerOpts), ("eik_request_info_t", E.RequestInfo),
             ("eik_tune_result_t", E.TuneResult)]
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "eikonal_amd.h"',
             'int main(void) {']
    for cname, py in pairs:
        lines.append('  printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for fname, _ in py._fields_:
            lines.append('  printf("%s %s %%zu\\n", offsetof(%s, %s));' % (cname, fname, cname, fname))
    lines += ['  return 0;', '}']
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in pairs:
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[(cname, fname)] == getattr(py, fname).offset, (cname, fname)


def test_negative_wait_limit_is_rejected_before_any_device_work(E):
    m = E.Mesh.square(8)
    m.partition(16)
    with pytest.raises(E.EikError) as ei:
        m.solve([0], wait_limit_us=-1)
    assert ei.value.code == E.EIK_ERR_ARG and "wait_limit_us" in str(ei.value)


# ------------------------------------------------------------------------------------------
# readers
# ------------------------------------------------------------------------------------------
@pytest.mark.parametrize("name", ["sphere8", "sphere13", "sphere334", "CubeMesh_size16"])
def test_tetgen_reader_matches_oracle_reader(E, O, name):
    m = E.Mesh.load_tetgen(golden(name))
    xyz, tets = O.read_tetgen(golden(name))
    np.testing.assert_allclose(m.vertices(), xyz, rtol=1e-6)
    np.testing.assert_array_equal(m.elements(), tets)
    assert m.info()["dim"] == 3


@pytest.mark.parametrize("name", ["sphere_12verts.ply", "sphere_74verts.ply",
                                  "sphere_1154verts.ply", "SquareMesh_size16.ply",
                                  "SquareMesh_size32.ply"])
def test_ply_reader_matches_oracle_reader(E, O, name):
    """SquareMesh_size16.ply numbers its vertices from 1: both readers re-base it."""
    m = E.Mesh.load_ply(golden(name))
    xyz, f = O.read_ply_ascii(golden(name))
    np.testing.assert_allclose(m.vertices(), xyz, rtol=1e-6)
    np.testing.assert_array_equal(m.elements(), f)
    assert f.min() == 0 and f.max() == xyz.shape[0] - 1


def test_missing_file_is_io_error(E):
    """The reference exits with code 1 here (Eikonal.cu:102-115, 3D-sanity.cc)."""
    for fn in (E.Mesh.load_tetgen, E.Mesh.load_ply):
        with pytest.raises(E.EikError) as ei:
            fn("notarealfile")
        assert ei.value.code == E.EIK_ERR_IO
        assert "notarealfile" in str(ei.value)


@pytest.mark.parametrize("node,ele", [
    ("4 3 0 0\n0 0 0 0\n1 1 0 0\n2 0 abc 0\n3 0 0 1\n", "1 4 0\n0 0 1 2 3\n"),  # garbage token
    ("4 3 0 0\n0 0 0 0\n1 1e50 0 0\n2 0 1 0\n3 0 0 1\n", "1 4 0\n0 0 1 2 3\n"),  # float overflow
    ("4 3 0 0\n0 0 0 0\n1 1 0 0\n2 0 1 0\n3 0 0 1\n", "1 4 0\n0 0 1 2 99999999999\n"),  # int overflow
    ("999999999999999 3 0 0\n0 0 0 0\n", "1 4 0\n0 0 1 2 3\n"),  # count too large to allocate
])
def test_malformed_tetgen_is_io_error(E, tmp_path, node, ele):
    """A malformed file is an EIK_ERR_IO return, never an exception through the C ABI (which
    would terminate the calling process)."""
    base = str(tmp_path / "bad")
    open(base + ".node", "w").write(node)
    open(base + ".ele", "w").write(ele)
    with pytest.raises(E.EikError) as ei:
        E.Mesh.load_tetgen(base)
    assert ei.value.code in (E.EIK_ERR_IO, E.EIK_ERR_ARG)


@pytest.mark.parametrize("body", [
    "ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\n"
    "property float z\nelement face 1\nproperty list uchar int vertex_indices\nend_header\n"
    "0 0 0\n1 zz 0\n0 1 0\n3 0 1 2\n",
    "ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\n"
    "property float z\nelement face 1\nproperty list uchar int vertex_indices\nend_header\n"
    "0 0 0\n1 1e60 0\n0 1 0\n3 0 1 2\n",
    "ply\nformat ascii 1.0\nelement vertex 99999999999999\nproperty float x\nproperty float y\n"
    "property float z\nend_header\n0 0 0\n",
])
def test_malformed_ply_is_io_error(E, tmp_path, body):
    path = str(tmp_path / "bad.ply")
    open(path, "w").write(body)
    with pytest.raises(E.EikError) as ei:
        E.Mesh.load_ply(path)
    assert ei.value.code in (E.EIK_ERR_IO, E.EIK_ERR_ARG)


# ------------------------------------------------------------------------------------------
# generators
# ------------------------------------------------------------------------------------------
def test_square_generator():
    import eikonal_amd as E
    n = 8
    m = E.Mesh.square(n)
    xyz, f = m.vertices(), m.elements()
    assert xyz.shape == ((n + 1) ** 2, 3) and f.shape == (2 * n * n, 3)
    j, i = np.divmod(np.arange((n + 1) ** 2), n + 1)
    np.testing.assert_allclose(xyz[:, 0], i / n, atol=1e-7)
    np.testing.assert_allclose(xyz[:, 1], j / n, atol=1e-7)
    # every cell is split along its (i,j)-(i+1,j+1) diagonal
    for q in range(0, len(f), 2):
        c = q // 2
        jj, ii = divmod(c, n)
        a, d = jj * (n + 1) + ii, (jj + 1) * (n + 1) + ii + 1
        assert a in f[q] and d in f[q] and a in f[q + 1] and d in f[q + 1]
    e1 = xyz[f[:, 1]] - xyz[f[:, 0]]
    e2 = xyz[f[:, 2]] - xyz[f[:, 0]]
    area = 0.5 * np.abs(np.cross(e1, e2)[:, 2])
    assert np.isclose(area.sum(), 1.0)


def _tet_volumes(xyz, t):
    x = xyz.astype(np.float64)
    return np.einsum("ij,ij->i", np.cross(x[t[:, 1]] - x[t[:, 0]], x[t[:, 2]] - x[t[:, 0]]),
                     x[t[:, 3]] - x[t[:, 0]]) / 6.0


def test_kuhn_cube_generator(E):
    n = 6
    m = E.Mesh.kuhn_cube(n, 0.2, seed=3)
    xyz, t = m.vertices(), m.elements()
    assert xyz.shape == ((n + 1) ** 3, 3) and t.shape == (6 * n ** 3, 4)
    g = np.arange(n + 1) / n
    k, j, i = np.meshgrid(g, g, g, indexing="ij")
    grid = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1)
    dev = np.abs(xyz - grid)
    boundary = np.any((grid == 0) | (grid == 1), axis=1)
    assert np.all(dev[boundary] < 1e-7)  # boundary vertices are not jittered (fp32 grid)
    assert np.all(dev <= 0.2 / n + 1e-7) and dev[~boundary].max() > 0.1 / n
    vol = np.abs(_tet_volumes(xyz, t))
    assert np.all(vol > 0) and np.isclose(vol.sum(), 1.0, rtol=1e-5)
    assert np.all(np.sign(_tet_volumes(xyz, t)) == np.sign(_tet_volumes(grid, t)))
    # every (i,j,k) cell holds 6 tets along its main diagonal
    assert np.all(np.isin(t[:, 0], [(c * (n + 1) + b) * (n + 1) + a for c in range(n)
                                    for b in range(n) for a in range(n)]))
    # seeded
    np.testing.assert_array_equal(E.Mesh.kuhn_cube(n, 0.2, seed=3).vertices(), xyz)
    assert not np.array_equal(E.Mesh.kuhn_cube(n, 0.2, seed=4).vertices(), xyz)


def test_kuhn_cube_centre_vertex(E):
    """The source of the 64^3 single solve: the centre vertex of the 64^3 cube is id 137,312 at (0.5, 0.5, 0.5)."""
    assert center_vertex(64) == 137312
    m = E.Mesh.kuhn_cube(4, 0.0)
    np.testing.assert_allclose(m.vertices()[center_vertex(4)], [0.5, 0.5, 0.5])


def test_icosphere_generator(E):
    for level in range(4):
        m = E.Mesh.icosphere(level)
        xyz, f = m.vertices(), m.elements()
        assert xyz.shape[0] == 10 * 4 ** level + 2 and f.shape[0] == 20 * 4 ** level
        np.testing.assert_allclose(np.linalg.norm(xyz, axis=1), 1.0, rtol=1e-6)
        # closed, consistently oriented outward: signed volume positive, every edge in 2 faces
        x = xyz.astype(np.float64)
        vol = np.einsum("ij,ij->i", np.cross(x[f[:, 0]], x[f[:, 1]]), x[f[:, 2]]).sum() / 6
        assert 0.6 * 4 / 3 * np.pi < vol <= 4 / 3 * np.pi
        e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
        _, cnt = np.unique(e, axis=0, return_counts=True)
        assert np.all(cnt == 2)


# ------------------------------------------------------------------------------------------
# speeds
# ------------------------------------------------------------------------------------------
PAIRS3 = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
PAIRS2 = [(0, 1), (1, 2), (2, 0)]


def _records_by_element(m):
    """{sorted global corner tuple: record L} over all blocks (3D: tets are unique)."""
    lay = m.layout()
    out = {}
    for b in range(lay["nblocks"]):
        vstart, nown, _, hstart, nhalo, estart, nelem, _ = lay["blocks"][b]
        loc = np.r_[lay["perm"][vstart:vstart + nown], lay["perm"][lay["halo"][hstart:hstart + nhalo]]]
        for r in lay["elems"][estart:estart + nelem]:
            g = loc[r["v"].astype(np.int64)]
            out.setdefault(tuple(g), r["L"].copy())
    return out


@pytest.mark.parametrize("kind", ["none", "scalar", "tensor", "wrong_count"])
def test_tet_speed_semantics(E, kind):
    """TetMesh::init (tetmesh.cu:54-79): 6*nt -> tensors, nt -> v*I, anything else -> I."""
    m = E.Mesh.kuhn_cube(3, 0.2, seed=1)
    xyz, t = m.vertices().astype(np.float64), m.elements()
    nt = t.shape[0]
    rng = np.random.default_rng(2)
    M = np.tile(np.eye(3), (nt, 1, 1))
    if kind == "scalar":
        v = rng.uniform(0.5, 2.0, nt).astype(np.float32)
        m.set_speed(v)
        M = M * v[:, None, None].astype(np.float64)
    elif kind == "tensor":
        A = rng.normal(size=(nt, 3, 3))
        S = A @ A.transpose(0, 2, 1) + 0.5 * np.eye(3)
        six = np.stack([S[:, 0, 0], S[:, 0, 1], S[:, 0, 2], S[:, 1, 1], S[:, 1, 2], S[:, 2, 2]],
                       axis=1).astype(np.float32)
        m.set_speed(six)
        M = S
    elif kind == "wrong_count":
        m.set_speed(np.ones(nt + 1, np.float32) * 7)
    m.partition(16)
    recs = _records_by_element(m)
    for e in range(nt):
        L = recs[tuple(t[e])]
        for p, (a, b) in enumerate(PAIRS3):
            d = xyz[t[e, b]] - xyz[t[e, a]]
            assert L[p] == pytest.approx(d @ M[e] @ d, rel=2e-6)


def test_tet_tensor_must_be_spd(E):
    m = E.Mesh.kuhn_cube(2, 0.0)
    nt = m.info()["nelem"]
    bad = np.tile(np.array([1, 2, 0, 1, 0, 1], np.float32), nt)  # [[1,2],[2,1]] is indefinite
    with pytest.raises(E.EikError) as ei:
        m.set_speed(bad)
    assert ei.value.code == E.EIK_ERR_ARG


def test_tri_speed_semantics(E):
    """TriMesh_stats.cu:56-69: nf speeds give speedInv = 1/v; squared lengths scale by 1/v^2."""
    m = E.Mesh.square(4)
    xyz, f = m.vertices().astype(np.float64), m.elements()
    v = np.linspace(0.5, 2.0, f.shape[0]).astype(np.float32)
    m.set_speed(v)
    m.partition_structured(5, 5, 1, 5, 5, 1)
    lay = m.layout()
    seen = 0
    for r in lay["elems"]:
        if r["mask"] != 7:
            continue
        loc = lay["perm"][r["v"].astype(np.int64)]
        e = np.nonzero(np.all(np.sort(f, axis=1) == np.sort(loc), axis=1))[0][0]
        for p, (a, b) in enumerate(PAIRS2):
            d = xyz[loc[b]] - xyz[loc[a]]
            assert r["L"][p] == pytest.approx((d @ d) / float(v[e]) ** 2, rel=2e-6)
        seen += 1
    assert seen == f.shape[0]


# ------------------------------------------------------------------------------------------
# partitions
# ------------------------------------------------------------------------------------------
def test_partition_block_counts(E):
    """ceil(nv/vpb) blocks for tets (meshFIM3dEikonal.cu:160), floor for triangles
    (meshFIM2d_eikonal.cu:119), at least 2."""
    m3 = E.Mesh.kuhn_cube(6, 0.2)
    nv = m3.nv
    for vpb in (16, 50, 64, 100, 10000):
        nb = m3.partition(vpb)
        assert nb == max(2, -(-nv // vpb))
        lab = m3.labels()
        assert lab.min() == 0 and lab.max() == nb - 1
        sizes = np.bincount(lab, minlength=nb)
        # every block within 15% of the mean size (cuts move to gaps between vertex planes)
        mean = nv / nb
        assert sizes.min() >= np.floor(0.85 * mean) - 1 and sizes.max() <= np.ceil(1.15 * mean) + 1
    m2 = E.Mesh.load_ply(golden("sphere_1154verts.ply"))
    for vpb in (32, 100, 5000):
        assert m2.partition(vpb) == max(2, m2.nv // vpb)


def test_partition_cuts_between_lattice_planes(E):
    """On a lattice whose block count is a power of two the bisection reproduces the perfect
    4^3 tiling (1.714 element copies per tet at 64 vertices per block), jitter or not, and on the
    65^3 cube (4,292 blocks, not a power of two) it stays at or under 1.75 copies (plain
    proportional cuts through jittered planes gave 1.92)."""
    m = E.Mesh.kuhn_cube(63, 0.2, seed=1)
    m.partition(64)
    i = m.info()
    assert abs(i["block_elems"] / i["nelem"] - 1.714) < 1e-3
    assert i["max_block_verts"] == 64
    m = E.Mesh.kuhn_cube(64, 0.2, seed=20240607)
    m.partition(64)
    i = m.info()
    assert i["nblocks"] == 4292 and i["block_elems"] / i["nelem"] <= 1.75


def test_partition_blocks_are_compact(E):
    """Recursive bisection gives spatially compact blocks: the halo is a small fraction."""
    m = E.Mesh.kuhn_cube(16, 0.2, seed=1)
    m.partition(64)
    info = m.info()
    assert info["block_elems"] < 2.2 * info["nelem"]


def test_structured_partition_formula(E):
    """GraphPartition_Square (meshFIM3dEikonal.cu:255-259): vertex id k*L*W + i*L + j goes to
    block (k/bd)*nbl*nbw + (i/bw)*nbl + j/bl."""
    m = E.Mesh.load_tetgen(golden("CubeMesh_size16"))
    L = W = D = 16
    for bl, bw, bd in ((4, 4, 4), (8, 4, 2), (16, 16, 16)):
        nb = m.partition_structured(L, W, D, bl, bw, bd)
        v = np.arange(m.nv)
        k, rem = np.divmod(v, L * W)
        i, j = np.divmod(rem, L)
        nbl, nbw = L // bl, W // bw
        want = (k // bd) * nbl * nbw + (i // bw) * nbl + j // bl
        np.testing.assert_array_equal(m.labels(), want)
        assert nb == want.max() + 1


def test_set_labels(E):
    m = E.Mesh.kuhn_cube(3, 0.2)
    lab = np.arange(m.nv) % 3
    m.set_labels(lab, 3)
    np.testing.assert_array_equal(m.labels(), lab)
    with pytest.raises(E.EikError) as ei:
        m.set_labels(np.full(m.nv, 5), 3)
    assert ei.value.code == E.EIK_ERR_ARG


# ------------------------------------------------------------------------------------------
# block layout
# ------------------------------------------------------------------------------------------
def _check_layout(m, elems_global, k):
    lay = m.layout()
    nv = lay["nv"]
    perm, iperm = lay["perm"], lay["iperm"]
    np.testing.assert_array_equal(perm[iperm], np.arange(nv))
    labels = m.labels()
    blocks = lay["blocks"]
    # owned ranges are contiguous, block after block, interior before boundary
    assert blocks[0, 0] == 0 and np.all(blocks[1:, 0] == blocks[:-1, 0] + blocks[:-1, 1])
    assert blocks[:, 1].sum() == nv
    touching = {}
    for e, c in enumerate(elems_global):
        for b in set(labels[c].tolist()):
            touching.setdefault(b, set()).add(tuple(sorted(c.tolist())))
    bnd_seen = np.zeros(nv, bool)
    for b in range(lay["nblocks"]):
        vstart, nown, nbnd0, hstart, nhalo, estart, nelem, _ = blocks[b]
        own = perm[vstart:vstart + nown]
        assert np.all(labels[own] == b)
        halo = perm[lay["halo"][hstart:hstart + nhalo]]
        assert np.all(labels[halo] != b) and len(set(halo.tolist())) == nhalo
        loc = np.r_[own, halo]
        got, allc = set(), set()
        for r in lay["elems"][estart:estart + nelem]:
            v = r["v"][:k].astype(np.int64)
            assert v.max() < nown + nhalo
            bnd_seen[loc[v][v >= nown]] = True
            allc.update(loc[v].tolist())
            if k == 3 and r["mask"] != 7:
                # a virtual face sits in its obtuse (first) vertex's block, updating only it
                assert r["mask"] == 1 and labels[loc[v[0]]] == b
                continue
            got.add(tuple(sorted(loc[v].tolist())))
        assert got == touching.get(b, set()), b
        # halo vertices are exactly the foreign corners of the block's elements
        foreign = {x for x in allc if labels[x] != b}
        assert foreign == set(halo.tolist())
    # boundary vertices (owned, index >= nbnd0) are exactly those some other block reads
    for b in range(lay["nblocks"]):
        vstart, nown, nbnd0 = blocks[b, :3]
        own = perm[vstart:vstart + nown]
        np.testing.assert_array_equal(bnd_seen[own], np.arange(nown) >= nbnd0)
    # dependents are the transpose of suppliers
    sup = {(c, int(s)) for c in range(lay["nblocks"])
           for s in lay["sup"][lay["sup_start"][c]:lay["sup_start"][c + 1]]}
    dep = {(int(c), b) for b in range(lay["nblocks"])
           for c in lay["dep"][lay["dep_start"][b]:lay["dep_start"][b + 1]]}
    assert sup == dep
    return lay


def test_layout_invariants_tets(E):
    m = E.Mesh.kuhn_cube(6, 0.2, seed=2)
    m.partition(20)
    _check_layout(m, m.elements(), 4)


def test_layout_invariants_triangles(E):
    m = E.Mesh.load_ply(golden("sphere_1154verts.ply"))
    m.partition(64)
    _check_layout(m, m.elements(), 3)


@pytest.mark.parametrize("kind", ["tet", "tri"])
def test_layout_is_deterministic(E, kind):
    """The layout is built on several host threads (per-element metrics, then per-block records
    and halos); two builds of the same partitioned mesh are byte-identical, and the records of a
    block do not depend on which thread built it."""
    def build():
        if kind == "tet":
            m = E.Mesh.kuhn_cube(18, 0.2, seed=4)
            m.partition(32)
        else:
            m = E.Mesh.icosphere(5)
            m.partition(48)
        return m.layout()
    a, b = build(), build()
    assert a.keys() == b.keys()
    for k in a:
        if hasattr(a[k], "tobytes"):
            assert a[k].tobytes() == b[k].tobytes(), k
        else:
            assert a[k] == b[k], k


def test_virtual_faces_match_oracle(E, O):
    """SplitFace / need_face_virtual_faces (TriMesh.h:191-313): the library's virtual faces and
    the oracle's are the same rows with the same originating faces."""
    for name in ("sphere_74verts.ply", "sphere_1154verts.ply", "sphere_4610verts.ply",
                 "SquareMesh_size32.ply"):
        m = E.Mesh.load_ply(golden(name))
        rows, origin = m.virtual_faces()
        xyz, f = O.read_ply_ascii(golden(name))
        orows, oorigin = O.tri_virtual_faces(xyz, f)
        a = sorted(zip(map(tuple, rows.tolist()), origin.tolist()))
        b = sorted(zip(map(tuple, orows.tolist()), oorigin.tolist()))
        assert a == b, name


@pytest.mark.parametrize("vpb", [8, 32, 100])
def test_blocked_oracle_over_layout_3d(E, O, vpb):
    """The oracle's blocked FIM run over the product's layout reaches the serial fixed point."""
    n = 7
    m = E.Mesh.kuhn_cube(n, 0.2, seed=vpb)
    nt = m.info()["nelem"]
    mult = lognormal_mult(nt, vpb)
    m.set_speed(mult)
    m.partition(vpb)
    lay = m.layout()
    c = center_vertex(n)
    with O.exact():
        Tb, it, visits = O.fim_blocked(lay, [c, 3])
    xyz, t = m.vertices().astype(np.float64), m.elements()
    # the serial oracle on the float32 metric lengths the layout carries: same fixed point
    with O.exact():
        Ts, _ = O.fim3d(xyz, t, [c, 3], M6=O.tet_metric(nt, mult.astype(np.float64)))
    np.testing.assert_allclose(Tb, Ts, rtol=2e-6)
    assert visits >= lay["nblocks"] // 2
    # the same schedule on several host threads (the multi-core CPU baseline) reaches the same
    # fixed point; each visit counts its block's owned vertices
    with O.exact():
        Tm, itm, vm, um = O.fim_blocked(lay, [c, 3], threads=4, counts=True)
    np.testing.assert_allclose(Tm, Ts, rtol=2e-6)
    assert vm >= lay["nblocks"] // 2 and um >= vm
    _, _, v1, u1 = O.fim_blocked(lay, [c, 3], max_inner=1, counts=True)
    assert u1 <= v1 * int(lay["blocks"][:, 1].max())


def test_blocked_fp32_cpu_baseline(E, O):
    """The FP32 blocked FIM (the multi-core CPU baseline: the oracle's restatement of the
    reference's local solver at FP32) solves the same problem: close to the exact FP64 fixed
    point (the reference's quadratic in FP32 loses a few digits to cancellation), identical on
    one and on several host threads up to the last bits, and every value above the exact
    minimum's rounding (FP32 can only miss minima of the same candidate set)."""
    n = 8
    m = E.Mesh.kuhn_cube(n, 0.2, seed=5)
    nt = m.info()["nelem"]
    mult = lognormal_mult(nt, 5)
    m.set_speed(mult)
    m.partition(32)
    lay = m.layout()
    c = center_vertex(n)
    T1, it1, v1 = O.fim_blocked(lay, [c], max_inner=2, fp32=True)
    T4, it4, v4 = O.fim_blocked(lay, [c], max_inner=2, threads=4, fp32=True)
    assert T1.dtype == np.float32
    xyz, t = m.vertices().astype(np.float64), m.elements()
    with O.exact():
        Te, _ = O.fim3d(xyz, t, [c], M6=O.tet_metric(nt, mult.astype(np.float64)))
    tau = 0.1 / n
    assert np.max(np.abs(T1 - Te) / np.maximum(Te, tau)) < 5e-3
    assert np.max(np.abs(T4 - Te) / np.maximum(Te, tau)) < 5e-3
    assert T1[c] == 0.0 and np.all(T1 < 1e8) and v1 >= lay["nblocks"]


def test_blocked_oracle_over_layout_2d_virtual_faces(E, O):
    m = E.Mesh.load_ply(golden("sphere_4610verts.ply"))
    m.partition(64)
    lay = m.layout()
    Tb, it, visits = O.fim_blocked(lay, [0])
    xyz, f = m.vertices(), m.elements()
    vf, origin = O.tri_virtual_faces(xyz, f)
    F = np.ones(f.shape[0])
    Ts, _ = O.fim2d(xyz.astype(np.float64), f, [0], F, 0, vfaces=vf, vparam=F[origin])
    np.testing.assert_allclose(Tb, Ts, rtol=2e-6)


# ------------------------------------------------------------------------------------------
# errors
# ------------------------------------------------------------------------------------------
def test_bad_arguments(E):
    with pytest.raises(E.EikError) as ei:
        E.Mesh.tet(np.zeros((3, 3)), np.array([[0, 1, 2, 3]]))
    assert ei.value.code == E.EIK_ERR_ARG
    with pytest.raises(E.EikError) as ei:
        E.Mesh.tri(np.zeros((3, 3)), np.array([[0, 1, 5]]))
    assert ei.value.code == E.EIK_ERR_ARG
    m = E.Mesh.kuhn_cube(2, 0.0)
    with pytest.raises(E.EikError) as ei:
        m.solve([0])  # not partitioned
    assert ei.value.code == E.EIK_ERR_STATE
    m.partition(8)
    with pytest.raises(E.EikError) as ei:
        m.solve([m.nv])  # seed out of range
    assert ei.value.code == E.EIK_ERR_ARG
    with pytest.raises(E.EikError) as ei:
        m.solve([0], seed_vals=[-1.0])
    assert ei.value.code == E.EIK_ERR_ARG
    with pytest.raises(E.EikError) as ei:
        m.solve([0], scheduler=1, record_history=1)
    assert ei.value.code == E.EIK_ERR_ARG


@pytest.mark.skipif(HAS_DEVICE, reason="checks the no-device error path")
def test_solve_without_device_fails_loudly(E):
    """No silent CPU fallback: without a HIP device a solve is an error."""
    m = E.Mesh.kuhn_cube(2, 0.0)
    m.partition(8)
    with pytest.raises(E.EikError) as ei:
        m.solve([0])
    assert ei.value.code == E.EIK_ERR_HIP


def test_capacity_error_for_huge_blocks(E):
    """Local vertex ids and work items are 16-bit: a block with more than 65535 local vertices
    or more than 16383 elements is refused."""
    for n in (260, 100):  # 68121 vertices; 10201 vertices but 20000 faces
        m = E.Mesh.square(n)
        with pytest.raises(E.EikError) as ei:
            m.set_labels(np.zeros(m.nv, np.int32), 1)
            m.layout()
        assert ei.value.code == E.EIK_ERR_CAPACITY
    m = E.Mesh.square(50)  # 5000 faces (and their virtual faces) in one block: accepted
    m.set_labels(np.zeros(m.nv, np.int32), 1)
    assert m.layout()["nblocks"] == 1


# ------------------------------------------------------------------------------------------
# corner heights (the pruning bound of the block-visit kernels)
# ------------------------------------------------------------------------------------------
def test_record_heights_are_tight_lower_bounds(E, O):
    """H[k] = metric distance from corner k to the plane (line) of the opposite face (edge),
    rounded down; and no local solve of corner k goes below min(other times) + H[k]."""
    m = E.Mesh.kuhn_cube(5, 0.2, seed=4)
    nt = m.info()["nelem"]
    mult = lognormal_mult(nt, 4)
    m.set_speed(mult)
    m.partition(16)
    xyz, t = m.vertices().astype(np.float64), m.elements()
    recs = {}
    lay = m.layout()
    for b in range(lay["nblocks"]):
        vstart, nown, _, hstart, nhalo, estart, nelem, _ = lay["blocks"][b]
        loc = np.r_[lay["perm"][vstart:vstart + nown], lay["perm"][lay["halo"][hstart:hstart + nhalo]]]
        for r in lay["elems"][estart:estart + nelem]:
            recs.setdefault(tuple(loc[r["v"].astype(np.int64)]), (r["L"].copy(), r["H"].copy()))
    rng = np.random.default_rng(5)
    worst = 0.0
    with O.exact():
        for e in range(nt):
            L, H = recs[tuple(t[e])]
            for k in range(4):
                o = [j for j in range(4) if j != k]
                A, B, C = xyz[t[e, o[0]]], xyz[t[e, o[1]]], xyz[t[e, o[2]]]
                n = np.cross(B - A, C - A)
                n /= np.linalg.norm(n)
                h = abs((xyz[t[e, k]] - A) @ n) * np.sqrt(mult[e])
                assert H[k] <= h * (1 + 1e-6) and H[k] >= h * (1 - 1e-4)
            P = O.tet_products_from_len(L.astype(np.float64))
            T4 = rng.uniform(0.0, 0.2, 4)
            for k in range(4):
                cand = O.tet_corner(T4, P, k)
                bound = np.min(np.delete(T4, k)) + H[k]
                worst = max(worst, bound - cand)
    assert worst <= 1e-12


def test_triangle_heights(E):
    m = E.Mesh.load_ply(golden("sphere_74verts.ply"))
    m.partition(16)
    lay = m.layout()
    xyz = m.vertices().astype(np.float64)
    for r in lay["elems"][:200]:
        if r["mask"] != 7:
            continue
        L = r["L"].astype(np.float64)
        for k in range(3):
            a, b = (k + 1) % 3, (k + 2) % 3
            la = [L[0], L[1], L[2]]
            # {01, 12, 20}: squared lengths from k to a, k to b, a to b
            l = {(0, 1): la[0], (1, 2): la[1], (0, 2): la[2]}
            ka, kb, ab = l[tuple(sorted((k, a)))], l[tuple(sorted((k, b)))], l[tuple(sorted((a, b)))]
            area2 = ka * kb - (0.5 * (ka + kb - ab)) ** 2
            h = np.sqrt(area2 / ab)
            assert r["H"][k] <= h * (1 + 1e-6) and r["H"][k] >= h * (1 - 1e-4)


def test_info_reports_partition_quality_at_once(E):
    """The block statistics are there right after the partition call (the layout is built with
    it) and survive a speed change, which only invalidates the element records."""
    m = E.Mesh.kuhn_cube(16, 0.2, seed=1)
    nb = m.partition(64)
    i = m.info()
    assert i["nblocks"] == nb == -(-17 ** 3 // 64)
    assert i["block_elems"] > i["nelem"]
    assert 0 < i["max_block_verts"] <= 74 < i["max_block_local"]
    assert i["block_halo"] > 0
    m.set_speed(np.full(i["nelem"], 2.0, np.float32))
    j = m.info()
    assert (j["block_elems"], j["block_halo"], j["max_block_elems"]) == \
        (i["block_elems"], i["block_halo"], i["max_block_elems"])
