"""ctypes front-end of the FP64 oracle (oracle/fim_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg as the reference to compare with; the product never imports it.

Besides the C entry points this module restates, in numpy, the host-side conventions the
solve depends on:
  * speed / metric multipliers -> per-tet 3x3 metric (src/core3d/cuda/tetmesh.cu:54-79):
    6*nt values = upper-triangular tensors [0 1 2; 1 3 4; 2 4 5], nt values = diag(v, v, v),
    anything else = identity.  Edge cost is sqrt(e^T M e), so a scalar v means 1/speed^2.
  * 2D face speeds -> speedInv = 1/speed (src/core2d/cuda/TriMesh_stats.cu:64-68);
    a size mismatch means speedInv = 1.
  * mesh fixtures: TetGen .node/.ele (src/core3d/cuda/tetgen.cu:1993-2320, 0- or 1-based indices
    per src/core3d/cuda/tetmesh.cu:27-52) and ASCII PLY (src/core2d/cuda/TriMesh_io.cu:268-510).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libfim_oracle.so")
_lib = None

LARGENUM_3D = 1e9
LARGENUM_2D = 1e7

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_fp = ctypes.POINTER(ctypes.c_float)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build():
    """Compile the oracle with gcc (oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        d = ctypes.c_double
        L.orc_tet_local.restype = d
        L.orc_tet_local.argtypes = [d] * 9
        L.orc_tet_products.restype = None
        L.orc_tet_products.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_tet_corner.restype = d
        L.orc_tet_corner.argtypes = [_dp, _dp, ctypes.c_int]
        L.orc_tri_iso_corner.restype = d
        L.orc_tri_iso_corner.argtypes = [d] * 7
        L.orc_tri_aniso_corner.restype = d
        L.orc_tri_aniso_corner.argtypes = [d] * 5
        L.orc_tri_aniso_vertex.restype = d
        L.orc_tri_aniso_vertex.argtypes = [_dp, _dp, ctypes.c_int]
        L.orc_fim3d.restype = ctypes.c_int64
        L.orc_fim3d.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _ip, _dp, ctypes.c_int, _ip, _dp,
                                _dp, _i64p]
        L.orc_fim3d_sweep.restype = ctypes.c_int64
        L.orc_fim3d_sweep.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _ip, _dp, ctypes.c_int,
                                      _ip, _dp, _dp, ctypes.c_int]
        L.orc_fim2d.restype = ctypes.c_int64
        L.orc_fim2d.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _ip, _dp, ctypes.c_int,
                                ctypes.c_int, _ip, _dp, ctypes.c_int, _ip, _dp, _dp, _i64p]
        L.orc_fim2d_sweep.restype = ctypes.c_int64
        L.orc_fim2d_sweep.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _ip, _dp, ctypes.c_int,
                                      ctypes.c_int, _ip, _dp, ctypes.c_int, _ip, _dp, _dp,
                                      ctypes.c_int]
        L.orc_fim_blocked.restype = ctypes.c_int64
        L.orc_fim_blocked.argtypes = [ctypes.c_int, ctypes.c_int, _ip, ctypes.c_void_p, _ip, _ip,
                                      _ip, _ip, ctypes.c_int64, ctypes.c_int, _ip, _dp, _dp,
                                      ctypes.c_int, ctypes.c_int, _i64p, _i64p]
        L.orc_fim_blocked_f32.restype = ctypes.c_int64
        L.orc_fim_blocked_f32.argtypes = [ctypes.c_int, ctypes.c_int, _ip, ctypes.c_void_p, _ip,
                                          _ip, _ip, _ip, ctypes.c_int64, ctypes.c_int, _ip, _fp,
                                          _fp, ctypes.c_int, ctypes.c_int, _i64p, _i64p]
        L.orc_tet_products_from_len.restype = None
        L.orc_tet_products_from_len.argtypes = [_dp, _dp]
        L.orc_set_interior_reg.restype = None
        L.orc_set_interior_reg.argtypes = [ctypes.c_double]
        L.orc_get_interior_reg.restype = ctypes.c_double
        L.orc_get_interior_reg.argtypes = []
        L.orc_is_non_obtuse.restype = ctypes.c_int
        L.orc_is_non_obtuse.argtypes = [_fp, _ip, ctypes.c_int]
        L.orc_tri_across_edge.restype = None
        L.orc_tri_across_edge.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _ip]
        L.orc_tri_virtual_faces.restype = ctypes.c_int
        L.orc_tri_virtual_faces.argtypes = [ctypes.c_int, _fp, ctypes.c_int, _ip, ctypes.c_int,
                                            _ip, _ip]
        _lib = L
    return _lib


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


# ----------------------------------------------------------------------------------------------
# local solvers
# ----------------------------------------------------------------------------------------------
SMALLNUM = 1e-8


class interior_reg:
    """Context manager selecting the interior regularisation of localSolverTet1.

    reg = 1e-8 (default) is the reference's SMALLNUM (CUDA_kernels.cu:200); reg = 0 is the
    unregularised stationary point, i.e. the exact Hopf-Lax minimum."""

    def __init__(self, reg):
        self.reg = reg

    def __enter__(self):
        self.old = lib().orc_get_interior_reg()
        lib().orc_set_interior_reg(self.reg)

    def __exit__(self, *a):
        lib().orc_set_interior_reg(self.old)


def exact():
    return interior_reg(0.0)

def tet_local(TA, TB, TC, ACAC, ACBC, ACCD, BCBC, BCCD, CDCD):
    return lib().orc_tet_local(TA, TB, TC, ACAC, ACBC, ACCD, BCBC, BCCD, CDCD)


def tet_products(A, B, C, D, M6=(1, 0, 0, 1, 0, 1)):
    out = np.zeros(11)
    keep = [_d(A), _d(B), _d(C), _d(D), _d(M6)]
    lib().orc_tet_products(*[k[1] for k in keep], out.ctypes.data_as(_dp))
    return out


def tet_corner(T4, P11, k):
    T4, pT = _d(T4)
    P11, pP = _d(P11)
    return lib().orc_tet_corner(pT, pP, int(k))


def tri_iso_corner(TA, TB, TC, LenAB, LenBC, LenAC, F):
    return lib().orc_tri_iso_corner(TA, TB, TC, LenAB, LenBC, LenAC, F)


def tri_aniso_corner(TX, TY, XYXY, XYYZ, YZYZ):
    return lib().orc_tri_aniso_corner(TX, TY, XYXY, XYYZ, YZYZ)


def tri_aniso_vertex(T3, G3, k):
    T3, pT = _d(T3)
    G3, pG = _d(G3)
    return lib().orc_tri_aniso_vertex(pT, pG, int(k))


# ----------------------------------------------------------------------------------------------
# host-side conventions
# ----------------------------------------------------------------------------------------------
def tet_metric(nt, values):
    """Per-tet 6-entry metric from the reference's speedMtx vector (tetmesh.cu:54-79)."""
    values = np.asarray(values, dtype=np.float64).ravel() if values is not None else np.zeros(0)
    M = np.zeros((nt, 6))
    if values.size == 6 * nt and nt > 0:
        M[:] = values.reshape(nt, 6)
    elif values.size == nt and nt > 0:
        M[:, 0] = M[:, 3] = M[:, 5] = values
    else:
        M[:, 0] = M[:, 3] = M[:, 5] = 1.0
    return M


def face_speed_inv(nf, speeds):
    """speedInv per face (TriMesh_stats.cu:64-68)."""
    speeds = np.asarray(speeds, dtype=np.float64).ravel() if speeds is not None else np.zeros(0)
    if speeds.size != nf or nf == 0:
        return np.ones(nf)
    return 1.0 / speeds


def tri_gram(xyz, faces, metric=None, speed_inv=None, frame_faces=None):
    """Gram products {e1.e1, e1.e2, e2.e2} (e1 = v1-v0, e2 = v2-v0) per face.

    metric: None, or (nf, 3) values {m11, m12, m22} of a 2x2 SPD tensor in the face frame
    t1 = unit(v1 - v0), t2 = unit(n x t1), n = unit((v1-v0) x (v2-v0)); the edges are projected
    onto that frame.  frame_faces (optional, one row per face) takes the frame from other faces:
    a virtual face carries the tensor of the face it was split from, in that face's frame.
    speed_inv: optional isotropic factor F per face (metric F^2 I)."""
    xyz = np.asarray(xyz, dtype=np.float64)
    f = np.asarray(faces, dtype=np.int64)
    e1 = xyz[f[:, 1]] - xyz[f[:, 0]]
    e2 = xyz[f[:, 2]] - xyz[f[:, 0]]
    if metric is None:
        g = np.stack([np.einsum("ij,ij->i", e1, e1), np.einsum("ij,ij->i", e1, e2),
                      np.einsum("ij,ij->i", e2, e2)], axis=1)
    else:
        metric = np.asarray(metric, dtype=np.float64).reshape(-1, 3)
        ff = f if frame_faces is None else np.asarray(frame_faces, dtype=np.int64)
        f1 = xyz[ff[:, 1]] - xyz[ff[:, 0]]
        f2 = xyz[ff[:, 2]] - xyz[ff[:, 0]]
        n = np.cross(f1, f2)
        n /= np.linalg.norm(n, axis=1, keepdims=True)
        t1 = f1 / np.linalg.norm(f1, axis=1, keepdims=True)
        t2 = np.cross(n, t1)
        a = np.stack([np.einsum("ij,ij->i", e1, t1), np.einsum("ij,ij->i", e1, t2)], axis=1)
        b = np.stack([np.einsum("ij,ij->i", e2, t1), np.einsum("ij,ij->i", e2, t2)], axis=1)

        def q(u, v):
            return (u[:, 0] * v[:, 0] * metric[:, 0] + (u[:, 0] * v[:, 1] + u[:, 1] * v[:, 0]) *
                    metric[:, 1] + u[:, 1] * v[:, 1] * metric[:, 2])
        g = np.stack([q(a, a), q(a, b), q(b, b)], axis=1)
    if speed_inv is not None:
        g = g * (np.asarray(speed_inv, dtype=np.float64) ** 2)[:, None]
    return g


# ----------------------------------------------------------------------------------------------
# FIM drivers
# ----------------------------------------------------------------------------------------------
def fim3d(xyz, tets, seeds, M6=None, seed_vals=None, sweep=False, max_sweeps=100000):
    """Serial FP64 FIM on a tet mesh. Returns (T, stats dict)."""
    xyz, pxyz = _d(xyz)
    tets, ptets = _i(tets)
    seeds, pseeds = _i(np.atleast_1d(seeds))
    nv, nt = xyz.shape[0], tets.shape[0]
    pM = None
    if M6 is not None:
        M6, pM = _d(np.asarray(M6).reshape(nt, 6))
    psv = None
    if seed_vals is not None:
        seed_vals, psv = _d(np.atleast_1d(seed_vals))
    T = np.zeros(nv)
    ns = ctypes.c_int64(0)
    if sweep:
        n = lib().orc_fim3d_sweep(nv, pxyz, nt, ptets, pM, len(seeds), pseeds, psv,
                                  T.ctypes.data_as(_dp), max_sweeps)
    else:
        n = lib().orc_fim3d(nv, pxyz, nt, ptets, pM, len(seeds), pseeds, psv,
                            T.ctypes.data_as(_dp), ctypes.byref(ns))
    if n < 0:
        raise ValueError("oracle fim3d failed (%d)" % n)
    return T, {"updates": int(n), "solves": int(ns.value)}


def fim2d(xyz, faces, seeds, param, mode, vfaces=None, vparam=None, seed_vals=None,
          sweep=False, max_sweeps=100000):
    """Serial FP64 FIM on a triangle mesh.

    mode 0: isotropic reference solve, param = speedInv per face.
    mode 1: anisotropic, param = Gram products (nf, 3) (see tri_gram)."""
    xyz, pxyz = _d(xyz)
    faces, pf = _i(faces)
    nv, nf = xyz.shape[0], faces.shape[0]
    np_ = 1 if mode == 0 else 3
    param, pp = _d(np.asarray(param, dtype=np.float64).reshape(nf * np_))
    if vfaces is None or len(vfaces) == 0:
        vfaces = np.zeros((0, 3), np.int32)
        vparam = np.zeros(0)
    vfaces, pvf = _i(np.asarray(vfaces).reshape(-1, 3))
    nvf = vfaces.shape[0]
    vparam, pvp = _d(np.asarray(vparam, dtype=np.float64).reshape(nvf * np_ if nvf else 0))
    seeds, pseeds = _i(np.atleast_1d(seeds))
    psv = None
    if seed_vals is not None:
        seed_vals, psv = _d(np.atleast_1d(seed_vals))
    T = np.zeros(nv)
    ns = ctypes.c_int64(0)
    if sweep:
        n = lib().orc_fim2d_sweep(nv, pxyz, nf, pf, pp, mode, nvf, pvf, pvp, len(seeds), pseeds,
                                  psv, T.ctypes.data_as(_dp), max_sweeps)
    else:
        n = lib().orc_fim2d(nv, pxyz, nf, pf, pp, mode, nvf, pvf, pvp, len(seeds), pseeds, psv,
                            T.ctypes.data_as(_dp), ctypes.byref(ns))
    if n < 0:
        raise ValueError("oracle fim2d failed (%d)" % n)
    return T, {"updates": int(n), "solves": int(ns.value)}


def fim_blocked(layout, seeds, seed_vals=None, max_inner=1000, threads=1, counts=False,
                fp32=False):
    """Blocked FIM over a product block layout (see eikonal_amd.Mesh.layout()), FP64 or, with
    fp32=True, the same formulas in FP32 (the multi-core CPU baseline).

    seeds are caller vertex ids; returns (T in caller order, outer iterations, block visits),
    plus the vertex-update count when counts=True.  threads > 1 visits each outer iteration's
    active blocks concurrently on that many host threads.  Unreached vertices read
    LARGENUM_3D."""
    nb = layout["nblocks"]
    blocks, pb = _i(layout["blocks"].ravel())
    recs = np.ascontiguousarray(layout["elems"])
    halo, ph = _i(layout["halo"])
    sup_start, pss = _i(layout["sup_start"])
    sup, ps = _i(layout["sup"] if layout["sup"].size else np.zeros(1, np.int32))
    iperm = layout["iperm"]
    owner = np.zeros(layout["nv"], np.int32)
    for b in range(nb):
        owner[blocks[8 * b]: blocks[8 * b] + blocks[8 * b + 1]] = b
    owner, po = _i(owner)
    seeds_new, psn = _i(iperm[np.atleast_1d(seeds)])
    psv = None
    visits = ctypes.c_int64(0)
    updates = ctypes.c_int64(0)
    if fp32:
        if seed_vals is not None:
            seed_vals = np.ascontiguousarray(np.atleast_1d(seed_vals), dtype=np.float32)
            psv = seed_vals.ctypes.data_as(_fp)
        Tn = np.zeros(layout["nv"], np.float32)
        it = lib().orc_fim_blocked_f32(layout["dim"], nb, pb, recs.ctypes.data_as(ctypes.c_void_p),
                                       ph, pss, ps, po, layout["nv"], len(seeds_new), psn, psv,
                                       Tn.ctypes.data_as(_fp), max_inner, int(threads),
                                       ctypes.byref(visits), ctypes.byref(updates))
    else:
        if seed_vals is not None:
            seed_vals, psv = _d(np.atleast_1d(seed_vals))
        Tn = np.zeros(layout["nv"])
        it = lib().orc_fim_blocked(layout["dim"], nb, pb, recs.ctypes.data_as(ctypes.c_void_p), ph,
                                   pss, ps, po, layout["nv"], len(seeds_new), psn, psv,
                                   Tn.ctypes.data_as(_dp), max_inner, int(threads),
                                   ctypes.byref(visits), ctypes.byref(updates))
    if counts:
        return Tn[iperm], int(it), int(visits.value), int(updates.value)
    return Tn[iperm], int(it), int(visits.value)


def tet_products_from_len(L6):
    L6, pL = _d(L6)
    out = np.zeros(11)
    lib().orc_tet_products_from_len(pL, out.ctypes.data_as(_dp))
    return out


def tri_virtual_faces(xyz, faces, cap=None):
    """Virtual faces (rows with the obtuse vertex first) and their originating faces."""
    xyz = np.ascontiguousarray(xyz, dtype=np.float32)
    faces, pf = _i(faces)
    nf = faces.shape[0]
    cap = cap if cap is not None else 16 * nf + 16
    out = np.zeros((cap, 3), np.int32)
    origin = np.zeros(cap, np.int32)
    n = lib().orc_tri_virtual_faces(xyz.shape[0], xyz.ctypes.data_as(_fp), nf, pf, cap,
                                    out.ctypes.data_as(_ip), origin.ctypes.data_as(_ip))
    if n < 0:
        raise ValueError("virtual face capacity exceeded")
    return out[:n].copy(), origin[:n].copy()


def tri_across_edge(nv, faces):
    faces, pf = _i(faces)
    across = np.zeros((faces.shape[0], 3), np.int32)
    lib().orc_tri_across_edge(nv, faces.shape[0], pf, across.ctypes.data_as(_ip))
    return across


# ----------------------------------------------------------------------------------------------
# fixture readers (independent of the product's C++ readers)
# ----------------------------------------------------------------------------------------------
def _numbers(path):
    rows = []
    with open(path) as fh:
        for line in fh:
            line = line.split("#", 1)[0].strip()
            if line:
                rows.append(line.split())
    return rows


def read_tetgen(base):
    """TetGen .node/.ele pair -> (xyz float64 (nv,3), tets int32 (nt,4)), 0-based."""
    nodes = _numbers(base + ".node")
    nv = int(nodes[0][0])
    xyz = np.array([[float(x) for x in r[1:4]] for r in nodes[1:1 + nv]], dtype=np.float64)
    first_id = int(nodes[1][0])
    eles = _numbers(base + ".ele")
    nt = int(eles[0][0])
    tets = np.array([[int(x) for x in r[1:5]] for r in eles[1:1 + nt]], dtype=np.int64)
    # the reference's TetMesh::init re-bases on the minimum index (tetmesh.cu:27-52)
    if tets.size and tets.min() == 1:
        tets -= 1
    del first_id
    return xyz, tets.astype(np.int32)


def read_ply_ascii(path):
    """ASCII PLY with vertex x y z and a face vertex_indices list -> (xyz, faces)."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    nv = nf = 0
    i = 0
    while not lines[i].startswith("end_header"):
        tok = lines[i].split()
        if tok[:2] == ["element", "vertex"]:
            nv = int(tok[2])
        elif tok[:2] == ["element", "face"]:
            nf = int(tok[2])
        i += 1
    body = [l.split() for l in lines[i + 1:] if l.strip()]
    xyz = np.array([[float(x) for x in r[:3]] for r in body[:nv]], dtype=np.float64)
    faces = []
    for r in body[nv:nv + nf]:
        k = int(r[0])
        idx = [int(x) for x in r[1:1 + k]]
        for j in range(1, k - 1):  # fan-triangulate polygons
            faces.append([idx[0], idx[j], idx[j + 1]])
    faces = np.array(faces, dtype=np.int64)
    # indices running from an offset (e.g. 1..nv) are re-based (TriMesh_io.cu:1187-1219)
    lo, hi = int(faces.min()), int(faces.max())
    if not (lo == 0 and hi == nv - 1) and hi - lo == nv - 1:
        faces -= lo
    return xyz, faces.astype(np.int32)
