"""Python binding of the MI355X FIM Eikonal solver (ctypes over libeikonal_amd.so).

The C ABI is declared in include/eikonal_amd.h; this module mirrors it for tests and the
benchmark.  The solve runs only through the HIP library: if the library is missing the import
fails, and a solve without a GPU fails with the HIP error the library reports.

The package directory name contains a hyphen, so load this file by path:
    spec = importlib.util.spec_from_file_location("eikonal_amd", ".../eikonal_amd.py")
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libeikonal_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "eikonal_amd.h")

EIK_OK = 0
EIK_ERR_ARG = 1
EIK_ERR_IO = 2
EIK_ERR_STATE = 3
EIK_ERR_HIP = 4
EIK_ERR_CAPACITY = 5
EIK_ERR_NOMEM = 6
EIK_ERR_TIMEOUT = 7
LARGENUM_3D = 1.0e9
LARGENUM_2D = 1.0e7
SCHED_SYNC = 0
SCHED_ASYNC = 1
STAGE_AUTO = 0
STAGE_LDS = 1
STAGE_MEMORY = 2


class EikError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("eikonal_amd error %d: %s" % (code, msg))
        self.code = code


class MeshInfo(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("nblocks", ctypes.c_int32), ("nv", ctypes.c_int64),
                ("nelem", ctypes.c_int64), ("nvirtual", ctypes.c_int64),
                ("block_elems", ctypes.c_int64), ("max_block_verts", ctypes.c_int32),
                ("max_block_local", ctypes.c_int32), ("max_block_elems", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("block_halo", ctypes.c_int64)]


class LayoutView(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("nblocks", ctypes.c_int32), ("nv", ctypes.c_int64),
                ("nelems", ctypes.c_int64), ("nhalo", ctypes.c_int64),
                ("rec_bytes", ctypes.c_int32), ("max_local", ctypes.c_int32),
                ("perm", ctypes.POINTER(ctypes.c_int32)), ("iperm", ctypes.POINTER(ctypes.c_int32)),
                ("blocks", ctypes.POINTER(ctypes.c_int32)), ("halo", ctypes.POINTER(ctypes.c_int32)),
                ("elems", ctypes.c_void_p), ("sup_start", ctypes.POINTER(ctypes.c_int32)),
                ("sup", ctypes.POINTER(ctypes.c_int32)),
                ("dep_start", ctypes.POINTER(ctypes.c_int32)),
                ("dep", ctypes.POINTER(ctypes.c_int32))]


TET_REC = np.dtype([("v", np.uint16, 4), ("L", np.float32, 6), ("H", np.float32, 4)])
TRI_REC = np.dtype([("v", np.uint16, 3), ("mask", np.uint16), ("L", np.float32, 3),
                    ("H", np.float32, 3)])


class SolveOpts(ctypes.Structure):
    _fields_ = [("scheduler", ctypes.c_int32), ("max_iterations", ctypes.c_int32),
                ("max_inner", ctypes.c_int32), ("eps", ctypes.c_float),
                ("record_history", ctypes.c_int32), ("workgroups", ctypes.c_int32),
                ("stage_records", ctypes.c_int32), ("wait_limit_us", ctypes.c_int32)]


class SolveStats(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("wall_ms", ctypes.c_double),
                ("outer_iterations", ctypes.c_int64), ("block_visits", ctypes.c_int64),
                ("vertex_updates", ctypes.c_int64), ("inner_sweeps", ctypes.c_int64),
                ("local_solves", ctypes.c_int64), ("algo_bytes", ctypes.c_int64),
                ("history_len", ctypes.c_int64), ("visit_ms", ctypes.c_double),
                ("visit_launches", ctypes.c_int64), ("busy_ms", ctypes.c_double),
                ("wait_ms", ctypes.c_double), ("load_ms", ctypes.c_double),
                ("sweep_ms", ctypes.c_double), ("writeback_ms", ctypes.c_double),
                ("queue_ms", ctypes.c_double), ("reruns", ctypes.c_int64),
                ("abandoned_waits", ctypes.c_int64), ("algo_bytes_loaded", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_vp = ctypes.c_void_p
_h = ctypes.c_void_p
_hp = ctypes.POINTER(ctypes.c_void_p)
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64

# name -> (restype, argtypes); every symbol the header declares
SIGNATURES = {
    "eik_last_error": (ctypes.c_char_p, []),
    "eik_version": (ctypes.c_char_p, []),
    "eik_mesh_create_tet": (ctypes.c_int, [_fp, _i64, _ip, _i64, _hp]),
    "eik_mesh_create_tri": (ctypes.c_int, [_fp, _i64, _ip, _i64, _hp]),
    "eik_mesh_load_tetgen": (ctypes.c_int, [ctypes.c_char_p, _hp]),
    "eik_mesh_load_ply": (ctypes.c_int, [ctypes.c_char_p, _hp]),
    "eik_mesh_destroy": (ctypes.c_int, [_h]),
    "eik_mesh_info": (ctypes.c_int, [_h, ctypes.POINTER(MeshInfo)]),
    "eik_mesh_get_vertices": (ctypes.c_int, [_h, _fp]),
    "eik_mesh_get_elements": (ctypes.c_int, [_h, _ip]),
    "eik_mesh_stream": (ctypes.c_int, [_h, _hp]),
    "eik_mesh_set_speed": (ctypes.c_int, [_h, _fp, _i64]),
    "eik_mesh_set_metric_tri": (ctypes.c_int, [_h, _fp, _i64]),
    "eik_gen_square": (ctypes.c_int, [_i32, _hp]),
    "eik_gen_kuhn_cube": (ctypes.c_int, [_i32, ctypes.c_float, ctypes.c_uint64, _hp]),
    "eik_gen_icosphere": (ctypes.c_int, [_i32, _hp]),
    "eik_partition": (ctypes.c_int, [_h, _i32, _ip]),
    "eik_partition_structured": (ctypes.c_int, [_h, _i32, _i32, _i32, _i32, _i32, _i32, _ip]),
    "eik_partition_set_labels": (ctypes.c_int, [_h, _ip, _i32]),
    "eik_partition_labels": (ctypes.c_int, [_h, _ip]),
    "eik_solve_opts_default": (None, [ctypes.POINTER(SolveOpts)]),
    "eik_solve": (ctypes.c_int, [_h, _ip, _i64, _fp, ctypes.POINTER(SolveOpts), _fp,
                                 ctypes.POINTER(SolveStats)]),
    "eik_solve_noalloc": (ctypes.c_int, [_h, ctypes.c_void_p, _i64, ctypes.c_void_p,
                                         ctypes.c_void_p]),
    "eik_get_result": (ctypes.c_int, [_h, _fp]),
    "eik_get_history": (ctypes.c_int, [_h, _i64, _fp]),
    "eik_mesh_upload": (ctypes.c_int, [_h]),
    "eik_mesh_virtual_faces": (ctypes.c_int, [_h, _ip, _ip, _i64, _i64p]),
    "eik_mesh_layout": (ctypes.c_int, [_h, ctypes.POINTER(LayoutView)]),
    "eik_debug_local_solve": (ctypes.c_int, [_i32, _fp, _fp, _i64, _fp]),
    "eik_batch_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), _i32, _hp]),
    "eik_batch_destroy": (ctypes.c_int, [_h]),
    "eik_batch_info": (ctypes.c_int, [_h, _i64p, _i64p]),
    "eik_batch_solve": (ctypes.c_int, [_h, _i64p, _ip, _fp, ctypes.POINTER(SolveOpts), _fp,
                                       ctypes.POINTER(SolveStats)]),
    "eik_batch_set_graph": (ctypes.c_int, [_h, _i32]),
    "eik_batch_set_factors": (ctypes.c_int, [_h, _i32, _fp, _i64]),
    "eik_batch_get_result": (ctypes.c_int, [_h, _i32, _fp]),
    "eik_batch_autotune": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), _i32, _i64p, _ip, _ip,
                                          _i32, _ip, _i32, _i32, ctypes.c_void_p,
                                          ctypes.POINTER(ctypes.c_double)]),
}


class ServerOpts(ctypes.Structure):
    _fields_ = [("max_solves_per_launch", ctypes.c_int32), ("launches", ctypes.c_int32),
                ("budget_bytes", ctypes.c_int64), ("max_seeds", ctypes.c_int32),
                ("max_inner", ctypes.c_int32)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("launch", ctypes.c_int32),
                ("submit_ms", ctypes.c_double), ("start_ms", ctypes.c_double),
                ("end_ms", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


SIGNATURES.update({
    "eik_server_opts_default": (None, [ctypes.POINTER(ServerOpts)]),
    "eik_server_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), _i32,
                                         ctypes.POINTER(ServerOpts), _hp]),
    "eik_server_submit": (ctypes.c_int, [_h, _i32, _ip, _i32, _fp, _fp, _i64p]),
    "eik_server_poll": (ctypes.c_int, [_h, _i64, ctypes.POINTER(RequestInfo)]),
    "eik_server_wait": (ctypes.c_int, [_h, _i64, ctypes.POINTER(RequestInfo)]),
    "eik_server_launches": (ctypes.c_int, [_h, _i64p]),
    "eik_server_destroy": (ctypes.c_int, [_h]),
})


class TuneResult(ctypes.Structure):
    _fields_ = [("verts_per_block", ctypes.c_int32), ("max_inner", ctypes.c_int32),
                ("ms", ctypes.c_double)]

_lib = None


def build(force=False):
    """Compile the library in-tree (make -C sci-solver_eikonal_amd)."""
    args = ["make", "-s", "-C", HERE, "-j8"]
    if force:
        subprocess.run(["make", "-s", "-C", HERE, "clean"], check=True)
    subprocess.run(args, check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libeikonal_amd.so is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc != EIK_OK:
        raise EikError(rc, lib().eik_last_error().decode())


def _f32(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_fp)


def _i32a(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def default_opts(**kw):
    o = SolveOpts()
    lib().eik_solve_opts_default(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


class Mesh:
    """Owns an eik_mesh_t handle."""

    def __init__(self, handle):
        self.h = ctypes.c_void_p(handle)

    # construction -----------------------------------------------------------------------
    @classmethod
    def _make(cls, fn, *args):
        h = ctypes.c_void_p()
        _check(fn(*args, ctypes.byref(h)))
        return cls(h.value)

    @classmethod
    def tet(cls, xyz, tets):
        xyz, px = _f32(np.asarray(xyz).reshape(-1, 3))
        tets, pt = _i32a(np.asarray(tets).reshape(-1, 4))
        return cls._make(lib().eik_mesh_create_tet, px, xyz.shape[0], pt, tets.shape[0])

    @classmethod
    def tri(cls, xyz, faces):
        xyz, px = _f32(np.asarray(xyz).reshape(-1, 3))
        faces, pf = _i32a(np.asarray(faces).reshape(-1, 3))
        return cls._make(lib().eik_mesh_create_tri, px, xyz.shape[0], pf, faces.shape[0])

    @classmethod
    def load_tetgen(cls, base):
        return cls._make(lib().eik_mesh_load_tetgen, base.encode())

    @classmethod
    def load_ply(cls, path):
        return cls._make(lib().eik_mesh_load_ply, path.encode())

    @classmethod
    def square(cls, n):
        return cls._make(lib().eik_gen_square, n)

    @classmethod
    def kuhn_cube(cls, n, jitter=0.2, seed=0):
        return cls._make(lib().eik_gen_kuhn_cube, n, jitter, seed)

    @classmethod
    def icosphere(cls, level):
        return cls._make(lib().eik_gen_icosphere, level)

    def close(self):
        if self.h is not None and self.h.value:
            lib().eik_mesh_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # accessors --------------------------------------------------------------------------
    def info(self):
        i = MeshInfo()
        _check(lib().eik_mesh_info(self.h, ctypes.byref(i)))
        return {k: getattr(i, k) for k, _ in i._fields_ if k != "pad"}

    @property
    def nv(self):
        return self.info()["nv"]

    @property
    def dim(self):
        return self.info()["dim"]

    def vertices(self):
        i = self.info()
        out = np.zeros((i["nv"], 3), np.float32)
        _check(lib().eik_mesh_get_vertices(self.h, out.ctypes.data_as(_fp)))
        return out

    def elements(self):
        i = self.info()
        k = 4 if i["dim"] == 3 else 3
        out = np.zeros((i["nelem"], k), np.int32)
        _check(lib().eik_mesh_get_elements(self.h, out.ctypes.data_as(_ip)))
        return out

    def stream(self):
        s = ctypes.c_void_p()
        _check(lib().eik_mesh_stream(self.h, ctypes.byref(s)))
        return s.value

    # speeds -----------------------------------------------------------------------------
    def set_speed(self, values):
        values = np.zeros(0, np.float32) if values is None else values
        v, pv = _f32(np.asarray(values).ravel())
        _check(lib().eik_mesh_set_speed(self.h, pv, v.size))

    def set_metric_tri(self, m3):
        v, pv = _f32(np.asarray(m3).reshape(-1, 3))
        _check(lib().eik_mesh_set_metric_tri(self.h, pv, v.shape[0]))

    # partitions -------------------------------------------------------------------------
    def partition(self, verts_per_block=64):
        nb = ctypes.c_int32()
        _check(lib().eik_partition(self.h, verts_per_block, ctypes.byref(nb)))
        return nb.value

    def partition_structured(self, L, W, D, bl, bw, bd):
        nb = ctypes.c_int32()
        _check(lib().eik_partition_structured(self.h, L, W, D, bl, bw, bd, ctypes.byref(nb)))
        return nb.value

    def set_labels(self, labels, nblocks):
        lab, pl = _i32a(labels)
        _check(lib().eik_partition_set_labels(self.h, pl, nblocks))

    def labels(self):
        out = np.zeros(self.nv, np.int32)
        _check(lib().eik_partition_labels(self.h, out.ctypes.data_as(_ip)))
        return out

    def virtual_faces(self):
        n = ctypes.c_int64()
        _check(lib().eik_mesh_virtual_faces(self.h, None, None, 0, ctypes.byref(n)))
        rows = np.zeros((max(n.value, 1), 3), np.int32)
        origin = np.zeros(max(n.value, 1), np.int32)
        _check(lib().eik_mesh_virtual_faces(self.h, rows.ctypes.data_as(_ip),
                                            origin.ctypes.data_as(_ip), n.value,
                                            ctypes.byref(n)))
        return rows[:n.value], origin[:n.value]

    def layout(self):
        """Copy of the block layout (numpy arrays); see eik_mesh_layout."""
        v = LayoutView()
        _check(lib().eik_mesh_layout(self.h, ctypes.byref(v)))

        def arr(p, n):
            return np.ctypeslib.as_array(p, shape=(n,)).copy() if n > 0 else np.zeros(0, np.int32)
        nb = v.nblocks
        rec = TET_REC if v.dim == 3 else TRI_REC
        assert rec.itemsize == v.rec_bytes
        buf = (ctypes.c_char * (v.nelems * v.rec_bytes)).from_address(v.elems)
        elems = np.frombuffer(bytes(buf), dtype=rec)
        sup_start = arr(v.sup_start, nb + 1)
        dep_start = arr(v.dep_start, nb + 1)
        return {
            "dim": v.dim, "nblocks": nb, "nv": v.nv, "max_local": v.max_local,
            "perm": arr(v.perm, v.nv), "iperm": arr(v.iperm, v.nv),
            "blocks": arr(v.blocks, 8 * nb).reshape(nb, 8), "halo": arr(v.halo, v.nhalo),
            "elems": elems, "sup_start": sup_start, "sup": arr(v.sup, int(sup_start[-1])),
            "dep_start": dep_start, "dep": arr(v.dep, int(dep_start[-1])),
        }

    # solve ------------------------------------------------------------------------------
    def upload(self):
        _check(lib().eik_mesh_upload(self.h))

    def solve(self, seeds, seed_vals=None, out=True, **opts):
        seeds, ps = _i32a(np.atleast_1d(seeds))
        pv = None
        if seed_vals is not None:
            sv, pv = _f32(np.atleast_1d(seed_vals))
        o = default_opts(**opts)
        st = SolveStats()
        T = np.zeros(self.nv, np.float32) if out else None
        _check(lib().eik_solve(self.h, ps, seeds.size, pv, ctypes.byref(o),
                               T.ctypes.data_as(_fp) if out else None, ctypes.byref(st)))
        return T, st.as_dict()

    def solve_noalloc(self, seeds_dev_ptr, nseeds, out_dev_ptr, stream=None):
        """eik_solve_noalloc: seeds and output in device memory (raw pointers, e.g. a torch
        tensor's data_ptr()); enqueues the solve on `stream` (None: the mesh's stream) and
        returns without waiting."""
        _check(lib().eik_solve_noalloc(self.h, ctypes.c_void_p(seeds_dev_ptr), int(nseeds),
                                       ctypes.c_void_p(out_dev_ptr),
                                       ctypes.c_void_p(stream) if stream else None))

    def result(self):
        T = np.zeros(self.nv, np.float32)
        _check(lib().eik_get_result(self.h, T.ctypes.data_as(_fp)))
        return T

    def history(self, i):
        T = np.zeros(self.nv, np.float32)
        _check(lib().eik_get_history(self.h, i, T.ctypes.data_as(_fp)))
        return T


class Batch:
    """Several partitioned meshes solved by one persistent launch (eik_batch_*).  Keeps
    references to its meshes, which must stay partitioned as they were."""

    def __init__(self, meshes):
        self.meshes = list(meshes)
        arr = (ctypes.c_void_p * len(self.meshes))(*[m.h.value for m in self.meshes])
        h = ctypes.c_void_p()
        _check(lib().eik_batch_create(arr, len(self.meshes), ctypes.byref(h)))
        self.h = h
        self.nvs = [m.nv for m in self.meshes]

    def close(self):
        if self.h is not None and self.h.value:
            lib().eik_batch_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        a, b = ctypes.c_int64(), ctypes.c_int64()
        _check(lib().eik_batch_info(self.h, ctypes.byref(a), ctypes.byref(b)))
        return {"total_verts": a.value, "total_blocks": b.value}

    def set_graph(self, enable):
        _check(lib().eik_batch_set_graph(self.h, 1 if enable else 0))

    def set_factors(self, i, factors):
        if factors is None:
            _check(lib().eik_batch_set_factors(self.h, i, None, 0))
            return
        f, pf = _f32(np.asarray(factors).ravel())
        _check(lib().eik_batch_set_factors(self.h, i, pf, f.size))

    def result(self, i):
        """Times of solve i of the last solve (eik_batch_get_result)."""
        T = np.zeros(self.nvs[i] if 0 <= i < len(self.nvs) else 1, np.float32)
        _check(lib().eik_batch_get_result(self.h, i, T.ctypes.data_as(_fp)))
        return T

    def solve(self, seeds, seed_vals=None, out=True, **opts):
        """seeds: one sequence of caller vertex ids per mesh.  Returns (list of T arrays or
        None, stats dict); with out=False the times stay on the device (result(i))."""
        assert len(seeds) == len(self.meshes)
        counts = [len(np.atleast_1d(x)) for x in seeds]
        off = np.zeros(len(seeds) + 1, np.int64)
        off[1:] = np.cumsum(counts)
        flat, pflat = _i32a(np.concatenate([np.atleast_1d(x) for x in seeds]).astype(np.int32))
        pv = None
        if seed_vals is not None:
            sv, pv = _f32(np.concatenate([np.atleast_1d(x) for x in seed_vals]))
        o = default_opts(**opts)
        st = SolveStats()
        T = np.zeros(sum(self.nvs), np.float32) if out else None
        _check(lib().eik_batch_solve(self.h, off.ctypes.data_as(_i64p), pflat, pv, ctypes.byref(o),
                                     T.ctypes.data_as(_fp) if out else None, ctypes.byref(st)))
        if not out:
            return None, st.as_dict()
        parts = np.split(T, np.cumsum(self.nvs)[:-1])
        return parts, st.as_dict()


class Server:
    """Serving loop over resident meshes (eik_server_*).  submit() queues a solve and returns
    its id; wait(id) blocks until it is done and returns (T or None, info dict)."""

    def __init__(self, meshes, **opts):
        self.meshes = list(meshes)
        o = ServerOpts()
        lib().eik_server_opts_default(ctypes.byref(o))
        for k, v in opts.items():
            setattr(o, k, v)
        arr = (ctypes.c_void_p * len(self.meshes))(*[m.h.value for m in self.meshes])
        h = ctypes.c_void_p()
        _check(lib().eik_server_create(arr, len(self.meshes), ctypes.byref(o), ctypes.byref(h)))
        self.h = h
        self._keep = {}  # request id -> (seeds, factors, out) kept alive until wait

    def submit(self, mesh, seeds, factors=None, out=True):
        sd, ps = _i32a(np.atleast_1d(seeds))
        pf = None
        f = None
        if factors is not None:
            f, pf = _f32(np.asarray(factors).ravel())
        T = np.empty(self.meshes[mesh].nv, np.float32) if out else None
        rid = ctypes.c_int64()
        _check(lib().eik_server_submit(self.h, mesh, ps, sd.size, pf,
                                       T.ctypes.data_as(_fp) if out else None, ctypes.byref(rid)))
        self._keep[rid.value] = (sd, f, T)
        return rid.value

    def poll(self, rid):
        info = RequestInfo()
        _check(lib().eik_server_poll(self.h, rid, ctypes.byref(info)))
        return info.as_dict()

    def wait(self, rid):
        info = RequestInfo()
        try:
            _check(lib().eik_server_wait(self.h, rid, ctypes.byref(info)))
        finally:
            kept = self._keep.pop(rid, (None, None, None))
        return kept[2], info.as_dict()

    def launches(self):
        n = ctypes.c_int64()
        _check(lib().eik_server_launches(self.h, ctypes.byref(n)))
        return n.value

    def close(self):
        if self.h is not None and self.h.value:
            lib().eik_server_destroy(self.h)
        self.h = None
        self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune(meshes, seeds, vpb=(32, 64, 128, 192), inner=(1, 2, 4), reps=3):
    """eik_batch_autotune: returns ({verts_per_block, max_inner, ms}, {(vpb, inner): ms}); the
    meshes end partitioned with the winner."""
    arr = (ctypes.c_void_p * len(meshes))(*[m.h.value for m in meshes])
    counts = [len(np.atleast_1d(x)) for x in seeds]
    off = np.zeros(len(seeds) + 1, np.int64)
    off[1:] = np.cumsum(counts)
    flat, pflat = _i32a(np.concatenate([np.atleast_1d(x) for x in seeds]).astype(np.int32))
    v, pv = _i32a(np.asarray(vpb))
    k, pk = _i32a(np.asarray(inner))
    best = TuneResult()
    times = np.zeros(len(vpb) * len(inner), np.float64)
    _check(lib().eik_batch_autotune(arr, len(meshes), off.ctypes.data_as(_i64p), pflat, pv, v.size,
                                    pk, k.size, reps, ctypes.cast(ctypes.byref(best), ctypes.c_void_p),
                                    times.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    table = {(int(a), int(b)): float(times[i * len(inner) + j])
             for i, a in enumerate(vpb) for j, b in enumerate(inner)}
    return {"verts_per_block": best.verts_per_block, "max_inner": best.max_inner,
            "ms": best.ms}, table


def debug_local_solve(dim, T, L):
    """Run the device local solver on independent elements; see eik_debug_local_solve."""
    k = 4 if dim == 3 else 3
    nl = 6 if dim == 3 else 3
    T, pT = _f32(np.asarray(T).reshape(-1, k))
    L, pL = _f32(np.asarray(L).reshape(-1, nl))
    out = np.zeros_like(T)
    _check(lib().eik_debug_local_solve(dim, pT, pL, T.shape[0], out.ctypes.data_as(_fp)))
    return out


def header_symbols():
    """Function names declared in include/eikonal_amd.h."""
    import re
    src = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?(?:int|void|const char\*)\s+(eik_\w+)\s*\(",
                                 src, re.M)))
