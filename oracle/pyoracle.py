"""ctypes front end of the CPU oracle (oracle/mzoracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker / CPU baseline. The product path
(mizuba_amd) never imports this module.

`run(trajs, times, conj_thresh, conj_det_interval, otypes=None, ...)` mirrors
`mizuba.conjunctions(pj, conj_thresh, conj_det_interval, otypes)`
(mizuba/expose_conjunctions.cpp:82-98) and returns plain numpy arrays with the
reference's dtypes and layouts.
"""

from __future__ import annotations

import ctypes as ct
import os
import pathlib
import subprocess

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
_LIB_PATH = _HERE / "libmzoracle.so"

BVH_NODE_DTYPE = np.dtype(
    [("begin", "<u4"), ("end", "<u4"), ("left", "<i4"), ("right", "<i4"), ("lb", "<f4", (4,)), ("ub", "<f4", (4,))]
)
AABB_COLLISION_DTYPE = np.dtype([("i", "<u4"), ("j", "<u4")])
CONJ_DTYPE = np.dtype(
    [
        ("i", "<u4"),
        ("j", "<u4"),
        ("tca", "<f8"),
        ("dca", "<f8"),
        ("ri", "<f8", (3,)),
        ("vi", "<f8", (3,)),
        ("rj", "<f8", (3,)),
        ("vj", "<f8", (3,)),
    ]
)
assert BVH_NODE_DTYPE.itemsize == 48 and AABB_COLLISION_DTYPE.itemsize == 8 and CONJ_DTYPE.itemsize == 120

NP_REPORT_FIELDS = (
    "n_tot_conj_candidates",
    "n_dist2_check",
    "n_poly_roots",
    "n_fex_check",
    "n_poly_no_roots",
    "n_tot_dist_minima",
    "n_tot_discarded_dist_minima",
)


class _PJ(ct.Structure):
    _fields_ = [
        ("n_objs", ct.c_uint32),
        ("order", ct.c_uint32),
        ("traj", ct.POINTER(ct.c_double)),
        ("time", ct.POINTER(ct.c_double)),
        ("traj_off", ct.POINTER(ct.c_uint64)),
        ("nsteps", ct.POINTER(ct.c_uint64)),
        ("time_off", ct.POINTER(ct.c_uint64)),
        ("maxT", ct.c_double),
    ]


class _Result(ct.Structure):
    _fields_ = [
        ("status", ct.c_int),
        ("err", ct.c_char * 512),
        ("n_objs", ct.c_uint64),
        ("n_cd_steps", ct.c_uint64),
        ("step_begin", ct.c_uint64),
        ("step_end", ct.c_uint64),
        ("skip_cd", ct.c_int),
        ("aabbs", ct.POINTER(ct.c_float)),
        ("srt_aabbs", ct.POINTER(ct.c_float)),
        ("mcodes", ct.POINTER(ct.c_uint64)),
        ("srt_mcodes", ct.POINTER(ct.c_uint64)),
        ("srt_idx", ct.POINTER(ct.c_uint32)),
        ("cd_end_times", ct.POINTER(ct.c_double)),
        ("trees", ct.c_void_p),
        ("tree_offsets", ct.POINTER(ct.c_uint64)),
        ("bp", ct.c_void_p),
        ("bp_offsets", ct.POINTER(ct.c_uint64)),
        ("conj", ct.c_void_p),
        ("n_conj", ct.c_uint64),
        ("np_report", ct.c_uint64 * 7),
        ("stage_seconds", ct.c_double * 5),
    ]


_lib = None


def build() -> pathlib.Path:
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build()
        L = ct.CDLL(str(_LIB_PATH))
        dp = ct.POINTER(ct.c_double)
        L.mzo_conjunctions.argtypes = [
            ct.POINTER(_PJ),
            ct.c_double,
            ct.c_double,
            ct.POINTER(ct.c_int32),
            ct.c_uint64,
            ct.c_uint64,
            ct.c_int,
            ct.POINTER(_Result),
        ]
        L.mzo_conjunctions.restype = ct.c_int
        L.mzo_free_result.argtypes = [ct.POINTER(_Result)]
        L.mzo_free_result.restype = None
        L.mzo_poly_translate.argtypes = [dp, ct.c_uint32, ct.c_uint32, ct.c_double, dp, ct.c_uint32]
        L.mzo_poly_translate.restype = None
        L.mzo_cs_enclosure.argtypes = [dp, ct.c_uint32, ct.c_uint32, ct.c_double, ct.c_double, dp, dp]
        L.mzo_cs_enclosure.restype = None
        L.mzo_disc_single_coord.argtypes = [ct.c_float, ct.c_float, ct.c_float]
        L.mzo_disc_single_coord.restype = ct.c_uint64
        L.mzo_morton_encode4.argtypes = [ct.c_uint64] * 4
        L.mzo_morton_encode4.restype = ct.c_uint64
        L.mzo_cheby_nodes.argtypes = [ct.c_uint32, ct.c_double, dp]
        L.mzo_cheby_nodes.restype = None
        L.mzo_vm_interp.argtypes = [ct.c_uint32, dp, dp, dp]
        L.mzo_vm_interp.restype = None
        L.mzo_root_finding.argtypes = [
            dp,
            ct.c_uint32,
            ct.c_double,
            ct.c_int,
            dp,
            ct.POINTER(ct.c_uint32),
            ct.POINTER(ct.c_int),
        ]
        L.mzo_root_finding.restype = ct.c_int
        L.mzo_bracketed_root.argtypes = [dp, ct.c_uint32, ct.c_double, ct.c_double, dp, ct.POINTER(ct.c_uint32)]
        L.mzo_bracketed_root.restype = ct.c_int
        L.mzo_cd_begin_end.argtypes = [ct.c_double, ct.c_uint64, ct.c_double, ct.c_uint64, dp, dp]
        L.mzo_cd_begin_end.restype = ct.c_int
        L.mzo_state_meval.argtypes = [
            ct.POINTER(_PJ),
            dp,
            ct.c_int,
            ct.POINTER(ct.c_uint64),
            ct.c_uint64,
            ct.c_uint64,
            dp,
            ct.POINTER(ct.c_uint64),
        ]
        L.mzo_state_meval.restype = ct.c_int
        _lib = L
    return _lib


def _dptr(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


class OracleError(Exception):
    pass


class Flat:
    """Concatenated polyjectory arrays (same view as polyjectory.cpp:1006-1036)."""

    def __init__(self, trajs, times):
        trajs = [np.ascontiguousarray(t, dtype=np.float64) for t in trajs]
        times = [np.ascontiguousarray(t, dtype=np.float64) for t in times]
        if len(trajs) == 0 or len(trajs) != len(times):
            raise ValueError("inconsistent polyjectory data")
        op1 = trajs[0].shape[1]
        self.order = op1 - 1
        self.n_objs = len(trajs)
        self.nsteps = np.array([t.shape[0] for t in trajs], dtype=np.uint64)
        sizes = np.array([t.size for t in trajs], dtype=np.uint64)
        self.traj_off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.uint64)
        tsz = np.array([t.shape[0] + (1 if t.shape[0] else 0) for t in trajs], dtype=np.uint64)
        self.time_off = np.concatenate([[0], np.cumsum(tsz)[:-1]]).astype(np.uint64)
        self.traj = np.concatenate([t.reshape(-1) for t in trajs]) if sizes.sum() else np.zeros(1)
        tl = [tm[: int(n)] for tm, n in zip(times, tsz)]
        self.time = np.concatenate(tl) if tsz.sum() else np.zeros(1)
        self.maxT = max(float(tm[-1]) for tm, n in zip(times, tsz) if n > 0)

    @classmethod
    def from_arrays(cls, order, traj, time, nsteps, traj_off, time_off, maxT):
        self = cls.__new__(cls)
        self.order = int(order)
        self.n_objs = len(nsteps)
        self.nsteps = np.ascontiguousarray(nsteps, dtype=np.uint64)
        self.traj_off = np.ascontiguousarray(traj_off, dtype=np.uint64)
        self.time_off = np.ascontiguousarray(time_off, dtype=np.uint64)
        self.traj = np.ascontiguousarray(traj, dtype=np.float64)
        self.time = np.ascontiguousarray(time, dtype=np.float64)
        self.maxT = float(maxT)
        return self

    def view(self):
        return _PJ(
            self.n_objs,
            self.order,
            _dptr(self.traj),
            _dptr(self.time),
            self.traj_off.ctypes.data_as(ct.POINTER(ct.c_uint64)),
            self.nsteps.ctypes.data_as(ct.POINTER(ct.c_uint64)),
            self.time_off.ctypes.data_as(ct.POINTER(ct.c_uint64)),
            self.maxT,
        )


def run(pj, conj_thresh, conj_det_interval, otypes=None, step_begin=0, step_end=0, n_threads=0):
    """Run the oracle pipeline; pj is a Flat or a (trajs, times) tuple."""
    if not isinstance(pj, Flat):
        pj = Flat(*pj)
    L = lib()
    view = pj.view()
    res = _Result()
    ot = None
    if otypes is not None:
        ot_arr = np.ascontiguousarray(otypes, dtype=np.int32)
        if ot_arr.shape[0] != pj.n_objs:
            raise ValueError(
                "Invalid array of object types passed to the constructor of a conjunctions objects: the expected "
                f"size is {pj.n_objs}, but the actual size is {ot_arr.shape[0]} instead"
            )
        ot = ot_arr.ctypes.data_as(ct.POINTER(ct.c_int32))
    st = L.mzo_conjunctions(
        ct.byref(view), float(conj_thresh), float(conj_det_interval), ot, step_begin, step_end, n_threads, ct.byref(res)
    )
    try:
        if st != 0:
            msg = res.err.decode(errors="replace")
            if st == 1:
                raise ValueError(msg)
            raise OracleError(f"oracle failed with status {st}: {msg}")
        N = int(res.n_objs)
        ns = int(res.step_end - res.step_begin)

        def arr(ptr, ctype, count):
            if count == 0:
                return np.zeros(0, dtype=np.dtype(ctype))
            return np.ctypeslib.as_array(ptr, shape=(count,)).copy()

        out = {
            "n_cd_steps": int(res.n_cd_steps),
            "step_begin": int(res.step_begin),
            "step_end": int(res.step_end),
            "skip_cd": bool(res.skip_cd),
            "aabbs": arr(res.aabbs, np.float32, ns * (N + 1) * 8).reshape(ns, N + 1, 2, 4),
            "srt_aabbs": arr(res.srt_aabbs, np.float32, ns * (N + 1) * 8).reshape(ns, N + 1, 2, 4),
            "mcodes": arr(res.mcodes, np.uint64, ns * N).reshape(ns, N),
            "srt_mcodes": arr(res.srt_mcodes, np.uint64, ns * N).reshape(ns, N),
            "srt_idx": arr(res.srt_idx, np.uint32, ns * N).reshape(ns, N),
            "cd_end_times": arr(res.cd_end_times, np.float64, ns),
            "tree_offsets": arr(res.tree_offsets, np.uint64, 2 * ns).reshape(ns, 2),
            "bp_offsets": arr(res.bp_offsets, np.uint64, 2 * ns).reshape(ns, 2),
            "np_report": dict(zip(NP_REPORT_FIELDS, [int(x) for x in res.np_report])),
            "stage_seconds": list(res.stage_seconds),
        }
        n_nodes = int(out["tree_offsets"][:, 1].sum()) if ns else 0
        n_bp = int(out["bp_offsets"][:, 1].sum()) if ns else 0
        n_conj = int(res.n_conj)

        def sarr(ptr, dtype, count):
            if count == 0:
                return np.zeros(0, dtype=dtype)
            buf = (ct.c_char * (dtype.itemsize * count)).from_address(ptr)
            return np.frombuffer(bytes(buf), dtype=dtype).copy()

        out["trees"] = sarr(res.trees, BVH_NODE_DTYPE, n_nodes)
        out["bp"] = sarr(res.bp, AABB_COLLISION_DTYPE, n_bp)
        out["conjunctions"] = sarr(res.conj, CONJ_DTYPE, n_conj)
        return out
    finally:
        L.mzo_free_result(ct.byref(res))


def get_bvh_tree(out, k):
    off, size = out["tree_offsets"][k]
    return out["trees"][int(off) : int(off) + int(size)]


def get_aabb_collisions(out, k):
    off, size = out["bp_offsets"][k]
    return out["bp"][int(off) : int(off) + int(size)]


# ---- unit-level helpers ----------------------------------------------------


def poly_translate(cfs, a):
    cfs = np.ascontiguousarray(cfs, dtype=np.float64)
    out = np.zeros_like(cfs)
    lib().mzo_poly_translate(_dptr(cfs), len(cfs) - 1, 1, float(a), _dptr(out), 1)
    return out


def cs_enclosure(cfs, lb, ub):
    cfs = np.ascontiguousarray(cfs, dtype=np.float64)
    lo, hi = ct.c_double(), ct.c_double()
    lib().mzo_cs_enclosure(_dptr(cfs), len(cfs) - 1, 1, float(lb), float(ub), ct.byref(lo), ct.byref(hi))
    return lo.value, hi.value


def disc_single_coord(x, mn, mx):
    return int(lib().mzo_disc_single_coord(float(x), float(mn), float(mx)))


def morton_encode4(n0, n1, n2, n3):
    return int(lib().mzo_morton_encode4(n0, n1, n2, n3))


def cheby_nodes(order, h):
    out = np.zeros(order + 1)
    lib().mzo_cheby_nodes(order, float(h), _dptr(out))
    return out


def vm_interp(alpha, f):
    alpha = np.ascontiguousarray(alpha, dtype=np.float64)
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros_like(f)
    lib().mzo_vm_interp(len(f) - 1, _dptr(alpha), _dptr(f), _dptr(out))
    return out


def root_finding(poly, rf_int, direction=1):
    poly = np.ascontiguousarray(poly, dtype=np.float64)
    order = len(poly) - 1
    roots = np.zeros(2 * order + 2)
    nr = ct.c_uint32()
    fex = ct.c_int()
    st = lib().mzo_root_finding(_dptr(poly), order, float(rf_int), direction, _dptr(roots), ct.byref(nr), ct.byref(fex))
    if st != 0:
        raise ValueError("root finding failed")
    return roots[: nr.value].copy(), bool(fex.value)


def bracketed_root(poly, lb, ub):
    poly = np.ascontiguousarray(poly, dtype=np.float64)
    r = ct.c_double()
    it = ct.c_uint32()
    st = lib().mzo_bracketed_root(_dptr(poly), len(poly) - 1, float(lb), float(ub), ct.byref(r), ct.byref(it))
    return st, r.value, it.value


def cd_begin_end(maxT, k, interval, n):
    b, e = ct.c_double(), ct.c_double()
    st = lib().mzo_cd_begin_end(float(maxT), int(k), float(interval), int(n), ct.byref(b), ct.byref(e))
    return st, b.value, e.value


def n_threads_default() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


TIME_SCALAR, TIME_PER_OBJECT, TIME_SHARED, TIME_PER_EVAL = 0, 1, 2, 3


def state_meval(pj, tm, time_mode, sel=None, n_rows=None, n_times=1):
    """Batched polyjectory evaluation (polyjectory.cpp:1094-1366).

    Returns (status, out[n_rows][n_times][7], flat index of the first failing
    element or None)."""
    if not isinstance(pj, Flat):
        pj = Flat(*pj)
    tm = np.ascontiguousarray(np.atleast_1d(tm), dtype=np.float64)
    if sel is not None:
        sel = np.ascontiguousarray(sel, dtype=np.uint64)
        n_rows = len(sel)
    elif n_rows is None:
        n_rows = pj.n_objs
    out = np.empty((n_rows, n_times, 7))
    bad = ct.c_uint64(0)
    view = pj.view()
    st = lib().mzo_state_meval(
        ct.byref(view),
        _dptr(tm),
        int(time_mode),
        sel.ctypes.data_as(ct.POINTER(ct.c_uint64)) if sel is not None else None,
        n_rows,
        n_times,
        _dptr(out),
        ct.byref(bad),
    )
    return st, out, (int(bad.value) if st else None)
