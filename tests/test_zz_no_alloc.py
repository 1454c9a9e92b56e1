"""The allocation-free root finder (mzh_root_finding_noalloc) through the C
ABI, against mzh_root_finding on the same polynomials: bit-identical roots,
counts and flags, on two calls in a row (the second smaller, with the slow
path forced), and a too-large request rejected."""

import ctypes
import pathlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = pathlib.Path(__file__).resolve().parent.parent


def _lib():
    lib = ctypes.CDLL(str(ROOT / "mizuba_amd" / "libmizuba_amd.so"))
    vp, u32, u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64
    lib.mzh_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    lib.mzh_destroy.argtypes = [vp]
    lib.mzh_last_error.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
    lib.mzh_device_alloc.argtypes = [vp, u64, ctypes.POINTER(vp)]
    lib.mzh_device_free.argtypes = [vp, vp]
    lib.mzh_copy_to_device.argtypes = [vp, vp, vp, u64]
    lib.mzh_copy_from_device.argtypes = [vp, vp, vp, u64]
    for f in (lib.mzh_root_finding, lib.mzh_root_finding_noalloc):
        f.argtypes = [vp, u32, u64, vp, vp, vp, vp, vp]
    return lib


class Dev:
    def __init__(self, lib, ctx):
        self.lib, self.ctx, self.bufs = lib, ctx, []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.lib.mzh_device_alloc(self.ctx, a.nbytes, ctypes.byref(p)) == 0
        assert self.lib.mzh_copy_to_device(self.ctx, p, a.ctypes.data, a.nbytes) == 0
        self.bufs.append(p)
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.lib.mzh_copy_from_device(self.ctx, out.ctypes.data, p, out.nbytes) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.lib.mzh_device_free(self.ctx, p)


def _err(lib, ctx):
    buf = ctypes.create_string_buffer(1024)
    lib.mzh_last_error(ctx, buf, 1024)
    return buf.value.decode()


def _host_call(lib, ctx, order, P, R):
    n = len(P)
    roots = np.zeros((n, 2 * order + 2))
    cnt = np.zeros(n, np.uint32)
    flags = np.zeros(n, np.uint32)
    st = lib.mzh_root_finding(ctx, order, n, P.ctypes.data, R.ctypes.data, roots.ctypes.data, cnt.ctypes.data,
                              flags.ctypes.data)
    assert st == 0, _err(lib, ctx)
    return roots, cnt, flags


def _noalloc_call(lib, ctx, order, P, R):
    n = len(P)
    d = Dev(lib, ctx)
    dp, dr = d.put(P), d.put(R)
    roots = np.full((n, 2 * order + 2), np.nan)
    cnt = np.full(n, 7, np.uint32)
    flags = np.full(n, 7, np.uint32)
    dro, dc, df = d.put(roots), d.put(cnt), d.put(flags)
    st = lib.mzh_root_finding_noalloc(ctx, order, n, dp, dr, dro, dc, df)
    assert st == 0, _err(lib, ctx)
    out = d.get(dro, roots), d.get(dc, cnt), d.get(df, flags)
    d.free()
    return out


def test_noalloc_root_finding_matches(monkeypatch):
    from tests.test_gpu_robustness import _poly_cases
    from oracle import pyoracle as po

    lib = _lib()
    ctx = ctypes.c_void_p()
    assert lib.mzh_create(0, ctypes.byref(ctx)) == 0
    try:
        order = 8
        P = _poly_cases(order, 3000, seed=77)
        R = np.full(len(P), 1.0)
        R[1::3] = 0.0021
        ok = []
        for k in range(len(P)):
            try:
                po.root_finding(P[k], R[k], 1)
                ok.append(k)
            except ValueError:
                pass
        P, R = np.ascontiguousarray(P[ok]), np.ascontiguousarray(R[ok])
        # first call: the whole set
        ref = _host_call(lib, ctx, order, P, R)
        got = _noalloc_call(lib, ctx, order, P, R)
        for a, b in zip(got, ref):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
        assert np.sum(ref[1]) > 0
        # second call, smaller, with the slow path (the scratch's work lists)
        # taken by most polynomials
        monkeypatch.setenv("MIZUBA_AMD_WL_FAST_CAP", "2")
        P2, R2 = P[:700][::-1].copy(), R[:700][::-1].copy()
        ref2 = _host_call(lib, ctx, order, P2, R2)
        got2 = _noalloc_call(lib, ctx, order, P2, R2)
        for a, b in zip(got2, ref2):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
        assert np.count_nonzero(got2[2] & 2) > 100
        # a failing polynomial: the same error as the allocating entry point
        bad = np.ascontiguousarray(P2[:4].copy())
        Rb = np.ascontiguousarray([1.0, 1.0, -1.0, 1.0])
        d = Dev(lib, ctx)
        st = lib.mzh_root_finding_noalloc(ctx, order, 4, d.put(bad), d.put(Rb), d.put(np.zeros((4, 18))),
                                          d.put(np.zeros(4, np.uint32)), None)
        d.free()
        assert st == 1 and "Invalid root-finding interval -1 for the polynomial at index 2" in _err(lib, ctx)
        # too many polynomials for the fixed scratch: rejected before any work
        st = lib.mzh_root_finding_noalloc(ctx, order, 10**8, ctypes.c_void_p(16), ctypes.c_void_p(16),
                                          ctypes.c_void_p(16), ctypes.c_void_p(16), None)
        assert st == 1 and "exceed the" in _err(lib, ctx) and "fixed scratch of 16 MiB" in _err(lib, ctx)
    finally:
        lib.mzh_destroy(ctx)
