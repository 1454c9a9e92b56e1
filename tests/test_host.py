"""Host-side tests that run without a GPU: the C-ABI library and its exports,
the polyjectory front end (validation with the reference's messages), the
synthetic catalog generator and bench.py's argument contract."""

import ctypes
import pathlib
import re
import subprocess
import sys

import numpy as np
import pytest

import mizuba_amd as mz

from . import helpers as H

ROOT = pathlib.Path(__file__).resolve().parent.parent


def _declared_functions():
    text = (ROOT / "include" / "mizuba_amd.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\bint\s+(mzh_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(ROOT / "mizuba_amd" / "libmizuba_amd.so"))
    names = _declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_c_abi_without_context():
    lib = ctypes.CDLL(str(ROOT / "mizuba_amd" / "libmizuba_amd.so"))
    n = ctypes.c_int(-1)
    assert lib.mzh_device_count(ctypes.byref(n)) == 0 and n.value >= 0
    # null context / arguments are rejected with MZH_EINVAL, never a crash
    assert lib.mzh_screen(None, ctypes.c_double(1.0), ctypes.c_double(1.0), None, 0) == 1
    assert lib.mzh_destroy(None) == 0
    buf = ctypes.create_string_buffer(256)
    assert lib.mzh_last_error(None, buf, 256) == 1


def test_screen_batch_argument_checks():
    lib = ctypes.CDLL(str(ROOT / "mizuba_amd" / "libmizuba_amd.so"))
    one = (ctypes.c_double * 2)(1.0, 1.0)
    st = (ctypes.c_int * 2)()
    # nothing to do
    assert lib.mzh_screen_batch(None, ctypes.c_uint64(0), None, None, None, 0, None) == 0
    # null context, repeated context, missing arrays: MZH_EINVAL before anything runs
    ctxs = (ctypes.c_void_p * 2)(None, None)
    assert lib.mzh_screen_batch(ctxs, ctypes.c_uint64(2), one, one, None, 0, st) == 1
    fake = ctypes.c_void_p(0x1000)
    ctxs = (ctypes.c_void_p * 2)(fake, fake)
    assert lib.mzh_screen_batch(ctxs, ctypes.c_uint64(2), one, one, None, 0, st) == 1
    assert lib.mzh_screen_batch(ctxs, ctypes.c_uint64(2), None, one, None, 0, st) == 1


def test_conjunctions_batch_length_mismatch():
    tcs, tms = H.planar_circ()
    pj = mz.polyjectory([tcs], [tms], [0])
    with pytest.raises(ValueError, match="lists of the same length"):
        mz.conjunctions_batch([pj, pj], [1.0], [0.1, 0.1])


@pytest.mark.skipif(mz.device_count() > 0, reason="checks the behaviour on a host without a HIP device")
def test_no_device_fails_loudly():
    tcs, tms = H.planar_circ()
    pj = mz.polyjectory([tcs], [tms], [0])
    with pytest.raises(RuntimeError, match="No HIP device is available"):
        mz.conjunctions(pj, 0.1, 0.1)
    lib = ctypes.CDLL(str(ROOT / "mizuba_amd" / "libmizuba_amd.so"))
    ctx = ctypes.c_void_p()
    assert lib.mzh_create(0, ctypes.byref(ctx)) == 3


def test_polyjectory_basics():
    tcs, tms = H.planar_circ()
    pj = mz.polyjectory([tcs, tcs[:3]], [tms, tms[:4]], [0, 5])
    assert pj.n_objs == 2 and pj.poly_order == 20 and pj.maxT == tms[-1]
    tr, tm, st = pj[1]
    assert tr.shape == (3, 21, 7) and np.array_equal(tm, tms[:4]) and st == 5
    assert np.array_equal(pj.status, [0, 5])
    with pytest.raises(ValueError):
        tr[0, 0, 0] = 1.0
    with pytest.raises(IndexError):
        pj[2]


@pytest.mark.parametrize(
    "trajs,times,status,msg",
    [
        ([], [], [], "Cannot initialise a polyjectory object from an empty list of trajectories"),
        ([np.zeros((1, 3, 6))], [np.array([0.0, 1.0])], [0], "must have a size of 7 in the third dimension"),
        ([np.zeros((1, 3))], [np.array([0.0, 1.0])], [0], "A trajectory array must have 3 dimensions"),
        ([np.zeros((1, 2, 7))], [np.array([0.0, 1.0])], [0], "is less than 2 - this is not allowed"),
        ([np.zeros((0, 3, 7))], [np.zeros(0)], [0], "number of steps equal to zero"),
        ([np.zeros((2, 3, 7))], [np.array([0.0, 1.0])], [0], "the number of times is 2"),
        ([np.zeros((1, 3, 7))], [np.array([0.0, np.nan])], [0], "A non-finite time coordinate"),
        ([np.zeros((1, 3, 7))], [np.array([-1.0, 1.0])], [0], "A negative time coordinate"),
        ([np.zeros((2, 3, 7))], [np.array([0.0, 1.0, 1.0])], [0], "is not monotonically increasing"),
        ([np.full((1, 3, 7), np.inf)], [np.array([0.0, 1.0])], [0], "A non-finite value was found in the trajectory"),
        ([np.zeros((1, 3, 7)), np.zeros((1, 4, 7))], [np.array([0.0, 1.0])] * 2, [0, 0], "is inconsistent with"),
        ([np.zeros((1, 3, 7))], [np.array([0.0, 1.0])] * 2, [0], "inconsistent with the number of objects deduced"),
    ],
)
def test_polyjectory_validation(trajs, times, status, msg):
    with pytest.raises(ValueError) as ei:
        mz.polyjectory(trajs, times, status)
    assert msg in str(ei.value)


def test_synthetic_generator():
    pj = mz.make_synthetic_polyjectory(200, 0.5, seed=4, frac_high=0.5)
    assert pj.n_objs == 200 and pj.poly_order == 8 and pj.maxT == 0.5
    for o in range(0, 200, 17):
        tr, tm, _ = pj[o]
        assert tm[0] == 0.0 and tm[-1] == 0.5 and np.all(np.diff(tm) > 0)
        # steps of about an eighth of a period (LEO / MEO-low: 90-130 min)
        assert np.median(np.diff(tm)) * 1440.0 > 5.0
        for s in range(len(tr) - 1):
            h = tm[s + 1] - tm[s]
            end = [np.polyval(tr[s][::-1, ch], h) for ch in range(7)]
            start = tr[s + 1][0]
            assert np.max(np.abs(np.array(end[:3]) - start[:3])) < 1e-6  # continuity, km
            assert abs(end[6] - start[6]) < 1e-6
        # r channel is the norm of the position
        pos = tr[:, 0, :3]
        np.testing.assert_allclose(tr[:, 0, 6], np.linalg.norm(pos, axis=1), rtol=1e-10)
        # km, km/s, LEO-MEO radii
        assert np.all((tr[:, 0, 6] > 6500) & (tr[:, 0, 6] < 9000))
        assert np.all(np.linalg.norm(tr[:, 0, 3:6], axis=1) < 9)
    # reproducible
    pj2 = mz.make_synthetic_polyjectory(200, 0.5, seed=4, frac_high=0.5)
    assert np.array_equal(pj._flat()[0], pj2._flat()[0])


def test_bench_rejects_multi_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_bench_reads_the_committed_profiles():
    # the roofline fields of every bench leg come from rows of the committed
    # rocprofv3 summaries; a row the parser misses would leave them null
    sys.path.insert(0, str(ROOT))
    import bench

    for k in ("k_bp", "k_aabbs", "k_lbvh_build"):
        assert bench.hbm_traffic(k) > 0
    assert bench.eval_hbm_traffic() > 0
    assert bench.sgp4_flops_per_launch() > 1e9


def test_check_screen_without_device():
    """The constructor's validation runs on the host before any device is
    touched: the reference's messages in the reference's order
    (conjunctions.cpp:267-332), then this build's size limit (objects times
    conjunction steps below 2^32)."""
    assert mz.check_screen(1000, 1.0, 5.0, 3.0 / 1440.0) == 480
    assert mz.check_screen(3, 1.0, 5.0, 0.3) == 4
    # threshold before interval before object types
    with pytest.raises(ValueError, match="The conjunction threshold must be finite and positive"):
        mz.check_screen(2, 1.0, -1.0, float("nan"), otypes=[7])
    with pytest.raises(ValueError, match="is too large and results in an overflow error"):
        mz.check_screen(2, 1.0, 1e200, 0.1)
    with pytest.raises(ValueError, match="The conjunction detection interval must be finite and positive"):
        mz.check_screen(2, 1.0, 1.0, 0.0, otypes=[7])
    with pytest.raises(ValueError, match="the expected size is 2, but the actual size is 1 instead"):
        mz.check_screen(2, 1.0, 1.0, 0.1, otypes=[7])
    with pytest.raises(ValueError, match="must be one of \\[1, 2, 4\\], but a value of 7 was detected"):
        mz.check_screen(2, 1.0, 1.0, 0.1, otypes=[1, 7])
    with pytest.raises(ValueError, match="must not exceed 4294967295"):
        mz.check_screen(100000, 30.0, 5.0, 1.0 / 86400.0)
    # conjunctions() raises the same before it needs a device
    tcs, tms = H.planar_circ()
    pj = mz.polyjectory([tcs, tcs], [tms, tms], [0, 0])
    with pytest.raises(ValueError, match="2 objects over 3000000000 conjunction steps is too large"):
        mz.conjunctions(pj, 0.1, pj.maxT / 3.0e9)
    with pytest.raises(ValueError, match="must not exceed 4294967295"):
        mz.conjunctions_batch([pj], [0.1], [pj.maxT / 3.0e9])
