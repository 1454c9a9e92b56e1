"""GPU polyjectory evaluation (state_eval / state_meval) against the oracle
and the reference's known answers (test/test_polyjectory.py:792-1260).

Evaluations are bit-exact against the oracle (same step search, same Horner
sequence, separately rounded operations): compared bit for bit, NaN included.
"""

import numpy as np
import pytest

import mizuba_amd as mz
from oracle import pyoracle as po

from . import helpers as H

pytestmark = pytest.mark.gpu

OBJ0_011 = [0.89, 0.0, 0.0, -1.0, 0.0, 0.0, 0.0]
OBJ0_012 = [0.88, 0.0, 0.0, -1.0, 0.0, 0.0, 0.0]
OBJ1_120 = [0.20, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0]
OBJ1_121 = [0.21, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0]


def close(a, b):
    np.testing.assert_allclose(a, b, rtol=0.0, atol=1e-15)


@pytest.fixture(scope="module")
def pj2():
    trajs, times = H.eval_pj_arrays()
    return mz.polyjectory(trajs, times, [0, 0])


def test_reference_known_answers(pj2):
    pj = pj2
    res = pj.state_eval(time=np.array([0.11, 0.11]))
    close(res[0], OBJ0_011)
    assert np.all(np.isnan(res[1]))
    out = np.zeros((2, 7))
    res = pj.state_eval(time=np.array([0.11, 0.11]), out=out)
    assert res is out
    close(res[0], OBJ0_011)
    assert np.all(np.isnan(res[1]))
    res = pj.state_meval(time=np.array([0.11, 0.12]))
    close(res[0], [OBJ0_011, OBJ0_012])
    assert np.all(np.isnan(res[1]))
    out = np.zeros((2, 2, 7))
    res = pj.state_meval(out=out, time=np.array([0.11, 0.12]))
    assert res is out
    close(res[0], [OBJ0_011, OBJ0_012])
    res = pj.state_meval(time=np.array([[0.11], [0.12]]))
    close(res[0], [OBJ0_011])
    assert np.all(np.isnan(res[1]))
    res = pj.state_eval(time=np.array([0.11, 0.12] * 1000), obj_idx=[0, 1] * 1000)
    close(res[::2], [OBJ0_011] * 1000)
    assert np.all(np.isnan(res[1::2]))
    res = pj.state_meval(time=np.array([0.11, 0.12] * 1000), obj_idx=[0, 1] * 1000)
    close(res[::2], [[OBJ0_011, OBJ0_012] * 1000] * 1000)
    assert np.all(np.isnan(res[1::2]))
    res = pj.state_meval(time=np.array([[0.11, 0.12] * 1000, [1.2, 1.21] * 1000]), obj_idx=[0, 1])
    close(res[0], [OBJ0_011, OBJ0_012] * 1000)
    close(res[1], [OBJ1_120, OBJ1_121] * 1000)
    res = pj.state_meval(time=np.array([[0.11, 0.12] * 1000, [1.2, 1.21] * 1000] * 1000), obj_idx=[0, 1] * 1000)
    close(res[::2], [[OBJ0_011, OBJ0_012] * 1000] * 1000)
    close(res[1::2], [[OBJ1_120, OBJ1_121] * 1000] * 1000)
    res = pj.state_eval(time=np.array([1.2, 1.21]))
    assert np.all(np.isnan(res[0]))
    close(res[1], OBJ1_121)
    res = pj.state_eval(time=np.array([0.11, 1.21]))
    close(res[0], OBJ0_011)
    close(res[1], OBJ1_121)
    # a scalar time and a single index
    close(pj.state_eval(0.11, obj_idx=0)[0], OBJ0_011)


def test_reference_errors(pj2):
    pj = pj2
    with pytest.raises(ValueError) as ei:
        pj.state_eval(time=np.full((2,), float("inf")))
    assert "An non-finite evaluation time of " in str(ei.value)
    with pytest.raises(ValueError) as ei:
        pj.state_eval(float("nan"))
    assert "An non-finite evaluation time of " in str(ei.value)
    with pytest.raises(IndexError) as ei:
        pj.state_eval(time=0.11, obj_idx=2)
    assert "Invalid object index 2 specified - the total number of objects in the polyjectory is only 2" in str(
        ei.value)
    with pytest.raises(IndexError) as ei:
        pj.state_eval(time=0.11, obj_idx=[45])
    assert "Invalid object index 45 specified - the total number of objects in the polyjectory is only 2" in str(
        ei.value)
    # the first failing evaluation is reported, and its time before its index
    with pytest.raises(IndexError) as ei:
        pj.state_meval(time=np.array([0.1, 0.2]), obj_idx=[0, 7, 9])
    assert "Invalid object index 7 specified" in str(ei.value)
    with pytest.raises(ValueError) as ei:
        pj.state_meval(time=np.array([[0.1, 0.2], [np.nan, 0.3], [-np.inf, 0.0]]), obj_idx=[0, 9, 1])
    assert "An non-finite evaluation time of nan" in str(ei.value)
    with pytest.raises(IndexError) as ei:
        pj.state_meval(time=np.array([[0.1, 0.2], [0.3, np.nan], [-np.inf, 0.0]]), obj_idx=[0, 9, 1])
    assert "Invalid object index 9 specified" in str(ei.value)


@pytest.fixture(scope="module")
def synth():
    pj = mz.make_synthetic_polyjectory(3000, 1.0, seed=5, frac_high=0.2)
    traj, time, traj_off, nsteps, time_off = pj._flat()
    flat = po.Flat.from_arrays(pj.poly_order, traj, time, nsteps, traj_off, time_off, pj.maxT)
    return pj, flat


def bits_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


def test_parity_all_modes(synth):
    pj, flat = synth
    rng = np.random.default_rng(3)
    N = pj.n_objs
    # times across the whole span, past both ends, and exactly at step starts
    t0 = np.concatenate([rng.uniform(-0.05, 1.05, 40), [0.0, 1.0, 0.5]])
    # state_eval, one scalar time
    for t in (0.0, 0.3, 1.0, 0.9999999):
        st, ref, _ = po.state_meval(flat, [t], po.TIME_SCALAR)
        assert st == 0
        assert bits_equal(pj.state_eval(t), ref[:, 0])
    # state_eval, one time per object
    tpo = rng.uniform(0, 1, N)
    st, ref, _ = po.state_meval(flat, tpo, po.TIME_PER_OBJECT)
    assert bits_equal(pj.state_eval(tpo), ref[:, 0])
    # state_meval, shared times
    st, ref, _ = po.state_meval(flat, t0, po.TIME_SHARED, n_times=len(t0))
    assert bits_equal(pj.state_meval(t0), ref)
    # state_meval, per-object times, with a selector (repeats, any order)
    sel = rng.integers(0, N, 700).astype(np.uintp)
    t2 = rng.uniform(0, 1, (700, 13))
    st, ref, _ = po.state_meval(flat, t2, po.TIME_PER_EVAL, sel=sel, n_times=13)
    assert bits_equal(pj.state_meval(t2, obj_idx=sel), ref)


def test_step_starts_and_ends(synth):
    pj, flat = synth
    # at a step start the state is the step's constant coefficients; at the
    # final time it is NaN (the range is [t_begin, t_end))
    for o in (0, 17, 2999):
        tr, tm, _ = pj[o]
        res = pj.state_meval(np.ascontiguousarray(tm), obj_idx=o)[0]
        for s in range(len(tm) - 1):
            assert np.array_equal(res[s], tr[s, 0, :])
        assert np.all(np.isnan(res[-1]))


def test_large_batch(synth):
    pj, flat = synth
    # 3000 objects x 256 shared times: 768 k evaluations
    t = np.linspace(0.0, 1.0, 256, endpoint=False)
    res = pj.state_meval(t)
    st, ref, _ = po.state_meval(flat, t, po.TIME_SHARED, n_times=256)
    assert bits_equal(res, ref)
    # positions stay within the fitted orbits (radius channel = |r|)
    r = np.sqrt(np.sum(res[:, :, :3] ** 2, axis=2))
    np.testing.assert_allclose(r, res[:, :, 6], rtol=1e-9)


def test_device_entry_point(synth):
    # mzh_state_eval_device: times, selector and results in HBM (buffers from
    # mzh_device_alloc)
    pj, flat = synth
    scr = mz.Screener(0)
    scr.upload(pj)
    t = np.linspace(0.0, 1.0, 33)
    d_tm = scr.device_buffer(t.nbytes)
    d_tm.write(t)
    d_out = scr.device_buffer(pj.n_objs * 33 * 7 * 8)
    scr.state_eval_device(d_tm.ptr, 2, 0, pj.n_objs, 33, d_out.ptr)
    got = np.empty((pj.n_objs, 33, 7))
    d_out.read(got)
    assert bits_equal(got, pj.state_meval(t))
    assert scr.eval_time() > 0
    rng = np.random.default_rng(9)
    sel = rng.integers(0, pj.n_objs, 500).astype(np.uint64)
    t2 = rng.uniform(0, 1, (500, 7))
    d_sel = scr.device_buffer(sel.nbytes)
    d_sel.write(sel)
    d_t2 = scr.device_buffer(t2.nbytes)
    d_t2.write(t2)
    d_o2 = scr.device_buffer(500 * 7 * 7 * 8)
    scr.state_eval_device(d_t2.ptr, 3, d_sel.ptr, 500, 7, d_o2.ptr)
    st, ref, _ = po.state_meval(flat, t2, po.TIME_PER_EVAL, sel=sel, n_times=7)
    got2 = np.empty((500, 7, 7))
    d_o2.read(got2)
    assert bits_equal(got2, ref)
    # errors read back from device memory
    sel[123] = pj.n_objs + 5
    d_sel.write(sel)
    with pytest.raises(IndexError) as ei:
        scr.state_eval_device(d_t2.ptr, 3, d_sel.ptr, 500, 7, d_o2.ptr)
    assert f"Invalid object index {pj.n_objs + 5} specified" in str(ei.value)
