"""Pins of the oracle's polyjectory evaluation (polyjectory.cpp:1094-1366).

Known answers are those of the reference's test_eval
(test/test_polyjectory.py:792-1260): two objects on the x axis moving at unit
speed, the first over [0, 0.9], the second over [1.1, 2.0]. The arithmetic is
further pinned by numpy.polynomial.polynomial.polyval on random per-step
polynomials, with step boundaries, the first and last instants and
out-of-range times (NaN) as edge cases.
"""

import numpy as np
import pytest

from oracle import pyoracle as po

from . import helpers as H

EV_OK, EV_EINVAL, EV_ERANGE = 0, 1, 2


def eval_pj():
    return H.eval_pj_arrays()


OBJ0_011 = [0.89, 0.0, 0.0, -1.0, 0.0, 0.0, 0.0]
OBJ0_012 = [0.88, 0.0, 0.0, -1.0, 0.0, 0.0, 0.0]
OBJ1_120 = [0.20, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0]
OBJ1_121 = [0.21, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0]


def test_reference_known_answers():
    pj = eval_pj()
    st, out, _ = po.state_meval(pj, [0.11, 0.11], po.TIME_PER_OBJECT)
    assert st == EV_OK
    np.testing.assert_allclose(out[0, 0], OBJ0_011, rtol=0, atol=1e-15)
    assert np.all(np.isnan(out[1]))
    st, out, _ = po.state_meval(pj, [0.11, 0.12], po.TIME_SHARED, n_times=2)
    np.testing.assert_allclose(out[0], [OBJ0_011, OBJ0_012], rtol=0, atol=1e-15)
    assert np.all(np.isnan(out[1]))
    st, out, _ = po.state_meval(pj, [[0.11], [0.12]], po.TIME_PER_EVAL, n_times=1)
    np.testing.assert_allclose(out[0, 0], OBJ0_011, rtol=0, atol=1e-15)
    assert np.all(np.isnan(out[1]))
    st, out, _ = po.state_meval(pj, [0.11, 0.12] * 1000, po.TIME_PER_OBJECT, sel=[0, 1] * 1000)
    np.testing.assert_allclose(out[::2, 0], [OBJ0_011] * 1000, rtol=0, atol=1e-15)
    assert np.all(np.isnan(out[1::2]))
    tm2 = np.array([[0.11, 0.12] * 1000, [1.2, 1.21] * 1000])
    st, out, _ = po.state_meval(pj, tm2, po.TIME_PER_EVAL, sel=[0, 1], n_times=2000)
    np.testing.assert_allclose(out[0], [OBJ0_011, OBJ0_012] * 1000, rtol=0, atol=1e-15)
    np.testing.assert_allclose(out[1], [OBJ1_120, OBJ1_121] * 1000, rtol=0, atol=1e-15)
    st, out, _ = po.state_meval(pj, [1.2, 1.21], po.TIME_PER_OBJECT)
    assert np.all(np.isnan(out[0]))
    np.testing.assert_allclose(out[1, 0], OBJ1_121, rtol=0, atol=1e-15)
    st, out, _ = po.state_meval(pj, [0.11, 1.21], po.TIME_PER_OBJECT)
    np.testing.assert_allclose(out[0, 0], OBJ0_011, rtol=0, atol=1e-15)
    np.testing.assert_allclose(out[1, 0], OBJ1_121, rtol=0, atol=1e-15)


def test_reference_errors():
    pj = eval_pj()
    for bad_t in (np.inf, np.nan):
        st, _, e = po.state_meval(pj, [bad_t, bad_t], po.TIME_PER_OBJECT)
        assert (st, e) == (EV_EINVAL, 0)
    st, _, e = po.state_meval(pj, [0.11], po.TIME_SCALAR, sel=[2])
    assert (st, e) == (EV_ERANGE, 0)
    st, _, e = po.state_meval(pj, [0.11], po.TIME_SCALAR, sel=[0, 1, 45])
    assert (st, e) == (EV_ERANGE, 2)
    # the time is checked before the index (polyjectory.cpp:1097-1104)
    st, _, e = po.state_meval(pj, [np.nan], po.TIME_SCALAR, sel=[45])
    assert (st, e) == (EV_EINVAL, 0)


def random_pj(rng, n_objs=5, order=6):
    trajs, times = [], []
    for o in range(n_objs):
        ns = int(rng.integers(0, 6)) if o else 4
        t = np.cumsum(np.concatenate([[rng.uniform(0, 1)], rng.uniform(0.1, 1.0, ns)]))
        trajs.append(rng.standard_normal((ns, order + 1, 7)))
        times.append(t if ns else np.zeros(0))
    return trajs, times


def test_random_polynomials_against_polyval():
    rng = np.random.default_rng(7)
    trajs, times = random_pj(rng)
    pj = (trajs, times)
    for o, (tr, t) in enumerate(zip(trajs, times)):
        ns = tr.shape[0]
        if ns == 0:
            st, out, _ = po.state_meval(pj, [0.5], po.TIME_SCALAR, sel=[o])
            assert np.all(np.isnan(out))
            continue
        # interior points, every step start, the first instant; the end and outside are NaN
        tq = np.concatenate([rng.uniform(t[0], t[-1], 50), t[:-1], [t[0]]])
        st, out, _ = po.state_meval(pj, tq, po.TIME_SHARED, sel=[o], n_times=len(tq))
        assert st == EV_OK
        for k, tv in enumerate(tq):
            s = int(np.searchsorted(t[1:], tv, side="right"))
            h = tv - t[s]
            exp = [np.polynomial.polynomial.polyval(h, tr[s, :, ch]) for ch in range(7)]
            np.testing.assert_allclose(out[0, k], exp, rtol=1e-13, atol=1e-13)
            if h == 0.0:
                assert np.array_equal(out[0, k], tr[s, 0, :])  # value at a step start is the constant term
        st, out, _ = po.state_meval(pj, [t[-1], np.nextafter(t[0], -np.inf), t[-1] + 1.0], po.TIME_SHARED,
                                    sel=[o], n_times=3)
        assert st == EV_OK and np.all(np.isnan(out))


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_time_modes_agree(mode):
    rng = np.random.default_rng(11)
    pj = random_pj(rng, n_objs=6)
    sel = np.array([3, 0, 0, 5, 2], dtype=np.uint64)
    nt = 4
    base = rng.uniform(0, 3, (len(sel), nt))
    tm = {0: base[:1, :1].ravel(), 1: base[:, 0], 2: base[0], 3: base}[mode]
    st, out, _ = po.state_meval(pj, tm, mode, sel=sel, n_times=1 if mode in (0, 1) else nt)
    assert st == EV_OK
    for r, o in enumerate(sel):
        for k in range(out.shape[1]):
            t = (lambda: tm[0], lambda: tm[r], lambda: tm[k], lambda: base[r, k])[mode]()
            _, one, _ = po.state_meval(pj, [t], po.TIME_SCALAR, sel=[o])
            assert np.array_equal(out[r, k], one[0, 0], equal_nan=True)
