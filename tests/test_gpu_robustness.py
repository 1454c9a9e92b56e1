"""GPU tests of the paths that ordinary screenings rarely take.

- The root finder (run_poly_root_finding, poly_utils.cpp:272-460) on
  free-standing polynomials, against the oracle bit for bit: with the private
  48-entry work list, and with that list cut to two entries so that almost
  every isolation continues on the slow path with the reference's 250-entry
  list in global memory.
- Screenings whose growable buffers start at one entry, so that every
  resize-and-rerun path runs (broad-phase overflow pool, step-pair,
  candidate, slow-path and conjunction queues), with the narrow phase in
  passes and the slow root-isolation path forced: bit-identical pairs,
  conjunctions and counters.
"""

import numpy as np
import pytest

import mizuba_amd as mz

pytestmark = pytest.mark.gpu

TINY = {
    "MIZUBA_AMD_INIT_CONJ_CAP": "1",
    "MIZUBA_AMD_INIT_CAND_CAP": "1",
    "MIZUBA_AMD_INIT_STEPS_CAP": "1",
    "MIZUBA_AMD_INIT_SLOW_CAP": "1",
    "MIZUBA_AMD_INIT_OVF_CAP": "1",
    "MIZUBA_AMD_NARROW_PASS_PAIRS": "3001",
    "MIZUBA_AMD_WL_FAST_CAP": "2",
}


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def _poly_cases(order, n, seed):
    """Polynomials like the derivative of a distance square: a cluster of two
    real roots near the right end of [0, 1) (separations down to 1e-17),
    times quadratic factors without real roots; and polynomials with several
    real roots spread over [0, 1)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        c = np.array([1.0])
        if k % 2 == 0:
            r = 1.0 - 10.0 ** rng.uniform(-16, -4)
            sep = 10.0 ** rng.uniform(-17, -6)
            for rt in (r, r - sep):
                c = np.convolve(c, [1.0, -rt])
        else:
            for rt in rng.uniform(0.0, 1.0, rng.integers(2, min(order, 6) + 1)):
                c = np.convolve(c, [1.0, -rt])
        while len(c) + 2 <= order + 1:
            a, b = rng.uniform(-3, 3), rng.uniform(0.1, 2)
            c = np.convolve(c, [1.0, -2 * a, a * a + b * b])
        p = np.zeros(order + 1)
        p[: len(c)] = c[::-1] * rng.uniform(0.5, 2.0)
        out.append(p)
    return np.array(out)


@pytest.mark.parametrize("order", [5, 8, 20])
def test_root_finding_matches_oracle(order, monkeypatch):
    from oracle import pyoracle as po

    polys = _poly_cases(order, 400, seed=order)
    rf = np.full(len(polys), 1.0)
    rf[1::3] = 0.0021  # a 3-minute step in days
    ok, ref = [], []
    for k, p in enumerate(polys):
        try:
            ref.append(po.root_finding(p, rf[k], 1)[0])
            ok.append(k)
        except ValueError:
            pass
    assert len(ok) > 300
    scr = mz.Screener(0)
    P, R = polys[ok], rf[ok]
    roots, cnt, flags = scr.root_finding(P, R)
    for k in range(len(ok)):
        assert cnt[k] == len(ref[k]), k
        assert np.array_equal(_bits(roots[k, : cnt[k]]), _bits(ref[k])), k
    assert not np.any(flags & 2)
    # the same with the private list cut to two entries: the slow path
    monkeypatch.setenv("MIZUBA_AMD_WL_FAST_CAP", "2")
    roots2, cnt2, flags2 = scr.root_finding(P, R)
    assert np.count_nonzero(flags2 & 2) > len(ok) // 4
    assert np.array_equal(cnt2, cnt)
    assert np.array_equal(_bits(roots2), _bits(roots))
    # the oracle's failures fail on the device too, with the same kind of error
    bad = [k for k in range(len(polys)) if k not in set(ok)]
    for k in bad[:10]:
        with pytest.raises(ValueError):
            scr.root_finding(polys[k : k + 1], rf[k : k + 1])


def _dense_duplicates():
    pj = mz.make_synthetic_polyjectory(400, 0.1, seed=9, frac_high=0.3)
    traj, tm, toff, ns, tmo = pj._flat()
    op1 = pj.poly_order + 1
    rng = np.random.default_rng(1)
    idx = np.concatenate([np.arange(400), rng.integers(0, 400, 300), np.repeat([3, 17], 12)])
    trajs, times = [], []
    for o in idx:
        n = int(ns[o])
        trajs.append(traj[int(toff[o]) : int(toff[o]) + n * op1 * 7].reshape(n, op1, 7))
        times.append(tm[int(tmo[o]) : int(tmo[o]) + n + 1])
    return mz.polyjectory(trajs, times, [0] * len(idx)), 100.0


@pytest.mark.parametrize("case", ["leo1000", "dense_duplicates"])
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_tiny_buffers_rerun(case, mode, monkeypatch):
    if case == "leo1000":
        pj, thresh = mz.make_synthetic_polyjectory(1000, 1.0, seed=11, frac_high=0.0), 5.0
    else:
        pj, thresh = _dense_duplicates()
    interval = 3.0 / 1440.0
    c1 = mz.conjunctions(pj, thresh, interval, mode=mode)
    for k, v in TINY.items():
        monkeypatch.setenv(k, v)
    c2 = mz.conjunctions(pj, thresh, interval, mode=mode)
    assert len(c1.conjunctions) > 0
    assert np.array_equal(_bits(c1.conjunctions), _bits(c2.conjunctions))
    assert c1.np_report == c2.np_report
    for k in range(c1.n_cd_steps):
        assert np.array_equal(_bits(c1.get_aabb_collisions(k)), _bits(c2.get_aabb_collisions(k))), k


def test_many_conjunctions_in_one_step():
    """500 objects on straight lines through a ball of radius 1 at t = 1, in
    random directions: every pair has its closest approach within the
    threshold, so one conjunction step holds about 1.2e5 conjunctions. They
    must come out as the oracle orders them (tca, then lexicographically)."""
    from oracle import pyoracle as po

    rng = np.random.default_rng(5)
    n = 500
    p0 = rng.uniform(-1.0, 1.0, (n, 3)) / np.sqrt(3.0)
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    tm = np.array([0.0, 1.0, 2.0])
    trajs = []
    for k in range(n):
        steps = []
        for s in range(2):
            c = np.zeros((4, 7))
            at = p0[k] + v[k] * (tm[s] - 1.0)  # position at the step's start
            c[0, :3] = at
            c[1, :3] = v[k]
            c[0, 3:6] = v[k]
            # radius channel: a rough enclosure is enough for the boxes; keep it
            # the exact norm at the start and its derivative
            c[0, 6] = np.linalg.norm(at)
            c[1, 6] = float(at @ v[k]) / max(np.linalg.norm(at), 1e-12)
            steps.append(c)
        trajs.append(np.array(steps))
    pj = mz.polyjectory(trajs, [tm] * n, [0] * n)
    c = mz.conjunctions(pj, 3.0, 2.0)
    ref = po.run((trajs, [tm] * n), 3.0, 2.0)
    g, r = c.conjunctions, ref["conjunctions"]
    assert c.n_cd_steps == 1 and len(r) > 100000
    assert len(g) == len(r)
    assert np.array_equal(_bits(g), _bits(r))
    assert np.all(np.diff(g["tca"]) >= 0)


def test_shared_device_copy_and_hint_release():
    """Screenings of one polyjectory share its device copy; hint_release moves
    a screening's results to host memory and frees the device copies, after
    which every accessor still returns the same arrays; releasing the
    polyjectory's device copy makes the next screening upload it again."""
    pj = mz.make_synthetic_polyjectory(600, 0.3, seed=17, frac_high=0.2)
    c1 = mz.conjunctions(pj, 10.0, 3.0 / 1440.0)
    c2 = mz.conjunctions(pj, 10.0, 3.0 / 1440.0, mode="mixed")
    ref = {
        name: getattr(c1, name).copy()
        for name in ("aabbs", "srt_aabbs", "mcodes", "srt_mcodes", "srt_idx", "conjunctions", "cd_end_times")
    }
    trees = [c1.get_bvh_tree(k).copy() for k in range(c1.n_cd_steps)]
    pairs = [c1.get_aabb_collisions(k).copy() for k in range(c1.n_cd_steps)]
    rep = c1.np_report
    c3 = mz.conjunctions(pj, 10.0, 3.0 / 1440.0)
    c3.hint_release()
    for name, a in ref.items():
        assert np.array_equal(_bits(getattr(c3, name)), _bits(a)), name
    for k in range(c3.n_cd_steps):
        assert np.array_equal(_bits(c3.get_bvh_tree(k)), _bits(trees[k]))
        assert np.array_equal(_bits(c3.get_aabb_collisions(k)), _bits(pairs[k]))
    assert c3.np_report == rep
    with pytest.raises(IndexError):
        c3.get_bvh_tree(c3.n_cd_steps)
    c3.hint_release()  # again: nothing left to release
    assert np.array_equal(_bits(c2.conjunctions), _bits(ref["conjunctions"]))
    pj.hint_release()
    c4 = mz.conjunctions(pj, 10.0, 3.0 / 1440.0)
    assert np.array_equal(_bits(c4.conjunctions), _bits(ref["conjunctions"]))
    # the device copy left by make_sgp4_polyjectory serves the screening
    g = mz.make_synthetic_gpes(2000, seed=3, frac_high=0.2, epoch_jd=2460496.5)
    spj = mz.make_sgp4_polyjectory(g, 2460496.5, 2460496.75)[0]
    a = mz.conjunctions(spj, 10.0, 3.0 / 1440.0).conjunctions
    scr = mz.Screener(0)
    scr.upload(spj, copy=True)  # an upload of the host copy
    scr.screen(10.0, 3.0 / 1440.0)
    assert len(a) > 0 and np.array_equal(_bits(a), _bits(scr.conjunctions()))
