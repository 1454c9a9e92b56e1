"""GPU runs of the configurations at their full sizes, on catalogs made the
configurations' way (seeded SGP4 mean elements through make_sgp4_polyjectory)
and, for the dense-shell case, also on the Keplerian stand-in catalog:
1 000 LEO objects over a day and the FP64 drop-in at 20 000 objects, checked
against the oracle on every step; 30 000 objects over 7 days (3 360
conjunction steps), checked against the oracle on every step window by
window; the dense-shell stress case (100 000 objects, 70 000 of them in five
near-circular shells, 10 km, 2-minute steps).

The CPU oracle cannot screen these in a test's time, so they are checked
through properties that hold at any size, and against the oracle on windows
of conjunction steps (the oracle screens any step range on its own):
- the FP64 screen equals the oracle bit for bit on the sampled windows (boxes,
  codes, sort order, trees, pairs, conjunctions);
- the mixed-precision conjunction list equals the FP64 one bit for bit;
- pairs are sorted, i < j; conjunctions have DCA below the threshold and DCA
  equal to the distance of their recorded positions;
- the narrow-phase counters add up: candidates = distance rejects + root
  findings; minima - discarded minima = conjunctions, except those found at
  the ends of trajectories (boundary minima, which the reference does not
  count as minima: detect_conjunctions_narrow_phase.cpp:629-714).
"""

import numpy as np
import pytest

import mizuba_amd as mz

from . import helpers as H

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def check_window(c, pj, thresh, interval, k0, k1):
    from oracle import pyoracle as po

    ref = po.run(H.flat_of(pj), thresh, interval, step_begin=k0, step_end=k1)
    for name in ("aabbs", "srt_aabbs", "mcodes", "srt_mcodes", "srt_idx"):
        g = getattr(c, name)[k0:k1]
        assert np.array_equal(_bits(g), _bits(ref[name])), (name, k0, k1)
    for k in range(k1 - k0):
        assert np.array_equal(_bits(c.get_bvh_tree(k0 + k)), _bits(po.get_bvh_tree(ref, k))), k0 + k
        assert np.array_equal(_bits(c.get_aabb_collisions(k0 + k)), _bits(po.get_aabb_collisions(ref, k))), k0 + k
    # conjunctions found in the window's steps
    t0 = 0.0 if k0 == 0 else c.cd_end_times[k0 - 1]
    t1 = c.cd_end_times[k1 - 1]
    cj = c.conjunctions
    # the last step also owns conjunctions at the very end of the time range
    upper = (cj["tca"] <= t1) if k1 == c.n_cd_steps else (cj["tca"] < t1)
    g = cj[(cj["tca"] >= t0) & upper]
    assert np.array_equal(_bits(g), _bits(ref["conjunctions"])), (k0, k1, len(g), len(ref["conjunctions"]))
    return len(g)


def check_properties(c64, cmx, thresh):
    # mixed = FP64, bit for bit
    assert np.array_equal(_bits(c64.conjunctions), _bits(cmx.conjunctions))
    cj = c64.conjunctions
    assert len(cj) > 0
    assert np.all(cj["i"] < cj["j"])
    assert np.all(cj["dca"] < thresh)
    d = np.sqrt(np.sum((cj["ri"] - cj["rj"]) ** 2, axis=1))
    np.testing.assert_allclose(cj["dca"], d, rtol=1e-12, atol=1e-12)
    rep = c64.np_report
    assert rep["n_tot_conj_candidates"] == rep["n_dist2_check"] + rep["n_poly_roots"]
    kept = rep["n_tot_dist_minima"] - rep["n_tot_discarded_dist_minima"]
    at_ends = int(np.count_nonzero((cj["tca"] == 0.0) | (cj["tca"] == c64.cd_end_times[-1])))
    assert kept <= len(cj) <= kept + at_ends
    # pairs of a few steps: sorted, i < j
    C = c64.n_cd_steps
    for k in (0, C // 3, C - 1):
        p = c64.get_aabb_collisions(k)
        assert np.all(p["i"] < p["j"])
        key = p["i"].astype(np.uint64) << np.uint64(32) | p["j"].astype(np.uint64)
        assert np.all(np.diff(key.astype(np.int64)) > 0)


def check_all_steps(c, pj, thresh, interval, window):
    """The whole screening against the oracle, window by window (the oracle
    screens any step range on its own): every step's boxes, codes, sort
    order, tree and pairs, and every conjunction."""
    C = c.n_cd_steps
    n_conj = 0
    for k0 in range(0, C, window):
        k1 = min(C, k0 + window)
        n_conj += check_window(c, pj, thresh, interval, k0, k1)
    assert n_conj == len(c.conjunctions)


def test_config_1000_sgp4_full_day():
    # the CPU configuration: 1 000 LEO objects from seeded SGP4 mean elements,
    # 1 day, 5 km, 3 minutes; every step and counter against the oracle
    from oracle import pyoracle as po

    pj = _sgp4_catalog(1000, 1.0, seed=2024, frac_high=0.0)
    thresh, interval = 5.0, 3.0 / 1440.0
    c64 = mz.conjunctions(pj, thresh, interval, mode="fp64")
    cmx = mz.conjunctions(pj, thresh, interval, mode="mixed")
    assert c64.n_cd_steps == 480
    ref = po.run(H.flat_of(pj), thresh, interval)
    for name in ("aabbs", "srt_aabbs", "mcodes", "srt_mcodes", "srt_idx"):
        assert np.array_equal(_bits(getattr(c64, name)), _bits(ref[name])), name
    for k in range(c64.n_cd_steps):
        assert np.array_equal(_bits(c64.get_bvh_tree(k)), _bits(po.get_bvh_tree(ref, k))), k
        assert np.array_equal(_bits(c64.get_aabb_collisions(k)), _bits(po.get_aabb_collisions(ref, k))), k
    assert np.array_equal(_bits(c64.conjunctions), _bits(ref["conjunctions"]))
    assert np.array_equal(_bits(cmx.conjunctions), _bits(ref["conjunctions"]))
    assert c64.np_report == ref["np_report"]


def test_config_30k_7days_sgp4_every_step():
    # the precision configuration: 30 000 objects (80% LEO, 20% higher) from
    # seeded SGP4 mean elements over 7 days (3 360 steps of 3 minutes), 5 km;
    # the FP64 screen equals the oracle on every step, mixed mode gives the
    # same conjunctions
    pj = _sgp4_catalog(30000, 7.0, seed=2026)
    thresh, interval = 5.0, 3.0 / 1440.0
    c64 = mz.conjunctions(pj, thresh, interval, mode="fp64")
    cmx = mz.conjunctions(pj, thresh, interval, mode="mixed")
    assert c64.n_cd_steps == 3360
    check_properties(c64, cmx, thresh)
    del cmx
    check_all_steps(c64, pj, thresh, interval, 420)


def test_config_dense_shells_100k():
    pj = mz.make_synthetic_polyjectory(100000, 1.0, seed=2027, frac_high=0.0, n_shell=70000)
    thresh, interval = 10.0, 2.0 / 1440.0
    c64 = mz.conjunctions(pj, thresh, interval, mode="fp64")
    cmx = mz.conjunctions(pj, thresh, interval, mode="mixed")
    assert c64.n_cd_steps == 720
    check_properties(c64, cmx, thresh)
    check_window(c64, pj, thresh, interval, 0, 2)
    check_window(c64, pj, thresh, interval, 718, 720)


def _sgp4_catalog(n, days, seed, n_shell=0, frac_high=0.2):
    jd0 = 2460496.5
    gpes = mz.make_synthetic_gpes(n, seed=seed, frac_high=frac_high, n_shell=n_shell, epoch_jd=jd0)
    pj = mz.make_sgp4_polyjectory(gpes, jd0, jd0 + days)[0]
    assert np.all(pj.status == 0)
    return pj


def test_config_20k_sgp4_fp64_every_step():
    # the all-FP64 drop-in configuration: 20 000 objects (80% LEO, 20% higher),
    # 1 day, order 8; boxes, codes, sort order, trees, pairs and conjunctions
    # equal the oracle's bit for bit on all 480 steps, and mixed mode gives
    # the same conjunctions
    from oracle import pyoracle as po

    pj = _sgp4_catalog(20000, 1.0, seed=2028)
    thresh, interval = 5.0, 3.0 / 1440.0
    c64 = mz.conjunctions(pj, thresh, interval, mode="fp64")
    cmx = mz.conjunctions(pj, thresh, interval, mode="mixed")
    check_properties(c64, cmx, thresh)
    ref = po.run(H.flat_of(pj), thresh, interval)
    for name in ("aabbs", "srt_aabbs", "mcodes", "srt_mcodes", "srt_idx"):
        assert np.array_equal(_bits(getattr(c64, name)), _bits(ref[name])), name
    for k in range(c64.n_cd_steps):
        assert np.array_equal(_bits(c64.get_bvh_tree(k)), _bits(po.get_bvh_tree(ref, k))), k
        assert np.array_equal(_bits(c64.get_aabb_collisions(k)), _bits(po.get_aabb_collisions(ref, k))), k
    assert np.array_equal(_bits(c64.conjunctions), _bits(ref["conjunctions"]))


def test_config_dense_shells_100k_sgp4():
    pj = _sgp4_catalog(100000, 1.0, seed=2029, n_shell=70000, frac_high=0.0)
    thresh, interval = 10.0, 2.0 / 1440.0
    c64 = mz.conjunctions(pj, thresh, interval, mode="fp64")
    cmx = mz.conjunctions(pj, thresh, interval, mode="mixed")
    assert c64.n_cd_steps == 720
    check_properties(c64, cmx, thresh)
    check_window(c64, pj, thresh, interval, 0, 2)
    check_window(c64, pj, thresh, interval, 359, 361)
    check_window(c64, pj, thresh, interval, 718, 720)

