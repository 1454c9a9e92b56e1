"""GPU parity: the HIP pipeline against the CPU oracle on the same inputs.

Bar (FP64 drop-in mode): AABBs, Morton codes, sort order, BVH nodes and
broad-phase pairs bit-exact; conjunction (i, j) lists identical with
|dTCA| <= 1e-3 s and |dDCA| <= 1e-3 km (and, in practice, bit-exact too --
checked separately). Mixed mode: boxes contain the FP64 boxes, the pair set is
a superset and the conjunction list is identical.
"""

import numpy as np
import pytest

import mizuba_amd as mz

from . import helpers as H

pytestmark = pytest.mark.gpu

DAY_S = 86400.0
TCA_TOL_DAYS = 1e-3 / DAY_S  # 1e-3 s
DCA_TOL_KM = 1e-3


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint8)


def assert_screen_equal(c, ref, trees=True, pairs=True):
    C = c.n_cd_steps
    assert C == ref["n_cd_steps"]
    np.testing.assert_array_equal(c.cd_end_times, ref["cd_end_times"])
    for name in ("aabbs", "srt_aabbs", "mcodes", "srt_mcodes", "srt_idx"):
        g, r = getattr(c, name), ref[name]
        assert g.shape == r.shape, name
        if not np.array_equal(_bits(g), _bits(r)):
            bad = np.argwhere(g != r)
            raise AssertionError(f"{name} differs at {bad[:5].tolist()} ({len(bad)} entries)")
    if trees:
        from oracle import pyoracle as po

        for k in range(C):
            g, r = c.get_bvh_tree(k), po.get_bvh_tree(ref, k)
            assert len(g) == len(r), (k, len(g), len(r))
            assert np.array_equal(_bits(g), _bits(r)), f"tree of step {k} differs"
    if pairs:
        from oracle import pyoracle as po

        for k in range(C):
            g, r = c.get_aabb_collisions(k), po.get_aabb_collisions(ref, k)
            assert len(g) == len(r), (k, len(g), len(r))
            assert np.array_equal(_bits(g), _bits(r)), f"pairs of step {k} differ"


def assert_conj_equal(g, r, bitwise=True):
    assert len(g) == len(r), (len(g), len(r))
    np.testing.assert_array_equal(g["i"], r["i"])
    np.testing.assert_array_equal(g["j"], r["j"])
    if len(g):
        assert np.max(np.abs(g["tca"] - r["tca"])) <= TCA_TOL_DAYS
        assert np.max(np.abs(g["dca"] - r["dca"])) <= DCA_TOL_KM
    if bitwise:
        assert np.array_equal(_bits(g), _bits(r)), "conjunction records differ in their bits"


def run_both(trajs, times, thresh, interval, otypes=None, mode="fp64", status=None):
    from oracle import pyoracle as po

    pj = mz.polyjectory(trajs, times, status if status is not None else [0] * len(trajs))
    c = mz.conjunctions(pj, thresh, interval, otypes=otypes, mode=mode)
    ref = po.run((trajs, times), thresh, interval, otypes=otypes)
    return c, ref


# ---- reference closed-form cases -------------------------------------------


def test_zero_aabbs_exact():
    trajs, times = H.zero_aabbs_pj()
    c, ref = run_both(trajs, times, float(np.finfo(float).eps), 0.1)
    a = c.aabbs
    assert np.all(a[:, :, 0, :3] == np.nextafter(np.single(1), np.single("-inf")))
    assert np.all(a[:, :, 1, :3] == np.nextafter(np.single(1), np.single("+inf")))
    assert np.all(a[:, :, 0, 3] == np.nextafter(np.single(np.sqrt(3.0)), np.single("-inf")))
    assert np.all(a[:, :, 1, 3] == np.nextafter(np.single(np.sqrt(3.0)), np.single("+inf")))
    assert_screen_equal(c, ref)


@pytest.mark.parametrize("interval", [0.01, 0.1, 0.5, 2.0, 5.0, 7.0, 42.0])
def test_planar_circ(interval):
    tcs, tms = H.planar_circ()
    c, ref = run_both([tcs], [tms], 0.1, interval)
    assert c.cd_end_times[-1] == tms[-1]
    a = c.aabbs
    assert np.all(a[:, 0] == a[:, 1])
    assert np.all(a == c.srt_aabbs)
    assert np.all(a[:, 0, 0, 2] >= -0.05001) and np.all(a[:, 0, 1, 2] <= 0.05001)
    assert np.all(a[:, 0, 0, 3] >= 1 - 0.05001) and np.all(a[:, 0, 1, 3] <= 1 + 0.05001)
    assert len(c.conjunctions) == 0
    assert_screen_equal(c, ref)


@pytest.mark.parametrize("case", range(6))
def test_cd_begin_end(case):
    trajs, times, exp = H.cd_begin_end_case(case)
    c, ref = run_both(trajs, times, 1e-6, 2.0 / 3.0)
    cj = c.conjunctions
    assert_conj_equal(cj, ref["conjunctions"])
    assert_screen_equal(c, ref)
    if exp == 0:
        assert len(cj) == 0
    else:
        assert len(cj) >= 1
        assert np.all(cj["i"] == 0) and np.all(cj["j"] == 1)
        assert abs(cj["tca"][0] - 1.0) < 1e-15
        assert abs(cj["dca"][0]) < 1e-15
        assert np.allclose(cj["vi"][0], [-1, 0, 0], atol=1e-15, rtol=0.0)
        assert np.allclose(cj["vj"][0], [1, 0, 0], atol=1e-15, rtol=0.0)


@pytest.mark.parametrize("shift", [0.0, 1.1])
def test_boundary_incoming(shift):
    trajs, times = H.boundary_incoming(shift)
    c, ref = run_both(trajs, times, 0.25, 0.03)
    cj = c.conjunctions
    assert len(cj) == 1
    assert abs(cj["tca"][0] - (0.9 + shift)) < 1e-15
    assert abs(cj["dca"][0] - 0.2) < 1e-15
    assert np.allclose(cj["ri"][0], [0.1, 0, 0], rtol=1e-15, atol=0.0)
    assert np.allclose(cj["rj"][0], [-0.1, 0, 0], rtol=1e-15, atol=0.0)
    assert_conj_equal(cj, ref["conjunctions"])
    assert_screen_equal(c, ref)


@pytest.mark.parametrize("shift", [0.0, 1.1])
def test_boundary_outgoing(shift):
    trajs, times = H.boundary_outgoing(shift)
    c, ref = run_both(trajs, times, 0.25, 0.03)
    cj = c.conjunctions
    assert len(cj) == 1
    assert abs(cj["tca"][0] - shift) < 1e-15
    assert abs(cj["dca"][0] - 0.2) < 1e-15
    assert np.allclose(cj["vi"][0], [1, 0, 0], rtol=1e-15, atol=0.0)
    assert_conj_equal(cj, ref["conjunctions"])
    assert_screen_equal(c, ref)


def test_bvh_cases():
    t = np.zeros((6, 7))
    t[1, :] = 0.1
    c, ref = run_both([np.array([t])], [np.array([0.0, 1.0])], 1e-16, 1.0)
    assert len(c.get_bvh_tree(0)) == 1
    with pytest.raises(IndexError) as ei:
        c.get_bvh_tree(1)
    assert (
        "Cannot fetch the BVH tree for the conjunction timestep at index 1: the total number of conjunction steps is only 1"
        in str(ei.value)
    )
    c, ref = run_both([np.array([t]), np.array([t])], [np.array([0.0, 1.0])] * 2, 1e-16, 1.0)
    tr = c.get_bvh_tree(0)
    assert len(tr) == 1 and tr[0]["begin"] == 0 and tr[0]["end"] == 2 and tr[0]["left"] == -1
    assert_screen_equal(c, ref)
    objs = H.axis_objects()
    t8 = np.zeros((6, 7))
    t8[0, :] = 1
    t8[0, 0] = 1.0 - 2.1 / 2**16
    objs.append(t8)
    c, ref = run_both([np.array([o]) for o in objs], [np.array([0.0, 1.0])] * 9, 1e-16, 1.0)
    assert c.mcodes[0, -2] == 2**64 - 1
    assert c.mcodes[0, -1] == 2**64 - 2
    assert c.srt_idx[0, -1] == 7 and c.srt_idx[0, -2] == 8
    assert_screen_equal(c, ref)


def test_no_traj_data():
    objs = H.axis_objects()
    trajs = [np.array([objs[0]])] * 10 + [np.array([o, o]) for o in objs]
    times = [np.array([0.0, 1.0])] * 10 + [np.array([0.0, 1.0, 2.0])] * 8
    c, ref = run_both(trajs, times, 1e-16, 1.0)
    assert np.all(np.isinf(c.aabbs[1, :10]))
    assert np.all(np.isinf(c.srt_aabbs[1, -11:-1]))
    assert np.all(c.mcodes[1, :10] == 2**64 - 1)
    assert np.all(c.srt_mcodes[1, -11:] == 2**64 - 1)
    assert_screen_equal(c, ref)


def test_empty_and_late_trajectories():
    tcs, tms = H.planar_circ()
    empty = np.zeros((0, tcs.shape[1], 7))
    for trajs, times in [
        ([tcs, empty], [tms, np.zeros(0)]),
        ([empty, tcs, empty], [np.zeros(0), tms, np.zeros(0)]),
        ([empty, empty, tcs], [np.zeros(0), np.zeros(0), tms]),
        ([tcs], [tms + 1.0]),
    ]:
        c, ref = run_both(trajs, times, 1.0, 0.1)
        assert len(c.conjunctions) == 0
        assert_screen_equal(c, ref)


# ---- synthetic catalogs ------------------------------------------------------


@pytest.fixture(scope="module")
def leo1000():
    return mz.make_synthetic_polyjectory(1000, 1.0, seed=11, frac_high=0.0)


def test_synthetic_1000_full_day(leo1000):
    from oracle import pyoracle as po

    pj = leo1000
    c = mz.conjunctions(pj, 5.0, 3.0 / 1440.0)
    ref = po.run(H.flat_of(pj), 5.0, 3.0 / 1440.0)
    assert c.n_cd_steps == 480
    assert_screen_equal(c, ref)
    assert_conj_equal(c.conjunctions, ref["conjunctions"])
    assert c.np_report == ref["np_report"]
    assert len(c.conjunctions) > 0


def test_synthetic_mix_otypes():
    from oracle import pyoracle as po

    pj = mz.make_synthetic_polyjectory(3000, 0.25, seed=5, frac_high=0.2)
    rng = np.random.default_rng(3)
    ot = rng.choice([1, 2, 4], size=pj.n_objs, p=[0.5, 0.4, 0.1]).astype(np.int32)
    c = mz.conjunctions(pj, 10.0, 2.0 / 1440.0, otypes=ot)
    ref = po.run(H.flat_of(pj), 10.0, 2.0 / 1440.0, otypes=ot)
    assert_screen_equal(c, ref)
    assert_conj_equal(c.conjunctions, ref["conjunctions"])
    cj = c.conjunctions
    assert np.all(ot[cj["i"]] + ot[cj["j"]] <= 3)
    for ot_all in (np.full(pj.n_objs, 2, np.int32), np.full(pj.n_objs, 4, np.int32)):
        c2 = mz.conjunctions(pj, 10.0, 2.0 / 1440.0, otypes=ot_all)
        assert all(len(c2.get_aabb_collisions(k)) == 0 for k in range(c2.n_cd_steps))
        assert len(c2.conjunctions) == 0


def test_dense_duplicates_many_partners():
    """Copies of the same trajectory (runs of equal Morton codes, i.e.
    multi-object leaves) and a 100 km threshold: many objects have more
    partners than are kept inline, and pairs inside a run of equal codes must
    come out as the reference's (i, j) order gives them."""
    from oracle import pyoracle as po

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
    ot = rng.choice([1, 2], size=len(idx), p=[0.7, 0.3]).astype(np.int32)
    for otypes in (None, ot):
        c, ref = run_both(trajs, times, 100.0, 3.0 / 1440.0, otypes=otypes)
        assert_screen_equal(c, ref)
        assert_conj_equal(c.conjunctions, ref["conjunctions"])
        counts = np.bincount(c.get_aabb_collisions(0)["i"], minlength=len(idx))
        assert counts.max() > 8


@pytest.mark.parametrize("copies", [300, 900, 4300])
def test_sort_large_buckets(copies):
    """Hundreds, then thousands, of copies of one object share a Morton code:
    the sort must handle a bucket far larger than a wave and, past 1024 keys,
    the radix-sort path; equal codes stay in index order. The copies are
    masked so that the pair count stays small."""
    from oracle import pyoracle as po

    pj = mz.make_synthetic_polyjectory(200, 0.02, seed=4, frac_high=0.3)
    traj, tm, toff, ns, tmo = pj._flat()
    op1 = pj.poly_order + 1
    idx = np.concatenate([np.arange(200), np.full(copies, 7)])
    trajs, times = [], []
    for o in idx:
        n = int(ns[o])
        trajs.append(traj[int(toff[o]) : int(toff[o]) + n * op1 * 7].reshape(n, op1, 7))
        times.append(tm[int(tmo[o]) : int(tmo[o]) + n + 1])
    ot = np.ones(len(idx), np.int32)
    ot[200:] = 4
    c, ref = run_both(trajs, times, 5.0, 3.0 / 1440.0, otypes=ot)
    assert_screen_equal(c, ref)
    assert_conj_equal(c.conjunctions, ref["conjunctions"])


def test_batch_equals_individual():
    """Problems of uneven size, threshold and horizon screened together give
    exactly what screening them one by one gives; a failing problem reports
    its own error."""
    rng = np.random.default_rng(8)
    pjs, threshs, intervals, ots = [], [], [], []
    for k, (n, days) in enumerate([(300, 0.3), (2500, 0.2), (900, 0.5), (50, 0.1), (1500, 0.25), (4000, 0.1)]):
        pjs.append(mz.make_synthetic_polyjectory(n, days, seed=100 + k, frac_high=0.2))
        threshs.append(float(rng.uniform(1, 10)))
        intervals.append(float(rng.choice([1.0, 2.0, 3.0])) / 1440.0)
        ots.append(rng.choice([1, 2, 4], size=n, p=[0.6, 0.3, 0.1]).astype(np.int32) if k % 2 else None)
    for mode in ("fp64", "mixed"):
        batch = mz.conjunctions_batch(pjs, threshs, intervals, otypes=ots, mode=mode)
        assert len(batch) == len(pjs)
        for pj, th, iv, ot, cb in zip(pjs, threshs, intervals, ots, batch):
            ci = mz.conjunctions(pj, th, iv, otypes=ot, mode=mode)
            assert cb.n_cd_steps == ci.n_cd_steps
            assert np.array_equal(_bits(cb.srt_aabbs), _bits(ci.srt_aabbs))
            assert np.array_equal(_bits(cb.srt_idx), _bits(ci.srt_idx))
            for k in (0, ci.n_cd_steps - 1):
                assert np.array_equal(_bits(cb.get_bvh_tree(k)), _bits(ci.get_bvh_tree(k)))
                assert np.array_equal(_bits(cb.get_aabb_collisions(k)), _bits(ci.get_aabb_collisions(k)))
            assert np.array_equal(_bits(cb.conjunctions), _bits(ci.conjunctions))
            assert cb.np_report == ci.np_report
    with pytest.raises(ValueError, match="The conjunction threshold must be finite and positive"):
        mz.conjunctions_batch(pjs[:2], [threshs[0], -1.0], intervals[:2])


def test_narrow_phase_in_passes(leo1000, monkeypatch):
    """The narrow phase split into many passes over the pair list (a few
    thousand pairs each, so steps straddle passes) gives the same records and
    counters as one pass."""
    c1 = mz.conjunctions(leo1000, 5.0, 3.0 / 1440.0)
    monkeypatch.setenv("MIZUBA_AMD_NARROW_PASS_PAIRS", "3001")
    c2 = mz.conjunctions(leo1000, 5.0, 3.0 / 1440.0)
    assert len(c2.conjunctions) > 0
    assert np.array_equal(_bits(c1.conjunctions), _bits(c2.conjunctions))
    assert c1.np_report == c2.np_report


def test_mixed_mode_superset(leo1000):
    pj = leo1000
    c64 = mz.conjunctions(pj, 5.0, 3.0 / 1440.0, mode="fp64")
    cmx = mz.conjunctions(pj, 5.0, 3.0 / 1440.0, mode="mixed")
    a64, amx = c64.aabbs, cmx.aabbs
    fin = np.isfinite(a64)
    assert np.array_equal(fin, np.isfinite(amx))
    assert np.all(amx[:, :, 0][fin[:, :, 0]] <= a64[:, :, 0][fin[:, :, 0]])
    assert np.all(amx[:, :, 1][fin[:, :, 1]] >= a64[:, :, 1][fin[:, :, 1]])
    n64 = sum(len(c64.get_aabb_collisions(k)) for k in range(c64.n_cd_steps))
    nmx = sum(len(cmx.get_aabb_collisions(k)) for k in range(cmx.n_cd_steps))
    assert nmx >= n64
    assert_conj_equal(cmx.conjunctions, c64.conjunctions)


def test_mixed_prescreen_near_threshold():
    """Mixed mode's narrow-phase pre-screen drops step pairs only where the
    FP64 narrow phase cannot find a conjunction: SGP4 objects with copies
    shifted along track (mean anomaly) and across (node) so that their
    distances sit around the threshold, from 0.6 to 1.4 times it. The mixed
    list must equal the FP64 list bit for bit, with minima just under the
    threshold among them."""
    thresh = 5.0
    jd0 = 2460496.5
    base = mz.make_synthetic_gpes(150, seed=77, frac_high=0.2, epoch_jd=jd0)
    a = (398600.8 / (base["n0"] / 60.0) ** 2) ** (1.0 / 3.0)
    parts = [base]
    nid = base["norad_id"].max() + 1
    for f in (0.6, 0.9, 0.97, 0.995, 1.0, 1.005, 1.03, 1.1, 1.4):
        g = base.copy()
        g["m0"] = np.mod(g["m0"] + f * thresh / a, 2 * np.pi)
        g["norad_id"] = np.arange(nid, nid + len(g))
        nid += len(g)
        parts.append(g)
        h = base.copy()
        h["node0"] = np.mod(h["node0"] + f * thresh / (a * np.maximum(np.sin(h["i0"]), 0.05)), 2 * np.pi)
        h["norad_id"] = np.arange(nid, nid + len(h))
        nid += len(h)
        parts.append(h)
    gpes = np.concatenate(parts)
    pj = mz.make_sgp4_polyjectory(gpes, jd0, jd0 + 0.25)[0]
    c64 = mz.conjunctions(pj, thresh, 3.0 / 1440.0, mode="fp64")
    cmx = mz.conjunctions(pj, thresh, 3.0 / 1440.0, mode="mixed")
    assert_conj_equal(cmx.conjunctions, c64.conjunctions)
    cj = c64.conjunctions
    assert np.count_nonzero(cj["dca"] > 0.95 * thresh) > 10
    r64, rmx = c64.np_report, cmx.np_report
    # the pre-screen's drops are counted as distance rejections
    assert rmx["n_dist2_check"] >= r64["n_dist2_check"] - (rmx["n_tot_conj_candidates"] - r64["n_tot_conj_candidates"])
    assert rmx["n_tot_conj_candidates"] == rmx["n_dist2_check"] + rmx["n_poly_roots"]


def test_mixed_mode_ill_conditioned_step():
    """A sliver trajectory step whose coefficients overflow FP32 takes the FP64
    enclosure in mixed mode: boxes still contain the FP64 ones."""
    tcs, tms = H.planar_circ()
    sliver = np.zeros((1,) + tcs.shape[1:])
    sliver[0, 0, :] = tcs[-1, 0, :]
    sliver[0, -1, 0] = 1e41  # far outside the float range
    trajs = [np.concatenate([tcs, sliver]), tcs]
    times = [np.append(tms, tms[-1] + 1e-9), tms]
    c64 = mz.conjunctions(mz.polyjectory(trajs, times, [0, 0]), 0.5, 0.7, mode="fp64")
    cmx = mz.conjunctions(mz.polyjectory(trajs, times, [0, 0]), 0.5, 0.7, mode="mixed")
    a64, amx = c64.aabbs, cmx.aabbs
    fin = np.isfinite(a64)
    assert np.array_equal(fin, np.isfinite(amx))
    assert np.all(amx[:, :, 0][fin[:, :, 0]] <= a64[:, :, 0][fin[:, :, 0]])
    assert np.all(amx[:, :, 1][fin[:, :, 1]] >= a64[:, :, 1][fin[:, :, 1]])
    assert_conj_equal(cmx.conjunctions, c64.conjunctions)


def test_errors():
    tcs, tms = H.planar_circ()
    pj = mz.polyjectory([tcs], [tms], [0])
    with pytest.raises(ValueError, match="The conjunction threshold must be finite and positive"):
        mz.conjunctions(pj, 0.0, 0.1)
    with pytest.raises(ValueError, match="is too large and results in an overflow error"):
        mz.conjunctions(pj, float(np.finfo(float).max), 0.1)
    with pytest.raises(ValueError, match="The conjunction detection interval must be finite and positive"):
        mz.conjunctions(pj, 1.0, float("nan"))
    with pytest.raises(ValueError) as ei:
        mz.conjunctions(pj, 0.1, 0.1, otypes=[])
    assert (
        "Invalid array of object types passed to the constructor of a conjunctions objects: the expected size is 1, "
        "but the actual size is 0 instead" in str(ei.value)
    )
    with pytest.raises(ValueError) as ei:
        mz.conjunctions(pj, 0.1, 0.1, otypes=[5])
    assert "The value of an object type must be one of [1, 2, 4], but a value of 5 was detected instead" in str(ei.value)
    c = mz.conjunctions(pj, 0.1, 0.1)
    with pytest.raises(IndexError, match="Cannot fetch the list of AABB collisions"):
        c.get_aabb_collisions(c.n_cd_steps)


def test_accessors_readonly_and_refcount():
    import sys

    tcs, tms = H.planar_circ()
    pj = mz.polyjectory([tcs], [tms], [0])
    c = mz.conjunctions(pj, 1.0, 0.1)
    for name in ("aabbs", "cd_end_times", "srt_aabbs", "mcodes", "srt_mcodes", "srt_idx", "otypes"):
        rc = sys.getrefcount(c)
        arr = getattr(c, name)
        assert sys.getrefcount(c) == rc + 1
        with pytest.raises(ValueError):
            arr[:] = arr
        with pytest.raises(AttributeError):
            setattr(c, name, arr)
        del arr
    assert isinstance(c.bvh_node, np.dtype) and c.bvh_node.itemsize == 48
    assert c.aabb_collision.itemsize == 8 and c.conj.itemsize == 120
    assert c.conj_thresh == 1.0 and c.conj_det_interval == 0.1
    c.hint_release()


def test_cross_interval_consistency():
    """The conjunctions do not depend on the conjunction-step length
    (test/test_conjunctions.py:786-972: 9000 objects over 15 minutes, 10 km,
    steps of 1 minute, 5 minutes and 1 day). Every interval gives the oracle's
    list bit for bit, and the same (i, j) pairs. Steps of 1 and 5 minutes agree
    on TCA to rtol 1e-11 (absolute floor 1e-13 days for TCAs seconds after
    t = 0), with positions within 1e-7 km. A single step over the whole span
    interpolates dist^2 over whole 12-minute trajectory steps, which moves TCAs
    by up to about 1e-9 days on these orbits; the bound there is the
    reference's cross-interval validation bound, 5e-8 days
    (validation/celestrak_on_orbit_20241101_consistency.py:76-114)."""
    from oracle import pyoracle as po

    pj = mz.make_synthetic_polyjectory(9000, 15.0 / 1440.0, seed=21, frac_high=0.2)
    flat = H.flat_of(pj)
    res = []
    for interval in (1.0 / 1440.0, 5.0 / 1440.0, 1.0):
        c = mz.conjunctions(pj, 10.0, interval, mode="fp64")
        assert_conj_equal(c.conjunctions, po.run(flat, 10.0, interval)["conjunctions"])
        assert np.all(np.diff(c.conjunctions["tca"]) >= 0)
        cj = np.sort(c.conjunctions, order=["tca", "i", "j"])
        d = np.sqrt(np.sum((cj["ri"] - cj["rj"]) ** 2, axis=1))
        np.testing.assert_allclose(cj["dca"], d, rtol=1e-14, atol=0)
        if interval < 1.0:
            assert all(len(c.get_aabb_collisions(k)) > 0 for k in range(c.n_cd_steps))
        res.append(cj)
    assert len(res[0]) > 0
    for other in res[1:]:
        assert np.array_equal(res[0]["i"], other["i"]) and np.array_equal(res[0]["j"], other["j"])
    np.testing.assert_allclose(res[0]["tca"], res[1]["tca"], rtol=1e-11, atol=1e-13)
    for x in ("ri", "rj"):
        assert np.max(np.linalg.norm(res[0][x] - res[1][x], axis=1)) < 1e-7
    assert np.max(np.abs(res[0]["tca"] - res[2]["tca"])) < 5e-8
