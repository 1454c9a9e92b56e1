"""The batching configuration on the GPU: 64 independent screening problems
(object counts log-uniform in [500, 20 000], thresholds U[1, 10] km, horizons
U[0.5, 3] days, 3-minute steps, seeded SGP4 catalogs made with
make_sgp4_polyjectory, mixed precision) through one batch handle: tuned launch
shapes, one eager run, then a replay of the captured graph. Every problem's
conjunctions equal those of screening it alone, on the replay as on the first
run, and the oracle's on windows of steps at both ends of its horizon; an FP64
batch of some of the problems equals the oracle there box for box and pair
for pair. Also: problems of different polynomial orders and object types in
one batch, and the errors of a batch."""

import numpy as np
import pytest

import mizuba_amd as mz

from . import helpers as H

pytestmark = pytest.mark.gpu

JD0 = 2460496.5
INTERVAL = 3.0 / 1440.0


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def config3(seed=2030, n_prob=64):
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(500), np.log(20000), n_prob)).astype(int)
    threshs = [float(x) for x in rng.uniform(1.0, 10.0, n_prob)]
    days = [float(x) for x in rng.uniform(0.5, 3.0, n_prob)]
    pjs = []
    for k in range(n_prob):
        g = mz.make_synthetic_gpes(int(ns[k]), seed=seed + 1 + k, frac_high=0.2, epoch_jd=JD0)
        pjs.append(mz.make_sgp4_polyjectory(g, JD0, JD0 + days[k])[0])
    return pjs, threshs


@pytest.fixture(scope="module")
def c3():
    return config3()


def _window_conj(cj, cd_end, k0, k1, C):
    t0 = 0.0 if k0 == 0 else cd_end[k0 - 1]
    t1 = cd_end[k1 - 1]
    upper = (cj["tca"] <= t1) if k1 == C else (cj["tca"] < t1)
    return cj[(cj["tca"] >= t0) & upper]


def test_config3_batch_graph_tuned(c3):
    from oracle import pyoracle as po

    pjs, threshs = c3
    n = len(pjs)
    scrs = [mz.Screener(0) for _ in pjs]
    for s, pj in zip(scrs, pjs):
        s.upload(pj)
    b = mz.Batch(scrs, threshs, [INTERVAL] * n, mode="mixed", graph=True, tune=True)
    b.run()
    assert b.shape["bp_block"] in (128, 256, 512) and b.shape["bucket_bits"] in (10, 11, 12)
    first = [s.conjunctions() for s in scrs]
    assert sum(len(x) for x in first) > 1000
    b.run()
    assert b.replays == 1
    for k, s in enumerate(scrs):
        assert np.array_equal(_bits(s.conjunctions()), _bits(first[k])), k
    # every problem on its own
    for k, (s, th) in enumerate(zip(scrs, threshs)):
        s.screen(th, INTERVAL, mode="mixed")
        assert np.array_equal(_bits(s.conjunctions()), _bits(first[k])), k
    # the oracle on the first and last steps of every problem
    for k, (pj, th) in enumerate(zip(pjs, threshs)):
        C = int(np.ceil(pj.maxT / INTERVAL))
        cd_end = np.array([po.cd_begin_end(pj.maxT, c, INTERVAL, C)[2] for c in range(C)])
        for k0, k1 in ((0, 4), (C - 4, C)):
            ref = po.run(H.flat_of(pj), th, INTERVAL, step_begin=k0, step_end=k1)
            g = _window_conj(first[k], cd_end, k0, k1, C)
            assert np.array_equal(_bits(g), _bits(ref["conjunctions"])), (k, k0, k1)
    # after the single screenings the batch describes its problems again
    b.run()
    for k, s in enumerate(scrs):
        assert np.array_equal(_bits(s.conjunctions()), _bits(first[k])), k


def test_config3_fp64_batch_against_oracle(c3):
    from oracle import pyoracle as po

    pjs, threshs = c3
    pick = [0, 7, 21, 33, 50, 63]
    scrs = [mz.Screener(0) for _ in pick]
    for s, k in zip(scrs, pick):
        s.upload(pjs[k])
    b = mz.Batch(scrs, [threshs[k] for k in pick], [INTERVAL] * len(pick), mode="fp64", graph=False)
    b.run()
    for s, k in zip(scrs, pick):
        pj, th = pjs[k], threshs[k]
        c = mz.conjunctions(pj, th, INTERVAL)  # the object's own screening
        assert np.array_equal(_bits(s.conjunctions()), _bits(c.conjunctions)), k
        C = c.n_cd_steps
        for k0, k1 in ((0, 3), (C // 2, C // 2 + 3), (C - 3, C)):
            ref = po.run(H.flat_of(pj), th, INTERVAL, step_begin=k0, step_end=k1)
            for name in ("aabbs", "srt_aabbs", "mcodes", "srt_mcodes", "srt_idx"):
                assert np.array_equal(_bits(getattr(c, name)[k0:k1]), _bits(ref[name])), (k, name)
            for q in range(k1 - k0):
                assert np.array_equal(_bits(c.get_bvh_tree(k0 + q)), _bits(po.get_bvh_tree(ref, q)))
                assert np.array_equal(_bits(c.get_aabb_collisions(k0 + q)), _bits(po.get_aabb_collisions(ref, q)))
            g = _window_conj(c.conjunctions, c.cd_end_times, k0, k1, C)
            assert np.array_equal(_bits(g), _bits(ref["conjunctions"]))


def test_batch_mixed_orders_otypes_and_graph_sizes():
    """Problems of polynomial orders 5, 8 and 12 with object types in one
    batch equal the problems screened alone; a replay after a capacity-
    growing change of one problem's polyjectory is redone and stays exact."""
    specs = [(1500, 0.3, 8), (400, 0.2, 5), (2500, 0.25, 8), (900, 0.4, 12), (60, 0.1, 5)]
    rng = np.random.default_rng(4)
    pjs = [mz.make_synthetic_polyjectory(n, d, seed=40 + k, frac_high=0.2, order=o)
           for k, (n, d, o) in enumerate(specs)]
    ths = [float(x) for x in rng.uniform(4, 12, len(pjs))]
    ivs = [float(rng.choice([1.0, 2.0, 3.0])) / 1440.0 for _ in pjs]
    ots = [rng.choice([1, 2, 4], size=pj.n_objs, p=[0.6, 0.3, 0.1]).astype(np.int32) if k % 2 else None
           for k, pj in enumerate(pjs)]
    scrs = [mz.Screener(0) for _ in pjs]
    for s, pj in zip(scrs, pjs):
        s.upload(pj)
    b = mz.Batch(scrs, ths, ivs, otypes=ots, mode="fp64", graph=True)
    b.run()
    got = [s.conjunctions() for s in scrs]
    for k, (pj, th, iv, ot) in enumerate(zip(pjs, ths, ivs, ots)):
        c = mz.conjunctions(pj, th, iv, otypes=ot)
        assert np.array_equal(_bits(got[k]), _bits(c.conjunctions)), k
    b.run()
    assert b.replays == 1
    for k, s in enumerate(scrs):
        assert np.array_equal(_bits(s.conjunctions()), _bits(got[k])), k
    # a denser catalog in problem 0's context: more pairs than the graph's
    # capacities; the run is redone eagerly and captured again
    dense = mz.make_synthetic_polyjectory(6000, 0.3, seed=99, frac_high=0.0)
    scrs[0].upload(dense)
    b.run()
    c = mz.conjunctions(dense, ths[0], ivs[0], otypes=None)
    assert np.array_equal(_bits(scrs[0].conjunctions()), _bits(c.conjunctions))
    for k in range(1, len(scrs)):
        assert np.array_equal(_bits(scrs[k].conjunctions()), _bits(got[k])), k


def test_batch_errors():
    tcs, tms = H.planar_circ()
    pj = mz.polyjectory([tcs], [tms], [0])
    s1, s2 = mz.Screener(0), mz.Screener(0)
    s1.upload(pj)
    s2.upload(pj)
    with pytest.raises(ValueError, match="The conjunction threshold must be finite and positive"):
        mz.Batch([s1, s2], [1.0, -1.0], [0.1, 0.1])
    with pytest.raises(ValueError, match="must be distinct"):
        mz.Batch([s1, s1], [1.0, 1.0], [0.1, 0.1])
    with pytest.raises(ValueError, match="lists of the same length"):
        mz.Batch([s1, s2], [1.0], [0.1, 0.1])
    b = mz.Batch([s1, s2], [1.0, 2.0], [0.1, 0.2])
    b.run()
    assert s1.n_cd_steps() == int(np.ceil(pj.maxT / 0.1)) and s2.n_cd_steps() == int(np.ceil(pj.maxT / 0.2))
