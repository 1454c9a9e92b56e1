"""Pins of the CPU oracle (oracle/mzoracle.c) -- runs without a GPU.

The reference cannot be built or imported here, so every oracle function is
checked against (a) the known answers of the reference's own tests
(mizuba/test/test_conjunctions.py, test_boundary_conjunctions.py), and (b)
what the mathematics fixes: closed forms, brute force on small inputs,
invariants that a plausible slip (wrong index, sign, dropped term, transposed
operand) would break.
"""

import math

import numpy as np
import pytest

from . import helpers as H

DAY_S = 86400.0


# ---------------------------------------------------------------------------
# numeric kernels (conjunctions_jit.cpp, poly_utils.cpp)
# ---------------------------------------------------------------------------


def _pval(c, x):
    return np.polyval(np.asarray(c)[::-1], x)


def test_poly_translate_matches_substitution(oracle):
    rng = np.random.default_rng(0)
    for order in (2, 3, 8, 20):
        c = rng.normal(size=order + 1)
        a = rng.uniform(-0.7, 0.7)
        t = oracle.poly_translate(c, a)
        xs = rng.uniform(-0.5, 0.5, 20)
        np.testing.assert_allclose(_pval(t, xs), _pval(c, xs + a), rtol=1e-12, atol=1e-12)
        back = oracle.poly_translate(t, -a)
        np.testing.assert_allclose(back, c, rtol=1e-9, atol=1e-11)
    # translation by zero is the identity, bit for bit
    c = rng.normal(size=9)
    assert np.array_equal(oracle.poly_translate(c, 0.0), c)


def test_cs_enclosure(oracle):
    rng = np.random.default_rng(1)
    # linear polynomial: the enclosure is exact
    lo, hi = oracle.cs_enclosure([1.0, 2.0, 0.0], 0.25, 0.75)
    assert (lo, hi) == (1.5, 2.5)
    # constant polynomial: degenerate interval
    lo, hi = oracle.cs_enclosure([3.0, 0.0, 0.0, 0.0], 0.0, 5.0)
    assert lo == hi == 3.0
    # x^2 on [-1, 1]: Bernstein coefficients 1, -1, 1
    lo, hi = oracle.cs_enclosure([0.0, 0.0, 1.0], -1.0, 1.0)
    assert (lo, hi) == (-1.0, 1.0)
    for order in (2, 5, 8, 20):
        for _ in range(20):
            c = rng.normal(size=order + 1)
            a, b = np.sort(rng.uniform(-1, 1, 2))
            lo, hi = oracle.cs_enclosure(c, a, b)
            v = _pval(c, np.linspace(a, b, 257))
            assert lo <= v.min() + 1e-12 and hi >= v.max() - 1e-12
            # end values are Bernstein coefficients b_0 and b_n: the bound is attained there
            assert lo <= min(v[0], v[-1]) + 1e-12 and hi >= max(v[0], v[-1]) - 1e-12
    # NaN propagates
    lo, hi = oracle.cs_enclosure([1.0, float("nan"), 0.0], 0.0, 1.0)
    assert math.isnan(lo) and math.isnan(hi)


def test_cheby_nodes(oracle):
    for order in (2, 8, 20):
        n = oracle.cheby_nodes(order, 2.5)
        op1 = order + 1
        ref = np.sort((np.cos((2 * np.arange(op1) + 1) / (2 * op1) * np.pi) + 1) / 2 * 2.5)
        np.testing.assert_allclose(n, ref, rtol=0, atol=1e-15)
        assert np.all(np.diff(n) > 0) and 0 < n[0] and n[-1] < 2.5


def test_vm_interp_reproduces_polynomials(oracle):
    rng = np.random.default_rng(2)
    for order in (2, 4, 8, 12):
        c = rng.normal(size=order + 1)
        x = oracle.cheby_nodes(order, 1.0)
        out = oracle.vm_interp(x, _pval(c, x))
        np.testing.assert_allclose(out, c, rtol=1e-7, atol=1e-9)
        # the interpolant passes through the data
        np.testing.assert_allclose(_pval(out, x), _pval(c, x), rtol=1e-12, atol=1e-12)
    # data from a lower degree gives zero leading coefficients
    x = oracle.cheby_nodes(8, 1.0)
    out = oracle.vm_interp(x, 3.0 - 2.0 * x)
    np.testing.assert_allclose(out, [3.0, -2.0] + [0.0] * 7, atol=1e-10)


def test_morton_encode_brute_force(oracle):
    rng = np.random.default_rng(3)
    for _ in range(200):
        f = [int(v) for v in rng.integers(0, 65536, 4)]
        code = 0
        for b in range(16):
            for d in range(4):
                code |= ((f[d] >> b) & 1) << (4 * b + d)
        assert oracle.morton_encode4(*f) == code
    assert oracle.morton_encode4(65535, 65535, 65535, 65535) == 2**64 - 1
    assert oracle.morton_encode4(1, 0, 0, 0) == 1 and oracle.morton_encode4(0, 0, 0, 1) == 8


def test_disc_single_coord(oracle):
    assert oracle.disc_single_coord(-1.0, -1.0, 1.0) == 0
    assert oracle.disc_single_coord(1.0, -1.0, 1.0) == 65535
    assert oracle.disc_single_coord(0.0, -1.0, 1.0) == 32768
    assert oracle.disc_single_coord(-5.0, -1.0, 1.0) == 0  # clamped below
    assert oracle.disc_single_coord(7.0, -1.0, 1.0) == 65535  # clamped above
    # float32 arithmetic, not float64: (x - min) / size rounds in single precision
    x, mn, mx = np.float32(0.3), np.float32(-0.7), np.float32(1.9)
    rx = (x - mn) / (mx - mn)
    assert oracle.disc_single_coord(x, mn, mx) == int(np.float32(rx) * np.float32(65536))


def test_bracketed_root(oracle):
    st, r, it = oracle.bracketed_root([-2.0, 0.0, 1.0], 1.0, 2.0)
    assert st == 0 and abs(r - math.sqrt(2.0)) <= 4 * np.spacing(math.sqrt(2.0)) and it < 100
    st, r, it = oracle.bracketed_root([-0.3, 1.0, 0.0, 0.0], 0.0, 1.0)
    assert st == 0 and abs(r - 0.3) < 1e-15
    # no sign change: domain error (errno EDOM = 33)
    st, _, _ = oracle.bracketed_root([1.0, 0.0, 1.0], 0.0, 1.0)
    assert st == 33


def test_root_finding_against_numpy(oracle):
    rng = np.random.default_rng(4)
    for _ in range(60):
        order = int(rng.integers(3, 9))
        nr = int(rng.integers(1, order))
        roots = np.sort(rng.uniform(0.02, 0.98, nr))
        if nr > 1 and np.min(np.diff(roots)) < 1e-3:
            continue
        # polynomial with these roots plus complex-pair factors (no other real roots)
        p = np.poly1d(roots, r=True)
        for _k in range(order - nr):
            p = p * np.poly1d([1.0, 0.0, 0.5])  # x^2 + 0.5 > 0
            if p.order >= order:
                break
        c = p.coeffs[::-1].copy()
        rf = 1.0
        got, fex = oracle.root_finding(c, rf)
        dp = p.deriv()
        exp = [r for r in roots if dp(r) > 0]
        assert not fex
        assert len(got) == len(exp), (got, exp)
        np.testing.assert_allclose(np.sort(got), exp, atol=1e-12)
        # rescaled interval: p(x / 2) on [0, 2)
        c2 = c / (2.0 ** np.arange(len(c)))
        got2, _ = oracle.root_finding(c2, 2.0)
        np.testing.assert_allclose(np.sort(got2), 2 * np.array(exp), atol=1e-11)
    # fast exclusion: strictly positive polynomial on [0, h]
    got, fex = oracle.root_finding([1.0, 0.5, 0.25], 1.0)
    assert fex and len(got) == 0


def test_cd_begin_end(oracle):
    assert oracle.cd_begin_end(1.0, 0, 0.3, 4) == (0, 0.0, 0.3)
    st, b, e = oracle.cd_begin_end(1.0, 3, 0.3, 4)
    assert st == 0 and e == 1.0 and b == 0.3 * 3
    st, _, _ = oracle.cd_begin_end(1.0, 3, 0.5, 4)  # begin beyond maxT
    assert st != 0


# ---------------------------------------------------------------------------
# reference known answers through the whole oracle pipeline
# ---------------------------------------------------------------------------


def test_zero_aabbs(oracle):
    out = oracle.run(H.zero_aabbs_pj(), float(np.finfo(float).eps), 0.1)
    a = out["aabbs"]
    assert np.all(a[:, :, 0, :3] == np.nextafter(np.single(1), np.single("-inf")))
    assert np.all(a[:, :, 1, :3] == np.nextafter(np.single(1), np.single("+inf")))
    assert np.all(a[:, :, 0, 3] == np.nextafter(np.single(np.sqrt(3.0)), np.single("-inf")))
    assert np.all(a[:, :, 1, 3] == np.nextafter(np.single(np.sqrt(3.0)), np.single("+inf")))


@pytest.mark.parametrize("shift", [0.0, 1.0])
def test_planar_circ(oracle, shift):
    tcs, tms = H.planar_circ()
    tms = tms + shift
    rng = np.random.default_rng(42)
    for interval in [0.01, 0.1, 0.5, 2.0, 5.0, 7.0]:
        out = oracle.run(([tcs], [tms]), 0.1, interval)
        a = out["aabbs"]
        assert out["cd_end_times"][-1] == tms[-1]
        assert out["n_cd_steps"] == math.ceil(tms[-1] / interval)
        assert np.all(a[:, 0] == a[:, 1])
        assert np.all(a == out["srt_aabbs"])
        fin = np.isfinite(a[:, 0, 0, 0])
        assert np.all(a[fin, 0, 0, 2] >= -0.05001) and np.all(a[fin, 0, 1, 2] <= 0.05001)
        assert np.all(a[fin, 0, 0, 3] >= 1 - 0.05001) and np.all(a[fin, 0, 1, 3] <= 1 + 0.05001)
        assert len(out["bp"]) == 0 and len(out["conjunctions"]) == 0
        # containment of random evaluations (test_conjunctions.py:71-173)
        ends = out["cd_end_times"]
        for k in range(len(ends)):
            b = 0.0 if k == 0 else ends[k - 1]
            box = a[k, 0]
            if not np.isfinite(box[0, 0]):
                assert tms[0] >= ends[k] or tms[-1] <= b
                continue
            for t in rng.uniform(b, ends[k], 5):
                s = np.searchsorted(tms, t, side="right")
                if s == 0 or s == len(tms):
                    continue
                h = t - tms[s - 1]
                for ch, d in zip([0, 1, 2, 6], range(4)):
                    v = _pval(tcs[s - 1][:, ch], h)
                    assert box[0, d] < v < box[1, d]


@pytest.mark.parametrize("case", range(6))
def test_cd_begin_end_cases(oracle, case):
    trajs, times, exp = H.cd_begin_end_case(case)
    out = oracle.run((trajs, times), 1e-6, 2.0 / 3.0)
    cj = out["conjunctions"]
    if exp == 0:
        assert len(cj) == 0
        return
    assert len(cj) >= 1
    assert np.all(cj["i"] == 0) and np.all(cj["j"] == 1)
    assert abs(cj["tca"][0] - 1.0) < 1e-15
    assert abs(cj["dca"][0]) <= 1e-15
    assert np.allclose(cj["ri"][0], [0, 0, 0], atol=1e-15, rtol=0.0)
    assert np.allclose(cj["rj"][0], [0, 0, 0], atol=1e-15, rtol=0.0)
    assert np.allclose(cj["vi"][0], [-1, 0, 0], atol=1e-15, rtol=0.0)
    assert np.allclose(cj["vj"][0], [1, 0, 0], atol=1e-15, rtol=0.0)


@pytest.mark.parametrize("shift", [0.0, 1.1])
def test_boundary_conjunctions(oracle, shift):
    out = oracle.run(H.boundary_incoming(shift), 0.25, 0.03)
    cj = out["conjunctions"]
    assert len(cj) == 1
    assert abs(cj["tca"][0] - (0.9 + shift)) < 1e-15 and abs(cj["dca"][0] - 0.2) < 1e-15
    assert np.allclose(cj["ri"][0], [0.1, 0, 0], rtol=1e-15, atol=0.0)
    assert np.allclose(cj["rj"][0], [-0.1, 0, 0], rtol=1e-15, atol=0.0)
    assert np.allclose(cj["vi"][0], [-1, 0, 0], rtol=1e-15, atol=0.0)
    assert np.allclose(cj["vj"][0], [1, 0, 0], rtol=1e-15, atol=0.0)
    out = oracle.run(H.boundary_outgoing(shift), 0.25, 0.03)
    cj = out["conjunctions"]
    assert len(cj) == 1
    assert abs(cj["tca"][0] - shift) < 1e-15 and abs(cj["dca"][0] - 0.2) < 1e-15
    assert np.allclose(cj["vi"][0], [1, 0, 0], rtol=1e-15, atol=0.0)
    assert np.allclose(cj["vj"][0], [-1, 0, 0], rtol=1e-15, atol=0.0)


def test_bvh_known_trees(oracle):
    t = np.zeros((6, 7))
    t[1, :] = 0.1
    out = oracle.run(([np.array([t])], [np.array([0.0, 1.0])]), 1e-16, 1.0)
    assert len(oracle.get_bvh_tree(out, 0)) == 1
    out = oracle.run(([np.array([t]), np.array([t])], [np.array([0.0, 1.0])] * 2), 1e-16, 1.0)
    tr = oracle.get_bvh_tree(out, 0)
    assert len(tr) == 1 and tr[0]["begin"] == 0 and tr[0]["end"] == 2 and tr[0]["left"] == -1 and tr[0]["right"] == -1
    objs = H.axis_objects()
    t8 = np.zeros((6, 7))
    t8[0, :] = 1
    t8[0, 0] = 1.0 - 2.1 / 2**16
    objs.append(t8)
    out = oracle.run(([np.array([o]) for o in objs], [np.array([0.0, 1.0])] * 9), 1e-16, 1.0)
    assert out["mcodes"][0, -2] == 2**64 - 1 and out["mcodes"][0, -1] == 2**64 - 2
    assert out["srt_idx"][0, -1] == 7 and out["srt_idx"][0, -2] == 8


def test_no_traj_data(oracle):
    objs = H.axis_objects()
    trajs = [np.array([objs[0]])] * 10 + [np.array([o, o]) for o in objs]
    times = [np.array([0.0, 1.0])] * 10 + [np.array([0.0, 1.0, 2.0])] * 8
    out = oracle.run((trajs, times), 1e-16, 1.0)
    assert np.all(np.isinf(out["aabbs"][1, :10]))
    assert np.all(np.isinf(out["srt_aabbs"][1, -11:-1]))
    assert np.all(out["mcodes"][1, :10] == 2**64 - 1)
    assert out["mcodes"][1, -1] == 2**64 - 1
    assert np.all(out["srt_mcodes"][1, -11:] == 2**64 - 1)
    # the finite object with the all-ones code sorts before the objects without data
    assert out["srt_idx"][1, -11] == 17


def test_errors(oracle):
    tcs, tms = H.planar_circ()
    with pytest.raises(ValueError, match="The conjunction threshold must be finite and positive"):
        oracle.run(([tcs], [tms]), 0.0, 0.1)
    with pytest.raises(ValueError, match="is too large and results in an overflow error"):
        oracle.run(([tcs], [tms]), float(np.finfo(float).max), 0.1)
    with pytest.raises(ValueError, match="The conjunction detection interval must be finite and positive"):
        oracle.run(([tcs], [tms]), 1.0, 0.0)
    with pytest.raises(ValueError, match=r"must be one of \[1, 2, 4\], but a value of -5 was detected instead"):
        oracle.run(([tcs], [tms]), 1.0, 0.1, otypes=[-5])


# ---------------------------------------------------------------------------
# closed-form collisions on circular orbits (replaces the heyoka cross-check of
# test_heyoka_conjunctions.py, which is not runnable offline)
# ---------------------------------------------------------------------------


def _circle_traj(plane, t_grid, order=20):
    """Exact Taylor coefficients of a unit-radius circular motion."""
    steps = []
    for t0 in t_grid[:-1]:
        c = np.zeros((order + 1, 7))
        for k in range(order + 1):
            f = 1.0 / math.factorial(k)
            cos_k = [math.cos, lambda x: -math.sin(x), lambda x: -math.cos(x), math.sin][k % 4](t0) * f
            sin_k = [math.sin, math.cos, lambda x: -math.sin(x), lambda x: -math.cos(x)][k % 4](t0) * f
            c[k, 0] = cos_k
            c[k, 1 if plane == "xy" else 2] = sin_k
            if k + 1 <= order:
                fk1 = 1.0 / math.factorial(k)
                c[k, 3] = [lambda x: -math.sin(x), lambda x: -math.cos(x), math.sin, math.cos][k % 4](t0) * fk1
                c[k, 4 if plane == "xy" else 5] = [math.cos, lambda x: -math.sin(x), lambda x: -math.cos(x), math.sin][
                    k % 4
                ](t0) * fk1
        c[0, 6] = 1.0
        steps.append(c)
    return np.array(steps)


def test_circular_orbit_collisions(oracle):
    grid = np.linspace(0.3, 7.0, 15)
    a = _circle_traj("xy", grid)
    b = _circle_traj("xz", grid)
    for interval in (0.37, 1.0, 3.3):
        out = oracle.run(([a, b], [grid, grid]), 1e-3, interval)
        cj = out["conjunctions"]
        assert len(cj) == 2
        np.testing.assert_allclose(cj["tca"], [math.pi, 2 * math.pi], rtol=0, atol=1e-12)
        assert np.all(cj["dca"] < 1e-12)
        np.testing.assert_allclose(cj["ri"][0], [-1, 0, 0], atol=1e-12)
        np.testing.assert_allclose(cj["vi"][0], [0, -1, 0], atol=1e-12)
        np.testing.assert_allclose(cj["vj"][0], [0, 0, -1], atol=1e-12)


# ---------------------------------------------------------------------------
# structural invariants on a synthetic catalog (verify_bvh_tree,
# verify_broad_phase of the reference, restated and strengthened)
# ---------------------------------------------------------------------------


@pytest.fixture(scope="module")
def synth_out(oracle):
    import mizuba_amd as mz

    pj = mz.make_synthetic_polyjectory(400, 0.08, seed=9, frac_high=0.2)
    rng = np.random.default_rng(5)
    ot = rng.choice([1, 2, 4], size=pj.n_objs, p=[0.6, 0.3, 0.1]).astype(np.int32)
    flat = H.flat_of(pj)
    return pj, ot, oracle.run(flat, 50.0, 3.0 / 1440.0, otypes=ot)


def test_morton_and_sort_invariants(synth_out, oracle):
    pj, ot, out = synth_out
    N = pj.n_objs
    for k in range(out["n_cd_steps"]):
        a = out["aabbs"][k]
        g = a[N]
        fin = np.isfinite(a[:N, 0, 0])
        assert np.array_equal(g[0], a[:N][fin, 0].min(axis=0)) and np.array_equal(g[1], a[:N][fin, 1].max(axis=0))
        ctr = a[:N, 0] / np.float32(2) + a[:N, 1] / np.float32(2)
        for o in range(0, N, 37):
            if not fin[o]:
                assert out["mcodes"][k, o] == 2**64 - 1
                continue
            q = [oracle.disc_single_coord(ctr[o, d], g[0, d], g[1, d]) for d in range(4)]
            assert out["mcodes"][k, o] == oracle.morton_encode4(*q)
        idx = out["srt_idx"][k]
        assert np.array_equal(np.sort(idx), np.arange(N))
        key = np.lexsort((np.arange(N), out["mcodes"][k], ~fin))  # (isinf, code, index)
        assert np.array_equal(idx, key)
        assert np.array_equal(out["srt_mcodes"][k], out["mcodes"][k][idx])
        assert np.array_equal(out["srt_aabbs"][k, :N], a[:N][idx]) and np.array_equal(out["srt_aabbs"][k, N], a[N])


def _verify_tree(tree, codes, boxes, n_eff):
    assert tree[0]["begin"] == 0 and tree[0]["end"] == n_eff
    seen = np.zeros(n_eff, dtype=int)
    level_of = {0: 0}
    for idx, nd in enumerate(tree):
        b, e = int(nd["begin"]), int(nd["end"])
        assert b < e
        members = boxes[b:e]
        assert np.array_equal(nd["lb"], members[:, 0].min(axis=0))
        assert np.array_equal(nd["ub"], members[:, 1].max(axis=0))
        if nd["left"] == -1:
            assert nd["right"] == -1
            assert e - b == 1 or np.all(codes[b:e] == codes[b])
            seen[b:e] += 1
        else:
            L, R = tree[nd["left"]], tree[nd["right"]]
            assert nd["right"] == nd["left"] + 1  # adjacent siblings
            assert nd["left"] > idx
            assert L["begin"] == b and L["end"] == R["begin"] and R["end"] == e
            x = int(codes[b]) ^ int(codes[e - 1])
            bit = x.bit_length() - 1
            assert all((int(c) >> bit) & 1 == 0 for c in codes[b : L["end"]])
            assert all((int(c) >> bit) & 1 == 1 for c in codes[R["begin"] : e])
            level_of[int(nd["left"])] = level_of[idx] + 1
            level_of[int(nd["right"])] = level_of[idx] + 1
    assert np.all(seen == 1)  # every object in exactly one leaf
    lv = [level_of[k] for k in range(len(tree))]
    assert lv == sorted(lv)  # breadth-first numbering


def test_bvh_invariants(synth_out, oracle):
    pj, ot, out = synth_out
    N = pj.n_objs
    for k in range(out["n_cd_steps"]):
        n_eff = int(np.sum(np.isfinite(out["srt_aabbs"][k, :N, 0, 0])))
        tree = oracle.get_bvh_tree(out, k)
        if n_eff == 0:
            assert len(tree) == 0
            continue
        assert len(tree) <= 2 * n_eff - 1
        _verify_tree(tree, out["srt_mcodes"][k], out["srt_aabbs"][k], n_eff)


def test_broad_phase_closed_form(synth_out, oracle):
    pj, ot, out = synth_out
    N = pj.n_objs
    for k in range(out["n_cd_steps"]):
        tree = oracle.get_bvh_tree(out, k)
        idx = out["srt_idx"][k]
        a = out["aabbs"][k]
        leafbox = {}
        for nd in tree:
            if nd["left"] == -1:
                for p in range(nd["begin"], nd["end"]):
                    leafbox[int(idx[p])] = (nd["lb"], nd["ub"])
        exp = set()
        direct = set()
        for i in leafbox:
            for j in leafbox:
                if i >= j or ot[i] + ot[j] > 3:
                    continue
                lb, ub = leafbox[j]
                if np.all(a[i, 1] >= lb) and np.all(a[i, 0] <= ub):
                    exp.add((i, j))
                if np.all(a[i, 1] >= a[j, 0]) and np.all(a[i, 0] <= a[j, 1]):
                    direct.add((i, j))
        got = oracle.get_aabb_collisions(out, k)
        got_l = [(int(p["i"]), int(p["j"])) for p in got]
        assert got_l == sorted(set(got_l))  # sorted, unique
        assert set(got_l) == exp
        assert direct <= set(got_l)


def test_narrow_phase_invariants(synth_out, oracle):
    pj, ot, out = synth_out
    cj = out["conjunctions"]
    assert len(cj) > 0
    d = np.sqrt(np.sum((cj["ri"] - cj["rj"]) ** 2, axis=1))
    np.testing.assert_allclose(cj["dca"], d, rtol=1e-14)
    assert np.all(cj["dca"] < 50.0)
    assert np.all(ot[cj["i"]] + ot[cj["j"]] <= 3) and np.all(cj["i"] < cj["j"])
    ends = out["cd_end_times"]
    step = np.searchsorted(ends, cj["tca"], side="right")
    assert np.all(np.diff(step) >= 0)
    for s in np.unique(step):
        t = cj["tca"][step == s]
        assert np.all(np.diff(t) >= 0)
    # every conjunction pair was a broad-phase candidate of its step
    for c, s in zip(cj, step):
        bp = oracle.get_aabb_collisions(out, min(int(s), out["n_cd_steps"] - 1))
        assert np.any((bp["i"] == c["i"]) & (bp["j"] == c["j"]))


def test_cross_interval_consistency(oracle):
    """The conjunction list must not depend on the conjunction step length
    (test_conjunctions.py:905-923)."""
    import mizuba_amd as mz

    pj = mz.make_synthetic_polyjectory(600, 0.1, seed=21, frac_high=0.2)
    flat = H.flat_of(pj)
    outs = [oracle.run(flat, 20.0, iv)["conjunctions"] for iv in (1.0 / 1440.0, 5.0 / 1440.0, 1.0)]
    assert len(outs[0]) > 0

    def key(c):
        o = np.lexsort((c["j"], c["i"]))
        return c[o]

    base = key(outs[0])
    for other in outs[1:]:
        o = key(other)
        # conjunctions at step boundaries can be split/duplicated; compare the set of distinct minima
        assert set(zip(base["i"], base["j"])) == set(zip(o["i"], o["j"]))
        for i, j in set(zip(base["i"], base["j"])):
            t1 = np.sort(base["tca"][(base["i"] == i) & (base["j"] == j)])
            t2 = np.sort(o["tca"][(o["i"] == i) & (o["j"] == j)])
            for t in t1:
                assert np.min(np.abs(t2 - t)) < 1e-8
