"""GPU SGP4: propagation against the reference's vendored Vallado SGP4 and
polyjectory generation against the reference's test expectations
(test/test_make_sgp4_polyjectory.py).

The propagation is compared with the fixture tests/golden/sgp4_vallado.npz
(the reference's own SGP4 code run on the GPE files of its test suite). The
generated polyjectories are compared with the GPU propagation at random times
inside every step, at the reference's tolerances: 1e-8 km and 1e-11 km/s
(test_make_sgp4_polyjectory.py:31-67).
"""

import bisect

import numpy as np
import pytest

import mizuba_amd as mz

from . import helpers as H

pytestmark = pytest.mark.gpu

# the published SGP4 code runs in FP64 with the C library's sin, cos, pow and
# atan2; the GPU's differ in the last bits, which moves states by ~1e-12 km.
# Kepler's equation stops at a 1e-12 rad correction, so where the last bits
# change the iteration count the state moves by up to ~1e-12 rad * r.
POS_TOL_TYPICAL = 1e-9
VEL_TOL_TYPICAL = 1e-12

# the generated polynomials against SGP4 (test_make_sgp4_polyjectory.py:31-67)
POS_TOL = 1e-8
VEL_TOL = 1e-11

# TAI - UTC over the dates used here
TAI_2024 = 37.0 / 86400.0


def _compare_steps(pj, obj, gpe, rng, n_per_step=5, pos_tol=POS_TOL, vel_tol=VEL_TOL, ep_off=0.0):
    """Evaluate every step of object obj at random times and compare with the
    GPU SGP4 of gpe (epoch ep_off days from the polyjectory's epoch)."""
    cfs, times, status = pj[obj]
    ns = len(times) - 1
    assert ns > 0
    x = rng.uniform(0.0, 1.0, (ns, n_per_step)) * np.diff(times)[:, None]
    tsince = ((times[:-1, None] + x) - ep_off) * 1440.0
    st, err = mz.sgp4_propagate(gpe, tsince.reshape(-1))
    assert np.all(err == 0)
    st = st[0].reshape(ns, n_per_step, 6)
    worst_p = worst_v = 0.0
    for k in range(ns):
        ev = H.poly_eval(cfs[k], x[k])
        worst_p = max(worst_p, np.abs(ev[:, :3] - st[k, :, :3]).max())
        worst_p = max(worst_p, np.abs(ev[:, 6] - np.linalg.norm(st[k, :, :3], axis=1)).max())
        worst_v = max(worst_v, np.abs(ev[:, 3:6] - st[k, :, 3:6]).max())
    print(f"object {obj}: {ns} steps, worst position {worst_p:.2e} km, velocity {worst_v:.2e} km/s")
    assert worst_p < pos_tol, worst_p
    assert worst_v < vel_tol, worst_v
    return worst_p, worst_v


# ---- propagation ------------------------------------------------------------


def test_propagate_matches_vallado():
    g = H.sgp4_golden()
    gpes = H.gpe_array(g["gpes"])
    st, err = mz.sgp4_propagate(gpes, g["tsince"])
    deep = g["deep"] == 1
    ref_err = g["errors"]
    assert np.array_equal(err, ref_err)
    for kind, sel in (("near-earth", ~deep), ("deep-space", deep)):
        ok = sel[:, None] & (ref_err == 0)
        dp = np.abs(st[..., :3] - g["states"][..., :3])[ok]
        dv = np.abs(st[..., 3:] - g["states"][..., 3:])[ok]
        print(f"{kind}: {ok.sum()} states, position q99 {np.quantile(dp, 0.99):.2e} max {dp.max():.2e} km, "
              f"velocity q99 {np.quantile(dv, 0.99):.2e} max {dv.max():.2e} km/s")
        # nearly every state to the last bits; a few may sit across a change
        # of the Kepler iteration count (1e-12 rad times the radius)
        rmax = np.linalg.norm(g["states"][..., :3], axis=2)[ok].max()
        assert np.quantile(dp, 0.99) < POS_TOL_TYPICAL * rmax / 7000.0
        assert np.quantile(dv, 0.99) < VEL_TOL_TYPICAL * rmax / 7000.0
        assert dp.max() < 2e-12 * rmax + 1e-8 and dv.max() < 2e-11


def test_propagate_shapes_and_errors():
    gpes = H.golden_group("multi_gpes")
    st, err = mz.sgp4_propagate(gpes, np.array([0.0, 10.0]))
    assert st.shape == (len(gpes), 2, 6) and err.shape == (len(gpes), 2)
    st2, _ = mz.sgp4_propagate(gpes, np.tile([0.0, 10.0], (len(gpes), 1)))
    assert np.array_equal(st, st2)
    with pytest.raises(TypeError):
        mz.sgp4_propagate(np.zeros(3), [0.0])


# ---- generation (test_make_sgp4_polyjectory.py) -----------------------------


def test_single_gpe():
    rng = np.random.default_rng(42)
    gpes = H.golden_group("single_gpe")
    jd_begin = 2460669.0
    pj, ids = mz.make_sgp4_polyjectory(gpes, jd_begin, jd_begin + 1)
    assert pj[0][1][0] == 0.0
    assert list(ids) == [int(gpes["norad_id"][0])]
    assert pj.status[0] == 0 and pj.poly_order == 8
    # TAI epoch: jd_begin + 37 s, normalised (test_make_sgp4_polyjectory.py:108-113)
    hi = jd_begin + TAI_2024
    lo = (jd_begin - hi) + TAI_2024
    assert pj.epoch[0] == hi and abs(pj.epoch[1] - lo) < 1e-16
    e_off = (gpes["epoch_jd1"][0] - jd_begin) + gpes["epoch_jd2"][0]
    _compare_steps(pj, 0, gpes, rng, ep_off=e_off)
    # the steps cover [0, 1 day] without gaps
    times = pj[0][1]
    assert times[-1] == pytest.approx(1.0, abs=1e-12) and np.all(np.diff(times) > 0)


def test_multi_gpes():
    rng = np.random.default_rng(123)
    gpes = H.golden_group("multi_gpes")
    jd_begin = 2460669.0
    pj, ids = mz.make_sgp4_polyjectory(gpes, jd_begin, jd_begin + 1)
    assert pj.n_objs == len(ids) == len(gpes)
    assert np.array_equal(ids, gpes["norad_id"])
    for i in range(pj.n_objs):
        assert pj[i][1][0] == 0.0
        e_off = (gpes["epoch_jd1"][i] - jd_begin) + gpes["epoch_jd2"][i]
        _compare_steps(pj, i, gpes[i : i + 1], rng, ep_off=e_off)


def test_iss_gpes():
    # one satellite, several GPEs; the GPE in use at t is the latest with an
    # epoch at or before t (the first one before its epoch)
    rng = np.random.default_rng(123)
    gpes = H.golden_group("iss_gpes")
    epochs = list(zip(gpes["epoch_jd1"], gpes["epoch_jd2"]))
    for jd_begin, jd_end in [(2460667.0, 2460684.0), (2460672.0, 2460679.0)]:
        pj, ids = mz.make_sgp4_polyjectory(gpes, jd_begin, jd_end)
        assert list(ids) == [25544]
        assert pj[0][1][0] == 0.0
        tms = rng.uniform(0, jd_end - jd_begin, 100)
        st = pj.state_meval(tms)[0]
        offs = [(a - jd_begin) + b for a, b in epochs]
        for k, tm in enumerate(tms):
            idx = bisect.bisect_right(offs, tm)
            idx -= int(idx != 0)
            ref, err = mz.sgp4_propagate(gpes[idx : idx + 1], [(tm - offs[idx]) * 1440.0])
            assert err[0, 0] == 0
            assert np.allclose(ref[0, 0, :3], st[k, :3], rtol=0, atol=POS_TOL)
            assert np.allclose(ref[0, 0, 3:], st[k, 3:6], rtol=0, atol=VEL_TOL)
        pj.hint_release()


def test_leap_second():
    # a span over the leap second at the end of 2005 (test_make_sgp4_polyjectory.py:336-441):
    # time runs in TAI from jd_begin, so 10 UTC days are 10 days and 1 s
    row = H.gpe_from_tle(
        "1 00045U 60007A   05363.79166667  .00000504  00000-0  14841-3 0  9992",
        "2 00045  66.6943  81.3521 0257384 317.3173  40.8180 14.34783636277898",
        norad_id=1234,
    )
    gpes = H.gpe_array([row])
    jd_begin = row[1] + row[2]
    pj = mz.make_sgp4_polyjectory(gpes, jd_begin, jd_begin + 10)[0]
    cfs, times, status = pj[0]
    assert status == 0 and times[0] == 0.0
    assert times[-1] == pytest.approx(10.0 + 1.0 / 86400.0, abs=1e-9)
    # TAI - UTC is 32 s before the leap second
    assert pj.epoch[0] + pj.epoch[1] == pytest.approx(jd_begin + 32.0 / 86400.0, abs=1e-11)
    # the GPE's epoch is jd_begin up to the rounding of the sum: minutes since
    # the epoch are TAI minutes since jd_begin, less that rounding
    _compare_steps(pj, 0, gpes, np.random.default_rng(420), ep_off=(row[1] - jd_begin) + row[2])


def _s(norad, e1, e2):
    g = np.zeros(2, dtype=mz.gpe_dtype)
    g["norad_id"] = norad
    g["epoch_jd1"] = e1
    g["epoch_jd2"] = e2
    return g


def test_malformed_data():
    arr = np.zeros((1, 1), dtype=mz.gpe_dtype)
    with pytest.raises(ValueError, match=r"must have 1 dimension, but the number of dimensions is 2 instead"):
        mz.make_sgp4_polyjectory(arr, 0.0, 1.0)
    arr = np.zeros((), dtype=mz.gpe_dtype)
    with pytest.raises(ValueError, match=r"must have 1 dimension, but the number of dimensions is 0 instead"):
        mz.make_sgp4_polyjectory(arr, 0.0, 1.0)
    arr = np.zeros((10,), dtype=mz.gpe_dtype)
    with pytest.raises(ValueError, match=r"must be C contiguous and properly aligned"):
        mz.make_sgp4_polyjectory(arr[::2], 0.0, 1.0)
    with pytest.raises(ValueError, match=r"requires a non-empty array of GPEs in input"):
        mz.make_sgp4_polyjectory(np.zeros((0,), dtype=mz.gpe_dtype), 0.0, 1.0)
    arr = _s([1, 1], [123, 123], [1, 1])
    with pytest.raises(ValueError) as ei:
        mz.make_sgp4_polyjectory(arr, 0.0, 1.0)
    assert "Two GPEs with identical NORAD ID 1 and epoch (123, 1) were identified by make_sgp4_polyjectory() - this is not allowed" in str(ei.value)
    for arr in (_s([1, 0], [123, 123], [1, 1]), _s([1, 1], [123, 122], [1, 1])):
        with pytest.raises(ValueError) as ei:
            mz.make_sgp4_polyjectory(arr, 0.0, 1.0)
        assert "is not sorted correctly: it must be ordered first by NORAD ID and then by epoch" in str(ei.value)
    arr = np.zeros((2,), dtype=mz.gpe_dtype)
    for b, e in [(float("inf"), 1.0), (1.0, float("nan")), (1.0, 1.0), (2.0, 1.0)]:
        with pytest.raises(ValueError, match="Invalid Julian date interval "):
            mz.make_sgp4_polyjectory(arr, b, e)
    for kw in ({"reentry_radius": float("nan")}, {"exit_radius": float("nan")}):
        with pytest.raises(ValueError) as ei:
            mz.make_sgp4_polyjectory(arr, 0.0, 1.0, **kw)
        assert "The reentry/exit radiuses in make_sgp4_polyjectory() cannot be NaN" in str(ei.value)
    with pytest.raises(ValueError) as ei:
        mz.make_sgp4_polyjectory(arr, 0.0, 1.0, reentry_radius=2, exit_radius=2)
    assert "The reentry radius (2) must be less than the exit radius (2) in make_sgp4_polyjectory()" in str(ei.value)
    with pytest.raises(ValueError) as ei:
        mz.make_sgp4_polyjectory(arr, 0.0, 1.0, reentry_radius=3, exit_radius=2)
    assert "The reentry radius (3) must be less than the exit radius (2) in make_sgp4_polyjectory()" in str(ei.value)
    arr = H.golden_group("single_gpe").copy()
    arr["n0"][0] = float("nan")
    with pytest.raises(ValueError, match="Invalid GPE detected by make_sgp4_polyjectory()"):
        mz.make_sgp4_polyjectory(arr, 2460669.0, 2460670.0)


def test_single_gpe_ds():
    # a deep-space GPE (test_make_sgp4_polyjectory.py:116-153: 1e-7 km, 1e-10 km/s)
    gpes = H.golden_group("single_gpe_ds")
    jd_begin = 2460669.0
    pj = mz.make_sgp4_polyjectory(gpes, jd_begin, jd_begin + 1)[0]
    assert pj[0][1][0] == 0.0 and pj.status[0] == 0
    e_off = (gpes["epoch_jd1"][0] - jd_begin) + gpes["epoch_jd2"][0]
    _compare_steps(pj, 0, gpes, np.random.default_rng(42), pos_tol=1e-7, vel_tol=1e-10, ep_off=e_off)


def test_single_gpe_syncom():
    # a deep-space GPE whose SGP4 derivatives spike now and then
    # (test_make_sgp4_polyjectory.py:155-197: 10 days, 1e-6 km, 1e-10 km/s)
    gpes = H.golden_group("syncom_gpe")
    jd_begin = 2460666.5
    pj = mz.make_sgp4_polyjectory(gpes, jd_begin, jd_begin + 10)[0]
    assert pj[0][1][0] == 0.0 and pj.status[0] == 0
    e_off = (gpes["epoch_jd1"][0] - jd_begin) + gpes["epoch_jd2"][0]
    _compare_steps(pj, 0, gpes, np.random.default_rng(24), pos_tol=1e-6, vel_tol=1e-10, ep_off=e_off)


def test_near_and_deep_together():
    # near-earth and deep-space satellites in one polyjectory come out as
    # they do alone
    near = H.golden_group("single_gpe")
    ds = H.golden_group("single_gpe_ds")
    both = np.concatenate([near, ds]) if near["norad_id"][0] < ds["norad_id"][0] else np.concatenate([ds, near])
    pj, ids = mz.make_sgp4_polyjectory(both, 2460669.0, 2460670.0)
    assert np.array_equal(ids, both["norad_id"]) and np.all(pj.status == 0)
    for k, one in enumerate(both):
        alone = mz.make_sgp4_polyjectory(both[k : k + 1], 2460669.0, 2460670.0)[0]
        assert np.array_equal(alone[0][0], pj[k][0]) and np.array_equal(alone[0][1], pj[k][1])


_S_8000 = ("1 00011U 59001A   24187.51496924  .00001069  00000-0  55482-3 0  9992",
           "2 00011  32.8711 255.0638 1455653 332.1888  20.7734 11.88503118450690")
_S_DEC = ("1 04206U 69082BV  24187.08533867  .00584698  00000-0  52886-2 0  9990",
          "2 04206  69.8949  69.3024 0029370 203.3165 156.6698 15.65658911882875")


def test_exit_decay():
    # test_make_sgp4_polyjectory.py:733-753
    gpes = H.gpe_array([H.gpe_from_tle(*_S_8000), H.gpe_from_tle(*_S_DEC)])
    pj = mz.make_sgp4_polyjectory(gpes, 2460496.5 + 1.0 / 32, 2460496.5 + 7, exit_radius=8000.0)[0]
    assert list(pj.status) == [12, 0]
    # the exiting satellite stops where a step begins at or beyond 8000 km
    cfs, times, _ = pj[0]
    if len(times) > 1:
        end = H.poly_eval(cfs[-1], [times[-1] - times[-2]])[0]
        assert np.linalg.norm(end[:3]) >= 8000.0 - 1e-6
    gpes = H.gpe_array([H.gpe_from_tle(*_S_DEC)])
    pj = mz.make_sgp4_polyjectory(gpes, 2460496.5 + 30.0, 2460496.5 + 30.0 + 7)[0]
    assert list(pj.status) == [6]


def test_nz_status_and_catalog():
    # strack GPEs over 4 hours: the trajectories without data have a nonzero
    # status (test_make_sgp4_polyjectory.py:667-693), and the states agree with
    # SGP4 (test_strack, :755-817: < 1e-6 km at 10 times)
    g = H.sgp4_golden()
    gi = list(g["groups"]).index("strack_20240705")
    gpes = H.gpe_array(g["gpes"][g["group"] == gi])
    begin = 2460496.5
    pj = mz.make_sgp4_polyjectory(gpes, begin, begin + 1 / 6.0)[0]
    for cfs, times, status in pj:
        if times.shape[0] == 0:
            assert status != 0
    tspan = np.linspace(0.0, np.nextafter(1 / 6.0, -1.0), 10)
    ev = pj.state_meval(tspan)
    e_off = (gpes["epoch_jd1"] - begin) + gpes["epoch_jd2"]
    ref, err = mz.sgp4_propagate(gpes, (tspan[None, :] - e_off[:, None]) * 1440.0)
    d = np.linalg.norm(ev[:, :, :3] - ref[:, :, :3], axis=2)
    ok = ~np.isnan(d)
    # the last time may lie past the trajectories' end (jd_end is rounded)
    assert ok[:, :-1].sum() > 0.8 * ok[:, :-1].size
    assert d[ok].max() < 1e-6
    # and, where nothing failed, at the reference's step tolerance, scaled by
    # the radius for deep-space orbits (the published Kepler stopping rule
    # moves states by up to 1e-12 rad times the radius)
    good = (pj.status == 0)[:, None] & ok
    r = np.linalg.norm(ref[:, :, :3], axis=2)
    assert np.all(d[good] < POS_TOL * np.maximum(1.0, r[good] / 7000.0))


@pytest.mark.parametrize("name,begin", [("strack_20240705", 2460496.5), ("strack_20240917", 2460569.5)])
def test_strack_catalogs(name, begin):
    # the reference's full space-track catalogs over 4 hours
    # (test_make_sgp4_polyjectory.py:755-817): some decayed satellites hit the
    # bisection limit (status 14), and the states of the others agree with
    # SGP4 within 1e-6 km at 10 times; satellites without data have a status
    gpes = H.gpe_array(np.load(H.GOLDEN / f"{name}.npz")["gpes"])
    prop_time = 1 / 6.0
    pj, ids = mz.make_sgp4_polyjectory(gpes, begin, begin + prop_time)
    assert pj.n_objs == len(ids) == len(np.unique(gpes["norad_id"]))
    st = np.asarray(pj.status)
    assert np.any(st == 14)
    for k in np.nonzero(st != 0)[0][:50]:
        assert pj[int(k)][2] == st[k]
    tspan = np.linspace(0.0, np.nextafter(prop_time, -1.0), 10)
    ev = pj.state_meval(tspan)
    e_off = (gpes["epoch_jd1"] - begin) + gpes["epoch_jd2"]
    ref, err = mz.sgp4_propagate(gpes, (tspan[None, :] - e_off[:, None]) * 1440.0)
    d = np.linalg.norm(ev[:, :, :3] - ref[:, :, :3], axis=2)
    ok = ~np.isnan(d)
    print(name, dict(zip(*[a.tolist() for a in np.unique(st, return_counts=True)])), "max", d[ok].max())
    assert d[ok].max() < 1e-6
    assert np.all(st[np.all(~ok, axis=1)] != 0)


def test_disc_gpe_isolated():
    # a GPE with known discontinuities: inaccurate interpolation is confined
    # to steps shorter than 5 s (test_make_sgp4_polyjectory.py:609-665)
    gpes = H.golden_group("disc_gpe_01")
    jd_begin = 2460688.0  # 2025-01-12T12:00:00 UTC
    other = H.golden_group("single_gpe")
    assert other["norad_id"][0] < gpes["norad_id"][0]
    pj = mz.make_sgp4_polyjectory(np.concatenate([other, gpes]), jd_begin, jd_begin + 5.0)[0]
    cfs, times, status = pj[1]
    dur = np.diff(times) * 86400.0
    print("status", status, "steps", len(times) - 1, "shortest steps (s)", np.sort(dur)[:5])
    assert status == 0
    e_off = (gpes["epoch_jd1"][0] - jd_begin) + gpes["epoch_jd2"][0]
    assert pj.status[0] == 0
    for k in range(len(times) - 1):
        x = np.linspace(0.0, np.nextafter(times[k + 1] - times[k], -1.0), 200)
        st, err = mz.sgp4_propagate(gpes, (times[k] + x - e_off) * 1440.0)
        assert np.all(err == 0)
        ev = H.poly_eval(cfs[k], x)
        max_err = np.linalg.norm(ev[:, :3] - st[0, :, :3], axis=1).max()
        if max_err > 1e-6:
            assert (times[k + 1] - times[k]) * 86400.0 < 5.0


def test_input_forms_and_storage(tmp_path):
    # polars/pandas/pyarrow frames and Satrec-like lists give the same
    # polyjectory as the array (test_satrec_construction); data_dir, persist
    # and tmpdir as for polyjectory (test_custom_data_dir, test_persist, test_tmpdir)
    import pyarrow as pa

    gpes = H.golden_group("single_gpe")
    jd_begin = 2460669.0
    pj1 = mz.make_sgp4_polyjectory(gpes, jd_begin, jd_begin + 1)[0]
    table = pa.table({f: gpes[f] for f in H.GPE_FIELDS})
    pj2 = mz.make_sgp4_polyjectory(table, jd_begin, jd_begin + 1)[0]

    class Sat:
        pass

    s = Sat()
    for a, f in zip(["satnum", "jdsatepoch", "jdsatepochF", "no_kozai", "ecco", "inclo", "nodeo", "argpo", "mo", "bstar"],
                    H.GPE_FIELDS):
        setattr(s, a, gpes[f][0].item())
    pj3 = mz.make_sgp4_polyjectory([s], jd_begin, jd_begin + 1)[0]
    for p in (pj2, pj3):
        assert p.status[0] == pj1.status[0]
        assert np.array_equal(p[0][0], pj1[0][0]) and np.array_equal(p[0][1], pj1[0][1])
    with pytest.raises(TypeError, match="all elements of the list must be Satrec instances"):
        mz.make_sgp4_polyjectory([1, 2, 3], 0, 1.0)

    data_dir = (tmp_path / "data_dir").resolve()
    pj = mz.make_sgp4_polyjectory(gpes, jd_begin, jd_begin + 1, data_dir=data_dir)[0]
    assert data_dir == pj.data_dir
    pj.detach()
    assert not data_dir.exists()
    pj = mz.make_sgp4_polyjectory(gpes, jd_begin, jd_begin + 1, data_dir=data_dir, persist=True)[0]
    pj.detach()
    assert data_dir.is_dir()
    back = mz.polyjectory.mount(data_dir)
    assert np.array_equal(back[0][0], pj1[0][0]) and back.epoch == pj1.epoch
    tmpd = tmp_path / "tmp"
    tmpd.mkdir()
    pj = mz.make_sgp4_polyjectory(gpes, jd_begin, jd_begin + 1, tmpdir=tmpd)[0]
    assert any(tmpd.iterdir())
    pj.detach()


def test_screen_generated_catalog():
    # the generated polyjectory feeds the screening path: the GPU conjunctions
    # of the strack sample over 2 hours match the oracle's bit for bit
    from oracle import pyoracle as po

    g = H.sgp4_golden()
    gi = list(g["groups"]).index("strack_20240705")
    gpes = H.gpe_array(g["gpes"][g["group"] == gi])
    pj = mz.make_sgp4_polyjectory(gpes, 2460496.5, 2460496.5 + 2 / 24.0)[0]
    thresh, interval = 50.0, 1.0 / 1440.0
    c = mz.conjunctions(pj, thresh, interval, mode="fp64")
    ref = po.run(H.flat_of(pj), thresh, interval)
    assert np.array_equal(c.conjunctions.view(np.uint8), ref["conjunctions"].view(np.uint8))
    cm = mz.conjunctions(pj, thresh, interval, mode="mixed")
    assert np.array_equal(cm.conjunctions.view(np.uint8), c.conjunctions.view(np.uint8))


def test_screener_resident_generation():
    # generation into a screening context: the data never leaves HBM before
    # the screen; the result equals screening the copied-out polyjectory
    g = H.sgp4_golden()
    gi = list(g["groups"]).index("strack_20240705")
    gpes = H.gpe_array(g["gpes"][g["group"] == gi])
    sc = mz.Screener(0)
    status = sc.make_sgp4(gpes, 2460496.5, 2460496.5 + 2 / 24.0)
    pj = mz.make_sgp4_polyjectory(gpes, 2460496.5, 2460496.5 + 2 / 24.0)[0]
    assert np.array_equal(status, pj.status)
    c = mz.conjunctions(pj, 50.0, 1.0 / 1440.0, mode="fp64")
    sc.screen(50.0, 1.0 / 1440.0)
    assert np.array_equal(sc.conjunctions().view(np.uint8), c.conjunctions.view(np.uint8))
