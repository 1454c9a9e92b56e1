"""SGP4 pieces that need no GPU: the fixture against the reference's own
Vallado SGP4 (when oracle/_ref is built), the Python front end's checks and
types (mizuba/_sgp4_polyjectory.py), and the TLE helper of the tests."""

import ctypes
import pathlib

import numpy as np
import pytest

import mizuba_amd as mz

from . import helpers as H

ROOT = pathlib.Path(__file__).resolve().parent.parent
REF_SGP4 = ROOT / "oracle" / "_ref" / "libref_sgp4.so"


def test_fixture_shape():
    g = H.sgp4_golden()
    n, k = g["tsince"].shape
    assert g["gpes"].shape == (n, 10) and g["states"].shape == (n, k, 6) and g["errors"].shape == (n, k)
    assert np.all(g["tsince"][:, 0] == 0.0)
    assert set(g["groups"]) >= {"single_gpe", "single_gpe_ds", "multi_gpes", "iss_gpes", "disc_gpe_01"}
    # both kinds of element sets and some SGP4 errors are covered
    assert 0 < g["deep"].sum() < n and np.any(g["errors"] != 0)
    # the near-earth states are orbits: radius of LEO-to-MEO size, speed below escape
    ok = (g["deep"] == 0)[:, None] & (g["errors"] == 0)
    r = np.linalg.norm(g["states"][..., :3], axis=2)[ok]
    v = np.linalg.norm(g["states"][..., 3:], axis=2)[ok]
    assert r.min() > 6378.0 and r.max() < 20000.0 and v.max() < 11.2


def test_strack_fixtures():
    # the reference's two space-track catalogs, sorted by NORAD ID then epoch
    for name, n in (("strack_20240705", 26389), ("strack_20240917", 26905)):
        g = np.load(H.GOLDEN / f"{name}.npz")["gpes"]
        assert g.shape == (n, 10) and np.all(np.isfinite(g))
        ids = g[:, 0]
        assert np.all(np.diff(ids) >= 0)
        same = np.diff(ids) == 0
        ep = g[:, 1] + g[:, 2]
        assert np.all(np.diff(ep)[same] > 0)


@pytest.mark.skipif(not REF_SGP4.exists(), reason="oracle/_ref not built (make -C oracle ref)")
def test_fixture_is_the_reference_sgp4():
    # the fixture's states are the reference's vendored SGP4 bit for bit
    lib = ctypes.CDLL(str(REF_SGP4))
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    lib.ref_sgp4_propagate.argtypes = [dp, dp, ctypes.c_int64, dp, ip]
    lib.ref_sgp4_propagate.restype = ctypes.c_int
    g = H.sgp4_golden()
    for i in range(0, len(g["gpes"]), 7):
        gp = np.ascontiguousarray(g["gpes"][i])
        ts = np.ascontiguousarray(g["tsince"][i])
        out = np.zeros((len(ts), 6))
        err = np.zeros(len(ts), dtype=np.int32)
        rc = lib.ref_sgp4_propagate(gp.ctypes.data_as(dp), ts.ctypes.data_as(dp), len(ts), out.ctypes.data_as(dp),
                                    err.ctypes.data_as(ip))
        if rc != 0:
            assert np.all(g["errors"][i] == rc)
            continue
        assert np.array_equal(err, g["errors"][i])
        assert np.array_equal(out, g["states"][i])


def test_gpe_dtype_and_status_codes():
    # mizuba/_sgp4_polyjectory.py:31-47, make_sgp4_polyjectory.hpp:35-47
    assert mz.gpe_dtype.names == tuple(H.GPE_FIELDS)
    assert mz.gpe_dtype.itemsize == 80
    s = mz.sgp4_pj_status
    assert [s.OK, s.SGP4_ERROR_MEAN_ECC, s.SGP4_ERROR_MEAN_MEAN_MOTION, s.SGP4_ERROR_PERT_ECC,
            s.SGP4_ERROR_SEMI_LATUS_RECTUM, s.SGP4_DECAY] == [0, 1, 2, 3, 4, 6]
    assert [s.ERROR_NONFINITE, s.REENTRY, s.EXIT, s.ERROR_SMALL_STEP, s.ERROR_INTERPOLATION] == [10, 11, 12, 13, 14]


def test_front_end_type_checks():
    # the checks made before any native call (mizuba/_sgp4_polyjectory.py:141-200)
    arr = np.zeros((2,), dtype=mz.gpe_dtype)
    msg = "The jd_begin, jd_end, reentry_radius and exit_radius arguments to make_sgp4_polyjectory() must all be floats or ints"
    with pytest.raises(TypeError) as ei:
        mz.make_sgp4_polyjectory(arr, [], 1.0, reentry_radius=3, exit_radius=2)
    assert msg in str(ei.value)
    with pytest.raises(TypeError) as ei:
        mz.make_sgp4_polyjectory(arr, 0, 1.0, reentry_radius=3, exit_radius=[])
    assert msg in str(ei.value)
    with pytest.raises(TypeError) as ei:
        mz.make_sgp4_polyjectory(1.0, 0, 1.0)
    assert ("The 'gpes' argument to make_sgp4_polyjectory() must be either a polars dataframe, a NumPy array or a "
            "list of Satrec objects, but it is of type") in str(ei.value)
    with pytest.raises(TypeError) as ei:
        mz.make_sgp4_polyjectory(np.zeros((1,)), 0, 1.0)
    assert ("When passing the 'gpes' argument to make_sgp4_polyjectory() as a NumPy array, the dtype must be "
            "'gpe_dtype', but it is '") in str(ei.value)
    with pytest.raises(TypeError) as ei:
        mz.make_sgp4_polyjectory([1, 2, 3], 0, 1.0)
    assert "all elements of the list must be Satrec instances" in str(ei.value)
    # shape checks of the native front end come before the device is needed
    with pytest.raises(ValueError, match="must have 1 dimension, but the number of dimensions is 2 instead"):
        mz.make_sgp4_polyjectory(np.zeros((1, 1), dtype=mz.gpe_dtype), 0.0, 1.0)
    with pytest.raises(ValueError, match="must be C contiguous and properly aligned"):
        mz.make_sgp4_polyjectory(np.zeros((10,), dtype=mz.gpe_dtype)[::2], 0.0, 1.0)


def test_frames_convert_to_gpe_arrays():
    import pyarrow as pa

    gpes = H.golden_group("multi_gpes")
    t = pa.table({f: gpes[f] for f in H.GPE_FIELDS[::-1]})
    from mizuba_amd import _sgp4_polyjectory as sp

    assert sp._is_frame(t)
    out = sp._frame_to_array(t)
    assert out.dtype == mz.gpe_dtype and np.array_equal(out, gpes)
    with pytest.raises(ValueError, match="lacks the GPE column"):
        sp._frame_to_array(pa.table({"norad_id": gpes["norad_id"]}))


def test_tle_helper():
    # known Julian dates of January 1 and a TLE epoch (2005 day 363.79166667)
    assert H._jd_jan1(2005) == 2453371.5 and H._jd_jan1(2024) == 2460310.5 and H._jd_jan1(2000) == 2451544.5
    row = H.gpe_from_tle(
        "1 00045U 60007A   05363.79166667  .00000504  00000-0  14841-3 0  9992",
        "2 00045  66.6943  81.3521 0257384 317.3173  40.8180 14.34783636277898",
    )
    assert row[0] == 45 and row[1] == 2453733.5 and abs(row[2] - 0.79166667) < 1e-12
    assert row[3] == pytest.approx(14.34783636 * 2 * np.pi / 1440.0, rel=1e-15)
    assert row[4] == 0.0257384 and row[5] == pytest.approx(np.radians(66.6943), rel=1e-15)
    assert row[9] == pytest.approx(0.14841e-3, rel=1e-15)
    neg = H.gpe_from_tle(
        "1 00045U 60007A   05363.79166667  .00000504  00000-0 -14841-3 0  9992",
        "2 00045  66.6943  81.3521 0257384 317.3173  40.8180 14.34783636277898",
    )
    assert neg[9] == pytest.approx(-0.14841e-3, rel=1e-15)


def test_synthetic_gpes():
    # the configurations' element distributions, as SGP4 mean elements
    g = mz.make_synthetic_gpes(20000, seed=7)
    assert g.dtype == mz.gpe_dtype and np.array_equal(g["norad_id"], np.arange(1, 20001))
    a = (398600.8 / (g["n0"] / 60.0) ** 2) ** (1.0 / 3.0)
    hp = a * (1.0 - g["e0"]) - 6378.135
    assert hp.min() > 400.0 - 1e-6 and hp.max() < 2000.0 + 1e-6
    low = hp < 1200.0
    assert 0.75 < low.mean() < 0.85 and g["e0"][low].max() <= 0.01 and g["e0"].max() <= 0.05
    assert np.degrees(g["i0"]).max() <= 100.0 and g["bstar"].min() >= 1e-5 and g["bstar"].max() <= 1e-4
    assert np.array_equal(mz.make_synthetic_gpes(100, seed=7), mz.make_synthetic_gpes(100, seed=7))
    s = mz.make_synthetic_gpes(1000, seed=1, n_shell=720)
    a_s = (398600.8 / (s["n0"][:720] / 60.0) ** 2) ** (1.0 / 3.0)
    alt = a_s * (1.0 - s["e0"][:720]) - 6378.135
    assert np.allclose(np.sort(np.unique(np.round(alt))), [540, 550, 560, 570, 580])

