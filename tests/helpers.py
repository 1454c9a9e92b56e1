"""Shared test helpers: reference closed-form trajectories and comparisons."""

import numpy as np

GOLDEN = __import__("pathlib").Path(__file__).resolve().parent / "golden"


def planar_circ():
    g = np.load(GOLDEN / "planar_circ.npz")
    return g["tcs"], g["times"]


def linear_traj(tms, sgn):
    """Objects on the x axis at +-1 moving towards the origin at unit speed
    (test/test_conjunctions.py:1076-1324); order 3."""
    out = []
    for tm in tms[1:]:
        td = np.zeros((7, 4))
        td[0, 0] = sgn * (1.0 - (tm - 0.1))
        td[0, 1] = -1.0 * sgn
        td[3, 0] = -1.0 * sgn
        out.append(td)
    return np.ascontiguousarray(np.array(out).transpose((0, 2, 1)))


TM21 = np.array([round(0.1 * k, 10) for k in range(21)])

# (slice of object 0, slice of object 1, number of conjunctions expected)
CD_BEGIN_END_CASES = [
    ((0, 10), (11, 21), 0),
    ((0, 12), (9, 21), 1),
    ((11, 21), (12, 21), 0),
    ((9, 21), (8, 21), 1),
    ((0, 9), (0, 8), 0),
    ((0, 11), (0, 12), 1),
]


def cd_begin_end_case(k):
    (a0, b0), (a1, b1), exp = CD_BEGIN_END_CASES[k]
    t0, t1 = TM21[a0:b0], TM21[a1:b1]
    return [linear_traj(t0, 1), linear_traj(t1, -1)], [t0, t1], exp


def boundary_incoming(shift=0.0):
    """test/test_boundary_conjunctions.py:24-123."""
    t0 = np.array([0.0, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9])
    t1 = np.array([round(0.1 * k, 10) for k in range(14)])
    tr = [[], []]
    for tm in t1[1:]:
        td = np.zeros((7, 4))
        td[0, 0] = 1.0 - (tm - 0.1)
        td[0, 1] = -1.0
        td[3, 0] = -1.0
        if tm <= 0.9:
            tr[0].append(td.T)
        tr[1].append((-td).T)
    return [np.ascontiguousarray(np.array(x)) for x in tr], [t0 + shift, t1 + shift]


def boundary_outgoing(shift=0.0):
    """test/test_boundary_conjunctions.py:125-198."""
    tm_data = np.array([0.0, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7, 0.8, 0.9])
    tr = [[], []]
    for tm in tm_data[1:]:
        td = np.zeros((7, 4))
        td[0, 0] = 0.1 + (tm - 0.1)
        td[0, 1] = 1.0
        td[3, 0] = 1.0
        tr[0].append(td.T)
        tr[1].append((-td).T)
    return [np.ascontiguousarray(np.array(x)) for x in tr], [tm_data + shift, tm_data + shift]


def zero_aabbs_pj():
    """test/test_conjunctions.py:586-626."""
    tdata = np.zeros((7, 6))
    tdata[:3, 0] = 1.0
    tdata[6, 0] = np.sqrt(3.0)
    tdata = np.ascontiguousarray(tdata.transpose())
    return [np.array([tdata, tdata, tdata])], [np.array([0.0, 1.0, 2.0, 3.0])]


def axis_objects():
    """Eight order-5 constant objects of test_bvh / test_no_traj_data
    (test/test_conjunctions.py:628-784)."""
    objs = []
    for ch, v in [(0, 1.0), (0, -1.0), (1, 1.0), (1, -1.0), (2, 1.0), (2, -1.0)]:
        t = np.zeros((6, 7))
        t[0, ch] = v
        objs.append(t)
    objs.append(np.zeros((6, 7)))
    t7 = np.zeros((6, 7))
    t7[0, :] = 1
    objs.append(t7)
    return objs


def flat_of(pj):
    from oracle import pyoracle as po

    traj, time_, toff, ns, tmo = pj._flat()
    return po.Flat.from_arrays(pj.poly_order, traj, time_, ns, toff, tmo, pj.maxT)


def eval_pj_arrays():
    """The two-object polyjectory of the reference's test_eval
    (test/test_polyjectory.py:792-844): x(t) = +-(1 - (t - 0.1)) at unit speed,
    object 0 over [0, 0.9], object 1 over [1.1, 2.0]."""
    tm = np.array([round(0.1 * k, 10) for k in range(21)])
    return [linear_traj(tm[:10], 1), linear_traj(tm[11:], -1)], [tm[:10], tm[11:]]


# ---- SGP4 -------------------------------------------------------------------

GPE_FIELDS = ["norad_id", "epoch_jd1", "epoch_jd2", "n0", "e0", "i0", "node0", "omega0", "m0", "bstar"]


def sgp4_golden():
    """The SGP4 fixture (tests/golden/make_sgp4_golden.py): the reference's
    GPE files with the vendored Vallado SGP4's states."""
    return np.load(GOLDEN / "sgp4_vallado.npz")


def gpe_array(rows):
    """[n, 10] float64 GPE fields -> a gpe_dtype array."""
    import mizuba_amd as mz

    rows = np.asarray(rows, dtype=np.float64).reshape(-1, 10)
    out = np.zeros(len(rows), dtype=mz.gpe_dtype)
    for k, f in enumerate(GPE_FIELDS):
        out[f] = rows[:, k]
    return out


def golden_group(name):
    """The GPEs of one file of the SGP4 fixture, as a gpe_dtype array."""
    g = sgp4_golden()
    gi = list(g["groups"]).index(name)
    return gpe_array(g["gpes"][g["group"] == gi])


def _jd_jan1(year):
    # Julian date of January 1, 0h (the Gregorian calendar formula of
    # Vallado's jday with month 1, day 1)
    return 367.0 * year - np.floor(7.0 * (year + np.floor(10.0 / 12.0)) * 0.25) + np.floor(275.0 / 9.0) + 1.0 + 1721013.5


def gpe_from_tle(line1, line2, norad_id=None):
    """A GPE from a two-line element set (the published column layout: epoch
    year and day at 1:18-32, B* at 1:53-61, inclination, node, eccentricity,
    argument of perigee, mean anomaly and mean motion at 2:8-63)."""
    yy = int(line1[18:20])
    year = 2000 + yy if yy < 57 else 1900 + yy
    doy = float(line1[20:32])
    b = line1[53:61]
    bstar = float(b[0] + "0." + b[1:6].strip()) * 10.0 ** int(b[6:8])
    deg = np.pi / 180.0
    row = [
        float(int(line1[2:7]) if norad_id is None else norad_id),
        _jd_jan1(year) + np.floor(doy) - 1.0,
        doy - np.floor(doy),
        float(line2[52:63]) * 2.0 * np.pi / 1440.0,
        float("0." + line2[26:33].strip()),
        float(line2[8:16]) * deg,
        float(line2[17:25]) * deg,
        float(line2[34:42]) * deg,
        float(line2[43:51]) * deg,
        bstar,
    ]
    return row


def poly_eval(cfs, x):
    """Horner of the [op1, 7] coefficients of one step at x -> [len(x), 7]."""
    x = np.asarray(x, dtype=np.float64)
    r = np.zeros((len(x), 7)) + cfs[-1]
    for k in range(cfs.shape[0] - 2, -1, -1):
        r = cfs[k] + r * x[:, None]
    return r
