"""Write the SGP4 fixture tests/golden/sgp4_vallado.npz.

Inputs are the GPE data files of the reference's test suite
(mizuba/test/{single_gpe,single_gpe_ds,multi_gpes,iss_gpes,disc_gpe_01,
syncom_gpe,strack_20240705}.parquet, read as data with pyarrow; 256 GPEs of
the strack catalog picked with a fixed seed). Expected outputs come from the
reference's own vendored Vallado SGP4 (mizuba/detail/sgp4/SGP4.cpp), compiled
from where it lies into oracle/_ref/libref_sgp4.so by `make -C oracle ref`
and initialised as make_sgp4_polyjectory does (WGS-72, opsmode 'a',
make_sgp4_polyjectory.cpp:722-758).

Stored arrays:
  gpes       [n, 10] float64 fields (norad_id, epoch_jd1, epoch_jd2, n0, e0,
             i0, node0, omega0, m0, bstar), norad_id as float64 (exact)
  group      [n] int32 index into `groups` (the file each GPE comes from)
  groups     the file names
  tsince     [n, K] minutes from each GPE's epoch (0 and K - 1 seeded
             uniform values in [-4320, 4320])
  states     [n, K, 6] x, y, z (km), vx, vy, vz (km/s), TEME
  errors     [n, K] int32 SGP4 error codes (the sgp4init code for all K when
             initialisation fails)
  deep       [n] int32, 1 for a deep-space GPE

It also writes the two full space-track catalogs of the reference's tests
(strack_20240705.npz and strack_20240917.npz: `gpes`, [n, 10] float64 in the
field order above), the data of test_make_sgp4_polyjectory.py:755-817.

Run only where the reference tree is present; the .npz files are the
committed fixtures.
"""

import ctypes
import pathlib

import numpy as np
import pyarrow.parquet as pq

HERE = pathlib.Path(__file__).resolve().parent
REF_TEST = pathlib.Path("/root/reference/mizuba/test")
LIB = HERE.parent.parent / "oracle" / "_ref" / "libref_sgp4.so"
OUT = HERE / "sgp4_vallado.npz"

FIELDS = ["norad_id", "epoch_jd1", "epoch_jd2", "n0", "e0", "i0", "node0", "omega0", "m0", "bstar"]
GROUPS = ["single_gpe", "single_gpe_ds", "multi_gpes", "iss_gpes", "disc_gpe_01", "syncom_gpe", "strack_20240705"]
K = 9
N_STRACK = 256


def main():
    lib = ctypes.CDLL(str(LIB))
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    lib.ref_sgp4_propagate.argtypes = [dp, dp, ctypes.c_int64, dp, ip]
    lib.ref_sgp4_propagate.restype = ctypes.c_int
    lib.ref_sgp4_is_deep_space.argtypes = [dp]
    lib.ref_sgp4_is_deep_space.restype = ctypes.c_int
    rng = np.random.default_rng(20240705)
    rows, group = [], []
    for gi, name in enumerate(GROUPS):
        t = pq.read_table(REF_TEST / f"{name}.parquet")
        arr = np.stack([np.asarray(t[f].to_numpy(), dtype=np.float64) for f in FIELDS], axis=1)
        if name == "strack_20240705":
            arr = arr[np.sort(rng.choice(len(arr), N_STRACK, replace=False))]
        rows.append(arr)
        group += [gi] * len(arr)
    gpes = np.ascontiguousarray(np.concatenate(rows))
    n = len(gpes)
    tsince = np.zeros((n, K))
    tsince[:, 1:] = rng.uniform(-4320.0, 4320.0, (n, K - 1))
    states = np.zeros((n, K, 6))
    errors = np.zeros((n, K), dtype=np.int32)
    deep = np.zeros(n, dtype=np.int32)
    for i in range(n):
        g = np.ascontiguousarray(gpes[i])
        ts = np.ascontiguousarray(tsince[i])
        out = np.zeros((K, 6))
        err = np.zeros(K, dtype=np.int32)
        rc = lib.ref_sgp4_propagate(g.ctypes.data_as(dp), ts.ctypes.data_as(dp), K, out.ctypes.data_as(dp),
                                    err.ctypes.data_as(ip))
        deep[i] = lib.ref_sgp4_is_deep_space(g.ctypes.data_as(dp))
        if rc != 0:
            err[:] = rc
            out[:] = 0.0
        states[i] = out
        errors[i] = err
    np.savez_compressed(OUT, gpes=gpes, group=np.array(group, dtype=np.int32), groups=np.array(GROUPS),
                        tsince=tsince, states=states, errors=errors, deep=deep)
    for name in ("strack_20240705", "strack_20240917"):
        t = pq.read_table(REF_TEST / f"{name}.parquet")
        arr = np.stack([np.asarray(t[f].to_numpy(), dtype=np.float64) for f in FIELDS], axis=1)
        np.savez_compressed(HERE / f"{name}.npz", gpes=arr)
        print(HERE / f"{name}.npz", arr.shape)
    print(OUT, n, "GPEs,", int(deep.sum()), "deep space,", int((errors != 0).sum()), "error entries")


if __name__ == "__main__":
    main()
