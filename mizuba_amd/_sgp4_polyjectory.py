"""Polyjectories from SGP4 element sets, generated on the GPU.

Mirrors `mizuba.make_sgp4_polyjectory` (mizuba/_sgp4_polyjectory.py:31-220):
the same arguments, validation messages, status codes and return value
`(polyjectory, norad_ids)`. The trajectories are computed by the HIP SGP4
kernels behind `mzh_make_sgp4_polyjectory` (include/mizuba_amd.h).

Accepted forms of `gpes` (mizuba/_sgp4_polyjectory.py:150-200):
- a NumPy array of `gpe_dtype`;
- a dataframe with the GPE columns (`norad_id`, `epoch_jd1`, `epoch_jd2`,
  `n0`, `e0`, `i0`, `node0`, `omega0`, `m0`, `bstar`): polars or pandas
  DataFrames and pyarrow Tables;
- a list of `sgp4.api.Satrec`-like objects (attributes `satnum`,
  `jdsatepoch`, `jdsatepochF`, `no_kozai`, `ecco`, `inclo`, `nodeo`,
  `argpo`, `mo`, `bstar`).

The GPU generator propagates near-earth and deep-space element sets with
SGP4 / SDP4 (WGS-72, AFSPC mode) and interpolates them with order-8
polynomials.
"""

from __future__ import annotations

import enum as _enum
import os
from typing import Union

import numpy as np

from . import _core

gpe_dtype = _core.gpe_dtype


class sgp4_pj_status(_enum.IntEnum):
    """Per-satellite status of an SGP4 polyjectory (mizuba/_sgp4_polyjectory.py:31-47).

    1-6 are the SGP4 propagator's own error codes; 10-14 are the generator's.
    """

    OK = 0
    SGP4_ERROR_MEAN_ECC = 1
    SGP4_ERROR_MEAN_MEAN_MOTION = 2
    SGP4_ERROR_PERT_ECC = 3
    SGP4_ERROR_SEMI_LATUS_RECTUM = 4
    SGP4_DECAY = 6
    ERROR_NONFINITE = 10
    REENTRY = 11
    EXIT = 12
    ERROR_SMALL_STEP = 13
    ERROR_INTERPOLATION = 14


_gpe_fields = ["norad_id", "epoch_jd1", "epoch_jd2", "n0", "e0", "i0", "node0", "omega0", "m0", "bstar"]

_satrec_fields = ["satnum", "jdsatepoch", "jdsatepochF", "no_kozai", "ecco", "inclo", "nodeo", "argpo", "mo", "bstar"]


def _frame_to_array(gpes) -> np.ndarray:
    # Column-wise copy of a dataframe into gpe_dtype, with the column types
    # enforced as mizuba/_sgp4_polyjectory.py:150-164 does.
    names = list(gpes.columns) if not hasattr(gpes, "column_names") else list(gpes.column_names)
    missing = [f for f in _gpe_fields if f not in names]
    if missing:
        raise ValueError(
            f"The dataframe passed to make_sgp4_polyjectory() lacks the GPE column(s) {missing}"
        )
    out = np.zeros((len(gpes),), dtype=gpe_dtype)
    for f in _gpe_fields:
        col = gpes[f]
        if hasattr(col, "to_numpy"):
            col = col.to_numpy()
        out[f] = np.asarray(col).astype(gpe_dtype[f], copy=False)
    return out


def _is_frame(gpes) -> bool:
    mod = type(gpes).__module__
    return (mod.startswith("polars") or mod.startswith("pandas") or mod.startswith("pyarrow")) and hasattr(
        gpes, "__getitem__"
    )


def make_sgp4_polyjectory(
    gpes,
    jd_begin: float,
    jd_end: float,
    reentry_radius: float = 0.0,
    exit_radius: float = float("inf"),
    data_dir: Union[os.PathLike, str, None] = None,
    persist: bool = False,
    tmpdir: Union[os.PathLike, str, None] = None,
    *,
    device: int = 0,
):
    """Polyjectory of a set of GPEs over the UTC Julian dates [jd_begin, jd_end).

    GPEs must be sorted by NORAD id and then by epoch, with no duplicates.
    Returns `(pj, norad_ids)`: the polyjectory (one object per satellite, in
    NORAD id order, time origin `jd_begin` in TAI, time unit days, positions
    in km and velocities in km/s) and the sorted unique NORAD ids
    (mizuba/_sgp4_polyjectory.py:130-220, make_sgp4_polyjectory.cpp:1461-1600).
    """
    if any(not isinstance(x, (float, int)) for x in [jd_begin, jd_end, reentry_radius, exit_radius]):
        raise TypeError(
            "The jd_begin, jd_end, reentry_radius and exit_radius arguments to make_sgp4_polyjectory() must all be floats or ints"
        )

    if _is_frame(gpes):
        gpes_arr = _frame_to_array(gpes)
    elif isinstance(gpes, list):
        if any(not all(hasattr(sat, f) for f in _satrec_fields) for sat in gpes):
            raise TypeError(
                "When passing the 'gpes' argument to make_sgp4_polyjectory() as a list, all elements of the list must be Satrec instances"
            )
        gpes_arr = np.zeros((len(gpes),), dtype=gpe_dtype)
        gpes_arr[:] = [tuple(getattr(sat, f) for f in _satrec_fields) for sat in gpes]
    else:
        if not isinstance(gpes, np.ndarray):
            raise TypeError(
                f"The 'gpes' argument to make_sgp4_polyjectory() must be either a polars dataframe, a NumPy array or a list of Satrec objects, but it is of type '{type(gpes)}' instead"
            )
        if gpes.dtype != gpe_dtype:
            raise TypeError(
                f"When passing the 'gpes' argument to make_sgp4_polyjectory() as a NumPy array, the dtype must be 'gpe_dtype', but it is '{gpes.dtype}' instead"
            )
        gpes_arr = gpes

    pj = _core._make_sgp4_polyjectory(
        gpes_arr,
        float(jd_begin),
        float(jd_end),
        float(reentry_radius),
        float(exit_radius),
        None if data_dir is None else os.fspath(data_dir),
        bool(persist),
        None if tmpdir is None else os.fspath(tmpdir),
        8,
        int(device),
    )
    return pj, np.unique(gpes_arr["norad_id"])


def sgp4_propagate(gpes: np.ndarray, tsince, device: int = 0):
    """SGP4 states of GPEs at minutes since their epochs, computed on the GPU.

    `tsince` is 1-D (the same times for every GPE) or 2-D (one row per GPE).
    Returns `(states, errors)`: float64 [n_gpes, n_times, 6] (km, km/s, TEME)
    and int32 [n_gpes, n_times] SGP4 error codes (0: none).
    """
    if not isinstance(gpes, np.ndarray) or gpes.dtype != gpe_dtype:
        raise TypeError("sgp4_propagate() requires a NumPy array of gpe_dtype")
    return _core.sgp4_propagate(np.ascontiguousarray(gpes), np.asarray(tsince, dtype=np.float64), device)
