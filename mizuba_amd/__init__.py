"""mizuba_amd -- MI355X-native conjunction screening over polyjectories.

Drop-in for the conjunction-detection path of mizuba
(`mizuba.conjunctions(pj, conj_thresh, conj_det_interval, otypes=None, ...)`):
the same constructor arguments, accessors, array layouts and record dtypes,
computed by hand-written gfx950 HIP kernels behind the C ABI declared in
include/mizuba_amd.h.

Extra keyword arguments of `conjunctions`: `mode` ("fp64": the reference's
arithmetic, bit-exact screen; "mixed": FP32 enclosures with a rigorous margin,
same conjunction list) and `device` (HIP device index).
`conjunctions_batch(pjs, conj_threshs, conj_det_intervals, otypes=None, mode=...)`
screens several independent problems together and returns one `conjunctions`
object per problem, identical to constructing them one by one.

`make_sgp4_polyjectory(gpes, jd_begin, jd_end, ...)` generates the
polyjectory of a set of SGP4 element sets on the GPU (near-earth satellites).

The native library must be built (`python mizuba_amd/build_native.py`); there
is no CPU fallback: without a HIP device, constructing `conjunctions` raises
RuntimeError.
"""

from __future__ import annotations

import enum as _enum

try:
    from . import _core
except ImportError as _e:  # pragma: no cover - depends on the build
    raise ImportError(
        "mizuba_amd native extension is not built; run `python mizuba_amd/build_native.py` first"
    ) from _e

polyjectory = _core.polyjectory
conjunctions = _core.conjunctions
conjunctions_batch = _core.conjunctions_batch
Screener = _core.Screener
Batch = _core.Batch
screen_batch = _core.screen_batch
make_synthetic_polyjectory = _core.make_synthetic_polyjectory
device_count = _core.device_count
check_screen = _core.check_screen
get_tmpdir = _core.get_tmpdir
set_tmpdir = _core.set_tmpdir

from ._sgp4_polyjectory import gpe_dtype, make_sgp4_polyjectory, sgp4_pj_status, sgp4_propagate  # noqa: E402
from ._synthetic_gpes import make_synthetic_gpes  # noqa: E402


class otype(_enum.IntEnum):
    """Object types for conjunction screening (mizuba/__init__.py:61-66)."""

    PRIMARY = 1
    SECONDARY = 2
    MASKED = 4


__all__ = [
    "polyjectory",
    "conjunctions",
    "conjunctions_batch",
    "otype",
    "Screener",
    "Batch",
    "screen_batch",
    "make_synthetic_polyjectory",
    "device_count",
    "check_screen",
    "get_tmpdir",
    "set_tmpdir",
    "gpe_dtype",
    "make_sgp4_polyjectory",
    "sgp4_pj_status",
    "sgp4_propagate",
    "make_synthetic_gpes",
]
