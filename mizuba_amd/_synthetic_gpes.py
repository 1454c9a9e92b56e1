"""Seeded synthetic catalogs as SGP4 mean elements.

The element distributions are those of the synthetic Keplerian catalogs
(`make_synthetic_polyjectory`):
- 80% LEO: perigee altitude U[400, 1200] km, e U[0, 0.01];
- 20% higher: perigee altitude U[1200, 2000] km, e U[0, 0.05];
- inclination U[0, 100] deg; node, argument of perigee and mean anomaly
  U[0, 360) deg; B* U[1e-5, 1e-4] (1 / Earth radii).

`n_shell` objects first sit in 5 near-circular shells (altitudes
540 / 550 / 560 / 570 / 580 km, e < 1e-4, inclination 53 or 97.6 deg, 72 evenly
spaced planes, evenly phased within them).

All element sets share the epoch `epoch_jd` (UTC Julian date), and NORAD ids
run from 1. The mean motion is the Keplerian one of the sampled semimajor axis
(WGS-72), in rad/min. `make_sgp4_polyjectory` turns the result into a
polyjectory.
"""

from __future__ import annotations

import numpy as np

from . import _core

_MU = 398600.8  # km^3 / s^2 (WGS-72)
_RE = 6378.135  # km


def make_synthetic_gpes(
    n_objs: int, seed: int = 42, frac_high: float = 0.2, n_shell: int = 0, epoch_jd: float = 2460496.5
) -> np.ndarray:
    """A seeded synthetic catalog of `n_objs` GPEs (an array of `gpe_dtype`)."""
    if n_objs < 0 or n_shell < 0 or n_shell > n_objs:
        raise ValueError("make_synthetic_gpes() needs 0 <= n_shell <= n_objs")
    rng = np.random.default_rng(seed)
    n = int(n_objs)
    g = np.zeros(n, dtype=_core.gpe_dtype)
    g["norad_id"] = np.arange(1, n + 1, dtype=np.uint64)
    a = np.empty(n)
    e = np.empty(n)
    inc = np.empty(n)
    raan = np.empty(n)
    m0 = np.empty(n)
    deg = np.pi / 180.0
    ns = int(n_shell)
    if ns:
        k = np.arange(ns)
        shell, plane, slot = k % 5, (k // 5) % 72, k // 360
        per_plane = max(1, (ns + 359) // 360)
        alt = 540.0 + 10.0 * shell
        e[:ns] = rng.uniform(0.0, 1e-4, ns)
        a[:ns] = (_RE + alt) / (1.0 - e[:ns])
        inc[:ns] = np.where(plane % 2 == 1, 97.6, 53.0) * deg
        raan[:ns] = 2.0 * np.pi * plane / 72.0
        m0[:ns] = 2.0 * np.pi * slot / per_plane + 2.0 * np.pi * shell / (5.0 * per_plane)
    nr = n - ns
    high = rng.uniform(0.0, 1.0, nr) < frac_high
    hp = np.where(high, rng.uniform(1200.0, 2000.0, nr), rng.uniform(400.0, 1200.0, nr))
    e[ns:] = np.where(high, rng.uniform(0.0, 0.05, nr), rng.uniform(0.0, 0.01, nr))
    a[ns:] = (_RE + hp) / (1.0 - e[ns:])
    inc[ns:] = rng.uniform(0.0, 100.0, nr) * deg
    raan[ns:] = rng.uniform(0.0, 2.0 * np.pi, nr)
    m0[ns:] = rng.uniform(0.0, 2.0 * np.pi, nr)
    g["n0"] = np.sqrt(_MU / a**3) * 60.0
    g["e0"] = e
    g["i0"] = inc
    g["node0"] = raan
    g["omega0"] = rng.uniform(0.0, 2.0 * np.pi, n)
    g["m0"] = m0
    g["bstar"] = rng.uniform(1e-5, 1e-4, n)
    g["epoch_jd1"] = np.floor(epoch_jd - 0.5) + 0.5
    g["epoch_jd2"] = epoch_jd - g["epoch_jd1"]
    return g
