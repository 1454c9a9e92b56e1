"""The data-dir formats on both sides of the path, checked against the
oracle on the GPU.

- Input (polyjectory.cpp:91-236, 652-948): a polyjectory written to a
  persistent data dir and mounted again, and one built by from_datafiles from
  raw trajectory and time files, are screened; boxes, codes, sort order,
  trees, pairs and conjunctions are those the oracle computes from the
  original arrays, bit for bit.
- Output (detect_conjunctions.cpp:94-126, 466-534): every result file of a
  conjunctions object with a data dir holds the oracle's arrays byte for
  byte, with the reference's sentinels (one zeroed pair record, a one-byte
  conjunctions file) when there are no pairs or no conjunctions.
"""

import os
import tempfile
from pathlib import Path

import numpy as np
import pytest

import mizuba_amd as mz

from . import helpers as H
from .test_gpu_parity import assert_conj_equal, assert_screen_equal

pytestmark = pytest.mark.gpu

INTERVAL = 3.0 / 1440.0


def _catalog(n=1500, days=0.3, seed=31):
    pj = mz.make_synthetic_polyjectory(n, days, seed=seed, frac_high=0.2)
    trajs = [np.array(pj[i][0]) for i in range(pj.n_objs)]
    times = [np.array(pj[i][1]) for i in range(pj.n_objs)]
    status = [0] * pj.n_objs
    # objects without data, as a generator leaves failed satellites
    for k in (5, n // 2, n - 1):
        trajs[k] = np.ascontiguousarray(trajs[k][:0])
        times[k] = np.ascontiguousarray(times[k][:0])
        status[k] = 3
    return trajs, times, status


def test_screen_mounted_and_from_datafiles():
    from oracle import pyoracle as po

    trajs, times, status = _catalog()
    thresh = 8.0
    ref = po.run((trajs, times), thresh, INTERVAL)
    assert len(ref["conjunctions"]) > 0
    with tempfile.TemporaryDirectory() as d:
        dd = Path(d) / "pj"
        written = mz.polyjectory(trajs, times, status, data_dir=dd, persist=True)
        del written
        mounted = mz.polyjectory.mount(dd)
        assert mounted.persist and list(mounted.status) == status
        c = mz.conjunctions(mounted, thresh, INTERVAL)
        assert_screen_equal(c, ref)
        assert_conj_equal(c.conjunctions, ref["conjunctions"])
        assert c.np_report == ref["np_report"]
        # raw files moved into a data dir by from_datafiles
        op1 = trajs[0].shape[1]
        sizes = [t.shape[0] * op1 * 7 for t in trajs]
        offs = np.zeros(len(trajs), dtype=mz.polyjectory.traj_offset)
        offs["offset"] = np.concatenate([[0], np.cumsum(sizes)[:-1]])
        offs["n_steps"] = [t.shape[0] for t in trajs]
        tf, mf = Path(d) / "traj_raw", Path(d) / "time_raw"
        np.concatenate([t.ravel() for t in trajs]).tofile(tf)
        np.concatenate(times).tofile(mf)
        pjf = mz.polyjectory.from_datafiles(traj_file=tf, time_file=mf, order=op1 - 1, traj_offsets=offs,
                                            status=status, data_dir=Path(d) / "pjf")
        assert not tf.exists() and not mf.exists()
        c2 = mz.conjunctions(pjf, thresh, INTERVAL, mode="mixed")
        assert_conj_equal(c2.conjunctions, ref["conjunctions"])
        c3 = mz.conjunctions(pjf, thresh, INTERVAL)
        assert_screen_equal(c3, ref)
        assert_conj_equal(c3.conjunctions, ref["conjunctions"])


def _file(dd, name):
    return (dd / name).read_bytes()


def check_files_against_oracle(dd, ref, boxes=True):
    assert sorted(os.listdir(dd)) == sorted(
        ["aabbs", "srt_aabbs", "mcodes", "srt_mcodes", "vidx", "bvh", "bp", "conjunctions"])
    if boxes:
        assert _file(dd, "aabbs") == ref["aabbs"].tobytes()
        assert _file(dd, "srt_aabbs") == ref["srt_aabbs"].tobytes()
        assert _file(dd, "mcodes") == ref["mcodes"].tobytes()
        assert _file(dd, "srt_mcodes") == ref["srt_mcodes"].tobytes()
        assert _file(dd, "vidx") == ref["srt_idx"].tobytes()
        assert _file(dd, "bvh") == ref["trees"].tobytes()
        assert _file(dd, "bp") == (ref["bp"].tobytes() if len(ref["bp"]) else bytes(8))
    assert _file(dd, "conjunctions") == (ref["conjunctions"].tobytes() if len(ref["conjunctions"]) else b"\0")
    for f in os.listdir(dd):
        assert (dd / f).stat().st_mode & 0o777 == 0o400


def test_result_files_against_oracle():
    from oracle import pyoracle as po

    trajs, times, status = _catalog(600, 0.2, seed=7)
    pj = mz.polyjectory(trajs, times, status)
    ref = po.run((trajs, times), 25.0, INTERVAL)
    assert len(ref["conjunctions"]) > 0 and len(ref["bp"]) > 0
    with tempfile.TemporaryDirectory() as d:
        dd = (Path(d) / "res").resolve()
        c = mz.conjunctions(pj, 25.0, INTERVAL, data_dir=dd)
        check_files_against_oracle(dd, ref)
        del c
        assert not dd.exists()
        # mixed mode: boxes differ by design, the conjunctions file does not
        c = mz.conjunctions(pj, 25.0, INTERVAL, tmpdir=d, mode="mixed")
        check_files_against_oracle(c.data_dir, ref, boxes=False)
        del c
    # no pairs and no conjunctions: the reference's sentinels
    tcs, tms = H.planar_circ()
    ref0 = po.run(([tcs], [tms]), 0.1, 0.01)
    assert len(ref0["bp"]) == 0 and len(ref0["conjunctions"]) == 0
    with tempfile.TemporaryDirectory() as d:
        c = mz.conjunctions(mz.polyjectory([tcs], [tms], [0]), 0.1, 0.01, tmpdir=d)
        check_files_against_oracle(c.data_dir, ref0)
        assert _file(c.data_dir, "bp") == bytes(8) and _file(c.data_dir, "conjunctions") == b"\0"
