"""Result files of a conjunctions object with a data dir
(detect_conjunctions.cpp:91-126, 455-534; test/test_conjunctions.py:500-580):
the eight files hold exactly what the accessors return, with the reference's
sentinels when there are no pairs or no conjunctions, and the dir goes with
the object."""

import os
import tempfile
from pathlib import Path

import numpy as np
import pytest

import mizuba_amd as mz

from . import helpers as H

pytestmark = pytest.mark.gpu


def _bytes(a):
    return np.ascontiguousarray(a).view(np.uint8).tobytes()


def check_files(c, dd):
    assert sorted(os.listdir(dd)) == sorted(
        ["aabbs", "srt_aabbs", "mcodes", "srt_mcodes", "vidx", "bvh", "bp", "conjunctions"])
    assert (dd / "aabbs").read_bytes() == _bytes(c.aabbs)
    assert (dd / "srt_aabbs").read_bytes() == _bytes(c.srt_aabbs)
    assert (dd / "mcodes").read_bytes() == _bytes(c.mcodes)
    assert (dd / "srt_mcodes").read_bytes() == _bytes(c.srt_mcodes)
    assert (dd / "vidx").read_bytes() == _bytes(c.srt_idx)
    trees = b"".join(_bytes(c.get_bvh_tree(k)) for k in range(c.n_cd_steps))
    assert (dd / "bvh").read_bytes() == trees
    pairs = b"".join(_bytes(c.get_aabb_collisions(k)) for k in range(c.n_cd_steps))
    assert (dd / "bp").read_bytes() == (pairs if pairs else bytes(8))
    cj = _bytes(c.conjunctions)
    assert (dd / "conjunctions").read_bytes() == (cj if cj else b"\0")
    for f in os.listdir(dd):
        assert (dd / f).stat().st_mode & 0o777 == 0o400


def test_data_dir_files():
    pj = mz.make_synthetic_polyjectory(400, 0.25, seed=11)
    with tempfile.TemporaryDirectory() as d:
        dd = (Path(d) / "data_dir").resolve()
        c = mz.conjunctions(pj, 25.0, 3.0 / 1440.0, data_dir=dd, mode="mixed")
        assert c.data_dir == dd and dd.is_dir()
        assert len(c.conjunctions) > 0
        check_files(c, dd)
        del c
        assert not dd.exists()


def test_empty_sentinels_and_tmpdir():
    # one object: no pairs, no conjunctions
    tcs, times = H.planar_circ()
    pj = mz.polyjectory([tcs], [times], [0])
    with tempfile.TemporaryDirectory() as d:
        c = mz.conjunctions(pj, conj_thresh=0.1, conj_det_interval=0.01, tmpdir=d)
        dd = c.data_dir
        assert dd.parent == Path(d).resolve() and dd.name.startswith("mizuba_conjunctions-")
        assert (dd / "bp").read_bytes() == bytes(8)
        assert (dd / "conjunctions").read_bytes() == b"\0"
        check_files(c, dd)
        del c
        assert not any(Path(d).iterdir())
    # no data dir asked for: results stay in memory
    c = mz.conjunctions(pj, conj_thresh=0.1, conj_det_interval=0.01, data_dir="")
    assert c.data_dir is None
