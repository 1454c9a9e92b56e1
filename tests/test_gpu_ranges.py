"""GPU: screenings in ranges of conjunction steps (mzh_screen_steps).

A problem beyond one batch's size limit, or beyond the device's free memory,
is screened in ranges of steps with the results kept in host memory, as the
reference processes its steps in chunks (detect_conjunctions.cpp:61-79,
184-242). The results must be the one-batch screening's bit for bit: every
per-step array, tree and pair list, the conjunctions and the narrow-phase
counters. MIZUBA_AMD_CHUNK_STEPS caps the range size to force the path at
small sizes.
"""

import numpy as np
import pytest

import mizuba_amd as mz

from . import helpers as H

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint8)


def assert_same_results(a, b):
    assert a.n_cd_steps == b.n_cd_steps
    for name in ("aabbs", "srt_aabbs", "mcodes", "srt_mcodes", "srt_idx", "cd_end_times", "otypes"):
        assert np.array_equal(_bits(getattr(a, name)), _bits(getattr(b, name))), name
    for k in range(a.n_cd_steps):
        assert np.array_equal(_bits(a.get_bvh_tree(k)), _bits(b.get_bvh_tree(k))), f"tree {k}"
        assert np.array_equal(_bits(a.get_aabb_collisions(k)), _bits(b.get_aabb_collisions(k))), f"pairs {k}"
    assert np.array_equal(_bits(a.conjunctions), _bits(b.conjunctions))
    assert a.np_report == b.np_report


@pytest.fixture(scope="module")
def leo1000():
    return mz.make_synthetic_polyjectory(1000, 1.0, seed=11, frac_high=0.0)


@pytest.mark.parametrize("chunk", [7, 128])
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_ranges_equal_one_batch(leo1000, chunk, mode, monkeypatch):
    """480 steps in ranges of 7 (the last one of 4) and of 128 steps, against
    the one-batch screening; the FP64 result also against the oracle."""
    one = mz.conjunctions(leo1000, 5.0, 3.0 / 1440.0, mode=mode)
    assert one.chunk_steps == 0
    monkeypatch.setenv("MIZUBA_AMD_CHUNK_STEPS", str(chunk))
    rng = mz.conjunctions(leo1000, 5.0, 3.0 / 1440.0, mode=mode)
    assert rng.chunk_steps == chunk
    assert_same_results(rng, one)
    assert len(rng.conjunctions) > 0
    with pytest.raises(IndexError, match="the total number of conjunction steps is only 480"):
        rng.get_bvh_tree(480)
    with pytest.raises(IndexError, match="the total number of conjunction steps is only 480"):
        rng.get_aabb_collisions(480)
    rng.hint_release()
    assert np.array_equal(_bits(rng.conjunctions), _bits(one.conjunctions))
    if mode == "fp64" and chunk == 7:
        from oracle import pyoracle as po

        ref = po.run(H.flat_of(leo1000), 5.0, 3.0 / 1440.0)
        assert np.array_equal(_bits(rng.conjunctions), _bits(ref["conjunctions"]))
        assert rng.np_report == ref["np_report"]


def test_ranges_with_object_types_and_files(tmp_path, monkeypatch):
    """Object types and the result files in a data dir: the files of a
    screening in ranges are byte-equal to those of one batch."""
    pj = mz.make_synthetic_polyjectory(1500, 0.3, seed=17, frac_high=0.2)
    ot = np.random.default_rng(5).choice([1, 2, 4], size=pj.n_objs, p=[0.6, 0.3, 0.1]).astype(np.int32)
    one = mz.conjunctions(pj, 20.0, 2.0 / 1440.0, otypes=ot, data_dir=tmp_path / "one")
    monkeypatch.setenv("MIZUBA_AMD_CHUNK_STEPS", "13")
    rng = mz.conjunctions(pj, 20.0, 2.0 / 1440.0, otypes=ot, data_dir=tmp_path / "rng")
    assert rng.chunk_steps == 13
    assert_same_results(rng, one)
    for name in ("aabbs", "srt_aabbs", "mcodes", "srt_mcodes", "vidx", "bvh", "bp", "conjunctions"):
        assert (tmp_path / "one" / name).read_bytes() == (tmp_path / "rng" / name).read_bytes(), name


def test_screen_steps_ranges(leo1000):
    """mzh_screen_steps on a Screener: each range's conjunctions are the
    one-batch list's rows whose step falls in the range; ranges outside the
    screening are rejected."""
    scr = mz.Screener(0)
    scr.upload(leo1000)
    scr.screen(5.0, 3.0 / 1440.0)
    whole = scr.conjunctions()
    C = scr.n_cd_steps()
    ends = np.minimum(3.0 / 1440.0 * np.arange(1, C + 1), leo1000.maxT)
    parts = []
    for k0, k1 in ((0, 100), (100, 101), (101, 333), (333, 480)):
        scr.screen(5.0, 3.0 / 1440.0, steps=(k0, k1))
        assert scr.n_cd_steps() == k1 - k0
        parts.append(scr.conjunctions())
    assert np.array_equal(_bits(np.concatenate(parts)), _bits(whole))
    # the first range's rows: conjunctions before the end of step 99
    assert np.all(parts[0]["tca"] <= ends[99])
    with pytest.raises(IndexError, match="Invalid range of conjunction steps \\[10, 481\\)"):
        scr.screen(5.0, 3.0 / 1440.0, steps=(10, 481))
    with pytest.raises(IndexError, match="Invalid range of conjunction steps \\[5, 5\\)"):
        scr.screen(5.0, 3.0 / 1440.0, steps=(5, 5))
