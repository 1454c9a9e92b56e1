"""Polyjectory data dirs: construction with data_dir / tmpdir / persist,
from_datafiles, mount, detach (test/test_polyjectory.py:350-720, 1300-1560),
and the byte layout of the six files (polyjectory.cpp:76-236). Host only."""

import os
import tempfile
from pathlib import Path

import numpy as np
import pytest

import mizuba_amd as mz
from mizuba_amd import polyjectory

DESC = np.dtype({"names": ["version", "persist", "n_objs", "poly_op1", "maxT", "epoch", "epoch2"],
                 "formats": ["<u4", "u1", "<u8", "<u4", "<f8", "<f8", "<f8"],
                 "offsets": [0, 4, 8, 16, 24, 32, 40], "itemsize": 48})


def raises(exc, text, f):
    with pytest.raises(exc) as ei:
        f()
    assert text in str(ei.value), str(ei.value)


def two_obj(state_data=None, **kw):
    sd = np.zeros((1, 8, 7)) if state_data is None else state_data
    return polyjectory(trajs=[sd, sd[1:]], times=[np.array([0.0, 1.0]), np.array([], dtype=float)],
                       status=np.array([0, 0], dtype=np.int32), **kw)


def write_files(d, traj, time):
    np.asarray(traj, dtype=float).tofile(Path(d) / "traj")
    np.asarray(time, dtype=float).tofile(Path(d) / "time")
    return str(Path(d) / "traj"), str(Path(d) / "time")


def test_constructor_contiguity_and_epoch():
    sd = np.zeros((1, 8, 7))
    raises(ValueError, "All trajectory arrays must be C contiguous and properly aligned",
           lambda: polyjectory([sd[:, ::2, :]], [[1.0]], [0]))
    raises(ValueError, "All time arrays must be C contiguous and properly aligned",
           lambda: polyjectory([sd, sd], [np.array([1.0, 2.0, 3.0, 4.0])[::2]], [0]))
    raises(ValueError, "The status array must be C contiguous and properly aligned",
           lambda: polyjectory([sd, sd], [np.array([1.0, 2.0])], np.array([0, 0, 0, 0], dtype=np.int32)[::2]))
    # the double-length epoch is normalised (test/test_polyjectory.py:309-321)
    pj = polyjectory(trajs=[sd, sd], times=[np.array([0.0, 1.0]), np.array([0.0, 3.0])],
                     status=np.array([0, 1], dtype=np.int32), epoch=42.0, epoch2=1.0)
    assert pj.epoch == (43.0, 0.0)
    assert pj.n_objs == 2 and pj.maxT == 3 and pj.poly_order == 7 and not pj.persist


def test_from_datafiles_argument_errors():
    TO = polyjectory.traj_offset
    with tempfile.NamedTemporaryFile() as tf, tempfile.NamedTemporaryFile() as mf:
        for order in (0, 1):
            raises(ValueError, f"Invalid polynomial order {order} specified during the construction of a polyjectory",
                   lambda: polyjectory.from_datafiles(traj_file=tf.name, time_file=mf.name, order=order,
                                                      traj_offsets=[], status=[]))
        base = dict(traj_file=tf.name, time_file=mf.name, order=2)
        raises(ValueError, "Invalid trajectory offsets vector passed to the constructor of a polyjectory: the vector "
                           "cannot be empty", lambda: polyjectory.from_datafiles(**base, traj_offsets=[], status=[]))
        raises(ValueError, "The array of trajectory offsets passed to 'from_datafiles()' must have 1 dimension, but "
                           "the number of dimensions is 2 instead",
               lambda: polyjectory.from_datafiles(**base, traj_offsets=np.zeros((1, 1), dtype=TO), status=[]))
        raises(ValueError, "Invalid status vector passed to the constructor of a polyjectory: the expected size is 1, "
                           "but the actual size is 0",
               lambda: polyjectory.from_datafiles(**base, traj_offsets=np.zeros((1,), dtype=TO), status=[]))
        raises(ValueError, "The initial epoch of a polyjectory must be finite, but instead a value of",
               lambda: polyjectory.from_datafiles(**base, traj_offsets=np.zeros((1,), dtype=TO), status=[1],
                                                  epoch=float("inf")))
        raises(ValueError, "The second component of the initial epoch of a polyjectory must be finite, but instead a "
                           "value of",
               lambda: polyjectory.from_datafiles(**base, traj_offsets=np.zeros((1,), dtype=TO), status=[1],
                                                  epoch=1.0, epoch2=float("inf")))
        raises(ValueError, "Invalid data file(s) passed to the constructor of a polyjectory: the trajectory data file "
                           "and the time data file are the same file",
               lambda: polyjectory.from_datafiles(traj_file=tf.name, time_file=tf.name, order=2,
                                                  traj_offsets=np.zeros((1,), dtype=TO), status=[1]))
        raises(ValueError, "Invalid trajectory offsets vector passed to the constructor of a polyjectory: the initial "
                           "offset value must be zero but it is 1 instead",
               lambda: polyjectory.from_datafiles(**base, traj_offsets=np.array([(1, 1)], dtype=TO), status=[1]))
        raises(ValueError, "Invalid trajectory offsets vector passed to the constructor of a polyjectory: the offset "
                           "of the object at index 2 is less than the offset of the previous object",
               lambda: polyjectory.from_datafiles(**base, traj_offsets=np.array([(0, 1), (21, 1), (0, 2)], dtype=TO),
                                                  status=[1, 3, 4]))
        raises(ValueError, "Inconsistent data detected in the trajectory offsets vector passed to the constructor of "
                           "a polyjectory: the offset of the object at index 1 is inconsistent with the offset and "
                           "number of steps of the previous object",
               lambda: polyjectory.from_datafiles(**base, traj_offsets=np.array([(0, 1), (2, 1)], dtype=TO),
                                                  status=[1, 3]))
        raises(ValueError, "All the trajectories in a polyjectory have a number of steps equal to zero: this is not "
                           "allowed",
               lambda: polyjectory.from_datafiles(**base, traj_offsets=np.zeros((1,), dtype=TO), status=[1]))
        # none of these moved the files
        assert os.path.exists(tf.name) and os.path.exists(mf.name)


def test_from_datafiles_file_errors():
    TO = polyjectory.traj_offset
    one = np.array([(0, 1)], dtype=TO)
    with tempfile.TemporaryDirectory() as d1, tempfile.TemporaryDirectory() as d2:
        sub = Path(d2) / "sub"
        sub.mkdir()
        (Path(d1) / "time").write_bytes(b"")
        raises(ValueError, "Invalid trajectory data file passed to the constructor of a polyjectory: the file is not a "
                           "regular file",
               lambda: polyjectory.from_datafiles(traj_file=sub, time_file=Path(d1) / "time", order=2,
                                                  traj_offsets=one, status=[1]))
    with tempfile.TemporaryDirectory() as d1, tempfile.TemporaryDirectory() as d2:
        sub = Path(d2) / "sub"
        sub.mkdir()
        (Path(d1) / "traj").write_bytes(b"")
        raises(ValueError, "Invalid time data file passed to the constructor of a polyjectory: the file is not a "
                           "regular file",
               lambda: polyjectory.from_datafiles(traj_file=Path(d1) / "traj", time_file=sub, order=2,
                                                  traj_offsets=one, status=[1]))
    cases = [
        (np.zeros(20), np.zeros(2), one,
         "Invalid trajectory data file passed to the constructor of a polyjectory: the expected size in bytes is"),
        (np.zeros(21), np.zeros(1), one,
         "Invalid time data file passed to the constructor of a polyjectory: the expected size in bytes is"),
        (np.full(21, np.inf), np.zeros(2), one, "A non-finite value was found in the trajectory at index 0"),
        (np.zeros(21), np.full(2, np.nan), one, "A non-finite time coordinate was found for the object at index 0"),
        (np.zeros(21), np.full(2, -1.0), one, "A negative time coordinate was found for the object at index 0"),
        (np.zeros(42), [0.0, 1.0, 1.0], np.array([(0, 2)], dtype=TO),
         "The sequence of times for the object at index 0 is not monotonically increasing"),
        (np.zeros(42), [0.0, 1.0, 0.9], np.array([(0, 2)], dtype=TO),
         "The sequence of times for the object at index 0 is not monotonically increasing"),
    ]
    for traj, time, offs, msg in cases:
        with tempfile.TemporaryDirectory() as d:
            tf, mf = write_files(d, traj, time)
            before = set(os.listdir(tempfile.gettempdir()))
            raises(ValueError, msg, lambda: polyjectory.from_datafiles(traj_file=tf, time_file=mf, order=2,
                                                                       traj_offsets=offs, status=[1], tmpdir=d))
            # the data dir made for the attempt is gone again
            assert [x for x in os.listdir(d) if x.startswith("mizuba_polyjectory-")] == []


def test_from_datafiles_success_and_layout():
    TO = polyjectory.traj_offset
    rng = np.random.default_rng(1)
    # three objects of order 2: 2 steps, none, 3 steps
    t0 = np.array([0.0, 1.0, 1.5])
    t2 = np.array([0.5, 0.75, 2.0, 2.25])
    tr0 = rng.standard_normal((2, 3, 7))
    tr2 = rng.standard_normal((3, 3, 7))
    offs = np.array([(0, 2), (42, 0), (42, 3)], dtype=TO)
    with tempfile.TemporaryDirectory() as d:
        tf, mf = write_files(d, np.concatenate([tr0.ravel(), tr2.ravel()]), np.concatenate([t0, t2]))
        dd = (Path(d) / "data_dir").resolve()
        pj = polyjectory.from_datafiles(traj_file=tf, time_file=mf, order=2, traj_offsets=offs, status=[5, 6, 7],
                                        epoch=10.0, epoch2=0.5, data_dir=dd)
        # the files were moved in
        assert not os.path.exists(tf) and not os.path.exists(mf)
        assert pj.data_dir == dd and not pj.persist
        assert pj.n_objs == 3 and pj.poly_order == 2 and pj.maxT == 2.25 and pj.epoch == (10.5, 0.0)
        assert np.array_equal(pj[0][0], tr0) and np.array_equal(pj[0][1], t0)
        assert pj[1][0].shape == (0, 3, 7) and pj[1][1].shape == (0,)
        assert np.array_equal(pj[2][0], tr2) and np.array_equal(pj[2][1], t2)
        assert list(pj.status) == [5, 6, 7]
        assert sorted(os.listdir(dd)) == ["desc", "status", "time", "time_offsets", "traj", "traj_offsets"]
        desc = np.fromfile(dd / "desc", dtype=DESC)
        assert len(desc) == 1 and (dd / "desc").stat().st_size == 48
        assert (desc["version"][0], desc["persist"][0], desc["n_objs"][0], desc["poly_op1"][0]) == (0, 0, 3, 3)
        assert (desc["maxT"][0], desc["epoch"][0], desc["epoch2"][0]) == (2.25, 10.5, 0.0)
        assert np.array_equal(np.fromfile(dd / "traj_offsets", dtype=TO), offs)
        assert np.array_equal(np.fromfile(dd / "time_offsets", dtype="<u8"), [0, 3, 3])
        assert np.array_equal(np.fromfile(dd / "status", dtype="<i4"), [5, 6, 7])
        # files are read-only, the dir is the owner's only
        assert (dd / "traj").stat().st_mode & 0o777 == 0o400
        assert dd.stat().st_mode & 0o777 == 0o700
        pj.detach()
        assert not dd.exists()


def test_custom_data_dir_and_persist():
    sd = np.zeros((1, 8, 7))
    with tempfile.TemporaryDirectory() as d:
        dd = Path(d) / "data_dir"
        os.mkdir(dd)
        raises(RuntimeError, "Error while creating the directory", lambda: two_obj(data_dir=dd))
    with tempfile.TemporaryDirectory() as d:
        dd = (Path(d) / "data_dir").resolve()
        pj = two_obj(data_dir=dd)
        assert pj.data_dir == dd
        pj.detach()
        assert not dd.exists()
    # an empty data dir path counts as none
    two_obj(data_dir="")
    with tempfile.TemporaryDirectory() as d:
        dd = (Path(d) / "data_dir").resolve()
        pj = two_obj(persist=True, data_dir=dd)
        assert dd.is_dir() and pj.persist
        pj.detach()
        assert dd.is_dir()
    # a temporary dir inside tmpdir
    with tempfile.TemporaryDirectory() as d:
        pj = two_obj(tmpdir=d)
        assert pj.data_dir.parent == Path(d).resolve() and pj.data_dir.name.startswith("mizuba_polyjectory-")
        dd = pj.data_dir
        del pj
        assert not dd.exists()
    raises(ValueError, "The 'data_dir' and 'tmpdir' construction arguments cannot be provided at the same time",
           lambda: two_obj(data_dir="a", tmpdir="b"))


def test_detached_throw():
    pj = two_obj()
    pj.detach()
    pj.detach()
    assert pj.is_detached
    for attr, args in (("n_objs", None), ("maxT", None), ("epoch", None), ("poly_order", None), ("data_dir", None),
                       ("__getitem__", (0,)), ("status", None), ("persist", None), ("state_eval", (0.0,)),
                       ("state_meval", ([0.0],))):
        with pytest.raises(ValueError) as ei:
            a = getattr(pj, attr)
            if args is not None:
                a(*args)
        assert "Cannot operate on a detached polyjectory" in str(ei.value)


def test_mount():
    sd = np.arange(56.0).reshape(1, 8, 7)
    with tempfile.TemporaryDirectory() as d:
        tmp = two_obj(sd, tmpdir=d)
        raises(ValueError, ": the data is not persistent", lambda: polyjectory.mount(tmp.data_dir))
        raises(ValueError, "could not be canonicalised (does it exist?)",
               lambda: polyjectory.mount(Path(d) / "foobar"))
        (Path(d) / "afile").write_bytes(b"")
        raises(ValueError, "the path is not a directory", lambda: polyjectory.mount(Path(d) / "afile"))
        dd = Path(d) / "foo"
        pj = two_obj(sd, data_dir=dd, persist=True, epoch=1.5)
        pj.detach()
        assert dd.is_dir()
        pj = polyjectory.mount(dd)
        assert np.array_equal(pj[0][0], sd) and np.array_equal(pj[0][1], [0.0, 1.0])
        assert np.array_equal(pj[1][0], sd[1:]) and pj[1][1].shape == (0,)
        assert pj.persist and pj.epoch == (1.5, 0.0) and pj.maxT == 1.0 and pj.poly_order == 7
        pj.detach()
        assert dd.is_dir()  # mounted data is never removed


def test_mount_synthetic_round_trip():
    pj = mz.make_synthetic_polyjectory(200, 0.5, seed=3)
    flat = pj._flat()
    trajs = [np.array(pj[i][0]) for i in range(pj.n_objs)]
    times = [np.array(pj[i][1]) for i in range(pj.n_objs)]
    with tempfile.TemporaryDirectory() as d:
        dd = Path(d) / "pj"
        pj2 = polyjectory(trajs, times, list(pj.status), data_dir=dd, persist=True)
        m = polyjectory.mount(dd)
        for a, b in zip(flat, m._flat()):
            assert np.array_equal(np.asarray(a), np.asarray(b))
        assert m.maxT == pj.maxT and m.n_objs == pj.n_objs


def test_global_tmpdir():
    orig = mz.get_tmpdir()
    try:
        with tempfile.TemporaryDirectory() as d:
            mz.set_tmpdir(d)
            assert mz.get_tmpdir() == Path(d)
            pj = two_obj(persist=False, tmpdir=None, data_dir=None)
            assert pj.data_dir is None  # in memory: no files asked for
            pj = polyjectory(trajs=[np.zeros((1, 8, 7))], times=[np.array([0.0, 1.0])], status=[0], persist=True)
            assert pj.data_dir.parent == Path(d).resolve()
            mz.set_tmpdir(None)
            assert mz.get_tmpdir() is None
            # a tmpdir argument wins over the global one
            mz.set_tmpdir(__file__)
            with tempfile.TemporaryDirectory() as d2:
                pj2 = two_obj(tmpdir=d2)
                assert pj2.data_dir.parent == Path(d2).resolve()
                del pj2
    finally:
        mz.set_tmpdir(orig)


def test_conjunctions_data_dir_errors_before_screening():
    trajs, times = [np.zeros((1, 3, 7))], [np.array([0.0, 1.0])]
    pj = polyjectory(trajs, times, [0])
    with tempfile.TemporaryDirectory() as d:
        dd = Path(d) / "data_dir"
        os.mkdir(dd)
        raises(RuntimeError, "Error while creating the directory",
               lambda: mz.conjunctions(pj, conj_thresh=0.1, conj_det_interval=0.01, data_dir=dd))
    raises(ValueError, "The 'data_dir' and 'tmpdir' construction arguments cannot be provided at the same time",
           lambda: mz.conjunctions(pj, conj_thresh=0.1, conj_det_interval=0.01, tmpdir="", data_dir=""))
