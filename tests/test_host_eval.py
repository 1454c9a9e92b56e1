"""Argument checks of polyjectory.state_eval / state_meval, with the
reference's messages (test/test_polyjectory.py:850-1085). All of them are
raised on the host before any device work, so they run without a GPU."""

import numpy as np
import pytest

import mizuba_amd as mz

from . import helpers as H


@pytest.fixture(scope="module")
def pj():
    trajs, times = H.eval_pj_arrays()
    return mz.polyjectory(trajs, times, [0, 0])


def raises(exc, text, f):
    with pytest.raises(exc) as ei:
        f()
    assert text in str(ei.value), str(ei.value)


def test_state_eval_argument_errors(pj):
    V = ValueError
    raises(V, "The time array passed to state_eval() must be C contiguous and properly aligned",
           lambda: pj.state_eval(np.zeros((5,))[::2]))
    raises(V, "The time array passed to state_eval() must have 1 dimension, but the number of dimensions is 2 instead",
           lambda: pj.state_eval(np.zeros((5, 5))))
    raises(V, "Invalid time array passed to state_eval(): the number of objects is 2 but the size of the array is 5 "
              "(the two numbers must be equal)", lambda: pj.state_eval(time=np.zeros((5,))))
    raises(V, "Invalid time array passed to state_eval(): the number of selected objects is 1 but the size of the "
              "array is 2 (the two numbers must be equal)",
           lambda: pj.state_eval(time=np.array([0.11, 0.11]), obj_idx=1))
    raises(V, "The selector array passed to state_eval() must be C contiguous and properly aligned",
           lambda: pj.state_eval(time=0.11, obj_idx=np.array([0, 0, 0, 0], dtype=np.uintp)[::2]))
    raises(V, "The selector array passed to state_eval() must have 1 dimension, but the number of dimensions is 2 "
              "instead", lambda: pj.state_eval(time=0.11, obj_idx=np.zeros((2, 2))))
    raises(V, "The output array passed to state_eval() must be C contiguous and properly aligned",
           lambda: pj.state_eval(time=0.11, out=np.array([1.0, 2.0, 3.0])[::2]))
    raises(V, "The output array passed to state_eval() must have 2 dimensions, but the number of dimensions is 1 "
              "instead", lambda: pj.state_eval(time=0.11, out=np.full((1,), 1.0)))
    raises(V, "The output array passed to state_eval() must have a size of 7 in the second dimension, but the size "
              "in the second dimension is 6 instead", lambda: pj.state_eval(time=0.11, out=np.full((1, 6), 1.0)))
    msg_n = ("Invalid output array passed to state_eval(): the number of objects is 2 but the size of the first "
             "dimension of the array is 3 (the two numbers must be equal)")
    msg_s = ("Invalid output array passed to state_eval(): the number of objects selected for evaluation is 1 but "
             "the size of the first dimension of the array is 3 (the two numbers must be equal)")
    raises(V, msg_n, lambda: pj.state_eval(time=0.11, out=np.full((3, 7), 1.0)))
    raises(V, msg_s, lambda: pj.state_eval(time=0.11, out=np.full((3, 7), 1.0), obj_idx=[1]))
    raises(V, msg_n, lambda: pj.state_eval(time=[0.11, 0.11], out=np.full((3, 7), 1.0)))
    raises(V, msg_s, lambda: pj.state_eval(time=[0.11, 0.11], out=np.full((3, 7), 1.0), obj_idx=[1]))
    out = np.zeros((2, 7))
    raises(V, "Potential memory overlap detected between the output array passed to state_eval() and the time array",
           lambda: pj.state_eval(time=out[:2], out=out))
    raises(V, "Potential memory overlap detected between the output array passed to state_eval() and the array of "
              "object indices", lambda: pj.state_eval(time=[1.1, 1.1], out=out, obj_idx=out[0, :2].view(np.uintp)))
    # the output array is not converted
    with pytest.raises(TypeError):
        pj.state_eval(time=0.11, out=np.zeros((2, 7), dtype=np.float32))


def test_state_meval_argument_errors(pj):
    V = ValueError
    raises(V, "The time array passed to state_meval() must be C contiguous and properly aligned",
           lambda: pj.state_meval(time=np.array([0.11, 0.11, 0.11, 0.11])[::2]))
    raises(V, "The time array passed to state_meval() must have either 1 or 2 dimensions, but the number of "
              "dimensions is 3 instead", lambda: pj.state_meval(time=np.zeros((3, 3, 3))))
    raises(V, "The output array passed to state_meval() must be C contiguous and properly aligned",
           lambda: pj.state_meval(time=[0.11, 0.11], out=np.array([1.0, 2.0, 0.3, 0.4])[::2]))
    raises(V, "The output array passed to state_meval() must have 3 dimensions, but the number of dimensions is 1 "
              "instead", lambda: pj.state_meval(time=[0.11, 0.11], out=np.array([1.0, 2.0, 0.3, 0.4])))
    raises(V, "The output array passed to state_meval() must have a size of 7 in the third dimension, but the size "
              "in the third dimension is 5 instead", lambda: pj.state_meval(time=[0.11, 0.11], out=np.zeros((3, 4, 5))))
    raises(V, "Invalid output array passed to state_meval(): the number of time evaluations per object is 2 but the "
              "size of the second dimension of the array is 4 (the two numbers must be equal)",
           lambda: pj.state_meval(time=[0.11, 0.11], out=np.zeros((3, 4, 7))))
    msg_s = ("Invalid output array passed to state_meval(): the number of selected objects is 1 but the size of the "
             "first dimension of the array is 2 (the two numbers must be equal)")
    raises(V, msg_s, lambda: pj.state_meval(time=[0.11, 0.11], out=np.zeros((2, 2, 7)), obj_idx=[1]))
    raises(V, msg_s, lambda: pj.state_meval(time=[[0.11, 0.11]], out=np.zeros((2, 2, 7)), obj_idx=[1]))
    raises(V, "Invalid time array passed to state_meval(): the number of selected objects is 1 but the size of the "
              "first dimension of the array is 2 (the two numbers must be equal)",
           lambda: pj.state_meval(time=[[0.11, 0.11], [0.11, 0.11]], out=np.zeros((1, 2, 7)), obj_idx=[1]))
    msg_n = ("Invalid output array passed to state_meval(): the number of objects is 2 but the size of the first "
             "dimension of the array is 3 (the two numbers must be equal)")
    raises(V, msg_n, lambda: pj.state_meval(time=[0.11, 0.11], out=np.zeros((3, 2, 7))))
    raises(V, msg_n, lambda: pj.state_meval(time=[[0.11, 0.11]], out=np.zeros((3, 2, 7))))
    raises(V, "Invalid time array passed to state_meval(): the number of objects is 2 but the size of the first "
              "dimension of the array is 1 (the two numbers must be equal)",
           lambda: pj.state_meval(time=[[0.11, 0.11]], out=np.zeros((2, 2, 7))))
    out = np.zeros((2, 2, 7))
    raises(V, "Potential memory overlap detected between the output array passed to state_meval() and the time array",
           lambda: pj.state_meval(time=out[0, 0, :2], out=out))
    raises(V, "Potential memory overlap detected between the output array passed to state_meval() and the array of "
              "object indices",
           lambda: pj.state_meval(time=[1.1, 1.1], out=out, obj_idx=out[0, 0, :2].view(np.uintp)))


def test_empty_selection_needs_no_device(pj):
    # nothing to evaluate: no device work, empty results of the right shape
    assert pj.state_eval(time=0.11, obj_idx=np.zeros(0, dtype=np.uintp)).shape == (0, 7)
    assert pj.state_meval(time=np.zeros(0)).shape == (2, 0, 7)
