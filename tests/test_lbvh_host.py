"""CPU check of the broad phase's walk tree (mizuba_amd/csrc/mzh_lbvh.hpp,
built for the host): Karras' construction over (code, position) keys with runs
of equal codes as leaves must give the reference's radix tree over runs
(detect_conjunctions_bvh.cpp:267-699) node for node in depth-first order, and
every node's escape index must name the node that follows its subtree. The
trees are compared with a recursive build from the definition (split where the
highest differing bit of the range flips) on adversarial code arrays, and with
the oracle's trees of a synthetic screening. The walk tree's grid boxes
(mizuba_amd/csrc/mzh_qbox.hpp) are checked against the float overlap test:
boxes that meet must meet on the grid, and the grid box of a union must be the
union of the grid boxes."""

import ctypes
import pathlib
import shutil
import subprocess

import numpy as np
import pytest

from . import helpers as H

ROOT = pathlib.Path(__file__).resolve().parent.parent
U32 = ctypes.POINTER(ctypes.c_uint32)


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not available")
    out = tmp_path_factory.mktemp("lbvh") / "liblbvh.so"
    subprocess.run(
        [gxx, "-O2", "-std=c++17", "-fPIC", "-shared", str(ROOT / "tests" / "lbvh_check.cpp"), "-o", str(out)],
        check=True,
    )
    L = ctypes.CDLL(str(out))
    L.lbvh_walk.restype = ctypes.c_int64
    L.lbvh_block_check.restype = ctypes.c_int64
    L.qbox_check.restype = ctypes.c_int64
    return L


def _walk(lib, codes):
    codes = np.ascontiguousarray(codes, dtype=np.uint64)
    n = len(codes)
    b, e, s = (np.zeros(max(2 * n, 1), np.uint32) for _ in range(3))
    m = lib.lbvh_walk(ctypes.c_uint32(n), codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                      b.ctypes.data_as(U32), e.ctypes.data_as(U32), s.ctypes.data_as(U32))
    assert m >= 0, "walk left the node array or did not end"
    return list(zip(b[:m].tolist(), e[:m].tolist(), s[:m].tolist()))


def _nodes(lib, codes):
    codes = np.ascontiguousarray(codes, dtype=np.uint64)
    n = len(codes)
    k = max(n - 1, 1)
    lo, hi, esc, sp, lc, rc = (np.zeros(k, np.uint32) for _ in range(6))
    run = np.zeros(k, np.uint8)
    lib.lbvh_nodes(ctypes.c_uint32(n), codes.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)), lo.ctypes.data_as(U32),
                   hi.ctypes.data_as(U32), esc.ctypes.data_as(U32), sp.ctypes.data_as(U32),
                   run.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), lc.ctypes.data_as(U32), rc.ctypes.data_as(U32))
    return tuple(a[: n - 1] for a in (lo, hi, esc, sp, run, lc, rc))


def _definition(codes):
    """Depth-first list of (begin, end, split) of the radix tree over runs,
    split 0 for a leaf, built recursively from its definition."""
    codes = [int(c) for c in codes]
    out = []
    stack = [(0, len(codes))] if codes else []
    while stack:
        b, e = stack.pop()
        x = codes[b] ^ codes[e - 1]
        if x == 0:
            out.append((b, e, 0))
            continue
        bit = x.bit_length() - 1
        s = next(p for p in range(b, e) if (codes[p] >> bit) & 1)
        out.append((b, e, s))
        stack.append((s, e))
        stack.append((b, s))
    return out


def _cases(rng):
    yield np.array([], np.uint64)
    yield np.array([5], np.uint64)
    yield np.array([7, 7], np.uint64)
    yield np.array([1, 2], np.uint64)
    yield np.full(37, 2**64 - 1, np.uint64)
    yield np.array([0, 2**64 - 2, 2**64 - 1, 2**64 - 1], np.uint64)
    for n in (3, 4, 5, 17, 64, 65, 300, 1500):
        for alphabet in (2, 3, 8, n, None):
            if alphabet is None:
                c = rng.integers(0, 2**63, n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, n, dtype=np.uint64)
            else:
                c = rng.choice(rng.integers(0, 2**64 - 1, alphabet, dtype=np.uint64, endpoint=True), n)
            yield np.sort(c)
        # long runs at both ends and in the middle
        c = np.sort(rng.integers(0, 1 << 20, n, dtype=np.uint64))
        c[: n // 3] = c[0]
        c[-(n // 4):] = c[-1]
        yield np.sort(c)


def _check(lib, codes):
    ref = _definition(codes)
    got = _walk(lib, codes)
    assert got == ref
    n = len(codes)
    if n < 2:
        return
    # the build's blocks: nodes that stay inside a block found from its codes
    # (and a halo) alone, the others recognised as leaving it
    arr = np.ascontiguousarray(codes, dtype=np.uint64)
    for bs in (1, 2, 3, 64, 512, 1024):
        assert lib.lbvh_block_check(ctypes.c_uint32(n), arr.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                    ctypes.c_uint32(bs)) == 0, bs
    # escape of every internal node: the node after its subtree
    pos = {(b, e): i for i, (b, e, _) in enumerate(ref)}
    size = [0] * len(ref)
    for i in range(len(ref) - 1, -1, -1):
        b, e, s = ref[i]
        size[i] = 1 if s == 0 else 1 + size[i + 1] + size[i + 1 + size[i + 1]]
    lo, hi, esc, sp, run, lc, rc = _nodes(lib, codes)

    def node_range(idx):
        if idx & 1:
            p = idx >> 1
            return (p, p + 1)
        k = idx >> 1
        return (int(lo[k]), int(hi[k]) + 1)

    for k in range(n - 1):
        rg = (int(lo[k]), int(hi[k]) + 1)
        if rg not in pos:
            assert run[k]  # inside a run's leaf: never reached by the walk
            continue
        i = pos[rg]
        nxt = i + size[i]
        if nxt == len(ref):
            assert esc[k] == 0xFFFFFFFF
        else:
            # the escape names the node after the subtree (a leaf inside a
            # run is never an escape target: runs are reached at their root)
            assert node_range(int(esc[k])) == ref[nxt][:2], (k, rg)
        assert bool(run[k]) == (ref[i][2] == 0)
        if not run[k]:
            # the split and both children (the walk goes to either directly)
            assert sp[k] == ref[i][2]
            assert node_range(int(lc[k])) == ref[i + 1][:2]
            assert node_range(int(rc[k])) == ref[i + 1 + size[i + 1]][:2]


def test_adversarial_codes(lib):
    rng = np.random.default_rng(11)
    for codes in _cases(rng):
        _check(lib, codes)


def test_oracle_trees(lib, oracle):
    import mizuba_amd as mz

    pj = mz.make_synthetic_polyjectory(400, 0.08, seed=9, frac_high=0.2)
    out = oracle.run(H.flat_of(pj), 50.0, 3.0 / 1440.0)
    N = pj.n_objs
    for k in range(out["n_cd_steps"]):
        n_eff = int(np.sum(np.isfinite(out["srt_aabbs"][k, :N, 0, 0])))
        codes = out["srt_mcodes"][k][:n_eff]
        tree = oracle.get_bvh_tree(out, k)
        # the oracle's breadth-first tree in depth-first order
        dfs, stack = [], [0] if len(tree) else []
        while stack:
            i = stack.pop()
            nd = tree[i]
            leaf = nd["left"] == -1
            dfs.append((int(nd["begin"]), int(nd["end"]), 0 if leaf else int(tree[nd["right"]]["begin"])))
            if not leaf:
                stack.append(int(nd["right"]))
                stack.append(int(nd["left"]))
        assert _walk(lib, codes) == dfs
        _check(lib, codes)


def _qbox_check(lib, boxes, g):
    boxes = np.ascontiguousarray(boxes, dtype=np.float32).reshape(-1, 8)
    g = np.ascontiguousarray(g, dtype=np.float32)
    mf, mg = ctypes.c_int64(0), ctypes.c_int64(0)
    F32 = ctypes.POINTER(ctypes.c_float)
    bad = lib.qbox_check(ctypes.c_uint32(len(boxes)), boxes.ctypes.data_as(F32), g.ctypes.data_as(F32),
                         ctypes.byref(mf), ctypes.byref(mg))
    return bad, mf.value, mg.value


def _global(boxes):
    b = np.asarray(boxes, np.float32).reshape(-1, 8)
    return np.concatenate([b[:, :4].min(axis=0), b[:, 4:].max(axis=0)])


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_grid_boxes_keep_every_overlap(lib, seed):
    """Random boxes of the sizes a screening makes (LEO positions over a
    3-minute step, km) and boxes sharing faces, edges and corners exactly: no
    pair that meets is lost on the grid, and grid unions are exact."""
    rng = np.random.default_rng(seed)
    n = 300
    c = rng.uniform(-8000.0, 8000.0, (n, 4)).astype(np.float32)
    c[:, 3] = rng.uniform(6700.0, 8400.0, n)
    h = rng.uniform(1.0, 900.0, (n, 4)).astype(np.float32)
    boxes = np.concatenate([c - h, c + h], axis=1).astype(np.float32)
    # boxes that touch: the low face of one at the high face of another
    for k in range(0, 60, 2):
        boxes[k + 1, :4] = boxes[k, 4:]
        boxes[k + 1, 4:] = boxes[k, 4:] + np.float32(1.0)
    # boxes on the global box's faces, and points
    boxes[100, :] = np.concatenate([boxes[:, :4].min(axis=0), boxes[:, :4].min(axis=0)])
    boxes[101, :] = np.concatenate([boxes[:, 4:].max(axis=0), boxes[:, 4:].max(axis=0)])
    boxes[102, 4:] = boxes[102, :4]
    bad, mf, mg = _qbox_check(lib, boxes, _global(boxes))
    assert bad == 0
    assert mf > n  # the pairs include overlaps beyond the diagonal
    assert mg < 0.05 * n * n  # the grid adds few pairs


def test_grid_boxes_degenerate_extents(lib):
    """A global box of zero width in some dimensions (all objects at one
    radius, or one object), and extents near the float limits."""
    rng = np.random.default_rng(5)
    boxes = rng.uniform(-1.0, 1.0, (50, 8)).astype(np.float32)
    boxes[:, 4:] = boxes[:, :4] + np.float32(0.5)
    boxes[:, 3] = boxes[:, 7] = 7000.0
    assert _qbox_check(lib, boxes, _global(boxes))[0] == 0
    one = np.array([[1.0, 2.0, 3.0, 4.0, 1.0, 2.0, 3.0, 4.0]], np.float32)
    assert _qbox_check(lib, one, _global(one))[0] == 0
    big = rng.uniform(-3e38, 3e38, (40, 4)).astype(np.float32)
    bb = np.concatenate([big, big + np.abs(big) * np.float32(1e-3)], axis=1)
    assert _qbox_check(lib, bb, _global(bb))[0] == 0
