"""Extract the planar circular-orbit fixture from the reference test suite.

The reference holds a heyoka-made polyjectory of one planar circular Keplerian
orbit (order 20, 7 steps) as numpy literals in mizuba/test/_planar_circ.py
(arrays at :23 and :1365). This script reads that file as TEXT, pulls the two
array literals out with `ast.literal_eval` (nothing is imported or executed)
and stores them as tests/golden/planar_circ.npz. Run it only where the
reference tree is present; the .npz is the committed fixture.
"""

import ast
import pathlib
import sys

import numpy as np

SRC = pathlib.Path("/root/reference/mizuba/test/_planar_circ.py")
OUT = pathlib.Path(__file__).resolve().parent / "planar_circ.npz"


def _literal(node):
    # _np.array([...]) -> the list literal
    assert isinstance(node, ast.Call) and node.args
    return ast.literal_eval(node.args[0])


def main():
    tree = ast.parse(SRC.read_text())
    found = {}
    for stmt in tree.body:
        if isinstance(stmt, ast.Assign) and len(stmt.targets) == 1 and isinstance(stmt.targets[0], ast.Name):
            name = stmt.targets[0].id
            if name in ("_planar_circ_tcs", "_planar_circ_times"):
                found[name] = np.array(_literal(stmt.value), dtype=np.float64)
    if set(found) != {"_planar_circ_tcs", "_planar_circ_times"}:
        sys.exit("fixture arrays not found")
    np.savez(OUT, tcs=found["_planar_circ_tcs"], times=found["_planar_circ_times"])
    print(OUT, found["_planar_circ_tcs"].shape, found["_planar_circ_times"].shape)


if __name__ == "__main__":
    main()
