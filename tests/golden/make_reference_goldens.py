"""Write tests/golden/reference_goldens.json from the reference's own tests.

Run once, in the build container (the reference tree is not present on the
GPU host):  python tests/golden/make_reference_goldens.py

The fixture holds only data: the numeric literals and index sets that the
reference's test files assert against.  Sources:

  G1  test/test_spline_grid.jl:29-53   2-D -> 2 grid, control points and eval
  G2  test/test_local_refinement.jl:48-67  18 refinement indices
  G3  test/test_knot_vector.jl:26-36   knots_all of KnotVector(5, 2)
  G4  test/test_nurbs_grid.jl:25-52    unit-circle NURBS inputs
"""
from __future__ import annotations

import json
import os
import re

import numpy as np

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                   "reference_goldens.json")


def _julia_array(literal: str) -> np.ndarray:
    """Parse a Julia array literal of the form [a b; c d;;; e f; g h].

    ';;;' separates slices along dim 3, ';' rows (dim 1), blanks dim 2."""
    body = literal.strip()
    assert body[0] == "[" and body[-1] == "]"
    body = body[1:-1]
    slices = []
    for sl in body.split(";;;"):
        rows = [r.split() for r in sl.split(";") if r.strip()]
        slices.append(np.array([[float(x) for x in r] for r in rows]))
    return np.stack(slices, axis=2)


def _literals_after(text: str, anchor: str, count: int) -> list[str]:
    pos = text.index(anchor)
    out = []
    for _ in range(count):
        start = text.index("[", pos)
        depth, i = 0, start
        while True:
            if text[i] == "[":
                depth += 1
            elif text[i] == "]":
                depth -= 1
                if depth == 0:
                    break
            i += 1
        out.append(text[start:i + 1])
        pos = i + 1
    return out


def main() -> None:
    with open(os.path.join(REF, "test/test_spline_grid.jl")) as f:
        sg = f.read()
    cp_lit, eval_lit = _literals_after(sg, "copyto!(\n        spline_grid.control_points", 1)[0], \
        _literals_after(sg, "adapt(backend,\n        [", 1)[0]
    cp = _julia_array(cp_lit)
    ev = _julia_array(eval_lit)
    assert cp.shape == (5, 6, 2) and ev.shape == (7, 9, 2)

    with open(os.path.join(REF, "test/test_local_refinement.jl")) as f:
        lr = f.read()
    m = re.search(r"Int32\[(.*?)\]", lr, re.S)
    pairs = [[int(v) for v in row.split()] for row in m.group(1).split(";")
             if row.strip()]
    assert len(pairs) == 18

    with open(os.path.join(REF, "test/test_nurbs_grid.jl")) as f:
        nb = f.read()
    cm = re.search(r"copyto!\(nurbs_grid\.control_points, \[(.*?)\]\)", nb, re.S)
    circ_cp = [[float(v) for v in row.split()] for row in cm.group(1).split(";")
               if row.strip()]
    assert len(circ_cp) == 9

    data = {
        "G1": {
            "source": "test/test_spline_grid.jl:29-53",
            "n_control_points": [5, 6], "degree": [3, 2],
            "n_sample_points": [7, 9], "Nout": 2,
            "control_points_colmajor": cp.ravel(order="F").tolist(),
            "eval_colmajor": ev.ravel(order="F").tolist(),
        },
        "G2": {
            "source": "test/test_local_refinement.jl:48-67",
            "n_control_points": [6, 6], "degree": [2, 2],
            "n_sample_points": [50, 50], "Nout": 3,
            "refined_n_control_points": [10, 10],
            "error_ranges_1based": [[20, 40], [10, 30]], "error_channel_1based": 2,
            "refinement_indices_1based": pairs,
        },
        "G3": {
            "source": "test/test_knot_vector.jl:26-36",
            "n": 5, "p": 2,
            "knot_values": [0.0, 1 / 3, 2 / 3, 1.0],
            "multiplicities": [3, 1, 1, 3],
            "knots_all": [0.0, 0.0, 0.0, 1 / 3, 2 / 3, 1.0, 1.0, 1.0],
            "extent2": [5.0, 7.0],
        },
        "G4": {
            "source": "test/test_nurbs_grid.jl:25-52",
            "n_control_points": 9, "degree": 2, "n_sample_points": 500,
            "knot_values": [0.0, float(np.float32(np.pi / 2)), float(np.float32(np.pi)),
                            float(np.float32(3 * np.pi / 2)), float(np.float32(2 * np.pi))],
            "multiplicities": [3, 2, 2, 2, 3],
            "weights_even_1based": float(1 / np.sqrt(2)),
            "control_points": circ_cp,
        },
    }
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
