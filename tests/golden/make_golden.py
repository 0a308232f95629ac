"""Writes the golden fixtures of the octree build from the CPU oracle.

Run after the oracle's pins (tests/test_oracle.py) pass:
    python tests/golden/make_golden.py
Cases (inputs from the package's seeded generators, box [0, 1024)):
  * uniform_64, uniform_1024, gmm_1024: every input and output array, raw
    (golden_small.npz);
  * uniform_307200 (the reference's own CPU case, 640 x 480 points, seed
    114514): counts and a SHA-256 digest of every input and output array
    (golden_307200.json).
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import __graft_entry__ as ge  # noqa: E402

# the arrays of one build, cut to their live lengths
FIELDS = ("morton_keys_alt", "unique", "prefix_n", "has_leaf_left", "has_leaf_right", "left_child",
          "parents", "edge_count", "edge_offset", "children", "corner", "cell_size",
          "child_node_mask", "child_leaf_mask")


def cases(rw):
    return {"uniform_64": rw.uniform_points(64, 114514),
            "uniform_1024": rw.uniform_points(1024, 114514),
            "gmm_1024": rw.gmm_points(1024, 2, 4, 8.0),
            "uniform_307200": rw.uniform_points(640 * 480, 114514)}


def live(o, counts, n, cap):
    nu, nb, no = counts
    rows = min(max(no, 1), cap)
    cut = {"morton_keys_alt": n, "unique": nu}
    out = {}
    for f in FIELDS:
        if f in cut:
            out[f] = o[f][:cut[f]]
        elif f in ("children", "corner", "cell_size", "child_node_mask", "child_leaf_mask"):
            out[f] = o[f][:rows]
        else:
            out[f] = o[f][:nb]
    return out


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    rw = ge.load_package()
    orc = ge.load_oracle()
    pool = orc.Pool(8)
    small = {}
    big = {}
    for name, pts in cases(rw).items():
        n = pts.shape[0]
        (nu, nb, no, rc), o, _ = orc.build(pool, pts)
        assert rc == 0, name
        arrays = live(o, (nu, nb, no), n, rw.octree_capacity(n))
        if n <= 4096:
            small[f"{name}/points"] = pts
            small[f"{name}/counts"] = np.array([nu, nb, no], np.uint32)
            for f, a in arrays.items():
                small[f"{name}/{f}"] = a
        else:
            big[name] = {"n": n, "counts": [nu, nb, no], "points_sha256": digest(pts),
                         "sha256": {f: digest(a) for f, a in arrays.items()}}
    np.savez_compressed(os.path.join(HERE, "golden_small.npz"), **small)
    with open(os.path.join(HERE, "golden_307200.json"), "w") as f:
        json.dump(big, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
