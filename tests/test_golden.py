"""Golden fixtures (tests/golden/, written by make_golden.py from the pinned
oracle): the seeded generators reproduce the fixture inputs, the oracle
reproduces the fixture outputs (CPU), and the HIP build does too (GPU).

Small cases hold every array raw; the reference's own CPU case (307,200
uniform points, seed 114514) holds counts and SHA-256 digests."""
import hashlib
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
SMALL = os.path.join(GOLDEN, "golden_small.npz")
BIG = os.path.join(GOLDEN, "golden_307200.json")
FIELDS = ("morton_keys_alt", "unique", "prefix_n", "has_leaf_left", "has_leaf_right", "left_child",
          "parents", "edge_count", "edge_offset", "children", "corner", "cell_size",
          "child_node_mask", "child_leaf_mask")
SLOT_FIELDS = ("children", "corner", "cell_size", "child_node_mask", "child_leaf_mask")


def inputs(rw, name):
    kind, n = name.rsplit("_", 1)
    n = int(n)
    return rw.uniform_points(n, 114514) if kind == "uniform" else rw.gmm_points(n, 2, 4, 8.0)


def small_cases():
    with np.load(SMALL) as z:
        names = sorted({k.split("/")[0] for k in z.files})
        return {nm: {k.split("/")[1]: z[k] for k in z.files if k.startswith(nm + "/")} for nm in names}


def live(rw, o, counts, n):
    nu, nb, no = counts
    rows = min(max(no, 1), rw.octree_capacity(n))
    out = {"morton_keys_alt": o["morton_keys_alt"][:n], "unique": o["unique"][:nu]}
    for f in FIELDS[2:]:
        out[f] = o[f][:rows] if f in SLOT_FIELDS else o[f][:nb]
    return out


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def check_small(rw, name, g, counts, outs):
    assert tuple(int(x) for x in g["counts"]) == tuple(counts), name
    got = live(rw, outs, counts, g["points"].shape[0])
    for f in FIELDS:
        assert np.array_equal(np.ascontiguousarray(got[f]).view(np.uint8),
                              np.ascontiguousarray(g[f]).view(np.uint8)), (name, f)


def check_big(rw, g, counts, outs, n):
    assert list(counts) == g["counts"]
    got = live(rw, outs, counts, n)
    for f in FIELDS:
        assert sha(got[f]) == g["sha256"][f], f


def test_generators_reproduce_fixture_inputs(rw):
    for name, g in small_cases().items():
        assert np.array_equal(inputs(rw, name), g["points"]), name
    big = json.load(open(BIG))
    for name, g in big.items():
        assert sha(inputs(rw, name)) == g["points_sha256"], name


def test_oracle_reproduces_fixtures(rw, orc, pool):
    for name, g in small_cases().items():
        (nu, nb, no, rc), o, _ = orc.build(pool, g["points"])
        check_small(rw, name, g, (nu, nb, no), o)
    for name, g in json.load(open(BIG)).items():
        pts = inputs(rw, name)
        (nu, nb, no, rc), o, _ = orc.build(pool, pts)
        check_big(rw, g, (nu, nb, no), o, pts.shape[0])


@pytest.mark.gpu
def test_hip_build_reproduces_fixtures(gpu):
    rw = gpu
    for name, g in small_cases().items():
        counts, o = rw.build_octree(g["points"])
        check_small(rw, name, g, counts, dict(o, unique=o["morton_keys"]))
    for name, g in json.load(open(BIG)).items():
        pts = inputs(rw, name)
        for v in range(len(rw.sort_variants())):
            print(f"case golden {name} sort_variant={v}", flush=True)
            counts, o = rw.build_octree(pts, sort_variant=v)
            check_big(rw, g, counts, dict(o, unique=o["morton_keys"]), pts.shape[0])
