"""Shared comparison of one build's outputs (GPU) with the oracle's."""
import numpy as np

PER_POINT = ("prefix_n", "has_leaf_left", "has_leaf_right", "left_child", "parents",
             "edge_count", "edge_offset")
PER_SLOT = ("children", "corner", "cell_size", "child_node_mask", "child_leaf_mask")


def bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint8)


def compare(rw, n, g_counts, g, o_counts, o, tag=""):
    nu, nb, no = g_counts
    assert (nu, nb, no) == tuple(o_counts[:3]), (tag, g_counts, o_counts)
    assert np.array_equal(g["morton_keys_alt"][:n], o["morton_keys_alt"][:n]), tag + " sorted"
    assert np.array_equal(g["morton_keys"][:nu] if "morton_keys" in g else g["unique"][:nu],
                          o["unique"][:nu]), tag + " unique"
    for k in PER_POINT:
        assert np.array_equal(g[k][:nb], o[k][:nb]), f"{tag} {k}"
    rows = min(no, rw.octree_capacity(n))
    for k in PER_SLOT:
        assert np.array_equal(bits(g[k][:rows]), bits(o[k][:rows])), f"{tag} {k}"
    if no == 0 and rw.octree_capacity(n) > 0 and nb > 0 and len(g["corner"]) > 0:
        # D8: with no contributing node the root cell stays at slot 0 (a
        # packed result carries no octree row then: its length is the count)
        assert np.array_equal(bits(g["corner"][:1]), bits(o["corner"][:1]))
