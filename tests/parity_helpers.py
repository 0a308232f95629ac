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


def far_fanout_tree(rw, orc, pool, n_points=12000, seed=78, fan=40):
    """The oracle's tree of a small cloud, then rewired (acyclic): fan nodes of
    edge count 0 in the first octree tile (1024 radix nodes), each pulled by
    its contributing parent, get as left child a node of count 0 far outside
    that tile's window (the last such node), their right child a leaf.  The
    first tile then queues more walks that leave its window than its queue
    holds (16), so both the queue and the redo of a whole walk from global
    memory run.  Returns (n_unique, n_brt, n_octree_nodes, outputs)."""
    pts = rw.uniform_points(n_points, seed)
    (nu, nb, no, rc), o, _ = orc.build(pool, pts)
    assert rc == 0
    cnt, par = o["edge_count"][:nb], o["parents"][:nb]
    far = next(i for i in range(nb - 2, 0, -1)
               if cnt[i] == 0 and not o["has_leaf_left"][i] and not o["has_leaf_right"][i])
    assert far > 1024 + 128 + 1024, (far, nb)
    qs = [i for i in range(1, 1024) if cnt[i] == 0 and cnt[par[i]] > 0][:fan]
    assert len(qs) == fan
    for q in qs:
        o["left_child"][q] = far
        o["has_leaf_left"][q] = 0
        o["has_leaf_right"][q] = 1
    return nu, nb, no, o
