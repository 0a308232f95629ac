"""The status word of a build: kernels that stop at one of their bounds set a
RWH_STATUS_* bit, and the read-back calls report RWH_ERR_OVERFLOW instead of
returning a wrong tree as a success.

The GPU cases force the conditions from the host with inputs every kernel
still finishes on (keys out of order; a tree whose left_child links form a
cycle, walked until the pull's step budget ends it; a child index past the
tree's end)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_status_bits_match_the_header(rw):
    text = open(os.path.join(ROOT, "include", "rwhip.h")).read()
    bits = dict(re.findall(r"#define RWH_STATUS_(\w+) (0x[0-9a-f]+)u", text))
    assert {k: int(v, 16) for k, v in bits.items()} == {
        "PAIR_CAP": rw.STATUS_PAIR_CAP, "CLIMB": rw.STATUS_CLIMB, "CROSS_CAP": rw.STATUS_CROSS_CAP,
        "KEY_ORDER": rw.STATUS_KEY_ORDER, "PULL": rw.STATUS_PULL, "LEAF": rw.STATUS_LEAF}
    assert int(re.search(r"#define RWH_ERR_OVERFLOW (\d+)", text).group(1)) == rw.RWH_ERR_OVERFLOW
    assert b"bound" in rw.lib().rwh_error_string(rw.RWH_ERR_OVERFLOW)


@pytest.mark.gpu
def test_read_counts_reports_a_set_status(gpu):
    rw = gpu
    L = rw.lib()
    pts = rw.uniform_points(5000, 1)
    ws = rw.Workspace(5000)
    counts, _ = rw.build_octree(pts, ws=ws)  # a clean build: status 0, no error
    flag = np.array([rw.STATUS_CROSS_CAP | rw.STATUS_PULL], np.uint32)
    rw.check(L.rwh_memcpy_h2d(ws.status_ptr(), flag.ctypes.data, 4, None))
    c3 = (C.c_uint32 * 3)()
    assert L.rwh_read_counts(ws.ptr, 1 << 20, c3, None) == rw.RWH_ERR_OVERFLOW
    msg = L.rwh_last_error()
    assert b"edge-crossing" in msg and b"link walk" in msg
    assert rw.build_octree(pts, ws=ws)[0] == counts  # the next build clears it


@pytest.mark.gpu
def test_read_batch_counts_names_the_cloud(gpu):
    rw = gpu
    L = rw.lib()
    clouds = [rw.uniform_points(n, 40 + i) for i, n in enumerate((3000, 9000, 500))]
    b = rw.Batch(clouds)
    b.run()
    b.counts()  # clean
    flag = np.array([rw.STATUS_LEAF], np.uint32)
    rw.check(L.rwh_memcpy_h2d(b.ws.status_ptr(2), flag.ctypes.data, 4, None))
    with pytest.raises(rw.RwhError) as e:
        b.counts()
    assert e.value.status == rw.RWH_ERR_OVERFLOW and "cloud 2" in str(e.value)
    b.run()  # the batch's first kernel clears every cloud's word
    b.counts()


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["duplicate", "descending"])
def test_radix_tree_reports_keys_out_of_order(gpu, kind):
    rw = gpu
    L = rw.lib()
    k = (np.arange(20000, dtype=np.uint32) * 7) << 4
    if kind == "duplicate":
        k[9000] = k[9001]
    else:
        k[12000:12010] = k[12000:12010][::-1].copy()
    nu = k.size
    ws = rw.Workspace(nu)
    d_k = rw.DeviceArray.from_numpy(k)
    d_nu = rw.DeviceArray.from_numpy(np.array([nu], np.uint32))
    out = [rw.DeviceArray(nu, dt) for dt in (np.uint8, np.uint8, np.uint8, np.int32, np.int32)]
    rw.check(L.rwh_radix_tree(ws.ptr, d_k.ptr, d_nu.ptr, nu, *[o.ptr for o in out], None))
    c3 = (C.c_uint32 * 3)()
    assert L.rwh_read_counts(ws.ptr, 1 << 20, c3, None) == rw.RWH_ERR_OVERFLOW
    assert b"not sorted and unique" in L.rwh_last_error()
    # sorted unique keys on the same workspace: the status is cleared again
    k = np.unique(k)
    d_k.upload(k)
    d_nu.upload(np.array([k.size], np.uint32))
    rw.check(L.rwh_radix_tree(ws.ptr, d_k.ptr, d_nu.ptr, k.size, *[o.ptr for o in out], None))
    rw.check(L.rwh_read_counts(ws.ptr, 1 << 20, c3, None))


def malformed_tree(rw, orc, pool, how):
    """The oracle's radix tree, edge counts and offsets of a small cloud with
    one node of edge count 0 (below a contributing parent) rewired: its left
    child points at itself ("cycle"), past the tree ("outside"), or at a node
    of count 0 far outside the node's tile, whose own left child is itself
    ("far_cycle": the walk leaves the tile's window before it cycles)."""
    pts = rw.uniform_points(6000, 77)
    (nu, nb, no, rc), o, _ = orc.build(pool, pts)
    cnt, par = o["edge_count"][:nb], o["parents"][:nb]
    q = next(i for i in range(1, nb) if cnt[i] == 0 and cnt[par[i]] > 0)
    if how == "far_cycle":
        r = next(i for i in range(nb - 2, 0, -1) if cnt[i] == 0)
        assert r - q > 2048
        o["left_child"][q] = r
        o["left_child"][r] = r
        o["has_leaf_left"][r] = 0
        o["has_leaf_right"][r] = 0
    else:
        o["left_child"][q] = q if how == "cycle" else nb + 100
    o["has_leaf_left"][q] = 0
    o["has_leaf_right"][q] = 1
    return nu, nb, no, o


@pytest.mark.gpu
@pytest.mark.parametrize("with_ws", [False, True])
@pytest.mark.parametrize("how", ["cycle", "outside", "far_cycle"])
def test_octree_reports_a_malformed_tree(gpu, orc, pool, how, with_ws):
    rw = gpu
    L = rw.lib()
    nu, nb, no, o = malformed_tree(rw, orc, pool, how)
    cap = rw.octree_capacity(nu)
    tb = rw.TreeBuild(nu, cap)
    tb["morton_keys"].upload(o["unique"][:nu])
    for f in ("prefix_n", "has_leaf_left", "has_leaf_right", "left_child", "parents", "edge_count",
              "edge_offset"):
        tb[f].upload(o[f][:nb])
    d_nu = rw.DeviceArray.from_numpy(np.array([nu], np.uint32))
    d_no = rw.DeviceArray.from_numpy(np.array([no], np.uint32))
    d_status = rw.DeviceArray(1, np.uint32, zero=True)
    # with a workspace, walks that leave a tile's window run in a second kernel
    ws = rw.Workspace(nu) if with_ws else None
    print(f"case malformed tree {how} n_brt={nb} workspace={with_ws}", flush=True)
    rw.check(L.rwh_octree(ws.ptr if ws else None, tb.ptr(), d_nu.ptr, d_no.ptr, nu, 0.0, 1024.0,
                          d_status.ptr, None))
    assert int(d_status.numpy()[0]) & rw.STATUS_PULL


@pytest.mark.gpu
def test_octree_queue_overflow_equals_walks_in_place(gpu, orc, pool):
    """rwh_octree with a workspace (far walks queued, the overflow redone from
    global memory) and without one (every walk in place) give the same rows,
    on a tree whose first tile has more far walks than its queue holds."""
    from parity_helpers import far_fanout_tree
    rw = gpu
    L = rw.lib()
    nu, nb, no, o = far_fanout_tree(rw, orc, pool)
    cap = rw.octree_capacity(nu)
    d_nu = rw.DeviceArray.from_numpy(np.array([nu], np.uint32))
    d_no = rw.DeviceArray.from_numpy(np.array([no], np.uint32))
    outs = []
    for with_ws in (False, True):
        tb = rw.TreeBuild(nu, cap)
        tb["morton_keys"].upload(o["unique"][:nu])
        for f in ("prefix_n", "has_leaf_left", "has_leaf_right", "left_child", "parents", "edge_count",
                  "edge_offset"):
            tb[f].upload(o[f][:nb])
        d_status = rw.DeviceArray(1, np.uint32, zero=True)
        ws = rw.Workspace(nu) if with_ws else None
        print(f"case far fan-out n_brt={nb} workspace={with_ws}", flush=True)
        rw.check(L.rwh_octree(ws.ptr if ws else None, tb.ptr(), d_nu.ptr, d_no.ptr, nu, 0.0, 1024.0,
                              d_status.ptr, None))
        assert int(d_status.numpy()[0]) == 0
        outs.append(tb.read(nu, min(no, cap)))
    for f in ("children", "corner", "cell_size", "child_node_mask", "child_leaf_mask"):
        assert np.array_equal(outs[0][f].view(np.uint8), outs[1][f].view(np.uint8)), f
