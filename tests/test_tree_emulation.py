"""The radix-tree kernels themselves, run on the CPU.

tests/cpp/bin/tree_emul compiles redwood-comp_amd/csrc/k_stages.hip
unmodified with g++ under a host emulation of the HIP runtime (one thread per
work-item, barriers for __syncthreads and the wave collectives) and
AddressSanitizer, and runs launch_radix_tree (k_tree_local, then
k_tree_cross) over sorted keys with every buffer sized exactly as the C ABI
documents.  An index past any buffer or LDS array aborts the run; the outputs
must equal the oracle's.  Cases: unique-key counts at the edges of the
4096-leaf tree tile, uniform, regular-stride, clustered, isolated first/last
keys, dense runs, keys with bit 31 set, nested prefixes crossing one tile
edge at every level, and keys out of order (reported, no access out of
bounds)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "cpp", "bin", "tree_emul")
TILE = 4096


@pytest.fixture(scope="module")
def emul():
    if not os.path.exists(EMU):
        subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "cpp"), "all"], check=True,
                       capture_output=True)
    return EMU


def run(emul, keys, tmp_path):
    kp, op = tmp_path / "keys.u32", tmp_path / "out.bin"
    np.ascontiguousarray(keys, np.uint32).tofile(kp)
    r = subprocess.run([emul, str(kp), str(op)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    b = op.read_bytes()
    nb = keys.size - 1
    out = {"status": int(np.frombuffer(b[:4], np.uint32)[0])}
    o = 4
    for name, dt in (("prefix_n", np.uint8), ("has_leaf_left", np.uint8), ("has_leaf_right", np.uint8),
                     ("left_child", np.int32), ("parents", np.int32), ("edge_count", np.int32)):
        w = np.dtype(dt).itemsize
        out[name] = np.frombuffer(b[o:o + nb * w], dt)
        o += nb * w
    out["tile_sum"] = np.frombuffer(b[o:], np.uint32)
    return out


def nested_keys(nu, edge):
    """Keys whose nodes of every prefix length cross the tile edge: keys
    c = 2m and c + 1 sit at indices edge - 1 and edge (they differ only in
    bit 0), and for each higher bit b the key c with bit b flipped joins the
    range on alternating sides, so the node of each prefix length holds both
    keys at the edge.  Filler keys below and above keep the indices in
    place."""
    c = 0x15555554  # bits alternate: flipped keys fall left and right in turn
    nested = {c, c + 1} | {c ^ (1 << b) for b in range(1, 29)}
    below = sorted(x for x in nested if x < c)
    above = sorted(x for x in nested if x > c + 1)
    n_lo = edge - 1 - len(below)
    n_hi = nu - edge - 1 - len(above)
    lo_max, hi_min = min(nested), max(nested)
    fill_lo = np.linspace(0, lo_max - 1, n_lo).astype(np.int64)
    fill_hi = np.linspace(hi_min + 1, (1 << 30) - 1, n_hi).astype(np.int64)
    k = np.concatenate([fill_lo, below, [c, c + 1], above, fill_hi]).astype(np.int64)
    assert k.size == nu and np.all(np.diff(k) > 0) and k[edge - 1] == c
    return k.astype(np.uint32)


def keys_of(kind, nu, seed):
    rng = np.random.default_rng(seed)
    if kind == "k30":
        k = rng.choice(1 << 30, size=nu, replace=False)
    elif kind == "stride":
        k = np.arange(nu, dtype=np.int64) << 7
    elif kind == "clustered":
        centers = rng.choice(1 << 10, size=16, replace=False).astype(np.int64) << 20
        per = -(-nu // 16)
        k = np.unique(np.concatenate([c + rng.choice(1 << 18, size=per, replace=False) for c in centers]))[:nu]
    elif kind == "lastfar":
        k = np.sort(rng.choice(1 << 20, size=nu - 1, replace=False))
        k = np.concatenate([k, [(1 << 30) - 1]])
    elif kind == "firstfar":
        k = np.sort(rng.choice(1 << 20, size=nu - 1, replace=False)) | (1 << 29)
        k = np.concatenate([[0], k])
    elif kind == "dense":
        k = np.arange(nu, dtype=np.int64) * 3 + (rng.integers(0, 3, size=nu) == 0)
    elif kind == "k32":
        k = rng.choice(1 << 32, size=nu, replace=False)
    elif kind == "nested":
        return nested_keys(nu, TILE)
    else:
        raise ValueError(kind)
    return np.sort(np.asarray(k, np.int64)).astype(np.uint32)


CASES = [(kind, nu) for nu in (4095, 4096, 4097, 4098, 8191, 8193, 12289)
         for kind in ("k30", "stride", "clustered", "lastfar", "firstfar", "dense")]
CASES += [("k32", 5000), ("k30", 40961), ("clustered", 40961), ("nested", 4096 + 200),
          ("nested", 2 * 4096 + 7)]


@pytest.mark.parametrize("kind,nu", CASES)
def test_tree_kernels_on_cpu_match_oracle(emul, orc, pool, tmp_path, kind, nu):
    k = keys_of(kind, nu, seed=nu + len(kind))
    assert k.size == nu and np.all(np.diff(k.astype(np.int64)) > 0)
    got = run(emul, k, tmp_path)
    assert got["status"] == 0
    want = orc.radix_tree(pool, k)
    for f in want:
        assert np.array_equal(got[f], want[f]), f
    if kind != "k32":  # the edge counts beside the tree (keys below 2^31)
        ec = orc.edge_count(pool, want["prefix_n"], want["parents"])
        assert np.array_equal(got["edge_count"], ec)
        sums = np.add.reduceat(ec.astype(np.int64), np.arange(0, ec.size, TILE))
        assert np.array_equal(got["tile_sum"][:sums.size].astype(np.int64), sums)


@pytest.mark.parametrize("how", ["duplicate", "descending", "all_equal"])
def test_tree_kernels_report_keys_out_of_order(emul, tmp_path, how):
    k = (np.arange(9000, dtype=np.uint32) * 5) << 3
    if how == "duplicate":
        k[4500] = k[4501]
    elif how == "descending":
        k[100:110] = k[100:110][::-1].copy()
    else:
        k[:] = 12345
    got = run(emul, k, tmp_path)
    assert got["status"] & 0x8  # RWH_STATUS_KEY_ORDER, and no access out of bounds


# ---------------------------------------------------------------------------
# S7 (k_oct_nodes) and leaf linking (k_link_leaf) on the CPU
# ---------------------------------------------------------------------------
OCT_EMU = os.path.join(ROOT, "tests", "cpp", "bin", "octree_emul")
NODE_FIELDS = (("prefix_n", np.uint8), ("has_leaf_left", np.uint8), ("has_leaf_right", np.uint8),
               ("left_child", np.int32), ("parents", np.int32), ("edge_count", np.int32),
               ("edge_offset", np.int32))


def run_octree(o, nu, no, cap, tmp_path, link_leaf=False, check=True, min_coord=0.0, range_=1024.0,
               queue=False):
    """queue: walks that leave a tile's LDS window go to the per-tile queue and
    the second kernel (k_oct_far), as in the whole builds; otherwise the node
    kernel walks them itself."""
    if not os.path.exists(OCT_EMU):
        subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "cpp"), "all"], check=True,
                       capture_output=True)
    nb = nu - 1
    ip, op = tmp_path / "tree.bin", tmp_path / "oct.bin"
    with open(ip, "wb") as f:
        f.write(np.array([nu, no, cap, int(link_leaf) | (0 if check else 2) | (4 if queue else 0)],
                         np.uint32).tobytes())
        f.write(np.array([min_coord, range_], np.float32).tobytes())
        f.write(np.ascontiguousarray(o["unique"][:nu], np.uint32).tobytes())
        for name, dt in NODE_FIELDS:
            f.write(np.ascontiguousarray(o[name][:nb], dt).tobytes())
    r = subprocess.run([OCT_EMU, str(ip), str(op)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    b = op.read_bytes()
    out = {"status": int(np.frombuffer(b[:4], np.uint32)[0])}
    o_ = 4
    for name, dt, w in (("children", np.int32, 8), ("corner", np.float32, 4), ("cell_size", np.float32, 1),
                        ("child_node_mask", np.int32, 1), ("child_leaf_mask", np.int32, 1)):
        cnt = cap * w
        a = np.frombuffer(b[o_:o_ + cnt * np.dtype(dt).itemsize], dt)
        out[name] = a.reshape(cap, w) if w > 1 else a
        o_ += cnt * np.dtype(dt).itemsize
    out["queued"] = int(np.frombuffer(b[o_:o_ + 4], np.uint32)[0])
    return out


def octree_cases(rw):
    yield "uniform", rw.uniform_points(20000, 5)
    yield "gmm", rw.gmm_points(30000, 6, 64, 20.48)
    p = rw.uniform_points(9000, 7)
    p[:, :3] = np.floor(p[:, :3] / 64) * 64 + 3
    yield "lattice", p
    for kind in ("lastfar", "firstfar"):
        p = rw.uniform_points(6000, 8)
        p[:, :3] = p[:, :3] / 64.0 + (0.0 if kind == "lastfar" else 900.0)
        p[-1, :3] = 1000.0 if kind == "lastfar" else 0.5
        yield kind, p


@pytest.mark.parametrize("link_leaf,check,queue", [(False, True, False), (False, False, False),
                                                   (True, True, False), (False, False, True),
                                                   (True, True, True)])
def test_octree_kernels_on_cpu_match_oracle(rw, orc, pool, tmp_path, link_leaf, check, queue):
    """check: the kernel of the stage entry point (a tree from the caller);
    without: the one the whole builds run on the tree they built.  queue: far
    walks run in the second kernel (the whole builds, and the stage entry
    point given a workspace)."""
    queued = 0
    for name, pts in octree_cases(rw):
        (nu, nb, no, rc), o, _ = orc.build(pool, pts)
        cap = rw.octree_capacity(pts.shape[0])
        assert rc == 0 and no <= cap, name
        got = run_octree(o, nu, no, cap, tmp_path, link_leaf, check, queue=queue)
        assert got["status"] == 0, name
        queued += got["queued"]
        if link_leaf:
            orc.link_leaf(pool, o, (nu, nb, no))
        for f in ("children", "corner", "cell_size", "child_node_mask", "child_leaf_mask"):
            assert np.array_equal(got[f][:no].view(np.uint8), o[f][:no].view(np.uint8)), (name, f)
    # the clouds have walks that leave a tile's window: the queue is exercised
    assert (queued > 0) == queue


@pytest.mark.parametrize("queue", [False, True])
@pytest.mark.parametrize("how", ["cycle", "outside", "far_cycle"])
def test_octree_kernel_on_cpu_reports_a_malformed_tree(rw, orc, pool, tmp_path, how, queue):
    from test_status import malformed_tree
    nu, nb, no, o = malformed_tree(rw, orc, pool, how)
    got = run_octree(o, nu, no, rw.octree_capacity(nu), tmp_path, link_leaf=True, queue=queue)
    assert got["status"] & 0x10  # RWH_STATUS_PULL; every walk ended, nothing read out of bounds
    if queue and how == "far_cycle":
        assert got["queued"] > 0  # the cycle was walked by the second kernel


def test_octree_queue_overflow_equals_walks_in_place(rw, orc, pool, tmp_path):
    """A tile with more far walks than its queue holds: the walks queued for
    k_oct_far and the ones redone from global memory after the pull loop give
    the same rows as every walk run in place (max and OR in any order)."""
    from parity_helpers import far_fanout_tree
    nu, nb, no, o = far_fanout_tree(rw, orc, pool)
    cap = rw.octree_capacity(nu)
    for check in (True, False):
        inplace = run_octree(o, nu, no, cap, tmp_path, check=check)
        queued = run_octree(o, nu, no, cap, tmp_path, check=check, queue=True)
        assert inplace["status"] == queued["status"] == 0
        assert queued["queued"] >= 16  # the first tile's queue is full
        for f in ("children", "corner", "cell_size", "child_node_mask", "child_leaf_mask"):
            assert np.array_equal(inplace[f][:no].view(np.uint8), queued[f][:no].view(np.uint8)), (check, f)
