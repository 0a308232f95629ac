"""Serving: batch planning (CPU) and the continuous-batching server (GPU).

The server's results (packed results, include/rwhip.h) must equal one plain
build of the same cloud, whatever batch a request landed in and whatever ran
beside it.
"""
import numpy as np
import pytest

CHUNK = 8192


def test_planner_takes_fifo_prefix_that_fits(rw):
    slot = 10 * CHUNK
    # footprints in chunks: 1, 2, 1, 5, 1 -> the first four fit 9 <= 10, the fifth would make 10
    sizes = [1, 2 * CHUNK, CHUNK, 5 * CHUNK - 7, 100]
    assert rw.plan_batch(slot, 512, sizes) == 5
    assert rw.plan_batch(slot, 512, sizes + [CHUNK]) == 5
    # arrival order is kept: a big head blocks smaller followers rather than being skipped
    assert rw.plan_batch(slot, 512, [9 * CHUNK, 2 * CHUNK, 1]) == 1
    assert rw.plan_batch(slot, 3, [1] * 10) == 3
    assert rw.plan_batch(slot, 512, []) == 0
    # an empty cloud still occupies one chunk
    assert rw.plan_batch(slot, 512, [0] * 11) == 10


def test_planner_octree_capacity_bound(rw):
    # the octree area of a slot is 0.6 * slot_points + 64 per cloud
    slot = 100 * CHUNK
    assert rw.plan_batch(slot, 4, [25 * CHUNK] * 4) == 4


@pytest.mark.gpu
def test_mix_equals_alone(gpu):
    rw = gpu
    sizes = [5000, 0, 70000, 1, 120000, 33, 8192, 250000, 9000, 3, 40000, 16385, 77777, 2, 150000]
    clouds = [rw.uniform_points(n, 700 + i) if i % 2 else rw.gmm_points(n, 700 + i, 16, 10.24)
              for i, n in enumerate(sizes)]
    srv = rw.Server(pool_bytes=8 << 30, n_slots=2, slot_points=1 << 19, max_clouds=8)
    res = srv.mix(clouds)
    assert max(b for _, b in res) > 1  # requests really were coalesced
    for i, c in enumerate(clouds):
        blob, _ = res[i]
        assert rw.result_check(blob) == rw.RWH_OK, i
        head, o = rw.unpack_result(blob)
        nu, nb, no = head["n_unique"], head["n_brt"], head["n_octree_nodes"]
        g_counts, g = rw.build_octree(c)
        assert (nu, nb, no) == g_counts, i
        n = c.shape[0]
        assert np.array_equal(o["morton_keys_alt"], g["morton_keys_alt"][:n])
        assert np.array_equal(o["unique"], g["morton_keys"][:nu])
        for k in ("prefix_n", "has_leaf_left", "has_leaf_right", "left_child", "parents",
                  "edge_count", "edge_offset"):
            assert np.array_equal(o[k], g[k][:nb]), (i, k)
        rows = min(no, rw.octree_capacity(n))
        for k in ("children", "corner", "cell_size", "child_node_mask", "child_leaf_mask"):
            assert len(o[k]) == rows, (i, k)
            assert np.array_equal(o[k].view(np.uint8), g[k][:rows].view(np.uint8)), (i, k)
    srv.close()


@pytest.mark.gpu
def test_oversized_request_is_rejected_and_server_continues(gpu):
    rw = gpu
    srv = rw.Server(pool_bytes=4 << 30, n_slots=1, slot_points=1 << 16, max_clouds=4)
    big = rw.uniform_points((1 << 16) + 1, 1)
    with pytest.raises(rw.RwhError) as e:
        srv.mix([big])
    assert "does not fit" in str(e.value)
    small = [rw.uniform_points(1000, 2), rw.uniform_points(3000, 3)]
    res = srv.mix(small, sections=rw.SECTIONS_OCTREE)
    head, arrays = rw.unpack_result(res[0][0])
    assert (head["n_unique"], head["n_brt"], head["n_octree_nodes"]) == rw.build_octree(small[0])[0]
    assert set(arrays) == {"children", "corner", "cell_size", "child_node_mask", "child_leaf_mask"}
    # a section bit outside the format fails the request through its future
    with pytest.raises(rw.RwhError):
        srv.mix(small, sections=1 << 14)
    srv.close()


@pytest.mark.gpu
def test_open_loop_load_reports(gpu):
    rw = gpu
    clouds = [rw.uniform_points(n, 50 + i) for i, n in enumerate([4096, 20000, 100000, 300000])]
    srv = rw.Server(pool_bytes=16 << 30, n_slots=3, slot_points=1 << 21, max_clouds=64)
    st = srv.load(clouds, rate=500.0, n_requests=200, seed=3)
    assert st["rejected"] == 0 and st["n_requests"] == 200
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    assert st["batches"] >= 1 and st["mean_batch"] >= 1
    srv.close()
