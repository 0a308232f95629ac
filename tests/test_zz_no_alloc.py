"""The allocation-free build (rwh_build_octree_noalloc, C ABI) against the
workspace build (rwh_build_octree): every output array and both counts equal
byte for byte, over consecutive calls of decreasing size that share the
library's scratch array."""
import numpy as np
import pytest

from parity_helpers import bits

pytestmark = pytest.mark.gpu


def noalloc_build(rw, pts):
    n = pts.shape[0]
    cap = rw.octree_capacity(n)
    tb = rw.TreeBuild(n, cap, rw.DeviceArray.from_numpy(pts.reshape(-1)))
    d_counts = rw.DeviceArray(3, np.uint32, zero=True)
    rw.check(rw.lib().rwh_build_octree_noalloc(tb.ptr(), n, 0.0, 1024.0, d_counts.ptr, None),
             "rwh_build_octree_noalloc")
    nu, no, status = (int(x) for x in d_counts.numpy())  # waits for the null stream
    assert status == 0
    return (nu, max(nu - 1, 0), no), tb.read(n, min(max(no, 1), cap))


def test_noalloc_equals_workspace_build(gpu):
    rw = gpu
    clouds = [rw.uniform_points(rw.NOALLOC_MAX_POINTS, 31),  # the largest size it takes
              rw.gmm_points(300001, 32, 64, 20.48),
              rw.uniform_points(5000, 33)]
    for pts in clouds:
        n = pts.shape[0]
        print(f"case noalloc n={n}", flush=True)
        counts, g = noalloc_build(rw, pts)
        want_counts, want = rw.build_octree(pts, sort_variant=5)
        assert counts == want_counts, n
        assert counts[2] <= rw.octree_capacity(n)
        for f in want:
            assert np.array_equal(bits(g[f]), bits(want[f])), (n, f)


def test_noalloc_rejects_sizes_above_its_maximum(gpu):
    rw = gpu
    L = rw.lib()
    d_counts = rw.DeviceArray(3, np.uint32, zero=True)
    tb = rw.TreeBuild(16, 9)
    st = L.rwh_build_octree_noalloc(tb.ptr(), rw.NOALLOC_MAX_POINTS + 1, 0.0, 1024.0, d_counts.ptr, None)
    assert st == rw.RWH_ERR_CAPACITY
    assert b"RWH_NOALLOC_MAX_POINTS" in L.rwh_last_error()
