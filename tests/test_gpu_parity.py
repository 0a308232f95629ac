"""GPU parity: every stage output of the HIP path equals the CPU oracle bit for
bit (integers, indices and floats alike), on the same seeded inputs.

All calls go through the C ABI (librwhip.so) or the C++ drop-in layer
(libredwood_tree.so).  Sizes reach 2^24 (the benchmark workload); there the
comparison is still element by element against the oracle.
"""
import numpy as np
import pytest

from parity_helpers import bits, compare

pytestmark = pytest.mark.gpu


# the benchmark's tile shape (1024 x 16)
BENCH_SORT_VARIANT = 5


def inputs(rw, kind, n, seed):
    if kind == "uniform":
        return rw.uniform_points(n, seed)
    if kind == "gmm":
        return rw.gmm_points(n, seed, 64, 20.48)
    if kind == "dups":  # a tiny lattice: most points collide
        p = rw.uniform_points(n, seed)
        p[:, :3] = np.floor(p[:, :3] / 256) * 256
        return p
    if kind in ("lastfar", "firstfar"):
        # all keys share a long prefix except one isolated at the top split
        # (the last key for lastfar: the case where the root's shortened
        # range (Q2) differs from the complete root)
        p = rw.uniform_points(n, seed)
        p[:, :3] = p[:, :3] / 64.0 + (0.0 if kind == "lastfar" else 900.0)
        p[-1, :3] = 1000.0 if kind == "lastfar" else 0.5
        return p
    raise ValueError(kind)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 64, 100, 1000, 4095, 4096, 4097, 8193, 65537, 307200])
@pytest.mark.parametrize("kind", ["uniform", "gmm", "dups", "lastfar", "firstfar"])
def test_build_matches_oracle(gpu, orc, pool, n, kind):
    """Whole builds with every compiled sort tile shape: the fused Morton +
    histogram sweep counts pass 0 per tile of the selected shape (tiles of
    4096, 8192 or 16384 keys), so each shape is its own code path."""
    rw = gpu
    pts = inputs(rw, kind, n, seed=n + 17)
    (nu, nb, no, rc), o, _ = orc.build(pool, pts)
    # sparse clouds can need more octree nodes than the reference's 0.6 n slots:
    # both sides then report it and keep the first 0.6 n slots (D11)
    assert rc == (-1 if no > rw.octree_capacity(n) else 0)
    # every tile shape with the default digit plan, every plan with the
    # benchmark's tile shape
    configs = [(v, 0) for v in range(len(rw.sort_variants()))]
    configs += [(BENCH_SORT_VARIANT, p) for p in range(1, len(rw.sort_plans()))]
    for v, p in configs:
        print(f"case build kind={kind} n={n} sort_variant={v} plan={p}", flush=True)
        g_counts, g = rw.build_octree(pts, strict=False, sort_variant=v, sort_plan=p)
        compare(rw, n, g_counts, g, (nu, nb, no), o, f"{kind}/{n}/v{v}p{p}")


@pytest.mark.parametrize("mn,rg", [(0.0, 1024.0), (0.25, 0.5), (-64.0, 2.0 ** 20), (-3.7, 1000.0),
                                   (1e-3, 3.0), (5.0, 2.0 ** -100)])
def test_build_box_matches_oracle(gpu, orc, pool, mn, rg):
    # power-of-two ranges take the product with the exact reciprocal, other
    # ranges divide; both must give the reference's quotient bit for bit
    rw = gpu
    n = 100003
    p = rw.uniform_points(n, 29)
    p[:, :3] = np.float32(mn) + p[:, :3] / np.float32(1024.0) * np.float32(rg)
    p[:64, :3] = np.float32(mn) - np.float32(abs(rg))   # below the box
    p[64:128, :3] = np.float32(mn) + np.float32(2.0) * np.float32(rg)   # beyond it
    print(f"case box mn={mn} rg={rg}", flush=True)
    g_counts, g = rw.build_octree(p, min_coord=mn, range_=rg, strict=False)
    (nu, nb, no, rc), o, _ = orc.build(pool, p, mn, rg)
    compare(rw, n, g_counts, g, (nu, nb, no), o, f"box {mn} {rg}")


@pytest.mark.parametrize("kind,n", [("uniform", 1 << 24), ("gmm", 1 << 24), ("gmm", 1 << 22)])
def test_full_size_matches_oracle(gpu, orc, kind, n):
    """The benchmarked configuration itself (2^24 uniform points, rwh_build_octree
    with the 1024 x 16 sort tile: the fused sweep's 16384-key pass-0 table),
    the clustered configurations, and the autotuner's shape and digit plan for
    the size (the pair the benchmark runs)."""
    rw = gpu
    assert rw.sort_variants()[BENCH_SORT_VARIANT] == (1024, 16)
    pts = inputs(rw, kind, n, seed=114514 if kind == "uniform" else 1)
    big = orc.Pool(16)
    (nu, nb, no, rc), o, _ = orc.build(big, pts)
    assert rc == 0
    ws = rw.Workspace(n)
    # the digit passes with each shape, then the bucket sort (the benchmark's
    # path), which the clustered points hand back to the digit passes
    runs = [(v, p, False) for v, p in sorted({(0, 0), (BENCH_SORT_VARIANT, 0), rw.best_sort_config(n)})]
    runs.append(rw.best_sort_config(n) + (True,))
    for v, p, bucket in runs:
        print(f"case full kind={kind} n={n} sort_variant={v} plan={p} bucket={bucket}", flush=True)
        g_counts, g = rw.build_octree(pts, ws=ws, sort_variant=v, sort_plan=p, bucket_sort=bucket)
        assert ws.sort_used_buckets() == (bucket and kind == "uniform")
        compare(rw, n, g_counts, g, (nu, nb, no), o, f"{kind}/{n}/v{v}p{p}b{bucket}")


@pytest.mark.parametrize("kind,n", [("uniform", 100003), ("uniform", (1 << 21) + 12345), ("gmm", 1 << 21),
                                    ("uniform", 1)])
def test_marked_build_matches_oracle(gpu, orc, pool, kind, n):
    """rwh_build_octree_marked (the build with stage-boundary events, which the
    benchmark times stage by stage) gives the plain build's output, on the
    digit-pass path, the bucket-sort path and its fallback; its stage times
    are non-negative and S5's is empty (written by the S4 kernels)."""
    rw = gpu
    pts = inputs(rw, kind, n, seed=n % 977 + 5)
    (nu, nb, no, rc), o, _ = orc.build(pool, pts)
    marks = [rw.Event() for _ in range(rw.BUILD_MARKS)]
    print(f"case marked kind={kind} n={n}", flush=True)
    g_counts, g = rw.build_octree(pts, strict=False, marks=marks)
    compare(rw, n, g_counts, g, (nu, nb, no), o, f"marked {kind}/{n}")
    ms = rw.stage_ms(marks)
    assert list(ms) == list(rw.BUILD_STAGES)
    assert all(v >= 0.0 for v in ms.values()), ms
    assert ms["S5_edge_count"] < 0.01, ms


@pytest.mark.parametrize("n", [(1 << 20) + 1, 1 << 21, (1 << 21) + 12345, 3 * (1 << 21) + 7, 1 << 23,
                               (1 << 23) + 1])
def test_bucket_sort_sizes(gpu, orc, n):
    """Whole builds through the bucket sort at the edges of its size range
    (2^20 < n <= 2^24: 256 to 2048 buckets, stable in-bucket passes of 10 down
    to 8 bits, last scatter tiles ending inside a wave) against the oracle."""
    rw = gpu
    pts = rw.uniform_points(n, seed=n % 1000 + 3)
    (nu, nb, no, rc), o, _ = orc.build(orc.Pool(16), pts)
    assert rc == 0
    ws = rw.Workspace(n)
    print(f"case bucket n={n}", flush=True)
    g_counts, g = rw.build_octree(pts, ws=ws)
    assert ws.sort_used_buckets()
    compare(rw, n, g_counts, g, (nu, nb, no), o, f"bucket/{n}")


def test_bucket_sort_out_of_range_runs_digit_passes(gpu):
    rw = gpu
    for n in (1 << 20, 1000):
        ws = rw.Workspace(n)
        rw.build_octree(rw.uniform_points(n, 5), ws=ws)
        assert not ws.sort_used_buckets()


def _cells_points(codes):
    """Points at the centres of the cells of 30-bit Morton codes."""
    c = codes.astype(np.uint64)
    pts = np.zeros((c.size, 4), np.float32)
    for a in range(3):
        v = (c >> np.uint64(a)) & np.uint64(0x09249249)
        v = (v | (v >> np.uint64(2))) & np.uint64(0x030C30C3)
        v = (v | (v >> np.uint64(4))) & np.uint64(0x0300F00F)
        v = (v | (v >> np.uint64(8))) & np.uint64(0x030000FF)
        v = (v | (v >> np.uint64(16))) & np.uint64(0x3FF)
        pts[:, a] = v.astype(np.float32) + np.float32(0.5)
    return pts


@pytest.mark.parametrize("extra", [0, 1])
def test_bucket_sort_capacity_edge(gpu, orc, pool, extra):
    """2^21 points in 256 buckets, bucket 5 holding exactly the 9216 keys an
    in-bucket sort holds (the bucket sort runs) or one more (the digit passes
    run): both builds equal the oracle's."""
    rw = gpu
    n = 1 << 21
    rng = np.random.default_rng(41 + extra)
    big = (np.uint64(5) << np.uint64(22)) + rng.integers(0, 1 << 22, 9216 + extra, dtype=np.uint64)
    rest = rng.integers(0, 1 << 30, 2 * n, dtype=np.uint64)
    rest = rest[(rest >> np.uint64(22)) != 5][:n - big.size]
    codes = np.concatenate([big, rest])
    rng.shuffle(codes)
    pts = _cells_points(codes)
    assert np.array_equal(orc.morton(pool, pts), codes.astype(np.uint32))
    (nu, nb, no, rc), o, _ = orc.build(orc.Pool(16), pts)
    ws = rw.Workspace(n)
    print(f"case bucket capacity extra={extra}", flush=True)
    g_counts, g = rw.build_octree(pts, ws=ws, strict=False)
    assert ws.sort_used_buckets() == (extra == 0)
    compare(rw, n, g_counts, g, (nu, nb, no), o, f"bucket cap+{extra}")


@pytest.mark.parametrize("n", [1, 3, 4, 1000, 4097, 123457, 1 << 20])
@pytest.mark.parametrize("key_bits", [30, 32, 8])
def test_sort_u32(gpu, n, key_bits):
    rw = gpu
    L = rw.lib()
    rng = np.random.default_rng(n * 7 + key_bits)
    k = rng.integers(0, 1 << key_bits, size=n, dtype=np.uint64).astype(np.uint32)
    d_in = rw.DeviceArray.from_numpy(k)
    d_out = rw.DeviceArray(n, np.uint32)
    ws = rw.Workspace(n)
    print(f"case sort n={n} bits={key_bits}", flush=True)
    rw.check(L.rwh_sort_u32(ws.ptr, d_in.ptr, d_out.ptr, n, key_bits, None))
    assert np.array_equal(d_out.numpy(), np.sort(k))
    assert np.array_equal(d_in.numpy(), k)  # input left unchanged
    # sort again with the same workspace: the look-back state is self-resetting
    rw.check(L.rwh_sort_u32(ws.ptr, d_in.ptr, d_out.ptr, n, key_bits, None))
    assert np.array_equal(d_out.numpy(), np.sort(k))


@pytest.mark.parametrize("n", [1, 4097, 100000, 1 << 20])
def test_sort_every_shape(gpu, n):
    """Every compiled tile shape sorts exactly (single cloud, 30-bit keys)."""
    rw = gpu
    L = rw.lib()
    rng = np.random.default_rng(n)
    k = rng.integers(0, 1 << 30, size=n, dtype=np.uint64).astype(np.uint32)
    d_in = rw.DeviceArray.from_numpy(k)
    d_out = rw.DeviceArray(n, np.uint32)
    ws = rw.Workspace(n)
    for v in range(L.rwh_sort_variant_count()):
        print(f"case sort shape {v} n={n}", flush=True)
        ws.set_sort_variant(v)
        rw.check(L.rwh_sort_u32(ws.ptr, d_in.ptr, d_out.ptr, n, 30, None))
        assert np.array_equal(d_out.numpy(), np.sort(k)), v


def test_sort_every_shape_full_size(gpu):
    """At 2^24 keys every tile shape has more tiles than resident workgroups,
    so each workgroup of a pass sorts several tiles in turn."""
    rw = gpu
    L = rw.lib()
    n = 1 << 24
    rng = np.random.default_rng(24)
    k = rng.integers(0, 1 << 30, size=n, dtype=np.uint64).astype(np.uint32)
    want = np.sort(k)
    d_in = rw.DeviceArray.from_numpy(k)
    d_out = rw.DeviceArray(n, np.uint32)
    ws = rw.Workspace(n)
    for v in range(L.rwh_sort_variant_count()):
        print(f"case sort shape {v} n={n}", flush=True)
        ws.set_sort_variant(v)
        rw.check(L.rwh_sort_u32(ws.ptr, d_in.ptr, d_out.ptr, n, 30, None))
        assert np.array_equal(d_out.numpy(), want), v


@pytest.mark.parametrize("order", ["sorted", "reversed", "runs"])
def test_sort_ordered_keys(gpu, order):
    """Already ordered keys: whole waves share a digit in the histogram sweep."""
    rw = gpu
    L = rw.lib()
    n = (1 << 22) + 77
    rng = np.random.default_rng(3)
    k = np.sort(rng.integers(0, 1 << 30, size=n, dtype=np.uint64).astype(np.uint32))
    if order == "reversed":
        k = k[::-1].copy()
    elif order == "runs":
        k = np.repeat(k[: n // 1000], 1000)[:n]
        k = np.concatenate([k, np.zeros(n - k.size, np.uint32)])
    d_in = rw.DeviceArray.from_numpy(k)
    d_out = rw.DeviceArray(n, np.uint32)
    ws = rw.Workspace(n)
    print(f"case sort ordered {order} n={n}", flush=True)
    rw.check(L.rwh_sort_u32(ws.ptr, d_in.ptr, d_out.ptr, n, 30, None))
    assert np.array_equal(d_out.numpy(), np.sort(k))


def test_build_morton_ordered_points(gpu, orc, pool):
    """Points already in Morton order (codes arrive sorted at the sweep)."""
    rw = gpu
    n = 1 << 21
    p = rw.uniform_points(n, 41)
    codes = orc.morton(pool, p)
    p = np.ascontiguousarray(p[np.argsort(codes, kind="stable")])
    print(f"case build morton-ordered n={n}", flush=True)
    g_counts, g = rw.build_octree(p, strict=False)
    (nu, nb, no, rc), o, _ = orc.build(pool, p)
    compare(rw, n, g_counts, g, (nu, nb, no), o, "morton-ordered")


def test_batch_sorts_many_tiles(gpu, orc, pool):
    """A batch large enough that its sort, compaction and scan launches loop
    over tiles: every cloud's sorted and unique codes equal the oracle's."""
    rw = gpu
    rng = np.random.default_rng(12)
    sizes = rng.integers(20000, 120000, size=150)
    clouds = [rw.uniform_points(int(n), 500 + s) if s % 3 else rw.gmm_points(int(n), 500 + s, 16, 10.24)
              for s, n in enumerate(sizes)]
    print(f"case batch many clouds={len(clouds)} points={int(sizes.sum())}", flush=True)
    b = rw.Batch(clouds)
    b.run()
    res = b.read()
    for s, c in enumerate(clouds):
        codes = np.sort(orc.morton(pool, c))
        u = np.unique(codes)
        assert np.array_equal(res[s]["morton_keys_alt"], codes), s
        assert res[s]["n_unique"] == u.size, s
        assert np.array_equal(res[s]["unique"], u), s


# sizes at the edges of the 16384-key compaction tiles, too
@pytest.mark.parametrize("n", [0, 1, 2, 4096, 4097, 16383, 16384, 16385, 49153, 100000])
def test_unique_u32(gpu, n):
    rw = gpu
    L = rw.lib()
    rng = np.random.default_rng(n)
    for hi in (1, 3, 1 << 14, 1 << 30):
        k = np.sort(rng.integers(0, hi, size=n).astype(np.uint32))
        d_in = rw.DeviceArray.from_numpy(k) if n else rw.DeviceArray(1, np.uint32)
        d_out = rw.DeviceArray(max(n, 1), np.uint32)
        ws = rw.Workspace(max(n, 1))
        rw.check(L.rwh_unique_u32(ws.ptr, d_in.ptr, n, d_out.ptr, ws.n_unique_ptr(), None))
        nu, _ = ws.counts()
        want = np.unique(k)
        assert int(nu[0]) == want.size
        assert np.array_equal(d_out.numpy(want.size), want)


def radix_keys(rng, kind, n):
    if kind == "k30":  # Morton-range keys
        k = rng.integers(0, 1 << 30, size=n, dtype=np.uint64)
    elif kind == "k32":  # bit 31 set on some keys: the reference search path
        k = rng.integers(0, 1 << 32, size=n, dtype=np.uint64)
    elif kind == "dense":  # consecutive runs: long chains of equal deltas
        k = np.arange(n, dtype=np.uint64) * 3 + (rng.integers(0, 3, size=n) == 0)
    elif kind == "lastfar":
        k = rng.integers(0, 1 << 20, size=n, dtype=np.uint64)
        k[-1] = (1 << 30) - 1
    else:  # firstfar
        k = rng.integers(0, 1 << 20, size=n, dtype=np.uint64) | (1 << 29)
        k[0] = 0
    return np.unique(k).astype(np.uint32)


@pytest.mark.parametrize("n", [2, 3, 5, 64, 1000, 5000, 70000, 600000])
@pytest.mark.parametrize("kind", ["k30", "k32", "dense", "lastfar", "firstfar"])
def test_radix_tree_stage(gpu, orc, pool, kind, n):
    """S4 alone on sorted unique keys (rwh_radix_tree), against the oracle."""
    rw = gpu
    L = rw.lib()
    rng = np.random.default_rng(n * 7 + len(kind))
    k = radix_keys(rng, kind, n)
    if k.size < 2:
        pytest.skip("fewer than two unique keys")
    nu = k.size
    d_k = rw.DeviceArray.from_numpy(k)
    d_nu = rw.DeviceArray.from_numpy(np.array([nu], np.uint32))
    out = {"prefix_n": rw.DeviceArray(nu, np.uint8), "has_leaf_left": rw.DeviceArray(nu, np.uint8),
           "has_leaf_right": rw.DeviceArray(nu, np.uint8), "left_child": rw.DeviceArray(nu, np.int32),
           "parents": rw.DeviceArray(nu, np.int32, zero=True)}
    ws = rw.Workspace(nu)
    rw.check(L.rwh_radix_tree(ws.ptr, d_k.ptr, d_nu.ptr, nu, out["prefix_n"].ptr, out["has_leaf_left"].ptr,
                              out["has_leaf_right"].ptr, out["left_child"].ptr, out["parents"].ptr,
                              None))
    rw.check(L.rwh_stream_sync(None))
    want = orc.radix_tree(pool, k)
    for name, arr in out.items():
        assert np.array_equal(arr.numpy(nu - 1), want[name]), name


def exact_keys(rng, kind, nu):
    """nu sorted unique keys below 2^30."""
    if kind == "k30":
        k = rng.choice(1 << 30, size=nu, replace=False)
    elif kind == "stride":  # a regular tree: every split at a power of two
        k = np.arange(nu, dtype=np.uint64) << 7
    else:  # clustered: 16 tight clusters, long chains of shared prefixes
        centers = rng.choice(1 << 10, size=16, replace=False).astype(np.uint64) << 20
        per = -(-nu // 16)
        k = np.concatenate([c + rng.choice(1 << 18, size=per, replace=False) for c in centers])
        k = np.unique(k)[:nu]
    return np.sort(np.asarray(k, np.uint64)).astype(np.uint32)


@pytest.mark.parametrize("nu", [4095, 4096, 4097, 4098, 8191, 8192, 8193, 12289, 40961, 2 ** 20 + 1])
@pytest.mark.parametrize("kind", ["k30", "stride", "clustered"])
def test_radix_tree_tile_edges(gpu, orc, pool, kind, nu):
    """S4 at unique-key counts around multiples of the tree tile (the last tile
    takes the extra leaf; nodes cross tile edges), against the oracle."""
    rw = gpu
    L = rw.lib()
    k = exact_keys(np.random.default_rng(nu + len(kind)), kind, nu)
    assert k.size == nu and np.all(np.diff(k.astype(np.int64)) > 0)
    d_k = rw.DeviceArray.from_numpy(k)
    d_nu = rw.DeviceArray.from_numpy(np.array([nu], np.uint32))
    ws = rw.Workspace(nu)
    out = {"prefix_n": rw.DeviceArray(nu, np.uint8), "has_leaf_left": rw.DeviceArray(nu, np.uint8),
           "has_leaf_right": rw.DeviceArray(nu, np.uint8), "left_child": rw.DeviceArray(nu, np.int32),
           "parents": rw.DeviceArray(nu, np.int32)}
    for _ in range(2):  # a second call over the first's outputs and lists
        rw.check(L.rwh_radix_tree(ws.ptr, d_k.ptr, d_nu.ptr, nu, out["prefix_n"].ptr,
                                  out["has_leaf_left"].ptr, out["has_leaf_right"].ptr,
                                  out["left_child"].ptr, out["parents"].ptr, None))
        rw.check(L.rwh_stream_sync(None))
        want = orc.radix_tree(pool, k)
        for name, arr in out.items():
            assert np.array_equal(arr.numpy(nu - 1), want[name]), name


def test_repeated_builds_are_identical(gpu):
    rw = gpu
    pts = rw.gmm_points(300000, seed=9, n_centers=16, sigma=10.24)
    ws = rw.Workspace(pts.shape[0])
    first = None
    for _ in range(3):
        c, g = rw.build_octree(pts, ws=ws)
        flat = [c] + [g[k].tobytes() for k in sorted(g)]
        if first is None:
            first = flat
        assert flat == first


def test_dropin_path_matches_oracle(gpu, orc, pool):
    rw = gpu
    pts = rw.uniform_points(200000, seed=5)
    print("case dropin", flush=True)
    g_counts, g = rw.dropin_build(pts)
    (nu, nb, no, rc), o, _ = orc.build(pool, pts)
    compare(rw, pts.shape[0], g_counts, g, (nu, nb, no), o, "dropin")


@pytest.mark.parametrize("stage", [1, 3, 6, 7])
def test_dropin_failure_surfaces_in_future(gpu, stage):
    rw = gpu
    pts = rw.uniform_points(5000, seed=1)
    with pytest.raises(rw.RwhError) as e:
        rw.dropin_build(pts, fail_stage=stage)
    assert e.value.status == stage
    assert "injected" in str(e.value)


@pytest.mark.parametrize("shape", [None, 5])
def test_batch_equals_oracle_per_cloud(gpu, orc, pool, shape):
    """shape 5 (1024 x 16, a tile of two chunks) is chosen for the workspace:
    the batched sort must fall back to a chunk-sized tile."""
    rw = gpu
    sizes = [0, 1, 2, 3, 1000, 8191, 8192, 8193, 30000, 70000, 5, 150000]
    clouds = []
    for s, n in enumerate(sizes):
        clouds.append(rw.uniform_points(n, 100 + s) if s % 2 == 0 else
                      rw.gmm_points(n, 100 + s, 16, 10.24))
    print(f"case batch shape={shape}", flush=True)
    b = rw.Batch(clouds)
    if shape is not None:
        b.ws.set_sort_variant(shape)
    b.run()
    res = b.read()
    for s, c in enumerate(clouds):
        (nu, nb, no, rc), o, _ = orc.build(pool, c)
        assert rc == (-1 if no > rw.octree_capacity(c.shape[0]) else 0)
        r = res[s]
        g = dict(r, morton_keys=r["unique"])
        compare(rw, c.shape[0], (r["n_unique"], max(r["n_unique"] - 1, 0), r["n_oct"]), g,
                (nu, nb, no), o, f"cloud{s}")


def test_batch_graph_replays_match_eager_and_oracle(gpu, orc, pool):
    """The batched build captured as a HIP graph (rwh_batch_graph_create) and
    replayed three times, with the program's own eager work between the
    replays (every output array and the per-cloud counts and status words
    poisoned by eager memsets, then a plain build of another cloud on the same
    stream): every replay's arrays and counts equal the eager batched build's
    (rwh_batch_prepare + rwh_build_octree_batch_enqueue), which equals the
    oracle cloud by cloud."""
    import ctypes as C
    rw = gpu
    L = rw.lib()
    sizes = [1000, 8193, 30000, 5, 70000, 0, 150000]
    clouds = [rw.uniform_points(n, 300 + s) if s % 2 == 0 else rw.gmm_points(n, 300 + s, 16, 10.24)
              for s, n in enumerate(sizes)]
    k = len(clouds)
    b = rw.Batch(clouds)
    st = rw.Stream()
    print("case batch graph eager", flush=True)
    rw.check(L.rwh_batch_prepare(b.ws.ptr, b.sizes.ctypes.data, b.pb.ctypes.data, b.ob.ctypes.data,
                                 b.oc.ctypes.data, k, st.ptr), "rwh_batch_prepare")
    rw.check(L.rwh_build_octree_batch_enqueue(b.ws.ptr, b.tb.ptr(), 0.0, 1024.0, st.ptr), "enqueue")
    st.sync()
    eager = b.read()
    for s, c in enumerate(clouds):
        (nu, nb, no, rc), o, _ = orc.build(pool, c)
        r = eager[s]
        compare(rw, c.shape[0], (r["n_unique"], max(r["n_unique"] - 1, 0), r["n_oct"]),
                dict(r, morton_keys=r["unique"]), (nu, nb, no), o, f"eager cloud{s}")
    g = C.c_void_p()
    rw.check(L.rwh_batch_graph_create(b.ws.ptr, b.tb.ptr(), 0.0, 1024.0, st.ptr, C.byref(g)), "graph")
    other = rw.uniform_points(20000, 9)
    try:
        for rep in range(3):
            print(f"case batch graph replay {rep}", flush=True)
            for name in rw.TREE_FIELDS:
                a = b.tb[name]
                rw.check(L.rwh_memset(a.ptr, 0x5A, a.nbytes, st.ptr), "poison")
            rw.check(L.rwh_memset(b.ws.counts_ptr, 0x5A, 4 * 3 * k, st.ptr), "poison counts")
            rw.build_octree(other, stream=st.ptr)
            rw.check(L.rwh_graph_launch(g, st.ptr), "rwh_graph_launch")
            st.sync()
            res = b.read()
            for s in range(k):
                for key, v in eager[s].items():
                    got = res[s][key]
                    same = got == v if np.isscalar(v) else np.array_equal(bits(got), bits(v))
                    assert same, f"replay {rep} cloud {s} {key}"
    finally:
        rw.check(L.rwh_graph_destroy(g), "rwh_graph_destroy")


def distinct_cells(n, seed):
    """n points in distinct cells of the 1024^3 lattice (cell centres): every
    Morton code differs, so n_unique = n and the radix tree has n - 1 nodes."""
    rng = np.random.default_rng(seed)
    c = rng.choice(1 << 30, size=n, replace=False).astype(np.uint64)
    p = np.empty((n, 4), np.float32)
    p[:, 0] = (c & 1023).astype(np.float32) + 0.5
    p[:, 1] = ((c >> 10) & 1023).astype(np.float32) + 0.5
    p[:, 2] = (c >> 20).astype(np.float32) + 0.5
    p[:, 3] = 0.0
    return p


# radix-tree node counts at the edges of the tree tiles (4096 nodes) and of
# the scan tiles (8192 and 16384), where the per-tile edge-count sums meet
TILE_EDGE_NODES = [4095, 4096, 4097, 8191, 8192, 8193, 12289, 16383, 16384, 16385, (1 << 20) + 1]


@pytest.mark.parametrize("nb", TILE_EDGE_NODES)
def test_build_tile_edge_sizes(gpu, orc, pool, nb):
    rw = gpu
    pts = distinct_cells(nb + 1, seed=nb)
    print(f"case build tile edges n_brt={nb}", flush=True)
    g_counts, g = rw.build_octree(pts, strict=False)
    assert g_counts[0] == nb + 1
    (nu, nb_o, no, rc), o, _ = orc.build(pool, pts)
    assert rc == (-1 if no > rw.octree_capacity(nb + 1) else 0)
    compare(rw, nb + 1, g_counts, g, (nu, nb_o, no), o, f"edges/{nb}")


def test_batch_tile_edge_sizes(gpu, orc, pool):
    rw = gpu
    clouds = [distinct_cells(nb + 1, seed=7 * nb) for nb in TILE_EDGE_NODES[:-1]]
    print("case batch tile edges", flush=True)
    b = rw.Batch(clouds)
    b.run()
    res = b.read()
    for s, c in enumerate(clouds):
        (nu, nb, no, rc), o, _ = orc.build(pool, c)
        r = res[s]
        assert r["n_unique"] == c.shape[0]
        g = dict(r, morton_keys=r["unique"])
        compare(rw, c.shape[0], (r["n_unique"], max(r["n_unique"] - 1, 0), r["n_oct"]), g,
                (nu, nb, no), o, f"cloud{s}")


def test_batch_repeat_and_reuse(gpu):
    """The same workspace serves batches of different shapes back to back."""
    rw = gpu
    ws = rw.Workspace(1 << 21, 64)
    for rep, sizes in enumerate(([5000, 70000, 12], [100000] * 10, [1, 2, 3])):
        clouds = [rw.uniform_points(n, rep * 100 + i) for i, n in enumerate(sizes)]
        b = rw.Batch(clouds, ws=ws)
        b.run()
        res = b.read()
        for s, c in enumerate(clouds):
            g_counts, g = rw.build_octree(c)
            assert res[s]["n_unique"] == g_counts[0] and res[s]["n_oct"] == g_counts[2]
            assert np.array_equal(res[s]["parents"], g["parents"][:max(g_counts[0] - 1, 0)])
            rows = min(g_counts[2], rw.octree_capacity(c.shape[0]))
            assert np.array_equal(bits(res[s]["children"][:rows]), bits(g["children"][:rows]))


@pytest.mark.parametrize("kind,n", [("uniform", 50000), ("gmm", 200000), ("dups", 30000)])
def test_link_leaf_matches_oracle(gpu, orc, pool, kind, n):
    """Opt-in leaf linking (the reference's never-launched k_LinkLeafNodes)."""
    rw = gpu
    pts = inputs(rw, kind, n, seed=n)
    g_counts, g = rw.build_octree(pts, link_leaf=True)
    (nu, nb, no, rc), o, _ = orc.build(pool, pts)
    before = o["children"].copy()
    orc.link_leaf(pool, o, (nu, nb, no))
    rows = min(no, rw.octree_capacity(n))
    assert not np.array_equal(before[:rows], o["children"][:rows])  # leaves were linked
    compare(rw, n, g_counts, g, (nu, nb, no), o, f"leaf {kind}/{n}")
