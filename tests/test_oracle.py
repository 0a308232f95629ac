"""Pins the CPU oracle (oracle/tree_oracle.cpp) before it is trusted.

Each stage is checked against something that is not the oracle's own formula:
known answers and an independent formulation (S1), numpy and the reference's
own func_sort.cpp compiled from source (S2), numpy.unique (S3), a brute-force
top-down construction and the reference's own func_brt.hpp (S4), the
reference's func_edge.hpp and a closed form (S5), numpy.cumsum (S6), and
structural invariants with closed forms (S7).
"""
import numpy as np
import pytest


# ---------------------------------------------------------------------------
# helpers: an independent Morton formulation (the reference's magic numbers,
# func_morton.hpp:15-22, vectorised) and a brute-force radix tree
# ---------------------------------------------------------------------------
def spread_magic(v):
    x = v.astype(np.uint64) & 0x3FF
    x = (x | (x << 16)) & 0x30000FF
    x = (x | (x << 8)) & 0x0300F00F
    x = (x | (x << 4)) & 0x30C30C3
    x = (x | (x << 2)) & 0x9249249
    return x.astype(np.uint32)


def morton_numpy(pts, min_coord=np.float32(0), range_=np.float32(1024)):
    q = []
    for a in range(3):
        t = (pts[:, a] - min_coord) / range_
        s = (t * np.float32(1024)).astype(np.float32)
        q.append(np.where(s > 0, np.floor(s), 0).astype(np.int64).astype(np.uint32))
    return spread_magic(q[0]) | (spread_magic(q[1]) << 1) | (spread_magic(q[2]) << 2)


def clz32(x):
    return 32 if x == 0 else 32 - int(x).bit_length()


def delta8(a, b):
    return (clz32(int(a) ^ int(b)) - 1) & 0xFF


def split_of(c, a, b):
    """Last index of [a, b] whose key has a 0 at the highest bit where c[a], c[b] differ."""
    if a == b:
        return a
    bit = 31 - clz32(int(c[a]) ^ int(c[b]))
    g = a
    for k in range(a, b + 1):
        if not (int(c[k]) >> bit) & 1:
            g = k
    return g


def brute_tree(c):
    """Radix tree over unique sorted keys c by top-down recursion (a second
    construction of the Karras tree), with the reference's quirks: node 0
    describes the range [0, n-1] (Q2), delta = uint8(clz - 1) (Q1), and a
    parent slot keeps the largest index that claims it (D7)."""
    n = len(c) - 1  # brt nodes
    P = np.zeros(n, np.uint8)
    HL = np.zeros(n, np.uint8)
    HR = np.zeros(n, np.uint8)
    LC = np.zeros(n, np.int32)
    claims = {}

    def node(idx, a, b):
        g = split_of(c, a, b)
        P[idx] = delta8(c[a], c[b])
        LC[idx] = g
        HL[idx] = 1 if a == g else 0
        HR[idx] = 1 if b == g + 1 else 0
        return g

    def rec(idx, a, b):  # the complete tree below the root; fills nodes != 0
        g = split_of(c, a, b)
        if idx != 0:
            node(idx, a, b)
        if g != a:
            claims.setdefault(g, []).append(idx)
            rec(g, a, g)
        if g + 1 != b:
            claims.setdefault(g + 1, []).append(idx)
            rec(g + 1, g + 1, b)

    rec(0, 0, n)
    # Q2 root over [0, n-1]
    g0 = node(0, 0, n - 1) if n - 1 > 0 else 0
    if n - 1 == 0:
        P[0] = delta8(c[0], c[0])
        LC[0] = 0
        HL[0] = 1
        HR[0] = 0
    if not HL[0]:
        claims.setdefault(int(LC[0]), []).append(0)
    if not HR[0]:
        claims.setdefault(int(LC[0]) + 1, []).append(0)
    par = np.zeros(max(n, 1), np.int32)
    for k, w in claims.items():
        if k < n:
            par[k] = max(w)
    return {"prefix_n": P, "has_leaf_left": HL, "has_leaf_right": HR, "left_child": LC,
            "parents": par[:n]}


def unique_keys(rng, n, bits=30):
    while True:
        k = np.unique(rng.integers(0, 1 << bits, size=n, dtype=np.uint64)).astype(np.uint32)
        if k.size >= 2:
            return k


# ---------------------------------------------------------------------------
# S1 Morton
# ---------------------------------------------------------------------------
def test_morton_known_answers(orc, pool):
    pts = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [1023, 1023, 1023, 0],
                    [1023.99994, 0, 0, 0], [1024, 0, 0, 0], [0.5, 0.99, 1.0, 0],
                    [512, 0, 0, 0], [0, 512, 0, 0], [0, 0, 512, 0]], np.float32)
    got = orc.morton(pool, pts)
    want = [1, 2, 4, (1 << 30) - 1, int(spread_magic(np.array([1023]))[0]), 0, 4,
            1 << 27, 1 << 28, 1 << 29]
    assert got.tolist() == want


def test_morton_matches_magic_numbers(orc, pool):
    rng = np.random.default_rng(1)
    pts = (rng.random((50000, 4)) * 1024).astype(np.float32)
    assert np.array_equal(orc.morton(pool, pts), morton_numpy(pts))
    # a non power-of-two box exercises the float sub/div/mul rounding
    mn, rg = np.float32(-3.25), np.float32(777.7)
    pts2 = (mn + rng.random((50000, 4)) * rg).astype(np.float32)
    assert np.array_equal(orc.morton(pool, pts2, float(mn), float(rg)), morton_numpy(pts2, mn, rg))


def test_pow2_range_product_equals_quotient():
    # the encode multiplies by 1/range when range is a power of two with a
    # normal reciprocal: the same real number, so the same IEEE float as the
    # quotient, for every dividend (subnormal, huge, infinite, NaN included)
    rng = np.random.default_rng(3)
    bits_ = rng.integers(0, 1 << 32, size=400000, dtype=np.uint64).astype(np.uint32)
    d = bits_.view(np.float32)
    d = np.concatenate([d, np.float32([0.0, -0.0, np.inf, -np.inf, 1e-45, 3e-39, 1e38, 1023.999])])
    with np.errstate(all="ignore"):
        for e in (-126, -20, -3, -1, 0, 1, 3, 10, 20, 126):
            r = np.float32(2.0 ** e)
            inv = np.float32(2.0 ** -e)
            q, m = d / r, d * inv
            same = (q.view(np.uint32) == m.view(np.uint32)) | (np.isnan(q) & np.isnan(m))
            assert same.all(), e
        # and a non power of two does not have the property (so the encode divides)
        r = np.float32(777.7)
        assert ((d / r) != (d * (np.float32(1) / r)))[np.isfinite(d)].any()


def test_morton_saturation_out_of_box(orc, pool):
    # below the box and NaN quantise to 0; far above saturates then keeps 10 bits (D10)
    pts = np.array([[-5, 0, 0, 0], [np.nan, 0, 0, 0], [1e12, 0, 0, 0]], np.float32)
    assert orc.morton(pool, pts).tolist() == [0, 0, int(spread_magic(np.array([0x3FF]))[0])]


# ---------------------------------------------------------------------------
# S2 sort, S3 unique
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 7, 1000, 307200])
def test_sort_matches_numpy(orc, pool, n):
    rng = np.random.default_rng(n)
    k = rng.integers(0, 1 << 32, size=n, dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(orc.sort(pool, k), np.sort(k))


def test_sort_matches_reference_source(orc, pool):
    if not orc.ref_available():
        pytest.skip("reference CPU kernels not built (no /root/reference here)")
    rng = np.random.default_rng(5)
    k = rng.integers(0, 1 << 30, size=100003, dtype=np.uint64).astype(np.uint32)
    keys = k.copy()
    alt = np.zeros_like(k)
    orc.ref().ref_sort(keys.ctypes.data, alt.ctypes.data, k.size, 4)
    assert np.array_equal(orc.sort(pool, k), keys)


def test_unique_matches_numpy(orc):
    rng = np.random.default_rng(3)
    for n in (0, 1, 2, 1000, 100000):
        k = np.sort(rng.integers(0, max(n // 3, 1), size=n).astype(np.uint32))
        assert np.array_equal(orc.unique(k), np.unique(k))


# ---------------------------------------------------------------------------
# S4 radix tree
# ---------------------------------------------------------------------------
def test_radix_tree_brute_force_tiny(orc, pool):
    rng = np.random.default_rng(11)
    cases = [np.array([0, 1], np.uint32), np.array([0, 1, 2], np.uint32),
             np.array([0b000, 0b001, 0b100, 0b101], np.uint32),
             np.array([1, 2, 3, 1 << 29], np.uint32)]
    for _ in range(400):
        n = int(rng.integers(2, 65))
        bits = int(rng.choice([4, 8, 12, 30]))
        if (1 << bits) < n:
            bits = 12
        cases.append(np.sort(rng.choice(1 << bits, size=n, replace=False)).astype(np.uint32))
    for c in cases:
        got = orc.radix_tree(pool, c)
        want = brute_tree(c)
        for k in want:
            assert np.array_equal(got[k], want[k]), (k, c.tolist())


def tree_from_adjacent_deltas(c):
    """Third construction of the radix tree, the one the GPU uses for keys
    below 2^31: with a[k] = delta8(c[k], c[k+1]), the node split at k covers
    keys [psv(k) + 1, nsv(k)] (nearest positions with a smaller a on the left /
    right), has prefix a[k], and is numbered by its endpoint next to the larger
    of a[psv], a[nsv]; node 0 reports the range [0, n-1] (Q2)."""
    c = [int(x) for x in c]
    n = len(c) - 1
    P = np.zeros(n, np.uint8)
    HL = np.zeros(n, np.uint8)
    HR = np.zeros(n, np.uint8)
    LC = np.zeros(n, np.int32)
    par = np.zeros(n, np.int32)
    if n == 1:
        P[0], HL[0] = delta8(c[0], c[0]), 1
        return {"prefix_n": P, "has_leaf_left": HL, "has_leaf_right": HR, "left_child": LC,
                "parents": par}
    a = [delta8(c[k], c[k + 1]) for k in range(n)]

    def put(i, v, k, ll, lr):
        P[i], LC[i], HL[i], HR[i] = v, k, int(ll), int(lr)

    for k in range(n):
        v = a[k]
        psv = next((q for q in range(k - 1, -1, -1) if a[q] < v), -1)
        nsv = next((q for q in range(k + 1, n) if a[q] < v), n)
        lo, hi = psv + 1, nsv
        idx = 0
        if psv >= 0 or nsv < n:
            ap = a[psv] if psv >= 0 else -1
            an = a[nsv] if nsv < n else -1
            idx = lo if ap > an else hi
            put(idx, v, k, lo == k, hi == k + 1)
        if psv == -1 and (nsv == n - 1 or (nsv == n and k < n - 1)):
            put(0, v, k, k == 0, k + 1 == n - 1)
        if lo != k:
            par[k] = idx
        if hi != k + 1:
            par[k + 1] = idx
    return {"prefix_n": P, "has_leaf_left": HL, "has_leaf_right": HR, "left_child": LC,
            "parents": par}


def test_radix_tree_adjacent_delta_formulation(orc, pool):
    rng = np.random.default_rng(14)
    cases = [np.array([0, 1], np.uint32), np.array([0, 1, 2], np.uint32),
             np.array([1, 2, 3, 1 << 29], np.uint32), np.array([0, 1 << 29, (1 << 29) + 1], np.uint32)]
    for s in range(1500):
        n = int(rng.integers(2, 90))
        bits = int(rng.choice([4, 8, 12, 30]))
        if (1 << bits) < n:
            bits = 12
        c = np.sort(rng.choice(1 << bits, size=n, replace=False)).astype(np.uint32)
        if s % 5 == 0:
            c[-1] = (1 << 30) - 1  # an isolated last key
        elif s % 5 == 1:
            c = np.concatenate([[0], c[1:] | (1 << 29)]).astype(np.uint32)  # isolated first key
        cases.append(np.unique(c))
    for c in cases:
        if c.size < 2:
            continue
        got = tree_from_adjacent_deltas(c)
        want = orc.radix_tree(pool, c)
        for k in want:
            assert np.array_equal(got[k], want[k]), (k, c.tolist())


def test_radix_tree_matches_reference_source(orc, pool):
    if not orc.ref_available():
        pytest.skip("reference CPU kernels not built (no /root/reference here)")
    rng = np.random.default_rng(12)
    for n in (3, 17, 1000, 200000):
        c = unique_keys(rng, n)
        nu = c.size
        r = {k: np.zeros(nu, dt) for k, dt in (("prefix_n", np.uint8), ("has_leaf_left", np.uint8),
                                                ("has_leaf_right", np.uint8), ("left_child", np.int32),
                                                ("parents", np.int32))}
        orc.ref().ref_radix_tree(c.ctypes.data, nu, r["prefix_n"].ctypes.data,
                                 r["has_leaf_left"].ctypes.data, r["has_leaf_right"].ctypes.data,
                                 r["left_child"].ctypes.data, r["parents"].ctypes.data)
        got = orc.radix_tree(pool, c)
        for k in got:
            assert np.array_equal(got[k], r[k][:nu - 1]), k


def test_radix_tree_invariants(orc, pool):
    rng = np.random.default_rng(13)
    c = unique_keys(rng, 50000)
    t = orc.radix_tree(pool, c)
    n = c.size - 1
    par = t["parents"]
    # a parent always describes a shorter common prefix than its child
    assert np.all(t["prefix_n"][par[1:]] <= t["prefix_n"][1:])
    assert par[0] == 0 and np.all((par >= 0) & (par < n))
    assert np.all((t["left_child"] >= 0) & (t["left_child"] < n))


# ---------------------------------------------------------------------------
# S5 edge count, S6 prefix sum
# ---------------------------------------------------------------------------
def test_edge_count_closed_form_and_reference(orc, pool):
    rng = np.random.default_rng(21)
    c = unique_keys(rng, 30000)
    t = orc.radix_tree(pool, c)
    got = orc.edge_count(pool, t["prefix_n"], t["parents"])
    want = t["prefix_n"].astype(np.int32) // 3 - t["prefix_n"][t["parents"]].astype(np.int32) // 3
    want[0] = 0
    assert np.array_equal(got, want)
    assert np.all(got >= 0)
    if orc.ref_available():
        r = np.zeros_like(got)
        orc.ref().ref_edge_count(t["prefix_n"].ctypes.data, t["parents"].ctypes.data, got.size,
                                 r.ctypes.data)
        assert np.array_equal(got, r)


def test_prefix_sum_matches_cumsum(orc):
    rng = np.random.default_rng(31)
    for n in (0, 1, 5, 100000):
        c = rng.integers(0, 11, size=n).astype(np.int32)
        off, total = orc.prefix_sum(c)
        assert total == int(c.sum())
        if n:
            assert np.array_equal(off, np.concatenate([[0], np.cumsum(c)[:-1]]).astype(np.int32))


# ---------------------------------------------------------------------------
# S7 octree: structural invariants with closed forms
# ---------------------------------------------------------------------------
def decode_numpy(code):
    d = []
    for a in range(3):
        v = 0
        for b in range(10):
            v |= ((int(code) >> (3 * b + a)) & 1) << b
        d.append(v)
    return d


def test_octree_invariants(orc, pool, rw):
    pts = rw.uniform_points(20000, seed=7)
    (nu, nb, no, rc), o, _ = orc.build(pool, pts)
    assert rc == 0
    codes = o["unique"][:nu]
    cnt = o["edge_count"][:nb]
    off = o["edge_offset"][:nb]
    pn = o["prefix_n"][:nb]
    root_level = int(pn[0]) // 3
    owner = np.full(no, -1)
    for i in range(1, nb):
        for j in range(cnt[i]):
            assert owner[off[i] + j] == -1
            owner[off[i] + j] = i
    assert np.all(owner >= 0)  # every slot has exactly one owner
    ch = o["children"][:no]
    mask = o["child_node_mask"][:no]
    # a child slot whose mask bit is clear was never linked (stays 0)
    for w in range(8):
        assert np.all(ch[(mask >> w) & 1 == 0, w] == 0)
    assert np.all(o["child_leaf_mask"][:no] == 0)  # Q6
    # corner = decoded prefix of the owner's code at the slot's level; cell closed form
    for s in range(no):
        i = owner[s]
        j = s - off[i]
        level = int(pn[i]) // 3 - j
        sh = (30 - 3 * level) & 31
        prefix = (int(codes[i]) >> sh) << sh
        x, y, z = decode_numpy(prefix)
        assert o["corner"][s].tolist() == [x * 1.0, y * 1.0, z * 1.0, 1.0]
        assert o["cell_size"][s] * (1 << (level - root_level)) == 1024.0


def slot_geometry(o, rows, root_level):
    """Level and 30-bit Morton prefix of every octree slot, read back from its
    corner and cell size (both pinned above by closed forms), for the box
    [0, 1024): cell = 1024 / 2^(level - root_level), corner = grid origin."""
    cell = o["cell_size"][:rows].astype(np.float64)
    k = np.log2(1024.0 / cell)
    assert np.all(k == np.round(k)), "cell sizes are powers of two of the box"
    level = root_level + np.round(k).astype(np.int64)
    g = o["corner"][:rows, :3]
    gi = g.astype(np.int64)
    assert np.all(g == gi), "corners are grid points"
    code = (spread_magic(gi[:, 0]).astype(np.int64) | spread_magic(gi[:, 1]).astype(np.int64) << 1 |
            spread_magic(gi[:, 2]).astype(np.int64) << 2)
    return level, code


def octant(code, level):
    """Octant of a cell (or key) at `level` within its parent: the three Morton
    bits of that level (the shift count taken mod 32 as the hardware does)."""
    return (int(code) >> ((30 - 3 * int(level)) & 31)) & 7


def contributing_ancestor(i, cnt, par):
    """Nearest node at or above i with a non-zero edge count, walking parent[]
    at most 31 steps (func_octree.hpp:83-92), for a walk that starts at i."""
    rt, steps = i, 0
    while cnt[rt] == 0:
        rt = par[rt]
        steps += 1
        if steps > 30:
            break
    return rt


def octree_clouds(rw):
    out = [rw.uniform_points(20000, 7), rw.gmm_points(30000, 3, 8, 5.0),
           rw.uniform_points(3000, 9)]
    p = rw.uniform_points(5000, 11)
    p[:, :3] = np.floor(p[:, :3] / 128) * 128 + 1  # dense lattice: few, long chains
    out.append(p)
    for kind in ("lastfar", "firstfar"):
        p = rw.uniform_points(4000, 13)
        p[:, :3] = p[:, :3] / 64.0 + (0.0 if kind == "lastfar" else 900.0)
        p[-1, :3] = 1000.0 if kind == "lastfar" else 0.5
        out.append(p)
    return out


def test_octree_links_follow_geometry(orc, pool, rw):
    """Pins S7's child links and octant assignment (func_octree.hpp:56-111).
    The expected rows are rebuilt from the pinned radix tree, edge counts and
    offsets, with each link's octant read from the CHILD slot's own corner and
    cell size (not from the oracle's shift of the radix node's code):
      * slot off+j+1 of a chain links slot off+j;
      * a chain's top links into the bottom slot of its nearest contributing
        ancestor (Q5 walk), at the top's octant;
      * collisions keep the largest child (D6), the mask ORs the octants.
    Geometry: a child is one level below its parent slot, and its cell lies
    inside the parent's cell except where the parent's prefix length is a
    multiple of 3 (Q3: the octree level takes one bit more than the keys
    share).  Slot 0 (Q4) is left out of the level checks."""
    checked = q3 = 0
    for pts in octree_clouds(rw):
        (nu, nb, no, rc), o, _ = orc.build(pool, pts)
        assert rc == 0 and nb > 1
        rows = no
        cnt, off = o["edge_count"][:nb], o["edge_offset"][:nb]
        par, pn = o["parents"][:nb], o["prefix_n"][:nb]
        root_level = int(pn[0]) // 3
        level, code = slot_geometry(o, rows, root_level)
        want = np.zeros((rows, 8), np.int64)
        want_mask = np.zeros(rows, np.int64)

        def link(parent_slot, child_slot, parent_node):
            nonlocal q3
            w = octant(code[child_slot], level[child_slot])
            want[parent_slot, w] = max(want[parent_slot, w], child_slot)
            want_mask[parent_slot] |= 1 << w
            if parent_slot != 0:
                assert level[child_slot] == level[parent_slot] + 1
                sh = (30 - 3 * int(level[parent_slot])) & 31
                if (code[child_slot] >> sh) != (code[parent_slot] >> sh):
                    assert int(pn[parent_node]) % 3 == 0, "outside the parent only under Q3"
                    q3 += 1

        for i in range(1, nb):
            for j in range(int(cnt[i]) - 1):
                link(off[i] + j + 1, off[i] + j, i)
            if cnt[i] > 0:
                t = contributing_ancestor(int(par[i]), cnt, par)
                link(int(off[t]), int(off[i]) + int(cnt[i]) - 1, t)
        assert np.array_equal(o["children"][:rows], want)
        assert np.array_equal(o["child_node_mask"][:rows], want_mask)
        checked += rows
    assert checked > 20000


def test_link_leaf_follows_geometry(orc, pool, rw):
    """Pins oracle_link_leaf (the reference's process_link_leaf,
    func_octree.hpp:114-163 / 07_octree.cu:116-173, never launched there).
    Every has_leaf flag of brt node i links key leaf = left_child + side into
    the bottom slot of the nearest contributing node at or above i (or is
    dropped when that slot lies outside the rows), at the octant of the
    leaf's key one level below the bottom slot's level read from its cell
    size; the child slot keeps the largest value (leaf indices and octree
    slots alike, D6), child_leaf_mask stays 0 (set_leaf clears the bit,
    func_octree.hpp:24).  The leaf's key lies in the bottom's cell except
    under Q3, and one level below it (slot 0 aside, Q4)."""
    links = q3 = 0
    for pts in octree_clouds(rw):
        (nu, nb, no, rc), o, _ = orc.build(pool, pts)
        assert rc == 0
        rows = no
        codes = o["unique"][:nu]
        cnt, off = o["edge_count"][:nb], o["edge_offset"][:nb]
        par, pn, lc = o["parents"][:nb], o["prefix_n"][:nb], o["left_child"][:nb]
        hl, hr = o["has_leaf_left"][:nb], o["has_leaf_right"][:nb]
        root_level = int(pn[0]) // 3
        level, code = slot_geometry(o, rows, root_level)
        want = o["children"][:rows].astype(np.int64).copy()
        dropped = 0
        for i in range(nb):
            for side, flag in ((0, hl[i]), (1, hr[i])):
                if not flag:
                    continue
                leaf = int(lc[i]) + side
                t = contributing_ancestor(i, cnt, par)
                b = int(off[t])
                if not 0 <= b < rows:
                    dropped += 1
                    continue
                w = octant(codes[leaf], level[b] + 1)
                want[b, w] = max(want[b, w], leaf)
                links += 1
                if b != 0:
                    assert level[b] == int(pn[i]) // 3  # the leaf sits one level below the bottom
                    sh = (30 - 3 * int(level[b])) & 31
                    if (int(codes[leaf]) >> sh) != (code[b] >> sh):
                        assert int(pn[t]) % 3 == 0, "outside the bottom's cell only under Q3"
                        q3 += 1
        n_flags = int(hl.sum()) + int(hr.sum())
        assert links <= n_flags + links and dropped <= n_flags
        orc.link_leaf(pool, o, (nu, nb, no))
        assert np.array_equal(o["children"][:rows], want)
        assert np.all(o["child_leaf_mask"][:rows] == 0)
        assert n_flags > 0
    assert links > 10000


def test_build_is_thread_count_invariant(orc, rw):
    pts = rw.gmm_points(40000, seed=3, n_centers=8, sigma=5.0)
    ref = None
    for t in (1, 3, 8):
        p = orc.Pool(t)
        counts, o, _ = orc.build(p, pts)
        flat = [counts] + [o[k].tobytes() for k in sorted(o) if k != "tmp"]
        if ref is None:
            ref = flat
        assert flat == ref


def pull_rows(o, nu, nb, no):
    """Second formulation of the octree links: every chain bottom pulls the
    chains reached from its radix-tree children through nodes of edge count 0
    (and slot 0 also pulls for node 0, the complete-range root).  Must equal the
    reference's push along parent[] (the oracle)."""
    codes = o["unique"][:nu].astype(np.int64)
    cnt = o["edge_count"][:nb].astype(np.int64)
    off = o["edge_offset"][:nb].astype(np.int64)
    pn = o["prefix_n"][:nb].astype(np.int64)
    lc = o["left_child"][:nb].astype(np.int64)
    hl, hr = o["has_leaf_left"][:nb], o["has_leaf_right"][:nb]
    n = nb
    rows = {}

    def link(c, row):
        top_level = pn[c] // 3 - cnt[c] + 1
        which = (int(codes[c]) >> ((30 - 3 * top_level) & 31)) & 7
        row[0][which] = max(row[0][which], off[c] + cnt[c] - 1)
        row[1] |= 1 << which

    def pull(stack, row, bounded):
        # stack entries: (node, depth below the pulling bottom; children are 1)
        while stack:
            q, d = stack.pop()
            if cnt[q] > 0:
                link(q, row)
                continue
            # below a node of count > 0, internal nodes three levels down always
            # have count > 0 (the GPU walks those three levels breadth-first)
            assert not (bounded and d >= 3), (q, d)
            if not hl[q]:
                stack.append((lc[q], d + 1))
            if not hr[q]:
                stack.append((lc[q] + 1, d + 1))

    def full_root_split():
        if n == 0:
            return 0
        bit = 31 - clz32(int(codes[0]) ^ int(codes[n]))
        return max(k for k in range(n + 1) if not (int(codes[k]) >> bit) & 1)

    for p in range(1, n):
        if cnt[p] <= 0 or off[p] >= no:
            continue
        row = [[0] * 8, 0]
        st = []
        if not hl[p]:
            st.append((lc[p], 1))
        if not hr[p]:
            st.append((lc[p] + 1, 1))
        pull(st, row, True)
        if off[p] == 0:
            g = full_root_split()
            st = ([(g, 1)] if g != 0 else []) + ([(g + 1, 1)] if g + 1 < n else [])
            pull(st, row, False)
        rows[off[p]] = row
    return rows


def test_octree_pull_equals_push(orc, pool, rw):
    """The GPU builds chain-bottom rows by pulling; the oracle pushes as the
    reference does.  Checked here on the CPU over random, clustered and
    adversarial (isolated first / last key) clouds."""
    rng = np.random.default_rng(77)
    clouds = []
    for s in range(60):
        n = int(rng.integers(2, 3000))
        kind = s % 5
        if kind == 0:
            p = rw.uniform_points(n, 1000 + s)
        elif kind == 1:
            p = rw.gmm_points(n, 1000 + s, 4, 3.0)
        elif kind == 2:
            p = rw.uniform_points(n, 1000 + s)
            p[:, :3] = np.floor(p[:, :3] / 128) * 128
        else:
            p = rw.uniform_points(n, 1000 + s)
            p[:, :3] = p[:, :3] / 64.0 + (0.0 if kind == 3 else 900.0)
            p[-1, :3] = 1000.0 if kind == 3 else 0.5
        clouds.append(p)
    checked = 0
    for p in clouds:
        (nu, nb, no, rc), o, _ = orc.build(pool, p)
        rows_n = min(no, o["cell_size"].shape[0])
        got = pull_rows(o, nu, nb, rows_n)
        # chain bottoms of the push result
        bottoms = {int(o["edge_offset"][i]) for i in range(1, nb)
                   if o["edge_count"][i] > 0 and o["edge_offset"][i] < rows_n}
        assert set(got) == bottoms
        for slot, (row, m) in got.items():
            assert o["children"][slot].tolist() == row, (slot, row)
            assert int(o["child_node_mask"][slot]) == m
            checked += 1
    assert checked > 1000
