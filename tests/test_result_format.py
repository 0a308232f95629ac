"""The packed result format (include/rwhip.h, version 1).

CPU: the packing kernel (csrc/k_result.hip) runs unmodified under the host
emulation of tests/cpp with AddressSanitizer, on sources holding exactly each
section's live bytes; its output is read back here with numpy against the
layout the header documents (a 256-byte header, sections in order at
multiples of 256 bytes, zero padding) and with the library's own reader
(rwh_result_check / rwh_result_section), which must also reject malformed
buffers.  GPU: packed results of real builds (single, batched cloud, the
allocation-free build, the C++ drop-in path) equal the oracle's outputs.
"""
import os
import subprocess

import numpy as np
import pytest

from parity_helpers import compare

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "cpp", "bin", "result_emul")

ALL, OCTREE = 0x3FFF, 0x3E00
ENTRY = [4, 4, 1, 1, 1, 4, 4, 4, 4, 32, 16, 4, 4, 4]
HDR = np.dtype([("magic", "<u4"), ("version", "<u4"), ("header_bytes", "<u4"), ("sections", "<u4"),
                ("status", "<u4"), ("flags", "<u4"), ("n_input", "<u4"), ("n_unique", "<u4"),
                ("n_brt", "<u4"), ("n_octree_nodes", "<u4"), ("oct_rows", "<u4"),
                ("oct_capacity", "<u4"), ("total_bytes", "<u8"), ("offset", "<u8", 14),
                ("count", "<u4", 14), ("entry_bytes", "<u2", 14), ("reserved", "<u4")])
assert HDR.itemsize == 256


def layout(n_input, n_unique, rows, sections):
    """Offsets, counts and total size as the header comment of rwhip.h states them."""
    nb = max(n_unique - 1, 0)
    counts = [n_input, n_unique] + [nb] * 7 + [rows] * 5
    at, off, cnt = 256, [0] * 14, [0] * 14
    for s in range(14):
        if (sections >> s) & 1:
            off[s], cnt[s] = at, counts[s]
            at = -(-(at + counts[s] * ENTRY[s]) // 256) * 256
    return off, cnt, at


def pattern(s, nbytes):
    j = np.arange(nbytes, dtype=np.uint64)
    return ((131 * j + 29 * s + 7) % 256).astype(np.uint8)


@pytest.fixture(scope="module")
def emul():
    if not os.path.exists(EMU):
        subprocess.run(["make", "-C", os.path.join(ROOT, "tests", "cpp"), "bin/result_emul"], check=True,
                       capture_output=True)
    return EMU


def run_emul(emul, tmp_path, n, nu, no, cap, sections, stride, status):
    out = tmp_path / "result.bin"
    r = subprocess.run([emul, str(n), str(nu), str(no), str(cap), str(sections), str(stride), str(status),
                        str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
    return np.fromfile(out, np.uint8)


CASES = [
    # n_input, n_unique, n_octree_nodes, capacity, sections, counts stride, status
    (1001, 777, 300, 600, ALL, 1, 0),
    (0, 0, 0, 0, ALL, 1, 0),
    (1, 1, 0, 0, ALL, 1, 0),
    (2, 2, 1, 1, ALL, 3, 0),
    (17, 5, 9, 10, ALL, 2, 0),
    (333, 300, 250, 199, ALL, 1, 0),            # octree above capacity: rows cut
    (4099, 4000, 2000, 2459, OCTREE, 1, 0x10),  # status word set
    (4099, 4000, 2000, 2459, (1 << 2) | (1 << 13), 5, 0),
    (1000, 1000, 0, 600, ALL, 1, 0),            # no octree node
    (70001, 65537, 30001, 42000, (1 << 0) | (1 << 3) | (1 << 10), 1, 0),
]


@pytest.mark.parametrize("case", CASES, ids=[f"n{c[0]}-u{c[1]}-o{c[2]}-s{c[4]:x}" for c in CASES])
def test_emulated_pack_follows_the_layout(rw, emul, tmp_path, case):
    n, nu, no, cap, sections, stride, status = case
    blob = run_emul(emul, tmp_path, *case)
    rows = min(no, cap)
    off, cnt, total = layout(n, nu, rows, sections)
    assert blob.size == total == rw.result_bytes(n, nu, no, cap, sections)
    h = blob[:256].view(HDR)[0]
    assert (h["magic"], h["version"], h["header_bytes"], h["sections"]) == (0x424F5752, 1, 256, sections)
    assert bytes(blob[:4]) == b"RWOB"
    assert (h["status"], h["flags"]) == (status, 1 if no > cap else 0)
    assert (h["n_input"], h["n_unique"], h["n_brt"], h["n_octree_nodes"], h["oct_rows"], h["oct_capacity"]) == \
        (n, nu, max(nu - 1, 0), no, rows, cap)
    assert h["total_bytes"] == total and h["reserved"] == 0
    assert list(h["offset"]) == off and list(h["count"]) == cnt
    assert list(h["entry_bytes"]) == [ENTRY[s] if (sections >> s) & 1 else 0 for s in range(14)]
    covered = np.zeros(total, bool)
    covered[:256] = True
    head, arrays = rw.unpack_result(blob)
    for s in range(14):
        if not (sections >> s) & 1:
            continue
        nbytes = cnt[s] * ENTRY[s]
        assert np.array_equal(blob[off[s]:off[s] + nbytes], pattern(s, nbytes)), s
        covered[off[s]:off[s] + nbytes] = True
        name = rw.RESULT_SECTIONS[s][0]
        assert arrays[name].view(np.uint8).size == nbytes
        assert np.array_equal(arrays[name].view(np.uint8).reshape(-1), pattern(s, nbytes)), name
    assert not blob[~covered].any()  # padding is zero
    want = rw.RWH_ERR_OVERFLOW if status else rw.RWH_ERR_CAPACITY if no > cap else rw.RWH_OK
    assert rw.result_check(blob) == want
    assert head["status"] == status and head["total_bytes"] == total


@pytest.fixture(scope="module")
def good_blob(emul, tmp_path_factory):
    return run_emul(emul, tmp_path_factory.mktemp("good"), 1001, 777, 300, 600, ALL, 1, 0)


def mutate(blob, field, value, index=None):
    b = blob.copy()
    h = b[:256].view(HDR)
    if index is None:
        h[0][field] = value
    else:
        h[0][field][index] = value
    return b


@pytest.mark.parametrize("field,value,index,message", [
    ("magic", 0x424F5753, None, "magic"),
    ("version", 2, None, "version 2"),
    ("header_bytes", 128, None, "header size"),
    ("reserved", 1, None, "reserved"),
    ("sections", ALL | (1 << 14), None, "sections"),
    ("sections", ALL & ~1, None, "total_bytes"),
    ("n_brt", 775, None, "counts disagree"),
    ("n_unique", 1002, None, "counts disagree"),
    ("flags", 1, None, "flags"),
    ("oct_rows", 299, None, "rows"),
    ("total_bytes", 0, None, "total_bytes"),
    ("count", 778, 1, "section 1"),
    ("offset", 512, 5, "section 5"),
    ("entry_bytes", 2, 2, "section 2"),
])
def test_reader_rejects_malformed_headers(rw, good_blob, field, value, index, message):
    assert rw.result_check(good_blob) == rw.RWH_OK
    bad = mutate(good_blob, field, value, index)
    assert rw.result_check(bad) == rw.RWH_ERR_INVALID_ARG
    assert message in rw.lib().rwh_last_error().decode()
    with pytest.raises(rw.RwhError):
        rw.unpack_result(bad)


def test_reader_rejects_short_buffers_and_absent_sections(rw, good_blob):
    assert rw.result_check(good_blob[:-1]) == rw.RWH_ERR_INVALID_ARG
    assert "shorter than total_bytes" in rw.lib().rwh_last_error().decode()
    assert rw.result_check(good_blob[:255]) == rw.RWH_ERR_INVALID_ARG
    # extra bytes after the result are allowed (a buffer sized for the largest result)
    assert rw.result_check(np.concatenate([good_blob, np.zeros(1000, np.uint8)])) == rw.RWH_OK
    import ctypes as C
    one = mutate(good_blob, "sections", 1 << 9)  # header says children only: layout then disagrees
    assert rw.result_check(one) == rw.RWH_ERR_INVALID_ARG
    p, cnt = C.c_void_p(), C.c_uint32()
    blob = np.ascontiguousarray(good_blob)
    assert rw.lib().rwh_result_section(blob.ctypes.data, blob.size, 14, C.byref(p), C.byref(cnt)) == \
        rw.RWH_ERR_INVALID_ARG
    assert rw.lib().rwh_result_section(blob.ctypes.data, blob.size, 9, C.byref(p), C.byref(cnt)) == rw.RWH_OK
    assert cnt.value == 300 and p.value - blob.ctypes.data == layout(1001, 777, 300, ALL)[0][9]


def test_status_set_result_still_reads(rw, emul, tmp_path):
    blob = run_emul(emul, tmp_path, 500, 400, 100, 300, ALL, 1, 0x8)
    assert rw.result_check(blob) == rw.RWH_ERR_OVERFLOW
    assert "keys not sorted" in rw.lib().rwh_last_error().decode()
    head, arrays = rw.unpack_result(blob)
    assert head["status"] == 0x8 and arrays["children"].shape == (100, 8)


# ---------------------------------------------------------------------------
# GPU: packed results of real builds against the oracle
# ---------------------------------------------------------------------------
def packed_as_outputs(arrays):
    return dict(arrays, morton_keys=arrays["unique"])


def header_counts(head):
    return head["n_unique"], head["n_brt"], head["n_octree_nodes"]


@pytest.mark.gpu
@pytest.mark.parametrize("n,kind", [(0, "u"), (1, "u"), (2, "u"), (3, "g"), (100, "u"), (500, "g"),
                                    (4097, "u"), (65537, "g"), (1 << 20, "u")])
def test_packed_build_matches_oracle(gpu, orc, pool, n, kind):
    rw = gpu
    p = rw.uniform_points(n, 900 + n) if kind == "u" else rw.gmm_points(n, 0 if n == 500 else 900 + n, 16, 10.24)
    print(f"case packed build n={n} kind={kind}", flush=True)
    cap = rw.octree_capacity(n)
    ws = rw.Workspace(max(n, 1))
    tb = rw.TreeBuild(n, cap, rw.DeviceArray.from_numpy(p.reshape(-1)))
    rw.check(rw.lib().rwh_build_octree(ws.ptr, tb.ptr(), n, 0.0, 1024.0, None))
    blob = rw.pack_result(tb, ws.counts_ptr, 1, n)
    head, arrays = rw.unpack_result(blob)
    (nu, nb, no, rc), o, _ = orc.build(pool, p)
    assert header_counts(head) == (nu, nb, no)
    assert blob.size == rw.result_bytes(n, nu, no, cap)
    assert rw.result_check(blob) == (rw.RWH_ERR_CAPACITY if no > cap else rw.RWH_OK)
    assert head["oct_rows"] == min(no, cap) and head["flags"] == (1 if no > cap else 0)
    compare(rw, n, header_counts(head), packed_as_outputs(arrays), (nu, nb, no), o, f"packed n={n}")
    # a subset of the sections carries the same bytes
    sub = rw.pack_result(tb, ws.counts_ptr, 1, n, sections=OCTREE | (1 << 2))
    h2, a2 = rw.unpack_result(sub)
    assert sub.size == rw.result_bytes(n, nu, no, cap, OCTREE | (1 << 2)) < blob.size or n == 0
    assert set(a2) == {"prefix_n", "children", "corner", "cell_size", "child_node_mask", "child_leaf_mask"}
    for k, v in a2.items():
        assert np.array_equal(v.view(np.uint8), arrays[k].view(np.uint8)), k


@pytest.mark.gpu
def test_packed_batch_clouds_match_oracle(gpu, orc, pool):
    rw = gpu
    sizes = [1, 5000, 0, 70000, 8192, 500, 33]
    clouds = [rw.uniform_points(n, 40 + i) if i % 2 == 0 else rw.gmm_points(n, 0 if n == 500 else 40 + i, 16, 10.24)
              for i, n in enumerate(sizes)]
    print(f"case packed batch clouds={len(clouds)}", flush=True)
    b = rw.Batch(clouds)
    b.run()
    for s, c in enumerate(clouds):
        n = c.shape[0]
        t = b.cloud_buffers(s)
        blob = rw.pack_result(b.tb, b.ws.counts_ptr + 4 * s, b.ws.max_segments, n, tree=t)
        head, arrays = rw.unpack_result(blob)
        (nu, nb, no, rc), o, _ = orc.build(pool, c)
        assert header_counts(head) == (nu, nb, no), s
        assert head["oct_capacity"] == int(b.oc[s])
        compare(rw, n, header_counts(head), packed_as_outputs(arrays), (nu, nb, no), o, f"cloud{s}")


@pytest.mark.gpu
def test_packed_noalloc_build_and_buffer_reuse(gpu, orc, pool):
    """The allocation-free build's 3-word counts (stride 1), packed twice into
    one device buffer, the second build smaller."""
    rw = gpu
    L = rw.lib()
    counts = rw.DeviceArray(3, np.uint32, zero=True)
    out = rw.DeviceArray(rw.result_bytes(300000, 300000, 180000, 180000), np.uint8)
    for n, seed in ((300000, 5), (12345, 6)):
        p = rw.uniform_points(n, seed)
        cap = rw.octree_capacity(n)
        tb = rw.TreeBuild(n, cap, rw.DeviceArray.from_numpy(p.reshape(-1)))
        print(f"case packed noalloc n={n}", flush=True)
        rw.check(L.rwh_build_octree_noalloc(tb.ptr(), n, 0.0, 1024.0, counts.ptr, None))
        blob = rw.pack_result(tb, counts.ptr, 1, n, out=out)
        head, arrays = rw.unpack_result(blob)
        (nu, nb, no, rc), o, _ = orc.build(pool, p)
        assert rw.result_check(blob) == rw.RWH_OK
        compare(rw, n, header_counts(head), packed_as_outputs(arrays), (nu, nb, no), o, f"noalloc n={n}")


@pytest.mark.gpu
def test_dropin_read_result_matches_oracle(gpu, orc, pool):
    rw = gpu
    for n, kind in ((7, "u"), (1000, "u"), (30000, "g")):
        p = rw.uniform_points(n, 3) if kind == "u" else rw.gmm_points(n, 3, 16, 10.24)
        print(f"case drop-in read_result n={n}", flush=True)
        blob = rw.dropin_packed(p)
        head, arrays = rw.unpack_result(blob)
        (nu, nb, no, rc), o, _ = orc.build(pool, p)
        assert rw.result_check(blob) == rw.RWH_OK
        compare(rw, n, header_counts(head), packed_as_outputs(arrays), (nu, nb, no), o, f"drop-in n={n}")


@pytest.mark.gpu
def test_pack_rejects_small_output_and_bad_sections(gpu):
    rw = gpu
    L = rw.lib()
    n = 1000
    tb = rw.TreeBuild(n, rw.octree_capacity(n))
    ws = rw.Workspace(n)
    small = rw.DeviceArray(rw.result_bytes(n, n, 600, 600) - 16, np.uint8)
    assert L.rwh_pack_result(tb.ptr(), ws.counts_ptr, 1, n, ALL, small.ptr, small.nbytes, None) == \
        rw.RWH_ERR_CAPACITY
    assert L.rwh_pack_result(tb.ptr(), ws.counts_ptr, 1, n, 1 << 14, small.ptr, small.nbytes, None) == \
        rw.RWH_ERR_INVALID_ARG
    assert L.rwh_pack_result(tb.ptr(), ws.counts_ptr, 0, n, ALL, small.ptr, small.nbytes, None) == \
        rw.RWH_ERR_INVALID_ARG
    assert L.rwh_pack_result(tb.ptr(), ws.counts_ptr, 1, n, ALL, small.ptr + 4, small.nbytes, None) == \
        rw.RWH_ERR_INVALID_ARG
