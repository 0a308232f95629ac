"""CPU checks of the boundary: the libraries load, export every function the
headers declare, reject bad arguments before touching a GPU, and the input
generators follow their definition."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(ROOT, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rwh_[a-z0-9_]+)\s*\(", text)))


def exported(path):
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_every_declared_symbol_is_exported(rw):
    names = declared("include/rwhip.h")
    assert len(names) >= 40
    syms = exported(rw.RWHIP_PATH)
    missing = [n for n in names if n not in syms]
    assert not missing, missing
    # and the Python binding declares exactly the header's functions
    assert sorted(rw.ABI) == names


def test_host_library_exports(rw):
    syms = exported(rw.HOST_PATH)
    for n in rw.HOST_ABI:
        assert n in syms, n


def test_argument_errors_need_no_gpu(rw):
    L = rw.lib()
    # sizes above the 2^30 limit are rejected before any HIP call
    assert L.rwh_morton(None, (1 << 30) + 1, 0.0, 1024.0, None, None) == rw.RWH_ERR_INVALID_ARG
    assert L.rwh_sort_u32(None, None, None, 5, 30, None) == rw.RWH_ERR_INVALID_ARG
    assert b"null workspace" in L.rwh_last_error()
    # a block that is not a workspace of this build (another build's layout,
    # or no workspace at all) is refused before anything reads its pointers
    forged = np.zeros(512, np.uint8)
    assert L.rwh_sort_u32(forged.ctypes.data, None, None, 5, 30, None) == rw.RWH_ERR_INVALID_ARG
    assert b"not a workspace of this build" in L.rwh_last_error()
    assert L.rwh_radix_tree(None, None, None, (1 << 30) + 1, None, None, None, None, None, None) == \
        rw.RWH_ERR_INVALID_ARG
    assert L.rwh_octree(None, None, None, None, 10, 0.0, 1024.0, None, None) == rw.RWH_ERR_INVALID_ARG
    assert L.rwh_build_octree_marked(None, None, 10, 0.0, 1024.0, None, None) == rw.RWH_ERR_INVALID_ARG
    assert L.rwh_error_string(rw.RWH_ERR_CAPACITY) == b"capacity exceeded"
    counts = np.zeros(4, np.uint32)
    n = np.array([1, 2], np.uint32)
    assert L.rwh_batch_layout(n.ctypes.data, 2, None, None, None, None, None) == rw.RWH_ERR_INVALID_ARG


def test_injected_failure_is_one_shot(rw):
    L = rw.lib()
    L.rwh_debug_fail_next(1)
    assert L.rwh_morton(None, 0, 0.0, 1024.0, None, None) == rw.RWH_ERR_INJECTED
    assert L.rwh_morton(None, 0, 0.0, 1024.0, None, None) == rw.RWH_OK  # n = 0: nothing to do


def test_batch_layout(rw):
    L = rw.lib()
    n = np.array([0, 1, 8192, 8193, 100000], np.uint32)
    pb, ob, oc = (np.zeros(5, np.uint32) for _ in range(3))
    space, ospace = C.c_uint32(0), C.c_uint32(0)
    rw.check(L.rwh_batch_layout(n.ctypes.data, 5, pb.ctypes.data, ob.ctypes.data, oc.ctypes.data,
                                C.byref(space), C.byref(ospace)))
    assert pb.tolist() == [0, 8192, 16384, 24576, 40960]
    assert space.value == 40960 + 13 * 8192
    assert oc.tolist() == [int(np.float32(v) * np.float32(0.6)) for v in n]
    assert np.all(ob[1:] >= ob[:-1] + oc[:-1]) and ospace.value >= ob[-1] + oc[-1]


def _mix64(z):
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def test_uniform_generator_definition(rw):
    # coordinate = top 24 bits of splitmix64(seed*phi + stream*c + counter) / 2^14
    n, seed = 5000, 114514
    with np.errstate(over="ignore"):
        ctr = np.arange(3 * n, dtype=np.uint64)
        x = np.uint64(seed) * np.uint64(0x9E3779B97F4A7C15) + ctr
        bits = (_mix64(x) >> np.uint64(32)).astype(np.uint32)
    coords = ((bits >> 8).astype(np.float32) / np.float32(16384)).reshape(n, 3)
    pts = rw.uniform_points(n, seed)
    assert np.array_equal(pts[:, :3], coords)
    assert np.all(pts[:, 3] == 0) and pts[:, :3].max() < 1024


def test_gmm_generator_shape(rw):
    pts = rw.gmm_points(20000, seed=1, n_centers=64, sigma=20.48)
    assert pts.shape == (20000, 4)
    assert pts[:, :3].min() >= 0 and pts[:, :3].max() < 1024
    # heavy duplicates at 10 bits per axis compared with uniform data
    q = np.floor(pts[:, :3]).astype(np.int64)
    assert np.unique(q[:, 0] * 1 << 20 | q[:, 1] << 10 | q[:, 2]).size < 20000


def test_sizes_generator(rw):
    s = rw.log_uniform_sizes(7, 512, 10, 20)
    assert s.min() >= 1024 and s.max() <= 1 << 20
    assert 60_000 < s.mean() < 250_000


def test_marked_build_arguments(rw):
    # the stage-marked build takes exactly RWH_BUILD_MARKS events, all set
    with pytest.raises(ValueError):
        rw.build_marked(None, None, 10, [None] * 3)
    assert rw.BUILD_MARKS == 8 and len(rw.BUILD_STAGES) == rw.BUILD_MARKS - 1
    h = open(os.path.join(ROOT, "include", "rwhip.h")).read()
    assert "#define RWH_BUILD_MARKS 8" in h
