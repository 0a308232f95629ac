"""TEST INFRASTRUCTURE ONLY — ctypes access to the CPU oracle (tree_oracle.h)
and to the reference's own CPU kernels compiled from source (_ref/).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(HERE, "build", "libtree_oracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_tree.so")

_P = C.c_void_p
_U32 = C.c_uint32
_I = C.c_int
_F = C.c_float
ABI = {
    "oracle_pool_create": (_P, [_I]),
    "oracle_pool_destroy": (None, [_P]),
    "oracle_pool_threads": (_I, [_P]),
    "oracle_morton": (None, [_P, _P, _U32, _F, _F, _P]),
    "oracle_sort": (None, [_P, _P, _P, _P, _U32]),
    "oracle_unique": (_U32, [_P, _U32, _P]),
    "oracle_radix_tree": (None, [_P, _P, _U32, _P, _P, _P, _P, _P]),
    "oracle_edge_count": (None, [_P, _P, _P, _U32, _P]),
    "oracle_prefix_sum": (C.c_int64, [_P, _U32, _P]),
    "oracle_octree": (C.c_int64, [_P, _P, _P, _P, _P, _P, _U32, _U32, _U32, _F, _F, _P, _P, _P, _P, _P]),
    "oracle_link_leaf": (C.c_int64, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _U32, _U32, _P, _P]),
    "oracle_build": (_I, [_P, _P, _U32, _F, _F] + [_P] * 11 + [_U32] + [_P] * 5 + [_P, _P]),
}
REF_ABI = {
    "ref_radix_tree": (None, [_P, _U32, _P, _P, _P, _P, _P]),
    "ref_edge_count": (None, [_P, _P, _U32, _P]),
    "ref_sort": (None, [_P, _P, _U32, _I]),
}

_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_PATH):
            raise RuntimeError(f"{ORACLE_PATH} missing: build oracle/ first")
        _lib = C.CDLL(ORACLE_PATH)
        for name, (res, args) in ABI.items():
            fn = getattr(_lib, name)
            fn.restype, fn.argtypes = res, args
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_PATH)
        for name, (res, args) in REF_ABI.items():
            fn = getattr(_ref, name)
            fn.restype, fn.argtypes = res, args
    return _ref


class Pool:
    def __init__(self, n_threads: int):
        self.ptr = lib().oracle_pool_create(n_threads)
        self.n_threads = n_threads

    def __del__(self):
        try:
            lib().oracle_pool_destroy(self.ptr)
        except Exception:
            pass


def _p(a: np.ndarray) -> int:
    assert a.flags.c_contiguous
    return a.ctypes.data


def morton(pool, pts, min_coord=0.0, range_=1024.0):
    pts = np.ascontiguousarray(pts, dtype=np.float32)
    n = pts.shape[0]
    out = np.empty(n, np.uint32)
    lib().oracle_morton(pool.ptr, _p(pts), n, min_coord, range_, _p(out))
    return out


def sort(pool, keys):
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    out = np.empty_like(keys)
    tmp = np.empty_like(keys)
    lib().oracle_sort(pool.ptr, _p(keys), _p(out), _p(tmp), keys.size)
    return out


def unique(sorted_keys):
    sorted_keys = np.ascontiguousarray(sorted_keys, dtype=np.uint32)
    out = np.empty_like(sorted_keys)
    m = lib().oracle_unique(_p(sorted_keys), sorted_keys.size, _p(out))
    return out[:m]


def radix_tree(pool, codes):
    codes = np.ascontiguousarray(codes, dtype=np.uint32)
    nu = codes.size
    m = max(nu, 1)
    r = {k: np.zeros(m, dt) for k, dt in (("prefix_n", np.uint8), ("has_leaf_left", np.uint8),
                                          ("has_leaf_right", np.uint8), ("left_child", np.int32),
                                          ("parents", np.int32))}
    lib().oracle_radix_tree(pool.ptr, _p(codes), nu, _p(r["prefix_n"]), _p(r["has_leaf_left"]),
                            _p(r["has_leaf_right"]), _p(r["left_child"]), _p(r["parents"]))
    nb = max(nu - 1, 0)
    return {k: v[:nb] for k, v in r.items()}


def edge_count(pool, prefix_n, parents):
    prefix_n = np.ascontiguousarray(prefix_n, np.uint8)
    parents = np.ascontiguousarray(parents, np.int32)
    out = np.empty(prefix_n.size, np.int32)
    lib().oracle_edge_count(pool.ptr, _p(prefix_n), _p(parents), prefix_n.size, _p(out))
    return out


def prefix_sum(counts):
    counts = np.ascontiguousarray(counts, np.int32)
    out = np.empty_like(counts)
    total = lib().oracle_prefix_sum(_p(counts), counts.size, _p(out))
    return out, int(total)


def octree(pool, offsets, counts, codes, prefix_n, parents, n_oct, oct_cap,
           min_coord=0.0, range_=1024.0):
    rows = max(min(n_oct, oct_cap), 1)
    r = {"children": np.zeros((rows, 8), np.int32), "corner": np.zeros((rows, 4), np.float32),
         "cell_size": np.zeros(rows, np.float32), "child_node_mask": np.zeros(rows, np.int32),
         "child_leaf_mask": np.zeros(rows, np.int32)}
    a = [np.ascontiguousarray(x) for x in (offsets, counts, codes, prefix_n, parents)]
    dropped = lib().oracle_octree(pool.ptr, *[_p(x) for x in a], a[0].size, n_oct, oct_cap,
                                  min_coord, range_, _p(r["children"]), _p(r["corner"]),
                                  _p(r["cell_size"]), _p(r["child_node_mask"]), _p(r["child_leaf_mask"]))
    return r, int(dropped)


def build(pool, pts, min_coord=0.0, range_=1024.0):
    """All seven stages; returns (counts, outputs, stage_ms)."""
    pts = np.ascontiguousarray(pts, dtype=np.float32)
    n = pts.shape[0]
    cap = int(np.float32(n) * np.float32(0.6))
    m = max(n, 1)
    o = {"morton_keys": np.zeros(m, np.uint32), "morton_keys_alt": np.zeros(m, np.uint32),
         "unique": np.zeros(m, np.uint32), "tmp": np.zeros(m, np.uint32),
         "prefix_n": np.zeros(m, np.uint8), "has_leaf_left": np.zeros(m, np.uint8),
         "has_leaf_right": np.zeros(m, np.uint8), "left_child": np.zeros(m, np.int32),
         "parents": np.zeros(m, np.int32), "edge_count": np.zeros(m, np.int32),
         "edge_offset": np.zeros(m, np.int32),
         "children": np.zeros((max(cap, 1), 8), np.int32), "corner": np.zeros((max(cap, 1), 4), np.float32),
         "cell_size": np.zeros(max(cap, 1), np.float32), "child_node_mask": np.zeros(max(cap, 1), np.int32),
         "child_leaf_mask": np.zeros(max(cap, 1), np.int32)}
    counts = np.zeros(3, np.uint32)
    ms = np.zeros(7, np.float64)
    rc = lib().oracle_build(pool.ptr, _p(pts), n, min_coord, range_, _p(o["morton_keys"]),
                            _p(o["morton_keys_alt"]), _p(o["unique"]), _p(o["tmp"]),
                            _p(o["prefix_n"]), _p(o["has_leaf_left"]), _p(o["has_leaf_right"]),
                            _p(o["left_child"]), _p(o["parents"]), _p(o["edge_count"]),
                            _p(o["edge_offset"]), cap, _p(o["children"]), _p(o["corner"]),
                            _p(o["cell_size"]), _p(o["child_node_mask"]), _p(o["child_leaf_mask"]),
                            _p(counts), _p(ms))
    return (int(counts[0]), int(counts[1]), int(counts[2]), rc), o, ms


def link_leaf(pool, o, counts):
    """Runs oracle_link_leaf in place on the outputs o of build()."""
    nu, nb, no = counts[:3]
    lib().oracle_link_leaf(pool.ptr, _p(o["edge_offset"]), _p(o["edge_count"]), _p(o["unique"]),
                           _p(o["has_leaf_left"]), _p(o["has_leaf_right"]), _p(o["prefix_n"]),
                           _p(o["parents"]), _p(o["left_child"]), nb,
                           min(no, o["cell_size"].shape[0]), _p(o["children"]), _p(o["child_leaf_mask"]))
    return o
