"""Python access to the gfx950 octree-build backend.

ctypes bindings of the C ABI (include/rwhip.h, lib/librwhip.so) and of the
host layer (lib/libredwood_tree.so: input generators, the drop-in
Engine/AppData/run_stage path), plus small numpy helpers used by the tests and
the benchmark.  Nothing here computes on the CPU: every stage runs in the HIP
library, and loading fails loudly when it is missing.

The directory name is not a Python identifier; load it with
``load_package()`` of ``__graft_entry__`` (it registers the module as
``redwood_comp_amd``).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
RWHIP_PATH = os.path.join(LIB_DIR, "librwhip.so")
HOST_PATH = os.path.join(LIB_DIR, "libredwood_tree.so")

MIN_COORD = 0.0
RANGE = 1024.0
MEMORY_RATIO = np.float32(0.6)

RWH_OK = 0
RWH_ERR_INVALID_ARG = 1
RWH_ERR_HIP = 2
RWH_ERR_CAPACITY = 3
RWH_ERR_NO_DEVICE = 4
RWH_ERR_INJECTED = 5
RWH_ERR_OVERFLOW = 6
# status bits of a build (RWH_STATUS_* of include/rwhip.h)
STATUS_PAIR_CAP, STATUS_CLIMB, STATUS_CROSS_CAP = 0x1, 0x2, 0x4
STATUS_KEY_ORDER, STATUS_PULL, STATUS_LEAF = 0x8, 0x10, 0x20
NOALLOC_MAX_POINTS = 1 << 22  # RWH_NOALLOC_MAX_POINTS

# Every function of include/rwhip.h: name -> (restype, argtypes).
_P = C.c_void_p
_U32 = C.c_uint32
_I = C.c_int
_F = C.c_float
_SZ = C.c_size_t
ABI = {
    "rwh_error_string": (C.c_char_p, [_I]),
    "rwh_last_error": (C.c_char_p, []),
    "rwh_version": (_I, []),
    "rwh_device_count": (_I, [C.POINTER(_I)]),
    "rwh_device_name": (_I, [C.c_char_p, _SZ]),
    "rwh_malloc": (_I, [C.POINTER(_P), _SZ]),
    "rwh_free": (_I, [_P]),
    "rwh_host_alloc": (_I, [C.POINTER(_P), _SZ]),
    "rwh_host_free": (_I, [_P]),
    "rwh_memcpy_h2d": (_I, [_P, _P, _SZ, _P]),
    "rwh_memcpy_d2h": (_I, [_P, _P, _SZ, _P]),
    "rwh_memcpy_d2d": (_I, [_P, _P, _SZ, _P]),
    "rwh_memset": (_I, [_P, _I, _SZ, _P]),
    "rwh_stream_create": (_I, [C.POINTER(_P)]),
    "rwh_stream_destroy": (_I, [_P]),
    "rwh_stream_sync": (_I, [_P]),
    "rwh_device_sync": (_I, []),
    "rwh_event_create": (_I, [C.POINTER(_P)]),
    "rwh_event_destroy": (_I, [_P]),
    "rwh_event_record": (_I, [_P, _P]),
    "rwh_event_sync": (_I, [_P]),
    "rwh_event_query": (_I, [_P]),
    "rwh_event_elapsed_ms": (_I, [C.POINTER(_F), _P, _P]),
    "rwh_workspace_create": (_I, [_U32, _U32, C.POINTER(_P)]),
    "rwh_workspace_destroy": (_I, [_P]),
    "rwh_workspace_bytes": (_SZ, [_P]),
    "rwh_workspace_counts": (_P, [_P]),
    "rwh_workspace_max_segments": (_U32, [_P]),
    "rwh_sort_variant_count": (_I, []),
    "rwh_sort_variant_shape": (_I, [_I, C.POINTER(_U32), C.POINTER(_U32)]),
    "rwh_workspace_set_sort_variant": (_I, [_P, _I]),
    "rwh_sort_plan_count": (_I, []),
    "rwh_sort_plan_widths": (_I, [_I, _P]),
    "rwh_workspace_set_sort_plan": (_I, [_P, _I]),
    "rwh_workspace_set_bucket_sort": (_I, [_P, _I]),
    "rwh_workspace_sort_used_buckets": (_I, [_P, _P, C.POINTER(_I)]),
    "rwh_morton": (_I, [_P, _U32, _F, _F, _P, _P]),
    "rwh_sort_u32": (_I, [_P, _P, _P, _U32, _U32, _P]),
    "rwh_unique_u32": (_I, [_P, _P, _U32, _P, _P, _P]),
    "rwh_radix_tree": (_I, [_P, _P, _P, _U32, _P, _P, _P, _P, _P, _P]),
    "rwh_edge_count": (_I, [_P, _P, _P, _U32, _P, _P]),
    "rwh_prefix_sum": (_I, [_P, _P, _P, _U32, _P, _P, _P]),
    "rwh_octree": (_I, [_P, _P, _P, _P, _U32, _F, _F, _P, _P]),
    "rwh_link_leaf": (_I, [_P, _P, _P, _U32, _P, _P]),
    "rwh_build_octree": (_I, [_P, _P, _U32, _F, _F, _P]),
    "rwh_build_octree_marked": (_I, [_P, _P, _U32, _F, _F, _P, _P]),
    "rwh_build_octree_noalloc": (_I, [_P, _U32, _F, _F, _P, _P]),
    "rwh_read_counts": (_I, [_P, _U32, C.POINTER(_U32), _P]),
    "rwh_batch_layout": (_I, [_P, _U32, _P, _P, _P, C.POINTER(_U32), C.POINTER(_U32)]),
    "rwh_build_octree_batch": (_I, [_P, _P, _P, _P, _P, _P, _U32, _F, _F, _P]),
    "rwh_read_batch_counts": (_I, [_P, _U32, _P, _P]),
    "rwh_batch_prepare": (_I, [_P, _P, _P, _P, _P, _U32, _P]),
    "rwh_build_octree_batch_enqueue": (_I, [_P, _P, _F, _F, _P]),
    "rwh_batch_graph_create": (_I, [_P, _P, _F, _F, _P, C.POINTER(_P)]),
    "rwh_graph_launch": (_I, [_P, _P]),
    "rwh_graph_destroy": (_I, [_P]),
    "rwh_debug_fail_next": (None, [_I]),
    "rwh_result_bytes": (C.c_uint64, [_U32, _U32, _U32, _U32, _U32]),
    "rwh_pack_result": (_I, [_P, _P, _U32, _U32, _U32, _P, C.c_uint64, _P]),
    "rwh_result_check": (_I, [_P, C.c_uint64]),
    "rwh_result_section": (_I, [_P, C.c_uint64, _I, C.POINTER(_P), C.POINTER(_U32)]),
    "rwh_result_header_get": (_I, [_P, C.c_uint64, _P]),
}

HOST_ABI = {
    "rwt_last_error": (C.c_char_p, []),
    "rwt_bits32": (_U32, [C.c_uint64, C.c_uint64, C.c_uint64]),
    "rwt_generate_uniform": (None, [C.c_uint64, C.c_uint64, _P]),
    "rwt_generate_gmm": (None, [C.c_uint64, C.c_uint64, _U32, C.c_double, _P]),
    "rwt_log_uniform_sizes": (None, [C.c_uint64, _U32, C.c_double, C.c_double, _P]),
    "rwt_dropin_build": (_I, [_P, _U32, _I] + [_P] * 14 + [_P]),
    "rwt_dropin_packed": (_I, [_P, _U32, _U32, _P, C.c_uint64, _P]),
    "rwt_plan_batch": (_U32, [C.c_uint64, _U32, _P, _U32]),
    "rwt_best_sort_variant": (_I, [_U32]),
    "rwt_best_sort_config": (None, [_U32, _I, C.POINTER(_I), C.POINTER(_I)]),
    "rwt_measure_sort_configs": (_I, [_U32, _I, _I, _P, _P, _P, _I]),
    "rwt_tune_lookup": (_I, [_U32, _I, C.POINTER(_I), C.POINTER(_I)]),
    "rwt_tune_store": (None, [_U32, _I, _I, _I]),
    "rwt_server_create": (_P, [C.c_uint64, _U32, _U32, _U32, _I]),
    "rwt_server_destroy": (None, [_P]),
    "rwt_server_slots": (_U32, [_P]),
    "rwt_serve_load": (_I, [_P, _U32, _P, _P, C.c_double, _U32, C.c_uint64, _P]),
    "rwt_serve_mix": (_I, [_P, _U32, _P, _P, _U32, _P, _P, _P, _P]),
}


class RwhError(RuntimeError):
    def __init__(self, status: int, what: str):
        super().__init__(what)
        self.status = status


class TreeBuffers(C.Structure):
    """rwh_tree_buffers of include/rwhip.h."""

    _fields_ = [(name, _P) for name in (
        "points", "morton_keys", "morton_keys_alt", "prefix_n", "has_leaf_left",
        "has_leaf_right", "left_child", "parents", "edge_count", "edge_offset", "children",
        "corner", "cell_size", "child_node_mask", "child_leaf_mask")] + [("oct_capacity", _U32)]


_lib = None
_host = None


def _declare(lib, table):
    for name, (res, args) in table.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def lib():
    """The loaded librwhip.so (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(RWHIP_PATH):
            raise RuntimeError(f"{RWHIP_PATH} is missing: run __graft_entry__.build() first")
        _lib = C.CDLL(RWHIP_PATH)
        _declare(_lib, ABI)
    return _lib


def host():
    """The loaded libredwood_tree.so (host layer)."""
    global _host
    if _host is None:
        lib()
        if not os.path.exists(HOST_PATH):
            raise RuntimeError(f"{HOST_PATH} is missing: run __graft_entry__.build() first")
        _host = C.CDLL(HOST_PATH)
        _declare(_host, HOST_ABI)
    return _host


def check(status: int, what: str = "rwhip"):
    if status != RWH_OK:
        L = lib()
        raise RwhError(status, f"{what}: {L.rwh_error_string(status).decode()} "
                               f"({L.rwh_last_error().decode()})")


def device_count() -> int:
    n = _I(0)
    check(lib().rwh_device_count(C.byref(n)), "rwh_device_count")
    return n.value


def device_name() -> str:
    buf = C.create_string_buffer(256)
    check(lib().rwh_device_name(buf, 256), "rwh_device_name")
    return buf.value.decode()


def octree_capacity(n: int) -> int:
    """floor(n * 0.6f) in float arithmetic, as the reference sizes it."""
    return int(np.float32(n) * MEMORY_RATIO)


# ---------------------------------------------------------------------------
# inputs (the counter-based generators of the host layer)
# ---------------------------------------------------------------------------
def uniform_points(n: int, seed: int = 114514) -> np.ndarray:
    out = np.empty((n, 4), dtype=np.float32)
    host().rwt_generate_uniform(seed, n, out.ctypes.data)
    return out


def gmm_points(n: int, seed: int, n_centers: int = 64, sigma: float = 20.48) -> np.ndarray:
    out = np.empty((n, 4), dtype=np.float32)
    host().rwt_generate_gmm(seed, n, n_centers, sigma, out.ctypes.data)
    return out


def log_uniform_sizes(seed: int, count: int, lo_log2: float, hi_log2: float) -> np.ndarray:
    out = np.empty(count, dtype=np.uint32)
    host().rwt_log_uniform_sizes(seed, count, lo_log2, hi_log2, out.ctypes.data)
    return out


def batch_mix_clouds(count: int = 512, seed: int = 7, lo_log2: float = 10, hi_log2: float = 20):
    """The batched workload: count clouds with sizes log-uniform in
    [2^lo, 2^hi] (seeded), even clouds uniform in [0,1024)^3, odd clouds a
    16-cluster Gaussian mixture with sigma 10.24 (1% of the box).  Defaults:
    512 clouds of 2^10..2^20 points, seed 7."""
    sizes = log_uniform_sizes(seed, count, lo_log2, hi_log2)
    clouds = [uniform_points(int(n), (seed << 20) + s) if s % 2 == 0 else
              gmm_points(int(n), (seed << 20) + s, 16, 10.24) for s, n in enumerate(sizes)]
    return sizes, clouds


def serving_mix_clouds(count: int = 32, seed: int = 11, lo_log2: float = 12, hi_log2: float = 22):
    """The serving workload's distinct requests: count clouds with sizes
    log-uniform in [2^lo, 2^hi] (seeded), even clouds uniform, odd clouds the
    clustered kind of the single-build workload (64-component Gaussian mixture,
    sigma 20.48).  Defaults: 32 clouds of 2^12..2^22 points, seed 11."""
    sizes = log_uniform_sizes(seed, count, lo_log2, hi_log2)
    clouds = [uniform_points(int(n), (seed << 20) + s) if s % 2 == 0 else
              gmm_points(int(n), (seed << 20) + s, 64, 20.48) for s, n in enumerate(sizes)]
    return sizes, clouds


# ---------------------------------------------------------------------------
# device memory
# ---------------------------------------------------------------------------
class DeviceArray:
    """A device allocation with a numpy dtype and length."""

    def __init__(self, count: int, dtype, zero: bool = False):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = self.count * self.dtype.itemsize
        p = _P()
        check(lib().rwh_malloc(C.byref(p), max(self.nbytes, 16)), "rwh_malloc")
        self.ptr = p.value
        if zero:
            check(lib().rwh_memset(self.ptr, 0, max(self.nbytes, 16), None), "rwh_memset")

    @classmethod
    def from_numpy(cls, a: np.ndarray, stream=None) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a, stream)
        return d

    def upload(self, a: np.ndarray, stream=None):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size <= self.count
        check(lib().rwh_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes, stream), "rwh_memcpy_h2d")
        check(lib().rwh_stream_sync(stream), "rwh_stream_sync")

    def numpy(self, count: int | None = None, stream=None) -> np.ndarray:
        count = self.count if count is None else int(count)
        out = np.empty(count, dtype=self.dtype)
        if count:
            check(lib().rwh_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes, stream), "rwh_memcpy_d2h")
            check(lib().rwh_stream_sync(stream), "rwh_stream_sync")
        return out

    def free(self):
        if self.ptr:
            lib().rwh_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Workspace:
    def __init__(self, max_points: int, max_segments: int = 1, sort_variant: int | None = None,
                 sort_plan: int | None = None, bucket_sort: bool | None = None):
        p = _P()
        check(lib().rwh_workspace_create(max_points, max_segments, C.byref(p)), "rwh_workspace_create")
        self.ptr = p.value
        self.max_segments = max_segments
        if sort_variant is not None:
            self.set_sort_variant(sort_variant)
        if sort_plan is not None:
            self.set_sort_plan(sort_plan)
        if bucket_sort is not None:
            self.set_bucket_sort(bucket_sort)

    def set_sort_variant(self, v: int):
        check(lib().rwh_workspace_set_sort_variant(self.ptr, v), "rwh_workspace_set_sort_variant")

    def set_sort_plan(self, p: int):
        check(lib().rwh_workspace_set_sort_plan(self.ptr, p), "rwh_workspace_set_sort_plan")

    def set_bucket_sort(self, enable: bool):
        """Whole single-cloud builds of (2^20, 2^24] points sort with the bucket
        sort (True, the default) or always with the digit passes (False)."""
        check(lib().rwh_workspace_set_bucket_sort(self.ptr, 1 if enable else 0), "rwh_workspace_set_bucket_sort")

    def sort_used_buckets(self, stream=None) -> bool:
        """Whether the last whole build on this workspace ran the bucket sort."""
        used = _I()
        check(lib().rwh_workspace_sort_used_buckets(self.ptr, stream, C.byref(used)),
              "rwh_workspace_sort_used_buckets")
        return bool(used.value)

    @property
    def counts_ptr(self) -> int:
        return lib().rwh_workspace_counts(self.ptr)

    def n_unique_ptr(self, s: int = 0) -> int:
        return self.counts_ptr + 4 * s

    def n_oct_ptr(self, s: int = 0) -> int:
        return self.counts_ptr + 4 * (self.max_segments + s)

    def status_ptr(self, s: int = 0) -> int:
        return self.counts_ptr + 4 * (2 * self.max_segments + s)

    def counts(self, n_clouds: int = 1) -> tuple[np.ndarray, np.ndarray]:
        """(n_unique, n_octree_nodes) of the first n_clouds clouds."""
        out = np.empty(3 * self.max_segments, dtype=np.uint32)
        check(lib().rwh_memcpy_d2h(out.ctypes.data, self.counts_ptr, out.nbytes, None), "rwh_memcpy_d2h")
        check(lib().rwh_stream_sync(None), "rwh_stream_sync")
        return out[:n_clouds], out[self.max_segments:self.max_segments + n_clouds]

    def free(self):
        if self.ptr:
            lib().rwh_workspace_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = _P()
        check(lib().rwh_stream_create(C.byref(p)), "rwh_stream_create")
        self.ptr = p.value

    def sync(self):
        check(lib().rwh_stream_sync(self.ptr), "rwh_stream_sync")

    def __del__(self):
        try:
            lib().rwh_stream_destroy(self.ptr)
        except Exception:
            pass


class Event:
    def __init__(self):
        p = _P()
        check(lib().rwh_event_create(C.byref(p)), "rwh_event_create")
        self.ptr = p.value

    def record(self, stream=None):
        check(lib().rwh_event_record(self.ptr, stream), "rwh_event_record")

    def sync(self):
        check(lib().rwh_event_sync(self.ptr), "rwh_event_sync")

    def elapsed_ms(self, stop: "Event") -> float:
        ms = _F(0)
        check(lib().rwh_event_elapsed_ms(C.byref(ms), self.ptr, stop.ptr), "rwh_event_elapsed_ms")
        return ms.value

    def __del__(self):
        try:
            lib().rwh_event_destroy(self.ptr)
        except Exception:
            pass


# ---------------------------------------------------------------------------
# one build's device buffers (the AppData layout)
# ---------------------------------------------------------------------------
TREE_FIELDS = {
    # name: (dtype, elements per point or per octree slot, per-point?)
    "morton_keys": (np.uint32, 1, True),
    "morton_keys_alt": (np.uint32, 1, True),
    "prefix_n": (np.uint8, 1, True),
    "has_leaf_left": (np.uint8, 1, True),
    "has_leaf_right": (np.uint8, 1, True),
    "left_child": (np.int32, 1, True),
    "parents": (np.int32, 1, True),
    "edge_count": (np.int32, 1, True),
    "edge_offset": (np.int32, 1, True),
    "children": (np.int32, 8, False),
    "corner": (np.float32, 4, False),
    "cell_size": (np.float32, 1, False),
    "child_node_mask": (np.int32, 1, False),
    "child_leaf_mask": (np.int32, 1, False),
}


class TreeBuild:
    """Device buffers of one build (or a batch index space) and its tree struct."""

    def __init__(self, n_index: int, oct_space: int, points: DeviceArray | None = None):
        self.n_index = n_index
        self.oct_space = oct_space
        self.points = points if points is not None else DeviceArray(4 * n_index, np.float32, zero=True)
        self.arrays = {}
        for name, (dt, k, per_point) in TREE_FIELDS.items():
            self.arrays[name] = DeviceArray(k * (n_index if per_point else oct_space), dt, zero=True)
        self.struct = TreeBuffers(points=self.points.ptr, oct_capacity=oct_space,
                                  **{k: v.ptr for k, v in self.arrays.items()})

    def ptr(self) -> int:
        return C.addressof(self.struct)

    def __getitem__(self, name) -> DeviceArray:
        return self.arrays[name]

    def read(self, n: int, n_rows: int) -> dict:
        """Host copies: per-point arrays over [0, n), octree arrays over [0, n_rows)."""
        out = {}
        for name, (dt, k, per_point) in TREE_FIELDS.items():
            cnt = k * (n if per_point else n_rows)
            a = self.arrays[name].numpy(cnt)
            out[name] = a.reshape(-1, k) if k > 1 else a
        return out


class PinnedArray:
    """Page-locked host memory viewed as a numpy array (fast H2D uploads)."""

    def __init__(self, shape, dtype=np.float32):
        dt = np.dtype(dtype)
        nbytes = int(np.prod(shape)) * dt.itemsize
        p = _P()
        check(lib().rwh_host_alloc(C.byref(p), max(nbytes, 16)), "rwh_host_alloc")
        self.ptr = p.value
        buf = (C.c_char * max(nbytes, 16)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=dt, count=int(np.prod(shape))).reshape(shape)

    def free(self):
        if self.ptr:
            self.array = None
            lib().rwh_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


BUILD_MARKS = 8  # RWH_BUILD_MARKS
BUILD_STAGES = ("S1_morton", "S2_sort", "S3_unique", "S4_radix_tree", "S5_edge_count", "S6_prefix_sum",
                "S7_octree")


def build_marked(ws, tb, n: int, marks: list, stream=None, min_coord: float = MIN_COORD,
                 range_: float = RANGE) -> int:
    """rwh_build_octree_marked on resident buffers; returns its status."""
    if len(marks) != BUILD_MARKS:
        raise ValueError(f"marks must hold {BUILD_MARKS} events")
    arr = (_P * BUILD_MARKS)(*[e.ptr for e in marks])
    return lib().rwh_build_octree_marked(ws.ptr, tb.ptr(), n, min_coord, range_, arr, stream)


def stage_ms(marks: list) -> dict:
    """Stage times (ms) between the marks of a completed marked build."""
    marks[-1].sync()
    return {name: marks[i].elapsed_ms(marks[i + 1]) for i, name in enumerate(BUILD_STAGES)}


def build_octree(points: np.ndarray, ws: Workspace | None = None, stream=None,
                 min_coord: float = MIN_COORD, range_: float = RANGE, link_leaf: bool = False,
                 strict: bool = True, sort_variant: int | None = None, sort_plan: int | None = None,
                 bucket_sort: bool | None = None, marks: list | None = None):
    """Runs S1..S7 on the GPU for host points (n, 4); returns (counts, host outputs).
    link_leaf also runs the opt-in leaf linking after S7.  strict=False returns the
    (capacity-clamped) outputs when n_octree_nodes exceeds the 0.6 n capacity
    instead of raising.  sort_variant picks the sort's tile shape (a new
    workspace's default otherwise); bucket_sort False forces the digit passes.
    marks: RWH_BUILD_MARKS Events recorded at the build's stage boundaries
    (rwh_build_octree_marked); stage_ms(marks) turns them into stage times."""
    n = points.shape[0]
    cap = octree_capacity(n)
    ws = ws or Workspace(max(n, 1))
    if sort_variant is not None:
        ws.set_sort_variant(sort_variant)
    if sort_plan is not None:
        ws.set_sort_plan(sort_plan)
    if bucket_sort is not None:
        ws.set_bucket_sort(bucket_sort)
    tb = TreeBuild(n, cap, DeviceArray.from_numpy(points.astype(np.float32).reshape(-1)))
    if marks is None:
        check(lib().rwh_build_octree(ws.ptr, tb.ptr(), n, min_coord, range_, stream), "rwh_build_octree")
    else:
        check(build_marked(ws, tb, n, marks, stream, min_coord, range_), "rwh_build_octree_marked")
    if link_leaf:
        check(lib().rwh_link_leaf(tb.ptr(), ws.n_unique_ptr(), ws.n_oct_ptr(), n, ws.status_ptr(), stream),
              "rwh_link_leaf")
    counts = (_U32 * 3)()
    st = lib().rwh_read_counts(ws.ptr, cap, counts, stream)
    if strict or st != RWH_ERR_CAPACITY:
        check(st, "rwh_read_counts")
    n_unique, n_brt, n_oct = counts[0], counts[1], counts[2]
    # octree rows [0, n_oct); slot 0 also when no node contributes (it holds the root cell)
    out = tb.read(n, min(max(n_oct, 1), cap))
    return (n_unique, n_brt, n_oct), out


# ---------------------------------------------------------------------------
# packed results (format version 1 of include/rwhip.h)
# ---------------------------------------------------------------------------
RESULT_MAGIC = 0x424F5752
RESULT_VERSION = 1
RESULT_HEADER_BYTES = 256
# section s of a packed result: (name, dtype, elements per entry); names as the
# keys of build_octree's outputs ("unique" for S3's keys)
RESULT_SECTIONS = (("morton_keys_alt", np.uint32, 1), ("unique", np.uint32, 1),
                   ("prefix_n", np.uint8, 1), ("has_leaf_left", np.uint8, 1),
                   ("has_leaf_right", np.uint8, 1), ("left_child", np.int32, 1),
                   ("parents", np.int32, 1), ("edge_count", np.int32, 1),
                   ("edge_offset", np.int32, 1), ("children", np.int32, 8),
                   ("corner", np.float32, 4), ("cell_size", np.float32, 1),
                   ("child_node_mask", np.int32, 1), ("child_leaf_mask", np.int32, 1))
SECTIONS_ALL = 0x3FFF
SECTIONS_OCTREE = 0x3E00


class ResultHeader(C.Structure):
    """rwh_result_header of include/rwhip.h (256 bytes)."""

    _fields_ = [(f, _U32) for f in ("magic", "version", "header_bytes", "sections", "status", "flags",
                                    "n_input", "n_unique", "n_brt", "n_octree_nodes", "oct_rows",
                                    "oct_capacity")] + [
        ("total_bytes", C.c_uint64), ("offset", C.c_uint64 * 14), ("count", _U32 * 14),
        ("entry_bytes", C.c_uint16 * 14), ("reserved", _U32)]


assert C.sizeof(ResultHeader) == RESULT_HEADER_BYTES


def result_bytes(n_input: int, n_unique: int, n_octree_nodes: int, oct_capacity: int,
                 sections: int = SECTIONS_ALL) -> int:
    """Size of the packed result of a build with these counts."""
    return int(lib().rwh_result_bytes(n_input, n_unique, n_octree_nodes, oct_capacity, sections))


def _host_bytes(buf):
    a = np.frombuffer(buf, np.uint8) if not isinstance(buf, np.ndarray) else buf.view(np.uint8).reshape(-1)
    return np.ascontiguousarray(a)


def result_check(buf) -> int:
    """rwh_result_check of a host buffer (bytes or numpy): the status code."""
    a = _host_bytes(buf)
    return int(lib().rwh_result_check(a.ctypes.data, a.size))


def unpack_result(buf):
    """(header fields as a dict, {section name: array}) of a packed result in
    host memory, read through the library's checked reader (raises RwhError
    if the buffer is malformed; the build's status is in the header)."""
    a = _host_bytes(buf)
    h = ResultHeader()
    check(lib().rwh_result_header_get(a.ctypes.data, a.size, C.byref(h)), "rwh_result_header_get")
    head = {f: getattr(h, f) for f, _ in ResultHeader._fields_ if f not in ("offset", "count", "entry_bytes")}
    head["offset"], head["count"] = list(h.offset), list(h.count)
    head["entry_bytes"] = list(h.entry_bytes)
    out = {}
    for s, (name, dt, width) in enumerate(RESULT_SECTIONS):
        if not (h.sections >> s) & 1:
            continue
        p, cnt = _P(), _U32()
        check(lib().rwh_result_section(a.ctypes.data, a.size, s, C.byref(p), C.byref(cnt)),
              "rwh_result_section")
        off = p.value - a.ctypes.data
        arr = a[off:off + cnt.value * width * np.dtype(dt).itemsize].view(dt)
        out[name] = arr.reshape(-1, width) if width > 1 else arr
    return head, out


def pack_result(tb: "TreeBuild", counts_ptr: int, counts_stride: int, n_input: int,
                sections: int = SECTIONS_ALL, stream=None, out: DeviceArray | None = None,
                tree: "TreeBuffers | None" = None) -> np.ndarray:
    """Packs a finished build on the device (rwh_pack_result) and copies the
    whole result to the host in one transfer; returns it as uint8."""
    t = tree if tree is not None else tb.struct
    cap = t.oct_capacity
    need = result_bytes(n_input, n_input, cap, cap, sections)
    if need == 0:
        check(RWH_ERR_INVALID_ARG, "pack_result: sections")
    dev = out if out is not None else DeviceArray(need, np.uint8)
    check(lib().rwh_pack_result(C.addressof(t), counts_ptr, counts_stride, n_input, sections, dev.ptr,
                                dev.nbytes, stream), "rwh_pack_result")
    head = dev.numpy(RESULT_HEADER_BYTES, stream)
    total = int(head[48:56].view(np.uint64)[0])
    return dev.numpy(min(total, dev.nbytes), stream)


# ---------------------------------------------------------------------------
# batched build
# ---------------------------------------------------------------------------
def batch_layout(sizes):
    sizes = np.ascontiguousarray(sizes, dtype=np.uint32)
    k = sizes.size
    pb, ob, oc = (np.zeros(k, np.uint32) for _ in range(3))
    space, ospace = _U32(0), _U32(0)
    check(lib().rwh_batch_layout(sizes.ctypes.data, k, pb.ctypes.data, ob.ctypes.data,
                                 oc.ctypes.data, C.byref(space), C.byref(ospace)), "rwh_batch_layout")
    return pb, ob, oc, space.value, ospace.value


class Batch:
    """Device buffers and layout of a batch of clouds (one shared index space)."""

    def __init__(self, clouds, ws: Workspace | None = None):
        self.sizes = np.array([c.shape[0] for c in clouds], np.uint32)
        self.pb, self.ob, self.oc, self.space, self.ospace = batch_layout(self.sizes)
        pts = np.zeros((self.space, 4), np.float32)
        for s, c in enumerate(clouds):
            pts[self.pb[s]:self.pb[s] + c.shape[0]] = c
        self.tb = TreeBuild(self.space, max(self.ospace, 1), DeviceArray.from_numpy(pts.reshape(-1)))
        self.ws = ws or Workspace(self.space, len(clouds))

    def run(self, stream=None, min_coord=MIN_COORD, range_=RANGE):
        check(lib().rwh_build_octree_batch(
            self.ws.ptr, self.tb.ptr(), self.sizes.ctypes.data, self.pb.ctypes.data,
            self.ob.ctypes.data, self.oc.ctypes.data, self.sizes.size, min_coord, range_, stream),
            "rwh_build_octree_batch")

    def counts(self, stream=None):
        k = self.sizes.size
        out = np.zeros(2 * k, np.uint32)
        check(lib().rwh_read_batch_counts(self.ws.ptr, k, out.ctypes.data, stream), "rwh_read_batch_counts")
        return out[:k], out[k:]

    def cloud_buffers(self, s: int) -> TreeBuffers:
        """rwh_tree_buffers of cloud s alone (pointers offset by its bases,
        oct_capacity = its capacity), e.g. for rwh_pack_result."""
        b, ob = int(self.pb[s]), int(self.ob[s])
        fields = {}
        for name, (dt, k, per_point) in TREE_FIELDS.items():
            fields[name] = self.tb.arrays[name].ptr + (b if per_point else ob) * k * np.dtype(dt).itemsize
        return TreeBuffers(points=self.tb.points.ptr + 16 * b, oct_capacity=int(self.oc[s]), **fields)

    def read(self):
        """Per-cloud host outputs, cut to each cloud's counts."""
        nu, no = self.counts()
        full = self.tb.read(self.space, max(self.ospace, 1))
        res = []
        for s in range(self.sizes.size):
            b, n, ob = int(self.pb[s]), int(self.sizes[s]), int(self.ob[s])
            nb = max(int(nu[s]) - 1, 0)
            # slot 0 too when no node contributes (it holds the root cell)
            rows = min(max(int(no[s]), 1), int(self.oc[s]))
            r = {"n_unique": int(nu[s]), "n_oct": int(no[s]),
                 "morton_keys_alt": full["morton_keys_alt"][b:b + n],
                 "unique": full["morton_keys"][b:b + int(nu[s])]}
            for k in ("prefix_n", "has_leaf_left", "has_leaf_right", "left_child", "parents",
                      "edge_count", "edge_offset"):
                r[k] = full[k][b:b + nb]
            for k in ("children", "corner", "cell_size", "child_node_mask", "child_leaf_mask"):
                r[k] = full[k][ob:ob + rows]
            res.append(r)
        return res


def dropin_packed(points: np.ndarray, sections: int = SECTIONS_ALL) -> np.ndarray:
    """One build through the drop-in path, read back with the C++
    redwood::hip::read_result (a packed result, uint8)."""
    n = points.shape[0]
    cap = octree_capacity(n)
    buf = np.zeros(max(result_bytes(n, n, cap, cap, sections), 16), np.uint8)
    size = C.c_uint64(0)
    pts = np.ascontiguousarray(points, np.float32)
    rc = host().rwt_dropin_packed(pts.ctypes.data, n, sections, buf.ctypes.data, buf.size, C.byref(size))
    if rc != 0:
        raise RwhError(rc, host().rwt_last_error().decode())
    return buf[:size.value]


def dropin_build(points: np.ndarray, fail_stage: int = 0):
    """One build through the C++ drop-in path (Engine + AppData + run_stage1..7)."""
    n = points.shape[0]
    cap = octree_capacity(n)
    m, c = max(n, 1), max(cap, 1)
    o = {"morton_keys_alt": np.zeros(m, np.uint32), "unique": np.zeros(m, np.uint32),
         "prefix_n": np.zeros(m, np.uint8), "has_leaf_left": np.zeros(m, np.uint8),
         "has_leaf_right": np.zeros(m, np.uint8), "left_child": np.zeros(m, np.int32),
         "parents": np.zeros(m, np.int32), "edge_count": np.zeros(m, np.int32),
         "edge_offset": np.zeros(m, np.int32), "children": np.zeros((c, 8), np.int32),
         "corner": np.zeros((c, 4), np.float32), "cell_size": np.zeros(c, np.float32),
         "child_node_mask": np.zeros(c, np.int32), "child_leaf_mask": np.zeros(c, np.int32)}
    counts = np.zeros(3, np.uint32)
    pts = np.ascontiguousarray(points, np.float32)
    rc = host().rwt_dropin_build(pts.ctypes.data, n, fail_stage,
                                 *[o[k].ctypes.data for k in o], counts.ctypes.data)
    if rc != 0:
        raise RwhError(rc, host().rwt_last_error().decode())
    return tuple(int(x) for x in counts), o


# ---------------------------------------------------------------------------
# serving (the C++ BuildServer of host/build_server.hpp)
# ---------------------------------------------------------------------------
SERVE_STATS = ("rate", "n_requests", "makespan_s", "throughput_rps", "mpts_per_s", "p50_ms",
               "p99_ms", "mean_ms", "max_ms", "mean_batch", "rejected", "batches")


def plan_batch(slot_points: int, max_clouds: int, sizes) -> int:
    sizes = np.ascontiguousarray(sizes, np.uint32)
    return int(host().rwt_plan_batch(slot_points, max_clouds, sizes.ctypes.data, sizes.size))


class Server:
    def __init__(self, pool_bytes=64 << 30, n_slots=3, slot_points=1 << 26, max_clouds=512,
                 plain=False):
        self.ptr = host().rwt_server_create(pool_bytes, n_slots, slot_points, max_clouds, int(plain))
        if not self.ptr:
            raise RwhError(-1, host().rwt_last_error().decode())
        self.n_slots = host().rwt_server_slots(self.ptr)

    def load(self, clouds, rate: float, n_requests: int, seed: int = 1) -> dict:
        ptrs = (C.c_void_p * len(clouds))(*[c.ctypes.data for c in clouds])
        sizes = np.array([c.shape[0] for c in clouds], np.uint32)
        stats = np.zeros(12, np.float64)
        rc = host().rwt_serve_load(self.ptr, len(clouds), ptrs, sizes.ctypes.data, rate,
                                   n_requests, seed, stats.ctypes.data)
        if rc != 0:
            raise RwhError(rc, host().rwt_last_error().decode())
        return dict(zip(SERVE_STATS, stats.tolist()))

    def mix(self, clouds, sections: int = SECTIONS_ALL):
        """All clouds submitted together, each asking for a packed result of
        the given sections; per request (packed result as uint8, batch size)."""
        k = len(clouds)
        bufs = []
        for c in clouds:
            n = c.shape[0]
            cap = octree_capacity(n)
            bufs.append(np.zeros(max(result_bytes(n, n, cap, cap, sections), 16), np.uint8))
        outs = (C.c_void_p * k)(*[b.ctypes.data for b in bufs])
        caps = np.array([b.size for b in bufs], np.uint64)
        sizes_out = np.zeros(k, np.uint64)
        ptrs = (C.c_void_p * k)(*[c.ctypes.data for c in clouds])
        sizes = np.array([c.shape[0] for c in clouds], np.uint32)
        bsz = np.zeros(k, np.uint32)
        rc = host().rwt_serve_mix(self.ptr, k, ptrs, sizes.ctypes.data, sections, outs, caps.ctypes.data,
                                  sizes_out.ctypes.data, bsz.ctypes.data)
        if rc != 0:
            raise RwhError(rc, host().rwt_last_error().decode())
        return [(bufs[r][:int(sizes_out[r])], int(bsz[r])) for r in range(k)]

    def close(self):
        if self.ptr:
            host().rwt_server_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sort_variants():
    """(block, keys_per_thread) of every compiled sort tile shape."""
    out = []
    for v in range(lib().rwh_sort_variant_count()):
        b, k = _U32(0), _U32(0)
        check(lib().rwh_sort_variant_shape(v, C.byref(b), C.byref(k)), "rwh_sort_variant_shape")
        out.append((b.value, k.value))
    return out


def sort_plans():
    """Digit widths (least significant pass first) of every sort plan of 30-bit keys."""
    out = []
    for p in range(lib().rwh_sort_plan_count()):
        w = np.zeros(4, np.uint32)
        check(lib().rwh_sort_plan_widths(p, w.ctypes.data), "rwh_sort_plan_widths")
        out.append(tuple(int(x) for x in w))
    return out


def best_sort_variant(n: int) -> int:
    return int(host().rwt_best_sort_variant(n))


def best_sort_config(n: int, chunk: bool = False) -> tuple[int, int]:
    """The tuned (tile shape, digit plan) for n keys (chunk: a batch's sort)."""
    v, p = _I(0), _I(0)
    host().rwt_best_sort_config(n, int(chunk), C.byref(v), C.byref(p))
    return v.value, p.value


def measure_sort_configs(n: int, reps: int = 5, chunk: bool = False) -> list:
    """[(variant, plan, ms)] for every configuration at size n."""
    cap = 256
    ms = np.zeros(cap, np.float64)
    vs = np.zeros(cap, np.int32)
    ps = np.zeros(cap, np.int32)
    k = host().rwt_measure_sort_configs(n, reps, int(chunk), ms.ctypes.data, vs.ctypes.data, ps.ctypes.data, cap)
    return [(int(vs[i]), int(ps[i]), float(ms[i])) for i in range(k)]


def tune_lookup(n: int, chunk: bool = False):
    """The cached (variant, plan) of n's size bucket, or None (no measuring)."""
    v, p = _I(0), _I(0)
    return (v.value, p.value) if host().rwt_tune_lookup(n, int(chunk), C.byref(v), C.byref(p)) else None


def tune_store(n: int, variant: int, plan: int, chunk: bool = False):
    host().rwt_tune_store(n, int(chunk), variant, plan)
