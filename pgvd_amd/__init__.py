"""pgvd_amd -- Python binding of libpgvd.so, the MI355X (gfx950) octree
construction path of edwardsjohnmartin/pgvd.

The binding is ctypes over the C ABI declared in include/pgvd.h.  Method
names follow the reference's stage drivers (namespace Kernels,
opencl/Kernels.h:26-50) and facade (Karras.h:23-36):

    Device.quantize           Karras::Quantize             (Karras.cpp:69-99)
    Device.points_to_morton   Kernels::PointsToMorton_p    (opencl/Kernels.cpp:47-60)
    Device.radix_sort         Kernels::RadixSortBigUnsigned(opencl/Kernels.cpp:219-247)
    Device.unique_sorted      Kernels::UniqueSorted        (opencl/Kernels.cpp:198-217)
    Device.stream_scan        Kernels::StreamScan_p        (opencl/Kernels.cpp:102-123)
    Device.build_brt          Kernels::BuildBinaryRadixTree_p (opencl/Kernels.cpp:249-264)
    Device.local_splits       Kernels::ComputeLocalSplits_p   (opencl/Kernels.cpp:275-301)
    Device.brt_to_octree      Kernels::BinaryRadixToOctree_p  (opencl/Kernels.cpp:333-368)
    Device.build_octree       Kernels::BuildOctree_p       (opencl/Kernels.cpp:413-429)

Every call runs HIP kernels; there is no CPU fallback.  Importing works
without a GPU (tests check the exported symbols); creating a Device needs
one.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# PGVD_LIB selects another build of the library (e.g. to compare tile shapes).
LIB_PATH = os.environ.get("PGVD_LIB") or os.path.join(HERE, "libpgvd.so")
ROOT = os.path.dirname(HERE)
HEADER = os.path.join(ROOT, "include", "pgvd.h")

STATUS = {
    0: "PGVD_OK", 1: "PGVD_ERR_INVALID", 2: "PGVD_ERR_EMPTY", 3: "PGVD_ERR_BAD_COORD",
    4: "PGVD_ERR_NO_DEVICE", 5: "PGVD_ERR_ALLOC", 6: "PGVD_ERR_HIP", 7: "PGVD_ERR_CAPACITY",
    8: "PGVD_ERR_TOO_LARGE", 9: "PGVD_ERR_DUP_KEY", 10: "PGVD_ERR_INTERNAL",
    11: "PGVD_NOT_READY", 12: "PGVD_ERR_GEOMETRY",
}


class PgvdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


_LIB = None


class OctreeResult(C.Structure):
    _fields_ = [
        ("n", C.c_int64), ("m", C.c_int64), ("octree_size", C.c_int64),
        ("dim", C.c_int), ("bits", C.c_int), ("mbits", C.c_int), ("key_bytes", C.c_int),
        ("malformed", C.c_int), ("bucket_sort", C.c_int),
        ("keys", C.c_void_p), ("brt", C.c_void_p), ("parent", C.c_void_p),
        ("prefix_sums", C.c_void_p), ("octree", C.c_void_p),
        ("key_conflict", C.c_void_p), ("node_conflict", C.c_void_p),
    ]


class BuildOpts(C.Structure):
    _fields_ = [
        ("use_custom_bb", C.c_int), ("bb_min", C.c_float * 3), ("bb_max", C.c_float * 3),
        ("clamped", C.c_int), ("quantized_input", C.c_int), ("no_host_wait", C.c_int),
    ]


class BatchOpts(C.Structure):
    _fields_ = [("quantized_input", C.c_int), ("clamped", C.c_int), ("use_graph", C.c_int),
                ("sort_ipt", C.c_int), ("brt_rows", C.c_int), ("digit_bits", C.c_int),
                ("sort_waves", C.c_int)]


class BatchResult(C.Structure):
    _fields_ = [
        ("count", C.c_int64), ("dim", C.c_int), ("bits", C.c_int), ("mbits", C.c_int),
        ("key_bytes", C.c_int), ("octree_stride", C.c_int), ("total_octree", C.c_int64),
        ("m", C.POINTER(C.c_int64)), ("octree_size", C.POINTER(C.c_int64)),
        ("key_offset", C.POINTER(C.c_int64)), ("octree_offset", C.POINTER(C.c_int64)),
        ("keys", C.c_void_p), ("brt", C.c_void_p), ("parent", C.c_void_p),
        ("prefix_sums", C.c_void_p), ("octree", C.c_void_p),
    ]


class Request(C.Structure):
    _fields_ = [
        ("host_pts", C.c_void_p), ("n", C.c_int64), ("dim", C.c_int), ("bits", C.c_int),
        ("arrival_s", C.c_double), ("admit_s", C.c_double), ("done_s", C.c_double),
        ("m", C.c_int64), ("octree_size", C.c_int64), ("octree_digest", C.c_uint64),
        ("status", C.c_int), ("slot", C.c_int),
    ]


class ServerStats(C.Structure):
    _fields_ = [("elapsed_s", C.c_double), ("max_in_flight", C.c_int),
                ("max_in_flight_bytes", C.c_size_t), ("completed", C.c_int64),
                ("max_device_bytes", C.c_size_t)]


class ServedResult(C.Structure):
    _fields_ = [("status", C.c_int), ("dim", C.c_int), ("bits", C.c_int), ("slot", C.c_int),
                ("n", C.c_int64), ("m", C.c_int64), ("octree_size", C.c_int64),
                ("octree_digest", C.c_uint64), ("arrival_s", C.c_double),
                ("admit_s", C.c_double), ("done_s", C.c_double), ("octree", C.c_void_p)]


PGVD_NOT_READY = 11


class GvdResult(C.Structure):
    _fields_ = [("iterations", C.c_int), ("n_qpoints", C.c_int64), ("octree_size", C.c_int64),
                ("n_segments", C.c_int64), ("n_extra", C.c_int64), ("n_boxes", C.c_int64),
                ("qpoints", C.c_void_p), ("octree", C.c_void_p), ("ci", C.c_void_p),
                ("extra", C.c_void_p), ("samples", C.c_void_p), ("origins", C.c_void_p),
                ("lengths", C.c_void_p)]


def lib():
    """Load libpgvd.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m pgvd_amd.build` "
                              "(or __graft_entry__.build()) first")
        L = C.CDLL(LIB_PATH)
        vp, i64, i32 = C.c_void_p, C.c_int64, C.c_int
        sig = {
            "pgvd_last_error": ([], C.c_char_p),
            "pgvd_version": ([], C.c_char_p),
            "pgvd_next_pow2": ([i32], i32),
            "pgvd_ctx_create": ([i32, C.POINTER(vp)], i32),
            "pgvd_ctx_destroy": ([vp], i32),
            "pgvd_ctx_stream": ([vp, C.POINTER(vp)], i32),
            "pgvd_ctx_set_pool_limit": ([vp, C.c_size_t], i32),
            "pgvd_ctx_pool_bytes": ([vp, C.POINTER(C.c_size_t)], i32),
            "pgvd_debug_fail_launch": ([C.c_char_p], i32),
            "pgvd_debug_gather_queue": ([i32], i32),
            "pgvd_synchronize": ([vp, vp], i32),
            "pgvd_alloc": ([vp, C.c_size_t, C.POINTER(vp)], i32),
            "pgvd_free": ([vp, vp], i32),
            "pgvd_memcpy_htod": ([vp, vp, vp, C.c_size_t, vp], i32),
            "pgvd_memcpy_dtoh": ([vp, vp, vp, C.c_size_t, vp], i32),
            "pgvd_copy_kernel": ([vp, vp, vp, C.c_size_t, vp], i32),
            "pgvd_digest": ([vp, vp, C.c_size_t, vp, C.POINTER(C.c_uint64)], i32),
            "pgvd_timing_enable": ([vp, i32], i32),
            "pgvd_timing_reset": ([vp], i32),
            "pgvd_timing_read": ([vp, C.c_char_p, C.POINTER(C.c_double), C.POINTER(i64)], i32),
            "pgvd_bbox": ([vp, vp, i64, i32, vp, vp], i32),
            "pgvd_quantize": ([vp, vp, i64, i32, i32, vp, i32, vp, C.POINTER(i64), vp], i32),
            "pgvd_points_to_morton": ([vp, vp, i32, i64, i32, i32, vp, i32, vp], i32),
            "pgvd_find_leaf": ([vp, vp, i64, i32, i32, vp, i64, vp, vp, vp, vp, vp, vp], i32),
            "pgvd_cell_walk": ([vp, vp, i64, i32, vp, i64, vp, vp, i64, C.POINTER(i64), vp, vp],
                               i32),
            "pgvd_octree_format_text": ([vp, vp, i64, i32, vp, i64, C.POINTER(i64), vp], i32),
            "pgvd_octree_parse_text": ([vp, vp, i64, i32, vp, i64, C.POINTER(i64), vp], i32),
            "pgvd_octree_save": ([vp, vp, i64, i32, C.c_char_p, i32], i32),
            "pgvd_octree_load": ([vp, C.c_char_p, i32, vp, i64, C.POINTER(i64)], i32),
            "pgvd_radix_sort": ([vp, vp, vp, i64, i32, i32, vp], i32),
            "pgvd_unique_sorted": ([vp, vp, vp, i64, i32, C.POINTER(i64), vp], i32),
            "pgvd_stream_scan": ([vp, vp, vp, i64, vp], i32),
            "pgvd_build_brt": ([vp, vp, i64, i32, i32, vp, vp, vp], i32),
            "pgvd_local_splits": ([vp, vp, i64, i32, i32, vp, vp], i32),
            "pgvd_brt_to_octree": ([vp, vp, vp, i64, i32, i32, vp, i64, C.POINTER(i64), vp], i32),
            "pgvd_build_octree": ([vp, vp, i64, i32, i32, vp, C.POINTER(BuildOpts), vp], i32),
            "pgvd_build_octree_finish": ([vp, C.POINTER(OctreeResult)], i32),
            "pgvd_build_octree_noalloc": ([vp, i64, i32, i32, vp, C.POINTER(BuildOpts), vp, i64,
                                           vp, vp, vp], i32),
            "pgvd_batch_create": ([vp, C.POINTER(i64), i64, i32, i32, C.POINTER(BatchOpts),
                                   C.POINTER(vp)], i32),
            "pgvd_batch_destroy": ([vp], i32),
            "pgvd_batch_run": ([vp, vp], i32),
            "pgvd_batch_finish": ([vp, C.POINTER(BatchResult)], i32),
            "pgvd_batch_tune": ([vp, vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32),
                                 C.POINTER(i32)], i32),
            "pgvd_batch_last_ms": ([vp, C.POINTER(C.c_float)], i32),
            "pgvd_server_create": ([i32, i32, C.c_size_t, C.POINTER(vp)], i32),
            "pgvd_server_destroy": ([vp], i32),
            "pgvd_server_run": ([vp, C.POINTER(Request), i64, C.POINTER(ServerStats)], i32),
            "pgvd_server_submit": ([vp, vp, i64, i32, i32, C.POINTER(i64)], i32),
            "pgvd_server_poll": ([vp, i64, i32, C.POINTER(ServedResult)], i32),
            "pgvd_server_release": ([vp, i64], i32),
            "pgvd_build_footprint": ([i64, i32, i32], C.c_size_t),
            "pgvd_keys_to_big_unsigned": ([vp, i32, i64, vp], i32),
            "pgvd_big_unsigned_to_keys": ([vp, i64, i32, vp], i32),
            "pgvd_compare_big_unsigned": ([vp, vp], i32),
            "pgvd_brt_to_ref_nodes": ([vp, vp, i64, i32, vp], i32),
            "pgvd_ref_nodes_to_brt": ([vp, i64, i32, vp, vp], i32),
            "pgvd_polyline_segments": ([vp, vp, C.POINTER(i64), i64, C.POINTER(C.c_float),
                                        C.c_float, i32, vp, vp, C.POINTER(i64), vp], i32),
            "pgvd_multi_cells": ([vp, vp, vp, i64, i32, vp, i64, vp, vp, vp], i32),
            "pgvd_fit_boxes": ([vp, vp, i64, C.c_float, vp, vp, vp, i64, vp, vp, C.POINTER(i64),
                                C.POINTER(i64), vp, vp], i32),
            "pgvd_find_multi_cells": ([vp, vp, i64, C.c_float, vp, vp, vp, vp, i64,
                                       C.POINTER(i64), C.POINTER(i64), vp], i32),
            "pgvd_gvd_refine": ([vp, vp, C.POINTER(i64), i64, i32, C.POINTER(C.c_float), i32,
                                 C.c_float, C.POINTER(i64), C.POINTER(GvdResult), vp], i32),
            "pgvd_host_alloc": ([C.c_size_t, C.POINTER(vp)], i32),
            "pgvd_host_free": ([vp], i32),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _LIB = L
    return _LIB


def exported_symbols() -> list[str]:
    """Function names declared in include/pgvd.h."""
    import re
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(pgvd_\w+)\s*\(", text, re.M)))


def key_bytes_for(mbits: int) -> int:
    return 4 if mbits <= 32 else (8 if mbits <= 64 else 16)


def key_dtype(key_bytes: int):
    return {4: np.uint32, 8: np.uint64, 16: np.uint64}[key_bytes]


BRT_DTYPES = {
    4: np.dtype([("lcp", "<u4"), ("left", "<i4"), ("lcp_length", "<u2"),
                 ("left_leaf", "u1"), ("right_leaf", "u1")]),
    8: np.dtype([("lcp", "<u8"), ("left", "<i4"), ("lcp_length", "<u2"),
                 ("left_leaf", "u1"), ("right_leaf", "u1")]),
    16: np.dtype([("lcp_lo", "<u8"), ("lcp_hi", "<u8"), ("left", "<i4"), ("lcp_length", "<u2"),
                  ("left_leaf", "u1"), ("right_leaf", "u1")]),
}
assert BRT_DTYPES[4].itemsize == 12 and BRT_DTYPES[8].itemsize == 16
# pgvd_brt_node128 is 8-byte aligned: 16 + 4 + 2 + 1 + 1 = 24 bytes
assert BRT_DTYPES[16].itemsize == 24


# The reference's records (C/BigUnsigned.h:8-16, C/BrtNode.h:10-21), as
# pgvd_big_unsigned / pgvd_ref_brt_node lay them out.
BIG_UNSIGNED = np.dtype([("len", "<i4"), ("blk", "u1", (11,)), ("pad", "u1")])
REF_BRT_NODE = np.dtype([("left", "<i4"), ("left_leaf", "u1"), ("right_leaf", "u1"),
                         ("pad", "u1", (2,)), ("lcp", BIG_UNSIGNED), ("lcp_length", "<i4"),
                         ("parent", "<i4")])
assert BIG_UNSIGNED.itemsize == 16 and REF_BRT_NODE.itemsize == 32


def to_big_unsigned(words: np.ndarray, key_bytes: int) -> np.ndarray:
    """Native key words (the device key layout) -> BigUnsigned records."""
    words = np.ascontiguousarray(words)
    n = words.size // (2 if key_bytes == 16 else 1)
    out = np.zeros(n, BIG_UNSIGNED)
    rc = lib().pgvd_keys_to_big_unsigned(words.ctypes.data, key_bytes, n, out.ctypes.data)
    if rc != 0:
        raise PgvdError(rc, lib().pgvd_last_error().decode())
    return out


def from_big_unsigned(bu: np.ndarray, key_bytes: int) -> np.ndarray:
    """BigUnsigned records -> native key words."""
    bu = np.ascontiguousarray(bu, BIG_UNSIGNED)
    out = np.zeros(bu.size * (2 if key_bytes == 16 else 1), key_dtype(key_bytes))
    rc = lib().pgvd_big_unsigned_to_keys(bu.ctypes.data, bu.size, key_bytes, out.ctypes.data)
    if rc != 0:
        raise PgvdError(rc, lib().pgvd_last_error().decode())
    return out


def to_ref_brt_nodes(brt: np.ndarray, parent: np.ndarray, key_bytes: int) -> np.ndarray:
    """Native BRT records + parent links -> the reference's BrtNode records."""
    brt = np.ascontiguousarray(brt)
    parent = np.ascontiguousarray(parent, np.int32)
    out = np.zeros(parent.size, REF_BRT_NODE)
    rc = lib().pgvd_brt_to_ref_nodes(brt.ctypes.data, parent.ctypes.data, parent.size, key_bytes,
                                     out.ctypes.data)
    if rc != 0:
        raise PgvdError(rc, lib().pgvd_last_error().decode())
    return out


def keys_to_words(lo: np.ndarray, hi, key_bytes: int) -> np.ndarray:
    """(lo, hi) u64 key halves -> the device key array layout."""
    lo = np.asarray(lo, np.uint64)
    if key_bytes == 4:
        return lo.astype(np.uint32)
    if key_bytes == 8:
        return lo.copy()
    hi = np.zeros_like(lo) if hi is None else np.asarray(hi, np.uint64)
    out = np.empty(lo.size * 2, np.uint64)
    out[0::2] = lo
    out[1::2] = hi
    return out


def words_to_keys(words: np.ndarray, key_bytes: int):
    """device key array -> (lo, hi) u64 halves."""
    if key_bytes == 16:
        return words[0::2].copy(), words[1::2].copy()
    lo = words.astype(np.uint64)
    return lo, np.zeros_like(lo)


def debug_gather_queue(entries: int) -> None:
    """Test hook: the octree gather's queues hold at most `entries` entries
    (0: full size); results must not change."""
    rc = lib().pgvd_debug_gather_queue(int(entries))
    if rc != 0:
        raise PgvdError(rc, lib().pgvd_last_error().decode())


def debug_fail_launch(stage) -> None:
    """Test hook: the next launch check of `stage` on this thread reports
    PGVD_ERR_HIP (None clears it)."""
    lib().pgvd_debug_fail_launch(stage.encode() if stage else None)


@dataclass
class DeviceBuffer:
    dev: "Device"
    ptr: int
    nbytes: int

    def free(self):
        if self.ptr:
            self.dev._check(lib().pgvd_free(self.dev.ctx, self.ptr))
            self.ptr = 0


class Device:
    """One pgvd context (one HIP stream and a keyed buffer pool) on a GPU."""

    def __init__(self, device: int = 0):
        L = lib()
        ctx = C.c_void_p()
        rc = L.pgvd_ctx_create(device, C.byref(ctx))
        if rc != 0:
            raise PgvdError(rc, L.pgvd_last_error().decode())
        self.ctx = ctx
        self._bufs: list[DeviceBuffer] = []

    # -- plumbing ---------------------------------------------------------
    def _check(self, rc: int):
        if rc != 0:
            raise PgvdError(rc, lib().pgvd_last_error().decode())

    def close(self):
        if self.ctx:
            for b in self._bufs:
                if b.ptr:
                    lib().pgvd_free(self.ctx, b.ptr)
                    b.ptr = 0
            lib().pgvd_ctx_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_pool_limit(self, nbytes: int) -> None:
        """Most device bytes this context's pool may hold (0: no limit);
        a call that would pass it raises PgvdError PGVD_ERR_ALLOC."""
        self._check(lib().pgvd_ctx_set_pool_limit(self.ctx, int(nbytes)))

    def pool_bytes(self) -> int:
        out = C.c_size_t()
        self._check(lib().pgvd_ctx_pool_bytes(self.ctx, C.byref(out)))
        return int(out.value)

    def alloc(self, nbytes: int) -> DeviceBuffer:
        p = C.c_void_p()
        self._check(lib().pgvd_alloc(self.ctx, max(int(nbytes), 1), C.byref(p)))
        b = DeviceBuffer(self, p.value, int(nbytes))
        self._bufs.append(b)
        return b

    def upload(self, arr: np.ndarray) -> DeviceBuffer:
        arr = np.ascontiguousarray(arr)
        b = self.alloc(arr.nbytes)
        if arr.nbytes:
            self._check(lib().pgvd_memcpy_htod(self.ctx, b.ptr, arr.ctypes.data, arr.nbytes, None))
        return b

    def download(self, ptr, dtype, count: int) -> np.ndarray:
        out = np.empty(int(count), dtype)
        p = ptr.ptr if isinstance(ptr, DeviceBuffer) else ptr
        if out.nbytes:
            self._check(lib().pgvd_memcpy_dtoh(self.ctx, out.ctypes.data, p, out.nbytes, None))
        return out

    def synchronize(self):
        self._check(lib().pgvd_synchronize(self.ctx, None))

    def digest(self, ptr, nbytes: int) -> int:
        """Position-dependent 64-bit digest of a device range (pgvd_digest)."""
        p = ptr.ptr if isinstance(ptr, DeviceBuffer) else int(ptr or 0)
        out = C.c_uint64()
        self._check(lib().pgvd_digest(self.ctx, p, int(nbytes), None, C.byref(out)))
        return out.value

    def copy_kernel(self, dst: DeviceBuffer, src: DeviceBuffer, nbytes: int):
        self._check(lib().pgvd_copy_kernel(self.ctx, dst.ptr, src.ptr, nbytes, None))

    def timing(self, on: bool = True, sort_span_only: bool = False, gather_only: bool = False):
        """Stage timing: every stage; or only one event pair per build around
        the radix-sort passes (sort_span_only) or the octree gather kernel
        (gather_only)."""
        mode = (3 if gather_only else 2 if sort_span_only else 1) if on else 0
        self._check(lib().pgvd_timing_enable(self.ctx, mode))

    def timing_reset(self):
        self._check(lib().pgvd_timing_reset(self.ctx))

    def timing_read(self, stage: str):
        ms = C.c_double()
        cnt = C.c_int64()
        self._check(lib().pgvd_timing_read(self.ctx, stage.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value

    # -- stages -----------------------------------------------------------
    def bbox(self, pts: np.ndarray):
        pts = np.ascontiguousarray(pts, np.float32)
        n, dim = pts.shape
        d = self.upload(pts)
        out = np.zeros(2 * dim, np.float32)
        self._check(lib().pgvd_bbox(self.ctx, d.ptr, n, dim, out.ctypes.data, None))
        d.free()
        return out[:dim], out[dim:]

    def quantize(self, pts: np.ndarray, bits: int, bb=None, clamped: bool = False) -> np.ndarray:
        pts = np.ascontiguousarray(pts, np.float32)
        n, dim = pts.shape
        d = self.upload(pts)
        q = self.alloc(n * dim * 8)
        nout = C.c_int64()
        cbb = None
        if bb is not None:
            cbb_arr = np.concatenate([np.asarray(bb[0], np.float32), np.asarray(bb[1], np.float32)])
            cbb = cbb_arr.ctypes.data
        try:
            self._check(lib().pgvd_quantize(self.ctx, d.ptr, n, dim, bits, cbb, int(clamped),
                                            q.ptr, C.byref(nout), None))
            out = self.download(q, np.int64, n * dim).reshape(n, dim)
        finally:
            d.free()
            q.free()
        return out[: nout.value]

    def points_to_morton(self, q: np.ndarray, bits: int, key_bytes: int | None = None):
        """Keys of integer points, padded with zero keys to next_pow2(n)."""
        q = np.ascontiguousarray(q)
        n, dim = q.shape
        cb = q.dtype.itemsize
        kb = key_bytes or key_bytes_for(dim * bits)
        padded = lib().pgvd_next_pow2(n)
        d = self.upload(q)
        k = self.alloc(padded * kb)
        try:
            self._check(lib().pgvd_points_to_morton(self.ctx, d.ptr, cb, n, dim, bits, k.ptr, kb, None))
            words = self.download(k, key_dtype(kb), padded * (2 if kb == 16 else 1))
        finally:
            d.free()
            k.free()
        return words_to_keys(words, kb)

    def find_leaf(self, q: np.ndarray, bits: int, children: np.ndarray, leaf: np.ndarray) -> dict:
        """OctreeUtils::FindLeaf for every row of q (int32 grid points) on an
        octree given as children (O x 2^dim) and leaf (O) host arrays."""
        q = np.ascontiguousarray(q, dtype=np.int32)
        n, dim = q.shape
        O = children.shape[0]
        oct_ = np.ascontiguousarray(np.concatenate([children.reshape(O, -1), leaf.reshape(O, 1)],
                                                   axis=1).astype(np.int32))
        dq, do = self.upload(q), self.upload(oct_)
        outs = {k: self.alloc(max(n, 1) * 4) for k in ("parent", "octant", "data", "level", "status")}
        try:
            self.find_leaf_device(dq, n, dim, bits, do, O, outs)
            res = {k: self.download(b, np.uint32 if k == "status" else np.int32, n)
                   for k, b in outs.items()}
        finally:
            for b in [dq, do, *outs.values()]:
                b.free()
        return res

    def find_leaf_device(self, dq: DeviceBuffer, n: int, dim: int, bits: int, octree,
                         octree_size: int, outs: dict) -> None:
        """pgvd_find_leaf on device buffers (asynchronous); octree is a
        DeviceBuffer or a device pointer (e.g. OctreeResult.octree)."""
        optr = octree.ptr if isinstance(octree, DeviceBuffer) else octree
        st = outs.get("status")
        self._check(lib().pgvd_find_leaf(self.ctx, dq.ptr, n, dim, bits, optr, octree_size,
                                         outs["parent"].ptr, outs["octant"].ptr,
                                         outs["data"].ptr, outs["level"].ptr,
                                         st.ptr if st is not None else None, None))

    def cell_walk(self, segs: np.ndarray, bits: int, children: np.ndarray,
                  leaf: np.ndarray) -> dict:
        """CellWalk for every row (ax, ay, bx, by) of segs on a 2D octree given as
        host arrays: offsets (n + 1), cells (rows of origin x, origin y, width,
        parent, octant) and status."""
        segs = np.ascontiguousarray(segs, dtype=np.float32).reshape(-1, 4)
        n = segs.shape[0]
        O = children.shape[0]
        oct_ = np.ascontiguousarray(np.concatenate([children.reshape(O, -1), leaf.reshape(O, 1)],
                                                   axis=1).astype(np.int32))
        ds, do = self.upload(segs), self.upload(oct_)
        doff, dst = self.alloc(8 * (n + 1)), self.alloc(4 * max(n, 1))
        total = C.c_int64()
        dc = None
        try:
            self._check(lib().pgvd_cell_walk(self.ctx, ds.ptr, n, bits, do.ptr, O, doff.ptr, None,
                                             0, C.byref(total), dst.ptr, None))
            dc = self.alloc(20 * max(total.value, 1))
            self._check(lib().pgvd_cell_walk(self.ctx, ds.ptr, n, bits, do.ptr, O, doff.ptr,
                                             dc.ptr, total.value, C.byref(total), dst.ptr, None))
            return {"offsets": self.download(doff, np.int64, n + 1),
                    "cells": self.download(dc, np.int32, 5 * total.value).reshape(-1, 5),
                    "status": self.download(dst, np.int32, n)}
        finally:
            for b in (ds, do, doff, dst, dc):
                if b is not None:
                    b.free()

    # ---- GVD refinement (2D; viewer/Octree2.cpp:127-190, 366-506) ----
    def polyline_segments(self, verts: np.ndarray, counts, bb_min, bbw: float, bits: int):
        """FindMultiCells' segments: obj2Oct of every vertex and vertex i to i+1
        of every polygon, labelled by polygon.  Returns (segs (k, 4), labels)."""
        v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 2)
        cnt = np.ascontiguousarray(counts, dtype=np.int64)
        nseg = int(np.maximum(cnt - 1, 0).sum())
        dv = self.upload(v)
        ds, dl = self.alloc(16 * max(nseg, 1)), self.alloc(4 * max(nseg, 1))
        n = C.c_int64()
        bmin = (C.c_float * 2)(*[float(x) for x in bb_min])
        try:
            self._check(lib().pgvd_polyline_segments(
                self.ctx, dv.ptr, cnt.ctypes.data_as(C.POINTER(C.c_int64)), cnt.size, bmin,
                float(bbw), bits, ds.ptr, dl.ptr, C.byref(n), None))
            return (self.download(ds, np.float32, 4 * n.value).reshape(-1, 4),
                    self.download(dl, np.int32, n.value))
        finally:
            for b in (dv, ds, dl):
                b.free()

    def multi_cells(self, segs: np.ndarray, labels, bits: int, children: np.ndarray,
                    leaf: np.ndarray) -> dict:
        """The CellIntersections of every node of a 2D octree (host arrays) for
        the labelled segments: labels (O, 4, 2), segs (O, 4, 2, 4), and each
        walk's status."""
        segs = np.ascontiguousarray(segs, dtype=np.float32).reshape(-1, 4)
        lab = np.ascontiguousarray(labels, dtype=np.int32)
        n = segs.shape[0]
        O = children.shape[0]
        oct_ = np.ascontiguousarray(np.concatenate([children.reshape(O, -1), leaf.reshape(O, 1)],
                                                   axis=1).astype(np.int32))
        ds, dl, do = self.upload(segs), self.upload(lab), self.upload(oct_)
        dci, dst = self.alloc(160 * O), self.alloc(4 * max(n, 1))
        try:
            self._check(lib().pgvd_multi_cells(self.ctx, ds.ptr, dl.ptr, n, bits, do.ptr, O,
                                               dci.ptr, dst.ptr, None))
            raw = self.download(dci, np.int32, 40 * O).reshape(O, 40)
            return {"labels": raw[:, :8].reshape(O, 4, 2).copy(),
                    "segs": raw[:, 8:].copy().view(np.float32).reshape(O, 4, 2, 4),
                    "status": self.download(dst, np.int32, n)}
        finally:
            for b in (ds, dl, do, dci, dst):
                b.free()

    def fit_boxes(self, pairs: np.ndarray, min_d: float = 1.0) -> dict:
        """Geom::FitBoxes for every row (A's ax ay bx by, B's) of pairs:
        samples, origins, lengths, per-pair offsets and status.  Pairs that
        fail keep their status; the call's PGVD_ERR_GEOMETRY is not raised."""
        pairs = np.ascontiguousarray(pairs, dtype=np.float32).reshape(-1, 8)
        n = pairs.shape[0]
        dp = self.upload(pairs)
        dso, dbo, dst = self.alloc(8 * (n + 1)), self.alloc(8 * (n + 1)), self.alloc(4 * max(n, 1))
        ns, nb = C.c_int64(), C.c_int64()
        bufs = [dp, dso, dbo, dst]
        try:
            rc = lib().pgvd_fit_boxes(self.ctx, dp.ptr, n, min_d, None, None, None, 0, dso.ptr,
                                      dbo.ptr, C.byref(ns), C.byref(nb), dst.ptr, None)
            if rc not in (0, 7, 12):
                self._check(rc)
            cap = max(ns.value, nb.value, 1)
            ds, do, dln = self.alloc(8 * cap), self.alloc(8 * cap), self.alloc(4 * cap)
            bufs += [ds, do, dln]
            rc = lib().pgvd_fit_boxes(self.ctx, dp.ptr, n, min_d, ds.ptr, do.ptr, dln.ptr, cap,
                                      dso.ptr, dbo.ptr, C.byref(ns), C.byref(nb), dst.ptr, None)
            if rc not in (0, 12):
                self._check(rc)
            return {"samples": self.download(ds, np.float32, 2 * ns.value).reshape(-1, 2),
                    "origins": self.download(do, np.float32, 2 * nb.value).reshape(-1, 2),
                    "lengths": self.download(dln, np.float32, nb.value),
                    "sample_offsets": self.download(dso, np.int64, n + 1),
                    "box_offsets": self.download(dbo, np.int64, n + 1),
                    "status": self.download(dst, np.int32, n)}
        finally:
            for b in bufs:
                b.free()

    def find_multi_cells(self, ci_labels: np.ndarray, ci_segs: np.ndarray,
                         min_d: float = 1.0) -> dict:
        """FindMultiCells' sampling of CellIntersections records (host arrays
        as multi_cells returns them): extra points, samples, origins, lengths."""
        lab = np.ascontiguousarray(ci_labels, dtype=np.int32).reshape(-1, 8)
        O = lab.shape[0]
        rec = np.concatenate([lab, np.ascontiguousarray(ci_segs, dtype=np.float32)
                              .reshape(O, 32).view(np.int32)], axis=1)
        dci = self.upload(np.ascontiguousarray(rec))
        ns, nb = C.c_int64(), C.c_int64()
        bufs = [dci]
        try:
            rc = lib().pgvd_find_multi_cells(self.ctx, dci.ptr, O, min_d, None, None, None, None, 0,
                                             C.byref(ns), C.byref(nb), None)
            if rc not in (0, 7):
                self._check(rc)
            cap = max(ns.value, nb.value, 1)
            de, ds, do, dln = (self.alloc(8 * cap), self.alloc(8 * cap), self.alloc(8 * cap),
                               self.alloc(4 * cap))
            bufs += [de, ds, do, dln]
            self._check(lib().pgvd_find_multi_cells(self.ctx, dci.ptr, O, min_d, de.ptr, ds.ptr,
                                                    do.ptr, dln.ptr, cap, C.byref(ns),
                                                    C.byref(nb), None))
            return {"extra": self.download(de, np.int32, 2 * ns.value).reshape(-1, 2),
                    "samples": self.download(ds, np.float32, 2 * ns.value).reshape(-1, 2),
                    "origins": self.download(do, np.float32, 2 * nb.value).reshape(-1, 2),
                    "lengths": self.download(dln, np.float32, nb.value)}
        finally:
            for b in bufs:
                b.free()

    def gvd_refine_device(self, dverts: DeviceBuffer, counts, bits: int, iterations: int,
                          custom_bb=None, min_d: float = 1.0):
        """pgvd_gvd_refine on device vertices: (GvdResult, extra points per
        iteration)."""
        cnt = np.ascontiguousarray(counts, dtype=np.int64)
        hist = np.zeros(max(iterations, 1), np.int64)
        bb = None if custom_bb is None else (C.c_float * 4)(*[float(x) for x in
                                                              np.asarray(custom_bb).reshape(-1)])
        r = GvdResult()
        self._check(lib().pgvd_gvd_refine(self.ctx, dverts.ptr,
                                          cnt.ctypes.data_as(C.POINTER(C.c_int64)), cnt.size,
                                          bits, bb, iterations, min_d,
                                          hist.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(r),
                                          None))
        return r, hist[: r.iterations].tolist()

    def gvd_refine(self, verts: np.ndarray, counts, bits: int, iterations: int,
                   custom_bb=None, min_d: float = 1.0) -> dict:
        """Octree2::build(const Polylines&) with FindMultiCells on the device;
        the last build's points, octree, CellIntersections and extra points
        as host arrays."""
        v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 2)
        dv = self.upload(v)
        try:
            r, hist = self.gvd_refine_device(dv, counts, bits, iterations, custom_bb, min_d)
            O = r.octree_size
            out = {"iterations": r.iterations, "extra_per_iteration": hist,
                   "qpoints": self.download(r.qpoints, np.int32, 2 * r.n_qpoints).reshape(-1, 2),
                   "n_segments": r.n_segments,
                   "extra": self.download(r.extra, np.int32, 2 * r.n_extra).reshape(-1, 2)
                   if r.n_extra else np.zeros((0, 2), np.int32),
                   "samples": self.download(r.samples, np.float32, 2 * r.n_extra).reshape(-1, 2)
                   if r.n_extra else np.zeros((0, 2), np.float32),
                   "origins": self.download(r.origins, np.float32, 2 * r.n_boxes).reshape(-1, 2)
                   if r.n_boxes else np.zeros((0, 2), np.float32),
                   "lengths": self.download(r.lengths, np.float32, r.n_boxes)
                   if r.n_boxes else np.zeros(0, np.float32),
                   "octree_size": O}
            if O:
                octv = self.download(r.octree, np.int32, 5 * O).reshape(O, 5)
                raw = self.download(r.ci, np.int32, 40 * O).reshape(O, 40)
                out.update(children=octv[:, :4].copy(), leaf=octv[:, 4].copy(),
                           ci_labels=raw[:, :8].reshape(O, 4, 2).copy(),
                           ci_segs=raw[:, 8:].copy().view(np.float32).reshape(O, 4, 2, 4))
            return out
        finally:
            dv.free()

    # ---- the octree's text and binary forms (C/OctNode.h:58-98) ----
    def format_octree_text(self, octree, O: int, dim: int) -> bytes:
        """operator<< of O device records (a DeviceBuffer or pointer) as bytes."""
        optr = octree.ptr if isinstance(octree, DeviceBuffer) else octree
        n = C.c_int64()
        self._check(lib().pgvd_octree_format_text(self.ctx, optr, O, dim, None, 0, C.byref(n), None))
        buf = self.alloc(n.value)
        try:
            self._check(lib().pgvd_octree_format_text(self.ctx, optr, O, dim, buf.ptr, n.value,
                                                      C.byref(n), None))
            return self.download(buf, np.uint8, n.value).tobytes()
        finally:
            buf.free()

    def parse_octree_text(self, text: bytes, dim: int) -> np.ndarray:
        """operator>> of text; returns the O x (2^dim + 1) int32 records."""
        stride = (1 << dim) + 1
        cap = len(text) // (2 * stride) + 1      # a record token takes two bytes at least
        dt = self.upload(np.frombuffer(text, np.uint8))
        do = self.alloc(cap * stride * 4)
        O = C.c_int64()
        try:
            self._check(lib().pgvd_octree_parse_text(self.ctx, dt.ptr, len(text), dim, do.ptr, cap,
                                                     C.byref(O), None))
            return self.download(do, np.int32, O.value * stride).reshape(O.value, stride)
        finally:
            dt.free()
            do.free()

    def save_octree(self, octree, O: int, dim: int, path: str, binary: bool = False) -> None:
        optr = octree.ptr if isinstance(octree, DeviceBuffer) else octree
        self._check(lib().pgvd_octree_save(self.ctx, optr, O, dim, os.fsencode(path),
                                           1 if binary else 0))

    def load_octree(self, path: str, dim: int) -> np.ndarray:
        """Reads an octree file (either form) to the device and returns its records."""
        stride = (1 << dim) + 1
        size = os.path.getsize(path)
        cap = max(size // (2 * stride), size // (4 * stride)) + 1
        do = self.alloc(cap * stride * 4)
        O = C.c_int64()
        try:
            self._check(lib().pgvd_octree_load(self.ctx, os.fsencode(path), dim, do.ptr, cap,
                                               C.byref(O)))
            return self.download(do, np.int32, O.value * stride).reshape(O.value, stride)
        finally:
            do.free()

    def radix_sort(self, lo, hi=None, mbits: int = 64, vals=None, key_bytes: int | None = None):
        kb = key_bytes or key_bytes_for(mbits)
        words = keys_to_words(lo, hi, kb)
        n = np.asarray(lo).size
        dk = self.upload(words)
        dv = self.upload(np.ascontiguousarray(vals, np.uint32)) if vals is not None else None
        try:
            self._check(lib().pgvd_radix_sort(self.ctx, dk.ptr, dv.ptr if dv else None, n, kb,
                                              mbits, None))
            out = words_to_keys(self.download(dk, key_dtype(kb), words.size), kb)
            outv = self.download(dv, np.uint32, n) if dv else None
        finally:
            dk.free()
            if dv:
                dv.free()
        return (out, outv) if vals is not None else out

    def unique_sorted(self, lo, hi=None, key_bytes: int = 8):
        words = keys_to_words(lo, hi, key_bytes)
        n = np.asarray(lo).size
        di = self.upload(words)
        do = self.alloc(max(words.nbytes, 1))
        m = C.c_int64()
        try:
            self._check(lib().pgvd_unique_sorted(self.ctx, di.ptr, do.ptr, n, key_bytes,
                                                 C.byref(m), None))
            out = self.download(do, key_dtype(key_bytes), m.value * (2 if key_bytes == 16 else 1))
        finally:
            di.free()
            do.free()
        return words_to_keys(out, key_bytes)

    def stream_scan(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.uint32)
        di = self.upload(x)
        do = self.alloc(max(x.nbytes, 4))
        try:
            self._check(lib().pgvd_stream_scan(self.ctx, di.ptr, do.ptr, x.size, None))
            return self.download(do, np.uint32, x.size)
        finally:
            di.free()
            do.free()

    def build_brt(self, lo, hi, mbits: int, key_bytes: int | None = None):
        kb = key_bytes or key_bytes_for(mbits)
        words = keys_to_words(lo, hi, kb)
        m = np.asarray(lo).size
        dk = self.upload(words)
        nn = max(m - 1, 1)
        db = self.alloc(nn * BRT_DTYPES[kb].itemsize)
        dp = self.alloc(nn * 4)
        try:
            self._check(lib().pgvd_build_brt(self.ctx, dk.ptr, m, kb, mbits, db.ptr, dp.ptr, None))
            raw = self.download(db, np.uint8, nn * BRT_DTYPES[kb].itemsize)
            brt = raw.view(BRT_DTYPES[kb])[: m - 1]
            parent = self.download(dp, np.int32, nn)[: m - 1]
        finally:
            dk.free()
            db.free()
            dp.free()
        return brt, parent

    def local_splits(self, brt: np.ndarray, m: int, dim: int, key_bytes: int) -> np.ndarray:
        db = self.upload(np.ascontiguousarray(brt).view(np.uint8))
        dl = self.alloc(m * 4)
        try:
            self._check(lib().pgvd_local_splits(self.ctx, db.ptr, m, key_bytes, dim, dl.ptr, None))
            return self.download(dl, np.uint32, m)
        finally:
            db.free()
            dl.free()

    def brt_to_octree(self, brt: np.ndarray, parent: np.ndarray, m: int, dim: int,
                      key_bytes: int, capacity: int | None = None):
        nc = 1 << dim
        cap = capacity if capacity is not None else max(2 * m, 16)
        db = self.upload(np.ascontiguousarray(brt).view(np.uint8) if m > 1 else np.zeros(64, np.uint8))
        dp = self.upload(np.ascontiguousarray(parent, np.int32) if m > 1 else np.zeros(1, np.int32))
        do = self.alloc(cap * (nc + 1) * 4)
        osz = C.c_int64()
        try:
            rc = lib().pgvd_brt_to_octree(self.ctx, db.ptr, dp.ptr, m, key_bytes, dim, do.ptr,
                                          cap, C.byref(osz), None)
            self._check(rc)
            raw = self.download(do, np.int32, osz.value * (nc + 1)).reshape(osz.value, nc + 1)
        finally:
            db.free()
            dp.free()
            do.free()
        return raw[:, :nc].copy(), raw[:, nc].copy()

    def build_octree_async(self, pts: DeviceBuffer, n: int, dim: int, bits: int,
                           labels: DeviceBuffer | None = None, custom_bb=None,
                           clamped: bool = False, quantized_input: bool = False,
                           no_host_wait: bool = False):
        opts = BuildOpts()
        opts.no_host_wait = int(no_host_wait)
        if custom_bb is not None:
            opts.use_custom_bb = 1
            for k in range(dim):
                opts.bb_min[k] = float(custom_bb[0][k])
                opts.bb_max[k] = float(custom_bb[1][k])
        opts.clamped = int(clamped)
        opts.quantized_input = int(quantized_input)
        self._check(lib().pgvd_build_octree(self.ctx, pts.ptr, n, dim, bits,
                                            labels.ptr if labels else None, C.byref(opts), None))

    def build_octree_finish(self) -> OctreeResult:
        r = OctreeResult()
        self._check(lib().pgvd_build_octree_finish(self.ctx, C.byref(r)))
        return r

    def fetch(self, r: OctreeResult) -> dict:
        """Download every output of a finished build into numpy arrays."""
        kb, m, O, dim = r.key_bytes, r.m, r.octree_size, r.dim
        nc = 1 << dim
        words = self.download(r.keys, key_dtype(kb), m * (2 if kb == 16 else 1))
        lo, hi = words_to_keys(words, kb)
        ni = max(m - 1, 0)
        brt = self.download(r.brt, np.uint8, ni * BRT_DTYPES[kb].itemsize).view(BRT_DTYPES[kb])
        parent = self.download(r.parent, np.int32, ni)
        ps = self.download(r.prefix_sums, np.uint32, m)
        raw = self.download(r.octree, np.int32, O * (nc + 1)).reshape(O, nc + 1)
        out = dict(m=m, octree_size=O, malformed=r.malformed, bucket_sort=r.bucket_sort,
                   keys_lo=lo, keys_hi=hi,
                   brt=brt, parent=parent, prefix_sums=ps,
                   children=raw[:, :nc].copy(), leaf=raw[:, nc].copy())
        if kb == 16:
            out["lcp_lo"], out["lcp_hi"] = brt["lcp_lo"].copy(), brt["lcp_hi"].copy()
        else:
            out["lcp_lo"], out["lcp_hi"] = brt["lcp"].astype(np.uint64), np.zeros(ni, np.uint64)
        out["left"] = brt["left"].copy()
        out["left_leaf"] = brt["left_leaf"].copy()
        out["right_leaf"] = brt["right_leaf"].copy()
        out["lcp_length"] = brt["lcp_length"].astype(np.int32)
        if r.key_conflict:
            bits = self.download(r.key_conflict, np.uint32, (m + 31) // 32)
            flags = np.unpackbits(bits.view(np.uint8), bitorder="little")[:m]
            out["key_conflict"] = flags.astype(np.uint8)
            out["node_conflict"] = self.download(r.node_conflict, np.uint32, O)
        return out

    def build_octree(self, pts: np.ndarray, bits: int, labels=None, custom_bb=None,
                     clamped: bool = False, no_host_wait: bool = False) -> dict:
        """Whole path from host points (float n x dim, or int32 grid points)."""
        pts = np.ascontiguousarray(pts)
        n, dim = pts.shape
        quant = pts.dtype == np.int32
        if not quant:
            pts = pts.astype(np.float32, copy=False)
        dp = self.upload(pts)
        dl = self.upload(np.ascontiguousarray(labels, np.uint32)) if labels is not None else None
        try:
            self.build_octree_async(dp, n, dim, bits, dl, custom_bb, clamped, quant, no_host_wait)
            r = self.build_octree_finish()
            return self.fetch(r)
        finally:
            dp.free()
            if dl:
                dl.free()


class Batch:
    """A batched-form plan (pgvd_batch_*): many independent point sets of
    uneven size through one segmented pipeline on one Device."""

    def __init__(self, dev: "Device", sizes, dim: int, bits: int, quantized: bool = False,
                 clamped: bool = False, use_graph: bool = True, sort_ipt: int = 0,
                 brt_rows: int = 0, digit_bits: int = 0, sort_waves: int = 0):
        self.dev = dev
        self.sizes = np.ascontiguousarray(sizes, np.int64)
        opts = BatchOpts(int(quantized), int(clamped), int(use_graph), int(sort_ipt),
                         int(brt_rows), int(digit_bits), int(sort_waves))
        h = C.c_void_p()
        dev._check(lib().pgvd_batch_create(
            dev.ctx, self.sizes.ctypes.data_as(C.POINTER(C.c_int64)), self.sizes.size, dim, bits,
            C.byref(opts), C.byref(h)))
        self.h = h
        self.dim, self.bits = dim, bits

    def close(self):
        if self.h:
            lib().pgvd_batch_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def run(self, pts):
        p = pts.ptr if isinstance(pts, DeviceBuffer) else int(pts)
        self.dev._check(lib().pgvd_batch_run(self.h, p))

    def finish(self) -> BatchResult:
        r = BatchResult()
        self.dev._check(lib().pgvd_batch_finish(self.h, C.byref(r)))
        return r

    def last_ms(self) -> float:
        ms = C.c_float()
        self.dev._check(lib().pgvd_batch_last_ms(self.h, C.byref(ms)))
        return ms.value

    def tune(self, pts) -> tuple[int, int, int, int]:
        """(sort_ipt, brt_rows, digit_bits, sort_waves) the tuner kept."""
        p = pts.ptr if isinstance(pts, DeviceBuffer) else int(pts)
        a, b, d, w = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self.dev._check(lib().pgvd_batch_tune(self.h, p, C.byref(a), C.byref(b), C.byref(d),
                                              C.byref(w)))
        return a.value, b.value, d.value, w.value

    @staticmethod
    def counts(r: BatchResult) -> dict:
        n = r.count
        return dict(m=np.ctypeslib.as_array(r.m, (n,)).copy(),
                    octree_size=np.ctypeslib.as_array(r.octree_size, (n,)).copy(),
                    key_offset=np.ctypeslib.as_array(r.key_offset, (n,)).copy(),
                    octree_offset=np.ctypeslib.as_array(r.octree_offset, (n,)).copy())

    def fetch_all(self, r: BatchResult) -> dict:
        """Download the whole batch output: per-slot arrays and the octree."""
        d, kb, nc = self.dev, r.key_bytes, 1 << r.dim
        slots = int(self.sizes.sum() + self.sizes.size)
        words = d.download(r.keys, key_dtype(kb), slots * (2 if kb == 16 else 1))
        brt = d.download(r.brt, np.uint8, slots * BRT_DTYPES[kb].itemsize).view(BRT_DTYPES[kb])
        parent = d.download(r.parent, np.int32, slots)
        ps = d.download(r.prefix_sums, np.uint32, slots)
        oct_ = d.download(r.octree, np.int32, r.total_octree * (nc + 1)).reshape(-1, nc + 1)
        return dict(words=words, brt=brt, parent=parent, ps=ps, octree=oct_, **self.counts(r))

    @staticmethod
    def instance(all_: dict, i: int, key_bytes: int, dim: int) -> dict:
        """Instance i of fetch_all() in the layout of Device.fetch()."""
        nc = 1 << dim
        k0, m = int(all_["key_offset"][i]), int(all_["m"][i])
        o0, O = int(all_["octree_offset"][i]), int(all_["octree_size"][i])
        w = all_["words"]
        words = w[2 * k0:2 * (k0 + m)] if key_bytes == 16 else w[k0:k0 + m]
        lo, hi = words_to_keys(words, key_bytes)
        ni = max(m - 1, 0)
        brt = all_["brt"][k0:k0 + ni]
        raw = all_["octree"][o0:o0 + O]
        out = dict(m=m, octree_size=O, keys_lo=lo, keys_hi=hi, brt=brt,
                   parent=all_["parent"][k0:k0 + ni].copy(), prefix_sums=all_["ps"][k0:k0 + m].copy(),
                   children=raw[:, :nc].copy(), leaf=raw[:, nc].copy(),
                   left=brt["left"].copy(), left_leaf=brt["left_leaf"].copy(),
                   right_leaf=brt["right_leaf"].copy(), lcp_length=brt["lcp_length"].astype(np.int32))
        if key_bytes == 16:
            out["lcp_lo"], out["lcp_hi"] = brt["lcp_lo"].copy(), brt["lcp_hi"].copy()
        else:
            out["lcp_lo"], out["lcp_hi"] = brt["lcp"].astype(np.uint64), np.zeros(ni, np.uint64)
        return out


class PinnedArray:
    """A numpy view of pinned host memory (pgvd_host_alloc)."""

    def __init__(self, shape, dtype=np.float32):
        self.dtype = np.dtype(dtype)
        nbytes = int(np.prod(shape)) * self.dtype.itemsize
        p = C.c_void_p()
        rc = lib().pgvd_host_alloc(max(nbytes, 1), C.byref(p))
        if rc != 0:
            raise PgvdError(rc, lib().pgvd_last_error().decode())
        self.ptr = p.value
        buf = (C.c_char * max(nbytes, 1)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=self.dtype, count=int(np.prod(shape))).reshape(shape)

    def free(self):
        if self.ptr:
            lib().pgvd_host_free(self.ptr)
            self.ptr = None
            self.array = None


class Server:
    """The served form (pgvd_server_*): `slots` contexts with pooled buffers,
    admission in arrival order against an HBM budget, overlapped builds."""

    def __init__(self, device: int = 0, slots: int = 4, hbm_budget: int = 64 << 30):
        h = C.c_void_p()
        rc = lib().pgvd_server_create(device, slots, hbm_budget, C.byref(h))
        if rc != 0:
            raise PgvdError(rc, lib().pgvd_last_error().decode())
        self.h = h

    def close(self):
        if self.h:
            lib().pgvd_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @staticmethod
    def footprint(n: int, dim: int, bits: int) -> int:
        return int(lib().pgvd_build_footprint(n, dim, bits))

    def submit(self, host_pts_ptr: int, n: int, dim: int, bits: int) -> int:
        """Queue a build of n points at host_pts_ptr (kept valid until the
        request is complete); returns its ticket."""
        t = C.c_int64()
        rc = lib().pgvd_server_submit(self.h, host_pts_ptr, n, dim, bits, C.byref(t))
        if rc != 0:
            raise PgvdError(rc, lib().pgvd_last_error().decode())
        return t.value

    def poll(self, ticket: int, wait: bool = False):
        """The ServedResult of a complete request, or None when not yet."""
        r = ServedResult()
        rc = lib().pgvd_server_poll(self.h, ticket, int(wait), C.byref(r))
        if rc == PGVD_NOT_READY:
            return None
        if rc != 0:
            raise PgvdError(rc, lib().pgvd_last_error().decode())
        return r

    def release(self, ticket: int):
        rc = lib().pgvd_server_release(self.h, ticket)
        if rc != 0:
            raise PgvdError(rc, lib().pgvd_last_error().decode())

    def run(self, reqs) -> ServerStats:
        """reqs: a ctypes array of Request sorted by arrival_s (filled in place)."""
        st = ServerStats()
        rc = lib().pgvd_server_run(self.h, reqs, len(reqs), C.byref(st))
        if rc != 0:
            raise PgvdError(rc, lib().pgvd_last_error().decode())
        return st

