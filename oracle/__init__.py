"""Test infrastructure only: ctypes access to the serial oracle and the
reference checker.

* ``liboracle.so`` -- the quirk-exact serial restatement of the reference's
  octree-construction path (pgvd_oracle.c).
* ``_ref/libpgvdref{2,3}.so`` -- the reference's own C sources (C/*.c of
  edwardsjohnmartin/pgvd) linked with a small harness that restates
  opencl/Kernels.cpp:388-411.  Built only where the reference tree exists.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package; the product library (pgvd_amd) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = {}

i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


def compile_libs(ref: bool = True) -> None:
    """Compile liboracle.so (and _ref when the reference is present)."""
    target = "all" if ref else "oracle"
    subprocess.run(["make", "-s", "-C", HERE, target], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            compile_libs(ref=False)
        L = C.CDLL(path)
        L.oracle_next_pow2.argtypes = [C.c_int]
        L.oracle_next_pow2.restype = C.c_int
        L.oracle_bbox.argtypes = [f32p, C.c_int64, C.c_int, f32p, f32p]
        L.oracle_bbox.restype = C.c_float
        L.oracle_quantize.argtypes = [f32p, C.c_int64, C.c_int, C.c_int64, C.c_int,
                                      C.c_void_p, C.c_void_p, C.c_int, i64p]
        L.oracle_quantize.restype = C.c_int64
        L.oracle_points_to_keys.argtypes = [i64p, C.c_int64, C.c_int, C.c_int, u64p, u64p]
        L.oracle_compute_lcp_length.argtypes = [C.c_uint64] * 4 + [C.c_int]
        L.oracle_compute_lcp_length.restype = C.c_int
        L.oracle_stream_scan.argtypes = [u32p, u32p, C.c_int]
        L.oracle_cell_walk.argtypes = [f32p, C.c_int64, C.c_int, i32p, C.c_int64, i32p, i32p,
                                       C.c_void_p, C.c_int64]
        L.oracle_cell_walk.restype = C.c_int64
        L.oracle_octree_format_text.argtypes = [i32p, C.c_int64, C.c_int, C.c_char_p]
        L.oracle_octree_format_text.restype = C.c_int64
        L.oracle_octree_parse_text.argtypes = [C.c_char_p, C.c_int64, C.c_int, i32p, C.c_int64,
                                               C.POINTER(C.c_int64)]
        L.oracle_find_leaf.argtypes = [i64p, C.c_int64, C.c_int, C.c_int, i32p, C.c_int64,
                                       i32p, i32p, i32p, i32p, i32p]
        L.oracle_build.argtypes = [i64p, C.c_int64, C.c_int, C.c_int, C.c_void_p]
        L.oracle_build.restype = C.c_void_p
        L.oracle_build_from_keys.argtypes = [u64p, u64p, C.c_int64, C.c_int, C.c_int]
        L.oracle_build_from_keys.restype = C.c_void_p
        for f in ("oracle_tree_m", "oracle_tree_octree_size"):
            getattr(L, f).argtypes = [C.c_void_p]
            getattr(L, f).restype = C.c_int64
        L.oracle_tree_keys.argtypes = [C.c_void_p, u64p, u64p]
        L.oracle_tree_brt.argtypes = [C.c_void_p, i32p, u8p, u8p, u64p, u64p, i32p, i32p]
        L.oracle_tree_splits.argtypes = [C.c_void_p, u32p, u32p]
        L.oracle_tree_octree.argtypes = [C.c_void_p, i32p, i32p]
        L.oracle_tree_conflicts.argtypes = [C.c_void_p, u8p, u32p]
        L.oracle_tree_conflicts.restype = C.c_int
        L.oracle_tree_free.argtypes = [C.c_void_p]
        L.oracle_run_pipeline.argtypes = [f32p, C.c_int64, C.c_int, C.c_int, C.POINTER(C.c_int64)]
        L.oracle_run_pipeline.restype = C.c_int64
        L.oracle_brt_only.argtypes = [u64p, u64p, C.c_int64, C.c_int, i32p, u8p, u8p, u64p,
                                      u64p, i32p, i32p]
        L.oracle_brt_only.restype = C.c_int
        L.oracle_multi_intersection.argtypes = [f32p] * 4
        L.oracle_multi_intersection.restype = C.c_int
        L.oracle_fit_boxes.argtypes = [f32p, f32p, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int)]
        L.oracle_fit_boxes.restype = C.c_int64
        L.oracle_fit_boxes_undefined.argtypes = [f32p, f32p, C.c_float]
        L.oracle_fit_boxes_undefined.restype = C.c_int
        L.oracle_multi_cells.argtypes = [f32p, i32p, C.c_int64, C.c_int, i32p, C.c_int64, i32p,
                                         f32p, i32p]
        L.oracle_find_multi_cells.argtypes = [i32p, f32p, C.c_int64, C.c_float, C.c_void_p,
                                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                              C.POINTER(C.c_int64), C.POINTER(C.c_int)]
        L.oracle_find_multi_cells.restype = C.c_int64
        L.oracle_obj2oct.argtypes = [f32p, C.c_int64, f32p, C.c_float, C.c_int64, f32p]
        L.oracle_obj2oct.restype = C.c_int
        L.oracle_find_intersections.argtypes = [f32p, f32p, i32p, C.c_int, f32p]
        L.oracle_find_intersections.restype = C.c_int
        _LIB = L
    return _LIB


def ref_available(dim: int = 2) -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", f"libpgvdref{dim}.so"))


def ref_geom_available() -> bool:
    return os.path.exists(os.path.join(HERE, "_ref", "libpgvdgeom2.so"))


_REFG = None


def ref_geom():
    """The reference's 2D geometry (opencl/Geom.cpp, OctreeUtils.cpp,
    CellIntersections.h) behind ref_geom.cpp."""
    global _REFG
    if _REFG is None:
        L = C.CDLL(os.path.join(HERE, "_ref", "libpgvdgeom2.so"))
        L.refg_multi_intersection.argtypes = [f32p] * 4
        L.refg_multi_intersection.restype = C.c_int
        L.refg_fit_boxes.argtypes = [f32p, f32p, C.c_float, f32p, f32p, f32p, C.c_int64,
                                     C.POINTER(C.c_int64), C.POINTER(C.c_int)]
        L.refg_fit_boxes.restype = C.c_int64
        L.refg_find_intersections.argtypes = [f32p, f32p, i32p, C.c_int, f32p]
        L.refg_find_intersections.restype = C.c_int
        L.refg_mc_new.argtypes = [C.c_int64]
        L.refg_mc_new.restype = C.c_void_p
        L.refg_mc_free.argtypes = [C.c_void_p]
        L.refg_mc_visit.argtypes = [C.c_void_p, i32p, f32p, f32p, C.c_int]
        L.refg_mc_visit.restype = C.c_int
        L.refg_mc_read.argtypes = [C.c_void_p, i32p, f32p]
        _REFG = L
    return _REFG


def ref(dim: int = 2):
    """The reference's own C sources (DIM 2 or 3) behind ref_harness.cpp."""
    if dim not in _REF:
        path = os.path.join(HERE, "_ref", f"libpgvdref{dim}.so")
        L = C.CDLL(path)
        L.ref_compute_lcp_length.argtypes = [C.c_uint64] * 4 + [C.c_int]
        L.ref_compute_lcp_length.restype = C.c_int
        L.ref_compute_lcp.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_int,
                                      C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.ref_stream_scan.argtypes = [u32p, u32p, C.c_int]
        L.ref_next_pow2.argtypes = [C.c_int]
        L.ref_next_pow2.restype = C.c_int
        if dim == 2:
            L.ref_xyz2z.argtypes = [C.c_int, C.c_int, C.c_int,
                                    C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
            L.ref_build_octree.argtypes = [i32p, C.c_int, C.c_int, C.c_int]
            L.ref_build_octree.restype = C.c_void_p
        L.ref_build_from_keys.argtypes = [u64p, u64p, C.c_int, C.c_int]
        L.ref_build_from_keys.restype = C.c_void_p
        for f in ("ref_tree_status", "ref_tree_m", "ref_tree_octree_size"):
            getattr(L, f).argtypes = [C.c_void_p]
            getattr(L, f).restype = C.c_int
        L.ref_tree_keys.argtypes = [C.c_void_p, u64p, u64p]
        L.ref_tree_brt.argtypes = [C.c_void_p, i32p, u8p, u8p, u64p, u64p, i32p, i32p]
        L.ref_tree_splits.argtypes = [C.c_void_p, u32p, u32p]
        L.ref_tree_octree.argtypes = [C.c_void_p, i32p, i32p]
        L.ref_tree_free.argtypes = [C.c_void_p]
        L.ref_tree_keys_raw.argtypes = [C.c_void_p, C.c_void_p]
        L.ref_tree_brt_raw.argtypes = [C.c_void_p, C.c_void_p]
        L.ref_layout.argtypes = [i32p]
        L.ref_bu_from_bits.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_void_p]
        L.ref_compare_bu.argtypes = [C.c_void_p, C.c_void_p]
        L.ref_compare_bu.restype = C.c_int
        L.ref_istream_octree.argtypes = [C.c_char_p, C.c_int64, C.c_void_p, C.c_int64,
                                         C.POINTER(C.c_int64)]
        L.ref_istream_octree.restype = C.c_int
        L.ref_brt_only.argtypes = [u64p, u64p, C.c_int, C.c_int, i32p, u8p, u8p, u64p, u64p,
                                   i32p, i32p]
        L.ref_brt_only.restype = C.c_int
        _REF[dim] = L
    return _REF[dim]


# ---------------------------------------------------------------------------
# numpy-level helpers
# ---------------------------------------------------------------------------

def next_pow2(n: int) -> int:
    return lib().oracle_next_pow2(int(n))


def bbox(pts: np.ndarray):
    pts = np.ascontiguousarray(pts, dtype=np.float32)
    dim = pts.shape[1]
    mn = np.zeros(dim, np.float32)
    mx = np.zeros(dim, np.float32)
    dw = lib().oracle_bbox(pts.reshape(-1), pts.shape[0], dim, mn, mx)
    return mn, mx, np.float32(dw)


def quantize(pts: np.ndarray, width: int, bb=None, clamped: bool = False) -> np.ndarray:
    """Karras.cpp:69-99.  Raises ValueError where the reference throws."""
    pts = np.ascontiguousarray(pts, dtype=np.float32)
    n, dim = pts.shape
    out = np.zeros(max(n, 1) * dim, np.int64)
    if bb is not None:
        mn = np.ascontiguousarray(bb[0], np.float32)
        mx = np.ascontiguousarray(bb[1], np.float32)
        r = lib().oracle_quantize(pts.reshape(-1), n, dim, int(width), 1,
                                  mn.ctypes.data, mx.ctypes.data, int(clamped), out)
    else:
        r = lib().oracle_quantize(pts.reshape(-1), n, dim, int(width), 0, None, None,
                                  int(clamped), out)
    if r < 0:
        raise ValueError("bad coordinate")
    return out[: r * dim].reshape(r, dim)


def morton(q: np.ndarray, bits: int):
    """xyz2z for every row of q (n x dim int64), no padding -> (lo, hi)."""
    q = np.ascontiguousarray(q, dtype=np.int64)
    n, dim = q.shape
    lo = np.zeros(n, np.uint64)
    hi = np.zeros(n, np.uint64)
    lib().oracle_points_to_keys(q.reshape(-1), n, dim, bits, lo, hi)
    return lo, hi


def lcp_length(a: int, b: int, mbits: int) -> int:
    M = (1 << 64) - 1
    return lib().oracle_compute_lcp_length(a & M, a >> 64, b & M, b >> 64, mbits)


def find_leaf(q: np.ndarray, bits: int, children: np.ndarray, leaf: np.ndarray) -> dict:
    """OctreeUtils::FindLeaf for every row of q (grid points) on an octree
    given as children (O x 2^dim) and leaf (O)."""
    q = np.ascontiguousarray(q, dtype=np.int64)
    n, dim = q.shape
    O = children.shape[0]
    oct_ = np.ascontiguousarray(np.concatenate([children.reshape(O, -1), leaf.reshape(O, 1)],
                                               axis=1).astype(np.int32))
    out = {k: np.zeros(n, np.int32) for k in ("parent", "octant", "data", "level", "status")}
    lib().oracle_find_leaf(q.reshape(-1), n, dim, bits, oct_.reshape(-1), O, out["parent"],
                           out["octant"], out["data"], out["level"], out["status"])
    return out


def cell_walk(segs: np.ndarray, bits: int, children: np.ndarray, leaf: np.ndarray) -> dict:
    """CellWalk for every row (ax, ay, bx, by) of segs on a 2D octree:
    per-segment counts, status, offsets, and the visited cells as rows of
    (origin x, origin y, width, parent, octant)."""
    segs = np.ascontiguousarray(segs, dtype=np.float32).reshape(-1, 4)
    n = segs.shape[0]
    O = children.shape[0]
    oct_ = np.ascontiguousarray(np.concatenate([children.reshape(O, -1), leaf.reshape(O, 1)],
                                               axis=1).astype(np.int32))
    counts, status = np.zeros(n, np.int32), np.zeros(n, np.int32)
    total = lib().oracle_cell_walk(segs.reshape(-1), n, bits, oct_.reshape(-1), O, counts, status,
                                   None, 0)
    cells = np.zeros((max(total, 1), 5), np.int32)
    lib().oracle_cell_walk(segs.reshape(-1), n, bits, oct_.reshape(-1), O, counts, status,
                           cells.ctypes.data, total)
    offsets = np.zeros(n + 1, np.int64)
    np.cumsum(counts, out=offsets[1:])
    return {"counts": counts, "status": status, "offsets": offsets, "cells": cells[:total]}


def octree_format_text(octree: np.ndarray, dim: int) -> bytes:
    """operator<< of the octree (O x (2^dim + 1) int32 records) as bytes."""
    o = np.ascontiguousarray(octree, dtype=np.int32).reshape(-1)
    O = o.size // ((1 << dim) + 1)
    n = lib().oracle_octree_format_text(o, O, dim, None)
    buf = C.create_string_buffer(int(n))
    lib().oracle_octree_format_text(o, O, dim, buf)
    return buf.raw[:n]


def octree_parse_text(text: bytes, dim: int, cap: int | None = None):
    """operator>> of an octree; returns the O x (2^dim + 1) records, or None
    when the text is malformed (-1) or holds more than cap nodes (-2)."""
    cap = len(text) if cap is None else cap
    out = np.zeros(max(cap, 1) * ((1 << dim) + 1), np.int32)
    O = C.c_int64()
    rc = lib().oracle_octree_parse_text(text, len(text), dim, out, cap, C.byref(O))
    if rc:
        return None
    return out[: O.value * ((1 << dim) + 1)].reshape(O.value, (1 << dim) + 1)


def stream_scan(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.uint32)
    out = np.zeros_like(x)
    lib().oracle_stream_scan(x, out, x.size)
    return out


def _collect(L, h, prefix: str, dim: int, labels: bool = False) -> dict:
    g = lambda name: getattr(L, prefix + name)
    m = int(g("tree_m")(h))
    O = int(g("tree_octree_size")(h))
    ni = max(m - 1, 0)
    lo = np.zeros(m, np.uint64)
    hi = np.zeros(m, np.uint64)
    g("tree_keys")(h, lo, hi)
    left = np.zeros(max(ni, 1), np.int32)
    ll = np.zeros(max(ni, 1), np.uint8)
    rl = np.zeros(max(ni, 1), np.uint8)
    lcl = np.zeros(max(ni, 1), np.uint64)
    lch = np.zeros(max(ni, 1), np.uint64)
    lcn = np.zeros(max(ni, 1), np.int32)
    par = np.zeros(max(ni, 1), np.int32)
    g("tree_brt")(h, left, ll, rl, lcl, lch, lcn, par)
    ls = np.zeros(m, np.uint32)
    ps = np.zeros(m, np.uint32)
    g("tree_splits")(h, ls, ps)
    nc = 1 << dim
    ch = np.zeros(max(O, 1) * nc, np.int32)
    lf = np.zeros(max(O, 1), np.int32)
    g("tree_octree")(h, ch, lf)
    out = dict(m=m, octree_size=O, keys_lo=lo, keys_hi=hi,
               left=left[:ni], left_leaf=ll[:ni], right_leaf=rl[:ni],
               lcp_lo=lcl[:ni], lcp_hi=lch[:ni], lcp_length=lcn[:ni], parent=par[:ni],
               local_splits=ls, prefix_sums=ps,
               children=ch[: O * nc].reshape(O, nc), leaf=lf[:O])
    if labels:
        kf = np.zeros(m, np.uint8)
        nm = np.zeros(max(O, 1), np.uint32)
        L.oracle_tree_conflicts(h, kf, nm)
        out["key_conflict"] = kf
        out["node_conflict"] = nm[:O]
    return out


def build(q: np.ndarray, bits: int, labels=None) -> dict:
    """BuildOctree_s (opencl/Kernels.cpp:388-411) from quantized points."""
    q = np.ascontiguousarray(q, dtype=np.int64)
    n, dim = q.shape
    L = lib()
    lab = None
    if labels is not None:
        lab = np.ascontiguousarray(labels, dtype=np.uint32)
    h = L.oracle_build(q.reshape(-1), n, dim, bits, None if lab is None else lab.ctypes.data)
    if not h:
        raise ValueError("oracle_build rejected its arguments")
    try:
        return _collect(L, h, "oracle_", dim, labels=lab is not None)
    finally:
        L.oracle_tree_free(h)


def build_from_keys(lo: np.ndarray, hi, dim: int, mbits: int) -> dict:
    lo = np.ascontiguousarray(lo, dtype=np.uint64)
    hi = np.zeros_like(lo) if hi is None else np.ascontiguousarray(hi, dtype=np.uint64)
    L = lib()
    h = L.oracle_build_from_keys(lo, hi, lo.size, dim, mbits)
    if not h:
        raise ValueError("oracle_build_from_keys rejected its arguments")
    try:
        return _collect(L, h, "oracle_", dim)
    finally:
        L.oracle_tree_free(h)


def ref_build_octree_2d(q: np.ndarray, bits: int) -> dict:
    """The reference's own BuildOctree_s path (2D, int32 points)."""
    q = np.ascontiguousarray(q, dtype=np.int32)
    L = ref(2)
    h = L.ref_build_octree(q.reshape(-1), q.shape[0], bits, 2 * bits)
    try:
        if L.ref_tree_status(h) != 0:
            raise RuntimeError("reference harness could not map its key arena")
        return _collect(L, h, "ref_", 2)
    finally:
        L.ref_tree_free(h)


def ref_build_from_keys(lo: np.ndarray, hi, dim: int, mbits: int) -> dict:
    lo = np.ascontiguousarray(lo, dtype=np.uint64)
    hi = np.zeros_like(lo) if hi is None else np.ascontiguousarray(hi, dtype=np.uint64)
    L = ref(dim)
    h = L.ref_build_from_keys(lo, hi, lo.size, mbits)
    try:
        if L.ref_tree_status(h) != 0:
            raise RuntimeError("reference harness could not map its key arena")
        return _collect(L, h, "ref_", dim)
    finally:
        L.ref_tree_free(h)


def ref_xyz2z_2d(x: int, y: int, bits: int) -> int:
    lo = C.c_uint64()
    hi = C.c_uint64()
    ref(2).ref_xyz2z(int(x), int(y), int(bits), C.byref(lo), C.byref(hi))
    return int(lo.value) | (int(hi.value) << 64)


def ref_lcp_length(a: int, b: int, mbits: int, dim: int = 2) -> int:
    M = (1 << 64) - 1
    return ref(dim).ref_compute_lcp_length(a & M, a >> 64, b & M, b >> 64, mbits)


def ref_istream_octree(text: bytes, dim: int, cap: int | None = None):
    """The reference's stream input of an octree (C/OctNode.h:69-98) on
    std::istringstream: the records, or None where the stream fails."""
    cap = len(text) if cap is None else cap
    out = np.zeros(max(cap, 1) * ((1 << dim) + 1), np.int32)
    O = C.c_int64()
    rc = ref(dim).ref_istream_octree(text, len(text), out.ctypes.data, cap, C.byref(O))
    if rc:
        return None
    return out[: O.value * ((1 << dim) + 1)].reshape(O.value, (1 << dim) + 1)


# ---------------------------------------------------------------------------
# GVD refinement (2D): multi-cells, FitBoxes and the Karras iteration
# ---------------------------------------------------------------------------

def _f2(p) -> np.ndarray:
    return np.ascontiguousarray(p, dtype=np.float32).reshape(-1)


def multi_intersection(s0, s1) -> bool:
    """Geom::multi_intersection (opencl/Geom.cpp:109-151) of segments
    (ax, ay, bx, by)."""
    s0, s1 = _f2(s0), _f2(s1)
    return bool(lib().oracle_multi_intersection(s0[:2].copy(), s0[2:].copy(), s1[:2].copy(),
                                                s1[2:].copy()))


def fit_boxes(s0, s1, min_d: float = 1.0) -> dict:
    """Geom::FitBoxes (opencl/Geom.cpp:667-716): samples (k, 2), box origins
    (b, 2), box widths (b,) and status (0, or the error it reports)."""
    a, b = _f2(s0), _f2(s1)
    L = lib()
    nb, st = C.c_int64(), C.c_int()
    k = L.oracle_fit_boxes(a, b, min_d, None, None, None, 0, C.byref(nb), C.byref(st))
    cap = max(k, nb.value, 1)
    smp = np.zeros(2 * cap, np.float32)
    org = np.zeros(2 * cap, np.float32)
    ln = np.zeros(cap, np.float32)
    L.oracle_fit_boxes(a, b, min_d, smp.ctypes.data, org.ctypes.data, ln.ctypes.data, cap,
                       C.byref(nb), C.byref(st))
    return {"samples": smp[: 2 * k].reshape(k, 2), "origins": org[: 2 * nb.value].reshape(-1, 2),
            "lengths": ln[: nb.value], "status": st.value}


def fit_boxes_undefined(s0, s1, min_d: float = 1.0) -> bool:
    """True where the reference's FitBoxes reads a segment it never set
    (opencl/Geom.cpp:684-697), so its result depends on stack contents."""
    return bool(lib().oracle_fit_boxes_undefined(_f2(s0), _f2(s1), min_d))


def obj2oct(verts: np.ndarray, bbmin, bbw: float, bits: int) -> np.ndarray:
    """Octree2::obj2Oct (viewer/Octree2.cpp:215-225) of every vertex."""
    v = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 2)
    out = np.zeros_like(v)
    if lib().oracle_obj2oct(v.reshape(-1), v.shape[0], _f2(bbmin), np.float32(bbw), 1 << bits,
                            out.reshape(-1)):
        raise ValueError("obj2Oct cannot have coord less than zero")
    return out


def polyline_segments(octv: np.ndarray, counts) -> tuple[np.ndarray, np.ndarray]:
    """FindMultiCells' segments (viewer/Octree2.cpp:442-450): vertex i to
    i+1 of every polygon, polygon by polygon, labelled with the polygon's
    index."""
    segs, labels, at = [], [], 0
    for j, c in enumerate(np.asarray(counts, dtype=np.int64)):
        p = octv[at: at + c]
        if c >= 2:
            segs.append(np.concatenate([p[:-1], p[1:]], axis=1))
            labels.append(np.full(c - 1, j, np.int32))
        at += c
    if not segs:
        return np.zeros((0, 4), np.float32), np.zeros(0, np.int32)
    return (np.ascontiguousarray(np.concatenate(segs), dtype=np.float32),
            np.ascontiguousarray(np.concatenate(labels), dtype=np.int32))


def multi_cells(segs, labels, bits: int, children: np.ndarray, leaf: np.ndarray) -> dict:
    """Every segment's cell walk with the MCCallback visitor into the
    CellIntersections of every node: labels (O, 4, 2) with -1 for an empty
    slot, segments (O, 4, 2, 4), status per segment."""
    segs = np.ascontiguousarray(segs, dtype=np.float32).reshape(-1, 4)
    labels = np.ascontiguousarray(labels, dtype=np.int32)
    O = children.shape[0]
    oct_ = np.ascontiguousarray(np.concatenate([children.reshape(O, -1), leaf.reshape(O, 1)],
                                               axis=1).astype(np.int32))
    lab = np.zeros(max(O, 1) * 8, np.int32)
    sg = np.zeros(max(O, 1) * 32, np.float32)
    st = np.zeros(max(segs.shape[0], 1), np.int32)
    lib().oracle_multi_cells(segs.reshape(-1), labels, segs.shape[0], bits, oct_.reshape(-1), O,
                             lab, sg, st)
    return {"labels": lab[: 8 * O].reshape(O, 4, 2), "segs": sg[: 32 * O].reshape(O, 4, 2, 4),
            "status": st[: segs.shape[0]]}


def find_multi_cells(ci_labels: np.ndarray, ci_segs: np.ndarray, min_d: float = 1.0) -> dict:
    """FindMultiCells' sampling: extra points (k, 2) int32, samples (k, 2),
    box origins and widths, status."""
    lab = np.ascontiguousarray(ci_labels, dtype=np.int32).reshape(-1)
    sg = np.ascontiguousarray(ci_segs, dtype=np.float32).reshape(-1)
    O = lab.size // 8
    if O == 0:
        lab, sg = np.full(8, -1, np.int32), np.zeros(32, np.float32)
    L = lib()
    nb, st = C.c_int64(), C.c_int()
    k = L.oracle_find_multi_cells(lab, sg, O, min_d, None, None, None, None, 0, C.byref(nb),
                                  C.byref(st))
    cap = max(k, nb.value, 1)
    ex = np.zeros(2 * cap, np.int32)
    smp = np.zeros(2 * cap, np.float32)
    org = np.zeros(2 * cap, np.float32)
    ln = np.zeros(cap, np.float32)
    L.oracle_find_multi_cells(lab, sg, O, min_d, ex.ctypes.data, smp.ctypes.data, org.ctypes.data,
                              ln.ctypes.data, cap, C.byref(nb), C.byref(st))
    return {"extra": ex[: 2 * k].reshape(k, 2), "samples": smp[: 2 * k].reshape(k, 2),
            "origins": org[: 2 * nb.value].reshape(-1, 2), "lengths": ln[: nb.value],
            "status": st.value}


def gvd_refine(verts: np.ndarray, counts, bits: int, iterations: int, custom_bb=None,
               min_d: float = 1.0) -> dict:
    """Octree2::build(const Polylines&) (viewer/Octree2.cpp:127-190) with
    its FindMultiCells call (:186) in place: quantize the vertices, then
    build the octree, find the multi-cells and add their FitBoxes samples as
    extra points, until `iterations` builds or no extra points."""
    verts = np.ascontiguousarray(verts, dtype=np.float32).reshape(-1, 2)
    if custom_bb is not None:
        bbmin = np.asarray(custom_bb[0], np.float32)
        bbmax = np.asarray(custom_bb[1], np.float32)
    else:
        bbmin, bbmax = verts.min(axis=0), verts.max(axis=0)
    bbw = np.float32(max(bbmax[0] - bbmin[0], bbmax[1] - bbmin[1]))
    q = quantize(verts, 1 << bits).astype(np.int64)
    segs, labels = polyline_segments(obj2oct(verts, bbmin, bbw, bits), counts)
    extra = np.zeros((0, 2), np.int64)
    it, history = 0, []
    tree = ci = fm = None
    while True:
        q = np.concatenate([q, extra])
        extra = np.zeros((0, 2), np.int64)
        tree = build(q, bits) if q.shape[0] > 1 else None
        if tree is not None:
            ci = multi_cells(segs, labels, bits, tree["children"], tree["leaf"])
            fm = find_multi_cells(ci["labels"], ci["segs"], min_d)
            if fm["status"]:
                raise RuntimeError(f"FitBoxes error {fm['status']}")
            extra = fm["extra"].astype(np.int64)
        it += 1
        history.append(extra.shape[0])
        if not (it < iterations and extra.shape[0] > 0):
            break
    return {"qpoints": q, "tree": tree, "ci": ci, "multi": fm, "extra": extra,
            "iterations": it, "extra_per_iteration": history, "segs": segs, "labels": labels}


def ref_multi_intersection(s0, s1) -> bool:
    s0, s1 = _f2(s0), _f2(s1)
    return bool(ref_geom().refg_multi_intersection(s0[:2].copy(), s0[2:].copy(), s1[:2].copy(),
                                                   s1[2:].copy()))


def ref_fit_boxes(s0, s1, min_d: float = 1.0, cap: int = 1 << 16) -> dict:
    """The reference's Geom::FitBoxes on the segments."""
    a, b = _f2(s0), _f2(s1)
    smp = np.zeros(2 * cap, np.float32)
    org = np.zeros(2 * cap, np.float32)
    ln = np.zeros(cap, np.float32)
    nb, st = C.c_int64(), C.c_int()
    k = ref_geom().refg_fit_boxes(a, b, min_d, smp, org, ln, cap, C.byref(nb), C.byref(st))
    if k > cap or nb.value > cap:
        raise ValueError("reference FitBoxes output exceeds the capacity")
    return {"samples": smp[: 2 * k].reshape(k, 2), "origins": org[: 2 * nb.value].reshape(-1, 2),
            "lengths": ln[: nb.value], "status": st.value}


def find_intersections(a, b, origin, width: int) -> np.ndarray:
    """OctreeUtils::FindIntersections: rows (t, x, y)."""
    out = np.zeros(12, np.float32)
    n = lib().oracle_find_intersections(_f2(a), _f2(b), np.ascontiguousarray(origin, dtype=np.int32),
                                        int(width), out)
    return out[: 3 * n].reshape(n, 3)


def ref_find_intersections(a, b, origin, width: int) -> np.ndarray:
    out = np.zeros(12, np.float32)
    n = ref_geom().refg_find_intersections(_f2(a), _f2(b),
                                           np.ascontiguousarray(origin, dtype=np.int32),
                                           int(width), out)
    return out[: 3 * n].reshape(n, 3)


def ref_multi_cells(segs, labels, walk_cells, offsets, O: int) -> dict:
    """MCCallback on the reference's FindIntersections and CellIntersections
    for the given walk (cells of segment s at offsets[s]..offsets[s+1])."""
    L = ref_geom()
    h = L.refg_mc_new(O)
    try:
        segs = np.ascontiguousarray(segs, dtype=np.float32).reshape(-1, 4)
        cells = np.ascontiguousarray(walk_cells, dtype=np.int32).reshape(-1, 5)
        for s in range(segs.shape[0]):
            a, b = segs[s, :2].copy(), segs[s, 2:].copy()
            for c in range(int(offsets[s]), int(offsets[s + 1])):
                L.refg_mc_visit(h, cells[c].copy(), a, b, int(labels[s]))
        lab = np.zeros(max(O, 1) * 8, np.int32)
        sg = np.zeros(max(O, 1) * 32, np.float32)
        L.refg_mc_read(h, lab, sg)
    finally:
        L.refg_mc_free(h)
    return {"labels": lab[: 8 * O].reshape(O, 4, 2), "segs": sg[: 32 * O].reshape(O, 4, 2, 4)}


def _brt_only(fn, lo, hi, mbits: int) -> dict:
    lo = np.ascontiguousarray(lo, dtype=np.uint64)
    hi = np.zeros_like(lo) if hi is None else np.ascontiguousarray(hi, dtype=np.uint64)
    m = lo.size
    n = max(m - 1, 1)
    out = dict(left=np.zeros(n, np.int32), left_leaf=np.zeros(n, np.uint8),
               right_leaf=np.zeros(n, np.uint8), lcp_lo=np.zeros(n, np.uint64),
               lcp_hi=np.zeros(n, np.uint64), lcp_length=np.zeros(n, np.int32),
               parent=np.zeros(n, np.int32))
    rc = fn(lo, hi, m, mbits, out["left"], out["left_leaf"], out["right_leaf"], out["lcp_lo"],
            out["lcp_hi"], out["lcp_length"], out["parent"])
    if rc:
        raise ValueError(f"BRT of {m} keys failed ({rc})")
    return {k: v[: m - 1] for k, v in out.items()}


def brt_only(lo, hi, mbits: int) -> dict:
    """BuildBinaryRadixTree_s alone on sorted keys that may repeat."""
    return _brt_only(lib().oracle_brt_only, lo, hi, mbits)


def ref_brt_only(lo, hi, mbits: int, dim: int = 2) -> dict:
    """The reference's BuildBinaryRadixTree on the same keys (<= 88 bits)."""
    return _brt_only(ref(dim).ref_brt_only, lo, hi, mbits)
