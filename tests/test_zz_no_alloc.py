"""The allocation-free whole path (pgvd_build_octree_noalloc) against the
pooled one (pgvd_build_octree + pgvd_build_octree_finish), through the C ABI.

Each case runs two builds in a row on the shared static workspace, the
second smaller than the first, and every output word must equal the pooled
build's: the octree records, the sizes, and (with labels) the per-node
conflict masks."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import pgvd_amd
from pgvd_amd.workloads import star_loops, uniform

pytestmark = pytest.mark.gpu

NOALLOC_BYTES = 64 << 20


def _stream(dev):
    s = C.c_void_p()
    dev._check(pgvd_amd.lib().pgvd_ctx_stream(dev.ctx, C.byref(s)))
    return s


def _noalloc(dev, pts: np.ndarray, bits: int, labels=None, cap: int | None = None,
             want_rc: int = 0):
    """Runs the allocation-free build; returns (state words, children, leaf,
    node_conflict) read back after the stream is done."""
    L = pgvd_amd.lib()
    n, dim = pts.shape
    stride = (1 << dim) + 1
    cap = cap if cap is not None else 2 * n + 1024
    dp = dev.upload(np.ascontiguousarray(pts, np.float32))
    dl = dev.upload(np.ascontiguousarray(labels, np.uint32)) if labels is not None else None
    doct = dev.alloc(cap * stride * 4)
    dnc = dev.alloc(cap * 4) if labels is not None else None
    dst = dev.alloc(16)
    try:
        print(f"noalloc build n={n} dim={dim} bits={bits} labels={labels is not None} cap={cap}",
              flush=True)
        rc = L.pgvd_build_octree_noalloc(dp.ptr, n, dim, bits, dl.ptr if dl else None, None,
                                         doct.ptr, cap, dnc.ptr if dnc else None, dst.ptr,
                                         _stream(dev))
        assert rc == want_rc, (rc, L.pgvd_last_error().decode())
        if rc != 0:
            return None
        dev.synchronize()
        state = dev.download(dst, np.int32, 4)
        O = int(state[2])
        if state[0] != 0:
            return state, None, None, None
        raw = dev.download(doct, np.int32, O * stride).reshape(O, stride)
        nc = dev.download(dnc, np.uint32, O) if dnc else None
        return state, raw[:, :stride - 1].copy(), raw[:, stride - 1].copy(), nc
    finally:
        for b in (dp, dl, doct, dnc, dst):
            if b:
                b.free()


def _check_pair(dev, pts, bits, labels=None):
    want = dev.build_octree(pts, bits, labels=labels)
    state, children, leaf, nc = _noalloc(dev, pts, bits, labels)
    assert state[0] == 0, f"status {state[0]}"
    assert int(state[1]) == want["m"]
    assert int(state[2]) == want["octree_size"]
    assert int(state[3]) == want["malformed"]
    np.testing.assert_array_equal(children, want["children"])
    np.testing.assert_array_equal(leaf, want["leaf"])
    if labels is not None:
        np.testing.assert_array_equal(nc, want["node_conflict"])


@pytest.mark.parametrize("dim,bits,sizes", [
    (3, 21, (1 << 20, 100_003)),      # 63-bit keys: the largest that fits, then a smaller one
    (2, 16, (65_536, 4_097)),         # 32-bit keys, zero-key padding on the second
])
def test_noalloc_matches_pooled(dev, dim, bits, sizes):
    for i, n in enumerate(sizes):
        _check_pair(dev, uniform(n, dim, seed=40 + i), bits)


def test_noalloc_labels_126bit(dev):
    # 126-bit keys with loop-id labels, then a smaller labelled build
    for i, n in enumerate((1 << 18, 30_001)):
        pts, lab = star_loops(n, loops=16, seed=50 + i)
        # a few points copied onto other loops so that some leaves conflict
        rng = np.random.default_rng(60 + i)
        k = n // 50
        src = rng.integers(0, n, k)
        dst = rng.integers(0, n, k)
        pts[dst] = pts[src]
        _check_pair(dev, pts, 42, labels=lab)


def test_noalloc_capacity_and_limit(dev):
    pts = uniform(50_000, 3, seed=70)
    want = dev.build_octree(pts, 21)
    # too small an octree buffer: reported on the device, nothing written
    state, *_ = _noalloc(dev, pts, 21, cap=want["octree_size"] - 1)
    assert state[0] == 7 and int(state[2]) == want["octree_size"]
    # the same workspace then serves an exact-capacity build
    state, children, leaf, _ = _noalloc(dev, pts, 21, cap=want["octree_size"])
    assert state[0] == 0
    np.testing.assert_array_equal(children, want["children"])
    # a build beyond the workspace is refused before anything is enqueued
    big = uniform(1 << 21, 3, seed=71)
    assert _noalloc(dev, big, 21, want_rc=8) is None
    assert "workspace" in pgvd_amd.lib().pgvd_last_error().decode()
