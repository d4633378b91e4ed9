"""Parity of the HIP path (through the C ABI) with the oracle and with the
reference's golden fixtures.  Integer, key, index and tree outputs must be
bit-exact; there is no floating-point output on this path."""
from __future__ import annotations

import ctypes as C
import hashlib
import os

import numpy as np
import pytest

import oracle
import pgvd_amd as P
from conftest import TREE_FIELDS, assert_tree_equal
from pgvd_amd import workloads

pytestmark = pytest.mark.gpu

DTYPES = {"keys_lo": np.uint64, "keys_hi": np.uint64, "left": np.int32, "left_leaf": np.uint8,
          "right_leaf": np.uint8, "lcp_lo": np.uint64, "lcp_hi": np.uint64,
          "lcp_length": np.int32, "parent": np.int32, "prefix_sums": np.uint32,
          "children": np.int32, "leaf": np.int32}


def _digest(a, dtype):
    return hashlib.sha256(np.ascontiguousarray(np.asarray(a).astype(dtype)).tobytes()).hexdigest()


def _pts(n, dim, seed):
    rng = np.random.default_rng(seed)
    return rng.random((n, dim), dtype=np.float32) * 2 - 1


# ------------------------------------------------------------ whole path ----
@pytest.mark.parametrize("n,dim,bits", [
    (1, 2, 16), (2, 2, 16), (3, 3, 21), (5, 2, 16), (7, 3, 21), (8, 2, 16), (9, 3, 21),
    (1000, 2, 16), (4096, 3, 21), (4097, 3, 21), (65536, 2, 16), (100003, 3, 21),
    (3000, 3, 42), (2000, 2, 64), (5000, 2, 30), (3000, 3, 10), (300, 2, 1), (50000, 3, 7),
])
def test_build_octree_matches_oracle(dev, n, dim, bits):
    pts = _pts(n, dim, seed=n * 7 + dim * 3 + bits)
    got = dev.build_octree(pts, bits)
    want = oracle.build(oracle.quantize(pts, 1 << bits), bits)
    assert_tree_equal(got, want, where=f"n={n} dim={dim} bits={bits}")


def test_duplicates_and_degenerate_inputs(dev):
    pts = _pts(20000, 3, 1)
    pts[10000:] = pts[:10000]                                  # every point twice
    want = oracle.build(oracle.quantize(pts, 1 << 21), 21)
    assert_tree_equal(dev.build_octree(pts, 21), want, where="dupes")
    same = np.full((777, 3), 0.3, np.float32)                  # zero-size box (Karras.cpp:84-88)
    got = dev.build_octree(same, 21)
    assert got["m"] == 1 and got["octree_size"] == 1
    assert_tree_equal(got, oracle.build(oracle.quantize(same, 1 << 21), 21), where="same")
    line = np.zeros((5000, 2), np.float32)                    # points on one axis
    line[:, 0] = np.linspace(-1, 1, 5000, dtype=np.float32)
    assert_tree_equal(dev.build_octree(line, 16),
                      oracle.build(oracle.quantize(line, 1 << 16), 16), where="line")


def test_custom_bbox_and_clamped(dev):
    pts = _pts(5000, 2, 2) * np.float32(0.5)
    bb = (np.array([-1, -1], np.float32), np.array([1, 1], np.float32))
    for clamped in (False, True):
        got = dev.build_octree(pts, 16, custom_bb=bb, clamped=clamped)
        want = oracle.build(oracle.quantize(pts, 1 << 16, bb=bb, clamped=clamped), 16)
        assert_tree_equal(got, want, where=f"custom clamped={clamped}")


def test_bad_coordinate_is_an_error(dev):
    """Karras.cpp:56-63 throws for a coordinate below the box; the C ABI
    returns PGVD_ERR_BAD_COORD."""
    pts = _pts(1000, 2, 3)
    bb = (np.array([0, 0], np.float32), np.array([1, 1], np.float32))
    with pytest.raises(P.PgvdError) as e:
        dev.build_octree(pts, 16, custom_bb=bb)
    assert e.value.code == 3
    # the context stays usable
    assert dev.build_octree(pts, 16)["m"] > 0


def test_empty_input_is_an_error(dev):
    with pytest.raises(P.PgvdError) as e:
        dev.build_octree(np.zeros((0, 3), np.float32), 21)
    assert e.value.code == 2                      # opencl/Kernels.cpp:389-391


@pytest.mark.parametrize("m", [2, 3, 256, 257, 258, 513, 769, 1024, 1025, 4097, 65537])
def test_tree_sizes_at_chunk_edges(dev, m):
    """m distinct grid points including the origin (no zero key prepended):
    trees whose last node starts, ends or straddles a 256-node chunk, where
    the gather derives the prefix sums from per-chunk totals."""
    rng = np.random.default_rng(m)
    flat = rng.choice(1 << 20, m - 1, replace=False) + 1
    q = np.stack([flat & 1023, flat >> 10], axis=1).astype(np.int32)
    q = np.concatenate([np.zeros((1, 2), np.int32), q])
    want = oracle.build(q, 10)
    assert want["m"] == m
    assert_tree_equal(dev.build_octree(q, 10), want, where=f"m={m}")


def test_octree_larger_than_first_buffer(dev):
    """Pairs of grid points one step apart, scattered: each pair hangs off a
    chain of single-child octree nodes, so the octree outgrows the 2m nodes
    first allocated and the finish call refills it from the kept BRT."""
    rng = np.random.default_rng(11)
    c = rng.integers(0, 1 << 20, (20000, 3)) * 2
    q = np.concatenate([c, c + np.array([1, 0, 0])]).astype(np.int32)
    want = oracle.build(q, 21)
    assert want["octree_size"] > 2 * (want["m"] + 1)
    assert_tree_equal(dev.build_octree(q, 21), want, where="refill")


def test_int32_grid_input_matches_oracle(dev, golden_small):
    """Kernels::BuildOctree_p takes integer points; the reference fixture at
    bits = 30 through that entry."""
    z = golden_small
    q = z["int30/q"].astype(np.int32)
    got = dev.build_octree(q, 30)
    want = {k: z[f"int30/{k}"] for k in DTYPES if f"int30/{k}" in z}
    want["m"] = want["keys_lo"].size
    want["octree_size"] = want["children"].shape[0]
    assert_tree_equal(got, want, where="int30 fixture")


@pytest.mark.parametrize("name", ["n1000", "n5", "n1", "n2", "n4096", "n3000_dupes"])
def test_reference_fixtures_2d(dev, golden_small, name):
    z = golden_small
    got = dev.build_octree(z[f"{name}/pts"], 16)
    want = {k: z[f"{name}/{k}"] for k in DTYPES}
    want["m"] = want["keys_lo"].size
    want["octree_size"] = want["children"].shape[0]
    assert_tree_equal(got, want, where=name)


def test_reference_c1_digest(dev, golden_digests):
    d = golden_digests["2d_c1_65536"]
    got = dev.build_octree(_pts(d["n"], 2, d["seed"]), 16)
    assert got["m"] == d["m"] and got["octree_size"] == d["octree_size"]
    for k, dt in DTYPES.items():
        assert _digest(got[k], dt) == d[k], k


@pytest.mark.parametrize("name", ["c2_uniform", "c2_gauss"])
def test_full_size_3d_digests(dev, golden_digests, name):
    """2^24 points, 63-bit keys: every output bit-exact with the reference
    (children, BRT, links) and the oracle (3D leaf)."""
    d = golden_digests[name]
    if name == "c2_uniform":
        pts = workloads.uniform(1 << 24, 3, 2)
    else:
        pts, _ = workloads.gaussian_mixture(1 << 24)
    assert _digest(pts, np.float32) == d["pts"]
    got = dev.build_octree(pts, 21)
    if name == "c2_uniform":
        assert got["bucket_sort"] == 1      # uniform keys fill every bucket evenly
    assert got["m"] == d["m"] and got["octree_size"] == d["octree_size"]
    for k, dt in DTYPES.items():
        assert _digest(got[k], dt) == d[k], k
    # size-independent properties
    k = got["keys_lo"]
    assert (k[1:] > k[:-1]).all()
    assert got["parent"][0] == -1


def test_bucket_sort_large_and_clustered(dev):
    """A build of 2^23 + 4097 64-bit keys: the keys scatter by their top bits
    into buckets that each sort inside one workgroup; buckets that are too
    full (points near the centre) are split further below their highest
    varying bit, or, when they hold more than a quarter of the keys, the LSD
    passes sort.  Against the oracle at a size that is not a power of two."""
    n = (1 << 23) + 4097
    pts = _pts(n, 3, seed=23)
    got = dev.build_octree(pts, 21)
    assert got["bucket_sort"] == 1
    assert_tree_equal(got, oracle.build(oracle.quantize(pts, 1 << 21), 21), where="buckets")
    # 150 copies of each of 8 points: their buckets hold a sub-bucket too
    # large for the counting pass and are sorted by the in-LDS radix passes
    dup = pts.copy()
    for g in range(8):
        dup[1000 * g + 1: 1000 * g + 150] = dup[1000 * g]
    got = dev.build_octree(dup, 21)
    assert got["bucket_sort"] == 1
    assert_tree_equal(got, oracle.build(oracle.quantize(dup, 1 << 21), 21), where="dup buckets")
    # a point on the box's minimum corner has the zero key, which merges with
    # the zero key prepended at this size (n is not a power of two)
    corner = pts.copy()
    corner[5] = pts.min(axis=0)
    got = dev.build_octree(corner, 21)
    assert got["bucket_sort"] == 1 and got["keys_lo"][0] == 0
    assert_tree_equal(got, oracle.build(oracle.quantize(corner, 1 << 21), 21), where="zero key")
    # an eighth of the points near the centre: their buckets go through the
    # levels; half of them: the LSD passes sort instead
    for frac, bucket in ((8, 1), (2, 0)):
        clustered = pts.copy()
        clustered[: n // frac] *= np.float32(1e-3)
        got = dev.build_octree(clustered, 21)
        assert got["bucket_sort"] == bucket, frac
        assert_tree_equal(got, oracle.build(oracle.quantize(clustered, 1 << 21), 21),
                          where=f"clustered 1/{frac}")
        assert (got["keys_lo"][1:] > got["keys_lo"][:-1]).all()


@pytest.mark.parametrize("dim,bits", [(2, 24), (3, 13)])
def test_bucket_sort_other_key_widths(dev, dim, bits):
    """The bucket sort on 48-bit 2D keys and 39-bit 3D keys (fewer bits left
    below the buckets' own), against the oracle."""
    n = (1 << 23) + 999
    pts = _pts(n, dim, seed=dim * 100 + bits)
    got = dev.build_octree(pts, bits)
    assert got["bucket_sort"] == 1
    assert_tree_equal(got, oracle.build(oracle.quantize(pts, 1 << bits), bits),
                      where=f"dim={dim} bits={bits}")


@pytest.mark.parametrize("dim", [2, 3])
def test_fp32_split_quirk_stage_path(dev, golden_digests, dim):
    """2^24 keys whose root split takes the reference's FP32 value: BRT via
    the stage entry, then BinaryRadixToOctree through the order-exact path."""
    from tests.golden.gen_golden import quirk_keys
    d = golden_digests[f"quirk_{dim}d"]
    keys, s_star = quirk_keys()
    brt, parent = dev.build_brt(keys, None, 42, key_bytes=8)
    assert int(brt["left"][0]) == d["root_left"] != s_star
    assert _digest(brt["left"], np.int32) == d["left"]
    assert _digest(parent, np.int32) == d["parent"]
    assert _digest(brt["lcp_length"], np.int32) == d["lcp_length"]
    assert _digest(brt["lcp"], np.uint64) == d["lcp_lo"]
    ls = dev.local_splits(brt, keys.size, dim, 8)
    assert _digest(ls, np.uint32) == d["local_splits"]
    ch, lf = dev.brt_to_octree(brt, parent, keys.size, dim, 8)
    assert ch.shape[0] == d["octree_size"]
    assert _digest(ch, np.int32) == d["children"]
    assert _digest(lf, np.int32) == d["leaf"]


def test_fp32_split_quirk_whole_build(dev, golden_digests):
    """The same 2^24 keys as 2D grid points (21 bits/axis, x in the even
    bits) through the whole build: the root split takes the reference's
    FP32 value, the tree is malformed, and parents, local splits, their scan
    and the octree come from the order-exact path inside the pipeline."""
    from tests.golden.gen_golden import quirk_keys
    d = golden_digests["quirk_2d"]
    keys, s_star = quirk_keys()
    q = np.zeros((keys.size, 2), np.int32)
    for i in range(21):
        for k in range(2):
            bit = (keys >> np.uint64(2 * i + k)) & np.uint64(1)
            q[:, k] |= bit.astype(np.int32) << i
    got = dev.build_octree(q, 21)
    assert got["malformed"] == 1
    assert np.array_equal(got["keys_lo"], keys)
    assert got["m"] == d["m"] and _digest(got["keys_lo"], np.uint64) == d["keys_lo"]
    assert _digest(got["prefix_sums"], np.uint32) == d["prefix_sums"]
    assert int(got["left"][0]) == d["root_left"] != s_star
    assert _digest(got["left"], np.int32) == d["left"]
    assert _digest(got["parent"], np.int32) == d["parent"]
    assert _digest(got["lcp_length"], np.int32) == d["lcp_length"]
    assert _digest(got["lcp_lo"], np.uint64) == d["lcp_lo"]
    assert got["octree_size"] == d["octree_size"]
    assert _digest(got["children"], np.int32) == d["children"]
    assert _digest(got["leaf"], np.int32) == d["leaf"]


def test_labels_conflict_cells(dev):
    """Point-label conflict pass: flags per unique key and per (node, octant)
    slot against the oracle (itself pinned by brute force)."""
    pts, labels = workloads.star_loops(1 << 16, loops=16, seed=9)
    pts[:2000] = pts[5000:7000]                     # overlapping samples of two loops
    for bits in (21, 42):
        got = dev.build_octree(pts, bits, labels=labels)
        want = oracle.build(oracle.quantize(pts, 1 << bits), bits, labels=labels)
        assert_tree_equal(got, want, where=f"labels bits={bits}")
        assert np.array_equal(got["key_conflict"], want["key_conflict"])
        assert np.array_equal(got["node_conflict"], want["node_conflict"])
        assert got["key_conflict"].sum() > 0


def test_conflict_cells_full_size(dev):
    """The conflict-cell workload at full size: 2^22 points on 64 star-shaped
    loops, 42 bits/axis (126-bit keys), loop ids as labels -- every output
    against the oracle.  At this resolution no two loops share a cell, so a
    second run copies 50,000 samples of some loops onto others (the same
    coordinates under another label) to exercise the flags at full size."""
    pts, labels = workloads.star_loops(1 << 22, loops=64, seed=4)
    rng = np.random.default_rng(11)
    shared = pts.copy()
    src = rng.choice(pts.shape[0], 50_000, replace=False)
    dst = rng.choice(pts.shape[0], 50_000, replace=False)
    shared[dst] = pts[src]
    for name, p in (("loops", pts), ("shared samples", shared)):
        got = dev.build_octree(p, 42, labels=labels)
        want = oracle.build(oracle.quantize(p, 1 << 42), 42, labels=labels)
        assert_tree_equal(got, want, where=name)
        assert np.array_equal(got["key_conflict"], want["key_conflict"]), name
        assert np.array_equal(got["node_conflict"], want["node_conflict"]), name
    assert want["key_conflict"].sum() > 1000


# ------------------------------------------------------------ stage APIs ----
def test_points_to_morton_reference_input(dev):
    """PointsToMorton on (i, i), bits = 30 (tests/KernelTests.cpp:57-84),
    including the zero-key padding to nextPow2."""
    import json
    import os
    from conftest import GOLDEN
    with open(os.path.join(GOLDEN, "morton_ref.json")) as f:
        d = json.load(f)
    q = np.stack([np.arange(1_000_000)] * 2, axis=1).astype(np.int32)
    lo, hi = dev.points_to_morton(q, 30)
    assert lo.size == oracle.next_pow2(1_000_000) == 1 << 20
    assert _digest(lo[:1_000_000], np.uint64) == d["keys_lo_sha256"]
    assert not lo[1_000_000:].any() and not hi.any()


@pytest.mark.parametrize("dim,bits", [(2, 16), (3, 21), (3, 42), (2, 64)])
def test_points_to_morton_matches_oracle(dev, dim, bits):
    rng = np.random.default_rng(bits)
    q = rng.integers(0, 2**62, (3001, dim), dtype=np.int64)
    lo, hi = dev.points_to_morton(q, bits)
    olo, ohi = oracle.morton(q, bits)
    assert np.array_equal(lo[:3001], olo)
    if dim * bits > 64:
        assert np.array_equal(hi[:3001], ohi)
    assert not lo[3001:].any()


@pytest.mark.parametrize("kb,mbits,n", [(4, 32, 100003), (8, 63, 300001), (8, 40, 5),
                                        (16, 126, 50001), (16, 90, 4096)])
def test_radix_sort(dev, kb, mbits, n):
    rng = np.random.default_rng(n)
    lo = rng.integers(0, 2**63, n, dtype=np.uint64)
    hi = rng.integers(0, 2**62, n, dtype=np.uint64)
    if mbits < 64:
        lo &= np.uint64((1 << mbits) - 1)
    if kb < 16:
        hi[:] = 0
    else:
        hi &= np.uint64((1 << (mbits - 64)) - 1)
    lo[: n // 3] = lo[n // 3: 2 * (n // 3)]           # duplicates: stability is observable
    hi[: n // 3] = hi[n // 3: 2 * (n // 3)]
    vals = np.arange(n, dtype=np.uint32)
    (slo, shi), svals = dev.radix_sort(lo, hi, mbits=mbits, vals=vals, key_bytes=kb)
    order = np.lexsort((vals, lo, hi))              # stable: ties keep input order
    assert np.array_equal(slo, lo[order])
    if kb == 16:
        assert np.array_equal(shi, hi[order])
    assert np.array_equal(svals, vals[order])


@pytest.mark.parametrize("kb", [4, 8, 16])
def test_unique_sorted(dev, kb):
    rng = np.random.default_rng(kb)
    lo = np.sort(rng.integers(0, 5000, 200000, dtype=np.uint64))
    hi = np.zeros_like(lo) if kb < 16 else (lo // np.uint64(7))
    ulo, uhi = dev.unique_sorted(lo, hi, key_bytes=kb)
    want = np.unique(lo)
    assert np.array_equal(ulo, want)


@pytest.mark.parametrize("n", [1, 2, 4095, 4096, 4097, 1 << 20, 3000017])
def test_stream_scan(dev, n):
    rng = np.random.default_rng(n)
    x = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)   # wraps mod 2^32
    assert np.array_equal(dev.stream_scan(x), oracle.stream_scan(x))


@pytest.mark.parametrize("mbits,dim", [(32, 2), (63, 3), (126, 3), (40, 2)])
def test_brt_and_octree_stages(dev, mbits, dim):
    rng = np.random.default_rng(mbits)
    m = 40000
    if mbits <= 64:
        lo = np.unique(rng.integers(0, 1 << min(mbits, 63), m, dtype=np.uint64))
        hi = None
    else:
        lo = rng.integers(0, 2**63, m, dtype=np.uint64)
        hi = rng.integers(0, 1 << (mbits - 64), m, dtype=np.uint64)
        order = np.lexsort((lo, hi))
        lo, hi = lo[order], hi[order]
        keep = np.ones(m, bool)
        keep[1:] = (lo[1:] != lo[:-1]) | (hi[1:] != hi[:-1])
        lo, hi = lo[keep], hi[keep]
    kb = P.key_bytes_for(mbits)
    want = oracle.build_from_keys(lo, hi, dim, mbits)
    brt, parent = dev.build_brt(lo, hi, mbits, key_bytes=kb)
    assert np.array_equal(brt["left"], want["left"])
    assert np.array_equal(brt["left_leaf"], want["left_leaf"])
    assert np.array_equal(brt["right_leaf"], want["right_leaf"])
    assert np.array_equal(brt["lcp_length"].astype(np.int32), want["lcp_length"])
    assert np.array_equal(parent, want["parent"])
    if kb == 16:
        assert np.array_equal(brt["lcp_lo"], want["lcp_lo"]) and np.array_equal(brt["lcp_hi"], want["lcp_hi"])
    else:
        assert np.array_equal(brt["lcp"].astype(np.uint64), want["lcp_lo"])
    ls = dev.local_splits(brt, lo.size, dim, kb)
    assert np.array_equal(ls, want["local_splits"])
    ch, lf = dev.brt_to_octree(brt, parent, lo.size, dim, kb)
    assert np.array_equal(ch, want["children"]) and np.array_equal(lf, want["leaf"])


@pytest.mark.parametrize("kb,mbits", [(4, 30), (8, 63), (16, 87), (16, 126)])
def test_brt_of_duplicate_keys_matches_oracle(dev, kb, mbits):
    """BuildBinaryRadixTree_p on sorted keys with runs of equal neighbours
    (pairs, triples, runs at both ends) returns the nodes and parent links
    of the reference's node loop (the oracle's, which equals the reference's
    own BuildBinaryRadixTree on such keys up to 88 bits)."""
    rng = np.random.default_rng(kb + mbits)
    for trial in range(40):
        m = int(rng.integers(2, 3000))
        lo = rng.integers(0, 2**63, m, dtype=np.uint64)
        hi = rng.integers(0, 2**63, m, dtype=np.uint64) if mbits > 64 else np.zeros(m, np.uint64)
        if mbits < 64:
            lo &= np.uint64((1 << mbits) - 1)
        if mbits > 64:
            hi &= np.uint64((1 << (mbits - 64)) - 1)
        for _ in range(m // 3):
            i = int(rng.integers(0, m))
            lo[(i + 1) % m], hi[(i + 1) % m] = lo[i], hi[i]
        order = np.lexsort((lo, hi))
        lo, hi = lo[order], hi[order]
        if trial % 4 == 0:
            lo[-3:], hi[-3:] = lo[-1], hi[-1]       # three equal keys at the end
        brt, parent = dev.build_brt(lo, hi if mbits > 64 else None, mbits, key_bytes=kb)
        want = oracle.brt_only(lo, hi, mbits)
        assert np.array_equal(brt["left"], want["left"]), trial
        assert np.array_equal(brt["left_leaf"], want["left_leaf"]), trial
        assert np.array_equal(brt["right_leaf"], want["right_leaf"]), trial
        assert np.array_equal(brt["lcp_length"], want["lcp_length"]), trial
        assert np.array_equal(parent, want["parent"]), trial
        if kb == 16:
            assert np.array_equal(brt["lcp_lo"], want["lcp_lo"]), trial
            assert np.array_equal(brt["lcp_hi"], want["lcp_hi"]), trial
        else:
            assert np.array_equal(brt["lcp"].astype(np.uint64), want["lcp_lo"]), trial


def test_octree_capacity_is_reported(dev):
    keys = np.unique(np.random.default_rng(1).integers(0, 1 << 40, 5000, dtype=np.uint64))
    brt, parent = dev.build_brt(keys, None, 40, key_bytes=8)
    with pytest.raises(P.PgvdError) as e:
        dev.brt_to_octree(brt, parent, keys.size, 2, 8, capacity=10)
    assert e.value.code == 7


# ------------------------------------------------------------ find leaf ----
@pytest.mark.parametrize("dim,bits,n", [(2, 16, 65536), (3, 21, 1 << 20), (3, 10, 200000),
                                        (2, 30, 50000), (2, 16, 1 << 21), (3, 21, 1 << 22),
                                        (3, 23, 1 << 22), (3, 31, 1 << 22), (2, 31, 1 << 21),
                                        (3, 5, 100000), (2, 2, 40)])
def test_find_leaf_matches_oracle(dev, dim, bits, n):
    """OctreeUtils::FindLeaf batched on the GPU against the oracle's
    restatement (pinned by brute force), on the octree the GPU built: the
    build's own points and random grid points, every output field.  Batches
    of at least twice the cell count at the table depth (2^21 cells in 3D,
    2^20 in 2D, fewer when bits is small) start from the depth table: here
    3D 21 bits at 2^22 points, 2D 16 bits at 2^21 points, and 3D 5 bits and
    2D 2 bits, where the table covers every level; 3D 23 bits (69-bit keys:
    the table index straddles the two key words) and 31 bits (93-bit keys:
    the index in the high word)."""
    rng = np.random.default_rng(dim * 31 + bits)
    pts = _pts(n, dim, seed=dim * 7 + bits)
    tree = dev.build_octree(pts, bits)
    q = oracle.quantize(pts, 1 << bits)
    queries = np.concatenate([q, rng.integers(0, 1 << bits, (n // 4, dim))]).astype(np.int32)
    got = dev.find_leaf(queries, bits, tree["children"], tree["leaf"])
    want = oracle.find_leaf(queries, bits, tree["children"], tree["leaf"])
    for k in ("parent", "octant", "data", "level", "status"):
        assert np.array_equal(got[k].astype(np.int64), want[k].astype(np.int64)), k
    assert (want["status"] == 0).all() and (want["level"] > 1).any()


@pytest.mark.parametrize("reps", [1, 4])
def test_find_leaf_rejects_foreign_child_index(dev, reps):
    """A child index outside the octree stops the descent with status 2 (the
    oracle agrees) instead of reading past the nodes; with 12 queries the
    descent starts from the depth table, which records the stop."""
    children = np.array([[3, -1, -1, -1], [-1, -1, -1, -1]], np.int32)   # node 0 -> node 3 of 2
    leaf = np.array([0b1110, 0b1111], np.int32)
    queries = np.tile(np.array([[0, 0], [1, 0], [1, 1]], np.int32), (reps, 1))
    got = dev.find_leaf(queries, 1, children, leaf)
    want = oracle.find_leaf(queries, 1, children, leaf)
    for k in want:
        assert np.array_equal(got[k].astype(np.int64), want[k].astype(np.int64)), k
    assert list(want["status"]) == [2, 0, 0] * reps


# ----------------------------------------------------- reference data ----
def _data_points():
    path = os.path.join(os.path.dirname(__file__), "golden", "data_points.npz")
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", ["data2_test1", "data2_test2", "data2_test3", "data2_test4",
                                  "data2_test5", "data2_test6", "data2_test7", "data2_test8",
                                  "data3_simple1", "data3_simple2", "data3_simple3", "data3_bun1"])
def test_reference_data_builds(dev, name):
    """The reference's own point sets (data2 polyline scenes, data3 meshes,
    read by pgvd_amd.io into the committed fixture), built at the viewer's
    resolution and at the widest key of each dimension."""
    pts = _data_points()[name]
    dim = pts.shape[1]
    for bits in ((16, 30) if dim == 2 else (21, 42)):
        got = dev.build_octree(pts, bits)
        want = oracle.build(oracle.quantize(pts, 1 << bits), bits)
        assert_tree_equal(got, want, where=f"{name} bits={bits}")


# ------------------------------------------------------------ cell walk ----
def _assert_walk_equal(got, want, where):
    assert np.array_equal(got["status"], want["status"]), where
    assert np.array_equal(got["offsets"], want["offsets"]), where
    assert np.array_equal(got["cells"], want["cells"]), where


def _obj2oct(poly, lo, bbw, width):
    """viewer/Octree2.cpp:216-226 in float32: ((v - min) / bbw) * width,
    clamped below width."""
    oct_ = ((poly - lo) / np.float32(bbw)) * np.float32(width)
    return np.where(oct_ >= width, np.float32(width - 0.0001), oct_).astype(np.float32)


@pytest.mark.parametrize("bits,npts,nseg,reach", [(10, 400, 3000, 1.0), (16, 65536, 20000, 0.05),
                                                  (20, 200000, 5000, 0.01), (3, 20, 500, 1.0),
                                                  (30, 100000, 3000, 0.01)])
def test_cell_walk_matches_oracle(dev, bits, npts, nseg, reach):
    """CellWalk on the GPU against the oracle's restatement (pinned by brute
    force) on a GPU-built quadtree: random segments, axis-parallel and
    degenerate ones, and segments along cell edges (integer coordinates),
    which take the reference's corner and decrement paths."""
    rng = np.random.default_rng(bits)
    tree = dev.build_octree(_pts(npts, 2, seed=bits), bits)
    W = 1 << bits
    a = rng.random((nseg, 2)) * (W - 1e-3)
    b = np.clip(a + (rng.random((nseg, 2)) * 2 - 1) * reach * W, 0, W - 1e-3)
    segs = np.concatenate([a, b], axis=1).astype(np.float32)
    k = nseg // 8
    segs[:k, 3] = segs[:k, 1]                            # horizontal
    segs[k:2 * k, 2] = segs[k:2 * k, 0]                  # vertical
    segs[2 * k:3 * k, 2:] = segs[2 * k:3 * k, :2]        # a == b
    segs[3 * k:4 * k] = np.floor(segs[3 * k:4 * k])      # on grid lines
    got = dev.cell_walk(segs, bits, tree["children"], tree["leaf"])
    want = oracle.cell_walk(segs, bits, tree["children"], tree["leaf"])
    _assert_walk_equal(got, want, f"bits={bits}")
    assert (np.diff(want["offsets"]) > 1).mean() > 0.3


def test_cell_walk_count_then_write(dev, monkeypatch):
    """With no scratch allowed for the first walk's early cells
    (PGVD_WALK_SCRATCH_BYTES=0), every segment is walked twice -- a count
    walk, then a write walk into the offsets the scan gave: the same cells
    as the oracle, and as the default path."""
    bits = 12
    rng = np.random.default_rng(31)
    tree = dev.build_octree(_pts(20000, 2, seed=31), bits)
    W = 1 << bits
    a = rng.random((4000, 2)) * (W - 1e-3)
    b = np.clip(a + (rng.random((4000, 2)) * 2 - 1) * 0.2 * W, 0, W - 1e-3)
    segs = np.concatenate([a, b], axis=1).astype(np.float32)
    default = dev.cell_walk(segs, bits, tree["children"], tree["leaf"])
    monkeypatch.setenv("PGVD_WALK_SCRATCH_BYTES", "0")
    got = dev.cell_walk(segs, bits, tree["children"], tree["leaf"])
    want = oracle.cell_walk(segs, bits, tree["children"], tree["leaf"])
    _assert_walk_equal(got, want, "count then write")
    _assert_walk_equal(got, default, "count then write vs default")
    assert (np.diff(want["offsets"]) > 32).any()


def test_cell_walk_reference_polylines(dev):
    """The data2 scenes walked as the viewer walks them (FindMultiCells,
    viewer/Octree2.cpp:430-455): each polyline's (file's) consecutive
    vertices, polyline by polyline (:442-450), mapped by obj2Oct into the
    octree of the scene's points."""
    pts = _data_points()
    for name in ("data2_test1", "data2_test2", "data2_test3", "data2_test4", "data2_test5"):
        v = pts[name]
        ends = np.cumsum(pts[name + "_counts"])
        for bits in (8, 16):
            tree = dev.build_octree(v, bits)
            lo = v.min(axis=0)
            bbw = float((v.max(axis=0) - lo).max())
            q = _obj2oct(v, lo, bbw, 1 << bits)
            segs = np.concatenate([np.concatenate([p[:-1], p[1:]], axis=1)
                                   for p in np.split(q, ends[:-1])])
            assert segs.shape[0] == v.shape[0] - ends.size
            got = dev.cell_walk(segs, bits, tree["children"], tree["leaf"])
            want = oracle.cell_walk(segs, bits, tree["children"], tree["leaf"])
            _assert_walk_equal(got, want, f"{name} bits={bits}")


def test_cell_walk_stops_after_10000_cells(dev):
    """A walk along a diagonal of unit cells passes 10000 cells: status 1 and
    the first 10000 cells, as the oracle (the reference throws there)."""
    bits = 16
    t = np.arange(0, 60000, dtype=np.float32)
    diag = np.stack([t, t], axis=1) / np.float32(1 << bits) * 2 - 1
    tree = dev.build_octree(diag.astype(np.float32), bits)
    segs = np.array([[0.5, 0.7, 59000.5, 59000.2], [10.5, 10.2, 20.5, 20.7]], np.float32)
    got = dev.cell_walk(segs, bits, tree["children"], tree["leaf"])
    want = oracle.cell_walk(segs, bits, tree["children"], tree["leaf"])
    _assert_walk_equal(got, want, "diagonal")
    assert list(want["status"]) == [1, 0] and want["offsets"][1] == 10000


def test_query_entry_points_reject_bad_arguments(dev):
    """Argument checks of the octree consumers: each bad call returns
    PGVD_ERR_INVALID (1) or PGVD_ERR_CAPACITY (7) with the size it needs,
    and launches nothing."""
    L = P.lib()
    tree = dev.build_octree(_pts(1000, 2, seed=5), 10)
    rec = np.concatenate([tree["children"], tree["leaf"][:, None]], axis=1).astype(np.int32)
    d = dev.upload(rec)
    q = dev.upload(np.zeros((4, 2), np.int32))
    outs = [dev.alloc(16) for _ in range(5)]
    O = rec.shape[0]
    ptrs = [o.ptr for o in outs]
    assert L.pgvd_find_leaf(dev.ctx, q.ptr, 4, 4, 10, d.ptr, O, *ptrs, None) == 1      # dim
    assert L.pgvd_find_leaf(dev.ctx, q.ptr, 4, 2, 0, d.ptr, O, *ptrs, None) == 1       # bits
    assert L.pgvd_find_leaf(dev.ctx, q.ptr, 4, 2, 10, d.ptr, 0, *ptrs, None) == 1      # empty octree
    assert L.pgvd_find_leaf(dev.ctx, q.ptr, 4, 2, 10, None, O, *ptrs, None) == 1
    segs = dev.upload(np.array([[1, 1, 900, 700]], np.float32))
    off = dev.alloc(16)
    tot = C.c_int64()
    assert L.pgvd_cell_walk(dev.ctx, segs.ptr, 1, 31, d.ptr, O, off.ptr, None, 0, C.byref(tot),
                            None, None) == 1                                         # bits
    assert L.pgvd_cell_walk(dev.ctx, segs.ptr, 1, 10, d.ptr, O, off.ptr, None, 0, None, None,
                            None) == 1                                               # total_out
    assert L.pgvd_cell_walk(dev.ctx, segs.ptr, 1, 10, d.ptr, O, off.ptr, None, 0, C.byref(tot),
                            None, None) == 0 and tot.value > 1                       # count only
    cells = dev.alloc(20)
    assert L.pgvd_cell_walk(dev.ctx, segs.ptr, 1, 10, d.ptr, O, off.ptr, cells.ptr, 1,
                            C.byref(tot), None, None) == 7 and tot.value > 1         # capacity
    ln = C.c_int64()
    small = dev.alloc(8)
    assert L.pgvd_octree_format_text(dev.ctx, d.ptr, O, 2, small.ptr, 8, C.byref(ln), None) == 7
    assert ln.value == len(oracle.octree_format_text(rec, 2))
    assert L.pgvd_octree_format_text(dev.ctx, d.ptr, O, 5, None, 0, C.byref(ln), None) == 1
    assert L.pgvd_octree_save(dev.ctx, d.ptr, O, 2, b"/nonexistent-dir/x.txt", 0) == 1
    assert L.pgvd_octree_save(dev.ctx, d.ptr, O, 2, b"/tmp/x.txt", 3) == 1               # format
    for b in [d, q, segs, off, cells, small, *outs]:
        b.free()


@pytest.mark.parametrize("cap", [1, 3, 16])
def test_gather_queue_overflow_matches_oracle(dev, cap):
    """The octree gather's per-wave queues shrunk to `cap` entries
    (pgvd_debug_gather_queue): the owners whose breadth-first work no longer
    fits go to the global-memory pass, and every output still equals the
    oracle's -- uniform and clustered 3D points and a 2D quadtree."""
    cases = [(_pts(1 << 18, 3, 31), 21), (_pts(1 << 17, 2, 32), 16)]
    g, _ = workloads.gaussian_mixture(1 << 17, seed=33)
    cases.append((g, 21))
    P.debug_gather_queue(cap)
    try:
        for pts, bits in cases:
            print(f"queue cap={cap} n={pts.shape[0]} dim={pts.shape[1]}", flush=True)
            got = dev.build_octree(pts, bits)
            want = oracle.build(oracle.quantize(pts, 1 << bits), bits)
            assert_tree_equal(got, want, where=f"cap={cap} n={pts.shape[0]} dim={pts.shape[1]}")
    finally:
        P.debug_gather_queue(0)
