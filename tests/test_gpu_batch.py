"""Batched form (pgvd_batch_*): every instance of a segmented batch equals the
plain whole-path build of that instance alone, which the oracle and the
reference digests pin (test_gpu_parity.py); a sample of instances is also
checked against the oracle directly."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import assert_tree_equal

import oracle
import pgvd_amd
from pgvd_amd import workloads

BATCH_FIELDS = ["keys_lo", "keys_hi", "left", "left_leaf", "right_leaf", "lcp_lo", "lcp_hi",
                "lcp_length", "parent", "prefix_sums", "children", "leaf"]

# Edge sizes: single points, the nextPow2 padding boundary (8, 2^k, 2^k+1),
# one sort tile and one past it, and an instance spanning two encode chunks.
EDGE_SIZES = [1, 2, 3, 7, 8, 9, 16, 17, 255, 256, 257, 1000, 4095, 4096, 4097, 65537, 70000]


def make_batch(sizes, dim, seed, clustered_every=2):
    rng = np.random.default_rng(seed)
    parts = []
    for i, n in enumerate(sizes):
        if i % clustered_every == 1:
            c = rng.uniform(-0.5, 0.5, (3, dim))
            pick = rng.integers(0, 3, n)
            p = c[pick] + rng.normal(0, 1e-4, (n, dim))
            parts.append(np.clip(p, -1, 1).astype(np.float32))
        else:
            parts.append((rng.random((n, dim), dtype=np.float32) * 2 - 1))
    return parts


def run_batch(dev, parts, dim, bits, **kw):
    sizes = [len(p) for p in parts]
    pts = np.concatenate(parts).astype(parts[0].dtype)
    d = dev.upload(pts)
    b = pgvd_amd.Batch(dev, sizes, dim, bits, quantized=pts.dtype == np.int32, **kw)
    try:
        b.run(d)
        r = b.finish()
        return b, r, b.fetch_all(r), d
    except Exception:
        b.close()
        d.free()
        raise


@pytest.mark.gpu
@pytest.mark.parametrize("dim,bits", [(2, 16), (3, 21), (3, 10), (2, 31)])
def test_batch_matches_plain_per_instance(dev, dim, bits):
    parts = make_batch(EDGE_SIZES, dim, seed=11 + dim)
    b, r, all_, d = run_batch(dev, parts, dim, bits)
    try:
        kb = r.key_bytes
        for i, p in enumerate(parts):
            got = pgvd_amd.Batch.instance(all_, i, kb, dim)
            want = dev.build_octree(p, bits)
            assert_tree_equal(got, want, BATCH_FIELDS, where=f"instance {i} (n={len(p)})")
    finally:
        b.close()
        d.free()


@pytest.mark.gpu
def test_batch_matches_oracle_sample(dev):
    dim, bits = 3, 21
    sizes = [1, 5, 64, 3000, 20000]
    parts = make_batch(sizes, dim, seed=5)
    b, r, all_, d = run_batch(dev, parts, dim, bits)
    try:
        for i, p in enumerate(parts):
            got = pgvd_amd.Batch.instance(all_, i, r.key_bytes, dim)
            want = oracle.build(oracle.quantize(p, 1 << bits), bits)
            assert_tree_equal(got, want, BATCH_FIELDS, where=f"instance {i}")
    finally:
        b.close()
        d.free()


@pytest.mark.gpu
def test_batch_wide_keys_and_int_input(dev):
    # 126-bit keys (42 bits/axis) from int32 grid points is out of int32
    # range, so wide keys use float input; int32 input with 63-bit keys.
    parts = make_batch([1, 9, 300, 5000], 3, seed=3)
    b, r, all_, d = run_batch(dev, parts, 3, 42)
    try:
        assert r.key_bytes == 16
        for i, p in enumerate(parts):
            got = pgvd_amd.Batch.instance(all_, i, 16, 3)
            assert_tree_equal(got, dev.build_octree(p, 42), BATCH_FIELDS, where=f"wide {i}")
    finally:
        b.close()
        d.free()
    rng = np.random.default_rng(4)
    q = [rng.integers(0, 1 << 21, (n, 3)).astype(np.int32) for n in (1, 8, 1234, 40000)]
    q[1][:] = q[1][0]                          # all-duplicate instance
    b, r, all_, d = run_batch(dev, q, 3, 21)
    try:
        for i, p in enumerate(q):
            got = pgvd_amd.Batch.instance(all_, i, r.key_bytes, 3)
            assert_tree_equal(got, dev.build_octree(p, 21), BATCH_FIELDS, where=f"int {i}")
    finally:
        b.close()
        d.free()


@pytest.mark.gpu
def test_batch_graph_replay_with_new_points(dev):
    """The captured graph is replayed: new contents at the same device
    address give the new results, every replay matching a graph-free run."""
    dim, bits = 3, 21
    sizes = [100, 5000, 33333, 7]
    sets = [make_batch(sizes, dim, seed=s) for s in (21, 22, 23)]
    pts0 = np.concatenate(sets[0])
    d = dev.upload(pts0)
    ref = pgvd_amd.Batch(dev, sizes, dim, bits, use_graph=False)
    b = pgvd_amd.Batch(dev, sizes, dim, bits, use_graph=True)
    try:
        for rep, parts in enumerate(sets + sets[:1]):
            pts = np.ascontiguousarray(np.concatenate(parts))
            dev._check(pgvd_amd.lib().pgvd_memcpy_htod(dev.ctx, d.ptr, pts.ctypes.data,
                                                         pts.nbytes, None))
            b.run(d)
            got = b.fetch_all(b.finish())
            ref.run(d)
            want = ref.fetch_all(ref.finish())
            for i in range(len(sizes)):
                g = pgvd_amd.Batch.instance(got, i, 8, dim)
                w = pgvd_amd.Batch.instance(want, i, 8, dim)
                assert_tree_equal(g, w, BATCH_FIELDS, where=f"replay {rep} instance {i}")
    finally:
        b.close()
        ref.close()
        d.free()


@pytest.mark.gpu
def test_batch_tuner_keeps_results(dev):
    dim, bits = 3, 21
    parts = make_batch([3000, 70000, 12, 9999], dim, seed=8)
    pts = np.concatenate(parts)
    d = dev.upload(pts)
    b = pgvd_amd.Batch(dev, [len(p) for p in parts], dim, bits)
    try:
        ipt, rows, dbits, waves = b.tune(d)
        assert ipt in (8, 12, 15, 16) and rows in (2, 4)
        assert dbits in (6, 7, 8) and waves in (2, 3, 4)
        b.run(d)
        all_ = b.fetch_all(b.finish())
        for i, p in enumerate(parts):
            got = pgvd_amd.Batch.instance(all_, i, 8, dim)
            assert_tree_equal(got, dev.build_octree(p, bits), BATCH_FIELDS, where=f"tuned {i}")
    finally:
        b.close()
        d.free()


@pytest.mark.gpu
@pytest.mark.parametrize("ipt,rows", [(8, 2), (12, 4), (15, 4), (16, 2)])
def test_batch_every_config(dev, ipt, rows):
    dim, bits = 3, 21
    parts = make_batch([1, 4097, 50000, 2], dim, seed=ipt * 10 + rows)
    b, r, all_, d = run_batch(dev, parts, dim, bits, sort_ipt=ipt, brt_rows=rows)
    try:
        for i, p in enumerate(parts):
            got = pgvd_amd.Batch.instance(all_, i, 8, dim)
            assert_tree_equal(got, dev.build_octree(p, bits), BATCH_FIELDS, where=f"cfg {i}")
    finally:
        b.close()
        d.free()


@pytest.mark.gpu
@pytest.mark.parametrize("dim,bits", [(3, 21), (2, 16), (3, 42)])
@pytest.mark.parametrize("dbits", [6, 7, 8])
@pytest.mark.parametrize("waves", [2, 3, 4])
def test_batch_every_digit_width_and_wave_count(dev, dim, bits, dbits, waves):
    """Every radix digit width and sort-kernel wave count the tuner can pick
    gives every instance's plain build, field by field (128-bit keys take
    8-bit digits only)."""
    if bits == 42 and dbits != 8:
        with pytest.raises(pgvd_amd.PgvdError):
            pgvd_amd.Batch(dev, [10], dim, bits, digit_bits=dbits)
        return
    parts = make_batch([1, 3000, 70000, 2, 9001], dim, seed=dbits * 10 + waves + bits)
    b, r, all_, d = run_batch(dev, parts, dim, bits, digit_bits=dbits, sort_waves=waves)
    kb = 16 if dim * bits > 64 else (8 if dim * bits > 32 else 4)
    try:
        for i, p in enumerate(parts):
            got = pgvd_amd.Batch.instance(all_, i, kb, dim)
            assert_tree_equal(got, dev.build_octree(p, bits), BATCH_FIELDS,
                              where=f"digits {dbits} waves {waves} instance {i}")
    finally:
        b.close()
        d.free()


@pytest.mark.gpu
def test_batch_c4_shaped_sample(dev):
    """A 64-instance sample of the 4096-instance workload's size law
    (log-uniform 1e3..1e6, half uniform, half clustered)."""
    sizes = workloads.batch_sizes(64, seed=5)
    parts = [workloads.batch_instance(i, int(n), seed=5) for i, n in enumerate(sizes)]
    b, r, all_, d = run_batch(dev, parts, 3, 21)
    try:
        for i in range(0, 64, 7):
            got = pgvd_amd.Batch.instance(all_, i, 8, 3)
            assert_tree_equal(got, dev.build_octree(parts[i], 21), BATCH_FIELDS, where=f"c4 {i}")
        # total octree size is the sum of the instances'
        assert r.total_octree == int(np.sum(all_["octree_size"]))
    finally:
        b.close()
        d.free()


@pytest.mark.gpu
def test_batch_rejects_bad_sizes(dev):
    with pytest.raises(pgvd_amd.PgvdError):
        pgvd_amd.Batch(dev, [10, 0, 5], 3, 21)
    with pytest.raises(pgvd_amd.PgvdError):
        pgvd_amd.Batch(dev, [(1 << 23) + 1], 3, 21)
    with pytest.raises(pgvd_amd.PgvdError):
        pgvd_amd.Batch(dev, [10], 3, 21, sort_ipt=7)
    with pytest.raises(pgvd_amd.PgvdError):
        pgvd_amd.Batch(dev, [10], 3, 21, digit_bits=5)
    with pytest.raises(pgvd_amd.PgvdError):
        pgvd_amd.Batch(dev, [10], 3, 21, sort_waves=5)


@pytest.mark.gpu
def test_digest_matches_definition(dev):
    rng = np.random.default_rng(0)
    w = rng.integers(0, 1 << 32, 100003, dtype=np.uint64).astype(np.uint32)
    d = dev.upload(w)
    try:
        got = dev.digest(d, w.nbytes)
    finally:
        d.free()
    z = w.astype(np.uint64) | (np.arange(w.size, dtype=np.uint64) << np.uint64(32))
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    z = z ^ (z >> np.uint64(31))
    assert got == int(z.sum(dtype=np.uint64))
