"""The batched and served workloads at their configured sizes.

* Batched: 4096 independent 3D point sets, sizes log-uniform in [1e3, 1e6]
  (seed 5; even instances uniform, odd ones 8 Gaussian clusters), 21
  bits/axis, through one segmented pipeline replayed as a HIP graph.  Every
  instance's keys, BRT records, parent links, prefix sums and octree equal
  the plain whole-path build of that instance alone (device digests), and
  16 instances chosen across the size range, the largest among them, equal
  the oracle field by field.
* Served: 520 requests arriving as a Poisson stream at 2000 per second,
  each 2D (16 bits/axis) or 3D (21 bits/axis) with equal probability, sizes
  log-uniform in [1e4, 4e6] (seed 6), under a 64 GB budget.  Every request's
  result equals the same build run alone, and 8 requests across the size
  range, the largest among them, equal the oracle's octree.
"""
from __future__ import annotations

import time

import numpy as np
import pytest

from conftest import assert_tree_equal

import oracle
import pgvd_amd
from pgvd_amd import BRT_DTYPES, DeviceBuffer, workloads
from test_gpu_serve import digest_np

BATCH_FIELDS = ["keys_lo", "keys_hi", "left", "left_leaf", "right_leaf", "lcp_lo", "lcp_hi",
                "lcp_length", "parent", "prefix_sums", "children", "leaf"]


def log(msg: str):
    print(msg, flush=True)


def spread_pick(sizes: np.ndarray, k: int) -> list[int]:
    """k indices at evenly spaced ranks of the sizes, the largest included."""
    order = np.argsort(sizes, kind="stable")
    ranks = np.unique(np.round(np.linspace(0, sizes.size - 1, k)).astype(np.int64))
    return [int(order[r]) for r in ranks]


def batch_instance_device(dev, r, cnt: dict, i: int, dim: int) -> dict:
    """Instance i of a batch result, downloaded range by range."""
    kb, nc = r.key_bytes, 1 << dim
    k0, m = int(cnt["key_offset"][i]), int(cnt["m"][i])
    o0, O = int(cnt["octree_offset"][i]), int(cnt["octree_size"][i])
    rec = BRT_DTYPES[kb]
    words = dev.download(r.keys + k0 * kb, np.uint64 if kb >= 8 else np.uint32,
                         m * (2 if kb == 16 else 1))
    lo, hi = pgvd_amd.words_to_keys(words, kb)
    ni = max(m - 1, 0)
    brt = dev.download(r.brt + k0 * rec.itemsize, np.uint8, ni * rec.itemsize).view(rec)
    raw = dev.download(r.octree + o0 * (nc + 1) * 4, np.int32, O * (nc + 1)).reshape(O, nc + 1)
    out = dict(m=m, octree_size=O, keys_lo=lo, keys_hi=hi,
               parent=dev.download(r.parent + k0 * 4, np.int32, ni),
               prefix_sums=dev.download(r.prefix_sums + k0 * 4, np.uint32, m),
               children=raw[:, :nc].copy(), leaf=raw[:, nc].copy(),
               left=brt["left"].copy(), left_leaf=brt["left_leaf"].copy(),
               right_leaf=brt["right_leaf"].copy(), lcp_length=brt["lcp_length"].astype(np.int32))
    if kb == 16:
        out["lcp_lo"], out["lcp_hi"] = brt["lcp_lo"].copy(), brt["lcp_hi"].copy()
    else:
        out["lcp_lo"], out["lcp_hi"] = brt["lcp"].astype(np.uint64), np.zeros(ni, np.uint64)
    return out


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_batch_c4_full_size(dev):
    dim, bits, count = 3, 21, 4096
    t0 = time.perf_counter()
    sizes = workloads.batch_sizes(count, seed=5)
    pts = workloads.batch_points(sizes, seed=5)
    offs = np.concatenate([[0], np.cumsum(sizes)])
    log(f"c4: {count} instances, {int(sizes.sum())} points generated in "
        f"{time.perf_counter() - t0:.1f} s")
    d = dev.upload(pts)
    b = pgvd_amd.Batch(dev, sizes, dim, bits, use_graph=True)
    try:
        b.run(d)             # captures the graph and launches it
        b.finish()
        b.run(d)             # replays it
        r = b.finish()
        cnt = pgvd_amd.Batch.counts(r)
        log(f"c4: batch done, {r.total_octree} octree nodes, {b.last_ms():.1f} ms")
        kb, rec = r.key_bytes, BRT_DTYPES[r.key_bytes].itemsize
        stride = ((1 << dim) + 1) * 4
        bad = []
        for i in range(count):
            n = int(sizes[i])
            view = DeviceBuffer(dev, d.ptr + int(offs[i]) * dim * 4, n * dim * 4)
            dev.build_octree_async(view, n, dim, bits)
            rr = dev.build_octree_finish()
            m, O = int(rr.m), int(rr.octree_size)
            if (m, O) != (int(cnt["m"][i]), int(cnt["octree_size"][i])):
                bad.append((i, "sizes"))
                continue
            k0, o0 = int(cnt["key_offset"][i]), int(cnt["octree_offset"][i])
            for name, a, bb, nbytes in (
                    ("keys", rr.keys, r.keys + k0 * kb, m * kb),
                    ("brt", rr.brt, r.brt + k0 * rec, (m - 1) * rec),
                    ("parent", rr.parent, r.parent + k0 * 4, (m - 1) * 4),
                    ("prefix_sums", rr.prefix_sums, r.prefix_sums + k0 * 4, m * 4),
                    ("octree", rr.octree, r.octree + o0 * stride, O * stride)):
                if dev.digest(a, nbytes) != dev.digest(bb, nbytes):
                    bad.append((i, name))
                    break
            if i % 512 == 511:
                log(f"c4: {i + 1} instances compared with their plain builds")
        assert not bad, f"instances differing from their plain builds: {bad[:10]}"
        picks = spread_pick(sizes, 16)
        assert int(np.argmax(sizes)) in picks
        for i in picks:
            part = pts[offs[i]:offs[i + 1]]
            want = oracle.build(oracle.quantize(part, 1 << bits), bits)
            got = batch_instance_device(dev, r, cnt, i, dim)
            assert_tree_equal(got, want, BATCH_FIELDS, where=f"c4 instance {i} (n={sizes[i]})")
        log(f"c4: oracle agrees on instances {picks}")
    finally:
        b.close()
        d.free()


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_served_c5_full_size(dev):
    count, rate, budget = 520, 2000.0, 64 << 30
    rng = np.random.default_rng(6)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, count))
    sizes = np.exp(rng.uniform(np.log(1e4), np.log(4e6), count)).astype(np.int64)
    dims = rng.choice([2, 3], count)
    pool_n, per_dim = 4_000_000, 4
    pools = {}
    try:
        for dd in (2, 3):
            for k in range(per_dim):
                pa = pgvd_amd.PinnedArray((pool_n, dd))
                pa.array[:] = rng.random((pool_n, dd), dtype=np.float32) * 2 - 1
                pools[(dd, k)] = pa
        reqs = (pgvd_amd.Request * count)()
        for i in range(count):
            dd = int(dims[i])
            q = reqs[i]
            q.host_pts = pools[(dd, i % per_dim)].ptr
            q.n, q.dim, q.bits = int(sizes[i]), dd, (16 if dd == 2 else 21)
            q.arrival_s = float(arrivals[i])
        log(f"c5: {count} requests, {int(sizes.sum())} points over {arrivals[-1]:.3f} s")
        with pgvd_amd.Server(0, slots=4, hbm_budget=budget) as srv:
            st = srv.run(reqs)
        assert st.completed == count
        assert st.max_in_flight > 1, "requests never overlapped"
        assert st.max_in_flight_bytes <= budget
        lat = np.array([reqs[i].done_s - reqs[i].arrival_s for i in range(count)])
        log(f"c5: served in {st.elapsed_s:.3f} s, p50 {np.percentile(lat, 50) * 1e3:.2f} ms, "
            f"p99 {np.percentile(lat, 99) * 1e3:.2f} ms, {st.max_in_flight} in flight at most")
        # every request against the same build alone (points already on the device)
        dpools = {key: dev.upload(pa.array) for key, pa in pools.items()}
        bad = []
        for i in range(count):
            q = reqs[i]
            assert q.status == 0, (i, q.status)
            dp = dpools[(q.dim, i % per_dim)]
            dev.build_octree_async(DeviceBuffer(dev, dp.ptr, q.n * q.dim * 4), q.n, q.dim, q.bits)
            rr = dev.build_octree_finish()
            dg = dev.digest(rr.octree, rr.octree_size * ((1 << q.dim) + 1) * 4)
            if (rr.m, rr.octree_size, dg) != (q.m, q.octree_size, q.octree_digest):
                bad.append(i)
        for dp in dpools.values():
            dp.free()
        assert not bad, f"requests differing from their builds alone: {bad[:10]}"
        # a spread of requests against the oracle
        picks = spread_pick(sizes, 8)
        for i in picks:
            q = reqs[i]
            part = np.ascontiguousarray(pools[(q.dim, i % per_dim)].array[: q.n])
            want = oracle.build(oracle.quantize(part, 1 << q.bits), q.bits)
            raw = np.concatenate([want["children"], want["leaf"][:, None]], axis=1)
            assert (q.m, q.octree_size) == (want["m"], want["octree_size"]), f"request {i}"
            assert q.octree_digest == digest_np(raw.astype(np.int32)), f"request {i} octree"
        log(f"c5: oracle agrees on requests {picks}")
    finally:
        for pa in pools.values():
            pa.free()
