"""Served form (pgvd_server_*): requests of mixed size and dimension,
admitted against an HBM budget and overlapped on the device, give exactly
the result of the same build run alone -- unique-key count, octree size and
a digest of every octree word (recomputed here in numpy from the alone
build's downloaded octree)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import pgvd_amd


def digest_np(words: np.ndarray) -> int:
    """The definition pgvd_digest implements: sum of splitmix64(w_i | i << 32)."""
    w = np.ascontiguousarray(words).reshape(-1).view(np.uint32).astype(np.uint64)
    z = w | (np.arange(w.size, dtype=np.uint64) << np.uint64(32))
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    z = z ^ (z >> np.uint64(31))
    return int(z.sum(dtype=np.uint64))


def make_trace(count, seed, lo=1e3, hi=2e5, spread_s=0.02):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), count)).astype(np.int64)
    dims = rng.choice([2, 3], count)
    arrivals = np.sort(rng.uniform(0, spread_s, count))
    pins, reqs = [], (pgvd_amd.Request * count)()
    for i in range(count):
        d = int(dims[i])
        pa = pgvd_amd.PinnedArray((int(sizes[i]), d))
        if i % 3 == 2:     # clustered: many duplicate keys
            c = rng.uniform(-0.5, 0.5, (4, d))
            pa.array[:] = np.clip(c[rng.integers(0, 4, sizes[i])] +
                                  rng.normal(0, 1e-4, (sizes[i], d)), -1, 1)
        else:
            pa.array[:] = rng.random((sizes[i], d), dtype=np.float32) * 2 - 1
        pins.append(pa)
        r = reqs[i]
        r.host_pts = pa.ptr
        r.n = int(sizes[i])
        r.dim = d
        r.bits = 16 if d == 2 else 21
        r.arrival_s = float(arrivals[i])
    return pins, reqs


def check_alone(dev, pins, reqs, idx):
    for i in idx:
        r = reqs[i]
        assert r.status == 0, (i, r.status)
        want = dev.build_octree(np.array(pins[i].array), r.bits)
        raw = np.concatenate([want["children"], want["leaf"][:, None]], axis=1).astype(np.int32)
        assert (r.m, r.octree_size) == (want["m"], want["octree_size"]), f"request {i}"
        assert r.octree_digest == digest_np(raw), f"request {i}: octree differs from the alone build"


@pytest.mark.gpu
def test_served_mix_equals_alone(dev):
    pins, reqs = make_trace(96, seed=6)
    try:
        with pgvd_amd.Server(0, slots=4, hbm_budget=8 << 30) as srv:
            st = srv.run(reqs)
        assert st.completed == len(reqs)
        assert st.max_in_flight > 1, "requests never overlapped"
        assert all(reqs[i].done_s >= reqs[i].admit_s >= reqs[i].arrival_s - 1e-6
                   for i in range(len(reqs)))
        check_alone(dev, pins, reqs, range(len(reqs)))
    finally:
        for p in pins:
            p.free()


@pytest.mark.gpu
def test_served_budget_limits_what_is_in_flight(dev):
    pins, reqs = make_trace(24, seed=7, lo=5e4, hi=2e5, spread_s=0.0)
    try:
        feet = [pgvd_amd.Server.footprint(r.n, r.dim, r.bits) for r in reqs]
        budget = int(max(feet) * 1.2)            # one large build at a time
        with pgvd_amd.Server(0, slots=4, hbm_budget=budget) as srv:
            st = srv.run(reqs)
        # what the server allocated (slot pools, points, results) stayed
        # within the budget, and so did the device's own count of used memory
        assert st.max_in_flight_bytes <= budget
        assert st.max_device_bytes <= budget + (64 << 20), (st.max_device_bytes, budget)
        assert st.completed == len(reqs)
        check_alone(dev, pins, reqs, range(0, len(reqs), 3))
    finally:
        for p in pins:
            p.free()


@pytest.mark.gpu
def test_served_rejects_a_request_over_budget():
    pins, reqs = make_trace(2, seed=8, lo=1e5, hi=1e5)
    try:
        with pgvd_amd.Server(0, slots=2, hbm_budget=1 << 20) as srv:
            with pytest.raises(pgvd_amd.PgvdError):
                srv.run(reqs)
    finally:
        for p in pins:
            p.free()


@pytest.mark.gpu
def test_served_submit_poll_release(dev):
    """Requests submitted online: polled until complete, each octree read
    from its device result (kept until released) and compared with the same
    build alone and, for three of them, with the oracle."""
    import oracle
    pins, reqs = make_trace(40, seed=9, lo=2e3, hi=3e5)
    try:
        with pgvd_amd.Server(0, slots=3, hbm_budget=16 << 30) as srv:
            tickets = [srv.submit(r.host_pts, r.n, r.dim, r.bits) for r in reqs]
            results = {}
            while len(results) < len(tickets):
                for i, t in enumerate(tickets):
                    if i not in results:
                        res = srv.poll(t)
                        if res is not None:
                            results[i] = res
            for i, t in enumerate(tickets):
                res = results[i]
                assert res.status == 0, (i, res.status)
                assert res.done_s >= res.admit_s >= res.arrival_s
                stride = (1 << res.dim) + 1
                raw = dev.download(res.octree, np.int32, res.octree_size * stride).reshape(-1, stride)
                want = dev.build_octree(np.array(pins[i].array), reqs[i].bits)
                assert (res.m, res.octree_size) == (want["m"], want["octree_size"]), i
                assert np.array_equal(raw[:, :-1], want["children"]), i
                assert np.array_equal(raw[:, -1], want["leaf"]), i
                assert res.octree_digest == digest_np(raw), i
                if i < 3:
                    q = oracle.quantize(np.array(pins[i].array), 1 << reqs[i].bits)
                    o = oracle.build(q, reqs[i].bits)
                    assert np.array_equal(raw[:, :-1], o["children"]), i
                    assert np.array_equal(raw[:, -1], o["leaf"]), i
                srv.release(t)
            with pytest.raises(pgvd_amd.PgvdError):
                srv.poll(tickets[0])
            with pytest.raises(pgvd_amd.PgvdError):
                srv.submit(reqs[0].host_pts, 1 << 29, 3, 21)    # footprint above the budget
    finally:
        for p in pins:
            p.free()
