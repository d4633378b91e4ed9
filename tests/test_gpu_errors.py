"""Error paths of the C ABI on the GPU: a device allocation that fails, a
pool limit that a build would pass, and launch failures injected at each
stage of the whole path (pgvd_debug_fail_launch).  Each must come back as
its status code with the context still usable: the same context then
builds the same points again and the result equals the oracle's."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import pgvd_amd as P
from conftest import assert_tree_equal

pytestmark = pytest.mark.gpu

ALLOC, INVALID, HIP = 5, 1, 6


def _pts(n, dim, seed):
    rng = np.random.default_rng(seed)
    return rng.random((n, dim), dtype=np.float32) * 2 - 1


@pytest.fixture(scope="module")
def dev():
    d = P.Device(0)
    yield d
    d.close()


def test_device_allocation_failure(dev):
    """An allocation far beyond the device's memory is PGVD_ERR_ALLOC; the
    failed hipMalloc leaves no error behind for the next call's launch
    checks to report."""
    with pytest.raises(P.PgvdError) as e:
        dev.alloc(1 << 46)
    assert e.value.code == ALLOC
    pts = _pts(5000, 3, 1)
    want = oracle.build(oracle.quantize(pts, 1 << 21), 21)
    assert_tree_equal(dev.build_octree(pts, 21), want, where="after a failed allocation")


def test_pool_limit_fails_a_build_with_alloc():
    """A build whose pool buffers would pass the context's limit fails with
    PGVD_ERR_ALLOC before allocating past it; no build is left in flight;
    without the limit, and then with a limit at what the build holds, the
    same context builds the oracle's tree."""
    pts = _pts(200000, 3, 2)
    want = oracle.build(oracle.quantize(pts, 1 << 21), 21)
    with P.Device(0) as d:
        d.set_pool_limit(1 << 20)
        with pytest.raises(P.PgvdError) as e:
            d.build_octree(pts, 21)
        assert e.value.code == ALLOC and "pool limit" in str(e.value)
        assert d.pool_bytes() <= 1 << 20
        with pytest.raises(P.PgvdError) as e2:
            d.build_octree_finish()
        assert e2.value.code == INVALID
        d.set_pool_limit(0)
        assert_tree_equal(d.build_octree(pts, 21), want, where="limit lifted")
        held = d.pool_bytes()
        assert held > 1 << 20
        d.set_pool_limit(held)
        assert_tree_equal(d.build_octree(pts, 21), want, where="limit at the footprint")
        assert d.pool_bytes() == held


STAGES = ["encode", "sort", "unique", "brt", "splits_scan", "malformed path", "octree"]


@pytest.mark.parametrize("stage", STAGES)
@pytest.mark.parametrize("n", [30000, 1 << 20])
def test_injected_launch_failure(dev, stage, n):
    """A launch failure at `stage` (after the stages before it were
    enqueued) is PGVD_ERR_HIP naming the stage; the next build on the same
    context is whole.  2^20 points take the bucket sort, 30000 the LSD
    passes."""
    pts = _pts(n, 3, 3)
    want = oracle.build(oracle.quantize(pts, 1 << 21), 21)
    print(f"inject stage={stage} n={n}", flush=True)
    P.debug_fail_launch(stage)
    try:
        with pytest.raises(P.PgvdError) as e:
            dev.build_octree(pts, 21)
    finally:
        P.debug_fail_launch(None)
    assert e.value.code == HIP and "injected" in str(e.value) and stage in str(e.value)
    assert_tree_equal(dev.build_octree(pts, 21), want, where=f"after {stage} failed")


def test_injected_failure_of_the_prep_stage(dev):
    """int32 grid input starts with the prep launch (no bbox pass)."""
    q = np.random.default_rng(4).integers(0, 1 << 16, (20000, 2), dtype=np.int32)
    want = oracle.build(q.astype(np.int64), 16)
    P.debug_fail_launch("prep")
    try:
        with pytest.raises(P.PgvdError) as e:
            dev.build_octree(q, 16)
    finally:
        P.debug_fail_launch(None)
    assert e.value.code == HIP and "prep" in str(e.value)
    assert_tree_equal(dev.build_octree(q, 16), want, where="after prep failed")


def test_injected_failure_of_a_stage_entry(dev):
    """A stage entry point (FindLeaf) reports the injected failure and works
    on the next call."""
    pts = _pts(20000, 3, 5)
    t = oracle.build(oracle.quantize(pts, 1 << 21), 21)
    q = np.random.default_rng(6).integers(0, 1 << 21, (4096, 3), dtype=np.int32)
    P.debug_fail_launch("find_leaf")
    try:
        with pytest.raises(P.PgvdError) as e:
            dev.find_leaf(q, 21, t["children"], t["leaf"])
    finally:
        P.debug_fail_launch(None)
    assert e.value.code == HIP
    got = dev.find_leaf(q, 21, t["children"], t["leaf"])
    want = oracle.find_leaf(q, 21, t["children"], t["leaf"])
    for k in ("parent", "octant", "data", "level", "status"):
        assert np.array_equal(np.asarray(got[k]).astype(np.int64), np.asarray(want[k]).astype(np.int64)), k
