"""The whole path's bucket sort (bucket.hip) through pgvd_build_octree,
against the oracle: every key width (32-, 64- and 128-bit keys), with and
without labels, at sizes that take each level-0 shape (one LDS item, one
scatter, two scatters), on clustered keys that need further levels, and on
runs of equal keys (copied rather than split)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import assert_tree_equal

import oracle
from pgvd_amd import workloads

pytestmark = pytest.mark.gpu


def _uniform(n, dim, seed):
    rng = np.random.default_rng(seed)
    return rng.random((n, dim), dtype=np.float32) * 2 - 1


def _clusters(n, dim, seed, k=8, sigma=2e-5):
    """k tight Gaussian clusters (many keys share long prefixes, so whole
    top-level buckets exceed one LDS item) plus the box corners."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-0.8, 0.8, (k, dim))
    pts = c[rng.integers(0, k, n)] + rng.standard_normal((n, dim)) * sigma
    pts = np.clip(pts, -1, 1).astype(np.float32)
    pts[0] = -1.0
    pts[1] = 1.0
    return pts


def _check(dev, pts, bits, labels=None, where="", bucket=True, **kw):
    print(f"bucket case: {where} n={pts.shape[0]} dim={pts.shape[1]} bits={bits} "
          f"labels={labels is not None}", flush=True)
    got = dev.build_octree(pts, bits, labels=labels, **kw)
    want = oracle.build(oracle.quantize(pts, 1 << bits), bits, labels=labels)
    if bucket is not None:
        assert got["bucket_sort"] == int(bucket), where
    assert_tree_equal(got, want, where=where)
    if labels is not None:
        assert np.array_equal(got["key_conflict"], want["key_conflict"]), where
        assert np.array_equal(got["node_conflict"], want["node_conflict"]), where
    return got


# (dim, bits): 32-bit keys (2D 16 bits), 63-bit (3D 21), 126-bit (3D 42)
@pytest.mark.parametrize("dim,bits", [(2, 16), (3, 21), (3, 42)])
@pytest.mark.parametrize("n", [1, 2, 9, 2000, 7000, 60_000, 1 << 20])
def test_bucket_uniform_sizes(dev, dim, bits, n):
    _check(dev, _uniform(n, dim, seed=n % 997 + dim), bits, where="uniform")


@pytest.mark.parametrize("dim,bits", [(2, 16), (3, 21), (3, 42)])
@pytest.mark.parametrize("n", [50_000, 1 << 20])
def test_bucket_clustered_levels(dev, dim, bits, n):
    """A sixth of the points in tight clusters: their top-level buckets
    exceed one LDS item and go through the levels."""
    pts = _uniform(n, dim, seed=n % 991 + dim)
    k = n // 6
    pts[:k] = _clusters(k, dim, seed=7 + dim)
    _check(dev, pts, bits, where="clusters")


@pytest.mark.parametrize("dim,bits", [(2, 16), (3, 21), (3, 42)])
def test_bucket_heavy_clusters(dev, dim, bits):
    """Every point in a tight cluster.  128-bit keys go through the levels;
    keys of at most 64 bits, with more than a quarter of them left to the
    levels, are sorted by the LSD passes instead."""
    pts = _clusters(1 << 20, dim, seed=17 + dim)
    got = _check(dev, pts, bits, where="heavy clusters", bucket=bits * dim > 64)
    assert got["m"] < pts.shape[0] or bits == 42     # duplicates at the coarser grids


def test_no_host_wait_takes_the_lsd_passes(dev):
    for dim, bits in ((2, 16), (3, 21), (3, 42)):
        _check(dev, _uniform(300_000, dim, seed=5), bits, where="no host wait", bucket=False,
               no_host_wait=True)


def test_bucket_equal_runs(dev):
    """Long runs of one point (a segment whose keys are all equal), beside
    a few others, in all three key widths."""
    for dim, bits in ((2, 16), (3, 21), (3, 42)):
        pts = _uniform(300_000, dim, seed=3)
        pts[1000:200_000] = pts[5]
        pts[250_000:260_000] = pts[7]
        # two thirds of the keys equal: narrower keys take the LSD passes
        _check(dev, pts, bits, where="equal runs", bucket=bits * dim > 64)


def test_bucket_labels(dev):
    """Labels travel with their keys through every level (unstable among
    equal keys, which no output observes): conflict flags against the
    oracle, on loops whose samples are copied onto other loops."""
    pts, lab = workloads.star_loops(1 << 19, loops=32, seed=12)
    rng = np.random.default_rng(13)
    src = rng.integers(0, pts.shape[0], 20_000)
    dst = rng.integers(0, pts.shape[0], 20_000)
    pts[dst] = pts[src]
    for bits, dim in ((42, 3), (21, 3)):
        # (63-bit keys on loops: bucket sort or LSD passes, by how clustered)
        _check(dev, pts, bits, labels=lab, where=f"labels {bits}", bucket=True if bits == 42 else None)
    p2 = _clusters(200_000, 2, seed=14, k=4, sigma=1e-3)
    l2 = np.random.default_rng(15).integers(0, 5, p2.shape[0]).astype(np.uint32)
    _check(dev, p2, 16, labels=l2, where="labels 2D", bucket=None)


def test_bucket_13_bit_level0(dev):
    """20 M uniform points: 2^13 top-level buckets (the encode kernel's larger
    bucket histogram), against the oracle."""
    _check(dev, _uniform(20_000_000, 3, seed=25), 21, where="2^13 buckets")
