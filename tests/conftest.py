"""Shared fixtures.  GPU tests are marked `gpu` and use one Device for the
session; CPU tests exercise the oracle, the reference checker build, the
golden fixtures and the C ABI's load/export surface."""
from __future__ import annotations

import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def dev():
    import pgvd_amd
    d = pgvd_amd.Device(0)
    yield d
    d.close()


@pytest.fixture(scope="session")
def golden_small():
    return np.load(os.path.join(GOLDEN, "ref_small.npz"))


@pytest.fixture(scope="session")
def golden_digests():
    with open(os.path.join(GOLDEN, "ref_digests.json")) as f:
        return json.load(f)


TREE_FIELDS = ["keys_lo", "keys_hi", "left", "left_leaf", "right_leaf", "lcp_lo", "lcp_hi",
               "lcp_length", "parent", "prefix_sums", "children", "leaf"]


def assert_tree_equal(got: dict, want: dict, fields=TREE_FIELDS, where: str = ""):
    assert got["m"] == want["m"], f"{where} m {got['m']} != {want['m']}"
    assert got["octree_size"] == want["octree_size"], \
        f"{where} octree_size {got['octree_size']} != {want['octree_size']}"
    for k in fields:
        a = np.asarray(got[k])
        b = np.asarray(want[k])
        if a.shape != b.shape or not np.array_equal(a.astype(np.int64) if a.dtype != np.uint64 else a,
                                                    b.astype(np.int64) if b.dtype != np.uint64 else b):
            bad = np.nonzero(a.reshape(-1) != b.reshape(-1))[0][:5] if a.shape == b.shape else "shape"
            raise AssertionError(f"{where} field {k} differs (first at {bad})")
