"""GPU parity of the GVD refinement step (2D) against the oracle, through the
C ABI: FindMultiCells' segments, walks and sampling, Geom::FitBoxes and the
Karras iteration of Octree2::build (viewer/Octree2.cpp:127-190).  Float
outputs are compared bit for bit."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

import oracle
import pgvd_amd
from pgvd_amd import workloads

sys.path.insert(0, os.path.dirname(__file__))
from test_gvd_oracle import KINDS, _pairs  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = pgvd_amd.Device(0)
    yield d
    d.close()


def _bits_equal(x, y) -> bool:
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    return x.shape == y.shape and np.array_equal(x.view(np.uint32), y.view(np.uint32))


def _data2(name):
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "data_points.npz"))
    return d[name], d[name + "_counts"]


@pytest.mark.parametrize("kind", KINDS)
def test_fit_boxes_pairs_match_oracle(dev, kind):
    """FitBoxes on 1500 pairs of each shape: every pair's samples, box
    origins and widths and its status (0, the reference's logic_errors, or
    the 4096-round bound) equal the oracle's."""
    pairs = _pairs(kind, 1500, seed=KINDS.index(kind) + 11).reshape(-1, 8)
    print(f"fit_boxes kind={kind} pairs={pairs.shape[0]}", flush=True)
    got = dev.fit_boxes(pairs)
    so, bo = got["sample_offsets"], got["box_offsets"]
    for i, p in enumerate(pairs):
        want = oracle.fit_boxes(p[:4], p[4:])
        assert got["status"][i] == want["status"], (i, p)
        assert _bits_equal(got["samples"][so[i]:so[i + 1]], want["samples"]), (i, p)
        assert _bits_equal(got["origins"][bo[i]:bo[i + 1]], want["origins"]), (i, p)
        assert _bits_equal(got["lengths"][bo[i]:bo[i + 1]], want["lengths"]), (i, p)


@pytest.mark.parametrize("bits", [8, 11, 12, 16, 24])
def test_polyline_segments_match_oracle(dev, bits):
    """obj2Oct and the per-polygon segments, including the far side of the
    box, which lands on the grid width itself from 2^12 up."""
    v, c = workloads.polyline_scene(30, 3, 4, 0.01)
    lo = v.min(axis=0)
    bbw = np.float32((v.max(axis=0) - lo).max())
    got_s, got_l = dev.polyline_segments(v, c, lo, bbw, bits)
    want_s, want_l = oracle.polyline_segments(oracle.obj2oct(v, lo, bbw, bits), c)
    assert _bits_equal(got_s, want_s)
    assert np.array_equal(got_l, want_l)


def test_polyline_segments_edge_counts(dev):
    """Polygons of 0 and 1 vertex give no segment; a vertex below the box
    is PGVD_ERR_BAD_COORD (obj2Oct throws)."""
    v = np.array([[0, 0], [1, 0], [1, 1], [0.5, 0.5], [0.2, 0.9], [0.3, 0.3]], np.float32)
    c = np.array([2, 0, 1, 3], np.int64)
    got_s, got_l = dev.polyline_segments(v, c, [0, 0], 1.0, 10)
    want_s, want_l = oracle.polyline_segments(oracle.obj2oct(v, np.zeros(2, np.float32),
                                                             np.float32(1), 10), c)
    assert _bits_equal(got_s, want_s) and np.array_equal(got_l, want_l)
    assert got_s.shape[0] == 3
    with pytest.raises(pgvd_amd.PgvdError) as e:
        dev.polyline_segments(v, c, [0.5, 0], 1.0, 10)
    assert e.value.code == 3


def _scene(loops, sub, seed, gap, bits):
    v, c = workloads.polyline_scene(loops, sub, seed, gap)
    lo = v.min(axis=0)
    bbw = np.float32((v.max(axis=0) - lo).max())
    tree = oracle.build(oracle.quantize(v, 1 << bits), bits)
    segs, labels = oracle.polyline_segments(oracle.obj2oct(v, lo, bbw, bits), c)
    return tree, segs, labels


SCENES = [(9, 1, 7, 0.02, 12), (25, 2, 7, 0.02, 16), (64, 2, 8, 0.01, 16), (16, 3, 3, 0.05, 8),
          (400, 3, 9, 0.002, 20), (2500, 2, 10, 0.001, 24)]


@pytest.mark.parametrize("loops,sub,seed,gap,bits", SCENES)
def test_multi_cells_match_oracle(dev, loops, sub, seed, gap, bits):
    """The CellIntersections of every node and octant: labels and kept
    parts, bit for bit."""
    tree, segs, labels = _scene(loops, sub, seed, gap, bits)
    print(f"multi_cells loops={loops} bits={bits} segs={segs.shape[0]} "
          f"nodes={tree['octree_size']}", flush=True)
    got = dev.multi_cells(segs, labels, bits, tree["children"], tree["leaf"])
    want = oracle.multi_cells(segs, labels, bits, tree["children"], tree["leaf"])
    assert np.array_equal(got["status"], want["status"])
    assert np.array_equal(got["labels"], want["labels"])
    assert _bits_equal(got["segs"], want["segs"])


@pytest.mark.parametrize("name", ["data2_test1", "data2_test2", "data2_test3", "data2_test4",
                                  "data2_test5"])
@pytest.mark.parametrize("bits", [8, 16])
def test_multi_cells_reference_scenes(dev, name, bits):
    """The reference's polyline scenes (data2/), walked polyline by polyline."""
    v, c = _data2(name)
    lo = v.min(axis=0)
    bbw = np.float32((v.max(axis=0) - lo).max())
    tree = oracle.build(oracle.quantize(v, 1 << bits), bits)
    segs, labels = oracle.polyline_segments(oracle.obj2oct(v, lo, bbw, bits), c)
    got = dev.multi_cells(segs, labels, bits, tree["children"], tree["leaf"])
    want = oracle.multi_cells(segs, labels, bits, tree["children"], tree["leaf"])
    assert np.array_equal(got["labels"], want["labels"])
    assert _bits_equal(got["segs"], want["segs"])


@pytest.mark.parametrize("loops,sub,seed,gap,bits", SCENES[:5])
def test_find_multi_cells_match_oracle(dev, loops, sub, seed, gap, bits):
    """Extra points, samples, box origins and widths in node, octant,
    sample order."""
    tree, segs, labels = _scene(loops, sub, seed, gap, bits)
    ci = oracle.multi_cells(segs, labels, bits, tree["children"], tree["leaf"])
    want = oracle.find_multi_cells(ci["labels"], ci["segs"])
    got = dev.find_multi_cells(ci["labels"], ci["segs"])
    assert want["status"] == 0
    assert np.array_equal(got["extra"], want["extra"])
    assert _bits_equal(got["samples"], want["samples"])
    assert _bits_equal(got["origins"], want["origins"])
    assert _bits_equal(got["lengths"], want["lengths"])


def _assert_refine_equal(got, want, what):
    assert got["iterations"] == want["iterations"], what
    assert got["extra_per_iteration"] == want["extra_per_iteration"], what
    assert np.array_equal(got["qpoints"], want["qpoints"].astype(np.int32)), what
    t = want["tree"]
    assert got["octree_size"] == t["octree_size"], what
    assert np.array_equal(got["children"], t["children"]), what
    assert np.array_equal(got["leaf"], t["leaf"]), what
    assert np.array_equal(got["ci_labels"], want["ci"]["labels"]), what
    assert _bits_equal(got["ci_segs"], want["ci"]["segs"]), what
    m = want["multi"]
    assert np.array_equal(got["extra"], m["extra"]), what
    assert _bits_equal(got["samples"], m["samples"]), what
    assert _bits_equal(got["origins"], m["origins"]), what
    assert _bits_equal(got["lengths"], m["lengths"]), what


@pytest.mark.parametrize("loops,sub,seed,gap,bits,iters", [
    (9, 1, 7, 0.02, 12, 8), (25, 2, 7, 0.02, 16, 8), (64, 2, 8, 0.01, 16, 2),
    (400, 3, 9, 0.002, 20, 8), (2500, 2, 10, 0.001, 24, 8)])
def test_gvd_refine_matches_oracle(dev, loops, sub, seed, gap, bits, iters):
    """The whole Karras loop on the device: every build's extra-point count,
    the last build's points, octree, CellIntersections and extra points."""
    v, c = workloads.polyline_scene(loops, sub, seed, gap)
    print(f"gvd_refine loops={loops} bits={bits} iters={iters}", flush=True)
    got = dev.gvd_refine(v, c, bits, iters)
    want = oracle.gvd_refine(v, c, bits, iters)
    _assert_refine_equal(got, want, f"loops={loops}")
    assert sum(want["extra_per_iteration"]) > 0


@pytest.mark.parametrize("name", ["data2_test1", "data2_test2", "data2_test3", "data2_test4"])
@pytest.mark.parametrize("bits", [10, 16])
def test_gvd_refine_reference_scenes(dev, name, bits):
    """The refinement of the reference's polyline scenes, also with a custom
    box (Octree2::build's customBB) around them."""
    v, c = _data2(name)
    for bb in (None, (v.min(axis=0) - 0.25, v.max(axis=0) + 0.5)):
        got = dev.gvd_refine(v, c, bits, 6, custom_bb=None if bb is None else np.concatenate(bb))
        want = oracle.gvd_refine(v, c, bits, 6, custom_bb=bb)
        _assert_refine_equal(got, want, f"{name} bb={bb is not None}")


def test_gvd_refine_single_point(dev):
    """A scene whose vertices all coincide quantizes to one point (Karras.cpp
    :84-88): no octree is built and nothing is walked."""
    v = np.ones((3, 2), np.float32)
    got = dev.gvd_refine(v, [3], 10, 4)
    assert got["iterations"] == 1 and got["octree_size"] == 0
    assert got["qpoints"].tolist() == [[0, 0]] and got["extra"].shape[0] == 0
