"""Point loaders (pgvd_amd.io) and the committed fixture of the reference's
data files."""
import os

import numpy as np
import pytest

from pgvd_amd import io

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "data_points.npz")


def test_read_polyline_pairs(tmp_path):
    """Numbers in pairs, any whitespace, a trailing odd number dropped
    (viewer/main_pgvd2.cpp:175-201 reads x and y until the end)."""
    p = tmp_path / "a.dat"
    p.write_text("0.0002 -0.8\n0.8\t-0.8\n\n  0.1 1e-3 7")
    got = io.read_polyline(str(p))
    want = np.array([[0.0002, -0.8], [0.8, -0.8], [0.1, 1e-3]], np.float64).astype(np.float32)
    assert got.dtype == np.float32 and np.array_equal(got, want)


def test_read_obj_vertices(tmp_path):
    p = tmp_path / "m.obj"
    p.write_text("# c\nvn 0 1 0\nv 1 2 3 0.5 0.5 0.5\nv -1 0.25 4\nf 1 2 3\nvt 0 0\n")
    got = io.read_obj_vertices(str(p))
    assert np.array_equal(got, np.array([[1, 2, 3], [-1, 0.25, 4]], np.float32))


def test_reference_data_fixture():
    with np.load(GOLDEN, allow_pickle=False) as z:
        d = {k: z[k] for k in z.files}
    sizes = {k: v.shape for k, v in d.items()}
    assert sizes["data2_test1"] == (24, 2) and sizes["data3_bun1"] == (8171, 3)
    for k, v in d.items():
        if k.endswith("_counts"):        # vertices per file of a data2 scene
            assert v.dtype == np.int64 and v.sum() == d[k[:-len("_counts")]].shape[0], k
            continue
        assert v.dtype == np.float32 and np.isfinite(v).all(), k
    assert list(d["data2_test1_counts"]) and d["data2_test1_counts"].size == 5
    # closed polylines repeat their first vertex: equal points reach unique
    t2 = d["data2_test1"]
    assert len({tuple(x) for x in t2.tolist()}) < len(t2)


@pytest.mark.skipif(not os.path.isdir("/root/reference/data2"), reason="reference data not present")
def test_fixture_matches_reference_files():
    """The fixture holds what the loaders read from the reference's files."""
    with np.load(GOLDEN, allow_pickle=False) as z:
        assert np.array_equal(z["data2_test5"], io.read_polyline("/root/reference/data2/test5.dat"))
        assert np.array_equal(z["data3_bun1"],
                              io.read_obj_vertices("/root/reference/data3/bunny/bun1.obj"))
