"""Writes tests/golden/data_points.npz: the vertices of the reference's own
data files, read with pgvd_amd.io, as plain float32 arrays.

  2D scenes (data2/): test1 (test1-1..5), test2, test3, test4 (their files
  in order), test5..test8, with each scene's vertex count per file
  (data2_<scene>_counts: one file is one polyline);
  3D meshes (data3/): simple1..3.obj and bunny/bun1.obj.

Run from the repository root with the reference checked out at
/root/reference:  python tests/golden/make_data_points.py
"""
import glob
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from pgvd_amd import io  # noqa: E402

REF = "/root/reference"


def main():
    out = {}
    d2 = os.path.join(REF, "data2")
    for scene in ("test1", "test2", "test3", "test4", "test5", "test6", "test7", "test8"):
        files = sorted(glob.glob(os.path.join(d2, scene + "-*.dat"))) or \
            [os.path.join(d2, scene + ".dat")]
        out["data2_" + scene] = io.read_scene(files)
        # vertices per file: one file is one polyline
        out["data2_" + scene + "_counts"] = np.array(
            [io.read_polyline(f).shape[0] for f in files], np.int64)
    d3 = os.path.join(REF, "data3")
    for name in ("simple1", "simple2", "simple3"):
        out["data3_" + name] = io.read_obj_vertices(os.path.join(d3, name + ".obj"))
    out["data3_bun1"] = io.read_obj_vertices(os.path.join(d3, "bunny", "bun1.obj"))
    path = os.path.join(ROOT, "tests", "golden", "data_points.npz")
    np.savez_compressed(path, **out)
    for k, v in out.items():
        print(k, v.shape)


if __name__ == "__main__":
    main()
