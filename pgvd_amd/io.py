"""Point loaders for the reference's data files.

read_polyline follows the polyline reader of viewer/main_pgvd2.cpp:175-201
(commented out there): whitespace-separated numbers read as doubles in
pairs until the end of the file, each pair one vertex converted to float2;
one file is one polyline, a scene is several files.  read_obj_vertices
takes the vertices of a Wavefront OBJ file (the data3/ meshes; the
reference has no reader for them): the first three numbers of every "v"
line, as float32; colours after them are ignored.
"""
from __future__ import annotations

import numpy as np


def read_polyline(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        vals = np.array(f.read().split(), dtype=np.float64)
    n = vals.size // 2                      # a trailing odd number is dropped, as in >> x >> y
    return vals[: 2 * n].reshape(n, 2).astype(np.float32)


def read_obj_vertices(path: str) -> np.ndarray:
    out = []
    with open(path, "rb") as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 4 and parts[0] == b"v":
                out.append([float(parts[1]), float(parts[2]), float(parts[3])])
    return np.array(out, dtype=np.float64).reshape(-1, 3).astype(np.float32)


def read_scene(paths: list[str]) -> np.ndarray:
    """The vertices of several polyline files, file after file."""
    return np.concatenate([read_polyline(p) for p in paths]) if paths else np.zeros((0, 2), np.float32)
