"""Seeded synthetic point sets for the benchmark and parity tests.

    uniform(n, dim, seed)        n points uniform in [-1,1]^dim
    gaussian_mixture(n, ...)     32 seeded Gaussian clusters, tuned for ~30%
                                 duplicate 63-bit keys at 21 bits/axis
    star_loops(n, ...)           64 closed star-shaped loops in random planes,
                                 sampled by arc length, with loop-id labels
    batch_sizes(count, ...)      log-uniform instance sizes for batched builds
    batch_instance(i, n, seed)   instance i of a batch: uniform (even i) or
                                 8 Gaussian clusters (odd i)
    batch_points(sizes, seed)    all instances, concatenated
"""
from __future__ import annotations

import numpy as np


def uniform(n: int, dim: int = 3, seed: int = 2) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.random((n, dim), dtype=np.float32) * 2 - 1


# Cluster widths.  At 21 bits/axis a grid cell is ~9.5e-7 wide, so clusters
# of sigma 0.01-0.1 give almost no duplicate keys (m/n = 0.91 even at
# 1e-5..1e-4).  Widths of 5e-6 .. 5e-5 (the 0.01-0.1 range scaled by 5e-4)
# give m/n = 0.723 at 2^24 points, i.e. ~28% duplicate keys.
GAUSS_SIGMA = (5e-6, 5e-5)


def gaussian_mixture(n: int, dim: int = 3, seed: int = 3, k: int = 32,
                     sigma=GAUSS_SIGMA) -> tuple[np.ndarray, str]:
    rng = np.random.default_rng(seed)
    centres = rng.uniform(-0.9, 0.9, (k, dim))
    sig = rng.uniform(sigma[0], sigma[1], k)
    comp = rng.integers(0, k, n)
    pts = centres[comp] + rng.standard_normal((n, dim)) * sig[comp, None]
    pts = np.clip(pts, -1, 1).astype(np.float32)
    # the bounding box spans the clusters; pin it to [-1,1]^dim with two
    # corner points so the grid is the same as the uniform workload's
    pts[0] = -1.0
    pts[1] = 1.0
    desc = (f"3D octree build, {n} float3 from a {k}-component Gaussian mixture "
            f"(centres U[-0.9,0.9]^3 seed {seed}, sigma U[{sigma[0]:g},{sigma[1]:g}], "
            f"corners pinned to [-1,1]^3)")
    return pts, desc


def star_loops(n: int, loops: int = 64, verts: int = 256, seed: int = 4):
    """Closed loops r(t) = r0 (1 + 0.3 sin(k t + phi)) in random planes,
    r0 ~ U[0.05, 0.3]; n points sampled uniformly by arc length over all
    loops (equal share per loop); labels = loop id (uint32)."""
    rng = np.random.default_rng(seed)
    per = np.full(loops, n // loops)
    per[: n - per.sum()] += 1
    pts = np.empty((n, 3), np.float64)
    labels = np.empty(n, np.uint32)
    off = 0
    for i in range(loops):
        r0 = rng.uniform(0.05, 0.3)
        kk = int(rng.integers(3, 9))
        phi = rng.uniform(0, 2 * np.pi)
        c = rng.uniform(-0.6, 0.6, 3)
        a = rng.standard_normal(3)
        a /= np.linalg.norm(a)
        b = rng.standard_normal(3)
        b -= a * (a @ b)
        b /= np.linalg.norm(b)
        t = np.arange(verts) * (2 * np.pi / verts)
        r = r0 * (1 + 0.3 * np.sin(kk * t + phi))
        V = c + (r * np.cos(t))[:, None] * a + (r * np.sin(t))[:, None] * b
        W = np.roll(V, -1, axis=0)
        seg = np.linalg.norm(W - V, axis=1)
        cum = np.concatenate([[0.0], np.cumsum(seg)])
        s = (np.arange(per[i]) + rng.random()) * (cum[-1] / per[i])
        j = np.clip(np.searchsorted(cum, s, side="right") - 1, 0, verts - 1)
        f = (s - cum[j]) / seg[j]
        pts[off: off + per[i]] = V[j] + f[:, None] * (W[j] - V[j])
        labels[off: off + per[i]] = i
        off += per[i]
    return pts.astype(np.float32), labels


def batch_sizes(count: int = 4096, lo: float = 1e3, hi: float = 1e6, seed: int = 5) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.exp(rng.uniform(np.log(lo), np.log(hi), count)).astype(np.int64)


def batch_instance(i: int, n: int, seed: int = 5, dim: int = 3) -> np.ndarray:
    """Instance i of a batch: even i uniform in [-1,1]^dim, odd i a mixture
    of 8 Gaussian clusters (centres U[-0.9,0.9], sigma U[0.01,0.1]),
    clipped to [-1,1]; seeded by (seed, i)."""
    rng = np.random.default_rng([seed, i])
    if i % 2 == 0:
        return rng.random((n, dim), dtype=np.float32) * 2 - 1
    centres = rng.uniform(-0.9, 0.9, (8, dim)).astype(np.float32)
    sig = rng.uniform(0.01, 0.1, 8).astype(np.float32)
    comp = rng.integers(0, 8, n)
    pts = centres[comp] + rng.standard_normal((n, dim), dtype=np.float32) * sig[comp, None]
    return np.clip(pts, -1, 1)


def batch_points(sizes, seed: int = 5, dim: int = 3) -> np.ndarray:
    """All instances of a batch, one after another (float32, sum(sizes) x dim)."""
    sizes = np.asarray(sizes, np.int64)
    out = np.empty((int(sizes.sum()), dim), np.float32)
    off = 0
    for i, n in enumerate(sizes):
        out[off: off + n] = batch_instance(i, int(n), seed, dim)
        off += int(n)
    return out



def polyline_scene(loops: int = 25, sub: int = 2, seed: int = 7, gap: float = 0.02,
                   jitter: float = 0.05):
    """A 2D scene of closed polylines for the GVD refinement: `loops`
    quadrilaterals on a g x g grid (g = ceil(sqrt(loops))) over [-1,1]^2
    whose column widths and row heights vary by up to +-25% (so that the
    gaps do not fall on the octree's power-of-two cell boundaries), each
    rotated by U[-jitter, jitter] radians and sized so that neighbours face
    each other across about `gap`; every side is cut into `sub` segments.
    Across each narrow gap two loops pass through the same empty octree
    cells, where the refinement adds points.  Every loop repeats its first
    vertex at the end, as the reference's polyline files close their loops.
    Returns (vertices float32 (sum counts, 2), counts int64 per loop)."""
    rng = np.random.default_rng(seed)
    g = int(np.ceil(np.sqrt(loops)))
    edges = []
    for _ in range(2):
        w = rng.uniform(0.75, 1.25, g)
        edges.append(np.concatenate([[0.0], np.cumsum(w / w.sum() * 2.0)]) - 1.0)
    unit = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], np.float64)
    out, counts = [], []
    for i in range(loops):
        x0, x1 = edges[0][i % g], edges[0][i % g + 1]
        y0, y1 = edges[1][i // g], edges[1][i // g + 1]
        half = np.array([(x1 - x0 - gap) / 2, (y1 - y0 - gap) / 2])
        rot = rng.uniform(-jitter, jitter)
        c, s_ = np.cos(rot), np.sin(rot)
        C = (unit * half) @ np.array([[c, s_], [-s_, c]]) / (abs(c) + abs(s_))
        P = np.array([C[e] + (C[(e + 1) % 4] - C[e]) * (t / sub) for e in range(4) for t in range(sub)])
        P = P + [(x0 + x1) / 2, (y0 + y1) / 2]
        P = np.concatenate([P, P[:1]])
        out.append(P)
        counts.append(P.shape[0])
    return np.concatenate(out).astype(np.float32), np.asarray(counts, np.int64)
