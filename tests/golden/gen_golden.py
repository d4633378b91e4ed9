"""Generate the golden fixtures under tests/golden/ from the reference's own
C sources (oracle/_ref: C/BigUnsigned.c, C/BuildBRT.c, C/BuildOctree.c,
C/ParallelAlgorithms.c of edwardsjohnmartin/pgvd, compiled unmodified, plus
the harness restating opencl/Kernels.cpp:388-411).

Runs only where the reference tree and oracle/_ref exist (this container):

    python tests/golden/gen_golden.py

Outputs:
  ref_small.npz    full outputs of small builds (2D through the reference's
                   own xyz2z; 3D BRT/octree from keys)
  ref_digests.json SHA-256 digests of every output array of large builds
                   (the 2D 65,536-point quadtree in full; the BRT and octree
                   children of the 2^24-point 3D workloads and of a
                   constructed 6,000,000-key input whose root split takes the
                   reference's FP32 split value)
  morton_ref.json  digest of the reference test's Morton input (i,i),
                   bits=30, i < 10^6 (tests/KernelTests.cpp:57-84)

Every large case is also run through the oracle and the script refuses to
write a digest on which the oracle and the reference disagree (the 3D
`leaf` field excepted: the reference reads leaf_masks out of bounds there).
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from pgvd_amd import workloads  # noqa: E402

FIELDS = {
    "keys_lo": np.uint64, "keys_hi": np.uint64, "left": np.int32, "left_leaf": np.uint8,
    "right_leaf": np.uint8, "lcp_lo": np.uint64, "lcp_hi": np.uint64, "lcp_length": np.int32,
    "parent": np.int32, "local_splits": np.uint32, "prefix_sums": np.uint32,
    "children": np.int32, "leaf": np.int32,
}


def digest(a, dtype) -> str:
    return hashlib.sha256(np.ascontiguousarray(np.asarray(a).astype(dtype)).tobytes()).hexdigest()


def digests(tree: dict, skip=()) -> dict:
    out = {"m": int(tree["m"]), "octree_size": int(tree["octree_size"])}
    for k, dt in FIELDS.items():
        if k in tree and k not in skip:
            out[k] = digest(tree[k], dt)
    return out


def agree(a: dict, b: dict, skip=()) -> None:
    for k in FIELDS:
        if k in skip:
            continue
        if not np.array_equal(np.asarray(a[k]), np.asarray(b[k])):
            raise SystemExit(f"oracle and reference disagree on {k}; not writing fixtures")
    if a["m"] != b["m"] or a["octree_size"] != b["octree_size"]:
        raise SystemExit("oracle and reference disagree on sizes")


def quirk_keys(m: int = 1 << 24):
    """Sorted unique keys whose root (l = m-1) has a split the FP32 step of
    C/BuildBRT.c:151 cannot reach: keys[i] = i for i <= s*, and
    2^40 + i above it, with s* the first split position (counting from the
    middle) that the FP32 binary search misses.  The FP32 step equals the
    exact ceil whenever l <= 2^23 (every intermediate is then an exactly
    representable integer), so m must exceed 2^23 + 1; at l = 2^24 - 1
    about a third of the split positions are unreachable."""
    l = m - 1

    def search(target, exact):
        s, den = 0, 2
        while den < 2 * l:
            if exact:
                t = (l + den - 1) // den
            else:
                t = int((np.float32(l) + np.float32(den) - np.float32(1)) / np.float32(den))
            if s + t <= target:
                s += t
            den *= 2
        return s

    for s_star in range(l // 2, l):
        if search(s_star, False) != s_star:
            assert search(s_star, True) == s_star
            break
    keys = np.arange(m, dtype=np.uint64)
    keys[s_star + 1:] += np.uint64(1 << 40)
    return keys, int(s_star)


def main() -> None:
    if not oracle.ref_available(2) or not oracle.ref_available(3):
        raise SystemExit("oracle/_ref is not built (needs the reference tree): make -C oracle ref")
    small = {}
    dig = {}
    # --- 2D builds through the reference's own xyz2z ----------------------
    cases = [("c1_65536", 65536, 1), ("n1000", 1000, 11), ("n5", 5, 12), ("n1", 1, 13),
             ("n2", 2, 14), ("n4096", 4096, 15), ("n3000_dupes", 3000, 16)]
    for name, n, seed in cases:
        rng = np.random.default_rng(seed)
        pts = rng.random((n, 2), dtype=np.float32) * 2 - 1
        if name.endswith("dupes"):
            pts[1000:2000] = pts[:1000]
        q = oracle.quantize(pts, 1 << 16)
        ref = oracle.ref_build_octree_2d(q, 16)
        orc = oracle.build(q, 16)
        agree(orc, ref)
        if n <= 5000:
            small[f"{name}/pts"] = pts
            small[f"{name}/q"] = q.astype(np.int32)
            for k in FIELDS:
                small[f"{name}/{k}"] = np.asarray(ref[k]).astype(FIELDS[k])
        dig[f"2d_{name}"] = dict(digests(ref), n=n, seed=seed, dim=2, bits=16)
        print(f"2d {name}: m={ref['m']} O={ref['octree_size']}", flush=True)
    # bits = 30 (the reference tests' resolution), integer points
    rng = np.random.default_rng(21)
    q = rng.integers(0, 1 << 30, (2000, 2)).astype(np.int32)
    ref = oracle.ref_build_octree_2d(q, 30)
    agree(oracle.build(q, 30), ref)
    small["int30/q"] = q
    for k in FIELDS:
        small[f"int30/{k}"] = np.asarray(ref[k]).astype(FIELDS[k])
    # --- 3D BRT / octree children from keys -------------------------------
    rng = np.random.default_rng(22)
    q3 = (rng.random((3000, 3), dtype=np.float32) * 2 - 1)
    qq = oracle.quantize(q3, 1 << 21)
    orc = oracle.build(qq, 21)
    ref = oracle.ref_build_from_keys(orc["keys_lo"], orc["keys_hi"], 3, 63)
    agree(orc, ref, skip=("leaf",))
    small["3d3000/pts"] = q3
    for k in FIELDS:
        if k != "leaf":
            small[f"3d3000/{k}"] = np.asarray(ref[k]).astype(FIELDS[k])
    np.savez_compressed(os.path.join(HERE, "ref_small.npz"), **small)

    # --- large 3D workloads: BRT + octree children from the reference -------
    for name, gen in [("c2_uniform", lambda: (workloads.uniform(1 << 24, 3, 2), "uniform")),
                      ("c2_gauss", lambda: (workloads.gaussian_mixture(1 << 24)[0], "gauss"))]:
        t0 = time.time()
        pts, _ = gen()
        q = oracle.quantize(pts, 1 << 21)
        orc = oracle.build(q, 21)
        ref = oracle.ref_build_from_keys(orc["keys_lo"], orc["keys_hi"], 3, 63)
        agree(orc, ref, skip=("leaf",))
        d = digests(orc)      # oracle == reference on every field but leaf
        d["pts"] = digest(pts, np.float32)
        d["leaf_source"] = "oracle (reference 3D leaf is undefined)"
        dig[name] = d
        print(f"{name}: m={orc['m']} O={orc['octree_size']} ({time.time() - t0:.0f}s)", flush=True)

    # --- constructed FP32-split case --------------------------------------
    t0 = time.time()
    keys, s_star = quirk_keys()
    for dim in (2, 3):
        orc = oracle.build_from_keys(keys, None, dim, 42)
        ref = oracle.ref_build_from_keys(keys, None, dim, 42)
        agree(orc, ref, skip=("leaf",) if dim == 3 else ())
        d = digests(orc)
        d.update(root_left=int(ref["left"][0]), exact_split=s_star, m=int(keys.size))
        dig[f"quirk_{dim}d"] = d
    print(f"quirk: m={keys.size} s*={s_star} root.left={dig['quirk_2d']['root_left']} "
          f"({time.time() - t0:.0f}s)", flush=True)
    with open(os.path.join(HERE, "ref_digests.json"), "w") as f:
        json.dump(dig, f, indent=1, sort_keys=True)

    # --- the reference test's Morton input -------------------------------
    lo = np.empty(1_000_000, np.uint64)
    for i in range(1_000_000):
        lo[i] = oracle.ref_xyz2z_2d(i, i, 30)
    q = np.stack([np.arange(1_000_000)] * 2, axis=1).astype(np.int64)
    olo, ohi = oracle.morton(q, 30)
    if not (np.array_equal(olo, lo) and not ohi.any()):
        raise SystemExit("oracle Morton disagrees with the reference's xyz2z")
    with open(os.path.join(HERE, "morton_ref.json"), "w") as f:
        json.dump({"input": "(i, i) for i < 1000000, bits = 30 (tests/KernelTests.cpp:57-84)",
                   "keys_lo_sha256": digest(lo, np.uint64),
                   "samples": {str(i): int(lo[i]) for i in (0, 1, 2, 3, 12345, 999999)}},
                  f, indent=1)
    print("done")


if __name__ == "__main__":
    main()
