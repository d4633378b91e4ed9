"""Pins for the oracle that do not come from the reference's fixtures:
closed forms, brute force on tiny inputs, and invariants a plausible mistake
(dropped term, wrong sign or index, transposed operand) would break."""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle


def bitlen(x: int) -> int:
    return x.bit_length()


# ---------------------------------------------------------------- delta ----
@pytest.mark.parametrize("mbits", [8, 32, 42, 63, 64, 87, 126, 128])
def test_lcp_length_is_mbits_minus_bitlen_of_xor(mbits):
    """compute_lcp_length's shift search (C/BuildBRT.c:54-88) equals
    mbits - bitlen(a ^ b): 2000 random pairs, pairs sharing long prefixes,
    and equal keys."""
    rng = np.random.default_rng(mbits)
    M = (1 << mbits) - 1
    for _ in range(2000):
        a = int(rng.integers(0, 2**62)) << 66 | int(rng.integers(0, 2**62)) << 4 | int(rng.integers(0, 16))
        a &= M
        share = int(rng.integers(0, mbits + 1))
        b = (a ^ (int(rng.integers(1, 2**62)) >> share)) & M if share < mbits else a
        assert oracle.lcp_length(a, b, mbits) == mbits - bitlen(a ^ b)
    assert oracle.lcp_length(M, M, mbits) == mbits
    assert oracle.lcp_length(0, M, mbits) == 0


@pytest.mark.skipif(not oracle.ref_available(2), reason="reference checker not built here")
def test_reference_lcp_length_matches_closed_form():
    """The reference's own compute_lcp_length, on BigUnsigned keys <= 88 bits."""
    rng = np.random.default_rng(3)
    for mbits in (32, 60, 63, 87):
        M = (1 << mbits) - 1
        for _ in range(3000):
            a = (int(rng.integers(0, 2**62)) << 30 | int(rng.integers(0, 2**30))) & M
            b = (a ^ (1 << int(rng.integers(0, mbits)))) if rng.random() < .5 else \
                (int(rng.integers(0, 2**62)) << 30 | int(rng.integers(0, 2**30))) & M
            assert oracle.ref_lcp_length(a, b, mbits) == mbits - bitlen(a ^ b)


# --------------------------------------------------------------- morton ----
def _interleave_brute(c, dim, bits):
    z = 0
    for i in range(bits):
        for k in range(dim):
            if (c[k] >> i) & 1:
                z |= 1 << (i * dim + k)
    return z


def _deinterleave(z, dim, bits):
    c = [0] * dim
    for i in range(bits):
        for k in range(dim):
            c[k] |= ((z >> (i * dim + k)) & 1) << i
    return c


@pytest.mark.parametrize("dim,bits", [(2, 16), (2, 30), (2, 64), (3, 21), (3, 42), (3, 10)])
def test_morton_is_the_bit_interleave(dim, bits):
    rng = np.random.default_rng(dim * 100 + bits)
    n = 500
    q = rng.integers(0, 2**62, (n, dim), dtype=np.int64)
    lo, hi = oracle.morton(q, bits)
    for i in range(n):
        z = int(lo[i]) | int(hi[i]) << 64
        c = [int(x) & ((1 << bits) - 1) for x in q[i]]     # bits >= `bits` ignored
        assert z == _interleave_brute(c, dim, bits)
        assert _deinterleave(z, dim, bits) == c
    # axis order: x is the lowest bit of each digit
    lo, _ = oracle.morton(np.array([[1] + [0] * (dim - 1)], np.int64), bits)
    assert int(lo[0]) == 1
    lo, _ = oracle.morton(np.array([[0] * (dim - 1) + [1]], np.int64), bits)
    assert int(lo[0]) == 1 << (dim - 1)


@pytest.mark.skipif(not oracle.ref_available(2), reason="reference checker not built here")
def test_reference_xyz2z_matches_oracle():
    rng = np.random.default_rng(4)
    q = rng.integers(0, 2**30, (300, 2))
    lo, hi = oracle.morton(q.astype(np.int64), 30)
    for i in range(300):
        assert oracle.ref_xyz2z_2d(int(q[i, 0]), int(q[i, 1]), 30) == int(lo[i])


# ------------------------------------------------------------ next_pow2 ----
def test_next_pow2():
    for n in list(range(1, 70)) + [1000, 65535, 65536, 65537, (1 << 24) - 1, 1 << 24, (1 << 28) + 1]:
        want = max(8, 1 << math.ceil(math.log2(n))) if n > 1 else 8
        assert oracle.next_pow2(n) == want, n


# ----------------------------------------------------------------- scan ----
def test_hillis_steele_scan_equals_prefix_sum():
    rng = np.random.default_rng(5)
    for n in (1, 2, 3, 7, 8, 9, 1000, 4097, 65536):
        x = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
        assert np.array_equal(oracle.stream_scan(x), np.cumsum(x, dtype=np.uint64).astype(np.uint32))


# ------------------------------------------------------------- quantize ----
def test_quantize_endpoints_and_rounding():
    # bbox [-1, 1] in x, W' = 2^16 - 1
    pts = np.array([[-1, -1], [1, 1], [0, 0], [-1, 1]], np.float32)
    q = oracle.quantize(pts, 1 << 16)
    assert q.tolist() == [[0, 0], [65535, 65535], [32768, 32768], [0, 65535]]
    # clamped uses W instead of W - 1 (Karras.cpp:46-49)
    q = oracle.quantize(pts, 1 << 16, clamped=True)
    assert q[1].tolist() == [65536, 65536]
    # half-way values round up (+0.5 then truncate), not to even:
    # t = 2.5/65535 * ... choose p so that W' * t is exactly k + 0.5
    W1 = 65535
    for k in (2, 3, 1000):
        p = np.float32(-1 + 2 * (k + 0.5) / W1)
        t = (np.float32(p) - np.float32(-1)) / np.float32(2)
        f = np.float32(W1) * np.float32(t)
        if float(f) == k + 0.5:
            assert oracle.quantize(np.array([[-1, -1], [1, 1], [p, p]], np.float32), 1 << 16)[2, 0] == k + 1


def test_quantize_is_float32_not_float64():
    """A point whose float32 product lands on .5 while the exact product
    is below it: the FP32 expression of Karras.cpp:53-55 rounds up, a
    float64 evaluation would not."""
    W1 = (1 << 21) - 1
    rng = np.random.default_rng(6)
    found = 0
    for p in rng.random(200000).astype(np.float32) * 2 - 1:
        t32 = (np.float32(p) - np.float32(-1)) / np.float32(2)
        f32 = np.float32(W1) * t32
        exact = W1 * (float(p) + 1) / 2
        v32 = int(float(f32) + 0.5)
        v64 = int(exact + 0.5)
        if v32 != v64:
            got = oracle.quantize(np.array([[-1, -1, -1], [1, 1, 1], [p, p, p]], np.float32), 1 << 21)
            assert got[2, 0] == v32 != v64
            found += 1
            if found >= 5:
                break
    assert found >= 1


def test_quantize_degenerate_and_errors():
    pts = np.full((10, 3), 0.25, np.float32)
    q = oracle.quantize(pts, 1 << 21)
    assert q.shape == (1, 3) and not q.any()           # Karras.cpp:84-88
    with pytest.raises(ValueError):                     # Karras.cpp:56-63
        oracle.quantize(np.array([[-5, 0], [0.5, 0.5]], np.float32), 1 << 16,
                        bb=(np.array([0, 0], np.float32), np.array([1, 1], np.float32)))
    # slightly below min still rounds to 0 (truncation toward zero of d+0.5)
    q = oracle.quantize(np.array([[-1e-6, 0.5]], np.float32), 1 << 16,
                        bb=(np.array([0, 0], np.float32), np.array([1, 1], np.float32)))
    assert q[0, 0] == 0


# ------------------------------------------------------ tree invariants ----
def _tree(n, dim, bits, seed, dupes=False):
    rng = np.random.default_rng(seed)
    pts = rng.random((n, dim), dtype=np.float32) * 2 - 1
    if dupes:
        pts[n // 2:] = pts[: n - n // 2]
    q = oracle.quantize(pts, 1 << bits)
    return q, oracle.build(q, bits)


def _delta(a, b, mbits):
    return mbits - (a ^ b).bit_length()


@pytest.mark.parametrize("n,dim,bits,seed", [(3000, 2, 16, 1), (2000, 3, 21, 2), (500, 3, 42, 3),
                                             (1000, 2, 16, 4)])
def test_brt_structure(n, dim, bits, seed):
    """Karras 2012: every internal node but the root is claimed exactly once,
    leaf children cover every key exactly once, a node's lcp is the common
    prefix of its range, and its split is at the first bit where the range
    differs (keys below the tree's FP32 zone, l <= 2^23)."""
    _, t = _tree(n, dim, bits, seed)
    m, mbits = t["m"], dim * bits
    keys = [int(a) | int(b) << 64 for a, b in zip(t["keys_lo"], t["keys_hi"])]
    claims = np.zeros(m, int)
    leaves = np.zeros(m, int)
    for i in range(m - 1):
        L = int(t["left"][i])
        for side, leaf in ((0, t["left_leaf"][i]), (1, t["right_leaf"][i])):
            (leaves if leaf else claims)[L + side] += 1
    assert claims[0] == 0 and (claims[1:m - 1] == 1).all()
    assert (leaves == 1).all()
    assert t["parent"][0] == -1
    for i in range(m - 1):
        L = int(t["left"][i])
        # range of i: scan children ranges via lcp: keys of the range share lcp_length bits
        ln = int(t["lcp_length"][i])
        assert _delta(keys[L], keys[L + 1], mbits) == ln       # split point differs right below lcp
        lcp = int(t["lcp_lo"][i]) | int(t["lcp_hi"][i]) << 64
        assert lcp == (keys[i] >> (mbits - ln) if ln else 0)


@pytest.mark.parametrize("n,dim,bits,seed,dupes", [(3000, 2, 16, 1, False), (2000, 3, 21, 2, True),
                                                   (700, 3, 42, 3, False), (9, 2, 16, 4, False)])
def test_octree_descent_reaches_every_key(n, dim, bits, seed, dupes):
    """Top-down walk of the octree by key digits, starting from node 0 at the
    root's level, ends in an empty slot for every unique key, no two keys
    share a slot, and `leaf` marks exactly the empty slots."""
    _, t = _tree(n, dim, bits, seed, dupes)
    mbits, nc = dim * bits, 1 << dim
    ch, lf = t["children"], t["leaf"]
    for o in range(t["octree_size"]):
        assert lf[o] == sum(1 << c for c in range(nc) if ch[o, c] == -1)
    if t["m"] < 2:
        return
    L0 = int(t["lcp_length"][0]) // dim
    seen = set()
    for lo, hi in zip(t["keys_lo"], t["keys_hi"]):
        z = int(lo) | int(hi) << 64
        node, lvl = 0, L0
        while True:
            o = (z >> (mbits - (lvl + 1) * dim)) & (nc - 1)
            c = int(ch[node, o])
            if c == -1:
                break
            node, lvl = c, lvl + 1
        assert (node, o) not in seen
        seen.add((node, o))


def test_conflict_flags_brute_force():
    """A unique key is in conflict when its points carry more than one label;
    the (node, octant) mask matches the descent slot of each such key."""
    rng = np.random.default_rng(8)
    n, dim, bits = 4000, 3, 12
    q = rng.integers(0, 40, (n, dim)).astype(np.int64)       # many duplicate keys
    labels = rng.integers(0, 3, n).astype(np.uint32)
    labels[q[:, 0] % 2 == 0] = 7                             # some keys single-labelled
    t = oracle.build(q, bits, labels=labels)
    lo, _ = oracle.morton(q, bits)
    groups = {}
    for z, lab in zip(lo.tolist(), labels.tolist()):
        groups.setdefault(z, set()).add(lab)
    keys = t["keys_lo"].tolist()
    for r, z in enumerate(keys):
        want = len(groups.get(z, set())) > 1
        assert bool(t["key_conflict"][r]) == want
    mbits, nc = dim * bits, 1 << dim
    ch = t["children"]
    L0 = int(t["lcp_length"][0]) // dim
    mask = np.zeros(t["octree_size"], np.uint32)
    for r, z in enumerate(keys):
        if not t["key_conflict"][r]:
            continue
        node, lvl = 0, L0
        while True:
            o = (z >> (mbits - (lvl + 1) * dim)) & (nc - 1)
            if ch[node, o] == -1:
                break
            node, lvl = int(ch[node, o]), lvl + 1
        mask[node] |= 1 << o
    assert np.array_equal(mask, t["node_conflict"])
    assert t["key_conflict"].sum() > 0


def test_padding_key_zero_present_iff_not_power_of_two():
    """Zero-key padding (opencl/Kernels.cpp:62-75): key 0 is among the unique
    keys when n is not a power of two or n < 8, even if no point maps to 0."""
    for n in (5, 8, 9, 16, 17):
        q = np.full((n, 2), 7, np.int64)
        q[:, 0] = np.arange(1, n + 1)
        t = oracle.build(q, 16)
        has0 = int(t["keys_lo"][0]) == 0
        assert has0 == (oracle.next_pow2(n) != n), n


def test_quirk_fixture_distinguishes_fp32_from_exact_ceil(golden_digests):
    """Planted mistake: an exact-ceil split search on the fixture's root finds
    the exact split, which the reference (and the fixture) do not."""
    from tests.golden.gen_golden import quirk_keys
    keys, s_star = quirk_keys()
    l = keys.size - 1
    s, den = 0, 2
    while den < 2 * l:
        t = (l + den - 1) // den
        if s + t <= s_star:
            s += t
        den *= 2
    assert s == s_star != golden_digests["quirk_2d"]["root_left"]


# ------------------------------------------------------------- find leaf ----
def _reachable_cells(children, dim):
    """Level and cell (the octant digits from node 0 down, as a tuple) of
    every node reachable from node 0, by walking child links top-down."""
    nc = 1 << dim
    cells = {0: (0, ())}
    stack = [0]
    while stack:
        node = stack.pop()
        lvl, path = cells[node]
        for o in range(nc):
            c = int(children[node, o])
            if c >= 0:
                cells[c] = (lvl + 1, path + (o,))
                stack.append(c)
    return cells


@pytest.mark.parametrize("dim,bits,n", [(2, 16, 3000), (3, 21, 2000), (3, 6, 1500), (2, 30, 800)])
def test_find_leaf_against_brute_force(dim, bits, n):
    """OctreeUtils::FindLeaf (oracle) against a brute force: label every
    reachable node with its level and the digits of its cell, and take for
    each query the deepest node whose cell holds it; its octant for the query
    must be a leaf, and that is the reference's result."""
    rng = np.random.default_rng(dim * 1000 + bits)
    pts = (rng.random((n, dim), dtype=np.float32) * 2 - 1)
    pts[: n // 3] *= np.float32(0.01)                      # a dense cluster too
    q = oracle.quantize(pts, 1 << bits)
    t = oracle.build(q, bits)
    queries = np.concatenate([q, rng.integers(0, 1 << bits, (400, dim))]).astype(np.int64)
    got = oracle.find_leaf(queries, bits, t["children"], t["leaf"])
    nc = 1 << dim
    cells = _reachable_cells(t["children"], dim)
    by_path = {path: node for node, (lvl, path) in cells.items()}

    def digit(p, lvl):   # the point's octant below a node of level lvl
        return sum(((int(p[k]) >> (bits - 1 - lvl)) & 1) << k for k in range(dim))

    for i, p in enumerate(queries):
        path, node, lvl = (), 0, 0
        while lvl < bits:
            nxt = path + (digit(p, lvl),)
            if nxt not in by_path:
                break
            path, node, lvl = nxt, by_path[nxt], lvl + 1
        o = digit(p, lvl)
        assert (int(t["leaf"][node]) >> o) & 1, (i, node, o)
        assert got["status"][i] == 0
        assert (got["parent"][i], got["octant"][i], got["level"][i]) == (node, o, lvl + 1), i
        assert got["data"][i] == t["children"][node, o] == -1
    # the octant mask matters: below the first digit the unmasked value of
    # the reference's line 37 would exceed 2^dim - 1 (an index past the
    # node); most queries descend further than that
    assert int((got["level"] > 2).sum()) > queries.shape[0] // 2


# ------------------------------------------------------ octree text form ----
def test_octree_text_known_answer():
    """operator<< (C/OctNode.h:58-87): the size and a space, then every node
    as its children and leaf, each value followed by one space."""
    octree = np.array([[1, -1, -1, -1, 14], [-1, -1, -1, -1, 15]], np.int32)
    assert oracle.octree_format_text(octree, 2) == b"2 1 -1 -1 -1 14 -1 -1 -1 -1 15 "
    ext = np.array([[-2147483648, 2147483647, 0, 10, 255]], np.int32)
    assert oracle.octree_format_text(ext, 2) == b"1 -2147483648 2147483647 0 10 255 "
    assert oracle.octree_format_text(np.zeros((0, 9), np.int32), 3) == b"0 "


def test_octree_text_parse_rules():
    """operator>> (C/OctNode.h:69-98): whitespace of any kind between
    integers, an optional sign, the leaf cast to unsigned char, and trailing
    tokens ignored; malformed or missing tokens are rejected."""
    want = np.array([[1, -1, -1, -1, 14], [-1, -1, -1, -1, 15]], np.int32)
    for text in (b"2 1 -1 -1 -1 14 -1 -1 -1 -1 15 ", b"2\n1\t-1  -1 -1 +14\r\n-1 -1 -1 -1 15",
                 b"  2 1 -1 -1 -1 14 -1 -1 -1 -1 15 7 7"):
        assert np.array_equal(oracle.octree_parse_text(text, 2), want), text
    got = oracle.octree_parse_text(b"1 0 0 0 0 300 ", 2)
    assert got[0, 4] == 300 % 256
    for bad in (b"2 1 -1 -1 -1 14 -1 -1 -1 -1", b"1 1 x -1 -1 14", b"1 2147483648 0 0 0 0",
                b"1 - 0 0 0 0", b"", b"-1 "):
        assert oracle.octree_parse_text(bad, 2) is None, bad
    assert oracle.octree_parse_text(b"1 -2147483648 0 0 0 0 ", 2)[0, 0] == -2147483648


def test_octree_text_parse_leading_zeros_and_signs():
    """Known answers of operator>> on int (std::num_get): leading zeros are
    skipped however many, a sign may precede them, "-0" is a count of zero
    nodes, and the value (not the digit count) decides overflow."""
    want = np.array([[1, -1, 0, 7, 14]], np.int32)
    for text in (b"1 00000000001 -1 0 +007 14", b"01 1 -0000000000000001 -0 7 0000000000000000000014",
                 b"+1 1 -1 +0 7 14 junk"):
        assert np.array_equal(oracle.octree_parse_text(text, 2), want), text
    assert oracle.octree_parse_text(b"-0 ", 2).shape == (0, 5)
    assert oracle.octree_parse_text(b"1 -000000000002147483648 0002147483647 0 0 0", 2)[0, :2].tolist() == \
        [-2147483648, 2147483647]
    for bad in (b"1 000000000002147483648 0 0 0 0", b"1 99999999999 0 0 0 0", b"-1 ", b"1 0 0 0 0 -",
                b"1 0 0 0 0 00x"):
        assert oracle.octree_parse_text(bad, 2) is None, bad


def _token_soup(rng, ntok):
    """Whitespace-separated integer tokens of the kinds operator>> meets:
    plain, zero-padded, signed, at and past the int32 limits, and a few
    non-numbers; every token ends at whitespace (a token with a numeric
    prefix and other bytes after it is not generated: the parser rejects
    it, operator>> reads the prefix and fails only on the next read)."""
    out = []
    for _ in range(ntok):
        k = rng.integers(0, 12)
        v = int(rng.integers(-2 ** 31, 2 ** 31))
        if k == 0:
            tok = str(v)
        elif k == 1:
            tok = ("-" if v < 0 else rng.choice(["", "+"])) + "0" * int(rng.integers(1, 30)) + str(abs(v))
        elif k == 2:
            tok = str(rng.choice([2 ** 31 - 1, -2 ** 31, 2 ** 31, -2 ** 31 - 1, 10 ** 10, 0]))
        elif k == 3:
            tok = str(rng.choice(["-0", "+0", "000", "-", "+", "x", "1x"[1:], "--1", "+-2"]))
        else:
            tok = str(int(rng.integers(-1, 300)))
        out.append(tok)
    seps = [" ", "\n", "\t", "  ", "\r\n", "\v", "\f"]
    return "".join(t + seps[int(rng.integers(0, len(seps)))] for t in out).encode()


@pytest.mark.skipif(not oracle.ref_available(2) or not oracle.ref_available(3),
                    reason="the reference tree was not built into oracle/_ref")
@pytest.mark.parametrize("dim", [2, 3])
def test_octree_text_parse_matches_istream(dim):
    """The oracle's parser against the reference's stream input of an
    octree (C/OctNode.h:69-98) run on std::istringstream: the same records,
    or failure on both, over token soups with and without bad tokens."""
    rng = np.random.default_rng(dim)
    stride = (1 << dim) + 1
    agree_ok = agree_bad = 0
    for trial in range(400):
        n = int(rng.integers(0, 6))
        body = _token_soup(rng, n * stride + int(rng.integers(-1, 3)))
        if trial % 2 == 0:    # valid tokens only, most of the time well formed
            body = " ".join(str(int(x)) if i % 3 else "0" * int(rng.integers(0, 4)) + str(abs(int(x)))
                            for i, x in enumerate(rng.integers(-1000, 1000, n * stride))).encode()
        head = rng.choice([str(n), "0" * int(rng.integers(1, 5)) + str(n), "+" + str(n)])
        text = head.encode() + b" " + body
        got = oracle.octree_parse_text(text, dim)
        want = oracle.ref_istream_octree(text, dim)
        if want is None:
            assert got is None, text
            agree_bad += 1
        else:
            assert got is not None and np.array_equal(got, want), text
            agree_ok += 1
    assert agree_ok > 100 and agree_bad > 50


@pytest.mark.parametrize("dim,n", [(2, 5000), (3, 20000)])
def test_octree_text_round_trip(dim, n):
    rng = np.random.default_rng(dim)
    t = oracle.build(rng.integers(0, 1 << 16, (n, dim)), 16)
    octree = np.concatenate([t["children"], t["leaf"][:, None]], axis=1)
    text = oracle.octree_format_text(octree, dim)
    assert np.array_equal(oracle.octree_parse_text(text, dim), octree)


# ------------------------------------------------------------ cell walk ----
def _leaf_boxes(t, bits, owners=False):
    """(origin x, origin y, width) of every leaf cell, labelled top-down;
    with owners, a dict from each box to (node, octant) of the node whose
    octant it is."""
    cells = _reachable_cells(t["children"], 2)
    boxes = {}
    for node, (lvl, path) in cells.items():
        for o in range(4):
            if (int(t["leaf"][node]) >> o) & 1:
                p = path + (o,)
                org = [sum(((d >> k) & 1) << (bits - 1 - l) for l, d in enumerate(p))
                       for k in range(2)]
                boxes[(org[0], org[1], 1 << (bits - len(p)))] = (node, o)
    return boxes if owners else list(boxes)


def _overlap(a, b, lo, hi):
    """Parameter interval of segment ab inside the box [lo, hi]^2 (Liang-Barsky,
    double precision), or None."""
    t0, t1 = 0.0, 1.0
    for k in range(2):
        d = float(b[k]) - float(a[k])
        if d == 0.0:
            if not lo[k] <= a[k] <= hi[k]:
                return None
            continue
        ta, tb = (lo[k] - float(a[k])) / d, (hi[k] - float(a[k])) / d
        t0, t1 = max(t0, min(ta, tb)), min(t1, max(ta, tb))
        if t0 > t1:
            return None
    return t0, t1


@pytest.mark.parametrize("bits,npts,seed", [(10, 400, 1), (12, 2000, 2), (8, 60, 3)])
def test_cell_walk_against_brute_force(bits, npts, seed):
    """CellWalk (viewer/Octree2.cpp:265-313, oracle) visits every leaf cell
    the segment crosses and no cell it misses: a cell whose box shrunk by
    delta meets the segment over a positive length must be visited, one whose
    box grown by delta misses it must not be (cells between are ambiguous at
    the reference's epsilons and are not checked).  The first cell holds a."""
    rng = np.random.default_rng(seed)
    t = oracle.build(rng.integers(0, 1 << bits, (npts, 2)), bits)
    boxes = _leaf_boxes(t, bits)
    segs = (rng.random((300, 4)) * ((1 << bits) - 1e-3)).astype(np.float32)
    w = oracle.cell_walk(segs, bits, t["children"], t["leaf"])
    delta = 1e-3
    checked = 0
    for s in range(segs.shape[0]):
        a, b = segs[s, :2], segs[s, 2:]
        got = w["cells"][w["offsets"][s]:w["offsets"][s + 1]]
        assert w["status"][s] == 0
        visited = {tuple(int(x) for x in c[:3]) for c in got}
        ox, oy, wd = got[0][:3]
        assert ox <= int(a[0]) < ox + wd and oy <= int(a[1]) < oy + wd
        for (x, y, wd) in boxes:
            inner = _overlap(a, b, (x + delta, y + delta), (x + wd - delta, y + wd - delta))
            outer = _overlap(a, b, (x - delta, y - delta), (x + wd + delta, y + wd + delta))
            if inner is not None and inner[1] > inner[0]:
                assert (x, y, wd) in visited, (s, (x, y, wd))
                checked += 1
            elif outer is None:
                assert (x, y, wd) not in visited, (s, (x, y, wd))
    assert checked > segs.shape[0]


def test_cell_walk_planted_mistake_fails():
    """The brute-force check catches a walk that stops one cell early."""
    rng = np.random.default_rng(4)
    t = oracle.build(rng.integers(0, 1 << 10, (400, 2)), 10)
    segs = (rng.random((50, 4)) * 1023).astype(np.float32)
    w = oracle.cell_walk(segs, 10, t["children"], t["leaf"])
    boxes = _leaf_boxes(t, 10)
    missed = 0
    for s in range(50):
        got = w["cells"][w["offsets"][s]:w["offsets"][s + 1] - 1]     # last cell dropped
        visited = {tuple(int(x) for x in c[:3]) for c in got}
        a, b = segs[s, :2], segs[s, 2:]
        for (x, y, wd) in boxes:
            inner = _overlap(a, b, (x + 1e-3, y + 1e-3), (x + wd - 1e-3, y + wd - 1e-3))
            if inner is not None and inner[1] > inner[0] and (x, y, wd) not in visited:
                missed += 1
    assert missed > 0


def _walk_violations(segs, w, owners, delta=1e-3):
    """What a brute force finds wrong with a walk: a cell that is not a leaf
    or whose (parent, octant) is not the node and octant holding it, a cell
    visited twice, cells out of order along the segment (the midpoints of
    their parameter intervals, on boxes grown by delta, must not decrease),
    a last cell whose closed box does not hold b, or offsets that disagree
    with the counts."""
    bad = []
    if not np.array_equal(np.diff(w["offsets"]), w["counts"]):
        bad.append("offsets")
    for s in range(segs.shape[0]):
        a, b = segs[s, :2], segs[s, 2:]
        got = w["cells"][w["offsets"][s]:w["offsets"][s + 1]]
        keys = [tuple(int(x) for x in c[:3]) for c in got]
        if len(set(keys)) != len(keys):
            bad.append((s, "repeat"))
        mids = []
        for c, k in zip(got, keys):
            if owners.get(k) != (int(c[3]), int(c[4])):
                bad.append((s, "owner", k))
            x, y, wd = k
            iv = _overlap(a, b, (x - delta, y - delta), (x + wd + delta, y + wd + delta))
            mids.append(-1.0 if iv is None else 0.5 * (iv[0] + iv[1]))
        length = float(np.hypot(*(b.astype(np.float64) - a)))
        tol = 4 * delta / max(length, 1e-9)
        if any(m1 < m0 - tol for m0, m1 in zip(mids, mids[1:])) or min(mids, default=0) < 0:
            bad.append((s, "order"))
        # b in the last cell's closed box (an end on a grid line, reached in
        # a decreasing direction, ends in the cell below or left of it)
        x, y, wd = keys[-1]
        if w["status"][s] == 0 and not (x <= b[0] <= x + wd and y <= b[1] <= y + wd):
            bad.append((s, "last"))
    return bad


@pytest.mark.parametrize("bits,npts,seed", [(10, 400, 5), (12, 3000, 6), (6, 40, 7), (16, 5000, 8)])
def test_cell_walk_order_owners_no_repeats(bits, npts, seed):
    """CellWalk (viewer/Octree2.cpp:265-313, oracle) visits its cells in order
    along the segment, never twice, ends in b's cell, and labels each cell
    with the node and octant that hold it (the OctCell's parent_idx and
    octant); random segments, axis-parallel ones and ones on grid lines."""
    rng = np.random.default_rng(seed)
    t = oracle.build(rng.integers(0, 1 << bits, (npts, 2)), bits)
    owners = _leaf_boxes(t, bits, owners=True)
    W = 1 << bits
    segs = (rng.random((400, 4)) * (W - 1e-3)).astype(np.float32)
    segs[:50, 3] = segs[:50, 1]
    segs[50:100, 2] = segs[50:100, 0]
    segs[100:150] = np.floor(segs[100:150]) + np.float32(0.5)
    segs[150:250] = np.floor(segs[150:250])              # on grid lines and corners
    segs[150:175, 3] = segs[150:175, 1]
    segs[175:200, 2] = segs[175:200, 0]
    w = oracle.cell_walk(segs, bits, t["children"], t["leaf"])
    assert (w["status"] == 0).all()
    assert _walk_violations(segs, w, owners) == []
    assert (w["counts"] > 2).mean() > 0.3


def test_cell_walk_order_check_catches_planted_mistakes():
    """The order / repeat / owner / offset checks each catch their mistake,
    planted in a copy of a correct walk."""
    rng = np.random.default_rng(9)
    t = oracle.build(rng.integers(0, 1 << 10, (400, 2)), 10)
    owners = _leaf_boxes(t, 10, owners=True)
    segs = (rng.random((60, 4)) * 1023).astype(np.float32)
    w = oracle.cell_walk(segs, 10, t["children"], t["leaf"])
    s = int(np.argmax(w["counts"]))
    assert w["counts"][s] >= 4
    o = int(w["offsets"][s])

    def planted(edit):
        v = {k: np.array(x) for k, x in w.items()}
        edit(v)
        return {str(x[1]) if isinstance(x, tuple) else x for x in _walk_violations(segs, v, owners)}

    def swap(v):
        v["cells"][[o + 1, o + 2]] = v["cells"][[o + 2, o + 1]]

    def repeat(v):
        v["cells"][o + 2] = v["cells"][o + 1]

    def octant(v):
        v["cells"][o + 1, 4] ^= 1

    def parent(v):
        v["cells"][o + 1, 3] += 1

    def offsets(v):
        v["offsets"][s + 1] -= 1

    assert "order" in planted(swap)
    assert "repeat" in planted(repeat)
    assert "owner" in planted(octant)
    assert "owner" in planted(parent)
    assert "offsets" in planted(offsets)


@pytest.mark.skipif(not oracle.ref_available(2), reason="reference checker not built here")
def test_brt_of_duplicate_keys_matches_reference():
    """The reference's BuildBinaryRadixTree on sorted keys with equal
    neighbours (pairs, triples, equal runs at the end) equals the oracle's
    node loop field by field, parents included (up to 88 bits)."""
    rng = np.random.default_rng(5)
    for t in range(400):
        mbits = int(rng.integers(8, 89))
        m = int(rng.integers(2, 300))
        k = rng.integers(0, 2**62, m, dtype=np.uint64)
        hi = None
        if mbits < 62:
            k >>= np.uint64(62 - mbits)
        elif mbits > 64:
            hi = rng.integers(0, 1 << (mbits - 64), m, dtype=np.uint64)
        for _ in range(m // 3):
            i = int(rng.integers(0, m))
            k[(i + 1) % m] = k[i]
            if hi is not None:
                hi[(i + 1) % m] = hi[i]
        order = np.lexsort((k,) if hi is None else (k, hi))
        k = k[order]
        hi = None if hi is None else hi[order]
        if t % 5 == 0:
            k[-3:] = k[-1]
            if hi is not None:
                hi[-3:] = hi[-1]
        a = oracle.brt_only(k, hi, mbits)
        b = oracle.ref_brt_only(k, hi, mbits)
        for f in a:
            assert np.array_equal(a[f], b[f]), (t, f)
