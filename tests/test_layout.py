"""Conversion between the library's native keys / BRT records and the
reference's BigUnsigned (C/BigUnsigned.h:8-16) and BrtNode
(C/BrtNode.h:10-21) layouts, pinned against the reference's own C code
(oracle/_ref: its structs, setBUBit/zapLeadingZeros, compareBU and the
BrtNodes its BuildBinaryRadixTree writes).  CPU only: the converters are
host functions of libpgvd.so."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import pgvd_amd
from pgvd_amd import BIG_UNSIGNED, REF_BRT_NODE, BRT_DTYPES

needs_ref = pytest.mark.skipif(not oracle.ref_available(2),
                               reason="reference checker build (oracle/_ref) not present")
M64 = (1 << 64) - 1


def value_words(vals, key_bytes):
    lo = np.array([v & M64 for v in vals], np.uint64)
    hi = np.array([v >> 64 for v in vals], np.uint64)
    return pgvd_amd.keys_to_words(lo, hi, key_bytes)


def meaningful(bu: np.ndarray) -> list:
    """len and the blocks below len: the bytes that carry a BigUnsigned's value."""
    return [(int(b["len"]), bytes(b["blk"][: max(int(b["len"]), 0)])) for b in bu]


def ref_bu(value: int, nbits: int) -> np.ndarray:
    out = np.zeros(1, BIG_UNSIGNED)
    oracle.ref(2).ref_bu_from_bits(value & M64, value >> 64, nbits, out.ctypes.data)
    return out


def edge_and_random(seed: int, count: int, bits: int) -> list[int]:
    rng = np.random.default_rng(seed)
    vals = [0, 1, 255, 256, 65535, (1 << 32) - 1, 1 << 32, M64, 1 << 64, (1 << bits) - 1,
            1 << (bits - 1)]
    vals = [v for v in vals if v < (1 << bits)]
    for _ in range(count):
        nb = int(rng.integers(1, bits + 1))
        v = int(rng.integers(0, 1 << min(nb, 63))) | (int(rng.integers(0, 1 << 25)) << 63 if nb > 63 else 0)
        vals.append(v & ((1 << nb) - 1))
    return vals


@needs_ref
def test_layout_matches_reference_structs():
    got = np.zeros(10, np.int32)
    oracle.ref(2).ref_layout(got)
    # sizeof(BigUnsigned), offsetof len, blk; sizeof(BrtNode), offsetof left,
    # left_leaf, right_leaf, lcp, lcp_length, parent
    want = [BIG_UNSIGNED.itemsize, BIG_UNSIGNED.fields["len"][1], BIG_UNSIGNED.fields["blk"][1],
            REF_BRT_NODE.itemsize] + [REF_BRT_NODE.fields[f][1] for f in
                                      ("left", "left_leaf", "right_leaf", "lcp", "lcp_length",
                                       "parent")]
    assert got.tolist() == want


@needs_ref
@pytest.mark.parametrize("key_bytes,bits", [(4, 32), (8, 64), (16, 88), (16, 87)])
def test_keys_to_big_unsigned_equals_reference_bits(key_bytes, bits):
    """Values built bit by bit with the reference's setBUBit and
    zapLeadingZeros (as tests/KernelTests.cpp:106-112 builds them) equal the
    converter's output in every byte that carries the value."""
    vals = edge_and_random(key_bytes * 7 + bits, 3000, bits)
    got = pgvd_amd.to_big_unsigned(value_words(vals, key_bytes), key_bytes)
    want = np.concatenate([ref_bu(v, 88) for v in vals])
    assert meaningful(got) == meaningful(want)
    # unused blocks and padding are written as zero
    for b in got:
        assert not b["blk"][int(b["len"]):].any() and b["pad"] == 0


@pytest.mark.parametrize("key_bytes,bits", [(4, 32), (8, 63), (8, 64), (16, 87), (16, 88)])
def test_round_trip(key_bytes, bits):
    vals = edge_and_random(bits, 5000, bits)
    words = value_words(vals, key_bytes)
    bu = pgvd_amd.to_big_unsigned(words, key_bytes)
    back = pgvd_amd.from_big_unsigned(bu, key_bytes)
    np.testing.assert_array_equal(back, words)
    # a value carries exactly its significant blocks
    for v, b in zip(vals, bu):
        assert int(b["len"]) == (v.bit_length() + 7) // 8


@needs_ref
def test_compare_matches_reference_compareBU():
    L, R = pgvd_amd.lib(), oracle.ref(2)
    vals = edge_and_random(5, 600, 88)
    rng = np.random.default_rng(6)
    bu = pgvd_amd.to_big_unsigned(value_words(vals, 16), 16)
    for _ in range(4000):
        i, j = (int(x) for x in rng.integers(0, len(vals), 2))
        if rng.random() < 0.1:
            j = i
        a, b = bu[i:i + 1], bu[j:j + 1]
        ours = L.pgvd_compare_big_unsigned(a.ctypes.data, b.ctypes.data)
        theirs = R.ref_compare_bu(a.ctypes.data, b.ctypes.data)
        numeric = (vals[i] > vals[j]) - (vals[i] < vals[j])
        assert ours == theirs == numeric, (vals[i], vals[j])


@needs_ref
@pytest.mark.parametrize("mbits,seed", [(87, 1), (63, 2), (32, 3)])
def test_brt_nodes_equal_reference_records(mbits, seed):
    """The BrtNodes the reference's BuildBinaryRadixTree writes (lcp as its
    compute_lcp leaves it) equal the converter applied to the oracle's BRT,
    field by field and in the lcp's value bytes; converting the reference's
    records back gives the native records."""
    rng = np.random.default_rng(seed)
    m = 3000
    hi_bits = max(mbits - 64, 0)
    lo = rng.integers(0, 1 << min(mbits, 63), m, dtype=np.uint64)
    if mbits > 63:
        lo |= rng.integers(0, 2, m, dtype=np.uint64) << np.uint64(63)
    hi = (rng.integers(0, 1 << hi_bits, m, dtype=np.uint64) if hi_bits else np.zeros(m, np.uint64))
    keys = sorted(set(zip(hi.tolist(), lo.tolist())))
    hi = np.array([k[0] for k in keys], np.uint64)
    lo = np.array([k[1] for k in keys], np.uint64)
    m = lo.size
    R = oracle.ref(2)
    h = R.ref_build_from_keys(lo, hi, m, mbits)
    try:
        raw = np.zeros(m - 1, REF_BRT_NODE)
        R.ref_tree_brt_raw(h, raw.ctypes.data)
        rkeys = np.zeros(m, BIG_UNSIGNED)
        R.ref_tree_keys_raw(h, rkeys.ctypes.data)
    finally:
        R.ref_tree_free(h)
    want = oracle.build_from_keys(lo, hi, 2, mbits)
    kb = pgvd_amd.key_bytes_for(mbits)
    rec = np.zeros(m - 1, BRT_DTYPES[kb])
    if kb == 16:
        rec["lcp_lo"], rec["lcp_hi"] = want["lcp_lo"], want["lcp_hi"]
    else:
        rec["lcp"] = want["lcp_lo"].astype(rec["lcp"].dtype)
    for f in ("left", "left_leaf", "right_leaf"):
        rec[f] = want[f]
    rec["lcp_length"] = want["lcp_length"]
    got = pgvd_amd.to_ref_brt_nodes(rec, want["parent"], kb)
    for f in ("left", "left_leaf", "right_leaf", "lcp_length", "parent"):
        np.testing.assert_array_equal(got[f], raw[f], err_msg=f)
    assert meaningful(got["lcp"]) == meaningful(raw["lcp"])
    # keys the reference holds as BigUnsigned equal the converted native keys
    words = pgvd_amd.keys_to_words(lo, hi, kb)
    assert meaningful(pgvd_amd.to_big_unsigned(words, kb)) == meaningful(rkeys)
    # and back: reference records -> native records
    back = np.zeros(m - 1, BRT_DTYPES[kb])
    par = np.zeros(m - 1, np.int32)
    rc = pgvd_amd.lib().pgvd_ref_nodes_to_brt(np.ascontiguousarray(raw).ctypes.data, m - 1, kb,
                                               back.ctypes.data, par.ctypes.data)
    assert rc == 0
    np.testing.assert_array_equal(back, rec)
    np.testing.assert_array_equal(par, want["parent"])


def test_conversion_errors():
    L = pgvd_amd.lib()
    big = value_words([1 << 88], 16)
    out = np.zeros(1, BIG_UNSIGNED)
    assert L.pgvd_keys_to_big_unsigned(big.ctypes.data, 16, 1, out.ctypes.data) == 8
    bad = np.zeros(1, BIG_UNSIGNED)
    bad["len"] = 12
    k = np.zeros(2, np.uint64)
    assert L.pgvd_big_unsigned_to_keys(bad.ctypes.data, 1, 16, k.ctypes.data) == 1
    wide = pgvd_amd.to_big_unsigned(value_words([1 << 40], 16), 16)
    k4 = np.zeros(1, np.uint32)
    assert L.pgvd_big_unsigned_to_keys(wide.ctypes.data, 1, 4, k4.ctypes.data) == 8
    assert L.pgvd_keys_to_big_unsigned(k4.ctypes.data, 5, 1, out.ctypes.data) == 1
