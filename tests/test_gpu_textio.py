"""The octree's text and binary forms on the GPU against the oracle's
operator<< / operator>> restatement (C/OctNode.h:58-98): byte-equal text,
equal records back, the reference's whitespace and leaf rules, and errors."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
import pgvd_amd as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = P.Device(0)
    yield d
    d.close()


def _records(O, dim, seed):
    """Random records with every decimal length and both signs."""
    rng = np.random.default_rng(seed)
    mag = rng.integers(0, 32, (O, (1 << dim) + 1))
    v = (rng.integers(0, 1 << 31, mag.shape, dtype=np.int64) >> (31 - mag)).astype(np.int64)
    v = np.where(rng.random(mag.shape) < 0.4, -v - 1, v)
    v[rng.random(mag.shape) < 0.3] = -1
    flat = v.reshape(-1)
    flat[: min(4, flat.size)] = [-2147483648, 2147483647, 0, -1][: min(4, flat.size)]
    return v.astype(np.int32)


@pytest.mark.parametrize("O,dim", [(1, 2), (255, 3), (256, 2), (257, 3), (100000, 3), (70001, 2)])
def test_format_matches_oracle(dev, O, dim):
    rec = _records(O, dim, seed=O + dim)
    d = dev.upload(rec)
    try:
        got = dev.format_octree_text(d, O, dim)
    finally:
        d.free()
    assert got == oracle.octree_format_text(rec, dim)


@pytest.mark.parametrize("dim,bits,n", [(2, 16, 65536), (3, 21, 1 << 20)])
def test_built_octree_text_round_trip(dev, dim, bits, n):
    rng = np.random.default_rng(dim)
    pts = (rng.random((n, dim), dtype=np.float32) * 2 - 1)
    tree = dev.build_octree(pts, bits)
    rec = np.concatenate([tree["children"], tree["leaf"][:, None]], axis=1).astype(np.int32)
    d = dev.upload(rec)
    try:
        text = dev.format_octree_text(d, rec.shape[0], dim)
    finally:
        d.free()
    assert text == oracle.octree_format_text(rec, dim)
    assert np.array_equal(dev.parse_octree_text(text, dim), rec)


def test_parse_whitespace_signs_leaf_and_trailing_tokens(dev):
    """Any whitespace run separates tokens, '+' is accepted, the leaf is read
    as unsigned char, and tokens after the last record are ignored; equal to
    the oracle's operator>> on the same bytes."""
    rng = np.random.default_rng(5)
    for dim, O in ((2, 3000), (3, 5000)):
        rec = _records(O, dim, seed=dim)
        rec[:, -1] = rng.integers(0, 100000, O)        # leaf values past 255
        toks = [str(O)] + [("+" if (x >= 0 and rng.random() < 0.2) else "") + str(x)
                           for x in rec.reshape(-1)]
        seps = rng.choice([" ", "\n", "\t", "  ", " \r\n", "\v", "\f"], len(toks))
        text = ("  \n" + "".join(t + s for t, s in zip(toks, seps)) + "7 8 9").encode()
        want = oracle.octree_parse_text(text, dim)
        assert want is not None and want.shape == rec.shape
        assert np.array_equal(want[:, -1], rec[:, -1] % 256)
        assert np.array_equal(dev.parse_octree_text(text, dim), want)


def test_parse_leading_zeros(dev):
    """Zero-padded tokens (0 to 300 zeros, so many run past the bytes a
    round stages and across round and tile edges) read as operator>> reads
    them, a zero-padded and a "-0" node count, and padded tokens past the
    int32 limits rejected; equal to the oracle's operator>>."""
    rng = np.random.default_rng(21)
    for dim, O in ((2, 4000), (3, 3000)):
        rec = _records(O, dim, seed=dim + 40)
        toks = []
        for x in rec.reshape(-1):
            z = "0" * int(rng.choice([0, 1, 5, 12, 40, 300], p=[.5, .15, .1, .1, .1, .05]))
            toks.append(("-" if x < 0 else "") + z + str(abs(int(x))))
        text = ("000" + str(O) + " " + " ".join(toks) + " ").encode()
        want = oracle.octree_parse_text(text, dim)
        assert want is not None and np.array_equal(want[:, :-1], rec[:, :-1])
        assert np.array_equal(dev.parse_octree_text(text, dim), want)
    assert dev.parse_octree_text(b"-0 ", 2).shape[0] == 0
    for bad in (b"1 0 0 0 0 " + b"0" * 100 + b"2147483648 ", b"1 0 -" + b"0" * 20 + b"2147483649 0 0 0 "):
        assert oracle.octree_parse_text(bad, 2) is None
        with pytest.raises(P.PgvdError):
            dev.parse_octree_text(bad, 2)


def test_parse_misaligned_text(dev):
    """A text that starts at an odd device address takes the byte loads."""
    rec = _records(4000, 3, seed=9)
    text = oracle.octree_format_text(rec, 3)
    buf = dev.upload(np.frombuffer(b"x" + text, np.uint8))
    out = dev.alloc(rec.nbytes)
    O = C.c_int64()
    try:
        rc = P.lib().pgvd_octree_parse_text(dev.ctx, buf.ptr + 1, len(text), 3, out.ptr, 4000,
                                            C.byref(O), None)
        assert rc == 0 and O.value == 4000
        want = oracle.octree_parse_text(text, 3)            # the leaf read as unsigned char
        assert np.array_equal(want[:, :-1], rec[:, :-1])
        assert np.array_equal(dev.download(out, np.int32, rec.size).reshape(rec.shape), want)
    finally:
        buf.free()
        out.free()


@pytest.mark.parametrize("bad", [b"2 1 -1 -1 -1 14 -1 -1 -1 -1", b"1 1 x -1 -1 14",
                                 b"1 2147483648 0 0 0 0", b"1 - 0 0 0 0", b"-1 ", b"   ",
                                 b"1 0 0 0 0 99999999999"])
def test_parse_rejects_what_the_oracle_rejects(dev, bad):
    assert oracle.octree_parse_text(bad, 2) is None
    with pytest.raises(P.PgvdError):
        dev.parse_octree_text(bad, 2)


def test_parse_capacity_reports_count(dev):
    text = oracle.octree_format_text(_records(10, 2, seed=1), 2)
    buf = dev.upload(np.frombuffer(text, np.uint8))
    out = dev.alloc(4 * 5 * 4)
    O = C.c_int64()
    try:
        rc = P.lib().pgvd_octree_parse_text(dev.ctx, buf.ptr, len(text), 2, out.ptr, 4,
                                            C.byref(O), None)
        assert rc == 7 and O.value == 10          # PGVD_ERR_CAPACITY, count reported
    finally:
        buf.free()
        out.free()


@pytest.mark.parametrize("binary", [False, True])
def test_save_load_files(dev, tmp_path, binary):
    rng = np.random.default_rng(3)
    pts = rng.random((50000, 3), dtype=np.float32)
    dev.build_octree_async(dev.upload(pts), 50000, 3, 21)
    res = dev.build_octree_finish()
    O = int(res.octree_size)
    rec = dev.download(res.octree, np.int32, O * 9).reshape(O, 9)
    path = str(tmp_path / ("o.bin" if binary else "o.txt"))
    dev.save_octree(res.octree, O, 3, path, binary=binary)
    raw = open(path, "rb").read()
    if binary:
        assert raw[:8] == b"PGVDOCT1" and len(raw) == 32 + rec.nbytes
    else:
        assert raw == oracle.octree_format_text(rec, 3)
    assert np.array_equal(dev.load_octree(path, 3), rec)
    if binary:
        with pytest.raises(P.PgvdError):
            dev.load_octree(path, 2)                # another dimension's records


def test_empty_octree_text(dev):
    assert dev.format_octree_text(0, 0, 2) == b"0 "
    assert dev.parse_octree_text(b"0 ", 3).shape == (0, 9)
