"""The C ABI on the CPU: the library loads, exports every function the
header declares, the header compiles as C and C++, record layouts match,
and argument errors come back as status codes without touching a GPU."""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

import pgvd_amd as P

HEADER = os.path.join(P.ROOT, "include", "pgvd.h")


def test_library_exports_every_declared_symbol():
    L = P.lib()
    declared = P.exported_symbols()
    assert len(declared) >= 20
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", P.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    defined = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(declared) <= defined


def test_library_is_gfx950_hip_code():
    out = subprocess.run(["strings", P.LIB_PATH], capture_output=True, text=True).stdout
    assert "gfx950" in out


@pytest.mark.parametrize("compiler,std", [("gcc", "-std=c99"), ("g++", "-std=c++11")])
def test_header_compiles_and_layouts(compiler, std):
    src = r"""
#include <stddef.h>
#include <stdio.h>
#include "pgvd.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\n", sizeof(pgvd_brt_node32), sizeof(pgvd_brt_node64),
         sizeof(pgvd_brt_node128), offsetof(pgvd_brt_node128, left), sizeof(pgvd_octree_result));
  return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        ext = ".c" if compiler == "gcc" else ".cpp"
        path = os.path.join(d, "t" + ext)
        with open(path, "w") as f:
            f.write(src)
        exe = os.path.join(d, "t")
        subprocess.run([compiler, std, "-Wall", "-Werror", "-I", os.path.dirname(HEADER), path,
                        "-o", exe], check=True)
        sizes = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    assert [int(s) for s in sizes[:4]] == [P.BRT_DTYPES[4].itemsize, P.BRT_DTYPES[8].itemsize,
                                           P.BRT_DTYPES[16].itemsize,
                                           P.BRT_DTYPES[16].fields["left"][1]]
    assert int(sizes[4]) == C.sizeof(P.OctreeResult)


def test_argument_errors_need_no_device():
    L = P.lib()
    assert L.pgvd_next_pow2(5) == 8 and L.pgvd_next_pow2(1000) == 1024
    assert L.pgvd_ctx_create(0, None) == 1                       # PGVD_ERR_INVALID
    assert L.pgvd_build_octree(None, None, 10, 3, 21, None, None, None) == 1
    assert L.pgvd_radix_sort(None, None, None, 10, 8, 63, None) == 1
    assert b"null" in L.pgvd_last_error()


def test_no_device_reports_status():
    """Without a visible GPU, context creation fails with PGVD_ERR_NO_DEVICE
    instead of falling back to anything."""
    import torch  # noqa: F401  (only to learn whether a GPU is visible here)
    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is visible")
    ctx = C.c_void_p()
    assert P.lib().pgvd_ctx_create(0, C.byref(ctx)) == 4
    with pytest.raises(P.PgvdError):
        P.Device(0)


def test_key_word_layout_roundtrip():
    rng = np.random.default_rng(1)
    lo = rng.integers(0, 2**63, 100, dtype=np.uint64)
    hi = rng.integers(0, 2**62, 100, dtype=np.uint64)
    for kb in (4, 8, 16):
        w = P.keys_to_words(lo, hi, kb)
        a, b = P.words_to_keys(w, kb)
        if kb == 16:
            assert np.array_equal(a, lo) and np.array_equal(b, hi)
        elif kb == 8:
            assert np.array_equal(a, lo)
        else:
            assert np.array_equal(a, lo & np.uint64(0xFFFFFFFF))
