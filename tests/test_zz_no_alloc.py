"""The allocation-free entry point ipr_hip_equal_noalloc against the entry
point it varies, ipr_hip_equal, through the C-ABI of libipregel_hip.so: the
same answers on two calls in a row, the second on a smaller input (the scratch
flag keeps the first call's contents, so it must be cleared), on equal and
unequal buffers, and its error for a size that is not a multiple of 4."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np
import pytest

LIB = Path(__file__).resolve().parents[1] / "ipregel_amd" / "lib" / "libipregel_hip.so"


def hip():
    L = C.CDLL(str(LIB))
    L.ipr_hip_init.argtypes = [C.c_int, C.c_char_p, C.c_size_t, C.POINTER(C.c_uint64)]
    L.ipr_hip_malloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.ipr_hip_free.argtypes = [C.c_void_p]
    L.ipr_hip_memcpy_h2d.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
    L.ipr_hip_stream_create.argtypes = [C.POINTER(C.c_void_p)]
    L.ipr_hip_stream_sync.argtypes = [C.c_void_p]
    L.ipr_hip_stream_destroy.argtypes = [C.c_void_p]
    for f in ("ipr_hip_equal", "ipr_hip_equal_noalloc"):
        getattr(L, f).argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(C.c_int), C.c_void_p]
    L.ipr_hip_last_error.restype = C.c_char_p
    name, hbm = C.create_string_buffer(128), C.c_uint64()
    assert L.ipr_hip_init(0, name, 128, C.byref(hbm)) == 0, L.ipr_hip_last_error()
    return L


@pytest.mark.gpu
def test_equal_noalloc_matches_equal_on_two_calls():
    L = hip()
    st = C.c_void_p()
    assert L.ipr_hip_stream_create(C.byref(st)) == 0
    n = (1 << 22) + 13                       # words: 16 MiB and a ragged tail
    rng = np.random.default_rng(7)
    x = rng.integers(0, 2**32, n, dtype=np.uint32)
    bufs = []
    try:
        for _ in range(3):
            p = C.c_void_p()
            assert L.ipr_hip_malloc(C.byref(p), n * 4) == 0
            bufs.append(p)
        a, b, c = bufs
        y = x.copy()
        y[n - 1] ^= 1                        # c differs from a in its last word only
        for p, h in ((a, x), (b, x), (c, y)):
            assert L.ipr_hip_memcpy_h2d(p, h.ctypes.data, n * 4, st) == 0
        assert L.ipr_hip_stream_sync(st) == 0

        def both(p, q, nbytes):
            r = []
            for f in (L.ipr_hip_equal, L.ipr_hip_equal_noalloc):
                eq = C.c_int(-1)
                assert f(p, q, nbytes, C.byref(eq), st) == 0, L.ipr_hip_last_error()
                r.append(eq.value)
            return r

        print("equal_noalloc: full size, unequal, then equal", flush=True)
        assert both(a, c, n * 4) == [0, 0]
        assert both(a, b, n * 4) == [1, 1]
        # the second call of a pair on a smaller input: the flag the unequal
        # full-size call raised must not leak into it
        print("equal_noalloc: unequal full size, then a smaller equal prefix", flush=True)
        assert both(a, c, n * 4) == [0, 0]
        assert both(a, c, (n - 1) * 4) == [1, 1]
        assert both(a, c, 4) == [1, 1]
        assert both(a, c, 0) == [1, 1]
        eq = C.c_int(-1)
        assert L.ipr_hip_equal_noalloc(a, b, 6, C.byref(eq), st) != 0
        assert b"multiple of 4" in L.ipr_hip_last_error()
    finally:
        for p in bufs:
            L.ipr_hip_free(p)
        L.ipr_hip_stream_destroy(st)
