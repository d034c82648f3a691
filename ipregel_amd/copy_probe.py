"""Device-to-device copy of a known size through libipregel_hip.so, used to
calibrate the FETCH_SIZE / WRITE_SIZE counters of rocprofv3 (bytes counted
per byte copied) before they are read as HBM traffic of the engine's kernels.

    python -m ipregel_amd.copy_probe [MiB]
"""
from __future__ import annotations

import ctypes as C
import sys
from pathlib import Path

LIB = Path(__file__).resolve().parent / "lib" / "libipregel_hip.so"


def main(mib: int = 1024) -> None:
    L = C.CDLL(str(LIB))
    L.ipr_hip_init.argtypes = [C.c_int, C.c_char_p, C.c_size_t, C.POINTER(C.c_uint64)]
    L.ipr_hip_malloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.ipr_hip_memcpy_d2d.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]
    L.ipr_hip_fill.argtypes = [C.c_void_p, C.c_int, C.c_size_t, C.c_void_p]
    L.ipr_hip_stream_create.argtypes = [C.POINTER(C.c_void_p)]
    L.ipr_hip_stream_sync.argtypes = [C.c_void_p]
    L.ipr_hip_free.argtypes = [C.c_void_p]
    name = C.create_string_buffer(128)
    hbm = C.c_uint64()
    if L.ipr_hip_init(0, name, 128, C.byref(hbm)):
        raise SystemExit("no device")
    n = mib << 20
    a, b, st = C.c_void_p(), C.c_void_p(), C.c_void_p()
    for p in (a, b):
        if L.ipr_hip_malloc(C.byref(p), n):
            raise SystemExit("allocation failed")
    L.ipr_hip_stream_create(C.byref(st))
    L.ipr_hip_fill(a, 1, n, st)
    L.ipr_hip_memcpy_d2d(b, a, n, st)
    L.ipr_hip_stream_sync(st)
    L.ipr_hip_free(a)
    L.ipr_hip_free(b)
    print(f"copied {n} bytes on {name.value.decode()}")


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 1024)
