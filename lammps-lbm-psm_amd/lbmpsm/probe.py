"""Measured HBM copy ceiling (libprobe.so): one streaming read and one
streaming write per element, timed with HIP events."""
import ctypes as C
import os

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libprobe.so")


def copy_bandwidth(nbytes=2 << 30, reps=10):
    """GB/s counting the bytes read plus the bytes written."""
    lib = C.CDLL(_PATH)
    lib.lbmpsm_copy_probe.restype = C.c_int
    lib.lbmpsm_copy_probe.argtypes = [C.c_int64, C.c_int, C.POINTER(C.c_double)]
    ms = C.c_double()
    rc = lib.lbmpsm_copy_probe(int(nbytes), int(reps), C.byref(ms))
    if rc != 0:
        raise RuntimeError(f"copy probe failed ({rc})")
    return 2.0 * nbytes / (ms.value * 1e-3) / 1e9, ms.value
