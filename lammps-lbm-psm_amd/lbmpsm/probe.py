"""Measured HBM copy ceiling (libprobe.so): streaming copies timed with HIP
events, counting the bytes read plus the bytes written."""
import ctypes as C
import os

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libprobe.so")
FORMS = {0: "8 B/thread, NT store", 1: "16 B/thread, NT store", 2: "16 B/thread, NT load+store",
         3: "4x16 B/thread grid-stride, NT store", 4: "4x16 B/thread grid-stride, NT load+store",
         5: "4x16 B/thread grid-stride, plain"}


def _lib():
    lib = C.CDLL(_PATH)
    lib.lbmpsm_copy_probe_form.restype = C.c_int
    lib.lbmpsm_copy_probe_form.argtypes = [C.c_int, C.c_int64, C.c_int, C.POINTER(C.c_double)]
    lib.lbmpsm_pull_probe.restype = C.c_int
    lib.lbmpsm_pull_probe.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.POINTER(C.c_double)]
    return lib


def copy_bandwidth(nbytes=2 << 30, reps=10, form=0):
    """GB/s of one copy form and its time per copy (ms)."""
    ms = C.c_double()
    rc = _lib().lbmpsm_copy_probe_form(int(form), int(nbytes), int(reps), C.byref(ms))
    if rc != 0:
        raise RuntimeError(f"copy probe failed ({rc})")
    return 2.0 * nbytes / (ms.value * 1e-3) / 1e9, ms.value


def copy_ceiling(nbytes=2 << 30, reps=10):
    """The best of all copy forms: (GB/s, form, {form: GB/s})."""
    res = {f: copy_bandwidth(nbytes, reps, f)[0] for f in FORMS}
    best = max(res, key=res.get)
    return res[best], best, res


def pull_pattern(sx, sy, sz, elem_bytes, reps=10):
    """Time (ms) of one pass of the D3Q19 collide-stream's access pattern
    alone (19 pulled and 19 stored elements per node, no arithmetic) over an
    sx x sy x sz box, sz a multiple of 64."""
    ms = C.c_double()
    rc = _lib().lbmpsm_pull_probe(int(sx), int(sy), int(sz), int(elem_bytes), int(reps),
                                  C.byref(ms))
    if rc != 0:
        raise RuntimeError(f"pull probe failed ({rc})")
    return ms.value
