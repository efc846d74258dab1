"""Device-rate calibration for the bench's roofline: global atomic throughput of
the patterns the pass kernels use (include/tsdf_calib.h, libtsdf_calib.so).

`python -m ros_tsdf_amd.calib` prints one JSON object with every pattern's rate;
under `rocprofv3 --pmc` the same run gives the atomic counters' reading per
lane atomic, which turns a kernel's counter value into lane atomics.
"""
import ctypes as C
import json
import os

from . import _build

PATTERNS = {"scatter_ret64": 0, "scatter_noret32": 1, "coalesced_ret32": 2, "same_ret32": 3}
# (workgroups of 256, steps per lane): ~33M lane atomics for the spread patterns (a few ms), 32k for the one
# contended word
SIZES = {"scatter_ret64": (8192, 16), "scatter_noret32": (8192, 16), "coalesced_ret32": (8192, 16),
         "same_ret32": (2048, 4)}

_lib = None


def load():
    global _lib
    if _lib is None:
        path = _build.CALIB_SO
        if not os.path.exists(path):
            raise RuntimeError("%s is not built: run __graft_entry__.build()" % path)
        L = C.CDLL(path)
        L.tsdf_calib_atomics.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int64),
                                         C.POINTER(C.c_double)]
        L.tsdf_calib_atomics.restype = C.c_int
        L.tsdf_calib_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def atomic_rate(pattern, workgroups=None, steps=None, device=0):
    """(lane atomics, seconds) of one timed launch of `pattern`."""
    L = load()
    wg, st = SIZES[pattern]
    ops, secs = C.c_int64(), C.c_double()
    r = L.tsdf_calib_atomics(device, PATTERNS[pattern], workgroups or wg, steps or st, C.byref(ops), C.byref(secs))
    if r != 0:
        raise RuntimeError("tsdf_calib_atomics(%s): %s" % (pattern, L.tsdf_calib_last_error().decode()))
    return ops.value, secs.value


def atomic_rates(device=0):
    """Every pattern: lane atomics per second, and ns per atomic for the contended word."""
    out = {}
    for name in PATTERNS:
        ops, secs = atomic_rate(name, device=device)
        out[name] = {"ops": ops, "seconds": secs, "ops_per_s": ops / secs}
    out["same_ret32"]["ns_per_op"] = 1e9 * out["same_ret32"]["seconds"] / out["same_ret32"]["ops"]
    return out


if __name__ == "__main__":
    print(json.dumps(atomic_rates()))
