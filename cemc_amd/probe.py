"""Roofline probes (libcemcprobe.so): measured HBM copy and LDS read rates on
the current GPU, and a copy kernel of known traffic for scaling the HBM byte
counters."""
from __future__ import annotations

import ctypes
import os

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libcemcprobe.so")
_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(_PATH):
            raise ImportError("libcemcprobe.so not found at %s: build it with `python -m cemc_amd.build`" % _PATH)
        lib = ctypes.CDLL(_PATH)
        lib.probe_last_error.restype = ctypes.c_char_p
        lib.probe_hbm_copy.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        lib.probe_lds_read.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double)]
        _lib = lib
    return _lib


def _check(lib, rc):
    if rc:
        raise RuntimeError(lib.probe_last_error().decode())


def hbm_copy(nbytes, reps=10):
    """Mean time (ms) of one copy of `nbytes` and the rate in GB/s counting
    the bytes read plus the bytes written."""
    lib = load()
    ms = ctypes.c_double()
    _check(lib, lib.probe_hbm_copy(int(nbytes), int(reps), ctypes.byref(ms)))
    nb = (int(nbytes) // 16) * 16
    return ms.value, 2.0 * nb / (ms.value * 1e-3) / 1e9


def lds_read(blocks=256 * 8, iters=1 << 16):
    """LDS read rate in GB/s over `blocks` workgroups of 256 lanes."""
    lib = load()
    ms, nb = ctypes.c_double(), ctypes.c_double()
    _check(lib, lib.probe_lds_read(int(blocks), int(iters), ctypes.byref(ms), ctypes.byref(nb)))
    return ms.value, nb.value / (ms.value * 1e-3) / 1e9
