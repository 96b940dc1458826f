"""ctypes view of libkiwi_io.so (include/kiwi_io.h): KIWI's SPH snapshot readers on the host.

    from kiwi_amd import io
    snap = io.load_snapshot("snap.txt")            # dict of numpy arrays: id x y z vx vy vz m h rho
    io.detect_format("snap.txt")                   # "plain_ascii" | "splash_ascii" | "phantom_binary"

Errors raise SnapshotError with the reference reader's message.  Raises if the library is not built
(python -m kiwi_amd.build).
"""
import ctypes
import os

import numpy as np

from . import IO_LIB_PATH

FIELDS = ("x", "y", "z", "vx", "vy", "vz", "m", "h", "rho")
_pd = ctypes.POINTER(ctypes.c_double)


class Snapshot(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("id", ctypes.POINTER(ctypes.c_int))] + [(f, _pd) for f in FIELDS]


class SnapshotError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(IO_LIB_PATH):
            raise RuntimeError(f"libkiwi_io.so not built at {IO_LIB_PATH}: run python -m kiwi_amd.build")
        L = ctypes.CDLL(IO_LIB_PATH)
        L.kiwi_io_last_error.restype = ctypes.c_char_p
        L.kiwi_io_detect_format.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.kiwi_io_load_snapshot.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_double, ctypes.POINTER(Snapshot)]
        L.kiwi_io_load_snapshot_ex.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_double,
                                               ctypes.POINTER(ctypes.c_double * 3), ctypes.c_int, ctypes.POINTER(Snapshot)]
        L.kiwi_io_free_snapshot.argtypes = [ctypes.POINTER(Snapshot)]
        L.kiwi_io_set_threads.argtypes = [ctypes.c_int]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise SnapshotError(rc, load().kiwi_io_last_error().decode(errors="replace"))


def set_threads(n):
    return load().kiwi_io_set_threads(int(n))


def detect_format(path):
    buf = ctypes.create_string_buffer(32)
    _check(load().kiwi_io_detect_format(os.fsencode(path), buf, 32))
    return buf.value.decode()


def load_snapshot(path, format="auto", hfact=1.2, units=None, allow_missing_velocity=False):
    """units: (length_cm, mass_g, velocity_cms) of a phantom_binary read, None for the defaults."""
    s = Snapshot()
    u = None if units is None else ctypes.byref((ctypes.c_double * 3)(*[float(v) for v in units]))
    _check(load().kiwi_io_load_snapshot_ex(os.fsencode(path), format.encode(), float(hfact), u,
                                           int(bool(allow_missing_velocity)), ctypes.byref(s)))
    try:
        n = s.n
        out = {"id": np.ctypeslib.as_array(s.id, (n,)).copy() if n else np.zeros(0, np.int32)}
        for f in FIELDS:
            out[f] = np.ctypeslib.as_array(getattr(s, f), (n,)).copy() if n else np.zeros(0)
        return out
    finally:
        load().kiwi_io_free_snapshot(ctypes.byref(s))
