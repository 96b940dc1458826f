"""ctypes view of libkiwi_io.so (include/kiwi_io.h): KIWI's SPH snapshot readers, run configuration and
provenance manifest on the host.

    from kiwi_amd import io
    snap = io.load_snapshot("snap.txt")            # dict of numpy arrays: id x y z vx vy vz m h rho
    io.detect_format("snap.txt")                   # "plain_ascii" | "splash_ascii" | "phantom_binary"
    cfg = io.load_config("run.json")               # dict: the Config with its derived values resolved
    io.write_manifest("run.json")                  # <run_dir>/manifest.json as the reference writes it

Errors raise SnapshotError / ConfigError with the reference's message.  Raises if the library is not
built (python -m kiwi_amd.build).
"""
import ctypes
import os

import numpy as np

from . import IO_LIB_PATH

FIELDS = ("x", "y", "z", "vx", "vy", "vz", "m", "h", "rho")
_pd = ctypes.POINTER(ctypes.c_double)


class Snapshot(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("id", ctypes.POINTER(ctypes.c_int))] + [(f, _pd) for f in FIELDS]


class PhantomUnits(ctypes.Structure):
    _fields_ = [("length_cm", ctypes.c_double), ("mass_g", ctypes.c_double), ("velocity_cms", ctypes.c_double)]


class Config(ctypes.Structure):
    _fields_ = [("run_name", ctypes.c_char * 256), ("input_snapshot", ctypes.c_char * 4096),
                ("input_snapshot_format", ctypes.c_char * 32), ("work_folder", ctypes.c_char * 4096),
                ("run_dir", ctypes.c_char * 4096), ("domain", ctypes.c_double * 6), ("kernel_id", ctypes.c_int),
                ("support_in_h", ctypes.c_double), ("hfact", ctypes.c_double), ("generic_path", ctypes.c_int),
                ("grid_nx", ctypes.c_int), ("grid_ny", ctypes.c_int), ("grid_nz", ctypes.c_int),
                ("grid_tabulated", ctypes.c_int), ("mapping_tabulated", ctypes.c_int),
                ("buffer_enabled", ctypes.c_int), ("buffer_strategy", ctypes.c_char * 64),
                ("coverage_spacing_fraction", ctypes.c_double), ("coverage_radius_factor", ctypes.c_double),
                ("lloyd_iterations", ctypes.c_int), ("cartesian_map_resolution", ctypes.c_int),
                ("block_scale_factor", ctypes.c_double), ("phantom_units", PhantomUnits),
                ("allow_missing_velocity", ctypes.c_int)]


class ConfigError(RuntimeError):
    pass


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
        L.kiwi_config_last_error.restype = ctypes.c_char_p
        L.kiwi_config_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(Config)]
        L.kiwi_config_json.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int64,
                                       ctypes.POINTER(ctypes.c_int64)]
        L.kiwi_manifest_write.argtypes = [ctypes.c_char_p]
        L.kiwi_manifest_check.argtypes = [ctypes.c_char_p]
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


def _config_check(rc):
    if rc != 0:
        raise ConfigError(load().kiwi_config_last_error().decode(errors="replace"))


def load_config(path):
    """Config::load_from_file with the derived values resolved.  Raises ConfigError with the text the
    reference prints after "[kiwi] error: "."""
    c = Config()
    _config_check(load().kiwi_config_load(os.fsencode(path), ctypes.byref(c)))
    out = {}
    for f, _ in Config._fields_:
        v = getattr(c, f)
        if isinstance(v, bytes):
            v = v.decode()
        elif f == "domain":
            v = list(v)
        elif f == "phantom_units":
            v = (v.length_cm, v.mass_g, v.velocity_cms)
        out[f] = v
    return out


def config_json(path):
    """Config::to_json() as compact JSON text (the manifest's "config" section)."""
    n = ctypes.c_int64()
    _config_check(load().kiwi_config_json(os.fsencode(path), None, 0, ctypes.byref(n)))
    buf = ctypes.create_string_buffer(n.value + 1)
    _config_check(load().kiwi_config_json(os.fsencode(path), buf, n.value + 1, ctypes.byref(n)))
    return buf.value.decode()


def write_manifest(config_path):
    _config_check(load().kiwi_manifest_write(os.fsencode(config_path)))


def check_manifest(config_path):
    """1 when the run directory's manifest records a different config (warning on stderr), else 0."""
    rc = load().kiwi_manifest_check(os.fsencode(config_path))
    if rc < 0:
        raise ConfigError(load().kiwi_config_last_error().decode(errors="replace"))
    return rc
