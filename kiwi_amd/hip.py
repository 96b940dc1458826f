"""ctypes view of libkiwi_hip (include/kiwi_hip.h).

This mirrors the reference's stage interface for the mapping loop: `sph_to_grid` takes the particle
columns of SPHSnapshot (include/kiwi/snapshot.hpp:70-77), the domain and resolved resolution
(src/cartesian_grid.cpp:89-97) and the kernel settings (include/kiwi/kernel_integrator.hpp:118-122),
and returns the un-normalised per-cell sums; `normalise` then applies src/cartesian_grid.cpp:294-318.
`voronoi_map` does the same over a flattened Voronoi tessellation (the particle loop of
run_mapping_stage, src/mapping.cpp:400-460), and `normalise_voronoi` applies src/mapping.cpp:462-506.
Errors from the library raise KiwiError carrying the library's message (the reference throws
std::runtime_error / ConfigError at the same points).

There is no fallback: if libkiwi_hip.so is missing or no GPU is present, loading raises.
"""
import ctypes
import os

import numpy as np

from . import LIB_PATH

FP64, MIXED = 0, 1
FP64_FALLBACK_CELLS = 16.0  # KIWI_FP64_FALLBACK_CELLS
KERNELS = ["M4", "M5", "M6", "WC2", "WC4", "WC6"]
ALIASES = {"cubic": "M4", "cubic_spline": "M4", "m4": "M4", "quartic": "M5", "quartic_spline": "M5", "m5": "M5",
           "quintic": "M6", "quintic_spline": "M6", "m6": "M6", "wendland_c2": "WC2", "wendlandc2": "WC2",
           "wc2": "WC2", "wendland_c4": "WC4", "wendlandc4": "WC4", "wc4": "WC4", "wendland_c6": "WC6",
           "wendlandc6": "WC6", "wc6": "WC6"}

_i64 = ctypes.c_int64
_d = ctypes.c_double
_i = ctypes.c_int
_pd = ctypes.POINTER(ctypes.c_double)

EXPORTED = ["kiwi_hip_abi_version", "kiwi_hip_create", "kiwi_hip_destroy", "kiwi_hip_last_error",
            "kiwi_hip_stream", "kiwi_hip_sph_to_grid", "kiwi_hip_voronoi_map", "kiwi_hip_sph_to_grid_adjoint",
            "kiwi_hip_sph_to_grid_batch", "kiwi_hip_set_tuning", "kiwi_hip_get_tuning", "kiwi_hip_autotune",
            "kiwi_hip_inverse_mapping", "kiwi_hip_cartesian_map", "kiwi_hip_cartesian_lattice",
            "kiwi_hip_coverage_buffer", "kiwi_hip_voronoi_centroids", "kiwi_hip_sph_to_grid_noalloc",
            "kiwi_hip_set_fp64_fallback", "kiwi_hip_get_fp64_fallback", "kiwi_hip_set_graphs",
            "kiwi_hip_get_graphs", "kiwi_hip_voronoi_tessellate", "kiwi_hip_voronoi_tessellation_fetch",
            "kiwi_hip_voronoi_home_cells"]


class KiwiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[kiwi] error: {msg} (code {code})")
        self.code = code


class Particles(ctypes.Structure):
    _fields_ = [("n", _i64), ("x", _pd), ("y", _pd), ("z", _pd), ("vx", _pd), ("vy", _pd), ("vz", _pd),
                ("m", _pd), ("h", _pd), ("on_device", _i)]


class Grid(ctypes.Structure):
    _fields_ = [("x_min", _d), ("x_max", _d), ("y_min", _d), ("y_max", _d), ("z_min", _d), ("z_max", _d),
                ("nx", _i), ("ny", _i), ("nz", _i)]


class Kernel(ctypes.Structure):
    _fields_ = [("kernel_id", _i), ("support_in_h", _d), ("generic_path", _i), ("tabulated", _i)]


class CellSums(ctypes.Structure):
    _fields_ = [("mass", _pd), ("px", _pd), ("py", _pd), ("pz", _pd), ("on_device", _i)]


class WarnRecord(ctypes.Structure):
    _fields_ = [("particle", _i64), ("cell", _i64), ("value", _d)]


class Stats(ctypes.Structure):
    _fields_ = [("skipped_outside", _i64), ("n_negative", _i64), ("n_negative_warn", _i64), ("pairs_bbox", _i64),
                ("pairs_sphere", _i64), ("particle_blocks", _i64), ("v_mean", _d * 3),
                ("device_ms_prep", _d), ("device_ms_lines", _d), ("device_ms_gather", _d), ("device_ms_accum", _d),
                ("device_ms_lines_kernel", _d),
                ("lattice_points", _i64 * 3), ("line_evals", _i64 * 3),
                ("plane_evals", _i64), ("fp64_fallback_blocks", _i64), ("fp64_fallback_pairs", _i64),
                ("fp64_fallback_points", _i64 * 3), ("device_ms_gather_fallback", _d), ("device_ms_gather_kernel", _d),
                ("n_warn_records", _i),
                ("warn", WarnRecord * 16)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_
             if f not in ("v_mean", "warn", "lattice_points", "line_evals", "fp64_fallback_points")}
        d["fp64_fallback_points"] = list(self.fp64_fallback_points)
        d["v_mean"] = list(self.v_mean)
        d["lattice_points"] = list(self.lattice_points)
        d["line_evals"] = list(self.line_evals)
        d["warn"] = [(w.particle, w.cell, w.value) for w in self.warn[: self.n_warn_records]]
        return d


_pll = ctypes.POINTER(ctypes.c_int64)
_pi = ctypes.POINTER(ctypes.c_int)


class BatchProblem(ctypes.Structure):
    _fields_ = [("particles", Particles), ("grid", Grid), ("kernel", Kernel), ("precision", _i), ("out", CellSums),
                ("stats", ctypes.POINTER(Stats))]


class InverseRows(ctypes.Structure):
    _fields_ = [("m_recovered", _pd), ("coverage_fraction", _pd), ("on_device", _i)]


class Sites(ctypes.Structure):
    _fields_ = [("n", _i64), ("x", _pd), ("y", _pd), ("z", _pd), ("id", _pi), ("density", _pd), ("vx", _pd),
                ("vy", _pd), ("vz", _pd), ("mapped", _pi), ("on_device", _i)]


class CartesianMapOut(ctypes.Structure):
    _fields_ = [("density", _pd), ("vx", _pd), ("vy", _pd), ("vz", _pd), ("cell_id", _pi), ("on_device", _i)]


class CartesianMapStats(ctypes.Structure):
    _fields_ = [("unresolved", _i64), ("unmapped", _i64), ("device_ms_prep", _d), ("device_ms_query", _d)]


class Lattice(ctypes.Structure):
    _fields_ = [("nx", _i), ("ny", _i), ("nz", _i), ("spacing", _d)]


class BufferSites(ctypes.Structure):
    _fields_ = [("x", _pd), ("y", _pd), ("z", _pd), ("id", _pi), ("capacity", _i64), ("on_device", _i)]


class CoverageStats(ctypes.Structure):
    _fields_ = [("candidates", _i64), ("kept", _i64), ("h_max", _d), ("particle_columns", _i64), ("device_ms", _d)]


class Tuning(ctypes.Structure):
    _fields_ = [("item_grab", _i), ("cursors", _i)]


class VoronoiCells(ctypes.Structure):
    _fields_ = [("n_cells", _i64), ("n_faces", _i64), ("n_face_vertices", _i64), ("cell_face_ptr", _pll),
                ("face_vert_ptr", _pll), ("vert_xyz", _pd), ("face_neighbor", _pi), ("domain", _d * 6),
                ("on_device", _i)]


class PointSet(ctypes.Structure):
    _fields_ = [("n", _i64), ("x", _pd), ("y", _pd), ("z", _pd), ("on_device", _i)]


class Tessellation(ctypes.Structure):
    _fields_ = [("cells", VoronoiCells), ("cell_site", _pi), ("n_dropped", _i64), ("device_ms_prep", _d),
                ("device_ms_count", _d), ("device_ms_fill", _d)]


class MapStats(ctypes.Structure):
    _fields_ = [("unresolved", _i64), ("unmapped", _i64), ("n_negative", _i64), ("n_negative_warn", _i64),
                ("pairs_candidate", _i64), ("pairs_visited", _i64), ("face_evals", _i64), ("v_mean", _d * 3),
                ("device_ms_prep", _d), ("device_ms_faces", _d), ("device_ms_accum", _d),
                ("n_warn_records", _i), ("warn", WarnRecord * 16)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_ if f not in ("v_mean", "warn")}
        d["v_mean"] = list(self.v_mean)
        d["warn"] = [(w.particle, w.cell, w.value) for w in self.warn[: self.n_warn_records]]
        return d


class ParticleGrads(ctypes.Structure):
    _fields_ = [("m", _pd), ("x", _pd), ("y", _pd), ("z", _pd), ("h", _pd), ("vx", _pd), ("vy", _pd), ("vz", _pd),
                ("on_device", _i)]


class AdjointStats(ctypes.Structure):
    _fields_ = [("skipped_outside", _i64), ("pairs_bbox", _i64), ("lattice_points", _i64), ("cells", _i64),
                ("device_ms_prep", _d), ("device_ms", _d)]


GRAD_FIELDS = ("m", "x", "y", "z", "h", "vx", "vy", "vz")

_lib = None


def lib_path():
    return LIB_PATH


def load():
    """Load libkiwi_hip.so.  Raises if it has not been built (python -m kiwi_amd.build)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libkiwi_hip.so not built at {LIB_PATH}: run python -m kiwi_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        L.kiwi_hip_abi_version.restype = _i
        L.kiwi_hip_create.restype = _i
        L.kiwi_hip_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), _i]
        L.kiwi_hip_destroy.restype = None
        L.kiwi_hip_destroy.argtypes = [ctypes.c_void_p]
        L.kiwi_hip_last_error.restype = ctypes.c_char_p
        L.kiwi_hip_last_error.argtypes = [ctypes.c_void_p]
        L.kiwi_hip_stream.restype = ctypes.c_void_p
        L.kiwi_hip_stream.argtypes = [ctypes.c_void_p]
        L.kiwi_hip_sph_to_grid.restype = _i
        L.kiwi_hip_sph_to_grid.argtypes = [ctypes.c_void_p, ctypes.POINTER(Particles), ctypes.POINTER(Grid),
                                           ctypes.POINTER(Kernel), _i, ctypes.POINTER(CellSums),
                                           ctypes.POINTER(Stats)]
        L.kiwi_hip_sph_to_grid_noalloc.restype = _i
        L.kiwi_hip_sph_to_grid_noalloc.argtypes = L.kiwi_hip_sph_to_grid.argtypes
        L.kiwi_hip_sph_to_grid_adjoint.restype = _i
        L.kiwi_hip_sph_to_grid_adjoint.argtypes = [ctypes.c_void_p, ctypes.POINTER(Particles), ctypes.POINTER(Grid),
                                                   ctypes.POINTER(Kernel), _i, ctypes.POINTER(CellSums),
                                                   ctypes.POINTER(ParticleGrads), ctypes.POINTER(AdjointStats)]
        L.kiwi_hip_sph_to_grid_batch.restype = _i
        L.kiwi_hip_sph_to_grid_batch.argtypes = [ctypes.c_void_p, _i, ctypes.POINTER(BatchProblem)]
        L.kiwi_hip_inverse_mapping.restype = _i
        L.kiwi_hip_inverse_mapping.argtypes = [ctypes.c_void_p, ctypes.POINTER(Particles), ctypes.c_void_p,
                                               ctypes.POINTER(VoronoiCells), ctypes.POINTER(Kernel), _i, _pd,
                                               ctypes.POINTER(InverseRows), ctypes.POINTER(MapStats)]
        L.kiwi_hip_cartesian_map.restype = _i
        L.kiwi_hip_cartesian_map.argtypes = [ctypes.c_void_p, ctypes.POINTER(Sites), ctypes.POINTER(_d), _i,
                                             ctypes.POINTER(CartesianMapOut), ctypes.POINTER(CartesianMapStats)]
        L.kiwi_hip_cartesian_lattice.argtypes = [ctypes.POINTER(_d), _d, ctypes.POINTER(Lattice)]
        L.kiwi_hip_coverage_buffer.argtypes = [ctypes.c_void_p, ctypes.POINTER(Particles), ctypes.POINTER(_d), _d, _d,
                                               ctypes.POINTER(BufferSites), ctypes.POINTER(CoverageStats)]
        L.kiwi_hip_set_tuning.restype = _i
        L.kiwi_hip_set_graphs.argtypes = [ctypes.c_void_p, _i]
        L.kiwi_hip_set_graphs.restype = _i
        L.kiwi_hip_get_graphs.argtypes = [ctypes.c_void_p, ctypes.POINTER(_i)]
        L.kiwi_hip_get_graphs.restype = _i
        L.kiwi_hip_set_fp64_fallback.argtypes = [ctypes.c_void_p, _d]
        L.kiwi_hip_set_fp64_fallback.restype = _i
        L.kiwi_hip_get_fp64_fallback.argtypes = [ctypes.c_void_p, ctypes.POINTER(_d)]
        L.kiwi_hip_get_fp64_fallback.restype = _i
        L.kiwi_hip_set_tuning.argtypes = [ctypes.c_void_p, ctypes.POINTER(Tuning)]
        L.kiwi_hip_get_tuning.restype = _i
        L.kiwi_hip_get_tuning.argtypes = [ctypes.c_void_p, ctypes.POINTER(Tuning)]
        L.kiwi_hip_autotune.restype = _i
        L.kiwi_hip_autotune.argtypes = [ctypes.c_void_p, _i, ctypes.POINTER(BatchProblem), ctypes.POINTER(Tuning),
                                        ctypes.POINTER(ctypes.c_double)]
        L.kiwi_hip_voronoi_centroids.restype = _i
        L.kiwi_hip_voronoi_centroids.argtypes = [ctypes.c_void_p, ctypes.POINTER(VoronoiCells), _pd, _pd, _pd]
        L.kiwi_hip_voronoi_tessellate.restype = _i
        L.kiwi_hip_voronoi_tessellate.argtypes = [ctypes.c_void_p, ctypes.POINTER(PointSet), ctypes.POINTER(_d),
                                                  ctypes.POINTER(Tessellation)]
        L.kiwi_hip_voronoi_home_cells.restype = _i
        L.kiwi_hip_voronoi_home_cells.argtypes = [ctypes.c_void_p, ctypes.POINTER(PointSet), _pi,
                                                  ctypes.POINTER(_i64)]
        L.kiwi_hip_voronoi_tessellation_fetch.restype = _i
        L.kiwi_hip_voronoi_tessellation_fetch.argtypes = [ctypes.c_void_p, _pll, _pll, _pd, _pi, _pi, _i]
        L.kiwi_hip_voronoi_map.restype = _i
        L.kiwi_hip_voronoi_map.argtypes = [ctypes.c_void_p, ctypes.POINTER(Particles), ctypes.c_void_p,
                                           ctypes.POINTER(VoronoiCells), ctypes.POINTER(Kernel), _i,
                                           ctypes.POINTER(CellSums), ctypes.POINTER(MapStats)]
        _lib = L
    return _lib


def cartesian_lattice(domain, spacing_fraction):
    """The coverage buffer's candidate lattice: (nx, ny, nz, spacing).  Raises ValueError on bad input."""
    lat = Lattice()
    dom = (_d * 6)(*[float(v) for v in domain])
    if load().kiwi_hip_cartesian_lattice(dom, float(spacing_fraction), ctypes.byref(lat)) != 0:
        raise ValueError("invalid domain or spacing_fraction")
    return lat.nx, lat.ny, lat.nz, lat.spacing


def kernel_spec(name="M4", support_in_h=0.0, generic_path=False, tabulated=False):
    key = ALIASES.get(str(name).lower().replace("-", "_").replace(" ", "_"), str(name).upper())
    if key not in KERNELS:
        raise ValueError(f"unknown kernel '{name}'")
    return Kernel(KERNELS.index(key), float(support_in_h), int(bool(generic_path)), int(bool(tabulated)))


def _ptr(a, ctype=_pd, dtype="float64"):
    """Host numpy array or device tensor of the given dtype -> (pointer, on_device)."""
    if isinstance(a, np.ndarray):
        assert a.dtype == np.dtype(dtype) and a.flags["C_CONTIGUOUS"], (a.dtype, dtype)
        return a.ctypes.data_as(ctype), 0
    # torch tensor (device memory plumbing only)
    assert str(a.dtype) == f"torch.{dtype}" and a.is_contiguous(), (a.dtype, dtype)
    return ctypes.cast(ctypes.c_void_p(a.data_ptr()), ctype), int(a.is_cuda)


def _device_inputs_ready(*objs):
    """Waits for torch's current stream when any of objs (tensors, or dicts / sequences of them) is a
    device tensor.  The library's streams do not order themselves after work torch has queued (a zero
    fill of an output, an optimiser step that wrote an input), so that work must be complete first."""
    stack = list(objs)
    while stack:
        o = stack.pop()
        if isinstance(o, dict):
            stack.extend(o.values())
        elif isinstance(o, (list, tuple)):
            stack.extend(o)
        elif getattr(o, "is_cuda", False):
            import torch
            torch.cuda.current_stream(o.device).synchronize()
            return


def _particles(parts):
    cols = [parts[c] for c in ("x", "y", "z", "vx", "vy", "vz", "m", "h")]
    ptrs = [_ptr(c) for c in cols]
    on_dev = {p[1] for p in ptrs}
    if len(on_dev) != 1:
        raise ValueError("particle columns must all be host arrays or all device tensors")
    return Particles(len(cols[0]), *[p[0] for p in ptrs], on_dev.pop())


class Context:
    """One libkiwi_hip context (device + stream).  Not thread-safe, like the C context."""

    def __init__(self, device=0):
        self.L = load()
        h = ctypes.c_void_p()
        rc = self.L.kiwi_hip_create(ctypes.byref(h), int(device))
        if rc != 0:
            raise KiwiError(rc, f"kiwi_hip_create(device={device}) failed")
        self.h = h
        self.device = int(device)

    def close(self):
        if self.h:
            self.L.kiwi_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def stream(self):
        return self.L.kiwi_hip_stream(self.h)

    def _check(self, rc):
        if rc != 0:
            raise KiwiError(rc, self.L.kiwi_hip_last_error(self.h).decode())

    def sph_to_grid(self, parts, domain, res, kernel=None, precision=FP64, out=None):
        """parts: dict of x y z vx vy vz m h (numpy float64 or device float64 tensors, all the same kind).
        domain: (x_min, x_max, y_min, y_max, z_min, z_max); res: (nx, ny, nz).
        out: None (host numpy results) or a tuple of 4 device float64 tensors of nx*ny*nz.
        Returns (mass, px, py, pz, Stats)."""
        kernel = kernel or kernel_spec()
        _device_inputs_ready(parts, out)
        P = _particles(parts)
        nx, ny, nz = (int(v) for v in res)
        G = Grid(*[float(v) for v in domain], nx, ny, nz)
        nc = nx * ny * nz
        if out is None:
            arrs = [np.zeros(nc) for _ in range(4)]
            O = CellSums(*[a.ctypes.data_as(_pd) for a in arrs], 0)
        else:
            arrs = list(out)
            O = CellSums(*[_ptr(a)[0] for a in arrs], 1)
        st = Stats()
        self._check(self.L.kiwi_hip_sph_to_grid(self.h, ctypes.byref(P), ctypes.byref(G), ctypes.byref(kernel),
                                                int(precision), ctypes.byref(O), ctypes.byref(st)))
        return arrs[0], arrs[1], arrs[2], arrs[3], st

    def sph_to_grid_noalloc(self, parts, domain, res, out, kernel=None, precision=FP64):
        """kiwi_hip_sph_to_grid_noalloc: as sph_to_grid with device particle columns and a tuple of 4 device
        float64 output tensors, computed without device allocation (in the library's fixed scratch area).
        Returns (mass, px, py, pz, Stats)."""
        kernel = kernel or kernel_spec()
        _device_inputs_ready(parts, out)
        P = _particles(parts)
        nx, ny, nz = (int(v) for v in res)
        G = Grid(*[float(v) for v in domain], nx, ny, nz)
        arrs = list(out)
        ptrs = [_ptr(a) for a in arrs]
        O = CellSums(*[q[0] for q in ptrs], int(all(q[1] for q in ptrs)))
        st = Stats()
        self._check(self.L.kiwi_hip_sph_to_grid_noalloc(self.h, ctypes.byref(P), ctypes.byref(G), ctypes.byref(kernel),
                                                        int(precision), ctypes.byref(O), ctypes.byref(st)))
        return arrs[0], arrs[1], arrs[2], arrs[3], st

    def _batch(self, problems):
        _device_inputs_ready([(q["parts"], q.get("out")) for q in problems])
        arr = (BatchProblem * max(len(problems), 1))()
        keep, results = [], []
        for i, q in enumerate(problems):
            P = _particles(q["parts"])
            nx, ny, nz = (int(v) for v in q["res"])
            G = Grid(*[float(v) for v in q["domain"]], nx, ny, nz)
            nc = nx * ny * nz
            if q.get("out") is None:
                outs = [np.zeros(nc) for _ in range(4)]
                O = CellSums(*[a.ctypes.data_as(_pd) for a in outs], 0)
            else:
                outs = list(q["out"])
                O = CellSums(*[_ptr(a)[0] for a in outs], 1)
            st = Stats()
            arr[i] = BatchProblem(P, G, q.get("kernel") or kernel_spec(), int(q.get("precision", FP64)), O,
                                  ctypes.pointer(st))
            keep.append((P, O, q["parts"]))
            results.append((outs, st))
        return arr, keep, results

    def sph_to_grid_batch(self, problems):
        """Many sph-to-grid problems in one call.  problems: list of dicts with keys parts, domain, res and
        optionally kernel, precision, out (as for sph_to_grid).  Returns a list of (mass, px, py, pz, Stats),
        one per problem."""
        arr, keep, results = self._batch(problems)
        self._check(self.L.kiwi_hip_sph_to_grid_batch(self.h, len(problems), arr))
        return [(o[0], o[1], o[2], o[3], st) for o, st in results]

    def set_tuning(self, item_grab=None, cursors=None):
        """The cell kernel's launch shape for later calls; no arguments restores the defaults."""
        if item_grab is None and cursors is None:
            self._check(self.L.kiwi_hip_set_tuning(self.h, None))
        else:
            cur = self.get_tuning()
            t = Tuning(int(item_grab if item_grab is not None else cur[0]),
                       int(cursors if cursors is not None else cur[1]))
            self._check(self.L.kiwi_hip_set_tuning(self.h, ctypes.byref(t)))

    def set_graphs(self, on):
        """Graph mode: sph-to-grid calls with device outputs replay a captured HIP graph of their pipeline."""
        self._check(self.L.kiwi_hip_set_graphs(self.h, 1 if on else 0))

    def get_graphs(self):
        v = _i()
        self._check(self.L.kiwi_hip_get_graphs(self.h, ctypes.byref(v)))
        return bool(v.value)

    def set_fp64_fallback(self, support_cells):
        """Mixed precision's FP64 fallback: particles whose support radius spans at least support_cells grid
        cells compute in FP64 (<= 0: off; default FP64_FALLBACK_CELLS)."""
        self._check(self.L.kiwi_hip_set_fp64_fallback(self.h, float(support_cells)))

    def get_fp64_fallback(self):
        v = _d()
        self._check(self.L.kiwi_hip_get_fp64_fallback(self.h, ctypes.byref(v)))
        return v.value

    def get_tuning(self):
        t = Tuning()
        self._check(self.L.kiwi_hip_get_tuning(self.h, ctypes.byref(t)))
        return t.item_grab, t.cursors

    def autotune(self, problems):
        """Time the batch (list as for sph_to_grid_batch) under every candidate launch shape, keep the
        fastest on the context.  Returns ((item_grab, cursors), ms of the fastest call)."""
        arr, keep, results = self._batch(problems)
        t, ms = Tuning(), ctypes.c_double()
        self._check(self.L.kiwi_hip_autotune(self.h, len(problems), arr, ctypes.byref(t), ctypes.byref(ms)))
        return (t.item_grab, t.cursors), ms.value

    def sph_to_grid_adjoint(self, parts, domain, res, grad_cells, kernel=None, precision=FP64, out=None):
        """Gradients of a loss L with respect to the particles' m, x, y, z, h, vx, vy, vz, given grad_cells =
        (dL/dmass, dL/dpx, dL/dpy, dL/dpz) per cell (numpy float64 or device float64 tensors of nx*ny*nz).
        out: None (host numpy results) or a dict of 8 device float64 tensors keyed by GRAD_FIELDS.
        Returns (dict of gradients, AdjointStats)."""
        kernel = kernel or kernel_spec()
        _device_inputs_ready(parts, grad_cells, out)
        P = _particles(parts)
        nx, ny, nz = (int(v) for v in res)
        G = Grid(*[float(v) for v in domain], nx, ny, nz)
        gp = [_ptr(a) for a in grad_cells]
        if len({q[1] for q in gp}) != 1:
            raise ValueError("cell gradients must all be host arrays or all device tensors")
        Gc = CellSums(*[q[0] for q in gp], gp[0][1])
        if out is None:
            grads = {f: np.zeros(P.n) for f in GRAD_FIELDS}
            O = ParticleGrads(*[grads[f].ctypes.data_as(_pd) for f in GRAD_FIELDS], 0)
        else:
            grads = dict(out)
            O = ParticleGrads(*[_ptr(grads[f])[0] for f in GRAD_FIELDS], 1)
        st = AdjointStats()
        self._check(self.L.kiwi_hip_sph_to_grid_adjoint(self.h, ctypes.byref(P), ctypes.byref(G),
                                                        ctypes.byref(kernel), int(precision), ctypes.byref(Gc),
                                                        ctypes.byref(O), ctypes.byref(st)))
        return grads, st

    def voronoi_map(self, parts, home, cells, kernel=None, precision=FP64, out=None):
        """The Voronoi map's per-cell sums.  parts: as sph_to_grid; home: int32 home cell per particle (-1:
        unresolved), host array or device tensor like the particle columns.  cells: dict with
        cell_face_ptr (int64, n_cells+1), face_vert_ptr (int64, n_faces+1), vert_xyz (float64, 3 per
        vertex), face_nbr (int32, n_faces) and domain (6 floats), all host arrays or all device tensors.
        Returns (mass, px, py, pz, MapStats) with n_cells entries each."""
        kernel = kernel or kernel_spec()
        _device_inputs_ready(parts, home, cells, out)
        P, hp, V, nc = self._map_inputs(parts, home, cells)
        if out is None:
            arrs = [np.zeros(max(nc, 0)) for _ in range(4)]
            O = CellSums(*[a.ctypes.data_as(_pd) for a in arrs], 0)
        else:
            arrs = list(out)
            O = CellSums(*[_ptr(a)[0] for a in arrs], 1)
        st = MapStats()
        self._check(self.L.kiwi_hip_voronoi_map(self.h, ctypes.byref(P), ctypes.cast(hp, ctypes.c_void_p),
                                                ctypes.byref(V), ctypes.byref(kernel), int(precision),
                                                ctypes.byref(O), ctypes.byref(st)))
        return arrs[0], arrs[1], arrs[2], arrs[3], st

    def inverse_mapping(self, parts, home, cells, cell_mass, kernel=None, precision=FP64, out=None):
        """The inverse-mapping test per particle (run_inverse_mapping_test_stage).  parts, home, cells: as
        voronoi_map; cell_mass: the forward map's density x volume per cell (host array, or a device tensor
        when out is given).  out: None (host results) or (m_recovered, coverage) device float64 tensors.
        Returns (m_recovered, coverage_fraction, MapStats)."""
        kernel = kernel or kernel_spec()
        _device_inputs_ready(parts, home, cells, cell_mass, out)
        P, hp, V, nc = self._map_inputs(parts, home, cells)
        if out is None:
            arrs = [np.zeros(max(P.n, 0)) for _ in range(2)]
            cm = np.ascontiguousarray(cell_mass, dtype=np.float64)
            R = InverseRows(arrs[0].ctypes.data_as(_pd), arrs[1].ctypes.data_as(_pd), 0)
            cmp = cm.ctypes.data_as(_pd)
        else:
            arrs = list(out)
            cm = cell_mass
            R = InverseRows(_ptr(arrs[0])[0], _ptr(arrs[1])[0], 1)
            cmp, cdev = _ptr(cm)
            if not cdev:
                raise ValueError("cell_mass must be a device tensor when the outputs are")
        st = MapStats()
        self._check(self.L.kiwi_hip_inverse_mapping(self.h, ctypes.byref(P), ctypes.cast(hp, ctypes.c_void_p),
                                                    ctypes.byref(V), ctypes.byref(kernel), int(precision), cmp,
                                                    ctypes.byref(R), ctypes.byref(st)))
        return arrs[0], arrs[1], st

    def cartesian_map(self, sites, domain, res, out=None):
        """The cartesian-map resample.  sites: dict of x, y, z (float64), id (int32), density, vx, vy, vz
        (float64, each site's mapping.dat row) and optionally mapped (int32, 0 = no row), all host arrays or
        all device tensors.  out: None (host results) or (density, vx, vy, vz, cell_id) device tensors of
        res^3.  Returns (density, vx, vy, vz, cell_id, CartesianMapStats)."""
        _device_inputs_ready(sites, out)
        cols = [_ptr(sites[k]) for k in ("x", "y", "z")]
        idp = _ptr(sites["id"], _pi, "int32")
        vals = [_ptr(sites[k]) for k in ("density", "vx", "vy", "vz")]
        mp = _ptr(sites["mapped"], _pi, "int32") if sites.get("mapped") is not None else (None, cols[0][1])
        if len({q[1] for q in cols + vals + [idp, mp]}) != 1:
            raise ValueError("site arrays must all be host arrays or all device tensors")
        n = len(sites["x"])
        if any(len(sites[k]) != n for k in sites if sites[k] is not None):
            raise ValueError("site arrays must all have the same length")
        S = Sites(n, *[q[0] for q in cols], idp[0], *[q[0] for q in vals], mp[0], cols[0][1])
        n3 = int(res) ** 3 if 0 < int(res) <= 2048 else 0  # the library rejects other resolutions
        if out is not None and any(len(a) != n3 for a in out):
            raise ValueError("out tensors must hold res^3 cells each")
        if out is None:
            arrs = [np.zeros(n3) for _ in range(4)] + [np.zeros(n3, dtype=np.int32)]
            O = CartesianMapOut(*[a.ctypes.data_as(_pd) for a in arrs[:4]], arrs[4].ctypes.data_as(_pi), 0)
        else:
            arrs = list(out)
            op = [_ptr(a) for a in arrs[:4]] + [_ptr(arrs[4], _pi, "int32")]
            if any(q[1] != 1 for q in op):
                raise ValueError("out must be device tensors")
            O = CartesianMapOut(*[q[0] for q in op], 1)
        dom = (_d * 6)(*[float(v) for v in domain])
        st = CartesianMapStats()
        self._check(self.L.kiwi_hip_cartesian_map(self.h, ctypes.byref(S), dom, int(res), ctypes.byref(O),
                                                  ctypes.byref(st)))
        return arrs[0], arrs[1], arrs[2], arrs[3], arrs[4], st

    def coverage_buffer(self, parts, domain, spacing_fraction=0.05, coverage_radius_factor=2.0, out=None):
        """The coverage-grid buffer.  parts: particle columns (x, y, z and h are read), host arrays or
        device tensors.  out: None (host results sized to the lattice) or (x, y, z, id) device tensors of
        a common length (the capacity).  Returns (ids, xyz (kept, 3) or the out tensors, CoverageStats)."""
        _device_inputs_ready(parts, out)
        cols = [_ptr(parts[k]) for k in ("x", "y", "z", "h")]
        if len({q[1] for q in cols}) != 1 or len({len(parts[k]) for k in ("x", "y", "z", "h")}) != 1:
            raise ValueError("x, y, z, h must be host arrays or device tensors of one length")
        P = Particles(len(parts["x"]), cols[0][0], cols[1][0], cols[2][0], None, None, None, None, cols[3][0],
                      cols[0][1])
        dom = (_d * 6)(*[float(v) for v in domain])
        st = CoverageStats()
        if out is None:
            try:
                nx, ny, nz, _ = cartesian_lattice(domain, spacing_fraction)
                cap = nx * ny * nz
            except ValueError:
                cap = 0  # the library reports the error
            arrs = [np.zeros(max(cap, 1)) for _ in range(3)] + [np.zeros(max(cap, 1), dtype=np.int32)]
            B = BufferSites(*[a.ctypes.data_as(_pd) for a in arrs[:3]], arrs[3].ctypes.data_as(_pi), cap, 0)
        else:
            arrs = list(out)
            op = [_ptr(a) for a in arrs[:3]] + [_ptr(arrs[3], _pi, "int32")]
            if any(q[1] != 1 for q in op) or len({len(a) for a in arrs}) != 1:
                raise ValueError("out must be device tensors of one length")
            B = BufferSites(*[q[0] for q in op], len(arrs[0]), 1)
        self._check(self.L.kiwi_hip_coverage_buffer(self.h, ctypes.byref(P), dom, float(spacing_fraction),
                                                    float(coverage_radius_factor), ctypes.byref(B), ctypes.byref(st)))
        if out is not None:
            return arrs[3], tuple(arrs[:3]), st
        k = st.kept
        return arrs[3][:k].copy(), np.stack([a[:k] for a in arrs[:3]], 1), st

    def voronoi_centroids(self, cells, sites, out=None):
        """The Lloyd step: each cell's centroid (its site's next position) and volume.  cells: as
        voronoi_map; sites: (n_cells, 3) float64, host array or device tensor like the cells.  out:
        (centroids (n_cells, 3), volumes (n_cells,)) arrays where the cells live, or None: new ones there
        (numpy arrays for host cells, device tensors for device cells).  Returns (centroids, volumes)."""
        _device_inputs_ready(cells, sites, out)
        cfp, d0 = _ptr(cells["cell_face_ptr"], _pll, "int64")
        fvp, d1 = _ptr(cells["face_vert_ptr"], _pll, "int64")
        vx, d2 = _ptr(cells["vert_xyz"])
        sp, d3 = _ptr(sites)
        if len({d0, d1, d2, d3}) != 1:
            raise ValueError("tessellation arrays and sites must all be host arrays or all device tensors")
        nc = len(cells["cell_face_ptr"]) - 1
        if int(np.prod(sites.shape)) != 3 * nc:
            raise ValueError("sites must hold 3 coordinates per cell")
        nf = len(cells["face_vert_ptr"]) - 1
        nv = int(np.prod(cells["vert_xyz"].shape)) // 3
        V = VoronoiCells(nc, nf, nv, cfp, fvp, vx, None, (_d * 6)(*[float(v) for v in cells["domain"]]), d0)
        if out is None and d0:  # the library writes the outputs where the cells live
            import torch
            dev = cells["cell_face_ptr"].device
            out = (torch.empty((max(nc, 0), 3), dtype=torch.float64, device=dev),
                   torch.empty(max(nc, 0), dtype=torch.float64, device=dev))
        if out is None:
            cen, vol = np.zeros((max(nc, 0), 3)), np.zeros(max(nc, 0))
            cp, vp = cen.ctypes.data_as(_pd), vol.ctypes.data_as(_pd)
        else:
            cen, vol = out
            (cp, e0), (vp, e1) = _ptr(cen), _ptr(vol)
            if e0 != d0 or e1 != d0:
                raise ValueError("centroid and volume outputs must live where the cells do")
        self._check(self.L.kiwi_hip_voronoi_centroids(self.h, ctypes.byref(V), sp, cp, vp))
        return cen, vol

    def voronoi_tessellate(self, sites, domain):
        """The Voronoi tessellation of the sites inside the domain box (kiwi_hip_voronoi_tessellate).
        sites: (n, 3) float64, host array or device tensor.  Returns (cells, info): cells as voronoi_map
        takes them, device tensors (cell_face_ptr, face_vert_ptr, vert_xyz (n_vertices, 3), face_nbr) plus
        cell_site (int32: the input index of each cell's site, voro++'s cell order) and domain; info: dict
        with n_dropped and the device times.  The tensors are copies (kiwi_hip_voronoi_tessellation_fetch):
        the library's own arrays are reused by its next tessellation."""
        import torch
        if isinstance(sites, torch.Tensor):
            if sites.dtype != torch.float64 or sites.dim() != 2 or sites.shape[1] != 3:
                raise ValueError("sites must be an (n, 3) float64 tensor")
            cols = [sites[:, k].contiguous() for k in range(3)]
            _device_inputs_ready(cols)  # the column copies are queued on torch's stream
            on_dev = int(sites.is_cuda)
            ptrs = [ctypes.cast(ctypes.c_void_p(c.data_ptr()), _pd) for c in cols]
        else:
            a = np.asarray(sites, dtype=np.float64)
            if a.ndim != 2 or a.shape[1] != 3:
                raise ValueError("sites must be an (n, 3) array")
            cols = [np.ascontiguousarray(a[:, k]) for k in range(3)]
            on_dev = 0
            ptrs = [c.ctypes.data_as(_pd) for c in cols]
        dom = (_d * 6)(*[float(v) for v in domain])
        S = PointSet(len(cols[0]), *ptrs, on_dev)
        T = Tessellation()
        self._check(self.L.kiwi_hip_voronoi_tessellate(self.h, ctypes.byref(S), dom, ctypes.byref(T)))
        V = T.cells
        dev = torch.device("cuda", self.device)
        nc, nf, nv = int(V.n_cells), int(V.n_faces), int(V.n_face_vertices)
        cells = {"cell_face_ptr": torch.empty(nc + 1, dtype=torch.int64, device=dev),
                 "face_vert_ptr": torch.empty(nf + 1, dtype=torch.int64, device=dev),
                 "vert_xyz": torch.empty((nv, 3), dtype=torch.float64, device=dev),
                 "face_nbr": torch.empty(nf, dtype=torch.int32, device=dev),
                 "cell_site": torch.empty(nc, dtype=torch.int32, device=dev)}
        ptr = {k: ctypes.c_void_p(t.data_ptr()) for k, t in cells.items()}
        self._check(self.L.kiwi_hip_voronoi_tessellation_fetch(
            self.h, ctypes.cast(ptr["cell_face_ptr"], _pll), ctypes.cast(ptr["face_vert_ptr"], _pll),
            ctypes.cast(ptr["vert_xyz"], _pd), ctypes.cast(ptr["face_nbr"], _pi), ctypes.cast(ptr["cell_site"], _pi), 1))
        cells["domain"] = [float(v) for v in domain]
        info = {"n_dropped": int(T.n_dropped), "device_ms_prep": T.device_ms_prep,
                "device_ms_count": T.device_ms_count, "device_ms_fill": T.device_ms_fill}
        return cells, info

    def voronoi_home_cells(self, points):
        """Each point's home cell in the context's last tessellation (kiwi_hip_voronoi_home_cells): int32,
        -1 where the container would not hold the point.  points: (m, 3) float64 host array or device
        tensor; the result lives there too.  Returns (home, unresolved)."""
        import torch
        if isinstance(points, torch.Tensor):
            if points.dtype != torch.float64 or points.dim() != 2 or points.shape[1] != 3:
                raise ValueError("points must be an (m, 3) float64 tensor")
            cols = [points[:, k].contiguous() for k in range(3)]
            home = torch.empty(points.shape[0], dtype=torch.int32, device=points.device)
            _device_inputs_ready(cols, home)
            ptrs = [ctypes.cast(ctypes.c_void_p(c.data_ptr()), _pd) for c in cols]
            hp = ctypes.cast(ctypes.c_void_p(home.data_ptr()), _pi)
            on_dev = int(points.is_cuda)
        else:
            a = np.asarray(points, dtype=np.float64)
            if a.ndim != 2 or a.shape[1] != 3:
                raise ValueError("points must be an (m, 3) array")
            cols = [np.ascontiguousarray(a[:, k]) for k in range(3)]
            home = np.zeros(len(a), dtype=np.int32)
            ptrs = [c.ctypes.data_as(_pd) for c in cols]
            hp = home.ctypes.data_as(_pi)
            on_dev = 0
        un = ctypes.c_int64(0)
        self._check(self.L.kiwi_hip_voronoi_home_cells(self.h, ctypes.byref(PointSet(len(cols[0]), *ptrs, on_dev)),
                                                       hp, ctypes.byref(un)))
        return home, int(un.value)

    def _map_inputs(self, parts, home, cells):
        P = _particles(parts)
        hp, hdev = _ptr(home, _pi, "int32")
        if P.n and hdev != P.on_device:
            raise ValueError("home must live where the particle columns do")
        cfp, d0 = _ptr(cells["cell_face_ptr"], _pll, "int64")
        fvp, d1 = _ptr(cells["face_vert_ptr"], _pll, "int64")
        vx, d2 = _ptr(cells["vert_xyz"])
        fnb, d3 = _ptr(cells["face_nbr"], _pi, "int32")
        if len({d0, d1, d2, d3}) != 1:
            raise ValueError("tessellation arrays must all be host arrays or all device tensors")
        nc = len(cells["cell_face_ptr"]) - 1
        nf = len(cells["face_vert_ptr"]) - 1
        nv = int(np.prod(cells["vert_xyz"].shape)) // 3
        V = VoronoiCells(nc, nf, nv, cfp, fvp, vx, fnb, (_d * 6)(*[float(v) for v in cells["domain"]]), d0)
        # the structures point into the caller's arrays, which outlive the call
        return P, hp, V, nc


def normalise_voronoi(mass, px, py, pz, volume, v_mean):
    """density = mass / volume (0 where volume <= 0), v = p / mass where mass > 0, else the mean velocity
    (src/mapping.cpp:489-499)."""
    pos = mass > 0
    safe = np.where(pos, mass, 1.0)
    dens = np.where(volume > 0, mass / np.where(volume > 0, volume, 1.0), 0.0)
    return (dens, np.where(pos, px / safe, v_mean[0]), np.where(pos, py / safe, v_mean[1]),
            np.where(pos, pz / safe, v_mean[2]))


def normalise(mass, px, py, pz, domain, res, v_mean):
    """density = mass / V and v = p / mass, or the mean velocity where no kernel reached the cell
    (src/cartesian_grid.cpp:303-318)."""
    nx, ny, nz = res
    V = ((domain[1] - domain[0]) / nx) * ((domain[3] - domain[2]) / ny) * ((domain[5] - domain[4]) / nz)
    dens = mass / V
    pos = mass > 0
    safe = np.where(pos, mass, 1.0)
    vx = np.where(pos, px / safe, v_mean[0])
    vy = np.where(pos, py / safe, v_mean[1])
    vz = np.where(pos, pz / safe, v_mean[2])
    return dens, vx, vy, vz
