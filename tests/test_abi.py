"""The C-ABI library: it loads without a GPU, exports every function include/kiwi_hip.h declares, and
its struct layouts agree with the Python view (no compute calls here)."""
import ctypes
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from kiwi_amd import LIB_PATH, hip  # noqa: E402

HEADER = os.path.join(ROOT, "include", "kiwi_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|const char|void\*)\s*\*?\s*(kiwi_hip_\w+)\s*\(", text, flags=re.M)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB_PATH):
        pytest.fail("libkiwi_hip.so is not built (python -m kiwi_amd.build)")
    return ctypes.CDLL(LIB_PATH)


def test_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert "kiwi_hip_sph_to_grid" in names and len(names) >= 6
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(hip.EXPORTED)


def test_abi_version(lib):
    lib.kiwi_hip_abi_version.restype = ctypes.c_int
    text = open(HEADER).read()
    want = int(re.search(r"#define KIWI_HIP_ABI_VERSION (\d+)", text).group(1))
    assert lib.kiwi_hip_abi_version() == want


def test_struct_layouts_match_header(tmp_path):
    """Compile a tiny C program against the header and compare sizeof/offsetof with ctypes."""
    src = tmp_path / "layout.c"
    fields = {"kiwi_particles": [f for f, _ in hip.Particles._fields_],
              "kiwi_grid": [f for f, _ in hip.Grid._fields_],
              "kiwi_kernel": [f for f, _ in hip.Kernel._fields_],
              "kiwi_cell_sums": [f for f, _ in hip.CellSums._fields_],
              "kiwi_stats": [f for f, _ in hip.Stats._fields_],
              "kiwi_voronoi_cells": [f for f, _ in hip.VoronoiCells._fields_],
              "kiwi_map_stats": [f for f, _ in hip.MapStats._fields_],
              "kiwi_particle_grads": [f for f, _ in hip.ParticleGrads._fields_],
              "kiwi_adjoint_stats": [f for f, _ in hip.AdjointStats._fields_],
              "kiwi_batch_problem": [f for f, _ in hip.BatchProblem._fields_],
              "kiwi_tuning": [f for f, _ in hip.Tuning._fields_],
              "kiwi_inverse_rows": [f for f, _ in hip.InverseRows._fields_],
              "kiwi_sites": [f for f, _ in hip.Sites._fields_],
              "kiwi_cartesian_map_out": [f for f, _ in hip.CartesianMapOut._fields_],
              "kiwi_cartesian_map_stats": [f for f, _ in hip.CartesianMapStats._fields_],
              "kiwi_lattice": [f for f, _ in hip.Lattice._fields_],
              "kiwi_buffer_sites": [f for f, _ in hip.BufferSites._fields_],
              "kiwi_coverage_stats": [f for f, _ in hip.CoverageStats._fields_],
              "kiwi_point_set": [f for f, _ in hip.PointSet._fields_],
              "kiwi_tessellation": [f for f, _ in hip.Tessellation._fields_]}
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void){"]
    for st, fs in fields.items():
        lines.append(f'printf("{st} %zu\\n", sizeof({st}));')
        for f in fs:
            lines.append(f'printf("{st}.{f} %zu\\n", offsetof({st}, {f}));')
    lines.append("return 0;}")
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", str(src), "-o", str(exe)], check=True)
    out = dict(l.rsplit(" ", 1) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                          check=True).stdout.split("\n") if l)
    for st, cls in (("kiwi_particles", hip.Particles), ("kiwi_grid", hip.Grid), ("kiwi_kernel", hip.Kernel),
                    ("kiwi_cell_sums", hip.CellSums), ("kiwi_stats", hip.Stats),
                    ("kiwi_voronoi_cells", hip.VoronoiCells), ("kiwi_map_stats", hip.MapStats),
                    ("kiwi_particle_grads", hip.ParticleGrads), ("kiwi_adjoint_stats", hip.AdjointStats),
                    ("kiwi_batch_problem", hip.BatchProblem), ("kiwi_tuning", hip.Tuning),
                    ("kiwi_inverse_rows", hip.InverseRows), ("kiwi_sites", hip.Sites),
                    ("kiwi_cartesian_map_out", hip.CartesianMapOut),
                    ("kiwi_cartesian_map_stats", hip.CartesianMapStats), ("kiwi_lattice", hip.Lattice),
                    ("kiwi_buffer_sites", hip.BufferSites), ("kiwi_coverage_stats", hip.CoverageStats)):
        assert int(out[st]) == ctypes.sizeof(cls), st
        for f, _ in cls._fields_:
            assert int(out[f"{st}.{f}"]) == getattr(cls, f).offset, (st, f)


def test_context_creation_fails_cleanly_without_gpu():
    """On a host without a GPU the library reports an error code instead of crashing."""
    try:
        import torch
        if torch.cuda.device_count() > 0:
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(hip.KiwiError):
        hip.Context(0)


def test_cartesian_lattice_matches_reference_sizes():
    """kiwi_hip_cartesian_lattice is a host computation (no GPU): the coverage buffer's lattice sizes and
    spacing (generate_cartesian_lattice, src/point_set.cpp:167-181) for domains where round() sits near
    a half, against the oracle's restatement, and its input checks."""
    sys.path.insert(0, ROOT)
    from oracle import oracle as O
    for dom, fr in (([0, 1, 0, 1, 0, 1], 0.05), ([-1.7, 1.7, -1.7, 1.7, -1.7, 1.7], 0.06), ([0, 2.5, 0, 1, 0, 3.5], 0.2),
                    ([0, 1.05, 0, 1, 0, 0.95], 0.1), ([0, 10, 0, 1, 0, 1], 1.0)):
        nx, ny, nz, sp = hip.cartesian_lattice(dom, fr)
        osp, on, _ = O.cartesian_lattice(dom, fr)
        assert (nx, ny, nz) == on and sp == osp, (dom, fr)
    with pytest.raises(ValueError):
        hip.cartesian_lattice([0, 1, 0, 1, 0, 1], 0.0)
    with pytest.raises(ValueError):
        hip.cartesian_lattice([0, 1, 1, 1, 0, 1], 0.05)


def test_io_library_exports_every_declared_symbol():
    """libkiwi_io.so (host snapshot readers) loads and exports every function include/kiwi_io.h declares;
    its struct layout matches the ctypes view."""
    from kiwi_amd import io
    L = io.load()
    text = open(os.path.join(ROOT, "include", "kiwi_io.h")).read()
    names = re.findall(r"^(?:int|void|const char\*)\s+(kiwi_io_\w+)\(", text, re.M)
    assert len(names) >= 6
    for n in names:
        assert hasattr(L, n), n
    assert L.kiwi_io_abi_version() == int(re.search(r"#define KIWI_IO_ABI_VERSION (\d+)", text).group(1))
    assert ctypes.sizeof(io.Snapshot) == 8 + 8 * 10
