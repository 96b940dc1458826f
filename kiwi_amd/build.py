"""Builds libkiwi_hip.so (gfx950, hipcc) and libkiwi_io.so (host snapshot readers, g++) in-tree.

    python -m kiwi_amd.build
"""
import os
import subprocess
import sys

from . import IO_LIB_PATH, LIB_PATH, PKG_DIR, ROOT_DIR

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are switched off for all device code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wall", "-Wno-unused-result"]


def _run(cmd):
    r = subprocess.run(cmd, cwd=ROOT_DIR, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(l for l in r.stderr.splitlines() if noise not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{err}")
    if err:
        sys.stderr.write(err + "\n")


SOURCES = ("kiwi_hip.hip", "kiwi_voronoi.hip", "kiwi_adjoint.hip", "kiwi_cartesian_map.hip", "kiwi_coverage.hip",
           "kiwi_lloyd.hip", "kiwi_tessellate.hip")
OBJ_DIR = os.path.join(PKG_DIR, "_build")


def _compile(src):
    obj = os.path.join(OBJ_DIR, os.path.splitext(os.path.basename(src))[0] + ".o")
    _run([HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT_DIR, "include"), "-c", src, "-o", obj + ".tmp"])
    os.replace(obj + ".tmp", obj)
    return obj


def build_lib():
    """Every translation unit compiled on its own (in parallel), then linked into one shared library."""
    from concurrent.futures import ThreadPoolExecutor
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = [os.path.join(PKG_DIR, "csrc", f) for f in SOURCES]
    with ThreadPoolExecutor(max_workers=min(len(srcs), os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB_PATH + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", *objs, "-o", tmp])
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def build_io():
    tmp = IO_LIB_PATH + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-fPIC", "-shared", "-pthread", "-I", os.path.join(ROOT_DIR, "include"),
          os.path.join(PKG_DIR, "csrc", "kiwi_io.cpp"), os.path.join(PKG_DIR, "csrc", "kiwi_config.cpp"), "-o", tmp])
    os.replace(tmp, IO_LIB_PATH)
    return IO_LIB_PATH


def build_all():
    build_lib()
    build_io()


if __name__ == "__main__":
    build_all()
