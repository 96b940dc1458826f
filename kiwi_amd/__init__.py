"""kiwi_amd: MI355X (gfx950) implementation of KIWI's exact kernel-integral mapping.

The product is the C-ABI library libkiwi_hip.so (include/kiwi_hip.h, sources in kiwi_amd/csrc) and the
host CLI built on it (kiwi_amd/host).  kiwi_amd.hip is the Python (ctypes) view of the same ABI used by
the tests and the benchmark; it raises if the library has not been built.
"""
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "libkiwi_hip.so")
CLI_PATH = os.path.join(PKG_DIR, "kiwi")
