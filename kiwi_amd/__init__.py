"""kiwi_amd: MI355X (gfx950) implementation of KIWI's exact kernel-integral mapping.

The product is the C-ABI library libkiwi_hip.so (include/kiwi_hip.h, sources in kiwi_amd/csrc) and the
host snapshot readers libkiwi_io.so (include/kiwi_io.h).  kiwi_amd.hip and kiwi_amd.io are the Python
(ctypes) views of the same ABIs used by the tests and the benchmark; they raise if a library has not
been built.
"""
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "libkiwi_hip.so")
IO_LIB_PATH = os.path.join(PKG_DIR, "libkiwi_io.so")
