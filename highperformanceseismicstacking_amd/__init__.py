"""MI355X (gfx950) semblance backend for HighPerformanceSeismicStacking's hot path.

The compute lives in ``libhss.so`` (hand-written HIP kernels behind the C ABI of
``include/hss.h``) and in the C++17 host library and command-line programs under
``csrc/host``.  This Python package is a thin ctypes binding used by the tests and the
benchmark; it never computes semblance itself and raises if the HIP library is missing.
"""
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
LIB_PATH = os.path.join(PKG_DIR, "libhss.so")
HOST_LIB_PATH = os.path.join(PKG_DIR, "libhss_host.so")
BIN_DIR = os.path.join(PKG_DIR, "bin")

CMP, ZOCRS, OCT = 0, 1, 2
F32, F16, BF16 = 0, 1, 2
EXACT, FAST, KAHAN = 0, 1, 2

MODEL_NAMES = {"cmp": CMP, "zocrs": ZOCRS, "oct": OCT}
MODEL_NPARAMS = {CMP: 1, ZOCRS: 3, OCT: 2}
