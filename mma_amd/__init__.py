"""MI355X-native implementation of the MMA hot path: the batched 1D-TDSE single-atom
solver (drop-in for the CTDSE library ``libsingleTDSE.so``) and the Hankel far-field
transform (drop-in for ``Hankel/Hankel_transform.py``).

The compute runs in hand-written HIP kernels for gfx950 (``mma_amd/csrc``), built into
``mma_amd/lib/libsingleTDSE.so`` by ``python -m mma_amd.build``.
"""
import ctypes
import os

__version__ = "0.1.0"

HERE = os.path.dirname(os.path.abspath(__file__))
# MMA_AMD_LIB_DIR: load the native libraries from another directory (e.g. a host-sanitizer
# build made by ``python -m mma_amd.build --sanitize DIR``, which puts them in DIR/lib)
LIB_DIR = os.environ.get("MMA_AMD_LIB_DIR") or os.path.join(HERE, "lib")
LIB_PATH = os.path.join(LIB_DIR, "libsingleTDSE.so")


class NativeLibraryMissing(RuntimeError):
    pass


_lib = None


def library_path():
    return LIB_PATH


def load_library():
    """The native library; raises NativeLibraryMissing if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing(
                "%s is missing: build it with `python -m mma_amd.build` (hipcc, gfx950)" % LIB_PATH)
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib
