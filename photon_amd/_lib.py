"""Locating and loading the in-tree native libraries (built by photon_amd/Makefile)."""
import ctypes as C
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
HOST_LIB = os.path.join(LIB_DIR, "libphoton_host.so")
HIP_LIB = os.path.join(LIB_DIR, "libphoton_hip.so")

_cache = {}


class NativeLibraryMissing(RuntimeError):
    pass


def load(path):
    """Load a shared library once; a missing library is an error, never a silent fallback."""
    lib = _cache.get(path)
    if lib is None:
        if not os.path.exists(path):
            raise NativeLibraryMissing(f"{path} is not built; run __graft_entry__.build() (make -C photon_amd)")
        lib = C.CDLL(path, mode=C.RTLD_GLOBAL)
        _cache[path] = lib
    return lib
