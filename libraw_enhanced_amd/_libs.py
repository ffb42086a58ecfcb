"""Locate and load the package's native libraries (built in-tree under _lib/)."""
from __future__ import annotations

import ctypes
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib")
_cache: dict[str, ctypes.CDLL] = {}


class NativeLibraryMissing(RuntimeError):
    """The in-tree native library has not been built (run __graft_entry__.build())."""


def load_library(name: str) -> ctypes.CDLL:
    lib = _cache.get(name)
    if lib is not None:
        return lib
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise NativeLibraryMissing(
            f"{path} is not built; build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    _cache[name] = lib
    return lib
