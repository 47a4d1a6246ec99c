"""Locate and load the in-tree native libraries of the package."""
from __future__ import annotations

import ctypes
import os

LIBDIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
_cache: dict = {}


def lib_path(name: str) -> str:
    return os.path.join(LIBDIR, name)


def load_library(name: str) -> ctypes.CDLL:
    """Load ``cup3d_amd/lib/<name>``; raise if it has not been built."""
    if name in _cache:
        return _cache[name]
    path = lib_path(name)
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build the native libraries first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(path)
    _cache[name] = lib
    return lib
