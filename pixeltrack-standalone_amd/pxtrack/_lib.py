"""Locates and loads the package's native libraries (built in-tree, never installed)."""
import ctypes
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PKG_DIR, "lib")
REPO_DIR = os.path.dirname(PKG_DIR)

_cache = {}


def lib_path(name):
    return os.path.join(LIB_DIR, name)


def load_library(name):
    """Loads lib/<name>; raises if it has not been built (no silent fallback)."""
    if name not in _cache:
        path = lib_path(name)
        if not os.path.exists(path):
            raise OSError(f"{path} is not built; run __graft_entry__.build() first")
        _cache[name] = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)
    return _cache[name]
