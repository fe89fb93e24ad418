"""Python bindings of the MI355X pixel reconstruction libraries.

The package directory holds the native code (HIP kernels and C ABI in csrc/,
the host framework and plugins in host/, the synthetic event source in gen/);
this module only loads the in-tree built libraries through ctypes.
"""
from ._lib import LIB_DIR, REPO_DIR, load_library  # noqa: F401
