"""MI355X-native merge-path tile SpMV (MERBIT's hot path, rebuilt for gfx950).

Device work runs only in libmerbit_amd.so (hand-written HIP); this package is
its Python front end.  See include/merbit_amd.h for the C ABI.
"""
from ._lib import LIB_PATH, MerbitError, lib, status_str  # noqa: F401

__all__ = ["LIB_PATH", "MerbitError", "lib", "status_str"]
