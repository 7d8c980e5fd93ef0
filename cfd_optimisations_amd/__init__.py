"""MI355X (gfx950) explicit unstructured finite-volume Euler residual.

The product is libhcfd.so (HIP kernels behind the C ABI in include/hcfd.h) and
the C++ drop-in classes of namespace MI355X (host/).  This package holds the
build script, the ctypes binding (``lib``) and the file-format helpers
(``formats``).
"""
from . import formats  # noqa: F401

__all__ = ["formats", "lib", "cases"]
