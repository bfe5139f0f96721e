"""honei_amd -- MI355X (gfx950) backend for HONEI's sparse solver path.

The compute lives in libhsolve.so (hand-written HIP kernels behind the
C-ABI in include/hsolve.h); this package only binds it.  Importing the
package does not load the library; the first call does, and fails loudly if
it is not built.
"""
from ._lib import HsolveError, LIB_PATH, exported_names, lib  # noqa: F401

__all__ = ["HsolveError", "LIB_PATH", "exported_names", "lib"]
