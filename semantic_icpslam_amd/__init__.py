"""MI355X-native dense SLAM hot path of semantic_icpSLAM.

The product is the HIP library `libsicp.so` (C ABI in include/sicp.h); this
package only locates it and exposes a ctypes binding (`binding`). Importing
the package does not load the library; `binding.lib()` does, and raises when
it has not been built.
"""
from . import binding  # noqa: F401

__all__ = ["binding"]
