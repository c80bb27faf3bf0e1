"""MI355X-native EarthBox hot path: Python host bindings of the C ABI.

The compute path is libearthbox_gpu.so (hand-written HIP for gfx950); this
package only marshals device pointers.  Importing fails loudly when the
library is not built.
"""
from ._lib import EBGError, LIB_PATH, lib  # noqa: F401  (raises if unbuilt)
from .grid import StaggeredGrid2D, uniform_grid  # noqa: F401
from .context import GPUContext  # noqa: F401
