"""TurboQuant encode + quantized-distance path on AMD MI355X (gfx950).

Drop-in for the reference package: ``from turboquant import TurboQuantSpace``
gives the same class, constructor and methods, backed by hand-written HIP
kernels through the C ABI in include/tq_hip.h.
"""
from . import _runtime  # noqa: F401  (must run before the native module loads)
from ._turboquant import PreparedQuery, PreparedSymCode, TurboQuantSpace, device_count

__all__ = ["TurboQuantSpace", "PreparedQuery", "PreparedSymCode", "device_count"]
__version__ = "0.1.0"
