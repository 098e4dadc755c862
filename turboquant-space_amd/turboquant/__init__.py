"""TurboQuant encode + quantized-distance path on AMD MI355X (gfx950).

Drop-in for the reference package: ``from turboquant import TurboQuantSpace``
gives the same class, constructor and methods, backed by hand-written HIP
kernels through the C ABI in include/tq_hip.h. Host (numpy) inputs behave as
in the reference; the hot methods also take GPU tensors (torch, or DLPack
from a ROCm device) and then work on them in place and return GPU tensors.
"""
from . import _runtime  # noqa: F401  (must run before the native module loads)
from . import _turboquant as _native
from .device import DeviceOps, as_device_tensor
from ._turboquant import PreparedQuery, PreparedSymCode, device_count

__all__ = ["TurboQuantSpace", "PreparedQuery", "PreparedSymCode", "device_count"]
__version__ = "0.1.0"


class TurboQuantSpace(_native.TurboQuantSpace):
    """TurboQuantSpace(dim, bits_per_coord=4, rot_seed=42, qjl_seed=137,
    num_threads=0) -- include/turboquant/space_turbo_quant.h:250-311,
    python/turboquant/bindings.cpp:60-65."""

    def __init__(self, dim, bits_per_coord=4, rot_seed=42, qjl_seed=137, num_threads=0):
        super().__init__(dim, bits_per_coord, rot_seed, qjl_seed, num_threads)
        self._dev = DeviceOps(dim, bits_per_coord, rot_seed, qjl_seed)

    # encode / encode_into / encode_batch / encode_batch_into (bindings.cpp:75-116)
    def encode(self, x):
        t = as_device_tensor(x)
        return self._dev.encode(t, single=True) if t is not None else super().encode(x)

    def encode_into(self, x, out):
        t, o = as_device_tensor(x), as_device_tensor(out)
        if t is None and o is None:
            return super().encode_into(x, out)
        self._both(t, o)
        self._dev.encode(t, out=o, single=True)

    def encode_batch(self, X):
        t = as_device_tensor(X)
        return self._dev.encode(t) if t is not None else super().encode_batch(X)

    def encode_batch_into(self, X, out):
        t, o = as_device_tensor(X), as_device_tensor(out)
        if t is None and o is None:
            return super().encode_batch_into(X, out)
        self._both(t, o)
        self._dev.encode(t, out=o)

    # distance / distance_1_to_n / distance_m_to_n (bindings.cpp:135-194)
    def distance(self, query, code):
        q, c = as_device_tensor(query), as_device_tensor(code)
        if q is None and c is None:
            return super().distance(query, code)
        self._both(q, c)
        return float(self._dev.one_to_n(q, c, single=True)[0].item())

    def distance_1_to_n(self, query, codes):
        q, c = as_device_tensor(query), as_device_tensor(codes)
        if q is None and c is None:
            return super().distance_1_to_n(query, codes)
        self._both(q, c)
        return self._dev.one_to_n(q, c)

    def distance_m_to_n(self, queries, codes, mode="fp32"):
        q, c = as_device_tensor(queries), as_device_tensor(codes)
        if q is None and c is None:
            return super().distance_m_to_n(queries, codes, mode=mode)
        self._both(q, c)
        return self._dev.m_to_n(q, c, self._mode(mode))

    # distance_symmetric / distance_m_to_n_symmetric (bindings.cpp:145-218)
    def distance_symmetric(self, code_a, code_b):
        a, b = as_device_tensor(code_a), as_device_tensor(code_b)
        if a is None and b is None:
            return super().distance_symmetric(code_a, code_b)
        self._both(a, b)
        return float(self._dev.symmetric(a, b, 0, single=True)[0, 0].item())

    def distance_m_to_n_symmetric(self, codes_a, codes_b, mode="fp32"):
        a, b = as_device_tensor(codes_a), as_device_tensor(codes_b)
        if a is None and b is None:
            return super().distance_m_to_n_symmetric(codes_a, codes_b, mode=mode)
        self._both(a, b)
        return self._dev.symmetric(a, b, self._mode(mode))

    # rotated_coords_batch (bindings.cpp:119-132)
    def rotated_coords_batch(self, X):
        t = as_device_tensor(X)
        return self._dev.rotated(t) if t is not None else super().rotated_coords_batch(X)

    @staticmethod
    def _both(a, b):
        if a is None or b is None:
            raise ValueError("inputs must all be on the same device")

    @staticmethod
    def _mode(mode):
        if mode == "fp32":
            return 0
        if mode == "mfma":
            return 1
        raise ValueError("mode must be 'fp32' or 'mfma'")
