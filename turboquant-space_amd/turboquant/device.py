"""Zero-copy device-tensor paths of TurboQuantSpace.

The reference takes host buffers only (its users call ``.numpy()`` on torch
tensors, python/examples/torch_encode.py:15-17). Here the hot methods also
accept GPU tensors -- torch tensors on the device, or any object exporting
DLPack for a ROCm/CUDA device -- and then run on the device buffers directly,
enqueued on torch's current stream, returning torch tensors on the same
device. Host inputs keep the reference's numpy behaviour.
"""
import ctypes

from . import capi

_DL_DEVICE_GPU = (2, 10)  # DLPack device types: kDLCUDA, kDLROCM


def as_device_tensor(x):
    """torch.Tensor on the GPU for a device input, else None."""
    try:
        import torch
    except ImportError:  # pragma: no cover - torch is the device-memory layer
        return None
    if isinstance(x, torch.Tensor):
        return x if x.is_cuda else None
    dev = getattr(x, "__dlpack_device__", None)
    if dev is not None and getattr(x, "__dlpack__", None) is not None:
        try:
            kind = int(dev()[0])
        except Exception:
            return None
        if kind in _DL_DEVICE_GPU:
            return torch.from_dlpack(x)
    return None


def _stream(torch, t):
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _float_rows(t, dim, name, ndim):
    """Validation of float inputs, with the host path's messages."""
    import torch
    if t.dtype != torch.float32:
        raise ValueError(f"{name}: expected float32")
    if t.dim() != ndim:
        raise ValueError(f"{name}: wrong ndim")
    if t.shape[-1] != dim:
        raise ValueError(f"{name}: wrong dim" if ndim == 2 else f"{name}: wrong length")
    if not t.is_contiguous():
        raise ValueError(f"{name}: must be C-contiguous float32")
    return t.shape[0] if ndim == 2 else 1


def _code_rows(t, cs, name="codes"):
    import torch
    if t.dtype != torch.uint8:
        raise ValueError(f"{name} must be uint8")
    if t.numel() % cs != 0:
        raise ValueError(f"{name} size not a multiple of code_size_bytes")
    if not t.is_contiguous():
        raise ValueError(f"{name}: must be C-contiguous uint8")
    return t.numel() // cs


def _same_device(*ts):
    devs = {t.device for t in ts}
    if len(devs) != 1:
        raise ValueError("inputs must all be on the same device")
    return ts[0].device


class DeviceOps:
    """Device-buffer calls of one space (its own tq_space handle on the device
    the first tensor lives on)."""

    def __init__(self, dim, bits, rot_seed, qjl_seed):
        self.args = (dim, bits, rot_seed, qjl_seed)
        self.sp = None
        self.device = None

    def space(self, device):
        import torch
        if self.sp is None or self.device != device:
            with torch.cuda.device(device):
                self.sp = capi.Space(*self.args)
            self.device = device
        return self.sp

    def encode(self, X, out=None, single=False):
        import torch
        sp = self.space(X.device)
        n = _float_rows(X, sp.dim, "input", 1 if single else 2)
        if out is None:
            out = torch.empty((sp.cs,) if single else (n, sp.cs), dtype=torch.uint8,
                              device=X.device)
        else:
            _same_device(X, out)
            if out.dtype != torch.uint8 or not out.is_contiguous():
                raise ValueError("codes: must be uint8")
            if out.numel() != n * sp.cs:
                raise ValueError("codes: wrong byte size" if not single else "code: wrong byte size")
        if n:
            with torch.cuda.device(X.device):
                sp.encode(X, n, out, mem=capi.TQ_MEM_DEVICE, stream=_stream(torch, X))
        return out

    def m_to_n(self, Q, codes, mode):
        import torch
        sp = self.space(_same_device(Q, codes))
        m = _float_rows(Q, sp.dim, "queries", 2)
        n = _code_rows(codes, sp.cs)
        out = torch.empty((m, n), dtype=torch.float32, device=Q.device)
        if m and n:
            with torch.cuda.device(Q.device):
                sp.m_to_n(Q, m, codes, n, out, mode=mode, mem=capi.TQ_MEM_DEVICE,
                          stream=_stream(torch, Q))
        return out

    def one_to_n(self, q, codes, single=False):
        """distance_1_to_n (or distance when `single`: one code of exactly
        code_size bytes) through tq_distance_1_to_n, as the host path."""
        import torch
        sp = self.space(_same_device(q, codes))
        _float_rows(q, sp.dim, "query", 1)
        if single:
            if codes.dtype != torch.uint8 or not codes.is_contiguous():
                raise ValueError("code: must be uint8")
            if codes.numel() != sp.cs:
                raise ValueError("code: wrong byte size")
            n = 1
        else:
            n = _code_rows(codes, sp.cs)
        out = torch.empty((n,), dtype=torch.float32, device=q.device)
        if n:
            with torch.cuda.device(q.device):
                sp.one_to_n(q, codes, n, out, mem=capi.TQ_MEM_DEVICE, stream=_stream(torch, q))
        return out

    def symmetric(self, A, B, mode, single=False):
        import torch
        sp = self.space(_same_device(A, B))
        if single:
            for t, name in ((A, "code_a"), (B, "code_b")):
                if t.dtype != torch.uint8 or not t.is_contiguous():
                    raise ValueError(f"{name}: must be uint8")
                if t.numel() != sp.cs:
                    raise ValueError(f"{name}: wrong byte size")
        m, n = _code_rows(A, sp.cs), _code_rows(B, sp.cs)
        out = torch.empty((m, n), dtype=torch.float32, device=A.device)
        if m and n:
            with torch.cuda.device(A.device):
                sp.symmetric(A, m, B, n, out, mem=capi.TQ_MEM_DEVICE, stream=_stream(torch, A),
                             mode=mode)
        return out

    def rotated(self, X):
        import torch
        sp = self.space(X.device)
        n = _float_rows(X, sp.dim, "input", 2)
        out = torch.empty((n, sp.D), dtype=torch.float32, device=X.device)
        if n:
            with torch.cuda.device(X.device):
                sp.rotated(X, n, out, mem=capi.TQ_MEM_DEVICE, stream=_stream(torch, X))
        return out
