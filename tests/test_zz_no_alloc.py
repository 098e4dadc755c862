"""The allocation-free distance_m_to_n (tq_distance_m_to_n_noalloc): its
scratch is a fixed per-device area inside the library. Bar: its distances
equal tq_distance_m_to_n's bit for bit (same kernels), on two calls in a row
with the second on a smaller input (the area keeps the first call's
contents), for both modes and both code layouts; oversized calls are refused
with TQ_ENOMEM before any device work."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("dim,bits", [(768, 4), (128, 8), (1000, 4), (64, 2)])
@pytest.mark.parametrize("mode", ["fp32", "mfma"])
def test_noalloc_equals_m_to_n(gpu, dim, bits, mode):
    import torch
    capi = gpu
    md = capi.TQ_MODE_MFMA if mode == "mfma" else capi.TQ_MODE_FP32
    say(f"[gpu] no-alloc m_to_n dim={dim} bits={bits} mode={mode}")
    sp = capi.Space(dim, bits)
    g = torch.Generator(device="cuda").manual_seed(dim + bits)
    n = 5000
    X = torch.randn((n, dim), generator=g, device="cuda")
    codes = torch.empty((n, sp.cs), dtype=torch.uint8, device="cuda")
    sp.encode(X, n, codes)
    for m, nn in ((200, n), (37, 3001)):
        Q = torch.randn((m, dim), generator=g, device="cuda")
        want = torch.empty((m, nn), dtype=torch.float32, device="cuda")
        got = torch.full((m, nn), -1.0, dtype=torch.float32, device="cuda")
        sp.m_to_n(Q, m, codes, nn, want, mode=md)
        torch.cuda.synchronize()
        sp.m_to_n_noalloc(Q, m, codes, nn, got, mode=md)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(got.cpu().numpy().view(np.uint32),
                                      want.cpu().numpy().view(np.uint32))
    sp.close()


def test_noalloc_refuses_oversized(gpu):
    import torch
    capi = gpu
    sp = capi.Space(1024, 4)
    m = 8000  # ~16 KB of scratch per query at padded dim 1024: past the area
    Q = torch.zeros((m, 1024), device="cuda")
    codes = torch.zeros((10, sp.cs), dtype=torch.uint8, device="cuda")
    out = torch.zeros((m, 10), device="cuda")
    with pytest.raises(capi.TQError, match="scratch bytes"):
        sp.m_to_n_noalloc(Q, m, codes, 10, out, mode=capi.TQ_MODE_MFMA)
    sp.close()
