"""Encode graphs (tq_encode_graph_*): every replay equals an eager encode of
the input buffer's contents at the time it runs, with eager work and
synchronisations between the replays; host buffers and empty batches are
rejected before anything is launched."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("dim,bits", [(768, 4), (200, 8), (1000, 3)])
def test_replays_equal_eager_encodes(gpu, dim, bits):
    import torch
    capi = gpu
    say(f"[gpu] encode graph dim={dim} bits={bits}")
    sp = capi.Space(dim, bits)
    n = 3000
    g = torch.Generator(device="cuda").manual_seed(dim + bits)
    X = torch.randn((n, dim), generator=g, device="cuda")
    codes_g = torch.zeros((n, sp.cs), dtype=torch.uint8, device="cuda")
    codes_e = torch.zeros_like(codes_g)
    eg = capi.EncodeGraph(sp, X, n, codes_g)
    for r in range(4):
        # new contents written in place by eager work between the replays
        X.copy_(torch.randn((n, dim), generator=g, device="cuda") * (1.0 + r))
        torch.cuda.synchronize()
        eg.run()
        sp.encode(X, n, codes_e)
        torch.cuda.synchronize()
        assert torch.equal(codes_g, codes_e), f"replay {r}"
    eg.close()
    sp.close()


def test_rejects_host_memory_and_empty_batches(gpu):
    import torch
    capi = gpu
    say("[gpu] encode graph argument checks")
    sp = capi.Space(128, 4)
    Xh = np.zeros((10, 128), np.float32)
    ch = np.zeros((10, sp.cs), np.uint8)
    with pytest.raises(capi.TQError):
        capi.EncodeGraph(sp, Xh, 10, ch)
    Xd = torch.zeros((10, 128), device="cuda")
    cd = torch.zeros((10, sp.cs), dtype=torch.uint8, device="cuda")
    with pytest.raises(capi.TQError):
        capi.EncodeGraph(sp, Xd, 10, ch)
    with pytest.raises(capi.TQError):
        capi.EncodeGraph(sp, Xd, 0, cd)
    sp.close()
