"""GPU tests of the drop-in's device-tensor paths: TurboQuantSpace methods
called with torch tensors on the GPU (or any DLPack exporter of a GPU
buffer) run on those buffers directly and return GPU tensors.

Bar: every result equals the host (numpy) call of the same method on the same
data bit for bit -- both go through the same C ABI entry points and kernels,
only the staging differs -- and codes equal the oracle's bytes.
"""
import sys

import numpy as np
import pytest

from _oracle import Oracle

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class DLPackOnly:
    """An object that exports a GPU buffer through DLPack only."""

    def __init__(self, t):
        self._t = t

    def __dlpack__(self, **kw):
        return self._t.__dlpack__(**kw)

    def __dlpack_device__(self):
        return self._t.__dlpack_device__()


def bits32(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("dim,bits", [(128, 4), (768, 4), (768, 8), (200, 3), (1536, 2)])
def test_device_paths_equal_host_paths(gpu, dim, bits):
    import torch
    import turboquant
    say(f"[gpu] device drop-in dim={dim} bits={bits}")
    s = turboquant.TurboQuantSpace(dim, bits_per_coord=bits)
    rng = np.random.default_rng(dim + bits)
    X = rng.standard_normal((1000, dim)).astype(np.float32)
    Q = rng.standard_normal((37, dim)).astype(np.float32)
    Xd, Qd = torch.from_numpy(X).cuda(), torch.from_numpy(Q).cuda()

    codes = s.encode_batch(X)
    np.testing.assert_array_equal(codes, Oracle(dim, bits).encode(X))
    cd = s.encode_batch(Xd)
    assert cd.is_cuda and cd.dtype == torch.uint8 and tuple(cd.shape) == codes.shape
    np.testing.assert_array_equal(cd.cpu().numpy(), codes)
    out = torch.zeros_like(cd)
    s.encode_batch_into(Xd, out)
    np.testing.assert_array_equal(out.cpu().numpy(), codes)
    one = s.encode(Xd[5])
    np.testing.assert_array_equal(one.cpu().numpy(), codes[5])
    o1 = torch.zeros(s.code_size_bytes(), dtype=torch.uint8, device="cuda")
    s.encode_into(Xd[6], o1)
    np.testing.assert_array_equal(o1.cpu().numpy(), codes[6])

    for mode in ("fp32", "mfma"):
        want = s.distance_m_to_n(Q, codes, mode=mode)
        got = s.distance_m_to_n(Qd, cd, mode=mode)
        assert got.is_cuda and tuple(got.shape) == want.shape
        np.testing.assert_array_equal(bits32(got.cpu().numpy()), bits32(want))
        wsym = s.distance_m_to_n_symmetric(codes[:40], codes, mode=mode)
        gsym = s.distance_m_to_n_symmetric(cd[:40], cd, mode=mode)
        np.testing.assert_array_equal(bits32(gsym.cpu().numpy()), bits32(wsym))

    w1 = s.distance_1_to_n(Q[3], codes)
    g1 = s.distance_1_to_n(Qd[3], cd)
    np.testing.assert_array_equal(bits32(g1.cpu().numpy()), bits32(w1))
    assert s.distance(Qd[3], cd[9]) == s.distance(Q[3], codes[9]) == w1[9]
    assert s.distance_symmetric(cd[1], cd[2]) == s.distance_symmetric(codes[1], codes[2])

    rot = s.rotated_coords_batch(X[:50])
    grot = s.rotated_coords_batch(Xd[:50])
    np.testing.assert_array_equal(bits32(grot.cpu().numpy()), bits32(rot))


def test_dlpack_inputs_and_side_stream(gpu):
    import torch
    import turboquant
    say("[gpu] device drop-in: DLPack exporters, side stream")
    dim = 768
    s = turboquant.TurboQuantSpace(dim)
    rng = np.random.default_rng(11)
    X = rng.standard_normal((2000, dim)).astype(np.float32)
    Q = rng.standard_normal((8, dim)).astype(np.float32)
    codes = s.encode_batch(X)
    want = s.distance_m_to_n(Q, codes, mode="mfma")
    Xd, Qd = torch.from_numpy(X).cuda(), torch.from_numpy(Q).cuda()
    cd = s.encode_batch(DLPackOnly(Xd))
    assert isinstance(cd, torch.Tensor) and cd.is_cuda
    np.testing.assert_array_equal(cd.cpu().numpy(), codes)
    got = s.distance_m_to_n(DLPackOnly(Qd), DLPackOnly(cd), mode="mfma")
    np.testing.assert_array_equal(bits32(got.cpu().numpy()), bits32(want))
    # Work is enqueued on torch's current stream, ordered after the producer.
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        Xs = torch.from_numpy(X).to("cuda", non_blocking=False) * 1.0
        cs = s.encode_batch(Xs)
        ds = s.distance_m_to_n(Qd, cs, mode="mfma")
    side.synchronize()
    np.testing.assert_array_equal(cs.cpu().numpy(), codes)
    np.testing.assert_array_equal(bits32(ds.cpu().numpy()), bits32(want))


def test_device_errors_and_empty(gpu):
    import torch
    import turboquant
    say("[gpu] device drop-in: errors and empty inputs")
    dim = 128
    s = turboquant.TurboQuantSpace(dim)
    cs = s.code_size_bytes()
    Xd = torch.randn(4, dim, device="cuda")
    cd = s.encode_batch(Xd)
    with pytest.raises(ValueError, match="same device"):
        s.distance_m_to_n(Xd, cd.cpu().numpy())
    with pytest.raises(ValueError, match="same device"):
        s.encode_batch_into(Xd, np.zeros((4, cs), np.uint8))
    with pytest.raises(ValueError, match="expected float32"):
        s.encode_batch(Xd.double())
    with pytest.raises(ValueError, match="wrong dim"):
        s.encode_batch(torch.zeros(3, dim + 1, device="cuda"))
    with pytest.raises(ValueError, match="must be C-contiguous"):
        s.encode_batch(torch.zeros(dim, 3, device="cuda").t())
    with pytest.raises(ValueError, match="multiple of code_size_bytes"):
        s.distance_1_to_n(Xd[0], cd.reshape(-1)[:-1])
    with pytest.raises(ValueError, match="codes must be uint8"):
        s.distance_1_to_n(Xd[0], cd.float())
    with pytest.raises(ValueError, match="codes: wrong byte size"):
        s.encode_batch_into(Xd, torch.zeros(3, cs, dtype=torch.uint8, device="cuda"))
    with pytest.raises(ValueError, match="code_b: wrong byte size"):
        s.distance_symmetric(cd[0], cd[1][:-1])
    with pytest.raises(ValueError, match="mode must be"):
        s.distance_m_to_n(Xd, cd, mode="bf16")
    e = s.encode_batch(torch.zeros(0, dim, device="cuda"))
    assert tuple(e.shape) == (0, cs)
    d = s.distance_m_to_n(Xd, torch.zeros(0, cs, dtype=torch.uint8, device="cuda"))
    assert tuple(d.shape) == (4, 0)
    d = s.distance_m_to_n(torch.zeros(0, dim, device="cuda"), cd, mode="mfma")
    assert tuple(d.shape) == (0, 4)
    torch.cuda.synchronize()
