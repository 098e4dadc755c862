"""The Python drop-in (`from turboquant import TurboQuantSpace`) keeps the
reference binding's behaviour: the invariants of the reference's own tests
(python/tests/test_turboquant.py) and its validation errors
(python/turboquant/bindings.cpp:14-49, :159-211)."""
import numpy as np
import pytest

from _oracle import Oracle, scale_rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def TQS():
    from turboquant import TurboQuantSpace
    return TurboQuantSpace


@pytest.mark.parametrize("dim", [32, 128, 512, 1024, 2048, 4096])
@pytest.mark.parametrize("bits", [4, 8])
def test_distance_approximates_l2(TQS, dim, bits):
    # test_turboquant.py:13-27
    space = TQS(dim, bits_per_coord=bits)
    rng = np.random.default_rng(dim * 10 + bits)
    q = rng.standard_normal(dim).astype(np.float32)
    x = rng.standard_normal(dim).astype(np.float32)
    tq = space.distance(q, space.encode(x))
    brute = float(np.linalg.norm(q - x) ** 2)
    tol = 0.25 if bits >= 8 else 0.5
    assert abs(tq - brute) < tol * brute + 1.0


def test_self_distance_near_zero(TQS):
    # test_turboquant.py:30-39
    space = TQS(512, bits_per_coord=8)
    x = np.random.default_rng(7).standard_normal(512).astype(np.float32)
    assert space.distance(x, space.encode(x)) < 0.1 * float(np.linalg.norm(x) ** 2)


def test_accessors(TQS):
    s = TQS(1000, bits_per_coord=4, rot_seed=1, qjl_seed=2, num_threads=3)
    assert (s.dim(), s.padded_dim(), s.padded(), s.code_size_bytes(), s.bits_per_coord(),
            s.num_threads()) == (1000, 1024, True, 524, 4, 3)
    s8 = TQS(768, 8)
    assert s8.code_size_bytes() == 1036 and s8.padded_dim() == 1024


def test_encode_into_and_batch_rows(TQS):
    # test_turboquant.py:60-82
    space = TQS(128, bits_per_coord=8)
    rng = np.random.default_rng(1)
    X = rng.standard_normal((17, 128)).astype(np.float32)
    codes = space.encode_batch(X)
    assert codes.shape == (17, space.code_size_bytes()) and codes.dtype == np.uint8
    for i in range(17):
        np.testing.assert_array_equal(codes[i], space.encode(X[i]))
    buf = np.empty(space.code_size_bytes(), np.uint8)
    space.encode_into(X[3], buf)
    np.testing.assert_array_equal(buf, codes[3])
    out = np.empty_like(codes)
    space.encode_batch_into(X, out)
    np.testing.assert_array_equal(out, codes)
    np.testing.assert_array_equal(codes, Oracle(128, 8).encode(X))


@pytest.mark.parametrize("bits", [4, 8])
def test_batch_distances_match_single(TQS, bits):
    # test_turboquant.py:101-146 (1e-4 absolute)
    dim = 96
    space = TQS(dim, bits_per_coord=bits)
    rng = np.random.default_rng(bits)
    X = rng.standard_normal((40, dim)).astype(np.float32)
    Q = rng.standard_normal((5, dim)).astype(np.float32)
    codes = space.encode_batch(X)
    d1 = space.distance_1_to_n(Q[0], codes)
    assert d1.shape == (40,) and d1.dtype == np.float32
    for i in range(40):
        assert abs(d1[i] - space.distance(Q[0], codes[i])) < 1e-4
    M = space.distance_m_to_n(Q, codes)
    assert M.shape == (5, 40)
    for i in range(5):
        for j in range(0, 40, 7):
            assert abs(M[i, j] - space.distance(Q[i], codes[j])) < 1e-4
    S = space.distance_m_to_n_symmetric(codes[:6], codes)
    assert S.shape == (6, 40)
    for i in range(6):
        for j in range(0, 40, 5):
            assert abs(S[i, j] - space.distance_symmetric(codes[i], codes[j])) < 1e-4
    # codes in any shape are accepted, as in the reference (total bytes only)
    np.testing.assert_array_equal(space.distance_1_to_n(Q[0], codes.reshape(-1)), d1)


def test_mfma_mode_close_to_fp64(TQS):
    dim, bits = 768, 4
    space = TQS(dim, bits_per_coord=bits)
    rng = np.random.default_rng(3)
    X = rng.standard_normal((700, dim)).astype(np.float32)
    Q = rng.standard_normal((33, dim)).astype(np.float32)
    codes = space.encode_batch(X)
    got = space.distance_m_to_n(Q, codes, mode="mfma")
    o = Oracle(dim, bits)
    qns = np.sum(Q.astype(np.float64) ** 2, 1)
    xns = o.trailer(codes)[:, 0].astype(np.float64) ** 2
    assert scale_rel_err(got, o.m_to_n_f64(Q, codes), qns, xns).max() <= 1e-4


def test_validation_messages(TQS):
    space = TQS(64, bits_per_coord=4)
    with pytest.raises(ValueError, match="input: expected float32"):
        space.encode(np.zeros(64, np.float64))
    with pytest.raises(ValueError, match="input: wrong length"):
        space.encode(np.zeros(63, np.float32))
    with pytest.raises(ValueError, match="input: wrong ndim"):
        space.encode_batch(np.zeros(64, np.float32))
    with pytest.raises(ValueError, match="input: wrong dim"):
        space.encode_batch(np.zeros((2, 65), np.float32))
    with pytest.raises(ValueError, match="must be C-contiguous float32"):
        space.encode_batch(np.zeros((4, 128), np.float32)[:, ::2])
    with pytest.raises(ValueError, match="codes size not a multiple of code_size_bytes"):
        space.distance_1_to_n(np.zeros(64, np.float32), np.zeros(45, np.uint8))
    with pytest.raises(ValueError, match="codes must be uint8"):
        space.distance_1_to_n(np.zeros(64, np.float32), np.zeros(11, np.float32))
    with pytest.raises(ValueError, match="code: wrong byte size"):
        space.encode_into(np.zeros(64, np.float32), np.zeros(3, np.uint8))
    with pytest.raises(ValueError, match="codes: wrong byte size"):
        space.encode_batch_into(np.zeros((2, 64), np.float32), np.zeros(3, np.uint8))
    with pytest.raises(ValueError):
        TQS(64, bits_per_coord=1)


def test_torch_cpu_tensors(TQS):
    # test_turboquant.py:42-57 (buffer protocol through .numpy())
    import torch
    space = TQS(128, bits_per_coord=4)
    torch.manual_seed(1)
    q, x = torch.randn(128), torch.randn(128)
    d = space.distance(q.numpy(), space.encode(x.numpy()))
    assert abs(d - float(((q - x) ** 2).sum())) < 0.5 * float(((q - x) ** 2).sum()) + 1.0
