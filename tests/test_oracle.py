"""Pins the CPU oracle (oracle/tq_oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible mistake in the restatement (a dropped
FMA, a wrong reduction order, a transposed butterfly, an off-by-one boundary)
fails: golden fixtures written by the reference build, the live reference
build where present, published known answers and closed forms.
"""
import hashlib
import json

import numpy as np
import pytest

from _oracle import (MANIFEST, Oracle, fp32_parity_excess, load_golden, olib, ptr, ref_available, rlib,
                     scale_rel_err)

GOLD = load_golden()
with open(MANIFEST) as f:
    META = json.load(f)
SHAPES = [(s["dim"], s["bits"]) for s in META["shapes"]]


def test_splitmix64_published_values():
    # Published splitmix64 outputs for seed 0 (Vigna's reference generator).
    st = __import__("ctypes").c_uint64(0)
    got = [olib().tqo_splitmix64_next(st) for _ in range(3)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


@pytest.mark.parametrize("seed", [42, 137, 0, 12345])
def test_signs_match_reference(seed):
    out = np.zeros(4096, np.float32)
    olib().tqo_signs(4096, seed, ptr(out))
    np.testing.assert_array_equal(out, GOLD[f"signs_{seed}"])


@pytest.mark.parametrize("mb", range(1, 9))
def test_codebook_bitwise_and_structure(mb):
    L = 1 << mb
    c = np.zeros(L, np.float32)
    b = np.zeros(L - 1, np.float32)
    olib().tqo_lloyd_max(mb, ptr(c), ptr(b))
    np.testing.assert_array_equal(c.view(np.uint32), GOLD[f"codebook_c_{mb}"].view(np.uint32))
    np.testing.assert_array_equal(b.view(np.uint32), GOLD[f"codebook_b_{mb}"].view(np.uint32))
    half = L // 2
    assert np.all(np.diff(c) > 0) and np.all(np.diff(b) > 0)
    np.testing.assert_array_equal(c[:half], -c[::-1][:half])       # symmetric
    assert b[half - 1] == 0.0
    mids = 0.5 * (c[:-1].astype(np.float64) + c[1:].astype(np.float64))
    np.testing.assert_allclose(b, mids, rtol=1e-6, atol=1e-7)       # midpoints


def test_codebook_textbook_values():
    # Max (1960) / Lloyd optimal quantizers for N(0,1): independent of the code.
    c1 = np.zeros(2, np.float32); olib().tqo_lloyd_max(1, ptr(c1), ptr(np.zeros(1, np.float32)))
    np.testing.assert_allclose(c1[1], np.sqrt(2 / np.pi), rtol=1e-6)
    c2 = np.zeros(4, np.float32); olib().tqo_lloyd_max(2, ptr(c2), ptr(np.zeros(3, np.float32)))
    np.testing.assert_allclose(c2[2:], [0.4528, 1.5104], atol=2e-4)
    c3 = np.zeros(8, np.float32); olib().tqo_lloyd_max(3, ptr(c3), ptr(np.zeros(7, np.float32)))
    np.testing.assert_allclose(c3[4:], [0.2451, 0.7560, 1.3440, 2.1520], atol=2e-4)


@pytest.mark.parametrize("mb", [1, 2, 3, 4, 5])
def test_codebook_is_fixed_point(mb):
    # Converged tables (1..5 MSE bits; 6..8 stop at the 1000-sweep cap and are
    # pinned by the fixture): each centroid is the N(0,1) mean over its cell.
    from scipy.stats import norm
    L = 1 << mb
    c = np.zeros(L, np.float32); b = np.zeros(L - 1, np.float32)
    olib().tqo_lloyd_max(mb, ptr(c), ptr(b))
    edges = np.concatenate([[-np.inf], b.astype(np.float64), [np.inf]])
    mass = norm.cdf(edges[1:]) - norm.cdf(edges[:-1])
    mean = (norm.pdf(edges[:-1]) - norm.pdf(edges[1:])) / mass
    np.testing.assert_allclose(c, mean, rtol=2e-5, atol=1e-6)


@pytest.mark.parametrize("D", [4, 8, 16, 32, 64])
def test_wht_closed_form(D):
    i = np.arange(D)
    H = (-1.0) ** np.array([[bin(a & b).count("1") for b in i] for a in i])
    x = np.random.default_rng(D).standard_normal(D).astype(np.float32)
    y = x.copy()
    olib().tqo_wht(ptr(y), D)
    np.testing.assert_allclose(y, H @ x.astype(np.float64) / np.sqrt(D), rtol=1e-5, atol=1e-6)
    # small integers: every butterfly is exact, and 1/sqrt(D) is exact for D = 4^k
    if D in (4, 16, 64):
        z = np.random.default_rng(D + 1).integers(-8, 8, D).astype(np.float32)
        w = z.copy()
        olib().tqo_wht(ptr(w), D)
        np.testing.assert_array_equal(w, (H @ z.astype(np.float64)) / np.sqrt(D))


def test_wht_is_an_involution_scaled():
    x = np.random.default_rng(3).standard_normal(256).astype(np.float32)
    y = x.copy()
    olib().tqo_wht(ptr(y), 256)
    olib().tqo_wht(ptr(y), 256)
    np.testing.assert_allclose(y, x, rtol=0, atol=1e-5)


def test_sumsq_order_matters_and_matches_closed_sum():
    x = np.random.default_rng(5).standard_normal(1003).astype(np.float32)
    s = olib().tqo_sumsq(ptr(x), len(x))
    np.testing.assert_allclose(s, np.sum(x.astype(np.float64) ** 2), rtol=1e-6)


@pytest.mark.parametrize("dim,bits", SHAPES)
def test_encode_matches_reference_fixture(dim, bits):
    o = Oracle(dim, bits)
    X = GOLD[f"X_{dim}_{bits}"]
    np.testing.assert_array_equal(o.encode(X), GOLD[f"codes_{dim}_{bits}"])


@pytest.mark.parametrize("dim,bits", SHAPES)
def test_prepare_matches_reference_fixture(dim, bits):
    o = Oracle(dim, bits)
    got = o.prepare(GOLD[f"Q_{dim}_{bits}"])[:, : 2 * o.D + 2]
    np.testing.assert_array_equal(got.view(np.uint32), GOLD[f"prep_{dim}_{bits}"].view(np.uint32))


@pytest.mark.parametrize("dim,bits", SHAPES)
def test_rotated_coords_match_reference_fixture(dim, bits):
    # rotatedCoordsBatch (space_turbo_quant.h:411-435); the reference build
    # vectorises both of its reductions the same way as encodeVector's.
    o = Oracle(dim, bits)
    X = GOLD[f"X_{dim}_{bits}"]
    got = np.zeros((len(X), o.D), np.float32)
    olib().tqo_rotated_coords_batch(o.h, ptr(X), len(X), ptr(got))
    np.testing.assert_array_equal(got.view(np.uint32), GOLD[f"rot_{dim}_{bits}"].view(np.uint32))


def test_encode_full_cpu_workload_digest():
    # 100,000 x 128, 4-bit, iid N(0,1) seed 0: sha256 of the whole code matrix.
    o = Oracle(128, 4)
    X = np.random.default_rng(0).standard_normal((100_000, 128)).astype(np.float32)
    codes = o.encode(X)
    np.testing.assert_array_equal(codes[:64], GOLD["cfg1_codes_head"])
    assert hashlib.sha256(codes.tobytes()).hexdigest() == META["cfg1_full_codes_sha256"]


@pytest.mark.parametrize("bits", [2, 4, 8, 9])
def test_encode_zero_vector_closed_form(bits):
    o = Oracle(64, bits)
    code = o.encode(np.zeros((1, 64), np.float32))
    half = (1 << (bits - 1)) // 2
    # sigma clamps to 1e-10, every normalized coordinate is 0 -> idx = half - 1
    norm, gamma, sigma = o.trailer(code)[0]
    assert norm == 0.0 and sigma == np.float32(1e-10)
    units = unpack(code[0], o)
    np.testing.assert_array_equal(units >> 1, (half - 1) & 0x7F if bits == 9 else half - 1)
    # residual is the constant a = -c[half-1]*1e-10 > 0; its RHT has the sign of
    # the integer sum_i s_i (-1)^popcount(i&k) (zero -> +0 -> bit 1)
    D = 64
    i = np.arange(D)
    H = (-1) ** np.array([[bin(a & b).count("1") for b in i] for a in i])
    want = ((H @ o.qjl.astype(np.int64)) >= 0).astype(np.uint8)
    np.testing.assert_array_equal(units & 1, want)
    a = np.float32(-o.centroids[half - 1] * np.float32(1e-10))
    np.testing.assert_allclose(gamma, np.sqrt(D) * a, rtol=1e-6)


@pytest.mark.parametrize("bits", [3, 4, 8])
def test_encode_basis_vector_closed_form(bits):
    # x = 2 e_5 in dim 16: every rotated coordinate is +-1/4 exactly, sigma = 1/4,
    # so every normalized coordinate is exactly +-1.
    o = Oracle(16, bits)
    x = np.zeros((1, 16), np.float32)
    x[0, 5] = 2.0
    code = o.encode(x)
    norm, gamma, sigma = o.trailer(code)[0]
    assert norm == 2.0 and sigma == 0.25
    i = np.arange(16)
    H = (-1) ** np.array([[bin(a & b).count("1") for b in i] for a in i])
    rot = H[:, 5] * o.rot[5] / 4.0                    # RHT(e_5)
    norm_c = rot / 0.25
    idx = np.searchsorted(o.boundaries, norm_c, side="right")
    np.testing.assert_array_equal(unpack(code[0], o) >> 1, idx)
    resid = rot - o.centroids[idx].astype(np.float64) * 0.25
    np.testing.assert_allclose(gamma, np.sqrt(np.sum(resid ** 2)), rtol=1e-6)


def unpack(code, o):
    if o.bits <= 4:
        p = code[: o.packed]
        return np.stack([p & 0xF, p >> 4], 1).reshape(-1)
    return code[: o.packed].copy()


@pytest.mark.parametrize("dim,bits", SHAPES)
def test_distances_match_reference_fixture(dim, bits):
    o = Oracle(dim, bits)
    tag = f"{dim}_{bits}"
    codes, Q = GOLD[f"codes_{tag}"], GOLD[f"Q_{tag}"]
    got = o.m_to_n(Q, codes)
    qns = GOLD[f"prep_{tag}"][:, 2 * o.D]
    xns = o.trailer(codes)[:, 0].astype(np.float64) ** 2
    assert scale_rel_err(got, GOLD[f"dist_{tag}"], qns, xns).max() <= 1e-5
    assert scale_rel_err(got[:1], GOLD[f"d1_{tag}"][None], qns[:1], xns).max() <= 1e-5
    A = codes[-4:]
    na = o.trailer(A)[:, 0].astype(np.float64) ** 2
    sym, gold = o.symmetric(A, codes), GOLD[f"sym_{tag}"]
    if o.D <= 4096:
        assert scale_rel_err(sym, gold, na, xns).max() <= 1e-5
    else:
        # A self-distance at padded dim 8192 sums 8,192 same-sign fp32 terms;
        # the reference build reassociates that sum (-ffast-math), the oracle
        # keeps the source's order. Every entry is within 1e-5 of scale of the
        # reference beyond the reference's own error, or within the textbook
        # bound of a recursive fp32 sum of D non-negative terms,
        # (D - 1) * 2^-24 * sum, of the exact (fp64) estimator.
        exact = o.symmetric_f64(A, codes)
        excess = fp32_parity_excess(sym, gold, exact, na, xns)
        bound = (o.D - 1) * 2.0 ** -24 * (na[:, None] + xns[None, :])
        assert np.all((excess <= 1e-5) | (np.abs(sym - exact) <= bound))


@pytest.mark.parametrize("dim,bits", [(128, 4), (768, 8), (1000, 4), (512, 3), (33, 7)])
def test_symmetric_full_and_light_match_reference_fixture(dim, bits):
    o = Oracle(dim, bits)
    tag = f"{dim}_{bits}"
    codes = GOLD[f"codes_{tag}"]
    A = codes[-4:]
    full = np.array([[olib().tqo_distance_symmetric_full(o.h, ptr(a), ptr(b)) for b in codes]
                     for a in A], np.float32)
    light = np.array([[olib().tqo_distance_symmetric_light(o.h, ptr(a), ptr(b)) for b in codes]
                      for a in A], np.float32)
    na = o.trailer(A)[:, 0].astype(np.float64) ** 2
    xns = o.trailer(codes)[:, 0].astype(np.float64) ** 2
    assert scale_rel_err(full, GOLD[f"symfull_{tag}"], na, xns).max() <= 1e-5
    assert scale_rel_err(light, GOLD[f"symlight_{tag}"], na, xns).max() <= 1e-5
    prep = np.array([[olib().tqo_distance_symmetric_full_prepared(o.h, ptr(a), ptr(b))
                      for b in codes] for a in A], np.float32)
    assert scale_rel_err(prep, GOLD[f"symfullprep_{tag}"], na, xns).max() <= 1e-5
    # The prepared form differs from the unprepared one by rounding only.
    assert scale_rel_err(GOLD[f"symfullprep_{tag}"], GOLD[f"symfull_{tag}"], na, xns).max() <= 1e-5


@pytest.mark.parametrize("dim,bits", [(128, 4), (768, 4), (768, 8), (1000, 4)])
def test_fp32_estimator_close_to_fp64_restatement(dim, bits):
    o = Oracle(dim, bits)
    rng = np.random.default_rng(dim + bits)
    X = rng.standard_normal((300, dim)).astype(np.float32)
    Q = rng.standard_normal((5, dim)).astype(np.float32)
    codes = o.encode(X)
    d32, d64 = o.m_to_n(Q, codes), o.m_to_n_f64(Q, codes)
    qns = np.sum(Q.astype(np.float64) ** 2, 1)
    xns = o.trailer(codes)[:, 0].astype(np.float64) ** 2
    assert scale_rel_err(d32, d64, qns, xns).max() <= 1e-5
    s32, s64 = o.symmetric(codes[:7], codes), o.symmetric_f64(codes[:7], codes)
    assert scale_rel_err(s32, s64, xns[:7], xns).max() <= 1e-5


@pytest.mark.parametrize("dim,bits", [(128, 4), (768, 8), (1000, 4), (512, 3), (33, 7)])
def test_fp64_symmetric_full_light_pinned_to_reference_fixture(dim, bits):
    # The FP64 restatements of the QJL-corrected symmetric estimators against
    # the reference build's fp32 values (full and its prepared form, which is
    # the same quantity, and light). A planted mistake (correction terms
    # dropped, i.e. the MSE-only estimator) must miss the full fixture.
    o = Oracle(dim, bits)
    tag = f"{dim}_{bits}"
    codes = GOLD[f"codes_{tag}"]
    A = codes[-4:]
    na = o.trailer(A)[:, 0].astype(np.float64) ** 2
    xns = o.trailer(codes)[:, 0].astype(np.float64) ** 2
    full, light = o.symmetric_full_f64(A, codes), o.symmetric_full_f64(A, codes, light=True)
    assert scale_rel_err(GOLD[f"symfull_{tag}"], full, na, xns).max() <= 1e-5
    assert scale_rel_err(GOLD[f"symfullprep_{tag}"], full, na, xns).max() <= 1e-5
    assert scale_rel_err(GOLD[f"symlight_{tag}"], light, na, xns).max() <= 1e-5
    mse_only = o.symmetric_f64(A, codes)
    assert scale_rel_err(GOLD[f"symfull_{tag}"], mse_only, na, xns).max() > 1e-4


def test_estimator_approximates_l2():
    # Reference test semantics (python/tests/test_turboquant.py:13-27).
    for dim in (32, 128, 1024):
        for bits in (4, 8):
            o = Oracle(dim, bits)
            rng = np.random.default_rng(dim * 10 + bits)
            q = rng.standard_normal((1, dim)).astype(np.float32)
            x = rng.standard_normal((1, dim)).astype(np.float32)
            tq = o.m_to_n(q, o.encode(x))[0, 0]
            brute = float(np.sum((q - x).astype(np.float64) ** 2))
            tol = 0.25 if bits >= 8 else 0.5
            assert abs(tq - brute) < tol * brute + 1.0


@pytest.mark.skipif(not ref_available(), reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("dim,bits", [(128, 4), (768, 8), (1000, 4), (96, 4), (1536, 4), (11, 9)])
def test_live_against_reference_build(dim, bits):
    o = Oracle(dim, bits)
    R = rlib()
    h = R.tqref_create(dim, bits, 42, 137, 1)
    try:
        rng = np.random.default_rng(77 + dim)
        X = rng.standard_normal((2000, dim)).astype(np.float32)
        X[::7] *= 1e-4
        want = np.zeros((len(X), o.cs), np.uint8)
        R.tqref_encode_batch(h, ptr(X), len(X), ptr(want))
        np.testing.assert_array_equal(o.encode(X), want)
        Q = rng.standard_normal((4, dim)).astype(np.float32)
        dref = np.zeros((4, len(X)), np.float32)
        R.tqref_distance_m_to_n(h, ptr(Q), 4, ptr(want), len(X), ptr(dref))
        qns = np.sum(Q.astype(np.float64) ** 2, 1)
        xns = o.trailer(want)[:, 0].astype(np.float64) ** 2
        assert scale_rel_err(o.m_to_n(Q, want), dref, qns, xns).max() <= 1e-5
    finally:
        R.tqref_destroy(h)
