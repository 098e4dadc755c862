"""Error of the GPU distance forms on samples of each benchmark workload.

For each workload's data recipe (dims, bits, Gaussian-mixture parameters and
seeds as the workloads define them; samples small enough for the FP64
restatement to finish in seconds) this prints, against the FP64 restatement
of the estimator:
  * MFMA: the relative error |d - d64| / max(d64, eps) on each query's 10
    nearest codes by d64 (the distances that decide a top-10) and over all
    pairs, eps = 1e-3 (||q||^2 + ||x||^2) (exact duplicates have d64 = 0), and
    the scale-relative error |d - d64| / (||q||^2 + ||x||^2) over all pairs;
  * fp32: the relative error against the fp32 oracle (the reference's
    arithmetic) on the same nearest codes, how many of those exceed 1e-5 and
    how many of them are excused because the GPU value is within 1e-5
    relative of d64 (the reference's own fp32 rounding sets them apart; its
    error against d64 is printed beside);
  * the top-10 overlap with the FP64 ranking, and the queries whose top-10
    ids differ (fp32 vs the fp32 oracle, MFMA vs FP64), each of which must be
    a near-tie at rank 10.
Bars: MFMA <= 1e-4 in all three measures; fp32 within 1e-5 relative of the
reference on the nearest codes except excused entries (counted), and within
1e-5 of scale beyond the oracle's own rounding (fp32_parity_excess).

Workloads (queries drawn from the same mixture with the next seed):
  w1  dim 128, 4-bit, iid N(0,1) rows (seed 0), 16 queries (seed 1);
      symmetric on 1,000 code pairs
  w2  dim 768, 8-bit, 1,024 centres, std 0.3, L2-normalised (seeds 2, 3)
  w3  dim 1000, 4-bit, 4,096 centres, std 0.5, not normalised (seed 4):
      symmetric
  w4  batched collections (96,4) (128,8) (384,4) (768,8) (1536,4), 512
      centres, std 0.3, normalised (seeds 10-14)
  w5  dim 768, 4-bit, 8,192 centres, std 0.3, normalised (seed 20)
"""
import sys

import numpy as np
import pytest

from _oracle import (EPS_FLOOR, Oracle, eps_rel_err, fp32_parity_excess, nearest,
                     scale_rel_err, topk_rel_err)

pytestmark = pytest.mark.gpu

MFMA_TOL = 1e-4
FP32_TOL = 1e-5
K = 10
TABLE = []


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def mixture(n, dim, k, std, seed, normalise, centres=None):
    rng = np.random.default_rng(seed)
    if centres is None:
        centres = rng.standard_normal((k, dim)).astype(np.float32)
    lab = rng.integers(0, k, n)
    x = centres[lab] + np.float32(std) * rng.standard_normal((n, dim)).astype(np.float32)
    if normalise:
        x /= np.linalg.norm(x, axis=1, keepdims=True)
    return np.ascontiguousarray(x, np.float32), centres


def tie_mismatches(got, ref, err_bound):
    """(queries whose top-K ids under `got` differ from those under `ref`,
    how many of them are not near-ties: rank K and K+1 of `ref` further apart
    than 2 * err_bound)."""
    g, r = nearest(got, K), nearest(ref, K)
    diff = bad = 0
    for i in range(ref.shape[0]):
        if set(g[i]) == set(r[i]):
            continue
        diff += 1
        srt = np.sort(np.partition(ref[i], K)[: K + 1])
        if srt[K] - srt[K - 1] > 2.0 * err_bound[i]:
            bad += 1
    return diff, bad


def overlap(got, ref):
    g, r = nearest(got, K), nearest(ref, K)
    return float(np.mean([len(set(a) & set(b)) / K for a, b in zip(g, r)]))


def fp32_rel(got, d32o, d64, idx, floor):
    """fp32 form on the nearest codes: (max relative error against the
    reference's arithmetic, entries above FP32_TOL, of those the excused ones,
    the reference's own max relative error against d64). An entry is excused
    when the GPU value is within FP32_TOL relative of the exact estimator d64:
    the reference's own fp32 sums are then what sets the two apart."""
    e = topk_rel_err(got, d32o, idx=idx, floor=floor)
    g = np.take_along_axis(got.astype(np.float64), idx, axis=1)
    o = np.take_along_axis(d32o.astype(np.float64), idx, axis=1)
    x = np.take_along_axis(d64, idx, axis=1)
    den = np.maximum(x, np.take_along_axis(floor, idx, axis=1))
    over = e > FP32_TOL
    excused = over & (np.abs(g - x) <= FP32_TOL * den)
    return e.max(), int(over.sum()), int(excused.sum()), float((np.abs(o - x) / den).max())


def report(label, shape, d_m, d_f, d32o, d64, qns, xns):
    """Measures both forms, records a table row and asserts the bars."""
    idx = nearest(d64, K)
    floor = EPS_FLOOR * (np.asarray(qns)[:, None] + np.asarray(xns)[None, :])
    m_top = topk_rel_err(d_m, d64, idx=idx, floor=floor)
    m_eps = eps_rel_err(d_m, d64, qns, xns)
    m_scale = scale_rel_err(d_m, d64, qns, xns)
    f_ref_max, f_over, f_excused, ref64 = fp32_rel(d_f, d32o, d64, idx, floor)
    f_top64 = topk_rel_err(d_f, d64, idx=idx, floor=floor)
    excess = fp32_parity_excess(d_f, d32o, d64, qns, xns)
    bound_f = FP32_TOL * (qns + xns.max())
    bound_m = MFMA_TOL * (qns + xns.max())
    f_diff, f_bad = tie_mismatches(d_f, d32o, bound_f)
    m_diff, m_bad = tie_mismatches(d_m, d64, bound_m)
    row = (f"{label:36s} {shape:18s} | MFMA top-10 rel max {m_top.max():.2e} mean {m_top.mean():.2e},"
           f" eps-rel max {m_eps.max():.2e}, scale-rel max {m_scale.max():.2e} mean"
           f" {m_scale.mean():.2e} | fp32 top-10 rel vs ref max {f_ref_max:.2e} (>1e-5: {f_over},"
           f" excused {f_excused}), vs fp64 max {f_top64.max():.2e} (reference vs fp64 max"
           f" {ref64:.2e}) | top-10 overlap with fp64:"
           f" MFMA {overlap(d_m, d64):.4f} fp32 {overlap(d_f, d64):.4f} | near-tie id"
           f" mismatches: fp32 {f_diff}, MFMA {m_diff}")
    TABLE.append(row)
    say(row)
    assert m_top.max() <= MFMA_TOL
    assert m_eps.max() <= MFMA_TOL
    assert m_scale.max() <= MFMA_TOL
    assert f_over == f_excused, f"{f_over - f_excused} fp32 entries off the reference and fp64 by >1e-5"
    assert excess.max() <= FP32_TOL
    assert f_bad == 0, f"{f_bad} queries with differing fp32 top-{K} ids and no near-tie"
    assert m_bad == 0, f"{m_bad} queries with differing MFMA top-{K} ids and no near-tie"


def gpu_forms(gpu, Q, codes, dim, bits):
    import torch
    capi = gpu
    sp = capi.Space(dim, bits)
    dQ = torch.from_numpy(Q).cuda()
    dC = torch.from_numpy(codes).cuda()
    out = {}
    for name, mode in (("fp32", capi.TQ_MODE_FP32), ("mfma", capi.TQ_MODE_MFMA)):
        dO = torch.empty((len(Q), len(codes)), dtype=torch.float32, device="cuda")
        sp.m_to_n(dQ, len(Q), dC, len(codes), dO, mode=mode)
        torch.cuda.synchronize()
        out[name] = dO.cpu().numpy()
    sp.close()
    return out


def check_asym(gpu, label, dim, bits, X, Q):
    o = Oracle(dim, bits)
    codes = o.encode(X)
    d64 = o.m_to_n_f64(Q, codes)
    d32o = o.m_to_n(Q, codes)
    qns = np.sum(Q.astype(np.float64) ** 2, 1)
    xns = o.trailer(codes)[:, 0].astype(np.float64) ** 2
    got = gpu_forms(gpu, Q, codes, dim, bits)
    report(label, f"m={len(Q)} n={len(X)}", got["mfma"], got["fp32"], d32o, d64, qns, xns)


def check_sym(gpu, label, dim, bits, A, B):
    import torch
    capi = gpu
    o = Oracle(dim, bits)
    ca, cb = o.encode(A), o.encode(B)
    d64 = o.symmetric_f64(ca, cb)
    d32o = o.symmetric(ca, cb)
    an = o.trailer(ca)[:, 0].astype(np.float64) ** 2
    bn = o.trailer(cb)[:, 0].astype(np.float64) ** 2
    sp = capi.Space(dim, bits)
    dA, dB = torch.from_numpy(ca).cuda(), torch.from_numpy(cb).cuda()
    dO = torch.empty((len(ca), len(cb)), dtype=torch.float32, device="cuda")
    sp.symmetric(dA, len(ca), dB, len(cb), dO)
    torch.cuda.synchronize()
    got = dO.cpu().numpy()
    sp.symmetric(dA, len(ca), dB, len(cb), dO, mode=capi.TQ_MODE_MFMA)
    torch.cuda.synchronize()
    got_m = dO.cpu().numpy()
    sp.close()
    report(label + " (symmetric)", f"m={len(A)} n={len(B)}", got_m, got, d32o, d64, an, bn)
    return got, d64


def test_w1_dim128_4bit(gpu):
    """The CPU-runnable workload at full size: 100,000 x 128 iid N(0,1) rows
    (seed 0), 16 queries from seed 1, then 1,000 code pairs drawn from the
    same generator after the queries."""
    say("[gpu] workload errors w1 dim 128 4-bit (full size)")
    X = np.random.default_rng(0).standard_normal((100_000, 128)).astype(np.float32)
    rng = np.random.default_rng(1)
    Q = rng.standard_normal((16, 128)).astype(np.float32)
    pairs = rng.integers(0, 100_000, (1000, 2))
    check_asym(gpu, "w1 128/4b iid N(0,1)", 128, 4, X, Q)
    # The pair (i, j) is entry [p, p] of the matrix between the picked rows.
    got, d64 = check_sym(gpu, "w1 128/4b pairs", 128, 4, X[pairs[:, 0]], X[pairs[:, 1]])
    diag = np.abs(np.diag(got).astype(np.float64) - np.diag(d64)) / np.maximum(np.diag(d64), 1e-30)
    say(f"  distance_symmetric on the 1,000 pairs: fp32 max rel err vs fp64 {diag.max():.2e}")


def test_w2_dim768_8bit(gpu):
    say("[gpu] workload errors w2 dim 768 8-bit")
    X, c = mixture(20_000, 768, 1024, 0.3, 2, True)
    Q, _ = mixture(64, 768, 1024, 0.3, 3, True, centres=c)
    check_asym(gpu, "w2 768/8b mixture 1024c std0.3 norm", 768, 8, X, Q)


def test_w3_dim1000_4bit_symmetric(gpu):
    say("[gpu] workload errors w3 dim 1000 4-bit symmetric")
    X, c = mixture(20_000, 1000, 4096, 0.5, 4, False)
    A, _ = mixture(64, 1000, 4096, 0.5, 5, False, centres=c)
    check_sym(gpu, "w3 1000/4b mixture 4096c std0.5", 1000, 4, A, X)


@pytest.mark.parametrize("i,dim,bits", [(0, 96, 4), (1, 128, 8), (2, 384, 4), (3, 768, 8), (4, 1536, 4)])
def test_w4_batched_collections(gpu, i, dim, bits):
    say(f"[gpu] workload errors w4 collection {i} dim {dim} bits {bits}")
    X, c = mixture(10_000, dim, 512, 0.3, 10 + i, True)
    Q, _ = mixture(32, dim, 512, 0.3, 100 + i, True, centres=c)
    check_asym(gpu, f"w4 {dim}/{bits}b mixture 512c std0.3 norm", dim, bits, X, Q)


def test_w5_dim768_4bit(gpu):
    say("[gpu] workload errors w5 dim 768 4-bit")
    X, c = mixture(20_000, 768, 8192, 0.3, 20, True)
    Q, _ = mixture(64, 768, 8192, 0.3, 21, True, centres=c)
    check_asym(gpu, "w5 768/4b mixture 8192c std0.3 norm", 768, 4, X, Q)
    say("workload error table:\n" + "\n".join(TABLE))
