"""Error of the GPU distance forms on samples of each benchmark workload.

For each workload's data recipe (dims, bits, Gaussian-mixture parameters and
seeds as the workloads define them; samples small enough for the FP64
restatement to finish in seconds) this prints the max and mean of
|d - d64| / (||q||^2 + ||x||^2) for the MFMA form and the plain fp32 form
against the FP64 restatement of the estimator, and the top-10 overlap with the
FP64 ranking. Bars: MFMA <= 1e-4; fp32 within 1e-5 beyond the fp32 oracle's
own rounding (fp32_parity_excess); the fp32 top-10 ids equal the fp32
oracle's on every query whose oracle ranking has no near-tie at rank 10.

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

from _oracle import Oracle, fp32_parity_excess, scale_rel_err

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


def topk(d, k):
    idx = np.argpartition(d, k, axis=1)[:, :k]
    return [set(r) for r in idx]


def ranking_checks(d_gpu32, d_o32, d64, bound):
    """(queries with differing fp32 top-10 ids that have no near-tie,
    mean top-10 overlap of d_gpu32 with the FP64 ranking)."""
    g, o, e = topk(d_gpu32, K), topk(d_o32, K), topk(d64, K)
    bad = 0
    for i in range(d_o32.shape[0]):
        if g[i] == o[i]:
            continue
        srt = np.sort(np.partition(d_o32[i], K)[: K + 1])
        if srt[K] - srt[K - 1] > 2.0 * bound[i]:
            bad += 1
    overlap = np.mean([len(a & b) / K for a, b in zip(g, e)])
    return bad, overlap


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
    e_m = scale_rel_err(got["mfma"], d64, qns, xns)
    e_f = scale_rel_err(got["fp32"], d64, qns, xns)
    excess = fp32_parity_excess(got["fp32"], d32o, d64, qns, xns)
    bad, ov32 = ranking_checks(got["fp32"], d32o, d64, FP32_TOL * (qns + xns.max()))
    _, ovm = ranking_checks(got["mfma"], d32o, d64, FP32_TOL * (qns + xns.max()))
    row = (f"{label:34s} m={len(Q):3d} n={len(X):6d}  mfma max {e_m.max():.2e} mean {e_m.mean():.2e}"
           f"  fp32 max {e_f.max():.2e} mean {e_f.mean():.2e}  top-10 overlap with fp64:"
           f" mfma {ovm:.4f} fp32 {ov32:.4f}")
    TABLE.append(row)
    say(row)
    assert e_m.max() <= MFMA_TOL
    assert excess.max() <= FP32_TOL
    assert bad == 0, f"{bad} queries with differing fp32 top-{K} ids and no near-tie"


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
    e = scale_rel_err(got, d64, an, bn)
    e_m = scale_rel_err(got_m, d64, an, bn)
    excess = fp32_parity_excess(got, d32o, d64, an, bn)
    row = (f"{label:34s} m={len(A):4d} n={len(B):6d}  symmetric fp32 max {e.max():.2e}"
           f" mean {e.mean():.2e}  mfma max {e_m.max():.2e} mean {e_m.mean():.2e}")
    TABLE.append(row)
    say(row)
    assert excess.max() <= FP32_TOL
    assert e_m.max() <= MFMA_TOL
    return got, d64


def test_w1_dim128_4bit(gpu):
    say("[gpu] workload errors w1 dim 128 4-bit")
    X = np.random.default_rng(0).standard_normal((20_000, 128)).astype(np.float32)
    Q = np.random.default_rng(1).standard_normal((16, 128)).astype(np.float32)
    check_asym(gpu, "w1 128/4b iid N(0,1)", 128, 4, X, Q)
    # symmetric on 1,000 random pairs: the pair (i, j) is entry [p, p] of
    # the matrix between the two picked row sets
    rng = np.random.default_rng(5)
    ia, ib = rng.integers(0, len(X), 1000), rng.integers(0, len(X), 1000)
    check_sym(gpu, "w1 128/4b symmetric pairs", 128, 4, X[ia], X[ib])


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
