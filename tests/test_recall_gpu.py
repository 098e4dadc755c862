"""Recall harness on the GPU: recall@10 of every quantized form against the
exact neighbours, on the reference's own recall setup (dim 128, 5000 base
and 100 query vectors of uniform integers in [0, 256), seed 42, bits 4 and 8;
python/tests/test_turboquant.py:286-296 and :684-690).

Parity: for every query whose oracle ranking has a clear gap at the k-th
place (the k-th and (k+1)-th oracle distances differ by more than twice the
form's error bound), the GPU's top-k set equals the oracle's, so recall
computed from GPU distances equals recall computed from the oracle's except
on near-ties. Error bounds: fp32 forms 1e-5 of (||q||^2 + ||x||^2), MFMA
forms 1e-4 of the same scale against the FP64 restatement. The fused top-k
search must return exactly the (distance, id)-ordered top-k of the MFMA
distance matrix. Recall values are printed (the reference's tests print
them too and assert no threshold).
"""
import sys

import numpy as np
import pytest

from _oracle import Oracle, olib, ptr

pytestmark = pytest.mark.gpu

K = 10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ref_setup():
    rng = np.random.default_rng(42)
    base = rng.integers(0, 256, size=(5000, 128)).astype(np.float32)
    queries = rng.integers(0, 256, size=(100, 128)).astype(np.float32)
    return base, queries


def brute_topk(queries, base, k):
    q = queries.astype(np.float64)
    x = base.astype(np.float64)
    d = (q * q).sum(1)[:, None] + (x * x).sum(1)[None, :] - 2.0 * q @ x.T
    from turboquant.recall import topk_from_distances
    return topk_from_distances(d, k)


def lex_topk(d, k):
    """Top-k ids by (distance, id), the fused search's order."""
    out = np.empty((d.shape[0], k), np.int64)
    ids = np.arange(d.shape[1])
    for i in range(d.shape[0]):
        out[i] = np.lexsort((ids, d[i]))[:k]
    return out


def check_unambiguous(got, want, k, bound_row):
    """Rows with a clear gap at rank k in `want` have equal top-k sets."""
    from turboquant.recall import topk_from_distances
    gk = topk_from_distances(got, k)
    wk = topk_from_distances(want, k)
    checked = 0
    for i in range(want.shape[0]):
        srt = np.partition(want[i], k)[: k + 1]
        srt.sort()
        if srt[k] - srt[k - 1] <= 2.0 * bound_row[i]:
            continue
        checked += 1
        assert set(gk[i]) == set(wk[i]), f"row {i}: top-{k} sets differ"
    return checked


@pytest.mark.parametrize("bits", [4, 8])
def test_recall_parity_reference_setup(gpu, bits):
    import turboquant
    from turboquant.recall import recall_at_k, topk_from_distances
    say(f"[gpu] recall harness bits={bits}")
    capi = gpu
    base, queries = ref_setup()
    gt = brute_topk(queries, base, K)
    s = turboquant.TurboQuantSpace(128, bits_per_coord=bits)
    o = Oracle(128, bits)
    codes = s.encode_batch(base)
    np.testing.assert_array_equal(codes, o.encode(base))
    qcodes = s.encode_batch(queries)
    qns = (queries.astype(np.float64) ** 2).sum(1)
    xns = o.trailer(codes)[:, 0].astype(np.float64) ** 2
    an = o.trailer(qcodes)[:, 0].astype(np.float64) ** 2
    b32 = 1e-5 * (qns + xns.max())
    b16 = 1e-4 * (qns + xns.max())
    bsym = 1e-5 * (an + xns.max())

    # Asymmetric, fp32 path against the oracle's fp32 distances.
    d32 = s.distance_m_to_n(queries, codes)
    o32 = o.m_to_n(queries, codes)
    n1 = check_unambiguous(d32, o32, K, b32)
    r_asym = recall_at_k(topk_from_distances(d32, K), gt, K)
    r_asym_o = recall_at_k(topk_from_distances(o32, K), gt, K)

    # Asymmetric, MFMA path against the FP64 restatement.
    dmf = s.distance_m_to_n(queries, codes, mode="mfma")
    o64 = o.m_to_n_f64(queries, codes)
    n2 = check_unambiguous(dmf, o64, K, b16)
    r_mfma = recall_at_k(topk_from_distances(dmf, K), gt, K)

    # Fused top-k search (MFMA) == (distance, id) top-k of the MFMA matrix.
    import torch
    dev = torch.device("cuda")
    Qd = torch.from_numpy(queries).to(dev)
    Cd = torch.from_numpy(codes).to(dev)
    od = torch.empty((len(queries), K), dtype=torch.float32, device=dev)
    oi = torch.empty((len(queries), K), dtype=torch.int64, device=dev)
    sp = capi.Space(128, bits)
    plan = capi.SearchPlan([(sp, Qd, len(queries), Cd, len(codes), od, oi)], K)
    plan.run()
    torch.cuda.synchronize()
    fused = oi.cpu().numpy()
    np.testing.assert_array_equal(fused, lex_topk(dmf, K))
    r_fused = recall_at_k(fused, gt, K)
    plan.close()
    sp.close()

    # Symmetric MSE-only, against the oracle.
    ds = s.distance_m_to_n_symmetric(qcodes, codes)
    os_ = o.symmetric(qcodes, codes)
    n3 = check_unambiguous(ds, os_, K, bsym)
    r_sym = recall_at_k(topk_from_distances(ds, K), gt, K)

    # Symmetric light / full (QJL-corrected), first 25 queries against the
    # oracle's pairwise functions.
    L = olib()
    A = qcodes[:25]
    res = {}
    for name, fn, meth in (("light", L.tqo_distance_symmetric_light, s.distance_m_to_n_symmetric_light),
                           ("full", L.tqo_distance_symmetric_full, s.distance_m_to_n_symmetric_full)):
        got = meth(A, codes)
        want = np.array([[fn(o.h, ptr(a), ptr(b)) for b in codes] for a in A], np.float32)
        check_unambiguous(got, want, K, bsym[:25])
        res[name] = recall_at_k(topk_from_distances(got, K), gt[:25], K)
    say(f"  recall@{K}: asym fp32 {r_asym:.4f} (oracle {r_asym_o:.4f}), asym mfma {r_mfma:.4f}, "
        f"fused search {r_fused:.4f}, sym {r_sym:.4f}; first 25 queries: light {res['light']:.4f}, "
        f"full {res['full']:.4f}; unambiguous rows checked {n1}/{n2}/{n3} of 100")
    assert n1 >= 90 and n3 >= 90
