"""Recall of quantized search against exact search.

The reference measures search quality as recall@k: the fraction of the exact
k nearest neighbours (squared L2 over the raw float vectors, ranked with
numpy argpartition + argsort) that a quantized ranking returns in its first
k (python/tests/test_turboquant.py:263-277 and :651-675;
python/benchmarks/run_benchmark.py:251-278). These helpers compute the same
quantity: `recall_at_k` is the reference's definition, `topk_from_distances`
its ranking of a distance matrix, and `exact_topk` the exact neighbours by an
fp64 brute force on the GPU (for collections too large for numpy).
"""
import numpy as np


def recall_at_k(pred, gt, k):
    """Mean over queries of |pred[i, :k] & gt[i, :k]| / k."""
    pred = np.asarray(pred)
    gt = np.asarray(gt)
    if pred.shape[0] != gt.shape[0]:
        raise ValueError("pred and gt must have the same number of queries")
    hits = 0
    for i in range(pred.shape[0]):
        hits += len(np.intersect1d(pred[i, :k], gt[i, :k]))
    return hits / (pred.shape[0] * k)


def topk_from_distances(dist, k):
    """(m, k) int64 ids of the k smallest entries of each row, ascending."""
    dist = np.asarray(dist)
    out = np.empty((dist.shape[0], k), np.int64)
    for i in range(dist.shape[0]):
        idx = np.argpartition(dist[i], k)[:k]
        out[i] = idx[np.argsort(dist[i, idx], kind="stable")]
    return out


def exact_topk(queries, base, k, chunk=1 << 18):
    """Exact k nearest neighbours by squared L2, fp64, on the device of
    `base` (torch tensors; `base` [n, dim], `queries` [m, dim]). Returns an
    (m, k) int64 torch tensor on that device, ascending by distance."""
    import torch
    q = queries.to(base.device, torch.float64)
    qn = (q * q).sum(1, keepdim=True)
    best_d = None
    best_i = None
    for s0 in range(0, base.shape[0], chunk):
        x = base[s0:s0 + chunk].to(torch.float64)
        d = qn + (x * x).sum(1)[None, :] - 2.0 * (q @ x.T)
        kk = min(k, d.shape[1])
        dv, di = torch.topk(d, kk, dim=1, largest=False)
        di = di + s0
        if best_d is None:
            best_d, best_i = dv, di
        else:
            cd = torch.cat([best_d, dv], 1)
            ci = torch.cat([best_i, di], 1)
            kk = min(k, cd.shape[1])
            best_d, sel = torch.topk(cd, kk, dim=1, largest=False)
            best_i = torch.gather(ci, 1, sel)
    order = torch.argsort(best_d, dim=1, stable=True)
    return torch.gather(best_i, 1, order)
