"""Top-k search on the GPU (tq_search_plan_*, tq_search, tq_topk_rows).

Bars:
  * the MFMA search, in both forms (fused persistent launch; blocked distance
    blocks + row top-k), returns exactly the top-k of the full MFMA distance
    matrix (same distances bit for bit, ids ordered by (distance, id)), and
    the two forms agree with each other;
  * the plain fp32 search returns exactly the top-k of the full fp32 matrix,
    and its ids agree with the oracle's top-k except where the oracle's own
    distances are within the fp32 tolerance of each other (near-ties);
  * a query's result is the same alone and inside a mixed multi-job plan;
  * graph replays read the job buffers when they run and equal eager runs;
  * empty collections and k > n give (+inf, -1) entries.
"""
import sys

import numpy as np
import pytest

from _oracle import Oracle

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    assert torch.cuda.is_available()
    return torch


def gm(rng, n, dim, centres, std=0.3):
    lab = rng.integers(0, len(centres), n)
    X = centres[lab] + std * rng.standard_normal((n, dim)).astype(np.float32)
    return (X / np.linalg.norm(X, axis=1, keepdims=True)).astype(np.float32)


def host_topk(D, k):
    """(distance, id) order per row, +inf / -1 past the row's length."""
    m, n = D.shape
    od = np.full((m, k), np.inf, np.float32)
    oi = np.full((m, k), -1, np.int64)
    for r in range(m):
        order = np.lexsort((np.arange(n), D[r]))[:k]
        od[r, : len(order)] = D[r, order]
        oi[r, : len(order)] = order
    return od, oi


class Coll:
    """A collection on the device: space, codes, plus its queries."""

    def __init__(self, torch, capi, dim, bits, n, seed, dup=0):
        self.sp = capi.Space(dim, bits)
        rng = np.random.default_rng(seed)
        centres = rng.standard_normal((64, dim)).astype(np.float32)
        X = gm(rng, n, dim, centres)
        if dup:  # exact duplicates -> exact distance ties
            X[n - dup:] = X[:dup]
        self.X = X
        self.centres = centres
        self.n = n
        dX = torch.from_numpy(X).cuda()
        self.codes = torch.empty((max(n, 1), self.sp.cs), dtype=torch.uint8, device="cuda")
        if n:
            self.sp.encode(dX, n, self.codes)
        self.rng = rng

    def queries(self, torch, m):
        Q = gm(self.rng, m, self.sp.dim, self.centres)
        return Q, torch.from_numpy(Q).cuda()


def full_matrix(torch, capi, c, dQ, m, mode):
    out = torch.empty((max(m, 1), max(c.n, 1)), dtype=torch.float32, device="cuda")
    if m and c.n:
        c.sp.m_to_n(dQ, m, c.codes, c.n, out, mode=mode)
    torch.cuda.synchronize()
    return out[:m, : c.n].cpu().numpy()


def outputs(torch, m, k):
    return (torch.empty((max(m, 1), k), dtype=torch.float32, device="cuda"),
            torch.empty((max(m, 1), k), dtype=torch.int64, device="cuda"))


CASES = [(128, 4), (96, 4), (768, 4), (1000, 4), (128, 8), (384, 8), (1536, 4)]


@pytest.mark.parametrize("dim,bits", CASES)
@pytest.mark.parametrize("form,qt", [(1, 32), (1, 64), (2, 0)])
def test_mfma_search_equals_topk_of_mfma_matrix(gpu, torch_cuda, dim, bits, form, qt):
    # form 1: the fused persistent launch with 32- or 64-query tiles;
    # form 2: blocked (distance blocks + row top-k)
    torch, capi = torch_cuda, gpu
    n = 3000 if dim < 1000 else 1300
    say(f"[gpu] MFMA search form={form} dim={dim} bits={bits} qt={qt} n={n}")
    c = Coll(torch, capi, dim, bits, n, seed=dim + bits, dup=200)
    for m in (1, 37, 70):
        Q, dQ = c.queries(torch, m)
        k = 10
        od, oi = outputs(torch, m, k)
        plan = capi.SearchPlan([(c.sp, dQ, m, c.codes, n, od, oi)], k, capi.TQ_MODE_MFMA,
                               query_tile=qt, items_per_wg=2 if qt else 0, form=form)
        plan.run()
        torch.cuda.synchronize()
        want_d, want_i = host_topk(full_matrix(torch, capi, c, dQ, m, capi.TQ_MODE_MFMA), k)
        got_d, got_i = od[:m].cpu().numpy(), oi[:m].cpu().numpy()
        np.testing.assert_array_equal(got_i, want_i)
        np.testing.assert_array_equal(got_d.view(np.uint32), want_d.view(np.uint32))
        plan.close()


@pytest.mark.parametrize("dim,bits", [(768, 4), (768, 8)])
def test_coarse_search_over_chunks_equals_topk_of_mfma_matrix(gpu, torch_cuda, dim, bits):
    # Coarse-form collections (padded dim 1024) of 64 clusters, so every query
    # has thousands of near pairs. 16 MB distance blocks split the blocked form
    # into 2-3 chunks: after the first, a block corrects only the near pairs
    # that can still beat the running k-th best; the fused form prunes against
    # its lists. Both must return the top-k of the fully corrected matrix.
    torch, capi = torch_cuda, gpu
    n = 150_000
    c = Coll(torch, capi, dim, bits, n, seed=4000 + bits, dup=500)
    for m in (40, 64):
        for form in (capi.TQ_FORM_BLOCKED, capi.TQ_FORM_FUSED):
            say(f"[gpu] coarse search over chunks dim={dim} bits={bits} m={m} form={form}")
            Q, dQ = c.queries(torch, m)
            k = 10
            od, oi = outputs(torch, m, k)
            plan = capi.SearchPlan([(c.sp, dQ, m, c.codes, n, od, oi)], k, capi.TQ_MODE_MFMA,
                                   form=form, block_mb=16)
            plan.run()
            torch.cuda.synchronize()
            want_d, want_i = host_topk(full_matrix(torch, capi, c, dQ, m, capi.TQ_MODE_MFMA), k)
            np.testing.assert_array_equal(oi[:m].cpu().numpy(), want_i)
            np.testing.assert_array_equal(od[:m].cpu().numpy().view(np.uint32),
                                          want_d.view(np.uint32))
            plan.close()


@pytest.mark.parametrize("form", [1, 2])
@pytest.mark.parametrize("k", [1, 16])
def test_mfma_search_k_bounds(gpu, torch_cuda, form, k):
    # k = 1 and k = 16 (the list length): the blocked form's running bound is
    # then the K-th entry of a full list, and k = 1 pops a single entry.
    torch, capi = torch_cuda, gpu
    say(f"[gpu] MFMA search form={form} k={k}")
    n, m = 9000, 33
    c = Coll(torch, capi, 128, 4, n, seed=90 + k + form, dup=300)
    Q, dQ = c.queries(torch, m)
    od, oi = outputs(torch, m, k)
    plan = capi.SearchPlan([(c.sp, dQ, m, c.codes, n, od, oi)], k, capi.TQ_MODE_MFMA, form=form)
    plan.run()
    torch.cuda.synchronize()
    want_d, want_i = host_topk(full_matrix(torch, capi, c, dQ, m, capi.TQ_MODE_MFMA), k)
    np.testing.assert_array_equal(oi[:m].cpu().numpy(), want_i)
    np.testing.assert_array_equal(od[:m].cpu().numpy().view(np.uint32), want_d.view(np.uint32))
    plan.close()


@pytest.mark.parametrize("dim,bits", [(128, 4), (768, 4), (128, 8), (64, 2)])
def test_fp32_search_equals_topk_of_fp32_matrix_and_oracle(gpu, torch_cuda, dim, bits):
    torch, capi = torch_cuda, gpu
    n, m, k = 2500, 9, 10
    say(f"[gpu] fp32 search dim={dim} bits={bits}")
    c = Coll(torch, capi, dim, bits, n, seed=7 * dim + bits, dup=100)
    Q, dQ = c.queries(torch, m)
    od, oi = outputs(torch, m, k)
    plan = capi.SearchPlan([(c.sp, dQ, m, c.codes, n, od, oi)], k, capi.TQ_MODE_FP32)
    plan.run()
    torch.cuda.synchronize()
    full = full_matrix(torch, capi, c, dQ, m, capi.TQ_MODE_FP32)
    want_d, want_i = host_topk(full, k)
    got_d, got_i = od[:m].cpu().numpy(), oi[:m].cpu().numpy()
    np.testing.assert_array_equal(got_i, want_i)
    np.testing.assert_array_equal(got_d.view(np.uint32), want_d.view(np.uint32))
    # Against the oracle's distances: the same ids except near-ties.
    o = Oracle(dim, bits)
    codes = c.codes[:n].cpu().numpy()
    ref = o.m_to_n(Q, codes)
    scale = (Q.astype(np.float64) ** 2).sum(1)[:, None] + 1.0
    ref_d, ref_i = host_topk(ref, k)
    for r in range(m):
        kth = ref_d[r, -1]
        for j in set(got_i[r]) ^ set(ref_i[r]):
            assert abs(ref[r, j] - kth) <= 2 * FP32_TOL * scale[r, 0], (r, j)


def test_mixed_plan_equals_each_job_alone(gpu, torch_cuda):
    """Jobs over several collections (dims, bit widths), uneven m, two jobs
    on one collection, an empty job and an empty collection, in one plan."""
    torch, capi = torch_cuda, gpu
    say("[gpu] mixed plan vs alone")
    colls = [Coll(torch, capi, 96, 4, 2000, 10), Coll(torch, capi, 128, 8, 1500, 11, dup=50),
             Coll(torch, capi, 384, 4, 2600, 12), Coll(torch, capi, 768, 8, 700, 13),
             Coll(torch, capi, 128, 4, 0, 14)]
    spec = [(0, 5), (1, 40), (2, 1), (0, 77), (3, 12), (4, 3), (2, 0), (1, 64), (3, 33)]
    k = 10
    jobs, alone, qs = [], [], []
    for ci, m in spec:
        c = colls[ci]
        Q, dQ = c.queries(torch, max(m, 1))
        od, oi = outputs(torch, m, k)
        jobs.append((c.sp, dQ, m, c.codes, c.n, od, oi))
        qs.append((ci, m, dQ))
    forms = {}
    for mode, form in ((capi.TQ_MODE_MFMA, capi.TQ_FORM_AUTO), (capi.TQ_MODE_MFMA, capi.TQ_FORM_FUSED),
                       (capi.TQ_MODE_FP32, capi.TQ_FORM_AUTO)):
        plan = capi.SearchPlan(jobs, k, mode, form=form)
        items, wgs = plan.stats()
        say(f"  mode={mode} form={form} items={items} workgroups={wgs}")
        plan.run()
        torch.cuda.synchronize()
        mixed = [(j[5][: j[2]].cpu().numpy(), j[6][: j[2]].cpu().numpy()) for j in jobs]
        forms[(mode, form)] = mixed
        for (ci, m, dQ), (md, mi) in zip(qs, mixed):
            c = colls[ci]
            od, oi = outputs(torch, m, k)
            p1 = capi.SearchPlan([(c.sp, dQ, m, c.codes, c.n, od, oi)], k, mode, form=form)
            p1.run()
            torch.cuda.synchronize()
            np.testing.assert_array_equal(oi[:m].cpu().numpy(), mi)
            np.testing.assert_array_equal(od[:m].cpu().numpy().view(np.uint32), md.view(np.uint32))
            want_d, want_i = host_topk(full_matrix(torch, capi, c, dQ, m, mode), k)
            np.testing.assert_array_equal(mi, want_i)
            if c.n == 0 and m:
                assert (mi == -1).all() and np.isinf(md).all()
            p1.close()
        plan.close()
    # The blocked and the fused form give the same bits.
    for (ad, ai), (fd, fi) in zip(forms[(capi.TQ_MODE_MFMA, capi.TQ_FORM_AUTO)],
                                  forms[(capi.TQ_MODE_MFMA, capi.TQ_FORM_FUSED)]):
        np.testing.assert_array_equal(ai, fi)
        np.testing.assert_array_equal(ad.view(np.uint32), fd.view(np.uint32))


def test_graph_replays_read_current_queries(gpu, torch_cuda):
    torch, capi = torch_cuda, gpu
    say("[gpu] graph replay")
    c1 = Coll(torch, capi, 768, 4, 4000, 21)
    c2 = Coll(torch, capi, 128, 8, 3000, 22)
    k = 10
    _, q1 = c1.queries(torch, 50)
    _, q2 = c2.queries(torch, 7)
    o1, o2 = outputs(torch, 50, k), outputs(torch, 7, k)
    jobs = [(c1.sp, q1, 50, c1.codes, 4000, *o1), (c2.sp, q2, 7, c2.codes, 3000, *o2)]
    for mode, form in ((capi.TQ_MODE_MFMA, capi.TQ_FORM_AUTO), (capi.TQ_MODE_MFMA, capi.TQ_FORM_FUSED),
                       (capi.TQ_MODE_FP32, capi.TQ_FORM_AUTO)):
        plan = capi.SearchPlan(jobs, k, mode, form=form)
        for rep in range(3):
            # New queries written in place between replays, with eager work
            # and a synchronisation in between.
            q1.copy_(torch.from_numpy(c1.queries(torch, 50)[0]).cuda())
            q2.copy_(torch.from_numpy(c2.queries(torch, 7)[0]).cuda())
            torch.cuda.synchronize()
            plan.run_graph()
            torch.cuda.synchronize()
            g = [t.clone() for t in (*o1, *o2)]
            for t in (*o1, *o2):
                t.fill_(0)
            plan.run()
            torch.cuda.synchronize()
            for a, b in zip(g, (*o1, *o2)):
                assert torch.equal(a, b), (mode, form, rep)
            want_d, want_i = host_topk(full_matrix(torch, capi, c1, q1, 50, mode), k)
            np.testing.assert_array_equal(o1[1].cpu().numpy(), want_i)
        plan.close()


def test_tuned_plan_gives_the_same_results(gpu, torch_cuda):
    torch, capi = torch_cuda, gpu
    say("[gpu] autotune")
    c = Coll(torch, capi, 384, 4, 6000, 31)
    _, dQ = c.queries(torch, 20)
    od, oi = outputs(torch, 20, 10)
    plan = capi.SearchPlan([(c.sp, dQ, 20, c.codes, c.n, od, oi)], 10, capi.TQ_MODE_MFMA)
    plan.run()
    torch.cuda.synchronize()
    before = (od.clone(), oi.clone())
    qt, ipw, form = plan.tune()
    assert qt in (32, 64) and ipw in (1, 2, 4, 8) and form in (capi.TQ_FORM_FUSED, capi.TQ_FORM_BLOCKED)
    t = plan.tuned
    assert t["wg_per_cu"] in (1, 2) and t["blocked_query_tile"] == 32  # 20 queries: one 32-row tile
    assert t["block_mb"] in (64, 128, 256)
    od.fill_(0)
    plan.run()
    torch.cuda.synchronize()
    assert torch.equal(before[0], od) and torch.equal(before[1], oi)
    plan.close()
    # A plan that forces the fused form takes the tuned fused knobs and gives
    # the same results.
    od.fill_(0)
    fused = capi.SearchPlan([(c.sp, dQ, 20, c.codes, c.n, od, oi)], 10, capi.TQ_MODE_MFMA,
                            form=capi.TQ_FORM_FUSED)
    fused.run()
    torch.cuda.synchronize()
    assert torch.equal(before[0], od) and torch.equal(before[1], oi)
    assert fused.stats()[0] > 0
    fused.close()


@pytest.mark.parametrize("knobs", [
    dict(form=2, blocked_query_tile=32, block_mb=16),   # 32-row MFMA tiles, 16 MiB blocks
    dict(form=2, blocked_query_tile=64, block_mb=64),
    dict(form=1, query_tile=64, items_per_wg=2, wg_per_cu=1),
    dict(form=1, query_tile=32, items_per_wg=8, wg_per_cu=2),
])
def test_every_knob_gives_the_same_bits(gpu, torch_cuda, knobs):
    # Each autotuner knob changes only how the work is cut: the top-k of a
    # plan built with it equals the default plan's bit for bit (70 queries:
    # three 32-row or two 64-row tiles; 130,000 codes: three chunks of a
    # 16 MiB block, 59,904 columns each).
    torch, capi = torch_cuda, gpu
    say(f"[gpu] knobs {knobs}")
    knobs = dict(knobs)
    knobs["form"] = capi.TQ_FORM_FUSED if knobs["form"] == 1 else capi.TQ_FORM_BLOCKED
    c = Coll(torch, capi, 768, 4, 130000, 43)
    _, dQ = c.queries(torch, 70)
    res = []
    for kw in ({}, knobs):
        od, oi = outputs(torch, 70, 10)
        plan = capi.SearchPlan([(c.sp, dQ, 70, c.codes, c.n, od, oi)], 10, capi.TQ_MODE_MFMA, **kw)
        plan.run()
        torch.cuda.synchronize()
        res.append((od.cpu(), oi.cpu()))
        plan.close()
    assert torch.equal(res[0][0], res[1][0]) and torch.equal(res[0][1], res[1][1])


def test_k_above_n_and_one_shot_search(gpu, torch_cuda):
    torch, capi = torch_cuda, gpu
    say("[gpu] k > n")
    c = Coll(torch, capi, 128, 4, 6, 41)
    _, dQ = c.queries(torch, 3)
    for mode in (capi.TQ_MODE_MFMA, capi.TQ_MODE_FP32):
        od, oi = outputs(torch, 3, 16)
        jobs = capi.make_jobs([(c.sp, dQ, 3, c.codes, 6, od, oi)])
        capi.check(capi.lib().tq_search(jobs, 1, 16, mode, None))
        got_i = oi.cpu().numpy()
        got_d = od.cpu().numpy()
        assert (got_i[:, 6:] == -1).all() and np.isinf(got_d[:, 6:]).all()
        want_d, want_i = host_topk(full_matrix(torch, capi, c, dQ, 3, mode), 16)
        np.testing.assert_array_equal(got_i, want_i)


@pytest.mark.parametrize("kind", ["ties", "uniform", "descending"])
def test_topk_rows_ties_by_column(gpu, torch_cuda, kind):
    # Segments of 8,192 columns go to the wavefronts as (row, segment) pairs,
    # so rows of 1-3 segments share workgroups; the shapes cover a lone
    # column in the last segment, partial sub-blocks of 2,048 columns, rows
    # shorter than one sub-block, and (ties, descending rows) more entries at
    # or below the first sub-block's bound than the packed list holds.
    torch, capi = torch_cuda, gpu
    say(f"[gpu] topk rows {kind}")
    rng = np.random.default_rng(5)
    for m, n, k in ((3, 100, 10), (20, 70000, 16), (1, 1, 1), (5, 4097, 7), (7, 8193, 16),
                    (9, 13568, 10), (33, 2047, 16), (6, 2049, 3), (4, 24577, 16),
                    (5, 70000, 100), (3, 8193, 128), (4, 50, 100), (2, 24577, 65),
                    (6, 3000, 17)):
        if kind == "ties":
            D = rng.integers(0, 50, (m, n)).astype(np.float32)  # many exact ties
        elif kind == "uniform":
            D = rng.random((m, n), dtype=np.float32)
        else:
            D = np.tile(np.arange(n, 0, -1, dtype=np.float32), (m, 1))
            D += rng.integers(0, 2, (m, 1)).astype(np.float32)
        dD = torch.from_numpy(D).cuda()
        od = torch.empty((m, k), dtype=torch.float32, device="cuda")
        oi = torch.empty((m, k), dtype=torch.int64, device="cuda")
        capi.topk_rows(dD, m, n, k, od, oi)
        torch.cuda.synchronize()
        want_d, want_i = host_topk(D, k)
        np.testing.assert_array_equal(oi.cpu().numpy(), want_i)
        np.testing.assert_array_equal(od.cpu().numpy(), want_d)


def test_search_rejects_bad_arguments(gpu, torch_cuda):
    torch, capi = torch_cuda, gpu
    c = Coll(torch, capi, 128, 4, 10, 51)
    _, dQ = c.queries(torch, 2)
    od, oi = outputs(torch, 2, 16)
    for k in (0, 129):
        with pytest.raises(capi.TQError, match="k must be"):
            capi.SearchPlan([(c.sp, dQ, 2, c.codes, 10, od, oi)], k)
    od17, oi17 = outputs(torch, 2, 17)
    with pytest.raises(capi.TQError, match="fused form takes k"):
        capi.SearchPlan([(c.sp, dQ, 2, c.codes, 10, od17, oi17)], 17, form=capi.TQ_FORM_FUSED)
    with pytest.raises(capi.TQError, match="query_tile"):
        capi.SearchPlan([(c.sp, dQ, 2, c.codes, 10, od, oi)], 10, query_tile=48)
    for bad in (dict(wg_per_cu=3), dict(blocked_query_tile=16), dict(block_mb=8),
                dict(block_mb=512)):
        with pytest.raises(capi.TQError):
            capi.SearchPlan([(c.sp, dQ, 2, c.codes, 10, od, oi)], 10, **bad)
    with pytest.raises(capi.TQError, match="NULL"):
        capi.SearchPlan([(c.sp, None, 2, c.codes, 10, od, oi)], 10)


@pytest.mark.parametrize("dim", [16, 5])
def test_mfma_plan_below_the_tile(gpu, torch_cuda, dim):
    # A collection of padded dim below 32 in an MFMA plan runs the plain fp32
    # scan and row top-k, next to an MFMA batch in the same plan: its result
    # equals the fp32 plan's, and the MFMA job's equals its own alone.
    torch, capi = torch_cuda, gpu
    say(f"[gpu] MFMA plan with a dim-{dim} collection")
    small = Coll(torch, capi, dim, 4, 5000, 61)
    big = Coll(torch, capi, 128, 4, 7000, 62)
    _, qs = small.queries(torch, 9)
    _, qb = big.queries(torch, 33)
    outs = {}
    for mode in (capi.TQ_MODE_MFMA, capi.TQ_MODE_FP32):
        od, oi = outputs(torch, 9, 10)
        bd, bi = outputs(torch, 33, 10)
        capi.check(capi.lib().tq_search(capi.make_jobs(
            [(small.sp, qs, 9, small.codes, small.n, od, oi),
             (big.sp, qb, 33, big.codes, big.n, bd, bi)]), 2, 10, mode, None))
        torch.cuda.synchronize()
        outs[mode] = (od.cpu().numpy(), oi.cpu().numpy(), bd.cpu().numpy(), bi.cpu().numpy())
    m, f = outs[capi.TQ_MODE_MFMA], outs[capi.TQ_MODE_FP32]
    np.testing.assert_array_equal(m[1], f[1])
    np.testing.assert_array_equal(m[0].view(np.uint32), f[0].view(np.uint32))
    want_d, want_i = host_topk(full_matrix(torch, capi, big, qb, 33, capi.TQ_MODE_MFMA), 10)
    np.testing.assert_array_equal(m[3], want_i)
    np.testing.assert_array_equal(m[2], want_d)
