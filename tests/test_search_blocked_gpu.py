"""MFMA search in the blocked form (MFMA distance blocks over chunks of codes,
row top-k per chunk, folded into one list per query; the plans' default) over
collections long enough to take several chunks.

Bars: the plan's top-k equals the (distance, id)-ordered top-k of the full
MFMA distance matrix, distances bit for bit, with one chunk and with several
(m = 5, 37, 100 over 2.3M codes: 1, 2 and 4 chunks), for 4-bit and 8-bit
codes; a plan holding a blocked and a fused batch gives each job what it gets
alone (blocked); graph replays equal eager runs; the server over such a
collection returns what a one-shot plan returns.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_BIG = 2_300_000
K = 10


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def big(gpu):
    import torch
    capi = gpu
    sp = capi.Space(128, 4)
    g = torch.Generator(device="cuda").manual_seed(7)
    codes = torch.empty((N_BIG, sp.cs), dtype=torch.uint8, device="cuda")
    for s0 in range(0, N_BIG, 500_000):
        rows = min(500_000, N_BIG - s0)
        x = torch.randn((rows, 128), generator=g, device="cuda")
        sp.encode(x, rows, codes[s0:s0 + rows])
    # a few exact duplicates, so equal distances occur and ties go by id
    codes[N_BIG - 1000:] = codes[:1000]
    torch.cuda.synchronize()
    yield sp, codes, g
    sp.close()


def matrix_topk(torch, capi, sp, Qd, m, codes, n, k):
    """(distance, id) top-k of the MFMA matrix, on the device: a stable sort
    keeps equal distances in id order."""
    D = torch.empty((m, n), dtype=torch.float32, device="cuda")
    sp.m_to_n(Qd, m, codes, n, D, mode=capi.TQ_MODE_MFMA)
    vals, idx = torch.sort(D, dim=1, stable=True)
    out = vals[:, :k].cpu().numpy(), idx[:, :k].cpu().numpy()
    del D, vals, idx
    torch.cuda.empty_cache()
    return out


def plan_topk(torch, capi, jobs, k, graph_replays=0, mode=None):
    outs = []
    spec = []
    for sp, Qd, m, codes, n in jobs:
        od = torch.empty((m, k), dtype=torch.float32, device="cuda")
        oi = torch.empty((m, k), dtype=torch.int64, device="cuda")
        outs.append((od, oi))
        spec.append((sp, Qd, m, codes, n, od, oi))
    plan = capi.SearchPlan(spec, k) if mode is None else capi.SearchPlan(spec, k, mode)
    plan.run()
    torch.cuda.synchronize()
    res = [(od.cpu().numpy(), oi.cpu().numpy()) for od, oi in outs]
    for _ in range(graph_replays):
        for od, oi in outs:
            od.fill_(-7.0)
            oi.fill_(-7)
        torch.cuda.synchronize()
        plan.run_graph()
        torch.cuda.synchronize()
        again = [(od.cpu().numpy(), oi.cpu().numpy()) for od, oi in outs]
        for (a_d, a_i), (b_d, b_i) in zip(res, again):
            np.testing.assert_array_equal(a_i, b_i)
            np.testing.assert_array_equal(a_d, b_d)
    plan.close()
    return res


@pytest.mark.parametrize("m", [5, 37, 100])
def test_blocked_equals_topk_of_matrix(gpu, big, m):
    import torch
    capi = gpu
    sp, codes, g = big
    say(f"[gpu] blocked search m={m} n={N_BIG}")
    Qd = torch.randn((m, 128), generator=g, device="cuda")
    want_d, want_i = matrix_topk(torch, capi, sp, Qd, m, codes, N_BIG, K)
    (got_d, got_i), = plan_topk(torch, capi, [(sp, Qd, m, codes, N_BIG)], K,
                                graph_replays=2 if m == 37 else 0)
    np.testing.assert_array_equal(got_i, want_i)
    np.testing.assert_array_equal(got_d, want_d)


def test_blocked_at_the_served_size(gpu):
    # The serving and headline collection size: 10,000,000 codes at dim 768,
    # 4-bit; 64 queries take ten 1M-code chunks. The plan's top-10 equals the
    # top-10 of the full [64, 10M] MFMA matrix, distances bit for bit.
    import torch
    capi = gpu
    say("[gpu] blocked search over 10M codes, dim 768 4-bit, m=64")
    sp = capi.Space(768, 4)
    n, m = 10_000_000, 64
    g = torch.Generator(device="cuda").manual_seed(23)
    codes = torch.empty((n, sp.cs), dtype=torch.uint8, device="cuda")
    for s0 in range(0, n, 1_000_000):
        x = torch.randn((1_000_000, 768), generator=g, device="cuda")
        sp.encode(x, 1_000_000, codes[s0:s0 + 1_000_000])
    del x
    Qd = torch.randn((m, 768), generator=g, device="cuda")
    want_d, want_i = matrix_topk(torch, capi, sp, Qd, m, codes, n, K)
    (got_d, got_i), = plan_topk(torch, capi, [(sp, Qd, m, codes, n)], K)
    np.testing.assert_array_equal(got_i, want_i)
    np.testing.assert_array_equal(got_d, want_d)
    del codes
    torch.cuda.empty_cache()
    sp.close()


@pytest.mark.parametrize("k", [1, 16, 17, 100, 128])
def test_blocked_k_bounds(gpu, big, k):
    # 100 queries over 2.3M codes: four chunks, so later chunks take the
    # running k-list's last entry as their bound. k = 100 is the reference
    # harness's recall@100 (run_benchmark.py:329-337); k > 16 runs blocked only.
    import torch
    capi = gpu
    sp, codes, g = big
    say(f"[gpu] blocked search k={k}")
    m = 100
    Qd = torch.randn((m, 128), generator=g, device="cuda")
    want_d, want_i = matrix_topk(torch, capi, sp, Qd, m, codes, N_BIG, k)
    (got_d, got_i), = plan_topk(torch, capi, [(sp, Qd, m, codes, N_BIG)], k)
    np.testing.assert_array_equal(got_i, want_i)
    np.testing.assert_array_equal(got_d, want_d)


@pytest.mark.parametrize("k", [10, 100])
@pytest.mark.parametrize("order", ["descending", "ascending"])
def test_blocked_adversarial_code_orders(gpu, big, order, k):
    # Codes sorted by their distance to one query: descending, every chunk
    # beats the running K-th best of that query's rows (each segment packs
    # more entries than its list holds and falls back); ascending, no later
    # chunk has anything at or below it (empty chunk lists).
    import torch
    capi = gpu
    sp, codes, g = big
    say(f"[gpu] blocked search, codes in {order} distance order, k={k}")
    q = torch.randn((3, 128), generator=g, device="cuda")
    d0 = torch.empty((1, N_BIG), dtype=torch.float32, device="cuda")
    sp.m_to_n(q[:1].contiguous(), 1, codes, N_BIG, d0, mode=capi.TQ_MODE_MFMA)
    perm = torch.argsort(d0[0], descending=(order == "descending"), stable=True)
    sorted_codes = codes.index_select(0, perm).contiguous()
    del d0, perm
    Qd = torch.cat([q[:1], q[:1], q[1:]]).contiguous()  # rows 0 and 1 the sorted-for query
    m = 4
    want_d, want_i = matrix_topk(torch, capi, sp, Qd, m, sorted_codes, N_BIG, k)
    (got_d, got_i), = plan_topk(torch, capi, [(sp, Qd, m, sorted_codes, N_BIG)], k)
    np.testing.assert_array_equal(got_i, want_i)
    np.testing.assert_array_equal(got_d, want_d)
    del sorted_codes
    torch.cuda.empty_cache()


@pytest.mark.parametrize("m", [9, 70])
def test_blocked_byte_codes(gpu, m):
    # 8-bit codes (the MFMA form's c_lo pass), 32- and 64-query tiles
    import torch
    capi = gpu
    say(f"[gpu] blocked search, 8-bit codes, m={m}")
    sp = capi.Space(128, 8)
    n = 2_100_000
    g = torch.Generator(device="cuda").manual_seed(11)
    codes = torch.empty((n, sp.cs), dtype=torch.uint8, device="cuda")
    for s0 in range(0, n, 700_000):
        x = torch.randn((700_000, 128), generator=g, device="cuda")
        sp.encode(x, 700_000, codes[s0:s0 + 700_000])
    Qd = torch.randn((m, 128), generator=g, device="cuda")
    want_d, want_i = matrix_topk(torch, capi, sp, Qd, m, codes, n, K)
    (got_d, got_i), = plan_topk(torch, capi, [(sp, Qd, m, codes, n)], K)
    np.testing.assert_array_equal(got_i, want_i)
    np.testing.assert_array_equal(got_d, want_d)
    sp.close()


def test_mixed_blocked_and_fused(gpu, big):
    # Two jobs over a second collection coalesce into one batch of 70,000
    # queries, more than the blocked form takes, so that batch runs fused in
    # the same plan as the blocked batches over the long collection; alone,
    # each of those jobs runs blocked. Every job gets the same bits either way.
    import torch
    capi = gpu
    sp, codes, g = big
    say("[gpu] blocked + fused batches in one plan")
    sp2 = capi.Space(96, 4)
    n2 = 60_000
    x2 = torch.randn((n2, 96), generator=g, device="cuda")
    codes2 = torch.empty((n2, sp2.cs), dtype=torch.uint8, device="cuda")
    sp2.encode(x2, n2, codes2)
    Q1 = torch.randn((20, 128), generator=g, device="cuda")
    Q2 = torch.randn((40_000, 96), generator=g, device="cuda")
    Q3 = torch.randn((7, 128), generator=g, device="cuda")
    Q4 = torch.randn((30_000, 96), generator=g, device="cuda")
    jobs = [(sp, Q1, 20, codes, N_BIG), (sp2, Q2, 40_000, codes2, n2), (sp, Q3, 7, codes, N_BIG),
            (sp2, Q4, 30_000, codes2, n2)]
    mixed = plan_topk(torch, capi, jobs, K)
    for job, (md, mi) in zip(jobs, mixed):
        (ad, ai), = plan_topk(torch, capi, [job], K)
        np.testing.assert_array_equal(mi, ai)
        np.testing.assert_array_equal(md, ad)
    sp2.close()


def test_server_over_blocked_collection(gpu, big):
    import torch
    capi = gpu
    sp, codes, g = big
    say("[gpu] server over a blocked collection")
    srv = capi.Server(mode=capi.TQ_MODE_MFMA, k=K, hbm_budget=8 << 30, slot_bytes=1 << 30,
                      max_inflight=2, max_batch_queries=4096)
    cid = srv.add_collection(sp, N_BIG)
    srv.load_codes(cid, codes, N_BIG)
    Qd = torch.randn((37, 128), generator=g, device="cuda")
    Q = Qd.cpu().numpy()
    od = np.empty((37, K), np.float32)
    oi = np.empty((37, K), np.int64)
    srv.wait(srv.search(cid, Q, od, oi))
    srv.close()
    (pd, pi), = plan_topk(torch, capi, [(sp, Qd, 37, codes, N_BIG)], K)
    np.testing.assert_array_equal(oi, pi)
    np.testing.assert_array_equal(od, pd)


@pytest.mark.parametrize("k", [10, 100])
def test_fp32_plan_large_k(gpu, big, k):
    # The plain path's plans: fp32 distances in row blocks, then the row top-k.
    import torch
    capi = gpu
    sp, codes, g = big
    say(f"[gpu] fp32 plan k={k}")
    m, n = 6, 300_000
    Qd = torch.randn((m, 128), generator=g, device="cuda")
    D = torch.empty((m, n), dtype=torch.float32, device="cuda")
    sp.m_to_n(Qd, m, codes, n, D, mode=capi.TQ_MODE_FP32)
    vals, idx = torch.sort(D, dim=1, stable=True)
    want_d, want_i = vals[:, :k].cpu().numpy(), idx[:, :k].cpu().numpy()
    (got_d, got_i), = plan_topk(torch, capi, [(sp, Qd, m, codes, n)], k, mode=capi.TQ_MODE_FP32)
    np.testing.assert_array_equal(got_i, want_i)
    np.testing.assert_array_equal(got_d, want_d)
