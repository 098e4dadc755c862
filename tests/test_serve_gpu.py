"""The request server (tq_server_*): continuous batching must not change
results. Each search's top-k is compared bit for bit with the same search run
alone, and with a search plan over the collection's codes as they stood when
the request was submitted; appends land in submission order and encode to the
same bytes as tq_encode_batch."""
import sys
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


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
    return np.ascontiguousarray((X / np.linalg.norm(X, axis=1, keepdims=True)).astype(np.float32))


def plan_topk(torch, capi, space, codes_ptr, n, Q, k, mode):
    m = len(Q)
    dQ = torch.from_numpy(Q).cuda()
    od = torch.empty((m, k), dtype=torch.float32, device="cuda")
    oi = torch.empty((m, k), dtype=torch.int64, device="cuda")
    p = capi.SearchPlan([(space, dQ, m, codes_ptr, n, od, oi)], k, mode)
    p.run()
    torch.cuda.synchronize()
    p.close()
    return od.cpu().numpy(), oi.cpu().numpy()


def make_server(capi, mode, specs, rng, **kw):
    srv = capi.Server(mode=mode, k=10, hbm_budget=4 << 30, slot_bytes=768 << 20, **kw)
    colls = []
    for dim, bits, n, cap in specs:
        sp = capi.Space(dim, bits)
        c = srv.add_collection(sp, cap)
        centres = rng.standard_normal((32, dim)).astype(np.float32)
        X = gm(rng, n, dim, centres)
        srv.wait(srv.append(c, X))
        colls.append((c, sp, centres, X))
    return srv, colls


SPECS = [(768, 4, 6000, 9000), (128, 8, 5000, 5000), (384, 4, 3000, 4000)]


@pytest.mark.parametrize("mode", [1, 0], ids=["mfma", "fp32"])
def test_results_same_alone_and_in_the_mix(gpu, torch_cuda, mode):
    torch, capi = torch_cuda, gpu
    say(f"[gpu] server alone vs mixed mode={mode}")
    rng = np.random.default_rng(100 + mode)
    srv, colls = make_server(capi, mode, SPECS, rng)
    reqs = []
    for i in range(48):
        ci = int(rng.integers(0, len(colls)))
        m = int(np.floor(np.exp(rng.uniform(0, np.log(65)))))
        c, sp, centres, _ = colls[ci]
        reqs.append((ci, gm(rng, m, sp.dim, centres)))

    def run(burst):
        outs = [(np.zeros((len(Q), 10), np.float32), np.zeros((len(Q), 10), np.int64))
                for _, Q in reqs]
        if burst:
            tickets = [srv.search(colls[ci][0], Q, *o) for (ci, Q), o in zip(reqs, outs)]
            lat = [srv.wait(t) for t in tickets]
        else:
            lat = [srv.wait(srv.search(colls[ci][0], Q, *o)) for (ci, Q), o in zip(reqs, outs)]
        assert all(x > 0 for x in lat)
        return outs

    alone = run(False)
    b0 = srv.stats()["batches"]
    mixed = run(True)
    st = srv.stats()
    say(f"  batches alone={b0} burst={st['batches'] - b0} max_batch_requests="
        f"{st['max_batch_requests']}")
    assert st["max_batch_requests"] > 1, "the burst was never batched"
    assert st["failed"] == 0
    for (ci, Q), (ad, ai), (md, mi) in zip(reqs, alone, mixed):
        np.testing.assert_array_equal(ai, mi)
        np.testing.assert_array_equal(ad.view(np.uint32), md.view(np.uint32))
        c, sp, _, _ = colls[ci]
        n, codes = srv.collection(c)
        pd, pi = plan_topk(torch, capi, sp, codes, n, Q, 10, mode)
        np.testing.assert_array_equal(ai, pi)
        np.testing.assert_array_equal(ad.view(np.uint32), pd.view(np.uint32))
    srv.close()


def test_appends_in_submission_order(gpu, torch_cuda):
    torch, capi = torch_cuda, gpu
    say("[gpu] server appends")
    rng = np.random.default_rng(7)
    srv = capi.Server(mode=capi.TQ_MODE_MFMA, k=10, hbm_budget=4 << 30, slot_bytes=512 << 20,
                      max_inflight=3)
    sp = capi.Space(768, 4)
    c = srv.add_collection(sp, 20000)
    centres = rng.standard_normal((16, 768)).astype(np.float32)
    parts = [gm(rng, n, 768, centres) for n in (3000, 1, 2500, 4000)]
    Q = gm(rng, 20, 768, centres)
    outs = [(np.zeros((20, 10), np.float32), np.zeros((20, 10), np.int64)) for _ in range(5)]
    # search sees 0, then each append in turn
    tickets = [srv.search(c, Q, *outs[0])]
    for i, X in enumerate(parts):
        tickets.append(srv.append(c, X))
        tickets.append(srv.search(c, Q, *outs[i + 1]))
    for t in tickets:
        srv.wait(t)
    n, codes = srv.collection(c)
    assert n == sum(len(p) for p in parts)
    # The appended codes are the encoder's bytes.
    Xall = np.concatenate(parts)
    dX = torch.from_numpy(Xall).cuda()
    want = torch.empty((n, sp.cs), dtype=torch.uint8, device="cuda")
    sp.encode(dX, n, want)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(srv.read_codes(c, 0, n, sp.cs), want.cpu().numpy())
    # Each search saw exactly the appends submitted before it.
    sizes = np.cumsum([0] + [len(p) for p in parts])
    for (od, oi), sz in zip(outs, sizes):
        if sz == 0:
            assert (oi == -1).all() and np.isinf(od).all()
            continue
        pd, pi = plan_topk(torch, capi, sp, codes, int(sz), Q, 10, capi.TQ_MODE_MFMA)
        np.testing.assert_array_equal(oi, pi)
        np.testing.assert_array_equal(od.view(np.uint32), pd.view(np.uint32))
        assert oi.max() < sz
    srv.close()


def test_concurrent_submitters(gpu, torch_cuda):
    torch, capi = torch_cuda, gpu
    say("[gpu] server concurrent submitters")
    rng = np.random.default_rng(9)
    srv, colls = make_server(capi, capi.TQ_MODE_MFMA, SPECS[:2], rng, max_batch_queries=256)
    work = []
    for t in range(4):
        mine = []
        for i in range(12):
            ci = int(rng.integers(0, 2))
            _, sp, centres, _ = colls[ci]
            m = int(rng.integers(1, 40))
            mine.append((ci, gm(rng, m, sp.dim, centres)))
        work.append(mine)
    results = [None] * 4

    def client(t):
        out = []
        for ci, Q in work[t]:
            od, oi = np.zeros((len(Q), 10), np.float32), np.zeros((len(Q), 10), np.int64)
            srv.wait(srv.search(colls[ci][0], Q, od, oi))
            out.append((od, oi))
        results[t] = out

    threads = [threading.Thread(target=client, args=(t,)) for t in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    for t in range(4):
        for (ci, Q), (od, oi) in zip(work[t], results[t]):
            c, sp, _, _ = colls[ci]
            n, codes = srv.collection(c)
            pd, pi = plan_topk(torch, capi, sp, codes, n, Q, 10, capi.TQ_MODE_MFMA)
            np.testing.assert_array_equal(oi, pi)
    srv.close()


def test_admission_limits(gpu, torch_cuda):
    torch, capi = torch_cuda, gpu
    with pytest.raises(capi.TQError, match="budget"):
        capi.Server(hbm_budget=1 << 20, slot_bytes=1 << 20, max_inflight=2)
    srv = capi.Server(hbm_budget=1 << 30, slot_bytes=256 << 20, max_inflight=2,
                      max_batch_queries=64)
    sp = capi.Space(768, 4)
    with pytest.raises(capi.TQError, match="budget"):
        srv.add_collection(sp, 2_000_000)  # 1.05 GB of codes
    c = srv.add_collection(sp, 100)
    X = np.zeros((101, 768), np.float32)
    with pytest.raises(capi.TQError, match="do not fit"):
        srv.append(c, X)
    Q = np.zeros((65, 768), np.float32)
    od, oi = np.zeros((65, 10), np.float32), np.zeros((65, 10), np.int64)
    with pytest.raises(capi.TQError, match="max_batch_queries"):
        srv.search(c, Q, od, oi)
    srv.close()


def test_add_collections_while_searches_run(gpu, torch_cuda):
    """Collections added while batches are in flight must not disturb the
    searches already submitted (the scheduler reads each request's collection
    without the server lock)."""
    torch, capi = torch_cuda, gpu
    say("[gpu] server add collections during searches")
    rng = np.random.default_rng(11)
    srv, colls = make_server(capi, capi.TQ_MODE_MFMA, SPECS[:1], rng, max_batch_queries=64)
    c, sp, centres, _ = colls[0]
    reqs = [gm(rng, int(rng.integers(1, 48)), sp.dim, centres) for _ in range(40)]
    outs = [(np.zeros((len(Q), 10), np.float32), np.zeros((len(Q), 10), np.int64))
            for Q in reqs]
    tickets = [srv.search(c, Q, *o) for Q, o in zip(reqs, outs)]
    extra = [srv.add_collection(capi.Space(128, 4), 16) for _ in range(64)]
    assert len(set(extra)) == 64
    for t in tickets:
        srv.wait(t)
    n, codes = srv.collection(c)
    for Q, (od, oi) in zip(reqs, outs):
        pd, pi = plan_topk(torch, capi, sp, codes, n, Q, 10, capi.TQ_MODE_MFMA)
        np.testing.assert_array_equal(oi, pi)
        np.testing.assert_array_equal(od.view(np.uint32), pd.view(np.uint32))
    assert srv.stats()["failed"] == 0
    srv.close()


def test_wrapper_holds_request_buffers(gpu, torch_cuda):
    """The native server has only the addresses of a request's host arrays:
    the Python wrapper must keep them alive until the ticket is waited for,
    also when the caller drops its own references right after submitting."""
    import gc
    import weakref
    capi = gpu
    say("[gpu] server wrapper holds request buffers")
    rng = np.random.default_rng(12)
    srv, colls = make_server(capi, capi.TQ_MODE_MFMA, SPECS[:1], rng, max_batch_queries=64)
    c, sp, centres, _ = colls[0]
    refs, tickets = [], []
    for _ in range(24):
        Q = gm(rng, int(rng.integers(1, 48)), sp.dim, centres)
        od = np.zeros((len(Q), 10), np.float32)
        oi = np.zeros((len(Q), 10), np.int64)
        tickets.append(srv.search(c, Q, od, oi))
        refs.append((weakref.ref(Q), weakref.ref(od), weakref.ref(oi)))
    X = gm(rng, 500, sp.dim, centres)
    xref = weakref.ref(X)
    tickets.append(srv.append(c, X))
    del Q, od, oi, X
    gc.collect()
    assert all(r() is not None for trio in refs for r in trio) and xref() is not None
    def check(t, trio):
        # Strong references taken before the wait, which releases the wrapper's.
        _, od, oi = (r() for r in trio)
        srv.wait(t)
        # Results written before the wait returned: every row holds k valid ids.
        assert (oi >= 0).all() and np.isfinite(od).all()

    for t, trio in zip(tickets, refs):
        check(t, trio)
    srv.wait(tickets[-1])
    gc.collect()
    assert all(r() is None for trio in refs for r in trio) and xref() is None
    assert srv.stats()["failed"] == 0
    srv.close()
