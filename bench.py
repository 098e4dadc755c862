"""Throughput of the TurboQuant distance path on one MI355X.

Workload (dim 768, 4-bit, 524-byte codes): one resident collection of
10,000,000 codes encoded on the GPU from a Gaussian mixture (8,192 centres
N(0,1), within-cluster std 0.3, L2-normalised, seed 20); each step scores one
batch of 64 queries drawn the same way (seed 21) against every code
(distance_m_to_n: query preparation + scan, output [64, 10M] float32 in HBM).
All inputs are resident in HBM before the timed region.

Printed: one JSON line. `value` is distances/s of the fastest path that meets
its accuracy bar (fp16 hi+lo MFMA, <= 1e-4 of the fp64 estimator, when built;
else the plain fp32 path); the plain fp32 path, distance_1_to_n and encode are
reported beside it, with the reference's own CPU build timed on this host,
followed by the recall, symmetric (dim 1000, 4,096 x 2M codes), 8-bit
(dim 768, 1,024 x 1M codes, fp32 and MFMA), batched
(256 uneven batches, 5 collections) and serving (fixed-QPS) measurements,
each skippable by a --no-* flag.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "turboquant-space_amd"))

HBM_PEAK_GBS = 8000.0
FP16_MFMA_PEAK_TFLOPS = 2500.0
DIM, BITS = 768, 4
N_CODES = 10_000_000
M_QUERIES = 64


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gaussian_mixture(torch, n, dim, k, std, seed, centres=None, chunk=1_000_000, normalise=True):
    """Rows of gm(n, dim, k, std) generated on the GPU in chunks, L2-normalised
    unless normalise is False."""
    g = torch.Generator(device="cuda").manual_seed(seed)
    if centres is None:
        centres = torch.randn((k, dim), generator=g, device="cuda")
    for start in range(0, n, chunk):
        rows = min(chunk, n - start)
        lab = torch.randint(0, k, (rows,), generator=g, device="cuda")
        x = centres[lab] + std * torch.randn((rows, dim), generator=g, device="cuda")
        if normalise:
            x = x / torch.linalg.vector_norm(x, dim=1, keepdim=True)
        yield start, x.contiguous(), centres


def event_time(torch, stream, fn, reps):
    """Average ms of fn() over reps launches, HIP events on `stream`."""
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record(stream)
    for _ in range(reps):
        fn()
    e.record(stream)
    e.synchronize()
    return s.elapsed_time(e) / reps


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    rocprofv3 FETCH_SIZE + WRITE_SIZE summary (profiles/traffic_bench.json,
    counters scaled by a 1 GiB copy); None if it has not been measured."""
    path = os.path.join(REPO, "profiles", "traffic_bench.json")
    try:
        with open(path) as f:
            return json.load(f)["kernels"][kernel]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


# Batched uneven launch: 5 resident collections, (dim, bits) and code sizes
# 96/4 -> 76 B, 128/8 -> 140 B, 384/4 -> 268 B, 768/8 -> 1036 B, 1536/4 -> 1036 B.
BATCH_COLLS = [(96, 4), (128, 8), (384, 4), (768, 8), (1536, 4)]
BATCH_N = 500_000
BATCH_COUNT = 256


def uneven_batches(seed=15, count=BATCH_COUNT, mmax=512):
    """m ~ floor(exp(U(0, ln 513))) capped at 512 and a uniform target
    collection per batch, both from one numpy generator."""
    rng = np.random.default_rng(seed)
    ms = np.minimum(np.floor(np.exp(rng.uniform(0.0, np.log(mmax + 1), count))), mmax)
    targets = rng.integers(0, len(BATCH_COLLS), count)
    return [(int(t), int(m)) for t, m in zip(targets, ms)]


def bench_batched(torch, capi, stream, sh, reps=5):
    """256 query batches of uneven m against 5 collections of 500,000 codes:
    one tuned MFMA search plan (top-10, HIP-graph replay) and the same plan
    with the fused persistent form forced, against the same batches as
    per-batch plain launches (prepare, fp32 distance_m_to_n, row top-10) and
    per-batch MFMA launches."""
    colls = []
    for i, (dim, bits) in enumerate(BATCH_COLLS):
        sp = capi.Space(dim, bits)
        codes = torch.empty((BATCH_N, sp.cs), dtype=torch.uint8, device="cuda")
        centres = None
        for start, x, centres in gaussian_mixture(torch, BATCH_N, dim, 512, 0.3, 10 + i,
                                                  chunk=250_000):
            sp.encode(x, x.shape[0], codes[start:start + x.shape[0]], stream=sh)
        colls.append((sp, codes, centres))
    batches = uneven_batches()
    jobs, queries = [], []
    for b, (t, m) in enumerate(batches):
        sp, codes, centres = colls[t]
        _, q, _ = next(gaussian_mixture(torch, m, sp.dim, 512, 0.3, 1000 + b, centres=centres))
        od = torch.empty((m, 10), dtype=torch.float32, device="cuda")
        oi = torch.empty((m, 10), dtype=torch.int64, device="cuda")
        queries.append(q)
        jobs.append((sp, q, m, codes, BATCH_N, od, oi))
    torch.cuda.synchronize()
    total_q = sum(m for _, m in batches)
    dists = total_q * BATCH_N
    flops = sum(4.0 * colls[t][0].D * m * BATCH_N for t, m in batches)

    out = {"batches": len(batches), "queries": total_q, "codes_per_collection": BATCH_N,
           "collections": [{"dim": d, "bits": b} for d, b in BATCH_COLLS], "k": 10}
    # The tuned plan: per collection the autotuner's fastest form (blocked or
    # fused) and fused knobs; then the same batches with the fused form forced
    # (one persistent launch over every collection, tuned knobs).
    plan = capi.SearchPlan(jobs, 10, capi.TQ_MODE_MFMA)
    t0 = time.perf_counter()
    qt, ipw, form = plan.tune(stream=sh)
    out["tune_s"] = time.perf_counter() - t0
    plan.run_graph(stream=sh)
    ms = event_time(torch, stream, lambda: plan.run_graph(stream=sh), reps)
    forms = {capi.TQ_FORM_FUSED: "fused", capi.TQ_FORM_BLOCKED: "blocked"}
    out["tuned_mfma_plan"] = {"ms": ms, "dist_per_s": dists / (ms / 1e3),
                              "queries_per_s": total_q / (ms / 1e3),
                              "tflops": flops / (ms / 1e3) / 1e12,
                              "form_of_costliest_collection": forms.get(form, form),
                              "tuned_fused_knobs": {"query_tile": qt, "items_per_wg": ipw},
                              "tuned_knobs_costliest_batch": plan.tuned}
    # The tuner's table entry of each collection (one batch per collection:
    # the plan coalesces the jobs aimed at it).
    per_coll = []
    for t, (sp, _, _) in enumerate(colls):
        mt = sum(m for tt, m in batches if tt == t)
        per_coll.append({"dim": sp.dim, "bits": BATCH_COLLS[t][1], "m": mt,
                         "knobs": capi.tuning_lookup(sp.D, BATCH_COLLS[t][1], mt)})
    out["tuned_mfma_plan"]["tuned_per_collection"] = per_coll
    ref_ids = [j[6].clone() for j in jobs]
    plan.close()
    log(f"batched: tuned MFMA plan graph {ms:.2f} ms ({dists / (ms / 1e3) / 1e9:.1f} G dist/s)")
    for pc in per_coll:
        log(f"batched: tuned knobs dim {pc['dim']} {pc['bits']}-bit m {pc['m']}: {pc['knobs']}")
    fplan = capi.SearchPlan(jobs, 10, capi.TQ_MODE_MFMA, form=capi.TQ_FORM_FUSED)
    items, wgs = fplan.stats()
    fplan.run_graph(stream=sh)
    msf = event_time(torch, stream, lambda: fplan.run_graph(stream=sh), reps)
    out["fused_persistent_mfma"] = {"ms": msf, "dist_per_s": dists / (msf / 1e3),
                                    "work_items": items, "workgroups": wgs}
    out["fused_ids_equal_tuned"] = all(torch.equal(a, j[6]) for a, j in zip(ref_ids, jobs))
    fplan.close()
    log(f"batched: fused persistent MFMA graph {msf:.2f} ms")

    plan32 = capi.SearchPlan(jobs, 10, capi.TQ_MODE_FP32)
    plan32.run_graph(stream=sh)
    ms32 = event_time(torch, stream, lambda: plan32.run_graph(stream=sh), max(1, reps // 2))
    out["persistent_fp32"] = {"ms": ms32, "dist_per_s": dists / (ms32 / 1e3)}
    plan32.close()
    log(f"batched: merged fp32 plan {ms32:.2f} ms")

    # Per-batch launches through the plain entry points.
    mmax = max(m for _, m in batches)
    dist = torch.empty((mmax, BATCH_N), dtype=torch.float32, device="cuda")
    prep = torch.empty((mmax, max(c[0].prep_stride for c in colls)), dtype=torch.float32,
                       device="cuda")

    def per_batch(mode):
        for (t, m), q, j in zip(batches, queries, jobs):
            sp, codes, _ = colls[t]
            sp.prepare(q, m, prep, stream=sh)
            sp.scan(prep, m, codes, BATCH_N, dist, mode=mode, stream=sh)
            capi.topk_rows(dist, m, BATCH_N, 10, j[5], j[6], stream=sh)

    for name, mode in (("per_batch_fp32", capi.TQ_MODE_FP32), ("per_batch_mfma", capi.TQ_MODE_MFMA)):
        per_batch(mode)
        msb = event_time(torch, stream, lambda: per_batch(mode), 1)
        out[name] = {"ms": msb, "dist_per_s": dists / (msb / 1e3)}
        log(f"batched: {name} {msb:.2f} ms")
    torch.cuda.synchronize()
    # The MFMA full matrices give the same top-10 ids as the tuned plan.
    out["per_batch_mfma_ids_equal_plan"] = all(torch.equal(a, j[6]) for a, j in zip(ref_ids, jobs))
    out["speedup_vs_per_batch_fp32"] = out["per_batch_fp32"]["ms"] / ms
    del dist, prep, colls, jobs, queries
    torch.cuda.empty_cache()
    return out


# Serving: 4 resident collections of 10,000,000 codes at dim 768, 4-bit.
SERVE_COLLS = 4
SERVE_N = 10_000_000


SERVE_SLO_MS = 50.0                       # p99 bar of the load curves' SLO figure
SERVE_QPS = {"mfma": [4, 25, 50, 75, 100, 125, 150, 200], "fp32": [0.5, 1, 2, 4, 8, 12]}


def bench_serving(torch, capi, sh, qps_by_mode, seconds, n_per_coll=SERVE_N, seed=21):
    """Load curves of the request server, one per form (MFMA, plain fp32):
    Poisson arrivals at each offered rate of `qps_by_mode` for `seconds`
    each; a request is a top-10 search of m ~ floor(exp(U(0, ln 257))) queries
    against a uniformly chosen collection, or with probability 0.05 an append
    of 10,000 vectors. The 4 collections are consecutive 10M-row slices of one
    Gaussian mixture over 40M vectors (8,192 centres, std 0.3, normalised,
    seed 20). Latency is submission to results-in-host-memory, measured by
    the server. A sweep stops past saturation (achieved QPS below 90% of the
    realized arrival rate, so the backlog grows for as long as the point
    runs). `slo_qps` is the highest offered rate served at p99 <=
    SERVE_SLO_MS with at least 95% of the realized rate achieved."""
    import threading
    results = {}
    sp = capi.Space(DIM, BITS)
    # The 40M-vector mixture, encoded once on the device, one buffer per slice.
    staged = [torch.empty((n_per_coll, sp.cs), dtype=torch.uint8, device="cuda")
              for _ in range(SERVE_COLLS)]
    centres = None
    for start, x, centres in gaussian_mixture(torch, SERVE_COLLS * n_per_coll, DIM, 8192, 0.3, 20):
        c, off = divmod(start, n_per_coll)
        rows = x.shape[0]
        while rows:  # a generator chunk may straddle two slices
            take = min(rows, n_per_coll - off)
            sp.encode(x[x.shape[0] - rows:x.shape[0] - rows + take].contiguous(), take,
                      staged[c][off:off + take], stream=sh)
            rows -= take
            c, off = c + 1, 0
    torch.cuda.synchronize()
    rng = np.random.default_rng(seed)
    longest = max(seconds * max(q) for q in qps_by_mode.values() if q)
    horizon = int(max(longest * 1.3, 50))
    # Request pool, drawn once: kinds and sizes; search queries are drawn per
    # request, append payloads from a pool of 16 (host memory stays bounded).
    reqs = []
    for i in range(horizon):
        c = int(rng.integers(0, SERVE_COLLS))
        if rng.random() < 0.05:
            reqs.append(("append", c, 10_000))
        else:
            m = int(min(np.floor(np.exp(rng.uniform(0.0, np.log(257)))), 256))
            reqs.append(("search", c, m))
    payload = {}
    appends = []
    for i in range(16):
        _, x, _ = next(gaussian_mixture(torch, 10_000, DIM, 8192, 0.3, 4000 + i, centres=centres))
        appends.append(np.ascontiguousarray(x.cpu().numpy()))
    for i, (kind, c, rows) in enumerate(reqs):
        if kind == "append":
            payload[i] = appends[i % 16]
            continue
        _, x, _ = next(gaussian_mixture(torch, rows, DIM, 8192, 0.3, 5000 + i, centres=centres))
        payload[i] = np.ascontiguousarray(x.cpu().numpy())

    def arrivals(qps):
        arr = np.cumsum(np.random.default_rng(seed + 1).exponential(1.0 / qps, len(reqs)))
        return arr, int(np.searchsorted(arr, seconds))

    def run_point(srv, ids, qps):
        arr, count = arrivals(qps)
        tickets = [None] * count
        done = []
        stop = [False]
        lock = threading.Condition()

        def waiter():
            for i in range(count):
                with lock:
                    while tickets[i] is None and not stop[0]:
                        lock.wait()
                    if tickets[i] is None:
                        return  # submission stopped early
                lat = srv.wait(tickets[i])
                done.append((i, lat, time.perf_counter()))

        st0 = srv.stats()
        th = threading.Thread(target=waiter, daemon=True)
        th.start()
        t_start = time.perf_counter()
        err = None
        try:
            for i in range(count):
                delay = t_start + arr[i] - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                kind, c, rows = reqs[i]
                if kind == "append":
                    t = srv.append(ids[c], payload[i])
                else:
                    od = np.empty((rows, 10), np.float32)
                    oi = np.empty((rows, 10), np.int64)
                    t = srv.search(ids[c], payload[i], od, oi)
                with lock:
                    tickets[i] = t
                    lock.notify()
        except Exception as e:  # a refused submission ends the point (and the run) loudly
            err = e
        finally:
            with lock:
                stop[0] = True
                lock.notify_all()
        th.join()
        if err is not None:
            raise err
        t_end = max(d[2] for d in done) if done else time.perf_counter()
        st = srv.stats()
        lat_s = np.array([d[1] for d in done if reqs[d[0]][0] == "search"])
        lat_a = np.array([d[1] for d in done if reqs[d[0]][0] == "append"])
        nq = sum(reqs[i][2] for i in range(count) if reqs[i][0] == "search")
        # Arrivals span `seconds`; the last completions may come later.
        wall = max(t_end - t_start, seconds)
        return {
            "qps_offered": qps, "qps_realized": count / seconds, "requests": count,
            "searches": int(len(lat_s)), "appends": int(len(lat_a)), "seconds": wall,
            "qps_achieved": count / wall,
            "queries_per_s": nq / wall, "dist_per_s": nq * n_per_coll / wall,
            "p50_ms": float(np.percentile(lat_s, 50) * 1e3),
            "p99_ms": float(np.percentile(lat_s, 99) * 1e3),
            "mean_ms": float(lat_s.mean() * 1e3),
            "append_p50_ms": float(np.percentile(lat_a, 50) * 1e3) if len(lat_a) else None,
            "batches": st["batches"] - st0["batches"],
            "requests_per_batch": count / max(1, st["batches"] - st0["batches"]),
            "max_batch_requests": st["max_batch_requests"],
            "failed": st["failed"] - st0["failed"]}

    for mode_name, points in qps_by_mode.items():
        if not points:
            continue  # this form's curve was not asked for
        mode = capi.TQ_MODE_MFMA if mode_name == "mfma" else capi.TQ_MODE_FP32
        # Capacity for every append the sweep can submit (points share the
        # request pool, and the collections keep what earlier points appended).
        extra = [0] * SERVE_COLLS
        for qps in points:
            for kind, c, rows in reqs[:arrivals(qps)[1]]:
                if kind == "append":
                    extra[c] += rows
        cap = n_per_coll + max(extra) + 10_000
        srv = capi.Server(mode=mode, k=10, hbm_budget=96 << 30, slot_bytes=2 << 30,
                          max_inflight=2, max_batch_queries=4096)
        ids = []
        for c in range(SERVE_COLLS):
            ids.append(srv.add_collection(sp, cap))
            srv.load_codes(ids[-1], staged[c], n_per_coll)
        curve = []
        for qps in points:
            pt = run_point(srv, ids, qps)
            curve.append(pt)
            log(f"serving {mode_name} @ {qps} QPS: {pt['requests']} requests, achieved "
                f"{pt['qps_achieved']:.1f} QPS, p50 {pt['p50_ms']:.1f} ms, p99 {pt['p99_ms']:.1f} ms, "
                f"{pt['requests_per_batch']:.2f} requests/batch")
            if pt["qps_achieved"] < 0.9 * pt["qps_realized"]:
                break  # past saturation: the backlog grows for as long as the point runs
        srv.close()
        ok = [pt["qps_offered"] for pt in curve
              if pt["p99_ms"] <= SERVE_SLO_MS and pt["qps_achieved"] >= 0.95 * pt["qps_realized"]]
        last = curve[-1]
        results[mode_name] = {"curve": curve, "slo_p99_ms": SERVE_SLO_MS,
                              "slo_qps": max(ok) if ok else None,
                              "saturated_at_qps": last["qps_offered"]
                              if last["qps_achieved"] < 0.9 * last["qps_realized"] else None,
                              "peak_queries_per_s": max(pt["queries_per_s"] for pt in curve)}
        log(f"serving {mode_name}: highest QPS at p99 <= {SERVE_SLO_MS:.0f} ms: "
            f"{results[mode_name]['slo_qps']}")
    del staged
    torch.cuda.empty_cache()
    return {"collections": SERVE_COLLS, "codes_per_collection": n_per_coll, "dim": DIM,
            "bits": BITS, "k": 10, "seconds_per_point": seconds,
            "data": "one Gaussian mixture over 40M vectors (8,192 centres, std 0.3, normalised, "
                    "seed 20), 4 consecutive slices", "modes": results}


RECALL_N = 1_000_000
RECALL_Q = 256


# Symmetric workload: dim 1000 (padded to 1024), 4-bit; 2,000,000 base
# vectors and 4,096 query vectors of one 4,096-centre mixture (std 0.5, not
# normalised; seeds 4 and 5).
SYM_DIM, SYM_BITS, SYM_N, SYM_M = 1000, 4, 2_000_000, 4096


def bench_symmetric(torch, capi, stream, sh, n=SYM_N, m=SYM_M, rows=512):
    """encode_batch of the 2M base vectors, then distance_m_to_n_symmetric of
    the 4,096 query codes against them, in blocks of `rows` query codes into
    one reused [rows, n] output: by the fp32 centroid-product LUT scan and by
    the MFMA form (decoded codes as fp16 hi+lo queries, two passes)."""
    sp = capi.Space(SYM_DIM, SYM_BITS)
    codes_b = torch.empty((n, sp.cs), dtype=torch.uint8, device="cuda")
    centres = None
    enc_ms = 0.0
    for start, x, centres in gaussian_mixture(torch, n, SYM_DIM, 4096, 0.5, 4, normalise=False):
        r = x.shape[0]
        enc_ms += event_time(torch, stream,
                             lambda: sp.encode(x, r, codes_b[start:start + r], stream=sh), 1)
    codes_a = torch.empty((m, sp.cs), dtype=torch.uint8, device="cuda")
    for start, x, _ in gaussian_mixture(torch, m, SYM_DIM, 4096, 0.5, 5, centres=centres,
                                        normalise=False):
        sp.encode(x, x.shape[0], codes_a[start:start + x.shape[0]], stream=sh)
    out = torch.empty((rows, n), dtype=torch.float32, device="cuda")

    def sweep(mode):
        for r0 in range(0, m, rows):
            rc = min(rows, m - r0)
            sp.symmetric(codes_a[r0:r0 + rc], rc, codes_b, n, out, stream=sh, mode=mode)

    ms = event_time(torch, stream, lambda: sweep(capi.TQ_MODE_FP32), 1)
    sweep(capi.TQ_MODE_MFMA)
    ms_mfma = event_time(torch, stream, lambda: sweep(capi.TQ_MODE_MFMA), 1)
    res = {"dim": SYM_DIM, "bits": SYM_BITS, "codes_b": n, "codes_a": m,
           "encode_vec_per_s": n / (enc_ms / 1e3),
           "symmetric_ms": ms, "dist_per_s": m * n / (ms / 1e3),
           "code_bytes_per_s": m * n * sp.cs / (ms / 1e3),
           "mfma": {"ms": ms_mfma, "dist_per_s": m * n / (ms_mfma / 1e3),
                    "tflops_useful": 2.0 * sp.D * m * n / (ms_mfma / 1e3) / 1e12}}
    log(f"symmetric {m} x {n} (dim {SYM_DIM}, {SYM_BITS}-bit): fp32 LUT {ms:.1f} ms, "
        f"{res['dist_per_s'] / 1e9:.2f} G dist/s; MFMA {ms_mfma:.1f} ms, "
        f"{res['mfma']['dist_per_s'] / 1e9:.1f} G dist/s; "
        f"encode {res['encode_vec_per_s'] / 1e6:.1f} M vec/s")
    del out, codes_a, codes_b
    torch.cuda.empty_cache()
    return res


# Byte-code workload: dim 768, 8-bit (1,036-byte codes); 1,000,000 base and
# 1,024 query vectors of one 1,024-centre mixture (std 0.3, L2-normalised;
# seeds 2 and 3).
A8_DIM, A8_BITS, A8_N, A8_M = 768, 8, 1_000_000, 1024


def bench_asym8(torch, capi, stream, sh, n=A8_N, m=A8_M):
    """encode_batch of the 1M base vectors, then distance_m_to_n of the 1,024
    queries against them (output [1024, 1M] in HBM) by the plain fp32 path and
    by the MFMA form (hi+lo fp16 queries, c_hi + c_lo centroids)."""
    sp = capi.Space(A8_DIM, A8_BITS)
    codes = torch.empty((n, sp.cs), dtype=torch.uint8, device="cuda")
    centres = None
    enc_ms = 0.0
    for start, x, centres in gaussian_mixture(torch, n, A8_DIM, 1024, 0.3, 2):
        r = x.shape[0]
        enc_ms += event_time(torch, stream, lambda: sp.encode(x, r, codes[start:start + r], stream=sh), 1)
    _, q, _ = next(gaussian_mixture(torch, m, A8_DIM, 1024, 0.3, 3, centres=centres))
    prep = torch.empty((m, sp.prep_stride), dtype=torch.float32, device="cuda")
    out = torch.empty((m, n), dtype=torch.float32, device="cuda")

    def run(mode):
        sp.prepare(q, m, prep, stream=sh)
        sp.scan(prep, m, codes, n, out, mode=mode, stream=sh)

    run(capi.TQ_MODE_MFMA)
    ms_m = event_time(torch, stream, lambda: run(capi.TQ_MODE_MFMA), 3)
    ms_f = event_time(torch, stream, lambda: run(capi.TQ_MODE_FP32), 1)
    res = {"dim": A8_DIM, "bits": A8_BITS, "codes": n, "queries": m,
           "encode_vec_per_s": n / (enc_ms / 1e3),
           "mfma_ms": ms_m, "mfma_dist_per_s": m * n / (ms_m / 1e3),
           "fp32_ms": ms_f, "fp32_dist_per_s": m * n / (ms_f / 1e3), "mfma_over_fp32": ms_f / ms_m}
    log(f"asym 8-bit {m} x {n}: mfma {ms_m:.2f} ms ({res['mfma_dist_per_s'] / 1e9:.1f} G dist/s), "
        f"fp32 {ms_f:.1f} ms ({res['fp32_dist_per_s'] / 1e9:.2f} G dist/s); "
        f"encode {res['encode_vec_per_s'] / 1e6:.1f} M vec/s")
    del out, codes, prep
    torch.cuda.empty_cache()
    return res


def bench_recall(torch, capi, sh, n=RECALL_N, m=RECALL_Q, k=100):
    """Recall@1, @10 and @100 of the quantized searches against the exact
    neighbours (squared L2 over the raw vectors, fp64 brute force on the GPU)
    on a sample of the bench workload: n vectors and m queries of the same
    recipe (dim 768, 8,192-centre Gaussian mixture, std 0.3, L2-normalised;
    seeds 40, 41). As the reference's harness does, one top-100 search gives
    all three: recall@r = |pred[:r] & exact[:r]| / r
    (python/benchmarks/run_benchmark.py:274-277, :329-337)."""
    from turboquant.recall import exact_topk, recall_at_k
    sp = capi.Space(DIM, BITS)
    X = torch.empty((n, DIM), device="cuda")
    codes = torch.empty((n, sp.cs), dtype=torch.uint8, device="cuda")
    centres = None
    for start, x, centres in gaussian_mixture(torch, n, DIM, 8192, 0.3, 40):
        X[start:start + x.shape[0]] = x
        sp.encode(x, x.shape[0], codes[start:start + x.shape[0]], stream=sh)
    _, Q, _ = next(gaussian_mixture(torch, m, DIM, 8192, 0.3, 41, centres=centres))
    torch.cuda.synchronize()
    gt = exact_topk(Q, X, k).cpu().numpy()
    del X
    od = torch.empty((m, k), device="cuda")
    oi = torch.empty((m, k), dtype=torch.int64, device="cuda")
    plan = capi.SearchPlan([(sp, Q, m, codes, n, od, oi)], k)
    plan.run(stream=sh)
    torch.cuda.synchronize()
    ids_mfma = oi.cpu().numpy()
    plan.close()
    dist = torch.empty((m, n), device="cuda")
    sp.m_to_n(Q, m, codes, n, dist, mode=capi.TQ_MODE_FP32, stream=sh)
    capi.topk_rows(dist, m, n, k, od, oi, stream=sh)
    torch.cuda.synchronize()
    ids_fp32 = oi.cpu().numpy()
    del dist, codes
    torch.cuda.empty_cache()
    res = {"vectors": n, "queries": m, "k": k, "dim": DIM, "bits": BITS}
    for name, ids in (("mfma_search_plan", ids_mfma), ("fp32_distances_topk", ids_fp32)):
        res[name] = {f"recall_at_{r}": recall_at_k(ids, gt, r) for r in (1, 10, 100)}
    log("recall@1/@10/@100 vs exact: MFMA search plan " +
        "/".join(f"{res['mfma_search_plan'][f'recall_at_{r}']:.4f}" for r in (1, 10, 100)) +
        ", fp32 path " +
        "/".join(f"{res['fp32_distances_topk'][f'recall_at_{r}']:.4f}" for r in (1, 10, 100)))
    return res


def cpu_info():
    """The host the CPU baseline runs on: CPU model, logical CPUs, the CPUs this
    process may run on, and physical cores (/proc/cpuinfo)."""
    model, cores = None, set()
    phys = core = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name" and model is None:
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                    cores.add((phys, core))
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count(),
            "usable_cpus": len(os.sched_getaffinity(0)), "physical_cores": len(cores) or None}


def cpu_child(path):
    """The CPU baseline's child process (no torch, so OpenMP starts with the
    environment the parent set): the reference build on the sample in `path`,
    three legs, min of 3 runs each; prints one JSON line."""
    d = np.load(path)
    codes, Q, X = d["codes"], np.ascontiguousarray(d["Q"]), np.ascontiguousarray(d["X"])
    L = ctypes.CDLL(os.path.join(REPO, "oracle", "_ref", "libtqref.so"))
    P, SZ = ctypes.c_void_p, ctypes.c_size_t
    L.tqref_create.restype = P
    L.tqref_create.argtypes = [SZ, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int]
    L.tqref_num_threads.restype = ctypes.c_int
    L.tqref_num_threads.argtypes = [P]
    L.tqref_distance_m_to_n.argtypes = [P, P, SZ, P, SZ, P]
    L.tqref_distance_1_to_n.argtypes = [P, P, P, SZ, P]
    L.tqref_encode_batch.argtypes = [P, P, SZ, P]
    L.tqref_destroy.argtypes = [P]
    h = L.tqref_create(DIM, BITS, 42, 137, 0)
    threads = L.tqref_num_threads(h)

    def best_of(fn, reps=3):
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            best = min(best, time.perf_counter() - t0)
        return best

    # m x n with m = 128 (> 64, so the reference's OpenMP loop over m
    # engages), n sized to ~8 s over the three runs.
    m = 128
    out = np.zeros((m, 20_000), np.float32)
    t = best_of(lambda: L.tqref_distance_m_to_n(h, Q.ctypes.data, m, codes.ctypes.data, 20_000,
                                                out.ctypes.data), 1)
    n = int(min(len(codes), max(20_000, 8.0 / 3 / (t / (m * 20_000)) / m)))
    out = np.zeros((m, n), np.float32)
    t_mn = best_of(lambda: L.tqref_distance_m_to_n(h, Q.ctypes.data, m, codes.ctypes.data, n,
                                                   out.ctypes.data))
    # distance_1_to_n over every sampled code (OpenMP over n).
    n1 = len(codes)
    one = np.zeros(n1, np.float32)
    t_1n = best_of(lambda: L.tqref_distance_1_to_n(h, Q.ctypes.data, codes.ctypes.data, n1,
                                                   one.ctypes.data), 5)
    # encode_batch of the raw sample (OpenMP over rows).
    cbuf = np.zeros((len(X), codes.shape[1]), np.uint8)
    t_enc = best_of(lambda: L.tqref_encode_batch(h, X.ctypes.data, len(X), cbuf.ctypes.data))
    L.tqref_destroy(h)
    print(json.dumps({"threads": threads, "m": m, "n": n, "t_mn": t_mn, "n1": n1, "t_1n": t_1n,
                      "n_enc": len(X), "t_enc": t_enc}), flush=True)
    return 0


def cpu_baseline(codes_host, Q_host, X_host):
    """The reference CPU build (oracle/_ref, compiled in place from the
    reference sources with its Release flags) on this host's cores, in a
    child process with OMP_PROC_BIND=close and OMP_PLACES=cores (threads:
    OMP_NUM_THREADS as set, else every usable CPU): distance_m_to_n
    (m = 128 over a prefix of the collection sized to ~8 s), distance_1_to_n
    over the whole sample and encode_batch of raw vectors of the same recipe;
    min of 3 (1-to-N: 5) runs."""
    import subprocess
    import tempfile
    if not os.path.exists(os.path.join(REPO, "oracle", "_ref", "libtqref.so")):
        return None
    info = cpu_info()
    env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores")
    env.setdefault("OMP_NUM_THREADS", str(info["usable_cpus"]))
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "cpu_sample.npz")
        np.savez(path, codes=codes_host, Q=Q_host, X=X_host)
        p = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child", path],
                           stdout=subprocess.PIPE, env=env, text=True, timeout=600)
    if p.returncode != 0:
        log(f"cpu baseline child failed ({p.returncode})")
        return None
    r = json.loads(p.stdout.strip().splitlines()[-1])
    info.update({"omp_num_threads": env["OMP_NUM_THREADS"], "omp_proc_bind": "close",
                 "omp_places": "cores"})
    return {"value": r["m"] * r["n"] / r["t_mn"], "unit": "distances/s", "cores": r["threads"],
            "kind": "reference",
            "sample": f"distance_m_to_n, m={r['m']} queries x n={r['n']} codes of the same "
                      f"collection (dim {DIM}, {BITS}-bit), min of 3 runs",
            "cpu": info,
            "legs": {
                "distance_1_to_n": {"dist_per_s": r["n1"] / r["t_1n"],
                                    "sample": f"1 query x {r['n1']} codes, min of 5 runs"},
                "encode": {"vec_per_s": r["n_enc"] / r["t_enc"],
                           "sample": f"encode_batch of {r['n_enc']} vectors of the same recipe, "
                                     "min of 3 runs"}}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--codes", type=int, default=N_CODES)
    ap.add_argument("--queries", type=int, default=M_QUERIES)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched uneven-launch measurement")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving measurement")
    ap.add_argument("--no-recall", action="store_true", help="skip the recall measurement")
    ap.add_argument("--no-symmetric", action="store_true",
                    help="skip the symmetric (dim 1000, 4,096 x 2M) measurement")
    ap.add_argument("--no-asym8", action="store_true",
                    help="skip the 8-bit (dim 768, 1,024 x 1M) measurement")
    ap.add_argument("--serve-seconds", type=float, default=5.0,
                    help="seconds per offered-load point of the serving curves")
    ap.add_argument("--qps-mfma", type=str, default=",".join(str(q) for q in SERVE_QPS["mfma"]),
                    help="offered QPS points of the MFMA serving curve (comma-separated)")
    ap.add_argument("--qps-fp32", type=str, default=",".join(str(q) for q in SERVE_QPS["fp32"]),
                    help="offered QPS points of the plain fp32 serving curve")
    ap.add_argument("--serve-codes", type=int, default=SERVE_N,
                    help="codes per resident collection (4 collections)")
    ap.add_argument("--cpu-child", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--calibrate", action="store_true",
                    help="run one 1 GiB device-to-device copy first (a kernel of known "
                         "traffic, used to scale rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.cpu_child:
        return cpu_child(args.cpu_child)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the distance path runs on a single MI355X)", file=sys.stderr)
        return 2

    import torch
    from turboquant import capi

    torch.cuda.set_device(0)
    if args.calibrate:
        a = torch.empty(1 << 28, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    stream = torch.cuda.current_stream()
    sh = ctypes.c_void_p(stream.cuda_stream)
    sp = capi.Space(DIM, BITS)
    n, m = args.codes, args.queries
    cs = sp.cs

    # ---- resident collection + queries (outside the timed region) --------
    codes = torch.empty((n, cs), dtype=torch.uint8, device="cuda")
    enc_ms = 0.0
    centres = None
    for start, x, centres in gaussian_mixture(torch, n, DIM, 8192, 0.3, 20):
        rows = x.shape[0]
        t = event_time(torch, stream, lambda: sp.encode(x, rows, codes[start:start + rows],
                                                        stream=sh), 1)
        enc_ms += t
    enc_vps = n / (enc_ms / 1e3)
    _, Qd, _ = next(gaussian_mixture(torch, max(m, 128), DIM, 8192, 0.3, 21, centres=centres))
    Qd = Qd.contiguous()
    prep = torch.empty((max(m, 128), sp.prep_stride), dtype=torch.float32, device="cuda")
    out = torch.empty((m, n), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    log(f"collection ready: {n} codes x {cs} B = {n * cs / 1e9:.2f} GB; encode {enc_vps / 1e6:.1f} M vec/s")

    def step(mode):
        sp.prepare(Qd, m, prep, stream=sh)
        sp.scan(prep, m, codes, n, out, mode=mode, stream=sh)

    results = {}
    for name, mode in (("fp32", capi.TQ_MODE_FP32), ("mfma", capi.TQ_MODE_MFMA)):
        try:
            step(mode)
        except capi.TQError as e:
            log(f"{name} path unavailable: {e}")
            continue
        for _ in range(args.warmup):
            step(mode)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        scan_ms = 0.0
        for _ in range(args.steps):
            sp.prepare(Qd, m, prep, stream=sh)
            scan_ms += event_time(torch, stream, lambda: sp.scan(prep, m, codes, n, out,
                                                                 mode=mode, stream=sh), 1)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / args.steps
        results[name] = {"ms_per_step": wall * 1e3, "scan_ms": scan_ms / args.steps,
                         "dist_per_s": m * n / wall}
        log(f"{name}: {wall * 1e3:.2f} ms/step, scan {scan_ms / args.steps:.2f} ms, "
            f"{m * n / wall / 1e9:.2f} G dist/s")

    # ---- distance_1_to_n (one prepared query, HBM-bound scan) -------------
    one = torch.empty(n, dtype=torch.float32, device="cuda")
    sp.scan(prep, 1, codes, n, one, stream=sh)
    t1 = event_time(torch, stream, lambda: sp.scan(prep, 1, codes, n, one, stream=sh), 10)
    bytes_1n = n * (cs + 4) + sp.prep_stride * 4
    one_to_n = {"dist_per_s": n / (t1 / 1e3), "kernel_ms": t1,
                "achieved_GBs": bytes_1n / (t1 / 1e3) / 1e9,
                "frac_hbm": bytes_1n / (t1 / 1e3) / 1e9 / HBM_PEAK_GBS,
                "traffic": measured_traffic("k_lut_wide_1ton")}
    log(f"1-to-N: {t1:.3f} ms, {one_to_n['achieved_GBs']:.0f} GB/s")

    best = "mfma" if "mfma" in results else "fp32"
    r = results[best]
    if best == "mfma":
        flops = 4.0 * sp.D * m * n
        achieved = flops / (r["scan_ms"] / 1e3) / 1e12
        roof = {"bound": "mfma", "achieved": achieved, "peak": FP16_MFMA_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP16_MFMA_PEAK_TFLOPS,
                "traffic": measured_traffic("k_mfma")}
        dtype = "f16 hi+lo MFMA, f32 accumulate"
    else:
        bytes_mn = m * n * (cs + 4)
        achieved = bytes_mn / (r["scan_ms"] / 1e3) / 1e9
        roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("k_lut_wide_mxn")}
        dtype = "f32"

    cpu = None
    if not args.no_cpu:
        n_cpu = min(n, 2_000_000)
        codes_host = codes[:n_cpu].cpu().numpy()
        Q_host = Qd[:128].cpu().numpy()
        _, Xs, _ = next(gaussian_mixture(torch, 100_000, DIM, 8192, 0.3, 22, centres=centres))
        cpu = cpu_baseline(codes_host, Q_host, Xs.cpu().numpy())
        del Xs
        if cpu:
            log(f"cpu baseline ({cpu['cpu']['model']}, {cpu['cores']} threads): m x n "
                f"{cpu['value'] / 1e6:.2f} M dist/s, 1-to-N "
                f"{cpu['legs']['distance_1_to_n']['dist_per_s'] / 1e6:.1f} M dist/s, encode "
                f"{cpu['legs']['encode']['vec_per_s'] / 1e3:.0f} k vec/s")

    # The collection and output of the headline measurement are released
    # before the batched and serving runs build their own.
    del codes, out, one
    torch.cuda.empty_cache()
    recall = None if args.no_recall else bench_recall(torch, capi, sh)
    symmetric = None if args.no_symmetric else bench_symmetric(torch, capi, stream, sh)
    asym8 = None if args.no_asym8 else bench_asym8(torch, capi, stream, sh)
    batched = None if args.no_batched else bench_batched(torch, capi, stream, sh)
    serving = None
    if not args.no_serving:
        pts = {k: [float(q) for q in v.split(",") if q]
               for k, v in (("mfma", args.qps_mfma), ("fp32", args.qps_fp32))}
        serving = bench_serving(torch, capi, sh, pts, args.serve_seconds,
                                n_per_coll=args.serve_codes)

    line = {
        "metric": "distances/s at dim768 4-bit, MFMA vs fp32 GPU path; p99 latency @ fixed QPS",
        "value": r["dist_per_s"], "unit": "distances/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": dtype,
        "data": "synthetic (Gaussian mixture, GPU-generated, encoded on the GPU)",
        "config": {"workload": f"distance_m_to_n dim {DIM} {BITS}-bit: {m} queries x "
                               f"{n} resident codes ({cs} B/code) per step",
                   "dim": DIM, "bits": BITS, "codes": n, "queries_per_step": m, "path": best},
        "roofline": roof,
        "cpu_baseline": cpu,
        "fp32_path_dist_per_s": results.get("fp32", {}).get("dist_per_s"),
        "mfma_path_dist_per_s": results.get("mfma", {}).get("dist_per_s"),
        "distance_1_to_n": one_to_n,
        "encode_vec_per_s": enc_vps,
        "recall": recall,
        "symmetric": symmetric,
        "asym_8bit": asym8,
        "batched_uneven": batched,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
