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
reported beside it, with the reference's own CPU build timed on this host.
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


def gaussian_mixture(torch, n, dim, k, std, seed, centres=None, chunk=1_000_000):
    """Rows of gm(n, dim, k, std, normalised) generated on the GPU in chunks."""
    g = torch.Generator(device="cuda").manual_seed(seed)
    if centres is None:
        centres = torch.randn((k, dim), generator=g, device="cuda")
    for start in range(0, n, chunk):
        rows = min(chunk, n - start)
        lab = torch.randint(0, k, (rows,), generator=g, device="cuda")
        x = centres[lab] + std * torch.randn((rows, dim), generator=g, device="cuda")
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


def cpu_baseline(codes_host, Q_host, budget_s=8.0):
    """The reference CPU build (oracle/_ref, compiled in place from the
    reference sources with its Release flags) on this host's cores:
    distance_m_to_n with m = 128 (> 64, so its OpenMP loop engages) over a
    prefix of the collection sized to the time budget; min of 3 runs."""
    lib_path = os.path.join(REPO, "oracle", "_ref", "libtqref.so")
    if not os.path.exists(lib_path):
        return None
    L = ctypes.CDLL(lib_path)
    P, SZ = ctypes.c_void_p, ctypes.c_size_t
    L.tqref_create.restype = P
    L.tqref_create.argtypes = [SZ, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int]
    L.tqref_num_threads.restype = ctypes.c_int
    L.tqref_num_threads.argtypes = [P]
    L.tqref_distance_m_to_n.argtypes = [P, P, SZ, P, SZ, P]
    L.tqref_destroy.argtypes = [P]
    h = L.tqref_create(DIM, BITS, 42, 137, 0)
    threads = L.tqref_num_threads(h)
    m = 128
    Q = np.ascontiguousarray(Q_host[:m])
    n = 20_000
    out = np.zeros((m, n), np.float32)
    t0 = time.perf_counter()
    L.tqref_distance_m_to_n(h, Q.ctypes.data, m, codes_host.ctypes.data, n, out.ctypes.data)
    per = (time.perf_counter() - t0) / (m * n)
    n = int(min(len(codes_host), max(20_000, budget_s / 3 / per / m)))
    out = np.zeros((m, n), np.float32)
    best = float("inf")
    for _ in range(3):
        t0 = time.perf_counter()
        L.tqref_distance_m_to_n(h, Q.ctypes.data, m, codes_host.ctypes.data, n, out.ctypes.data)
        best = min(best, time.perf_counter() - t0)
    L.tqref_destroy(h)
    return {"value": m * n / best, "unit": "distances/s", "cores": threads,
            "kind": "reference",
            "sample": f"distance_m_to_n, m={m} queries x n={n} codes of the same collection "
                      f"(dim {DIM}, {BITS}-bit), min of 3 runs, OMP_NUM_THREADS={threads}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--codes", type=int, default=N_CODES)
    ap.add_argument("--queries", type=int, default=M_QUERIES)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate", action="store_true",
                    help="run one 1 GiB device-to-device copy first (a kernel of known "
                         "traffic, used to scale rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
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
                "frac_hbm": bytes_1n / (t1 / 1e3) / 1e9 / HBM_PEAK_GBS}
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
                "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic("k_lut_nib")}
        dtype = "f32"

    cpu = None
    if not args.no_cpu:
        n_cpu = min(n, 2_000_000)
        codes_host = codes[:n_cpu].cpu().numpy()
        Q_host = Qd[:128].cpu().numpy()
        cpu = cpu_baseline(codes_host, Q_host)
        if cpu:
            log(f"cpu baseline: {cpu['value'] / 1e6:.2f} M dist/s on {cpu['cores']} threads")

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
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
