"""The MFMA form's precision design, emulated on the CPU.

The GPU's m x n MFMA form (tq_mfma.hip) computes the asymmetric estimator
(space_turbo_quant.h:659-686, :951-983) as fp16 matrix products with fp32
accumulation:
    MSE = q_hi*c_hi + q_lo*c_hi + q_hi*c_lo,   QJL = s_hi*(+-1) [+ s_lo*(+-1)]
with q_rot and s_q scaled by 2^10 and centroids by 2^8 before the fp16 split,
and the reference's fp32 epilogue. This file rebuilds that arithmetic in
numpy from the oracle's codes and prepared queries (fp16 products are exact
in fp32; the sums here run in fp64, so this is the design's error without the
accumulation-order part) and measures it against the oracle's FP64
restatement of the estimator on the samples the workloads' recipes give:
  * the relative error |d - d64| / d64 on each query's 10 nearest codes, the
    distances a top-10 decides on, must stay <= 1e-4 (the bar the GPU tests
    hold the MFMA form to) with room to spare;
  * dropping any lo pass the design keeps must break that bar on at least one
    sample, so every pass is shown to be needed.
Byte codes (5-9 bits) keep four passes: their QJL term is scaled by a ~15x
smaller residual norm, and the s_lo pass is shown not to be needed there.
"""
import numpy as np
import pytest

from _oracle import Oracle, nearest, topk_rel_err

BAR = 1e-4
F32 = np.float32


def mixture(n, dim, k, std, seed, normalise, centres=None):
    rng = np.random.default_rng(seed)
    if centres is None:
        centres = rng.standard_normal((k, dim)).astype(F32)
    lab = rng.integers(0, k, n)
    x = centres[lab] + F32(std) * rng.standard_normal((n, dim)).astype(F32)
    if normalise:
        x /= np.linalg.norm(x, axis=1, keepdims=True)
    return np.ascontiguousarray(x, F32), centres


def split16(v32):
    """fp32 -> (hi, lo) fp16 parts as exact float64 values."""
    hi = v32.astype(np.float16)
    lo = (v32 - hi.astype(F32)).astype(np.float16)
    return hi.astype(np.float64), lo.astype(np.float64)


def decode(o, codes):
    """Centroid values and +-1 QJL factors of every coordinate of every code."""
    packed = codes[:, :o.packed]
    if o.bits <= 4:
        units = np.stack([packed & 0xF, packed >> 4], axis=2).reshape(len(codes), -1)[:, :o.D]
    else:
        units = packed[:, :o.D]
    idx = (units >> 1).astype(np.int64)
    return o.centroids[idx], np.where(units & 1, 1.0, -1.0)


def emulate(o, Q, codes, passes):
    """The MFMA form's distances with the given set of passes among
    {"q_lo", "c_lo", "s_lo"} (q_hi*c_hi and s_hi*(+-1) always run)."""
    D = o.D
    prep = o.prepare(Q)
    q, s = prep[:, :D], prep[:, D:2 * D]
    qns, qn = prep[:, 2 * D], prep[:, 2 * D + 1]
    c, b = decode(o, codes)
    tr = o.trailer(codes)
    xn, gamma, sigma = tr[:, 0], tr[:, 1], tr[:, 2]
    qh, ql = split16(q * F32(1024))
    sh, sl = split16(s * F32(1024))
    ch, cl = split16(c * F32(256))
    mse = qh @ ch.T
    if "q_lo" in passes:
        mse += ql @ ch.T
    if "c_lo" in passes:
        mse += qh @ cl.T
    qjl = sh @ b.T
    if "s_lo" in passes:
        qjl += sl @ b.T
    fm = sigma * F32(1.0 / 262144.0)
    fq = F32(o.scale) * gamma * F32(1.0 / 1024.0)
    ip = (mse.astype(F32) * fm[None, :] + qjl.astype(F32) * fq[None, :]) * xn[None, :]
    ip = ip * qn[:, None]
    d = qns[:, None] + (xn * xn)[None, :] - F32(2) * ip
    return np.maximum(d, F32(0))


# (label, dim, bits, centres, seed): the workloads' mixture recipes (std 0.3,
# L2-normalised) with fewer centres than the full collections, so that a
# 4,000-code sample still gives every query same-cluster neighbours.
SAMPLES = [
    ("768/4-bit (dim 768 4-bit workload recipe)", 768, 4, 32, 20),
    ("96/4-bit (batched workload, first collection)", 96, 4, 32, 10),
    ("384/4-bit (batched workload, third collection)", 384, 4, 32, 12),
    ("1536/4-bit (batched workload, fifth collection)", 1536, 4, 32, 14),
    ("768/8-bit (8-bit workload recipe)", 768, 8, 32, 2),
    ("128/8-bit (batched workload, second collection)", 128, 8, 32, 11),
]
DESIGN = {True: {"q_lo", "c_lo", "s_lo"}, False: {"q_lo", "c_lo"}}
_cache = {}


def sample(dim, bits, k, seed):
    key = (dim, bits, k, seed)
    if key not in _cache:
        X, c = mixture(4000, dim, k, 0.3, seed, True)
        Q, _ = mixture(16, dim, k, 0.3, seed + 1000, True, centres=c)
        o = Oracle(dim, bits)
        codes = o.encode(X)
        d64 = o.m_to_n_f64(Q, codes)
        _cache[key] = (o, Q, codes, d64, nearest(d64, 10))
    return _cache[key]


@pytest.mark.parametrize("label,dim,bits,k,seed", SAMPLES)
def test_design_meets_the_bar(label, dim, bits, k, seed):
    o, Q, codes, d64, idx = sample(dim, bits, k, seed)
    d = emulate(o, Q, codes, DESIGN[bits <= 4])
    e = topk_rel_err(d, d64, idx=idx)
    print(f"{label}: design max {e.max():.2e} mean {e.mean():.2e}")
    assert e.max() <= BAR / 4


def test_every_kept_pass_is_needed():
    """Each lo pass of the nibble design, dropped alone, breaks the 1e-4 bar
    on some 4-bit sample; for byte codes the dropped s_lo pass does not."""
    worst = {}
    for label, dim, bits, k, seed in SAMPLES:
        o, Q, codes, d64, idx = sample(dim, bits, k, seed)
        full = DESIGN[bits <= 4]
        for p in sorted(full) + (["s_lo"] if bits > 4 else []):
            passes = full - {p} if p in full else full | {p}
            e = topk_rel_err(emulate(o, Q, codes, passes), d64, idx=idx).max()
            key = (bits <= 4, p)
            worst[key] = max(worst.get(key, 0.0), e)
            print(f"{label}: {'without' if p in full else 'with'} {p}: max {e:.2e}")
    for p in ("q_lo", "c_lo", "s_lo"):
        assert worst[(True, p)] > BAR, f"nibble design without {p} still meets the bar"
    assert worst[(False, "s_lo")] <= BAR / 4
