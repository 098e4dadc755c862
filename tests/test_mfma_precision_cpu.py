"""The MFMA form's precision design, emulated on the CPU.

The GPU's m x n MFMA form (tq_mfma.hip) computes the asymmetric estimator
(space_turbo_quant.h:659-686, :951-983) as fp16 matrix products with fp32
accumulation, q_rot and s_q scaled by 2^10 before the fp16 hi/lo split:
  * 4-bit codes at padded dim >= 1024, the coarse pass set: MSE = q_hi*c_hi
    + q_lo*c_hi with the centroids scaled by alpha = 413.354 (chosen at
    space creation so that fp16(c * alpha) / alpha is within 2^-15 of every
    centroid), QJL = s_hi*(+-1); every pair whose coarse distance is below
    half of (||q||^2 + ||x||^2) then gets the dropped terms added (q_hi times
    the exact fp32 centroid remainder, s_lo*(+-1)), which makes it the full
    pass set's value;
  * byte codes at padded dim >= 1024, the same coarse pass set with the
    centroids scaled by 2^8 (no single scale fits 128 centroids), and only
    q_hi times the exact remainder added to the near pairs (byte codes take
    no s_lo in either pass set);
  * 2-3-bit codes and smaller dims: MSE = q_hi*c_hi + q_lo*c_hi + q_hi*c_lo
    with centroids scaled by 2^8, QJL = s_hi*(+-1) [+ s_lo*(+-1) for 2-4
    bits];
and the reference's fp32 epilogue. This file rebuilds that arithmetic in
numpy from the oracle's codes and prepared queries (fp16 products are exact
in fp32; the sums here run in fp64, so this is the design's error without the
accumulation-order part) and measures it against the oracle's FP64
restatement of the estimator on the samples the workloads' recipes give:
  * the relative error |d - d64| / d64 on each query's 10 nearest codes, the
    distances a top-10 decides on, must stay <= 1e-4 (the bar the GPU tests
    hold the MFMA form to) with room to spare, and so must the relative error
    over every pair (|d - d64| / max(d64, 1e-3 (||q||^2 + ||x||^2)));
  * dropping the q_lo pass or the correction of near pairs from the
    coarse designs, or any lo pass from the full designs, must break that
    bar on at least one sample, so every part is shown to be needed; byte
    codes do not need s_lo (their QJL term is scaled by a ~15x smaller
    residual norm), and the coarse form at padded dim 128 (byte codes) keeps
    under half the bar's margin on far pairs.
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


def coarse_alpha(cent):
    """The coarse form's centroid scale, as the space's creation picks it:
    alpha in [256, 512) on a 2^-12 grid minimising the worst relative error
    of fp16(fl32(|c| * alpha)) * fl32(1 / (2^10 alpha)) * 2^10 (the first
    minimum)."""
    a = 256.0 + np.arange(1 << 20) * 2.0 ** -12
    u = (1.0 / (1024.0 * a)).astype(F32).astype(np.float64)
    worst = np.zeros_like(a)
    for c in np.abs(cent[len(cent) // 2:]).astype(F32):
        v = (c * a.astype(F32)).astype(np.float16).astype(np.float64)
        worst = np.maximum(worst, np.abs(v * u * 1024.0 - float(c)) / float(c))
    k = int(np.argmin(worst))
    return (a[k], worst[k]) if worst[k] <= 2.0 ** -15 else (256.0, worst[k])


def emulate(o, Q, codes, passes):
    """The MFMA form's distances with the given set of passes among
    {"q_lo", "c_lo", "s_lo"} (q_hi*c_hi and s_hi*(+-1) always run), and with
    "refine" the coarse form's fp32 recomputation of the near pairs (and its
    centroid scale)."""
    D = o.D
    prep = o.prepare(Q)
    q, s = prep[:, :D], prep[:, D:2 * D]
    qns, qn = prep[:, 2 * D], prep[:, 2 * D + 1]
    c, b = decode(o, codes)
    tr = o.trailer(codes)
    xn, gamma, sigma = tr[:, 0], tr[:, 1], tr[:, 2]
    nib = o.bits <= 4
    alpha = coarse_alpha(o.centroids)[0] if "refine" in passes and nib else 256.0
    cl_in = "c_lo" in passes and "refine" not in passes
    qh, ql = split16(q * F32(1024))
    sh, sl = split16(s * F32(1024))
    cs = (c * F32(alpha)).astype(F32)
    ch = cs.astype(np.float16).astype(np.float64)
    # c_lo: the fp16 remainder in the MFMA pass; the refinement adds the exact
    # fp32 remainder (refine_delta)
    cl = (cs - ch.astype(F32)).astype(np.float16).astype(np.float64)
    cl_exact = (cs - ch.astype(F32)).astype(np.float64)
    fm = sigma * F32(1.0 / (1024.0 * alpha))
    fq = F32(o.scale) * gamma * F32(1.0 / 1024.0)
    scale = qns[:, None] + (xn * xn)[None, :]

    def finish(mse, qjl):
        ip = (mse.astype(F32) * fm[None, :] + qjl.astype(F32) * fq[None, :]) * xn[None, :]
        ip = ip * qn[:, None]
        return np.maximum(qns[:, None] + (xn * xn)[None, :] - F32(2) * ip, F32(0))

    mse = qh @ ch.T
    if "q_lo" in passes:
        mse += ql @ ch.T
    if cl_in:
        mse += qh @ cl.T
    qjl = sh @ b.T
    if "s_lo" in passes:
        qjl += sl @ b.T
    d = finish(mse, qjl)
    if "refine" in passes:  # the near pairs get the dropped terms added
        near = d < F32(0.5) * scale
        d = np.where(near, finish(mse + qh @ cl_exact.T, qjl + sl @ b.T if nib else qjl), d)
    return d


def all_pairs_rel(d, d64, o, Q, codes):
    """|d - d64| / max(d64, 1e-3 (||q||^2 + ||x||^2)) over every pair."""
    qns = np.sum(Q.astype(np.float64) ** 2, 1)
    xns = o.trailer(codes)[:, 0].astype(np.float64) ** 2
    floor = 1e-3 * (qns[:, None] + xns[None, :])
    return np.abs(d - d64) / np.maximum(d64, floor)


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
    ("384/8-bit (padded dim 512, bytes with every lo pass but s_lo)", 384, 8, 32, 13),
]
_cache = {}


def design(o):
    """The pass set the GPU runs for the space (mfma_coarse): the coarse set
    with the near pairs corrected for 4-bit and byte codes at padded dim >=
    1024, else every lo pass (but s_lo for bytes)."""
    if o.D >= 1024 and o.bits >= 4:
        return {"q_lo", "refine"}
    return {"q_lo", "c_lo", "s_lo"} if o.bits <= 4 else {"q_lo", "c_lo"}


def kind(o):
    return ("4-bit" if o.bits <= 4 else "byte") + (" coarse" if "refine" in design(o) else "")


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
    d = emulate(o, Q, codes, design(o))
    e = topk_rel_err(d, d64, idx=idx)
    a = all_pairs_rel(d, d64, o, Q, codes)
    print(f"{label}: design top-10 max {e.max():.2e} mean {e.mean():.2e}, "
          f"all pairs max {a.max():.2e}")
    assert e.max() <= BAR / 4
    assert a.max() <= BAR / 2


def test_coarse_scale():
    """The coarse form's centroid scale: 413.354 for the 3-bit codebook,
    within 2^-15 of every centroid."""
    alpha, worst = coarse_alpha(Oracle(128, 4).centroids)
    assert abs(alpha - 413.354) < 1e-3 and worst <= 2.0 ** -15
    assert alpha * 4096.0 == int(alpha * 4096.0)  # on the 2^-12 grid


def test_every_kept_part_is_needed():
    """Without q_lo, or without the correction of near pairs, the coarse
    designs break the 1e-4 bar on some sample; so do the full designs
    without any of their lo passes. Byte codes do not need s_lo, and the
    coarse form at padded dim 128 (byte codes: c_lo swapped for the
    correction) leaves under half the bar's margin on far pairs."""
    worst = {}
    for label, dim, bits, k, seed in SAMPLES:
        o, Q, codes, d64, idx = sample(dim, bits, k, seed)
        full = design(o)
        trials = [(p, full - {p}) for p in sorted(full)]
        if bits > 4:
            trials.append(("s_lo", full | {"s_lo"}))
        if kind(o) == "byte" and o.D == 128:
            trials.append(("coarse", {"q_lo", "refine"}))
        for p, passes in trials:
            d = emulate(o, Q, codes, passes)
            e = topk_rel_err(d, d64, idx=idx).max()
            a = all_pairs_rel(d, d64, o, Q, codes).max()
            key = (kind(o), p)
            worst[key] = max(worst.get(key, 0.0), e, a)
            print(f"{label}: {sorted(passes)}: top-10 max {e:.2e}, all pairs max {a:.2e}")
    for p in ("q_lo", "refine"):
        for kd in ("4-bit coarse", "byte coarse"):
            assert worst[(kd, p)] > BAR, f"{kd} design without {p} still meets the bar"
    for p in ("q_lo", "c_lo", "s_lo"):
        assert worst[("4-bit", p)] > BAR, f"4-bit design without {p} still meets the bar"
    for p in ("q_lo", "c_lo"):
        assert worst[("byte", p)] > BAR, f"byte design without {p} still meets the bar"
    assert max(worst[("byte", "s_lo")], worst[("byte coarse", "s_lo")]) <= BAR / 2
    assert worst[("byte", "coarse")] > BAR / 2


def _units(o, codes):
    """Code units (index << 1 | QJL bit) of every coordinate."""
    packed = codes[:, :o.packed]
    if o.bits <= 4:
        return np.stack([packed & 0xF, packed >> 4], axis=2).reshape(len(codes), -1)[:, :o.D]
    return packed[:, :o.D].astype(np.int64)


def correction_sums(qh, clo_u, sl, sgn, nib):
    """refine_delta's totals for one query row against codes: per 8-coordinate
    group (element order of the correction row) an fp32 fused multiply-add
    chain of q_hi * c_lo and an fp32 sum of +-s_lo, each rounded to a 32-bit
    fixed-point integer (2^13, 2^22), the integers summed, the totals scaled
    back to fp32. clo_u, sgn: [codes, D] exact remainders and +-1 of the
    codes' units."""
    f32 = np.float32
    D = qh.shape[-1]
    perm = np.array([0, 2, 4, 6, 1, 3, 5, 7] if nib else list(range(8)))
    order = (np.arange(D // 8)[:, None] * 8 + perm[None, :]).reshape(-1)
    prod = (qh[None, order] * clo_u[:, order].astype(np.float64)).reshape(len(clo_u), -1, 8)
    pm = np.zeros(prod.shape[:2], f32)
    for e in range(8):
        pm = (pm.astype(np.float64) + prod[:, :, e]).astype(f32)
    dm = np.rint(np.ldexp(pm.astype(np.float64), 13)).sum(1).astype(f32) * f32(2.0 ** -13)
    if not nib:
        return dm, np.zeros(len(clo_u), f32)
    t = (sl[None, order] * sgn[:, order]).reshape(len(clo_u), -1, 8)
    pq = np.zeros(t.shape[:2], f32)
    for e in range(8):
        pq = (pq + t[:, :, e].astype(f32)).astype(f32)
    dq = np.rint(np.ldexp(pq.astype(np.float64), 22)).sum(1).astype(f32) * f32(2.0 ** -22)
    return dm, dq


def refine_limits(D, l1_qhi, l1_slo, clo_max, nib):
    """tq_mfma_tile.h refine_limits: bounds on |dm| and |dq| of a query row."""
    f32 = np.float32
    ng = f32(D // 8)
    dmb = (f32(l1_qhi) * f32(clo_max) + ng * f32(2.0 ** -14)) * f32(1 + 2.0 ** -10)
    dqb = (f32(l1_slo) + ng * f32(2.0 ** -23)) * f32(1 + 2.0 ** -10) if nib else f32(0)
    return dmb, dqb


def coarse_tables(o):
    """Scaled centroids' fp16 parts and exact fp32 remainders (cent_lo), and
    the largest remainder the code units can index (SpaceDev::clo_max)."""
    f32 = np.float32
    alpha = coarse_alpha(o.centroids)[0] if o.bits <= 4 else 256.0
    cs = (o.centroids * f32(alpha)).astype(f32)
    chi = cs.astype(np.float16).astype(f32)
    clo = (cs - chi).astype(f32)
    return alpha, chi.astype(np.float64), clo, float(np.abs(clo[:min(len(clo), 128)]).max())


@pytest.mark.parametrize("bits", [4, 8])
def test_correction_limits_hold_on_a_worst_case_row(bits):
    """refine_limits bounds |dm| by sum |q_hi| * max |c_lo|: a query row whose
    signs all match the remainder of the one centroid every coordinate of the
    code uses reaches that bound to within 1%, so a bound from the Euclidean
    norm of q_hi (32x smaller at D = 1024) would fail here; likewise |dq| <=
    sum |s_lo| when every QJL sign matches s_lo's."""
    f32 = np.float32
    o = Oracle(768, bits)
    D, nib = o.D, bits <= 4
    _, _, clo, clo_max = coarse_tables(o)
    kmax = int(np.argmax(np.abs(clo[:min(len(clo), 128)])))
    rng = np.random.default_rng(5)
    q = (np.sign(clo[kmax]) * rng.uniform(0.5, 1.5, D)).astype(f32)
    q /= np.linalg.norm(q)
    qh = split16(q * f32(1024))[0]
    s = rng.standard_normal(D).astype(f32)
    s /= np.linalg.norm(s)
    sl = split16(s * f32(1024))[1]
    clo_u = np.full((1, D), clo[kmax], f32)
    sgn = np.sign(sl)[None, :]
    sgn[sgn == 0] = 1.0
    dm, dq = correction_sums(qh, clo_u, sl, sgn, nib)
    dmb, dqb = refine_limits(D, np.abs(qh).sum(), np.abs(sl).sum(), clo_max, nib)
    assert abs(dm[0]) <= dmb and abs(dm[0]) >= 0.99 * dmb
    assert np.sqrt((qh ** 2).sum()) * clo_max < 0.05 * abs(dm[0])
    if nib:
        assert abs(dq[0]) <= dqb and abs(dq[0]) >= 0.99 * dqb


@pytest.mark.parametrize("label,dim,bits,k,seed", [s for s in SAMPLES if s[1] >= 768])
def test_pruning_bound_covers_the_correction(label, dim, bits, k, seed):
    """The search forms leave a near pair uncorrected when its coarse distance
    minus E (refine_bound, tq_mfma_tile.h) exceeds a bound on the query's
    k-th best; that is exact only if E >= |corrected d - coarse d| for every
    pair. This rebuilds the device arithmetic in fp32 (the coarse sums, the
    correction totals, the epilogue's distance) and E from the query's L1
    norms and the space's largest centroid remainder, and checks the
    inequality on every pair of the coarse-form samples (and |dm|, |dq|
    against their limits)."""
    o, Q, codes, d64, idx = sample(dim, bits, k, seed)
    assert design(o) == {"q_lo", "refine"}
    f32 = np.float32
    D, nib = o.D, o.bits <= 4
    prep = o.prepare(Q)
    q, s = prep[:, :D], prep[:, D:2 * D]
    qns, qn = prep[:, 2 * D], prep[:, 2 * D + 1]
    tr = o.trailer(codes)
    xn, gamma, sigma = tr[:, 0], tr[:, 1], tr[:, 2]
    alpha, chi, clo, clo_max = coarse_tables(o)
    qh, ql = split16(q * f32(1024))
    sh, sl = split16(s * f32(1024))
    units = _units(o, codes)
    idx_u, sgn = units >> 1, np.where(units & 1, 1.0, -1.0)
    am_c = ((qh + ql) @ chi[idx_u].T).astype(f32)
    aq_c = (sh @ sgn.T).astype(f32)
    fm = (sigma * f32(1.0 / (1024.0 * alpha))).astype(f32)
    fq = (f32(o.scale) * gamma * f32(1.0 / 1024.0)).astype(f32)
    xns = (xn * xn).astype(f32)

    def dist(am, aq, i):
        ip = (am * fm + aq * fq) * xn * qn[i]
        return np.maximum((qns[i] + xns) - f32(2) * ip, f32(0))

    worst = 0.0
    for i in range(len(Q)):
        dm, dq = correction_sums(qh[i], clo[idx_u], sl[i], sgn, nib)
        dmb, dqb = refine_limits(D, np.abs(qh[i]).sum(), np.abs(sl[i]).sum(), clo_max, nib)
        assert (np.abs(dm) <= dmb).all() and (np.abs(dq) <= dqb).all()
        d_c = dist(am_c[i], aq_c[i], i)
        d_r = dist((am_c[i] + dm).astype(f32), (aq_c[i] + dq).astype(f32), i)
        fa, ga = np.abs(fm), np.abs(fq)
        tt = (dmb * fa + dqb * ga) * xn * qn[i]
        pp = (np.abs(am_c[i]) * fa + np.abs(aq_c[i]) * ga) * xn * qn[i]
        E = (f32(2) * tt * f32(1 + 2.0 ** -18)
             + f32(2.0 ** -18) * (pp + qns[i] + xns)) * f32(1 + 2.0 ** -16)
        diff = np.abs(d_r.astype(np.float64) - d_c.astype(np.float64))
        assert (diff <= E).all(), f"{label}: query {i}: bound exceeded by {(diff - E).max():.3e}"
        worst = max(worst, float((diff / E).max()))
    print(f"{label}: max |corrected - coarse| / E = {worst:.3f}")
    assert worst > 1e-3  # the check compares against a correction that is really there
