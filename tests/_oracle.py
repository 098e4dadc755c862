"""ctypes view of oracle/liboracle.so (the CPU checker) and of the compiled
reference oracle/_ref/libtqref.so where it exists. Test infrastructure only."""
import ctypes
import os

import numpy as np

REPO = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
ORACLE_LIB = os.path.join(REPO, "oracle", "liboracle.so")
REF_LIB = os.path.join(REPO, "oracle", "_ref", "libtqref.so")
GOLDEN = os.path.join(REPO, "tests", "golden", "golden.npz")
MANIFEST = os.path.join(REPO, "tests", "golden", "manifest.json")

P = ctypes.c_void_p
SZ = ctypes.c_size_t

_o = None
_r = None


def ptr(a):
    return a.ctypes.data_as(P)


def olib():
    global _o
    if _o is None:
        L = ctypes.CDLL(ORACLE_LIB)
        L.tqo_create.restype = P
        L.tqo_create.argtypes = [SZ, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64]
        L.tqo_destroy.argtypes = [P]
        for f in ("tqo_padded_dim", "tqo_code_size", "tqo_packed_bytes"):
            getattr(L, f).restype = SZ
            getattr(L, f).argtypes = [P]
        L.tqo_scale.restype = ctypes.c_float
        L.tqo_scale.argtypes = [P]
        for f in ("tqo_centroids", "tqo_boundaries", "tqo_rot_signs", "tqo_qjl_signs"):
            getattr(L, f).restype = ctypes.POINTER(ctypes.c_float)
            getattr(L, f).argtypes = [P]
        L.tqo_splitmix64_next.restype = ctypes.c_uint64
        L.tqo_splitmix64_next.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
        L.tqo_signs.argtypes = [SZ, ctypes.c_uint64, P]
        L.tqo_lloyd_max.argtypes = [ctypes.c_int, P, P]
        L.tqo_sumsq.restype = ctypes.c_float
        L.tqo_sumsq.argtypes = [P, SZ]
        L.tqo_wht.argtypes = [P, SZ]
        L.tqo_rht.argtypes = [P, P, SZ]
        L.tqo_encode_batch.argtypes = [P, P, SZ, P]
        L.tqo_rotated_coords_batch.argtypes = [P, P, SZ, P]
        L.tqo_prepare_query.argtypes = [P, P, P, P, P, P]
        L.tqo_distance_1_to_n.argtypes = [P, P, P, SZ, P]
        L.tqo_distance_m_to_n.argtypes = [P, P, SZ, P, SZ, P]
        L.tqo_distance_m_to_n_symmetric.argtypes = [P, P, SZ, P, SZ, P]
        L.tqo_distance_symmetric_full.restype = ctypes.c_float
        L.tqo_distance_symmetric_full.argtypes = [P, P, P]
        L.tqo_distance_symmetric_light.restype = ctypes.c_float
        L.tqo_distance_symmetric_light.argtypes = [P, P, P]
        L.tqo_distance_symmetric_full_prepared.restype = ctypes.c_float
        L.tqo_distance_symmetric_full_prepared.argtypes = [P, P, P]
        L.tqo_distance_m_to_n_f64.argtypes = [P, P, SZ, P, SZ, P]
        L.tqo_distance_m_to_n_symmetric_f64.argtypes = [P, P, SZ, P, SZ, P]
        L.tqo_distance_m_to_n_symmetric_full_f64.argtypes = [P, P, SZ, P, SZ, ctypes.c_int, P]
        _o = L
    return _o


def ref_available():
    return os.path.exists(REF_LIB)


def rlib():
    global _r
    if _r is None:
        L = ctypes.CDLL(REF_LIB)
        L.tqref_create.restype = P
        L.tqref_create.argtypes = [SZ, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int]
        L.tqref_destroy.argtypes = [P]
        L.tqref_code_size.restype = SZ
        L.tqref_code_size.argtypes = [P]
        L.tqref_encode_batch.argtypes = [P, P, SZ, P]
        L.tqref_distance_1_to_n.argtypes = [P, P, P, SZ, P]
        L.tqref_distance_m_to_n.argtypes = [P, P, SZ, P, SZ, P]
        L.tqref_distance_m_to_n_symmetric.argtypes = [P, P, SZ, P, SZ, P]
        L.tqref_prepare_query.argtypes = [P, P, P, P, P, P]
        _r = L
    return _r


class Oracle:
    """One oracle space."""

    def __init__(self, dim, bits=4, rot_seed=42, qjl_seed=137):
        L = olib()
        self.h = L.tqo_create(dim, bits, rot_seed, qjl_seed)
        assert self.h, "oracle rejected the space"
        self.dim, self.bits = dim, bits
        self.D = L.tqo_padded_dim(self.h)
        self.cs = L.tqo_code_size(self.h)
        self.packed = L.tqo_packed_bytes(self.h)
        self.scale = L.tqo_scale(self.h)
        nl = 1 << (bits - 1)
        self.centroids = np.ctypeslib.as_array(L.tqo_centroids(self.h), (nl,)).copy()
        self.boundaries = np.ctypeslib.as_array(L.tqo_boundaries(self.h), (nl - 1,)).copy()
        self.rot = np.ctypeslib.as_array(L.tqo_rot_signs(self.h), (self.D,)).copy()
        self.qjl = np.ctypeslib.as_array(L.tqo_qjl_signs(self.h), (self.D,)).copy()

    def __del__(self):
        try:
            olib().tqo_destroy(self.h)
        except Exception:
            pass

    def encode(self, X):
        X = np.ascontiguousarray(X, np.float32)
        out = np.zeros((len(X), self.cs), np.uint8)
        olib().tqo_encode_batch(self.h, ptr(X), len(X), ptr(out))
        return out

    def prepare(self, Q):
        Q = np.ascontiguousarray(Q, np.float32)
        out = np.zeros((len(Q), 2 * self.D + 4), np.float32)
        for i in range(len(Q)):
            r = out[i]
            olib().tqo_prepare_query(self.h, ptr(Q[i]), r.ctypes.data, r[self.D:].ctypes.data,
                                     r[2 * self.D:].ctypes.data, r[2 * self.D + 1:].ctypes.data)
        return out

    def m_to_n(self, Q, codes):
        Q = np.ascontiguousarray(Q, np.float32)
        out = np.zeros((len(Q), len(codes)), np.float32)
        olib().tqo_distance_m_to_n(self.h, ptr(Q), len(Q), ptr(codes), len(codes), ptr(out))
        return out

    def m_to_n_f64(self, Q, codes):
        Q = np.ascontiguousarray(Q, np.float32)
        out = np.zeros((len(Q), len(codes)), np.float64)
        olib().tqo_distance_m_to_n_f64(self.h, ptr(Q), len(Q), ptr(codes), len(codes), ptr(out))
        return out

    def symmetric(self, A, B):
        out = np.zeros((len(A), len(B)), np.float32)
        olib().tqo_distance_m_to_n_symmetric(self.h, ptr(A), len(A), ptr(B), len(B), ptr(out))
        return out

    def symmetric_f64(self, A, B):
        out = np.zeros((len(A), len(B)), np.float64)
        olib().tqo_distance_m_to_n_symmetric_f64(self.h, ptr(A), len(A), ptr(B), len(B), ptr(out))
        return out

    def symmetric_full_f64(self, A, B, light=False):
        out = np.zeros((len(A), len(B)), np.float64)
        olib().tqo_distance_m_to_n_symmetric_full_f64(self.h, ptr(A), len(A), ptr(B), len(B),
                                                      1 if light else 0, ptr(out))
        return out

    def trailer(self, codes):
        """[n, 3] float32 (norm, gamma, sigma)."""
        return np.ascontiguousarray(codes[:, self.packed:self.packed + 12]).view(np.float32)


def load_golden():
    return np.load(GOLDEN, allow_pickle=False)


def scale_rel_err(got, want, qns, xns):
    """|got - want| / (||q||^2 + ||x||^2): the distance's natural scale (the
    estimator subtracts 2*ip from this sum, so it bounds every term)."""
    denom = np.asarray(qns)[:, None] + np.asarray(xns)[None, :]
    denom = np.maximum(denom, 1e-30)
    return np.abs(got.astype(np.float64) - want.astype(np.float64)) / denom


def fp32_parity_excess(got, oracle32, exact, qns, xns):
    """How far the GPU's fp32 result is from the fp32 oracle beyond the
    oracle's own rounding error, both measured against the fp64 restatement of
    the estimator, in units of ||q||^2 + ||x||^2. The fp32 bar is
    fp32_parity_excess <= 1e-5: the reference sums in fp32 too, and over long
    same-sign sums (self-distances at large D) its own error approaches 1e-5,
    while the GPU's tree-ordered sums stay well below it."""
    return (scale_rel_err(got, oracle32, qns, xns) - scale_rel_err(oracle32, exact, qns, xns))


# Floor of the relative error |d - d64| / max(d64, eps): eps = EPS_FLOOR *
# (||q||^2 + ||x||^2), i.e. distances below 0.1% of their scale count as that.
EPS_FLOOR = 1e-3


def eps_rel_err(got, want, qns, xns, floor=EPS_FLOOR):
    """|got - want| / max(want, floor * (||q||^2 + ||x||^2)), elementwise."""
    scale = np.asarray(qns)[:, None] + np.asarray(xns)[None, :]
    denom = np.maximum(want.astype(np.float64), floor * scale)
    denom = np.maximum(denom, 1e-30)
    return np.abs(got.astype(np.float64) - want.astype(np.float64)) / denom


def nearest(d, k=10):
    """[m, k] column indices of each row's k smallest entries (ascending)."""
    idx = np.argpartition(d, k, axis=1)[:, :k]
    order = np.argsort(np.take_along_axis(d, idx, axis=1), axis=1, kind="stable")
    return np.take_along_axis(idx, order, axis=1)


def topk_rel_err(got, want, k=10, idx=None, floor=None):
    """|got - want| / want on each row's k nearest columns by `want` (or the
    columns `idx`): the relative error of the distances that decide a top-k.
    With `floor` ([m, n], e.g. EPS_FLOOR * (||q||^2 + ||x||^2)) the
    denominator is max(want, floor) (exact duplicates have want = 0)."""
    if idx is None:
        idx = nearest(want, k)
    g = np.take_along_axis(got.astype(np.float64), idx, axis=1)
    w = np.take_along_axis(want.astype(np.float64), idx, axis=1)
    den = np.maximum(w, 1e-30)
    if floor is not None:
        den = np.maximum(den, np.take_along_axis(np.asarray(floor, np.float64), idx, axis=1))
    return np.abs(g - w) / den
