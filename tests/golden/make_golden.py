"""Write the golden fixtures for the oracle and the GPU path.

Runs only where the reference tree is present: it builds oracle/_ref/libtqref.so
(the reference header compiled in place with the reference's Release flags,
see oracle/Makefile) and records inputs and outputs of that build:

  golden.npz      codebooks (mse bits 1..8), sign vectors, encoded vectors and
                  trailers, prepared queries and distances for small seeded
                  inputs at every config shape plus edge cases;
  manifest.json   compiler, flags, numpy version, seeds, sha256 of the arrays,
                  and sha256 of the full 100,000 x 128 4-bit code matrix of the
                  CPU-runnable workload (X iid N(0,1), seed 0).

Usage: python tests/golden/make_golden.py
"""
import ctypes
import hashlib
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, "..", ".."))
REF_LIB = os.path.join(REPO, "oracle", "_ref", "libtqref.so")

P = ctypes.c_void_p
SZ = ctypes.c_size_t

# (dim, bits): every config shape, plus small/odd dims and every bit width.
SHAPES = [
    (128, 4), (768, 8), (1000, 4), (96, 4), (128, 8), (384, 4), (768, 4), (1536, 4),
    (1, 4), (2, 8), (3, 3), (4, 2), (5, 4), (7, 9), (8, 4), (13, 5), (16, 6), (33, 7),
    (64, 4), (100, 8), (256, 9), (512, 3), (2048, 8),
    # padded dims above 4096 (the reference pads any dim)
    (5000, 4), (7000, 8),
]


def load_ref():
    subprocess.check_call(["make", "-s", "-C", os.path.join(REPO, "oracle"), "ref"])
    L = ctypes.CDLL(REF_LIB)
    L.tqref_create.restype = P
    L.tqref_create.argtypes = [SZ, ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int]
    L.tqref_code_size.restype = SZ
    L.tqref_code_size.argtypes = [P]
    L.tqref_padded_dim.restype = SZ
    L.tqref_padded_dim.argtypes = [P]
    L.tqref_destroy.argtypes = [P]
    for name in ["tqref_encode_batch"]:
        getattr(L, name).argtypes = [P, P, SZ, P]
    L.tqref_prepare_query.argtypes = [P, P, P, P, P, P]
    L.tqref_distance_1_to_n.argtypes = [P, P, P, SZ, P]
    L.tqref_distance_m_to_n.argtypes = [P, P, SZ, P, SZ, P]
    L.tqref_distance_m_to_n_symmetric.argtypes = [P, P, SZ, P, SZ, P]
    L.tqref_distance_m_to_n_symmetric_full.argtypes = [P, P, SZ, P, SZ, P]
    L.tqref_distance_m_to_n_symmetric_light.argtypes = [P, P, SZ, P, SZ, P]
    L.tqref_distance_m_to_n_symmetric_full_prepared.argtypes = [P, P, SZ, P, SZ, P]
    L.tqref_codebook.argtypes = [ctypes.c_int, P, P]
    L.tqref_rotated_coords_batch.argtypes = [P, P, SZ, P]
    L.tqref_signs.argtypes = [SZ, ctypes.c_uint64, P]
    return L


def ptr(a):
    return a.ctypes.data_as(P)


def edge_rows(dim, rng):
    """Zero vector, scaled basis vectors, tiny, huge, constant, integer rows."""
    rows = [np.zeros(dim, np.float32)]
    e = np.zeros(dim, np.float32)
    e[dim // 3] = 2.0
    rows.append(e)
    e2 = np.zeros(dim, np.float32)
    e2[0] = -0.5
    rows.append(e2)
    rows.append((rng.standard_normal(dim) * 1e-7).astype(np.float32))
    rows.append((rng.standard_normal(dim) * 1e6).astype(np.float32))
    rows.append(np.full(dim, 0.25, np.float32))
    rows.append(rng.integers(0, 256, dim).astype(np.float32))
    return np.stack(rows)


def main():
    L = load_ref()
    out = {}
    # Codebooks for every MSE bit count used by bits_per_coord 2..9.
    for mb in range(1, 9):
        c = np.zeros(1 << mb, np.float32)
        b = np.zeros((1 << mb) - 1, np.float32)
        L.tqref_codebook(mb, ptr(c), ptr(b))
        out[f"codebook_c_{mb}"] = c
        out[f"codebook_b_{mb}"] = b
    for seed in (42, 137, 0, 12345):
        s = np.zeros(4096, np.float32)
        L.tqref_signs(4096, seed, ptr(s))
        out[f"signs_{seed}"] = s

    shapes_meta = []
    for k, (dim, bits) in enumerate(SHAPES):
        h = L.tqref_create(dim, bits, 42, 137, 1)
        cs = L.tqref_code_size(h)
        D = L.tqref_padded_dim(h)
        rng = np.random.default_rng(1000 + k)
        n = 24 if dim <= 256 else 8
        X = np.concatenate([edge_rows(dim, rng),
                            rng.standard_normal((n, dim)).astype(np.float32)])
        codes = np.zeros((len(X), cs), np.uint8)
        L.tqref_encode_batch(h, ptr(X), len(X), ptr(codes))
        m = 3
        Q = rng.standard_normal((m, dim)).astype(np.float32)
        prep = np.zeros((m, 2 * D + 2), np.float32)
        for i in range(m):
            row = prep[i]
            L.tqref_prepare_query(h, ptr(Q[i]), row.ctypes.data, row[D:].ctypes.data,
                                  row[2 * D:].ctypes.data, row[2 * D + 1:].ctypes.data)
        dist = np.zeros((m, len(X)), np.float32)
        L.tqref_distance_m_to_n(h, ptr(Q), m, ptr(codes), len(X), ptr(dist))
        d1 = np.zeros(len(X), np.float32)
        L.tqref_distance_1_to_n(h, ptr(Q[0]), ptr(codes), len(X), ptr(d1))
        sym = np.zeros((4, len(X)), np.float32)
        L.tqref_distance_m_to_n_symmetric(h, ptr(codes[-4:]), 4, ptr(codes), len(X), ptr(sym))
        symf = np.zeros((4, len(X)), np.float32)
        L.tqref_distance_m_to_n_symmetric_full(h, ptr(codes[-4:]), 4, ptr(codes), len(X),
                                               ptr(symf))
        syml = np.zeros((4, len(X)), np.float32)
        L.tqref_distance_m_to_n_symmetric_light(h, ptr(codes[-4:]), 4, ptr(codes), len(X),
                                                ptr(syml))
        symfp = np.zeros((4, len(X)), np.float32)
        L.tqref_distance_m_to_n_symmetric_full_prepared(h, ptr(codes[-4:]), 4, ptr(codes),
                                                        len(X), ptr(symfp))
        rot = np.zeros((len(X), D), np.float32)
        L.tqref_rotated_coords_batch(h, ptr(X), len(X), ptr(rot))
        tag = f"{dim}_{bits}"
        out[f"rot_{tag}"] = rot
        out[f"X_{tag}"] = X
        out[f"codes_{tag}"] = codes
        out[f"Q_{tag}"] = Q
        out[f"prep_{tag}"] = prep
        out[f"dist_{tag}"] = dist
        out[f"d1_{tag}"] = d1
        out[f"sym_{tag}"] = sym
        out[f"symfull_{tag}"] = symf
        out[f"symlight_{tag}"] = syml
        out[f"symfullprep_{tag}"] = symfp
        shapes_meta.append({"dim": dim, "bits": bits, "padded_dim": int(D), "code_size": int(cs),
                            "rows": int(len(X)), "queries": m})
        L.tqref_destroy(h)

    # The CPU-runnable workload in full: 100,000 x 128, 4-bit, seeds 42/137,
    # X iid N(0,1) from seed 0 -- only the digest is stored.
    h = L.tqref_create(128, 4, 42, 137, 0)
    X = np.random.default_rng(0).standard_normal((100_000, 128)).astype(np.float32)
    codes = np.zeros((100_000, 76), np.uint8)
    L.tqref_encode_batch(h, ptr(X), 100_000, ptr(codes))
    full_sha = hashlib.sha256(codes.tobytes()).hexdigest()
    out["cfg1_codes_head"] = codes[:64].copy()
    L.tqref_destroy(h)

    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    cxx = subprocess.run(["g++", "--version"], stdout=subprocess.PIPE, text=True).stdout
    manifest = {
        "generator": "tests/golden/make_golden.py",
        "reference_build": {
            "compiler": cxx.splitlines()[0],
            "flags": "-std=c++17 -O3 -ffast-math -fno-finite-math-only -march=x86-64-v3 "
                     "-fopenmp -DTURBOQUANT_HAVE_OPENMP -DNDEBUG",
            "sources": "include/turboquant/space_turbo_quant.h, include/turboquant/turbo_quant.h",
        },
        "numpy": np.__version__,
        "rot_seed": 42,
        "qjl_seed": 137,
        "shapes": shapes_meta,
        "cfg1_full_codes_sha256": full_sha,
        "arrays_sha256": {k: hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest()
                          for k, v in sorted(out.items())},
    }
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    print("wrote", os.path.join(HERE, "golden.npz"),
          os.path.getsize(os.path.join(HERE, "golden.npz")), "bytes")
    return 0


if __name__ == "__main__":
    sys.exit(main())
