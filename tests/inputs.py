"""Seeded synthetic inputs shared by the tests, the fixture generator and
the benchmark.  numpy's Philox counter-based generator; every matrix is
returned column-major (Fortran order), Eigen's default storage."""
import hashlib

import numpy as np


def rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def normal(seed, m, n=None, dtype=np.float64):
    shape = (m,) if n is None else (m, n)
    a = rng(seed).standard_normal(shape)
    return np.asfortranarray(a.astype(dtype))


def spd(seed, n, dtype=np.float64, shift=None):
    """A = B B^T + n I (the batched-LLT construction), B N(0,1)."""
    b = rng(seed).standard_normal((n, n))
    a = b @ b.T + (n if shift is None else shift) * np.eye(n)
    return np.asfortranarray(a.astype(dtype))


def sha256(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes(order="F")).hexdigest()


def packed_systems(seed, count, sizes, dtype, spd_systems, nrhs=1):
    """Packed variable-size systems: sizes drawn uniformly from `sizes` (a
    list) or uniform integers in [lo, hi] (a tuple).  Returns n (int32),
    offA (int64), A (flat), offB (int64), B (flat, n_i x nrhs col-major)."""
    g = rng(seed)
    if isinstance(sizes, tuple):
        n = g.integers(sizes[0], sizes[1] + 1, size=count).astype(np.int32)
    else:
        n = np.asarray(sizes, dtype=np.int32)[g.integers(0, len(sizes), size=count)]
    n64 = n.astype(np.int64)
    offA = np.zeros(count, dtype=np.int64)
    offA[1:] = np.cumsum(n64 * n64)[:-1]
    offB = np.zeros(count, dtype=np.int64)
    offB[1:] = np.cumsum(n64 * nrhs)[:-1]
    totA = int((n64 * n64).sum())
    A = np.empty(totA, dtype=dtype)
    for i in range(count):
        k = int(n[i])
        b = g.standard_normal((k, k))
        a = b @ b.T + k * np.eye(k) if spd_systems else b
        A[offA[i]:offA[i] + k * k] = a.astype(dtype).ravel(order="F")
    B = g.standard_normal(int((n64 * nrhs).sum())).astype(dtype)
    return n, offA, A, offB, B


def ldlt_matrix(kind, n):
    """Symmetric test matrices of the LDLT fixtures: "sym" (indefinite),
    "pos" (positive definite) or "neg" (negative definite)."""
    if kind == "sym":
        g = normal(7300 + n, n, n)
        return np.asfortranarray(g + g.T)
    if kind == "pos":
        return spd(7400 + n, n)
    return np.asfortranarray(-spd(7500 + n, n))
