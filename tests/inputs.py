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


# ------------------------------------------------- device-side generators --
def packed_spd_gpu(count, sizes, seed, device="cuda"):
    """Packed FP32 SPD systems A_i = B B^T + n_i I (B iid N(0,1)), b_i N(0,1);
    sizes drawn uniformly from `sizes`, generated on the GPU with torch.
    Returns n (int32), offA, A, offb, b (device tensors)."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    sz = torch.tensor(sizes, dtype=torch.int32, device=device)
    pick = torch.randint(0, len(sizes), (count,), generator=g, device=device)
    n = sz[pick]
    n64 = n.to(torch.int64)
    offA = torch.zeros(count, dtype=torch.int64, device=device)
    offA[1:] = torch.cumsum(n64 * n64, 0)[:-1]
    offb = torch.zeros(count, dtype=torch.int64, device=device)
    offb[1:] = torch.cumsum(n64, 0)[:-1]
    A = torch.empty(int((n64 * n64).sum()), dtype=torch.float32, device=device)
    for s in sizes:
        idx = torch.nonzero(n == s).flatten()
        k = idx.numel()
        if k == 0:
            continue
        for c0 in range(0, k, 65536):
            ii = idx[c0:c0 + 65536]
            Bm = torch.randn(ii.numel(), s, s, generator=g, device=device, dtype=torch.float32)
            Am = torch.bmm(Bm, Bm.transpose(1, 2).contiguous())
            Am += s * torch.eye(s, device=device, dtype=torch.float32)
            flat = offA[ii][:, None] + torch.arange(s * s, device=device)[None, :]
            A[flat.flatten()] = Am.transpose(1, 2).reshape(-1)
    b = torch.randn(int(n64.sum()), generator=g, device=device, dtype=torch.float32)
    return n, offA, A, offb, b


def packed_general_gpu(count, lo, hi, nrhs, seed, device="cuda"):
    """Packed FP64 general systems, A_i iid N(0,1) with n_i uniform integer in
    [lo, hi], B_i n_i x nrhs N(0,1), generated on the GPU with torch.
    Returns n (int32), offA, A, offB, B (device tensors)."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    n = torch.randint(lo, hi + 1, (count,), generator=g, device=device, dtype=torch.int64)
    offA = torch.zeros(count, dtype=torch.int64, device=device)
    offA[1:] = torch.cumsum(n * n, 0)[:-1]
    offB = torch.zeros(count, dtype=torch.int64, device=device)
    offB[1:] = torch.cumsum(n * nrhs, 0)[:-1]
    A = torch.randn(int((n * n).sum()), generator=g, device=device, dtype=torch.float64)
    B = torch.randn(int((n * nrhs).sum()), generator=g, device=device, dtype=torch.float64)
    return n.to(torch.int32), offA, A, offB, B
