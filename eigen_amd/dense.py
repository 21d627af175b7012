"""Eigen-shaped host interface over the C ABI (device-resident torch tensors).

Mirrors the reference's decomposition classes:

* ``PartialPivLU``  -- ``compute() / solve() / matrixLU() / permutationP() /
  determinant() / inverse() / rcond() / reconstructedMatrix()``
  (Eigen/src/LU/PartialPivLU.h:98-221)
* ``LLT``           -- ``compute() / solve() / solveInPlace() / matrixL() /
  matrixLLT() / info() / rcond()`` (Eigen/src/Cholesky/LLT.h:84-203)
* ``gemm`` (``C.noalias() = A*B``), ``triangular_solve``
  (``TriangularView::solveInPlace``), ``rank_update`` (``SelfAdjointView::
  rankUpdate``).

Matrices are column-major like Eigen's default ``Matrix``: a torch tensor
with strides (1, ld).  ``colmajor(t)`` converts any 2-D tensor.
"""
import ctypes as C

import torch

from . import _lib
from ._lib import check, lib

Success, NumericalIssue, NoConvergence, InvalidInput = 0, 1, 2, 3  # Constants.h:434-444

_tls = None


class Handle:
    """One library handle (HIP stream + cached workspace) per host thread."""

    def __init__(self):
        h = C.c_void_p()
        check("eamd_create", lib().eamd_create(C.byref(h)))
        self.ptr = h

    def bind_current_stream(self):
        s = torch.cuda.current_stream().cuda_stream
        check("eamd_set_stream", lib().eamd_set_stream(self.ptr, C.c_void_p(s)))
        return self.ptr

    def __del__(self):
        try:
            if self.ptr:
                lib().eamd_destroy(self.ptr)
        except Exception:
            pass


def handle():
    """Thread-local handle bound to torch's current stream."""
    import threading
    global _tls
    if _tls is None:
        _tls = threading.local()
    if getattr(_tls, "h", None) is None:
        _tls.h = Handle()
    return _tls.h.bind_current_stream()


def colmajor(t, dtype=None):
    """A column-major (Fortran-ordered) copy/view of a 2-D tensor on the GPU."""
    if t.dim() == 1:
        t = t.reshape(-1, 1)
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    if not t.is_cuda:
        t = t.cuda()
    if t.stride(0) == 1 and t.stride(1) >= max(1, t.shape[0]):
        return t
    return t.t().contiguous().t()


def _ld(t):
    if t.dim() == 2 and t.shape[1] > 1:
        return max(1, t.stride(1))
    return max(1, t.shape[0])  # a single column: any ld >= rows is valid


def _p(t):
    return C.c_void_p(t.data_ptr())


def gemm(A, B, C_=None, alpha=1.0, beta=0.0, transa="N", transb="N"):
    """C = alpha*op(A)*op(B) + beta*C (FP64), general_matrix_matrix_product."""
    A = colmajor(A, torch.float64)
    B = colmajor(B, torch.float64)
    m = A.shape[0] if transa == "N" else A.shape[1]
    k = A.shape[1] if transa == "N" else A.shape[0]
    n = B.shape[1] if transb == "N" else B.shape[0]
    if C_ is None:
        C_ = colmajor(torch.zeros(m, n, dtype=torch.float64, device=A.device))
        beta = 0.0
    else:
        C_ = colmajor(C_, torch.float64)
    check("eamd_dgemm", lib().eamd_dgemm(handle(), transa.encode(), transb.encode(), m, n, k,
                                         alpha, _p(A), _ld(A), _p(B), _ld(B), beta, _p(C_),
                                         _ld(C_)))
    return C_


def rank_update(C_, A, alpha=-1.0, uplo="L", beta=1.0):
    """C(uplo) = beta*C + alpha*A*A^T in place (SelfAdjointView::rankUpdate)."""
    A = colmajor(A, torch.float64)
    n, k = A.shape
    check("eamd_dsyrk", lib().eamd_dsyrk(handle(), uplo.encode(), b"N", n, k, alpha, _p(A),
                                         _ld(A), beta, _p(C_), _ld(C_)))
    return C_


def triangular_solve(T, B, side="L", uplo="L", trans="N", diag="N", alpha=1.0):
    """B <- alpha*op(T)^{-1} B (side L) or alpha*B op(T)^{-1} (side R), in place."""
    T = colmajor(T, torch.float64)
    m, n = B.shape
    check("eamd_dtrsm", lib().eamd_dtrsm(handle(), side.encode(), uplo.encode(), trans.encode(),
                                         diag.encode(), m, n, alpha, _p(T), _ld(T), _p(B),
                                         _ld(B)))
    return B


class PartialPivLU:
    """LU with partial pivoting, PA = LU (Eigen/src/LU/PartialPivLU.h)."""

    def __init__(self, matrix=None):
        self._lu = None
        self._t = None
        self._nbt = 0
        self._first_zero = -1
        self._initialized = False
        if matrix is not None:
            self.compute(matrix)

    def compute(self, matrix):
        lu = colmajor(matrix, torch.float64).clone() if not _is_colmajor_f64(matrix) else \
            matrix.clone(memory_format=torch.preserve_format)
        lu = colmajor(lu)
        n = lu.shape[0]
        if lu.shape[0] != lu.shape[1]:
            raise ValueError("PartialPivLU is only for square matrices")
        self._l1_norm = float(lu.abs().sum(dim=0).max()) if n else 0.0
        t = torch.empty(n, dtype=torch.int32, device=lu.device)
        nbt = C.c_int32(0)
        fz = C.c_int64(-1)
        check("eamd_dgetrf", lib().eamd_dgetrf(handle(), n, n, _p(lu), _ld(lu), _p(t),
                                               C.byref(nbt), C.byref(fz)))
        self._lu, self._t, self._nbt, self._first_zero = lu, t, nbt.value, fz.value
        self._initialized = True
        return self

    def _check(self):
        if not self._initialized:
            raise RuntimeError("PartialPivLU is not initialized.")

    def matrixLU(self):
        self._check()
        return self._lu

    def transpositionsP(self):
        self._check()
        return self._t

    def permutationP(self):
        """Indices of the permutation P (PermutationMatrix.h:82-88)."""
        self._check()
        n = self._t.numel()
        p = torch.empty(n, dtype=torch.int32, device=self._t.device)
        check("eamd_transpositions_to_permutation",
              lib().eamd_transpositions_to_permutation(handle(), n, _p(self._t), _p(p)))
        return p

    def first_zero_pivot(self):
        self._check()
        return self._first_zero

    def nb_transpositions(self):
        self._check()
        return self._nbt

    def solve(self, b, trans="N"):
        self._check()
        vec = b.dim() == 1
        x = colmajor(b, torch.float64).clone()
        x = colmajor(x)
        n = self._lu.shape[0]
        check("eamd_dgetrs", lib().eamd_dgetrs(handle(), trans.encode(), n, x.shape[1],
                                               _p(self._lu), _ld(self._lu), _p(self._t), _p(x),
                                               _ld(x)))
        return x.reshape(-1) if vec else x

    def determinant(self):
        self._check()
        sign = -1.0 if self._nbt % 2 else 1.0
        return sign * float(torch.prod(torch.diagonal(self._lu)))

    def inverse(self):
        """A^{-1} from the factors: the solve of the identity (PartialPivLU.h:197-202)."""
        self._check()
        n = self._lu.shape[0]
        return self.solve(torch.eye(n, dtype=torch.float64, device=self._lu.device))

    def rcond(self):
        """Estimate of 1 / (||A||_1 ||A^{-1}||_1) (PartialPivLU.h:184-188)."""
        self._check()
        n = self._lu.shape[0]
        return _rcond_estimate(self._l1_norm, n, self.solve,
                               lambda v: self.solve(v, trans="T"), self._lu.device)

    def reconstructedMatrix(self):
        """P^{-1} L U (PartialPivLU.h:538-566)."""
        self._check()
        lu = self._lu
        L = torch.tril(lu, -1) + torch.eye(lu.shape[0], dtype=lu.dtype, device=lu.device)
        U = torch.triu(lu)
        res = L @ U
        # P A = L U with (P x)[perm[i]] = x[i]  =>  A[i] = (L U)[perm[i]]
        return res[self.permutationP().long()]

    def rows(self):
        return self._lu.shape[0]

    def cols(self):
        return self._lu.shape[1]


def _is_colmajor_f64(t):
    return (t.is_cuda and t.dtype == torch.float64 and t.dim() == 2 and t.stride(0) == 1)


class LLT:
    """Cholesky A = L L^T (Eigen/src/Cholesky/LLT.h)."""

    def __init__(self, matrix=None, uplo="L"):
        self.uplo = uplo
        self._m = None
        self._info = InvalidInput
        self._fail = -1
        if matrix is not None:
            self.compute(matrix)

    def compute(self, matrix):
        m = colmajor(matrix, torch.float64).clone()
        m = colmajor(m)
        n = m.shape[0]
        # ||A||_1 of the selfadjoint matrix read from the stored triangle
        # (LLT.h:434-445)
        tri = torch.tril(m) if self.uplo == "L" else torch.triu(m)
        sym = tri + tri.t() - torch.diag(torch.diagonal(tri))
        self._l1_norm = float(sym.abs().sum(dim=0).max()) if n else 0.0
        fc = C.c_int64(-1)
        check("eamd_dpotrf", lib().eamd_dpotrf(handle(), self.uplo.encode(), n, _p(m), _ld(m),
                                               C.byref(fc)))
        self._m = m
        self._fail = fc.value
        self._info = Success if fc.value < 0 else NumericalIssue
        return self

    def info(self):
        return self._info

    def fail_column(self):
        return self._fail

    def matrixLLT(self):
        return self._m

    def matrixL(self):
        return torch.tril(self._m) if self.uplo == "L" else torch.triu(self._m).t()

    def matrixU(self):
        return self.matrixL().t()

    def solveInPlace(self, b):
        n = self._m.shape[0]
        check("eamd_dpotrs", lib().eamd_dpotrs(handle(), self.uplo.encode(), n, b.shape[1],
                                               _p(self._m), _ld(self._m), _p(b), _ld(b)))
        return b

    def solve(self, b):
        vec = b.dim() == 1
        x = colmajor(b, torch.float64).clone()
        x = colmajor(x)
        self.solveInPlace(x)
        return x.reshape(-1) if vec else x

    def rankUpdate(self, v, sigma=1.0):
        """LLT::rankUpdate (LLT.h:459-472): the factor becomes that of
        A + sigma v v^T, in place; info() is NumericalIssue when a downdate
        (sigma <= 0) leaves A not positive definite.  The stored 1-norm is
        left as it was, as in the reference."""
        from . import batched
        n = self._m.shape[0]
        if n > 256:
            raise ValueError("rankUpdate: sizes up to 256")
        dev = self._m.device
        vv = torch.as_tensor(v, dtype=torch.float64, device=dev).reshape(-1).contiguous()
        if vv.numel() != n:
            raise ValueError("rankUpdate: vector of size %d for a %d x %d factor"
                             % (vv.numel(), n, n))
        z = torch.zeros(1, dtype=torch.int64, device=dev)
        info = batched.llt_rank_update_batched(
            z, torch.tensor([n], dtype=torch.int32, device=dev), self._m, z, vv,
            torch.tensor([float(sigma)], dtype=torch.float64, device=dev), upper=self.uplo == "U")
        self._info = Success if int(info[0]) == -1 else NumericalIssue
        return self

    def rcond(self):
        """Estimate of 1 / (||A||_1 ||A^{-1}||_1) (LLT.h:161-166); the
        factorization must have succeeded."""
        if self._info != Success:
            raise RuntimeError("LLT is not initialized or the matrix is not positive definite.")
        n = self._m.shape[0]
        return _rcond_estimate(self._l1_norm, n, self.solve, self.solve, self._m.device)


class LDLT:
    """Robust Cholesky A = P^T L D L^T P with symmetric diagonal pivoting
    (Eigen/src/Cholesky/LDLT.h, lower storage), sizes up to 64: one system
    of the batched GPU factorization."""

    def __init__(self, matrix=None):
        self._m = None
        self._info = InvalidInput
        if matrix is not None:
            self.compute(matrix)

    def compute(self, matrix):
        """LDLT.h:491-522; the stored 1-norm is that of the selfadjoint
        matrix read from the lower triangle (LDLT.h:500-511)."""
        from . import batched
        m = colmajor(matrix, torch.float64).clone()
        m = colmajor(m)
        n = m.shape[0]
        if n > 64:
            raise ValueError("LDLT: sizes up to 64")
        tri = torch.tril(m)
        sym = tri + tri.t() - torch.diag(torch.diagonal(tri))
        self._l1_norm = float(sym.abs().sum(dim=0).max()) if n else 0.0
        dev = m.device
        z = torch.zeros(1, dtype=torch.int64, device=dev)
        nn = torch.tensor([n], dtype=torch.int32, device=dev)
        T, _, sign, info = batched.ldlt_batched(z, nn, m, z, max_n=n)
        self._m, self._t = m, T[:n]
        self._sign = int(sign[0])
        self._info = Success if int(info[0]) == 0 else NumericalIssue
        return self

    def info(self):
        return self._info

    def matrixLDLT(self):
        return self._m

    def matrixL(self):
        return torch.tril(self._m, -1) + torch.eye(self._m.shape[0], dtype=torch.float64,
                                                   device=self._m.device)

    def matrixU(self):
        return self.matrixL().t()

    def vectorD(self):
        return torch.diagonal(self._m).clone()

    def transpositionsP(self):
        """The 0-based transposition indices (LDLT.h:143-147)."""
        return self._t

    def isPositive(self):
        return self._sign in (0, 2)

    def isNegative(self):
        return self._sign in (1, 2)

    def solve(self, b):
        """x = A^{-1} b, one right-hand side or a column per system
        (LDLT.h:559-593; D^+ zeroes rows whose |d| is not above 1/DBL_MAX)."""
        from . import batched
        n = self._m.shape[0]
        dev = self._m.device
        vec = b.dim() == 1
        x = colmajor(b, torch.float64).clone()
        x = colmajor(x.reshape(n, -1))
        r = x.shape[1]
        z = torch.zeros(r, dtype=torch.int64, device=dev)
        offb = torch.arange(r, dtype=torch.int64, device=dev) * n
        nn = torch.full((r,), n, dtype=torch.int32, device=dev)
        batched.ldlt_solve_batched(z, nn, self._m, z, self._t, offb, x, max_n=n)
        return x.reshape(-1) if vec else x

    def rankUpdate(self, w, sigma=1.0):
        """The factor of A + sigma w w^T in place (LDLT.h:529-555, through
        ldlt_inplace<Lower>::update, :403-449); the transpositions and the
        sign stay as they are.  Only a computed factor is updated (the
        reference also starts from an empty one)."""
        from . import batched
        if self._m is None:
            raise ValueError("LDLT.rankUpdate: compute() first")
        n = self._m.shape[0]
        dev = self._m.device
        wv = colmajor(w, torch.float64).reshape(-1).contiguous()
        if wv.numel() != n:
            raise ValueError("LDLT.rankUpdate: w has %d entries, the factor %d" % (wv.numel(), n))
        z = torch.zeros(1, dtype=torch.int64, device=dev)
        nn = torch.tensor([n], dtype=torch.int32, device=dev)
        sg = torch.tensor([float(sigma)], dtype=torch.float64, device=dev)
        info = batched.ldlt_rank_update_batched(z, nn, self._m, z, self._t, z, wv, sg)
        if int(info[0]) != 0:
            raise RuntimeError("LDLT.rankUpdate: info %d" % int(info[0]))
        return self

    def reconstructedMatrix(self):
        """P^T L D L^T P (LDLT.h:620-643)."""
        L = self.matrixL()
        res = (L * self.vectorD()) @ L.t()
        # (P^T M P)_ij = M_{s(i), s(j)} with P e_i = e_{s(i)}: s(i) follows
        # index i through the transpositions in order
        s = []
        tr = self._t.tolist()
        for i in range(len(tr)):
            pos = i
            for k, j in enumerate(tr):
                if pos == k:
                    pos = j
                elif pos == j:
                    pos = k
            s.append(pos)
        idx = torch.tensor(s, dtype=torch.int64, device=self._m.device)
        return res[idx][:, idx]


def _rcond_estimate(norm1, n, solve, solve_adj, device):
    """1 / (norm1 * est(||A^{-1}||_1)) from solves with A and A^T: Hager's
    one-norm estimator with Higham's refinements (N. J. Higham, ACM TOMS 14,
    1988, Algorithms 4.1 and 5.1), as rcond_estimate_helper and
    rcond_invmatrix_L1_norm_estimate (Eigen/src/Core/ConditionEstimator.h:
    56-172) use it."""
    if n == 0:
        return 1.0
    if norm1 == 0.0:
        return 0.0
    if n == 1:
        return 1.0
    f64 = dict(dtype=torch.float64, device=device)
    v = solve(torch.full((n,), 1.0 / n, **f64))
    lower = float(v.abs().sum())
    old_lower, old_sign, old_idx = lower, None, -1
    for k in range(4):
        sign = torch.where(v < 0, -1.0, 1.0).to(torch.float64)
        if k > 0 and torch.equal(sign, old_sign):
            break
        z = solve_adj(sign)
        idx = int(torch.argmax(z.abs()))  # the first maximum
        if idx == old_idx:
            break
        e = torch.zeros(n, **f64)
        e[idx] = 1.0
        v = solve(e)
        lower = float(v.abs().sum())
        if lower <= old_lower:
            break
        old_sign, old_idx, old_lower = sign, idx, lower
    i = torch.arange(n, **f64)
    alt = torch.where(torch.arange(n, device=device) % 2 == 0, 1.0, -1.0).to(torch.float64)
    alt = alt * (1.0 + i / (n - 1))
    est = max(lower, 2.0 * float(solve(alt).abs().sum()) / (3.0 * n))
    return 0.0 if est == 0.0 else (1.0 / est) / norm1


def backward_error(A, X, B):
    """max_j ||b_j - A x_j||_inf / (||A||_inf ||x_j||_inf + ||b_j||_inf) (FP64, device)."""
    A = colmajor(A, torch.float64)
    X = colmajor(X, torch.float64)
    B = colmajor(B, torch.float64)
    out = C.c_double(0)
    check("eamd_dbackward_error", lib().eamd_dbackward_error(
        handle(), A.shape[0], X.shape[1], _p(A), _ld(A), _p(X), _ld(X), _p(B), _ld(B),
        C.byref(out)))
    return out.value
