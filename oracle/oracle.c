/*
 * oracle.c -- CPU restatement of Eigen's dense factor-and-solve path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the "port" CPU baseline).  The
 * product never links this file.  Each routine cites the Eigen 3.3.90
 * snapshot routine it restates; blocking constants, the panel recursion, the
 * pivot score / tie rule, the zero-pivot rule and the swap order are the
 * reference's.  Inner products are plain loops: the accumulation order of
 * Eigen's GEBP micro-kernel is not reproduced (nor needed: the pivots are
 * compared with a near-tie rule and values with error bounds).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(i, j, ld) ((size_t)(i) + (size_t)(j) * (size_t)(ld))

/* ------------------------------------------------------------------------- */
/* generic pieces, instantiated for double and float                          */
/* ------------------------------------------------------------------------- */

/* X <- L^{-1} X with L unit lower (bs x bs), X bs x nc.
 * A11.triangularView<UnitLower>().solveInPlace(A12), PartialPivLU.h:490 */
#define DEF_TRSM_LLU(T, SFX)                                                   \
  static void trsm_llu_##SFX(long bs, long nc, const T* L, long ldl, T* X,     \
                             long ldx) {                                       \
    for (long j = 0; j < nc; ++j)                                              \
      for (long k = 0; k < bs; ++k) {                                          \
        T b = X[IDX(k, j, ldx)];                                               \
        if (b != (T)0)                                                         \
          for (long i = k + 1; i < bs; ++i)                                    \
            X[IDX(i, j, ldx)] -= b * L[IDX(i, k, ldl)];                        \
      }                                                                        \
  }

/* C -= A*B, A m x k, B k x n (A22.noalias() -= A21*A12, PartialPivLU.h:492) */
#define DEF_GEMM_SUB(T, SFX)                                                   \
  static void gemm_sub_##SFX(long m, long n, long k, const T* A, long lda,     \
                             const T* B, long ldb, T* C, long ldc) {           \
    for (long j = 0; j < n; ++j)                                               \
      for (long p = 0; p < k; ++p) {                                           \
        T b = B[IDX(p, j, ldb)];                                               \
        if (b == (T)0) continue;                                               \
        const T* a = A + IDX(0, p, lda);                                       \
        T* c = C + IDX(0, j, ldc);                                             \
        for (long i = 0; i < m; ++i) c[i] -= a[i] * b;                         \
      }                                                                        \
  }

/* unblocked_lu, PartialPivLU.h:365-409.
 * Pivot: argmax_i |a_ik| over the column tail with a strict '>' scan that
 * starts from the first element (Visitor.h:44-53,179; UnaryFunctors.h:63-66),
 * so the first maximum wins and a NaN wins only from the first position.
 * Zero pivot: no swap, no scaling, remember the first such column. */
#define DEF_LU_UNBLOCKED(T, SFX)                                               \
  static long lu_unblocked_##SFX(long rows, long cols, T* a, long lda,         \
                                 int* transp, int* nbt, double* margin) {      \
    long size = rows < cols ? rows : cols;                                     \
    long first_zero = -1;                                                      \
    *nbt = 0;                                                                  \
    for (long k = 0; k < size; ++k) {                                          \
      long rrows = rows - k - 1, rcols = cols - k - 1;                         \
      long best_i = k;                                                         \
      T best = (T)fabs((double)a[IDX(k, k, lda)]);                             \
      for (long i = k + 1; i < rows; ++i) {                                    \
        T s = (T)fabs((double)a[IDX(i, k, lda)]);                              \
        if (s > best) { best = s; best_i = i; }                                \
      }                                                                        \
      if (margin) {                                                            \
        double second = -1.0;                                                  \
        for (long i = k; i < rows; ++i) {                                      \
          if (i == best_i) continue;                                           \
          double s = fabs((double)a[IDX(i, k, lda)]);                          \
          if (s > second) second = s;                                          \
        }                                                                      \
        margin[k] = (second < 0.0 || best == (T)0)                             \
                        ? 1.0                                                  \
                        : ((double)best - second) / (double)best;              \
      }                                                                        \
      transp[k] = (int)best_i;                                                 \
      if (best != (T)0) {                                                      \
        if (k != best_i) {                                                     \
          for (long j = 0; j < cols; ++j) {                                    \
            T t = a[IDX(k, j, lda)];                                           \
            a[IDX(k, j, lda)] = a[IDX(best_i, j, lda)];                        \
            a[IDX(best_i, j, lda)] = t;                                        \
          }                                                                    \
          ++*nbt;                                                              \
        }                                                                      \
        T piv = a[IDX(k, k, lda)];                                             \
        for (long i = k + 1; i < rows; ++i) a[IDX(i, k, lda)] /= piv;          \
      } else if (first_zero == -1) {                                           \
        first_zero = k;                                                        \
      }                                                                        \
      if (k < rows - 1)                                                        \
        for (long j = 0; j < rcols; ++j) {                                     \
          T u = a[IDX(k, k + 1 + j, lda)];                                     \
          T* c = a + IDX(k + 1, k + 1 + j, lda);                               \
          const T* l = a + IDX(k + 1, k, lda);                                 \
          for (long i = 0; i < rrows; ++i) c[i] -= l[i] * u;                   \
        }                                                                      \
    }                                                                          \
    return first_zero;                                                         \
  }

/* blocked_lu, PartialPivLU.h:426-496: block = clamp(((size/8)/16)*16, 8,
 * maxBlockSize); the (trows+bs) x bs panel is factored by this same routine
 * with maxBlockSize 16; its transpositions are shifted by k, applied to the
 * left block A_0 and, when there are trailing rows, to A_2; then
 * A12 <- L11^{-1} A12 and A22 -= A21*A12. */
#define DEF_LU_BLOCKED(T, SFX)                                                 \
  static long lu_blocked_##SFX(long rows, long cols, T* a, long lda,           \
                               int* transp, int* nbt, long max_block,          \
                               double* margin) {                               \
    long size = rows < cols ? rows : cols;                                     \
    if (size <= 16) return lu_unblocked_##SFX(rows, cols, a, lda, transp, nbt, \
                                              margin);                         \
    long bsz = size / 8;                                                       \
    bsz = (bsz / 16) * 16;                                                     \
    if (bsz < 8) bsz = 8;                                                      \
    if (bsz > max_block) bsz = max_block;                                      \
    *nbt = 0;                                                                  \
    long first_zero = -1;                                                      \
    for (long k = 0; k < size; k += bsz) {                                     \
      long bs = size - k < bsz ? size - k : bsz;                               \
      long trows = rows - k - bs;                                              \
      long tsize = size - k - bs;                                              \
      int nbt_panel = 0;                                                       \
      long ret = lu_blocked_##SFX(trows + bs, bs, a + IDX(k, k, lda), lda,     \
                                  transp + k, &nbt_panel, 16,                  \
                                  margin ? margin + k : NULL);                 \
      if (ret >= 0 && first_zero == -1) first_zero = k + ret;                  \
      *nbt += nbt_panel;                                                       \
      for (long i = k; i < k + bs; ++i) {                                      \
        long piv = (transp[i] += (int)k);                                      \
        if (piv != i)                                                          \
          for (long j = 0; j < k; ++j) {                                       \
            T t = a[IDX(i, j, lda)];                                           \
            a[IDX(i, j, lda)] = a[IDX(piv, j, lda)];                           \
            a[IDX(piv, j, lda)] = t;                                           \
          }                                                                    \
      }                                                                        \
      if (trows) {                                                             \
        for (long i = k; i < k + bs; ++i) {                                    \
          long piv = transp[i];                                                \
          if (piv != i)                                                        \
            for (long j = k + bs; j < k + bs + tsize; ++j) {                   \
              T t = a[IDX(i, j, lda)];                                         \
              a[IDX(i, j, lda)] = a[IDX(piv, j, lda)];                         \
              a[IDX(piv, j, lda)] = t;                                         \
            }                                                                  \
        }                                                                      \
        trsm_llu_##SFX(bs, tsize, a + IDX(k, k, lda), lda,                     \
                       a + IDX(k, k + bs, lda), lda);                          \
        gemm_sub_##SFX(trows, tsize, bs, a + IDX(k + bs, k, lda), lda,         \
                       a + IDX(k, k + bs, lda), lda,                           \
                       a + IDX(k + bs, k + bs, lda), lda);                     \
      }                                                                        \
    }                                                                          \
    return first_zero;                                                         \
  }

/* _solve_impl, PartialPivLU.h:226-244.  P*b applies the transpositions in
 * order (the permutation built by PermutationMatrix.h:82-88 maps row i to
 * perm[i]; the same map), then unit-lower and upper substitution. */
#define DEF_LU_SOLVE(T, SFX)                                                   \
  void orc_lu_solve_##SFX(long n, long nrhs, const T* lu, long lda,            \
                          const int* transp, T* b, long ldb) {                 \
    for (long c = 0; c < nrhs; ++c) {                                          \
      T* x = b + IDX(0, c, ldb);                                               \
      for (long k = 0; k < n; ++k) {                                           \
        long t = transp[k];                                                    \
        if (t != k) { T s = x[k]; x[k] = x[t]; x[t] = s; }                     \
      }                                                                        \
      for (long k = 0; k < n; ++k) {                                           \
        T v = x[k];                                                            \
        if (v != (T)0)                                                         \
          for (long i = k + 1; i < n; ++i) x[i] -= v * lu[IDX(i, k, lda)];     \
      }                                                                        \
      for (long k = n - 1; k >= 0; --k) {                                      \
        x[k] /= lu[IDX(k, k, lda)];                                            \
        T v = x[k];                                                            \
        if (v != (T)0)                                                         \
          for (long i = 0; i < k; ++i) x[i] -= v * lu[IDX(i, k, lda)];         \
      }                                                                        \
    }                                                                          \
  }

/* llt_inplace<.,Lower>::unblocked, LLT.h:304-328:
 * x = a_kk - ||A10||^2; x <= 0 -> return k (NaN passes the test, as in the
 * reference); L_kk = sqrt(x); A21 -= A20*A10^T; A21 /= L_kk. */
#define DEF_LLT_UNBLOCKED(T, SFX)                                              \
  long orc_llt_unblocked_##SFX(long n, T* a, long lda) {                       \
    for (long k = 0; k < n; ++k) {                                             \
      long rs = n - k - 1;                                                     \
      T x = a[IDX(k, k, lda)];                                                 \
      if (k > 0) {                                                             \
        T s = 0;                                                               \
        for (long j = 0; j < k; ++j) s += a[IDX(k, j, lda)] * a[IDX(k, j, lda)];\
        x -= s;                                                                \
      }                                                                        \
      if (x <= (T)0) return k;                                                 \
      x = (T)sqrt((double)x);                                                  \
      a[IDX(k, k, lda)] = x;                                                   \
      if (k > 0 && rs > 0)                                                     \
        for (long j = 0; j < k; ++j) {                                         \
          T l = a[IDX(k, j, lda)];                                             \
          for (long i = k + 1; i < n; ++i) a[IDX(i, k, lda)] -= a[IDX(i, j, lda)] * l;\
        }                                                                      \
      if (rs > 0)                                                              \
        for (long i = k + 1; i < n; ++i) a[IDX(i, k, lda)] /= x;               \
    }                                                                          \
    return -1;                                                                 \
  }

/* llt_inplace<.,Lower>::blocked, LLT.h:330-360: n < 32 -> unblocked;
 * block = clamp(((n/8)/16)*16, 8, 128); per block: unblocked(A11) (failure
 * returns k+ret at once), A21 <- A21 L11^{-T}, A22 -= A21 A21^T (lower). */
#define DEF_LLT_BLOCKED(T, SFX)                                                \
  long orc_llt_blocked_##SFX(long n, T* a, long lda) {                         \
    if (n < 32) return orc_llt_unblocked_##SFX(n, a, lda);                     \
    long bsz = n / 8;                                                          \
    bsz = (bsz / 16) * 16;                                                     \
    if (bsz < 8) bsz = 8;                                                      \
    if (bsz > 128) bsz = 128;                                                  \
    for (long k = 0; k < n; k += bsz) {                                        \
      long bs = bsz < n - k ? bsz : n - k;                                     \
      long rs = n - k - bs;                                                    \
      T* A11 = a + IDX(k, k, lda);                                             \
      T* A21 = a + IDX(k + bs, k, lda);                                        \
      T* A22 = a + IDX(k + bs, k + bs, lda);                                   \
      long ret = orc_llt_unblocked_##SFX(bs, A11, lda);                        \
      if (ret >= 0) return k + ret;                                            \
      if (rs > 0) {                                                            \
        /* X L11^T = A21  ->  column sweep over L11^T (upper) */               \
        for (long j = 0; j < bs; ++j) {                                        \
          for (long p = 0; p < j; ++p) {                                       \
            T l = A11[IDX(j, p, lda)];                                         \
            for (long i = 0; i < rs; ++i) A21[IDX(i, j, lda)] -= A21[IDX(i, p, lda)] * l;\
          }                                                                    \
          T d = A11[IDX(j, j, lda)];                                           \
          for (long i = 0; i < rs; ++i) A21[IDX(i, j, lda)] /= d;              \
        }                                                                      \
        for (long j = 0; j < rs; ++j)                                          \
          for (long p = 0; p < bs; ++p) {                                      \
            T l = A21[IDX(j, p, lda)];                                         \
            for (long i = j; i < rs; ++i) A22[IDX(i, j, lda)] -= A21[IDX(i, p, lda)] * l;\
          }                                                                    \
      }                                                                        \
    }                                                                          \
    return -1;                                                                 \
  }

/* LLT::_solve_impl, LLT.h:475-505: L y = b then L^T x = y */
#define DEF_LLT_SOLVE(T, SFX)                                                  \
  void orc_llt_solve_##SFX(long n, long nrhs, const T* l, long ldl, T* b,      \
                           long ldb) {                                         \
    for (long c = 0; c < nrhs; ++c) {                                          \
      T* x = b + IDX(0, c, ldb);                                               \
      for (long k = 0; k < n; ++k) {                                           \
        x[k] /= l[IDX(k, k, ldl)];                                             \
        T v = x[k];                                                            \
        for (long i = k + 1; i < n; ++i) x[i] -= v * l[IDX(i, k, ldl)];        \
      }                                                                        \
      for (long k = n - 1; k >= 0; --k) {                                      \
        T s = x[k];                                                            \
        for (long i = k + 1; i < n; ++i) s -= l[IDX(i, k, ldl)] * x[i];        \
        x[k] = s / l[IDX(k, k, ldl)];                                          \
      }                                                                        \
    }                                                                          \
  }

DEF_TRSM_LLU(double, d)
DEF_TRSM_LLU(float, f)
DEF_GEMM_SUB(double, d)
DEF_GEMM_SUB(float, f)
DEF_LU_UNBLOCKED(double, d)
DEF_LU_UNBLOCKED(float, f)
DEF_LU_BLOCKED(double, d)
DEF_LU_BLOCKED(float, f)
DEF_LU_SOLVE(double, d)
DEF_LU_SOLVE(float, f)
DEF_LLT_UNBLOCKED(double, d)
DEF_LLT_UNBLOCKED(float, f)
DEF_LLT_BLOCKED(double, d)
DEF_LLT_BLOCKED(float, f)
DEF_LLT_SOLVE(double, d)
DEF_LLT_SOLVE(float, f)

long orc_lu_unblocked_d(long rows, long cols, double* a, long lda, int* transp,
                        int* nb_transp, double* margin) {
  return lu_unblocked_d(rows, cols, a, lda, transp, nb_transp, margin);
}

long orc_lu_blocked_d(long rows, long cols, double* a, long lda, int* transp,
                      int* nb_transp, long max_block, double* margin) {
  return lu_blocked_d(rows, cols, a, lda, transp, nb_transp, max_block, margin);
}

long orc_lu_blocked_f(long rows, long cols, float* a, long lda, int* transp,
                      int* nb_transp, long max_block) {
  return lu_blocked_f(rows, cols, a, lda, transp, nb_transp, max_block, NULL);
}

/* PermutationMatrix::operator=(Transpositions), PermutationMatrix.h:82-88:
 * identity, then for k = n-1 .. 0 swap(indices[k], indices[t_k]). */
void orc_transpositions_to_perm(long n, const int* transp, int* perm) {
  for (long i = 0; i < n; ++i) perm[i] = (int)i;
  for (long k = n - 1; k >= 0; --k) {
    int t = perm[k];
    perm[k] = perm[transp[k]];
    perm[transp[k]] = t;
  }
}

/* ------------------------------------------------------------------------- */
/* level 3                                                                   */
/* ------------------------------------------------------------------------- */
static inline double opget(const double* M, long ld, char t, long i, long j) {
  return (t == 'N' || t == 'n') ? M[IDX(i, j, ld)] : M[IDX(j, i, ld)];
}

void orc_gemm_d(char transa, char transb, long m, long n, long k, double alpha,
                const double* A, long lda, const double* B, long ldb,
                double beta, double* C, long ldc) {
  for (long j = 0; j < n; ++j)
    for (long i = 0; i < m; ++i) {
      long double s = 0;
      for (long p = 0; p < k; ++p)
        s += (long double)opget(A, lda, transa, i, p) * opget(B, ldb, transb, p, j);
      double c = beta == 0.0 ? 0.0 : beta * C[IDX(i, j, ldc)];
      C[IDX(i, j, ldc)] = (double)(c + (long double)alpha * s);
    }
}

/* B <- alpha * op(A)^{-1} B (side L) or alpha * B op(A)^{-1} (side R).
 * Substitution order follows the column sweeps of triangular_solve_matrix
 * (TriangularSolverMatrix.h:112-146 left, 262-300 right). */
void orc_trsm_d(char side, char uplo, char transa, char diag, long m, long n,
                double alpha, const double* A, long lda, double* B, long ldb) {
  int left = (side == 'L' || side == 'l');
  int lower = (uplo == 'L' || uplo == 'l');
  int trans = !(transa == 'N' || transa == 'n');
  int unit = (diag == 'U' || diag == 'u');
  /* effective triangle of op(A): lower xor trans */
  int eff_lower = lower ^ trans;
  long na = left ? m : n;
#define TA(i, j) (trans ? A[IDX(j, i, lda)] : A[IDX(i, j, lda)])
  for (long j = 0; j < n; ++j)
    for (long i = 0; i < m; ++i) B[IDX(i, j, ldb)] *= alpha;
  if (left) {
    for (long j = 0; j < n; ++j) {
      double* x = B + IDX(0, j, ldb);
      if (eff_lower) {
        for (long k = 0; k < na; ++k) {
          if (!unit) x[k] /= TA(k, k);
          for (long i = k + 1; i < na; ++i) x[i] -= x[k] * TA(i, k);
        }
      } else {
        for (long k = na - 1; k >= 0; --k) {
          if (!unit) x[k] /= TA(k, k);
          for (long i = 0; i < k; ++i) x[i] -= x[k] * TA(i, k);
        }
      }
    }
  } else {
    /* X op(A) = B: column j of X depends on columns before (upper) or after
     * (lower) */
    if (!eff_lower) {
      for (long j = 0; j < na; ++j) {
        for (long p = 0; p < j; ++p) {
          double a = TA(p, j);
          for (long i = 0; i < m; ++i) B[IDX(i, j, ldb)] -= B[IDX(i, p, ldb)] * a;
        }
        if (!unit)
          for (long i = 0; i < m; ++i) B[IDX(i, j, ldb)] /= TA(j, j);
      }
    } else {
      for (long j = na - 1; j >= 0; --j) {
        for (long p = j + 1; p < na; ++p) {
          double a = TA(p, j);
          for (long i = 0; i < m; ++i) B[IDX(i, j, ldb)] -= B[IDX(i, p, ldb)] * a;
        }
        if (!unit)
          for (long i = 0; i < m; ++i) B[IDX(i, j, ldb)] /= TA(j, j);
      }
    }
  }
#undef TA
}

void orc_syrk_d(char uplo, char trans, long n, long k, double alpha,
                const double* A, long lda, double beta, double* C, long ldc) {
  int lower = (uplo == 'L' || uplo == 'l');
  int t = !(trans == 'N' || trans == 'n');
  for (long j = 0; j < n; ++j) {
    long i0 = lower ? j : 0, i1 = lower ? n : j + 1;
    for (long i = i0; i < i1; ++i) {
      long double s = 0;
      for (long p = 0; p < k; ++p) {
        double ai = t ? A[IDX(p, i, lda)] : A[IDX(i, p, lda)];
        double aj = t ? A[IDX(p, j, lda)] : A[IDX(j, p, lda)];
        s += (long double)ai * aj;
      }
      double c = beta == 0.0 ? 0.0 : beta * C[IDX(i, j, ldc)];
      C[IDX(i, j, ldc)] = (double)(c + (long double)alpha * s);
    }
  }
}

double orc_backward_error_d(long n, long nrhs, const double* A, long lda,
                            const double* X, long ldx, const double* B,
                            long ldb) {
  double anorm = 0.0;
  for (long i = 0; i < n; ++i) {
    long double s = 0;
    for (long j = 0; j < n; ++j) s += fabsl((long double)A[IDX(i, j, lda)]);
    if ((double)s > anorm) anorm = (double)s;
  }
  double worst = 0.0;
  for (long c = 0; c < nrhs; ++c) {
    const double* x = X + IDX(0, c, ldx);
    const double* b = B + IDX(0, c, ldb);
    double xn = 0.0, bn = 0.0, rn = 0.0;
    for (long i = 0; i < n; ++i) {
      if (fabs(x[i]) > xn) xn = fabs(x[i]);
      if (fabs(b[i]) > bn) bn = fabs(b[i]);
    }
    for (long i = 0; i < n; ++i) {
      long double s = -(long double)b[i];
      for (long j = 0; j < n; ++j) s += (long double)A[IDX(i, j, lda)] * x[j];
      double r = (double)fabsl(s);
      if (r > rn) rn = r;
    }
    double den = anorm * xn + bn;
    double e = den > 0.0 ? rn / den : rn;
    if (e > worst || e != e) worst = e;
  }
  return worst;
}

/* ------------------------------------------------------------------------- */
/* LLT rank-one update                                                       */
/* ------------------------------------------------------------------------- */
/* L L^T + sigma v v^T = L' L'^T in place on the lower factor (column-major,
 * ld); returns -1, or the column where a downdate (sigma <= 0) stops.
 * internal::llt_rank_update_lower (LLT.h:231-299): sigma > 0 by Givens
 * rotations (JacobiRotation::makeGivens for reals, Jacobi.h:226-262, applied
 * as x' = c x + s y, y' = -s x + c y, Jacobi.h:315-330); otherwise the
 * scalar downdate recurrence. */
long orc_llt_rank_update_d(long n, double* L, long ld, const double* v, double sigma) {
  double* t = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  long ret = -1;
  if (sigma > 0) {
    const double ss = sqrt(sigma);
    for (long i = 0; i < n; ++i) t[i] = ss * v[i];
    for (long i = 0; i < n; ++i) {
      const double p = L[IDX(i, i, ld)], q = -t[i];
      double c, s, r;
      if (q == 0.0) {
        c = p < 0.0 ? -1.0 : 1.0;
        s = 0.0;
        r = fabs(p);
      } else if (p == 0.0) {
        c = 0.0;
        s = q < 0.0 ? 1.0 : -1.0;
        r = fabs(q);
      } else if (fabs(p) > fabs(q)) {
        const double tt = q / p;
        double u = sqrt(1.0 + tt * tt);
        if (p < 0.0) u = -u;
        c = 1.0 / u;
        s = -tt * c;
        r = p * u;
      } else {
        const double tt = p / q;
        double u = sqrt(1.0 + tt * tt);
        if (q < 0.0) u = -u;
        s = -1.0 / u;
        c = -tt * s;
        r = q * u;
      }
      L[IDX(i, i, ld)] = r;
      /* the identity rotation is not applied (Jacobi.h:460-461) */
      if (c == 1.0 && s == 0.0) continue;
      for (long k = i + 1; k < n; ++k) {
        const double x = L[IDX(k, i, ld)], y = t[k];
        L[IDX(k, i, ld)] = c * x + s * y;
        t[k] = -s * x + c * y;
      }
    }
  } else {
    for (long i = 0; i < n; ++i) t[i] = v[i];
    double beta = 1.0;
    for (long j = 0; j < n; ++j) {
      const double Ljj = L[IDX(j, j, ld)], dj = Ljj * Ljj, wj = t[j];
      const double swj2 = sigma * wj * wj, gamma = dj * beta + swj2;
      const double x = dj + swj2 / beta;
      if (x <= 0.0) {
        ret = j;
        break;
      }
      const double nLjj = sqrt(x);
      L[IDX(j, j, ld)] = nLjj;
      beta += swj2 / dj;
      for (long k = j + 1; k < n; ++k) t[k] -= (wj / Ljj) * L[IDX(k, j, ld)];
      if (gamma != 0.0)
        for (long k = j + 1; k < n; ++k)
          L[IDX(k, j, ld)] = (nLjj / Ljj) * L[IDX(k, j, ld)] + (nLjj * sigma * wj / gamma) * t[k];
    }
  }
  free(t);
  return ret;
}

/* ------------------------------------------------------------------------- */
/* batched                                                                   */
/* ------------------------------------------------------------------------- */
void orc_batched_llt_solve_f(long count, const int64_t* offA, const int32_t* n,
                             const float* A, const int64_t* offb, float* b,
                             int32_t* info) {
  float* work = NULL;
  long cap = 0;
  for (long s = 0; s < count; ++s) {
    long ns = n[s];
    if (ns * ns > cap) {
      cap = ns * ns;
      free(work);
      work = (float*)malloc(sizeof(float) * (size_t)(cap ? cap : 1));
    }
    memcpy(work, A + offA[s], sizeof(float) * (size_t)(ns * ns));
    long r = orc_llt_blocked_f(ns, work, ns);
    info[s] = (int32_t)r;
    if (r < 0) orc_llt_solve_f(ns, 1, work, ns, b + offb[s], ns);
  }
  free(work);
}

void orc_batched_lu_solve_d(long count, const int64_t* offA, const int32_t* n,
                            const double* A, const int64_t* offb, long nrhs,
                            double* B, int32_t* info) {
  double* work = NULL;
  int* tp = NULL;
  long cap = 0, capn = 0;
  for (long s = 0; s < count; ++s) {
    long ns = n[s];
    if (ns * ns > cap) {
      cap = ns * ns;
      free(work);
      work = (double*)malloc(sizeof(double) * (size_t)(cap ? cap : 1));
    }
    if (ns > capn) {
      capn = ns;
      free(tp);
      tp = (int*)malloc(sizeof(int) * (size_t)capn);
    }
    memcpy(work, A + offA[s], sizeof(double) * (size_t)(ns * ns));
    int nbt = 0;
    long r = orc_lu_blocked_d(ns, ns, work, ns, tp, &nbt, 256, NULL);
    info[s] = (int32_t)r;
    orc_lu_solve_d(ns, nrhs, work, ns, tp, B + offb[s], ns);
  }
  free(work);
  free(tp);
}

/* ------------------------------------------------------------------------- */
/* LDLT (symmetric diagonal pivoting), lower storage                          */
/* ------------------------------------------------------------------------- */
/* In place on the lower triangle of A (column-major, ld): strictly lower L
 * (unit diagonal implied), D on the diagonal; trans[k] = the row exchanged
 * with k at step k; *sign = 0 PositiveSemiDef, 1 NegativeSemiDef,
 * 2 ZeroSign, 3 Indefinite (internal::SignMatrix, LDLT.h:22).  Returns 1 for
 * Success, 0 for NumericalIssue.  Restates internal::ldlt_inplace<Lower>::
 * unblocked (LDLT.h:290-395) with compute()'s sign = ZeroSign start
 * (LDLT.h:516): the pivot is the first largest |diagonal| of the trailing
 * block (maxCoeff's strict '>' scan, Visitor.h:173-185), the lower-only
 * symmetric swap, then the left-looking column
 * temp = D A10^T, a_kk -= A10 temp, A21 -= A20 temp, A21 /= a_kk. */
long orc_ldlt_d(long n, double* A, long ld, int* trans, int* sign) {
  int sg = 2;
  long ret = 1, found_zero = 0;
  if (n <= 1) {
    if (n == 1) {
      trans[0] = 0;
      sg = A[0] > 0.0 ? 0 : (A[0] < 0.0 ? 1 : 2);
    }
    *sign = sg;
    return 1;
  }
  double* temp = (double*)malloc(sizeof(double) * (size_t)n);
  for (long k = 0; k < n; ++k) {
    long p = k;
    double best = fabs(A[IDX(k, k, ld)]);
    for (long i = k + 1; i < n; ++i)
      if (fabs(A[IDX(i, i, ld)]) > best) {
        best = fabs(A[IDX(i, i, ld)]);
        p = i;
      }
    trans[k] = (int)p;
    if (p != k) {
      for (long j = 0; j < k; ++j) {
        double t = A[IDX(k, j, ld)];
        A[IDX(k, j, ld)] = A[IDX(p, j, ld)];
        A[IDX(p, j, ld)] = t;
      }
      for (long i = p + 1; i < n; ++i) {
        double t = A[IDX(i, k, ld)];
        A[IDX(i, k, ld)] = A[IDX(i, p, ld)];
        A[IDX(i, p, ld)] = t;
      }
      double t = A[IDX(k, k, ld)];
      A[IDX(k, k, ld)] = A[IDX(p, p, ld)];
      A[IDX(p, p, ld)] = t;
      for (long i = k + 1; i < p; ++i) {
        double u = A[IDX(i, k, ld)];
        A[IDX(i, k, ld)] = A[IDX(p, i, ld)];
        A[IDX(p, i, ld)] = u;
      }
    }
    if (k > 0) {
      double s = 0.0;
      for (long j = 0; j < k; ++j) {
        temp[j] = A[IDX(j, j, ld)] * A[IDX(k, j, ld)];
        s += A[IDX(k, j, ld)] * temp[j];
      }
      A[IDX(k, k, ld)] -= s;
      for (long i = k + 1; i < n; ++i) {
        double r = 0.0;
        for (long j = 0; j < k; ++j) r += A[IDX(i, j, ld)] * temp[j];
        A[IDX(i, k, ld)] -= r;
      }
    }
    const double akk = A[IDX(k, k, ld)];
    const int valid = fabs(akk) > 0.0;
    if (k == 0 && !valid) {
      for (long j = 0; j < n; ++j) {
        trans[j] = (int)j;
        for (long i = j + 1; i < n; ++i)
          if (!(A[IDX(i, j, ld)] == 0.0)) ret = 0;
      }
      free(temp);
      *sign = 2;
      return ret;
    }
    if (k + 1 < n && valid) {
      for (long i = k + 1; i < n; ++i) A[IDX(i, k, ld)] /= akk;
    } else if (k + 1 < n) {
      for (long i = k + 1; i < n; ++i)
        if (!(A[IDX(i, k, ld)] == 0.0)) ret = 0;
    }
    if (found_zero && valid) ret = 0;
    else if (!valid) found_zero = 1;
    if (sg == 0) {
      if (akk < 0.0) sg = 3;
    } else if (sg == 1) {
      if (akk > 0.0) sg = 3;
    } else if (sg == 2) {
      if (akk > 0.0) sg = 0;
      else if (akk < 0.0) sg = 1;
    }
  }
  free(temp);
  *sign = sg;
  return ret;
}

/* x <- A^{-1} b from the LDLT factors, nrhs columns (LDLT.h:559-593):
 * x = P b (the transpositions in order), L^{-1}, D^+ (rows whose |d| is not
 * above 1 / DBL_MAX are set to zero), L^{-T}, then P^T (in reverse order). */
/* LDLT rank-one update in place on ldlt's factor (strictly lower L, D on the
 * diagonal, column-major, ld; trans as orc_ldlt_d returns it): w (n values,
 * read only) is permuted by the transpositions in order k = 0, 1, ...
 * (Transpositions * vector, ProductEvaluators.h:1051-1056), then column by
 * column d_j += sigma w_j^2 / alpha, alpha += sigma w_j^2 / d_j,
 * w_i -= w_j l_ij and, if gamma = d_j alpha + sigma w_j^2 != 0,
 * l_ij += (sigma w_j / gamma) w_i, stopping once alpha is not finite.
 * Restates internal::ldlt_inplace<Lower>::update / updateInPlace
 * (LDLT.h:403-449), as LDLT::rankUpdate (LDLT.h:529-555) calls it on a
 * computed factor. */
void orc_ldlt_rank_update_d(long n, double* A, long ld, const int* trans, const double* w,
                            double sigma) {
  double* t = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (long i = 0; i < n; ++i) t[i] = w[i];
  for (long k = 0; k < n; ++k) {
    const long j = trans[k];
    if (j != k) {
      const double x = t[k];
      t[k] = t[j];
      t[j] = x;
    }
  }
  double alpha = 1.0;
  for (long j = 0; j < n; ++j) {
    if (!isfinite(alpha)) break;
    const double dj = A[j + j * ld], wj = t[j];
    const double swj2 = sigma * wj * wj;
    const double gamma = dj * alpha + swj2;
    A[j + j * ld] += swj2 / alpha;
    alpha += swj2 / dj;
    for (long i = j + 1; i < n; ++i) t[i] -= wj * A[i + j * ld];
    if (gamma != 0.0) {
      const double c = sigma * wj / gamma;
      for (long i = j + 1; i < n; ++i) A[i + j * ld] += c * t[i];
    }
  }
  free(t);
}

void orc_ldlt_solve_d(long n, long nrhs, const double* A, long ld, const int* trans, double* B,
                      long ldb) {
  const double tol = 1.0 / 1.7976931348623157e308;
  for (long c = 0; c < nrhs; ++c) {
    double* x = B + (size_t)c * (size_t)ldb;
    for (long k = 0; k < n; ++k) {
      double t = x[k];
      x[k] = x[trans[k]];
      x[trans[k]] = t;
    }
    for (long k = 0; k < n; ++k)
      for (long i = k + 1; i < n; ++i) x[i] -= A[IDX(i, k, ld)] * x[k];
    for (long i = 0; i < n; ++i) {
      const double d = A[IDX(i, i, ld)];
      x[i] = fabs(d) > tol ? x[i] / d : 0.0;
    }
    for (long k = n - 1; k >= 0; --k) {
      double s = x[k];
      for (long i = k + 1; i < n; ++i) s -= A[IDX(i, k, ld)] * x[i];
      x[k] = s;
    }
    for (long k = n - 1; k >= 0; --k) {
      double t = x[k];
      x[k] = x[trans[k]];
      x[trans[k]] = t;
    }
  }
}
