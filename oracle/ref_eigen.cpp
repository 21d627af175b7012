// ref_eigen.cpp -- a C-ABI shim over the reference (Eigen 3.3.90) itself.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile with g++ against the
// reference headers where they lie (-I <reference root>), into oracle/_ref/.
// It is used to pin the C restatement (oracle.c) and as the "reference" CPU
// baseline.  No reference source is copied: this file only calls Eigen's
// public classes and its internal entry points named in the comments.
//
// Build flags follow the survey's protocol: -O3 -DNDEBUG -mavx2 -mfma
// -fopenmp (the snapshot's AVX-512 path crashes in PartialPivLU).
#include <Eigen/Dense>
#include <cstdint>
#include <cstring>
#ifdef _OPENMP
#include <omp.h>
#endif

using namespace Eigen;
typedef Map<MatrixXd, 0, OuterStride<> > MapD;
typedef Map<MatrixXf, 0, OuterStride<> > MapF;

extern "C" {

int eref_version(void) {
  return EIGEN_WORLD_VERSION * 10000 + EIGEN_MAJOR_VERSION * 100 + EIGEN_MINOR_VERSION;
}

void eref_set_threads(int n) { Eigen::setNbThreads(n); }
int eref_get_threads(void) { return Eigen::nbThreads(); }

// internal::partial_lu_impl<double,ColMajor,int>::blocked_lu (PartialPivLU.h:426)
// -- the plug point; returns first_zero_pivot or -1.
long eref_lu_blocked_d(long rows, long cols, double* a, long lda, int* transp,
                       int* nb_transp) {
  int nbt = 0;
  long r = (long)internal::partial_lu_impl<double, ColMajor, int>::blocked_lu(
      rows, cols, a, lda, transp, nbt);
  *nb_transp = nbt;
  return r;
}

long eref_lu_blocked_f(long rows, long cols, float* a, long lda, int* transp,
                       int* nb_transp) {
  int nbt = 0;
  long r = (long)internal::partial_lu_impl<float, ColMajor, int>::blocked_lu(
      rows, cols, a, lda, transp, nbt);
  *nb_transp = nbt;
  return r;
}

// PartialPivLU<MatrixXd>: compute, then solve (PartialPivLU.h:128,173,226)
// lu_out (n x n, ld n) receives matrixLU(), perm_out permutationP().indices(),
// transp_out the transpositions, X (ld ldx) the solution of A X = B.
int eref_partialpivlu_d(long n, const double* A, long lda, long nrhs,
                        const double* B, long ldb, double* lu_out,
                        int* perm_out, double* X, long ldx, double* det) {
  MatrixXd Am = MapD(const_cast<double*>(A), n, n, OuterStride<>(lda));
  PartialPivLU<MatrixXd> lu(Am);
  if (lu_out) Map<MatrixXd>(lu_out, n, n) = lu.matrixLU();
  if (perm_out)
    for (long i = 0; i < n; ++i) perm_out[i] = lu.permutationP().indices()(i);
  if (X && nrhs > 0) {
    MatrixXd Bm = MapD(const_cast<double*>(B), n, nrhs, OuterStride<>(ldb));
    MapD(X, n, nrhs, OuterStride<>(ldx)) = lu.solve(Bm);
  }
  if (det) *det = lu.determinant();
  return 0;
}

// LLT<MatrixXd,Lower>: compute then solve (LLT.h:422,475); returns info
// (0 Success, 1 NumericalIssue); L_out receives matrixLLT()
int eref_llt_d(long n, const double* A, long lda, long nrhs, const double* B,
               long ldb, double* L_out, double* X, long ldx) {
  MatrixXd Am = MapD(const_cast<double*>(A), n, n, OuterStride<>(lda));
  LLT<MatrixXd> llt(Am);
  if (L_out) Map<MatrixXd>(L_out, n, n) = llt.matrixLLT();
  if (X && nrhs > 0 && llt.info() == Success) {
    MatrixXd Bm = MapD(const_cast<double*>(B), n, nrhs, OuterStride<>(ldb));
    MapD(X, n, nrhs, OuterStride<>(ldx)) = llt.solve(Bm);
  }
  return llt.info() == Success ? 0 : 1;
}

int eref_llt_f(long n, const float* A, long lda, long nrhs, const float* B,
               long ldb, float* L_out, float* X, long ldx) {
  MatrixXf Am = MapF(const_cast<float*>(A), n, n, OuterStride<>(lda));
  LLT<MatrixXf> llt(Am);
  if (L_out) Map<MatrixXf>(L_out, n, n) = llt.matrixLLT();
  if (X && nrhs > 0 && llt.info() == Success) {
    MatrixXf Bm = MapF(const_cast<float*>(B), n, nrhs, OuterStride<>(ldb));
    MapF(X, n, nrhs, OuterStride<>(ldx)) = llt.solve(Bm);
  }
  return llt.info() == Success ? 0 : 1;
}

// PartialPivLU<MatrixXd>::rcond() (PartialPivLU.h:184-188,
// ConditionEstimator.h:157-172) and inverse() (PartialPivLU.h:197-202)
double eref_lu_rcond_d(long n, const double* A, long lda) {
  MatrixXd Am = MapD(const_cast<double*>(A), n, n, OuterStride<>(lda));
  PartialPivLU<MatrixXd> lu(Am);
  return lu.rcond();
}
void eref_lu_inverse_d(long n, const double* A, long lda, double* Ainv) {
  MatrixXd Am = MapD(const_cast<double*>(A), n, n, OuterStride<>(lda));
  PartialPivLU<MatrixXd> lu(Am);
  Map<MatrixXd>(Ainv, n, n) = lu.inverse();
}
// LLT<MatrixXd,Lower>::rcond() (LLT.h:161-166): -1 when not SPD
double eref_llt_rcond_d(long n, const double* A, long lda) {
  MatrixXd Am = MapD(const_cast<double*>(A), n, n, OuterStride<>(lda));
  LLT<MatrixXd> llt(Am);
  return llt.info() == Success ? llt.rcond() : -1.0;
}

// internal::llt_inplace<double,Lower>::rankUpdate (LLT.h:231-299, 380-384):
// the lower factor (column-major, ld) updated in place; -1 or failing column
long eref_llt_rank_update_d(long n, double* L, long ld, const double* v, double sigma) {
  MapD Lm(L, n, n, OuterStride<>(ld));
  Map<const VectorXd> vv(v, n);
  return (long)internal::llt_inplace<double, Lower>::rankUpdate(Lm, vv, sigma);
}

// LDLT<MatrixXd, Lower>::compute (LDLT.h:491-522): the lower triangle of A
// (column-major, lda) replaced by matrixLDLT()'s, the transpositions' indices,
// *sign: 0 PositiveSemiDef, 1 NegativeSemiDef, 2 ZeroSign, 3 Indefinite
// (read back through isPositive()/isNegative(), LDLT.h:170-181); returns 1
// for Success, 0 for NumericalIssue
long eref_ldlt_d(long n, double* a, long lda, int* trans, int* sign) {
  MapD A(a, n, n, OuterStride<>(lda));
  LDLT<MatrixXd, Lower> f(n);
  f.compute(MatrixXd(A));
  const MatrixXd& m = f.matrixLDLT();
  for (long j = 0; j < n; ++j)
    for (long i = j; i < n; ++i) A(i, j) = m(i, j);
  for (long k = 0; k < n; ++k) trans[k] = (int)f.transpositionsP().indices()(k);
  const bool pos = f.isPositive(), neg = f.isNegative();
  *sign = pos && neg ? 2 : (pos ? 0 : (neg ? 1 : 3));
  return f.info() == Success ? 1 : 0;
}

// LDLT<MatrixXd, Lower>(A).rankUpdate(w, sigma) (LDLT.h:529-555): the lower
// triangle of A (column-major, lda) replaced by the updated matrixLDLT()'s,
// the transpositions' indices in trans
void eref_ldlt_rank_update_d(long n, double* a, long lda, const double* w, double sigma,
                             int* trans) {
  MapD A(a, n, n, OuterStride<>(lda));
  LDLT<MatrixXd, Lower> f(n);
  f.compute(MatrixXd(A));
  f.rankUpdate(Map<const VectorXd>(w, n), sigma);
  const MatrixXd& m = f.matrixLDLT();
  for (long j = 0; j < n; ++j)
    for (long i = j; i < n; ++i) A(i, j) = m(i, j);
  for (long k = 0; k < n; ++k) trans[k] = (int)f.transpositionsP().indices()(k);
}

// LDLT<MatrixXd, Lower>::compute(A).solve(B) (LDLT.h:559-593), B in place
void eref_ldlt_solve_d(long n, long nrhs, const double* a, long lda, double* b, long ldb) {
  MapD A(const_cast<double*>(a), n, n, OuterStride<>(lda));
  MapD Bm(b, n, nrhs, OuterStride<>(ldb));
  const MatrixXd Am = A;
  LDLT<MatrixXd, Lower> f(Am);
  const MatrixXd Bc = Bm;
  MatrixXd x = f.solve(Bc);
  Bm = x;
}

// internal::llt_inplace<double,Lower>::blocked (LLT.h:330): failing col or -1
long eref_llt_blocked_d(long n, double* a, long lda) {
  MapD A(a, n, n, OuterStride<>(lda));
  return (long)internal::llt_inplace<double, Lower>::blocked(A);
}

// C.noalias() = A * B  (GeneralMatrixMatrix.h:405-492)
void eref_gemm_d(long m, long n, long k, const double* A, long lda,
                 const double* B, long ldb, double* C, long ldc) {
  MapD Am(const_cast<double*>(A), m, k, OuterStride<>(lda));
  MapD Bm(const_cast<double*>(B), k, n, OuterStride<>(ldb));
  MapD Cm(C, m, n, OuterStride<>(ldc));
  Cm.noalias() = Am * Bm;
}

// TriangularView<..>::solveInPlace<Side> (SolveTriangular.h:165-183)
// side 'L': B <- op(T)^{-1} B ; side 'R': B <- B op(T)^{-1}
void eref_trsm_d(char side, char uplo, char trans, char diag, long m, long n,
                 const double* A, long lda, double* B, long ldb) {
  long na = (side == 'L') ? m : n;
  MapD T(const_cast<double*>(A), na, na, OuterStride<>(lda));
  MapD X(B, m, n, OuterStride<>(ldb));
  bool lower = uplo == 'L', unit = diag == 'U', tr = trans != 'N';
  if (side == 'L') {
    if (!tr) {
      if (lower) { if (unit) T.triangularView<UnitLower>().solveInPlace(X); else T.triangularView<Lower>().solveInPlace(X); }
      else       { if (unit) T.triangularView<UnitUpper>().solveInPlace(X); else T.triangularView<Upper>().solveInPlace(X); }
    } else {
      if (lower) { if (unit) T.transpose().triangularView<UnitUpper>().solveInPlace(X); else T.transpose().triangularView<Upper>().solveInPlace(X); }
      else       { if (unit) T.transpose().triangularView<UnitLower>().solveInPlace(X); else T.transpose().triangularView<Lower>().solveInPlace(X); }
    }
  } else {
    if (!tr) {
      if (lower) { if (unit) T.triangularView<UnitLower>().solveInPlace<OnTheRight>(X); else T.triangularView<Lower>().solveInPlace<OnTheRight>(X); }
      else       { if (unit) T.triangularView<UnitUpper>().solveInPlace<OnTheRight>(X); else T.triangularView<Upper>().solveInPlace<OnTheRight>(X); }
    } else {
      if (lower) { if (unit) T.transpose().triangularView<UnitUpper>().solveInPlace<OnTheRight>(X); else T.transpose().triangularView<Upper>().solveInPlace<OnTheRight>(X); }
      else       { if (unit) T.transpose().triangularView<UnitLower>().solveInPlace<OnTheRight>(X); else T.transpose().triangularView<Lower>().solveInPlace<OnTheRight>(X); }
    }
  }
}

// A22.selfadjointView<Lower>().rankUpdate(A21, alpha)  (SelfadjointProduct.h:121)
void eref_syrk_lower_d(long n, long k, double alpha, const double* A, long lda,
                       double* C, long ldc) {
  MapD Am(const_cast<double*>(A), n, k, OuterStride<>(lda));
  MapD Cm(C, n, n, OuterStride<>(ldc));
  Cm.selfadjointView<Lower>().rankUpdate(Am, alpha);
}

// Batched: one LLT<MatrixXf> per system, OpenMP parallel-for over systems
// (Eigen's own parallelism switches itself off when nested,
// Parallelizer.h:123).
void eref_batched_llt_solve_f(long count, const int64_t* offA, const int32_t* n,
                              const float* A, const int64_t* offb, float* b,
                              int32_t* info) {
#pragma omp parallel for schedule(dynamic, 256)
  for (long s = 0; s < count; ++s) {
    long ns = n[s];
    Map<const MatrixXf> As(A + offA[s], ns, ns);
    LLT<MatrixXf> llt(As);
    info[s] = llt.info() == Success ? -1 : 1;
    Map<VectorXf> bs(b + offb[s], ns);
    if (llt.info() == Success) bs = llt.solve(bs);
  }
}

// Batched: one PartialPivLU<MatrixXd> per system, nrhs columns (ld n_i)
void eref_batched_lu_solve_d(long count, const int64_t* offA, const int32_t* n,
                             const double* A, const int64_t* offb, long nrhs,
                             double* B, int32_t* info) {
#pragma omp parallel for schedule(dynamic, 4)
  for (long s = 0; s < count; ++s) {
    long ns = n[s];
    Map<const MatrixXd> As(A + offA[s], ns, ns);
    PartialPivLU<MatrixXd> lu(As);
    Map<MatrixXd> Bs(B + offb[s], ns, nrhs);
    Bs = lu.solve(Bs);
    info[s] = -1;
  }
}

int eref_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

}  // extern "C"
