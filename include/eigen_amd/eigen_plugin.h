/* eigen_plugin.h -- Eigen's own back-end plug points, specialised for
 * double (and, for the LU and Cholesky factorizations, float: FP32 storage,
 * FP64 arithmetic) to run on the MI355X through the eigen_amd C ABI.
 *
 * Include it after <Eigen/Dense> (or after <Eigen/LU> and <Eigen/Cholesky>)
 * and before the first use of the decompositions, exactly where Eigen's
 * LAPACKE/BLAS bindings are included (Eigen/LU:30-37, Eigen/Cholesky:34-41,
 * Eigen/Core:543-552).  Unmodified user code then factors on the GPU:
 *
 *   PartialPivLU<MatrixXd>(A)  -> partial_lu_impl<double, ColMajor, int>::blocked_lu
 *                                 (PartialPivLU.h:426-496; same plug point and
 *                                 conventions as PartialPivLU_LAPACKE.h:42-77)
 *   PartialPivLU<MatrixXf>(A)  -> partial_lu_impl<float, ColMajor, int>::blocked_lu
 *                                 (as PartialPivLU_LAPACKE.h:75 binds sgetrf)
 *   LLT<MatrixXd, Lower|Upper>  -> llt_inplace<double, Lower|Upper>::blocked
 *                                 (LLT.h:330-360; as LLT_LAPACKE.h:42-93)
 *   LLT<MatrixXf, Lower|Upper>  -> llt_inplace<float, Lower|Upper>::blocked
 *                                 (as LLT_LAPACKE.h:91 binds spotrf)
 *   C.noalias() += A * B        -> general_matrix_matrix_product<..., double, ..., ColMajor>::run
 *                                 (GeneralMatrixMatrix.h:49-199; as
 *                                 GeneralMatrixMatrix_BLAS.h:49-116), when
 *                                 EIGEN_AMD_USE_GEMM is defined
 *   S.selfadjointView<Lower|Upper>().rankUpdate(A), and C.triangularView<..>() += A * A^T
 *                              -> general_matrix_matrix_triangular_product<..., double, ...,
 *                                 ColMajor, UpLo, Specialized>::run
 *                                 (GeneralMatrixMatrixTriangular.h:56-125; as
 *                                 GeneralMatrixMatrixTriangular_BLAS.h:47-98:
 *                                 SYRK when both operands are the same array,
 *                                 Eigen's built-in kernel for other triangular
 *                                 products, as there), when EIGEN_AMD_USE_GEMM
 *                                 is defined
 *   T.solveInPlace(X), matrix X -> triangular_solve_matrix<double, ..., ColMajor>::run
 *                                 (TriangularSolverMatrix.h:38-329; as
 *                                 TriangularSolverMatrix_BLAS.h:41-157), when
 *                                 EIGEN_AMD_USE_TRSM is defined
 *   T.solveInPlace(x), vector x -> triangular_solve_vector<double, double, ..., OnTheLeft, ...>::run
 *                                 (TriangularSolverVector.h:88-139 / 30-85; the
 *                                 OnTheRight form maps onto these, :17-26), so
 *                                 PartialPivLU / LLT solve(VectorXd) runs on the
 *                                 GPU too, when EIGEN_AMD_USE_TRSM is defined
 *
 * Buffers are Eigen's host buffers: each call uploads, computes and
 * downloads through the eamd_host_* entry points on a per-thread handle
 * (handles are thread-confined, like Eigen objects).  The decompositions
 * hold their factors in host memory (m_lu, m_matrix), so each triangular
 * solve uploads the factor it solves with; callers that solve many times
 * with one factor keep it on the device through the C ABI (eamd_dgetrf /
 * eamd_dgetrs) instead.  A failure of the GPU path throws std::runtime_error
 * with the library's message: there is no silent CPU fallback.  Rank updates
 * of an existing LLT keep Eigen's own algorithm (llt_rank_update_lower,
 * LLT.h:231-299). */
#ifndef EIGEN_AMD_EIGEN_PLUGIN_H
#define EIGEN_AMD_EIGEN_PLUGIN_H

#include <stdexcept>
#include <string>
#include <vector>

#include "../eigen_amd.h"

namespace eigen_amd_plugin {

/* One handle per host thread, created on first use and destroyed with the
 * thread. */
inline eamd_handle_t handle() {
  struct Owner {
    eamd_handle_t h = nullptr;
    Owner() {
      if (eamd_create(&h) != 0) h = nullptr;
    }
    ~Owner() {
      if (h) eamd_destroy(h);
    }
  };
  thread_local Owner owner;
  if (!owner.h) throw std::runtime_error(std::string("eigen_amd: ") + eamd_last_error());
  return owner.h;
}

/* How many times each plug point reached the C ABI in this process (for
 * callers and tests that want to see which of Eigen's calls ran on the
 * GPU). */
struct CallCounts {
  long getrf = 0, potrf = 0, gemm = 0, syrk = 0, trsm = 0, trsv = 0;
};
inline CallCounts& call_counts() {
  static CallCounts c;
  return c;
}

inline void check(int status, const char* what) {
  if (status != 0) {
    const char* msg = eamd_last_error();
    throw std::runtime_error(std::string("eigen_amd: ") + what + " returned " +
                             std::to_string(status) + (msg && *msg ? std::string(": ") + msg : ""));
  }
}

/* A row-major block seen as the column-major transpose: copy it out as
 * column-major, run, copy back. */
template <typename T>
inline void to_colmajor(const T* src, Eigen::Index rows, Eigen::Index cols,
                        Eigen::Index stride, std::vector<T>& dst) {
  dst.resize((size_t)rows * (size_t)cols);
  for (Eigen::Index i = 0; i < rows; ++i)
    for (Eigen::Index j = 0; j < cols; ++j) dst[(size_t)j * rows + i] = src[i * stride + j];
}
template <typename T>
inline void from_colmajor(const std::vector<T>& src, Eigen::Index rows, Eigen::Index cols,
                          T* dst, Eigen::Index stride) {
  for (Eigen::Index i = 0; i < rows; ++i)
    for (Eigen::Index j = 0; j < cols; ++j) dst[i * stride + j] = src[(size_t)j * rows + i];
}

/* The host-buffer LU and Cholesky entry points by scalar type (FP32 ones
 * compute in FP64 and round their results once). */
inline int host_getrf(eamd_handle_t h, int64_t m, int64_t n, double* A, int64_t lda,
                      int32_t* t, int32_t* nbt, int64_t* fz) {
  return eamd_host_dgetrf(h, m, n, A, lda, t, nbt, fz);
}
inline int host_getrf(eamd_handle_t h, int64_t m, int64_t n, float* A, int64_t lda,
                      int32_t* t, int32_t* nbt, int64_t* fz) {
  return eamd_host_sgetrf(h, m, n, A, lda, t, nbt, fz);
}
inline int host_potrf(eamd_handle_t h, char uplo, int64_t n, double* A, int64_t lda,
                      int64_t* fail) {
  return eamd_host_dpotrf(h, uplo, n, A, lda, fail);
}
inline int host_potrf(eamd_handle_t h, char uplo, int64_t n, float* A, int64_t lda,
                      int64_t* fail) {
  return eamd_host_spotrf(h, uplo, n, A, lda, fail);
}

}  // namespace eigen_amd_plugin

namespace Eigen {
namespace internal {

/* PartialPivLU<Matrix<double|float,...>>: returns the first exactly-zero
 * pivot or -1; transpositions are 0-based, nb_transpositions counts
 * t[i] != i.  PartialPivLU_LAPACKE.h:74-75 binds both scalar types at this
 * seam. */
template <typename Scalar, int StorageOrder>
struct eigen_amd_lu {
  static Index blocked_lu(Index rows, Index cols, Scalar* lu_data, Index luStride,
                          int* row_transpositions, int& nb_transpositions) {
    int32_t nbt = 0;
    int64_t first_zero = -1;
    /* an empty matrix has no pivots; its stride may be 0, which the C ABI's
     * leading-dimension check would refuse */
    if (rows == 0 || cols == 0) {
      nb_transpositions = 0;
      return -1;
    }
    ++eigen_amd_plugin::call_counts().getrf;
    if (StorageOrder == ColMajor) {
      eigen_amd_plugin::check(
          eigen_amd_plugin::host_getrf(eigen_amd_plugin::handle(), rows, cols, lu_data, luStride,
                                       row_transpositions, &nbt, &first_zero),
          "eamd_host_getrf");
    } else {
      std::vector<Scalar> tmp;
      eigen_amd_plugin::to_colmajor(lu_data, rows, cols, luStride, tmp);
      eigen_amd_plugin::check(
          eigen_amd_plugin::host_getrf(eigen_amd_plugin::handle(), rows, cols, tmp.data(), rows,
                                       row_transpositions, &nbt, &first_zero),
          "eamd_host_getrf");
      eigen_amd_plugin::from_colmajor(tmp, rows, cols, lu_data, luStride);
    }
    nb_transpositions = nbt;
    return (Index)first_zero;
  }
};

template <int StorageOrder>
struct partial_lu_impl<double, StorageOrder, int> {
  static Index blocked_lu(Index rows, Index cols, double* lu_data, Index luStride,
                          int* row_transpositions, int& nb_transpositions,
                          Index maxBlockSize = 256) {
    EIGEN_UNUSED_VARIABLE(maxBlockSize);
    return eigen_amd_lu<double, StorageOrder>::blocked_lu(rows, cols, lu_data, luStride,
                                                          row_transpositions, nb_transpositions);
  }
};

template <int StorageOrder>
struct partial_lu_impl<float, StorageOrder, int> {
  static Index blocked_lu(Index rows, Index cols, float* lu_data, Index luStride,
                          int* row_transpositions, int& nb_transpositions,
                          Index maxBlockSize = 256) {
    EIGEN_UNUSED_VARIABLE(maxBlockSize);
    return eigen_amd_lu<float, StorageOrder>::blocked_lu(rows, cols, lu_data, luStride,
                                                         row_transpositions, nb_transpositions);
  }
};

/* LLT<Matrix<double|float,...>, Lower|Upper>: the failing column or -1
 * (LLT_LAPACKE.h:90-91 binds both scalar types at this seam). */
template <int UpLo>
struct eigen_amd_llt {
  template <typename MatrixType>
  static Index potrf(MatrixType& m) {
    eigen_assert(m.rows() == m.cols());
    const Index n = m.rows();
    int64_t fail = -1;
    if (n == 0) return -1; /* as the unblocked loop over no columns (LLT.h:304-328) */
    const char uplo = UpLo == Lower ? 'L' : 'U';
    ++eigen_amd_plugin::call_counts().potrf;
    if (!(MatrixType::Flags & RowMajorBit)) {
      eigen_amd_plugin::check(eigen_amd_plugin::host_potrf(eigen_amd_plugin::handle(), uplo, n,
                                                           &m.coeffRef(0, 0), m.outerStride(),
                                                           &fail),
                              "eamd_host_potrf");
    } else {
      /* row-major: the stored array is the column-major transpose, whose
       * opposite triangle holds the same factor */
      eigen_amd_plugin::check(eigen_amd_plugin::host_potrf(eigen_amd_plugin::handle(),
                                                           UpLo == Lower ? 'U' : 'L', n,
                                                           &m.coeffRef(0, 0), m.outerStride(),
                                                           &fail),
                              "eamd_host_potrf");
    }
    return (Index)fail;
  }
};

template <>
struct llt_inplace<double, Lower> {
  template <typename MatrixType>
  static Index blocked(MatrixType& m) {
    return eigen_amd_llt<Lower>::potrf(m);
  }
  template <typename MatrixType, typename VectorType>
  static Index rankUpdate(MatrixType& mat, const VectorType& vec,
                          const typename MatrixType::RealScalar& sigma) {
    return Eigen::internal::llt_rank_update_lower(mat, vec, sigma);
  }
};

template <>
struct llt_inplace<double, Upper> {
  template <typename MatrixType>
  static Index blocked(MatrixType& m) {
    return eigen_amd_llt<Upper>::potrf(m);
  }
  template <typename MatrixType, typename VectorType>
  static Index rankUpdate(MatrixType& mat, const VectorType& vec,
                          const typename MatrixType::RealScalar& sigma) {
    Transpose<MatrixType> matt(mat);
    return llt_inplace<double, Lower>::rankUpdate(matt, vec.conjugate(), sigma);
  }
};

template <>
struct llt_inplace<float, Lower> {
  template <typename MatrixType>
  static Index blocked(MatrixType& m) {
    return eigen_amd_llt<Lower>::potrf(m);
  }
  template <typename MatrixType, typename VectorType>
  static Index rankUpdate(MatrixType& mat, const VectorType& vec,
                          const typename MatrixType::RealScalar& sigma) {
    return Eigen::internal::llt_rank_update_lower(mat, vec, sigma);
  }
};

template <>
struct llt_inplace<float, Upper> {
  template <typename MatrixType>
  static Index blocked(MatrixType& m) {
    return eigen_amd_llt<Upper>::potrf(m);
  }
  template <typename MatrixType, typename VectorType>
  static Index rankUpdate(MatrixType& mat, const VectorType& vec,
                          const typename MatrixType::RealScalar& sigma) {
    Transpose<MatrixType> matt(mat);
    return llt_inplace<float, Lower>::rankUpdate(matt, vec.conjugate(), sigma);
  }
};

#ifdef EIGEN_AMD_USE_GEMM
/* res += alpha * op(lhs) * op(rhs), column-major result. */
template <typename Index_, int LhsStorageOrder, bool ConjugateLhs, int RhsStorageOrder,
          bool ConjugateRhs>
struct general_matrix_matrix_product<Index_, double, LhsStorageOrder, ConjugateLhs, double,
                                     RhsStorageOrder, ConjugateRhs, ColMajor> {
  typedef gebp_traits<double, double> Traits;
  static void run(Index_ rows, Index_ cols, Index_ depth, const double* lhs, Index_ lhsStride,
                  const double* rhs, Index_ rhsStride, double* res, Index_ resStride,
                  double alpha, level3_blocking<double, double>& /*blocking*/,
                  GemmParallelInfo<Index_>* /*info*/ = 0) {
    const char ta = LhsStorageOrder == RowMajor ? 'T' : 'N';
    const char tb = RhsStorageOrder == RowMajor ? 'T' : 'N';
    ++eigen_amd_plugin::call_counts().gemm;
    eigen_amd_plugin::check(eamd_host_dgemm(eigen_amd_plugin::handle(), ta, tb, rows, cols,
                                            depth, alpha, lhs, lhsStride, rhs, rhsStride, 1.0,
                                            res, resStride),
                            "eamd_host_dgemm");
  }
};
#endif

#ifdef EIGEN_AMD_USE_GEMM
/* res(UpLo) += alpha * lhs * rhs.  With lhs and rhs the same array in
 * opposite storage orders the product is op(A) op(A)^T, a SYRK; any other
 * triangular product keeps Eigen's built-in kernel, as the reference's BLAS
 * binding does (GeneralMatrixMatrixTriangular_BLAS.h:53-64). */
template <typename Index_, int LhsStorageOrder, bool ConjugateLhs, int RhsStorageOrder,
          bool ConjugateRhs, int UpLo>
struct general_matrix_matrix_triangular_product<Index_, double, LhsStorageOrder, ConjugateLhs,
                                                double, RhsStorageOrder, ConjugateRhs, ColMajor,
                                                UpLo, Specialized> {
  static void run(Index_ size, Index_ depth, const double* lhs, Index_ lhsStride,
                  const double* rhs, Index_ rhsStride, double* res, Index_ resStride,
                  const double& alpha, level3_blocking<double, double>& blocking) {
    if (lhs == rhs && lhsStride == rhsStride && LhsStorageOrder != RhsStorageOrder) {
      if (size == 0) return;
      /* a column-major lhs is A (size x depth): A A^T; a row-major lhs is the
       * column-major array A (depth x size) read transposed: A^T A */
      const char uplo = (UpLo & Lower) == Lower ? 'L' : 'U';
      const char trans = LhsStorageOrder == RowMajor ? 'T' : 'N';
      ++eigen_amd_plugin::call_counts().syrk;
      eigen_amd_plugin::check(eamd_host_dsyrk(eigen_amd_plugin::handle(), uplo, trans, size,
                                              depth, alpha, lhs, lhsStride, 1.0, res,
                                              resStride),
                              "eamd_host_dsyrk");
    } else {
      general_matrix_matrix_triangular_product<Index_, double, LhsStorageOrder, ConjugateLhs,
                                               double, RhsStorageOrder, ConjugateRhs, ColMajor,
                                               UpLo, BuiltIn>::run(size, depth, lhs, lhsStride,
                                                                   rhs, rhsStride, res,
                                                                   resStride, alpha, blocking);
    }
  }
};
#endif

#ifdef EIGEN_AMD_USE_TRSM
/* rhs <- tri^-1 rhs for one right-hand side vector (contiguous). */
template <typename Index_, int Mode, int TriStorageOrder>
struct eigen_amd_trsv {
  static void run(Index_ size, const double* tri, Index_ triStride, double* rhs) {
    if (size == 0) return;
    const bool lower = (Mode & Lower) == Lower;
    const char diag = (Mode & UnitDiag) ? 'U' : 'N';
    /* a row-major triangle is the transpose of the column-major array */
    const char uplo = (TriStorageOrder == RowMajor) ? (lower ? 'U' : 'L') : (lower ? 'L' : 'U');
    const char trans = TriStorageOrder == RowMajor ? 'T' : 'N';
    ++eigen_amd_plugin::call_counts().trsv;
    eigen_amd_plugin::check(eamd_host_dtrsm(eigen_amd_plugin::handle(), 'L', uplo, trans, diag,
                                            size, 1, 1.0, tri, triStride, rhs, size),
                            "eamd_host_dtrsm");
  }
};

template <typename Index_, int Mode, bool Conjugate>
struct triangular_solve_vector<double, double, Index_, OnTheLeft, Mode, Conjugate, ColMajor> {
  static void run(Index_ size, const double* lhs, Index_ lhsStride, double* rhs) {
    eigen_amd_trsv<Index_, Mode, ColMajor>::run(size, lhs, lhsStride, rhs);
  }
};

template <typename Index_, int Mode, bool Conjugate>
struct triangular_solve_vector<double, double, Index_, OnTheLeft, Mode, Conjugate, RowMajor> {
  static void run(Index_ size, const double* lhs, Index_ lhsStride, double* rhs) {
    eigen_amd_trsv<Index_, Mode, RowMajor>::run(size, lhs, lhsStride, rhs);
  }
};

/* other <- tri^-1 other (left) or other tri^-1 (right); other column-major. */
template <typename Index_, int Side, int Mode, bool Conjugate, int TriStorageOrder>
struct eigen_amd_trsm {
  static void run(Index_ size, Index_ otherSize, const double* tri, Index_ triStride,
                  double* other, Index_ otherStride) {
    const bool lower = (Mode & Lower) == Lower;
    const char diag = (Mode & UnitDiag) ? 'U' : 'N';
    /* a row-major triangle is the transpose of the column-major array */
    const char uplo = (TriStorageOrder == RowMajor) ? (lower ? 'U' : 'L') : (lower ? 'L' : 'U');
    const char trans = TriStorageOrder == RowMajor ? 'T' : 'N';
    const char side = Side == OnTheLeft ? 'L' : 'R';
    const Index_ m = Side == OnTheLeft ? size : otherSize;
    const Index_ n = Side == OnTheLeft ? otherSize : size;
    ++eigen_amd_plugin::call_counts().trsm;
    eigen_amd_plugin::check(eamd_host_dtrsm(eigen_amd_plugin::handle(), side, uplo, trans, diag,
                                            m, n, 1.0, tri, triStride, other, otherStride),
                            "eamd_host_dtrsm");
  }
};

template <typename Index_, int Mode, bool Conjugate, int TriStorageOrder>
struct triangular_solve_matrix<double, Index_, OnTheLeft, Mode, Conjugate, TriStorageOrder,
                               ColMajor> {
  static void run(Index_ size, Index_ otherSize, const double* tri, Index_ triStride,
                  double* other, Index_ otherStride, level3_blocking<double, double>&) {
    eigen_amd_trsm<Index_, OnTheLeft, Mode, Conjugate, TriStorageOrder>::run(
        size, otherSize, tri, triStride, other, otherStride);
  }
};

template <typename Index_, int Mode, bool Conjugate, int TriStorageOrder>
struct triangular_solve_matrix<double, Index_, OnTheRight, Mode, Conjugate, TriStorageOrder,
                               ColMajor> {
  static void run(Index_ size, Index_ otherSize, const double* tri, Index_ triStride,
                  double* other, Index_ otherStride, level3_blocking<double, double>&) {
    eigen_amd_trsm<Index_, OnTheRight, Mode, Conjugate, TriStorageOrder>::run(
        size, otherSize, tri, triStride, other, otherStride);
  }
};
#endif

}  // namespace internal
}  // namespace Eigen

#endif  // EIGEN_AMD_EIGEN_PLUGIN_H
