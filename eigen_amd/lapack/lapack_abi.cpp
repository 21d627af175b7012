// lapack_abi.cpp -- dgetrf_/dgetrs_/dpotrf_/dpotrs_ (and their FP32 s*
// forms), dgemm_/dtrsm_/dsyrk_ over the eigen_amd host-buffer entry points.  Argument checks and info codes
// follow lapack/lu.cpp:14-89, lapack/cholesky.cpp:14-72 and
// blas/level3_impl.h:12-179, 357-391;
// each host thread gets its own handle (handles are thread-confined).
#include "../../include/eigen_amd_lapack.h"

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <vector>

#include "../../include/eigen_amd.h"

namespace {

eamd_handle_t handle() {
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
  return owner.h;
}

int illegal(const char* name, int arg, int* info) {
  std::fprintf(stderr, "On entry to %s, parameter number %d had an illegal value\n", name, arg);
  if (info) *info = -arg;
  return 0;
}

int device_failure(const char* name, int status, int* info) {
  const char* msg = eamd_last_error();
  std::fprintf(stderr, "%s: eigen_amd status %d%s%s\n", name, status, msg && *msg ? ": " : "",
               msg ? msg : "");
  if (info) *info = -1000;
  return 0;
}

bool is_trans(char c) {
  c = (char)std::toupper((unsigned char)c);
  return c == 'N' || c == 'T' || c == 'C';
}
bool is_uplo(char c) {
  c = (char)std::toupper((unsigned char)c);
  return c == 'U' || c == 'L';
}
char up(char c) { return (char)std::toupper((unsigned char)c); }

// The LU and Cholesky stubs for both real types (the FP32 entry points keep
// FP32 storage and compute in FP64).
inline int host_getrf(eamd_handle_t h, int m, int n, double* a, int lda, int* ipiv, int32_t* nbt,
                      int64_t* fz) {
  return eamd_host_dgetrf(h, m, n, a, lda, ipiv, nbt, fz);
}
inline int host_getrf(eamd_handle_t h, int m, int n, float* a, int lda, int* ipiv, int32_t* nbt,
                      int64_t* fz) {
  return eamd_host_sgetrf(h, m, n, a, lda, ipiv, nbt, fz);
}
inline int host_getrs(eamd_handle_t h, char tr, int n, int nrhs, const double* a, int lda,
                      const int32_t* t, double* b, int ldb) {
  return eamd_host_dgetrs(h, tr, n, nrhs, a, lda, t, b, ldb);
}
inline int host_getrs(eamd_handle_t h, char tr, int n, int nrhs, const float* a, int lda,
                      const int32_t* t, float* b, int ldb) {
  return eamd_host_sgetrs(h, tr, n, nrhs, a, lda, t, b, ldb);
}
inline int host_potrf(eamd_handle_t h, char uplo, int n, double* a, int lda, int64_t* fail) {
  return eamd_host_dpotrf(h, uplo, n, a, lda, fail);
}
inline int host_potrf(eamd_handle_t h, char uplo, int n, float* a, int lda, int64_t* fail) {
  return eamd_host_spotrf(h, uplo, n, a, lda, fail);
}
inline int host_potrs(eamd_handle_t h, char uplo, int n, int nrhs, const double* a, int lda,
                      double* b, int ldb) {
  return eamd_host_dpotrs(h, uplo, n, nrhs, a, lda, b, ldb);
}
inline int host_potrs(eamd_handle_t h, char uplo, int n, int nrhs, const float* a, int lda,
                      float* b, int ldb) {
  return eamd_host_spotrs(h, uplo, n, nrhs, a, lda, b, ldb);
}

template <typename T>
int getrf_t(const char* nm, int* m, int* n, T* a, int* lda, int* ipiv, int* info) {
  *info = 0;
  if (*m < 0) return illegal(nm, 1, info);
  if (*n < 0) return illegal(nm, 2, info);
  if (*lda < std::max(1, *m)) return illegal(nm, 4, info);
  if (*m == 0 || *n == 0) return 0;
  eamd_handle_t h = handle();
  if (!h) return device_failure(nm, -1000, info);
  int32_t nbt = 0;
  int64_t fz = -1;
  const int st = host_getrf(h, *m, *n, a, *lda, ipiv, &nbt, &fz);
  if (st != 0) return device_failure(nm, st, info);
  for (int i = 0; i < std::min(*m, *n); ++i) ipiv[i] += 1;
  if (fz >= 0) *info = (int)fz + 1;
  return 0;
}

template <typename T>
int getrs_t(const char* nm, char* trans, int* n, int* nrhs, T* a, int* lda, int* ipiv, T* b,
            int* ldb, int* info) {
  *info = 0;
  if (!is_trans(*trans)) return illegal(nm, 1, info);
  if (*n < 0) return illegal(nm, 2, info);
  if (*nrhs < 0) return illegal(nm, 3, info);
  if (*lda < std::max(1, *n)) return illegal(nm, 5, info);
  if (*ldb < std::max(1, *n)) return illegal(nm, 8, info);
  if (*n == 0 || *nrhs == 0) return 0;
  eamd_handle_t h = handle();
  if (!h) return device_failure(nm, -1000, info);
  std::vector<int32_t> t(ipiv, ipiv + *n);
  for (auto& v : t) v -= 1;
  const char tr = up(*trans) == 'N' ? 'N' : 'T';  // real: adjoint = transpose
  const int st = host_getrs(h, tr, *n, *nrhs, a, *lda, t.data(), b, *ldb);
  if (st != 0) return device_failure(nm, st, info);
  return 0;
}

template <typename T>
int potrf_t(const char* nm, char* uplo, int* n, T* a, int* lda, int* info) {
  *info = 0;
  if (!is_uplo(*uplo)) return illegal(nm, 1, info);
  if (*n < 0) return illegal(nm, 2, info);
  if (*lda < std::max(1, *n)) return illegal(nm, 4, info);
  if (*n == 0) return 0;
  eamd_handle_t h = handle();
  if (!h) return device_failure(nm, -1000, info);
  int64_t fail = -1;
  const int st = host_potrf(h, up(*uplo), *n, a, *lda, &fail);
  if (st != 0) return device_failure(nm, st, info);
  if (fail >= 0) *info = (int)fail + 1;
  return 0;
}

template <typename T>
int potrs_t(const char* nm, char* uplo, int* n, int* nrhs, T* a, int* lda, T* b, int* ldb,
            int* info) {
  *info = 0;
  if (!is_uplo(*uplo)) return illegal(nm, 1, info);
  if (*n < 0) return illegal(nm, 2, info);
  if (*nrhs < 0) return illegal(nm, 3, info);
  if (*lda < std::max(1, *n)) return illegal(nm, 5, info);
  if (*ldb < std::max(1, *n)) return illegal(nm, 7, info);
  if (*n == 0 || *nrhs == 0) return 0;
  eamd_handle_t h = handle();
  if (!h) return device_failure(nm, -1000, info);
  const int st = host_potrs(h, up(*uplo), *n, *nrhs, a, *lda, b, *ldb);
  if (st != 0) return device_failure(nm, st, info);
  return 0;
}

}  // namespace

extern "C" {

int dgetrf_(int* m, int* n, double* a, int* lda, int* ipiv, int* info) {
  return getrf_t("DGETRF", m, n, a, lda, ipiv, info);
}
int sgetrf_(int* m, int* n, float* a, int* lda, int* ipiv, int* info) {
  return getrf_t("SGETRF", m, n, a, lda, ipiv, info);
}

int dgetrs_(char* trans, int* n, int* nrhs, double* a, int* lda, int* ipiv, double* b, int* ldb,
            int* info) {
  return getrs_t("DGETRS", trans, n, nrhs, a, lda, ipiv, b, ldb, info);
}
int sgetrs_(char* trans, int* n, int* nrhs, float* a, int* lda, int* ipiv, float* b, int* ldb,
            int* info) {
  return getrs_t("SGETRS", trans, n, nrhs, a, lda, ipiv, b, ldb, info);
}

int dpotrf_(char* uplo, int* n, double* a, int* lda, int* info) {
  return potrf_t("DPOTRF", uplo, n, a, lda, info);
}
int spotrf_(char* uplo, int* n, float* a, int* lda, int* info) {
  return potrf_t("SPOTRF", uplo, n, a, lda, info);
}

int dpotrs_(char* uplo, int* n, int* nrhs, double* a, int* lda, double* b, int* ldb, int* info) {
  return potrs_t("DPOTRS", uplo, n, nrhs, a, lda, b, ldb, info);
}
int spotrs_(char* uplo, int* n, int* nrhs, float* a, int* lda, float* b, int* ldb, int* info) {
  return potrs_t("SPOTRS", uplo, n, nrhs, a, lda, b, ldb, info);
}

int dgemm_(const char* opa, const char* opb, const int* m, const int* n, const int* k,
           const double* alpha, const double* a, const int* lda, const double* b, const int* ldb,
           const double* beta, double* c, const int* ldc) {
  int info = 0;
  const char ta = up(*opa), tb = up(*opb);
  if (!is_trans(ta)) info = 1;
  else if (!is_trans(tb)) info = 2;
  else if (*m < 0) info = 3;
  else if (*n < 0) info = 4;
  else if (*k < 0) info = 5;
  else if (*lda < std::max(1, ta == 'N' ? *m : *k)) info = 8;
  else if (*ldb < std::max(1, tb == 'N' ? *k : *n)) info = 10;
  else if (*ldc < std::max(1, *m)) info = 13;
  if (info) return illegal("DGEMM", info, nullptr);
  if (*m == 0 || *n == 0) return 0;
  eamd_handle_t h = handle();
  if (!h) return device_failure("DGEMM", -1000, nullptr);
  const int st = eamd_host_dgemm(h, ta == 'N' ? 'N' : 'T', tb == 'N' ? 'N' : 'T', *m, *n, *k,
                                 *alpha, a, *lda, b, *ldb, *beta, c, *ldc);
  if (st != 0) return device_failure("DGEMM", st, nullptr);
  return 0;
}

int dtrsm_(const char* side, const char* uplo, const char* opa, const char* diag, const int* m,
           const int* n, const double* alpha, const double* a, const int* lda, double* b,
           const int* ldb) {
  int info = 0;
  const char sd = up(*side), ul = up(*uplo), ta = up(*opa), dg = up(*diag);
  if (sd != 'L' && sd != 'R') info = 1;
  else if (!is_uplo(ul)) info = 2;
  else if (!is_trans(ta)) info = 3;
  else if (dg != 'U' && dg != 'N') info = 4;
  else if (*m < 0) info = 5;
  else if (*n < 0) info = 6;
  else if (*lda < std::max(1, sd == 'L' ? *m : *n)) info = 9;
  else if (*ldb < std::max(1, *m)) info = 11;
  if (info) return illegal("DTRSM", info, nullptr);
  if (*m == 0 || *n == 0) return 0;
  eamd_handle_t h = handle();
  if (!h) return device_failure("DTRSM", -1000, nullptr);
  const int st = eamd_host_dtrsm(h, sd, ul, ta == 'N' ? 'N' : 'T', dg, *m, *n, *alpha, a, *lda,
                                 b, *ldb);
  if (st != 0) return device_failure("DTRSM", st, nullptr);
  return 0;
}

int dsyrk_(const char* uplo, const char* op, const int* n, const int* k, const double* alpha,
           const double* a, const int* lda, const double* beta, double* c, const int* ldc) {
  int info = 0;
  const char ul = up(*uplo), ta = up(*op);
  if (!is_uplo(ul)) info = 1;
  else if (!is_trans(ta)) info = 2;
  else if (*n < 0) info = 3;
  else if (*k < 0) info = 4;
  else if (*lda < std::max(1, ta == 'N' ? *n : *k)) info = 7;
  else if (*ldc < std::max(1, *n)) info = 10;
  if (info) return illegal("DSYRK ", info, nullptr);
  if (*n == 0) return 0;
  eamd_handle_t h = handle();
  if (!h) return device_failure("DSYRK", -1000, nullptr);
  // k == 0 still scales the triangle by beta (blas/level3_impl.h:381-389)
  const int st = eamd_host_dsyrk(h, ul, ta == 'N' ? 'N' : 'T', *n, *k, *alpha, a, *lda, *beta, c,
                                 *ldc);
  if (st != 0) return device_failure("DSYRK", st, nullptr);
  return 0;
}

}  // extern "C"
