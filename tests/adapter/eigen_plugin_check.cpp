// eigen_plugin_check.cpp -- unmodified Eigen user code, compiled with
// include/eigen_amd/eigen_plugin.h so that PartialPivLU<MatrixXd>,
// LLT<MatrixXd>, C.noalias() += A*B and TriangularView::solveInPlace run
// through the eamd C ABI on the GPU.  Writes the inputs and the results as
// raw little-endian arrays into the directory given as argv[1]; the Python
// test compares them with the oracle.  Exit status 0 when every call
// returned, 1 on an exception from the GPU path.
#include <Eigen/Dense>

#define EIGEN_AMD_USE_GEMM
#define EIGEN_AMD_USE_TRSM
#include "../../include/eigen_amd/eigen_plugin.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>

using Eigen::Index;
using Eigen::MatrixXd;
using Eigen::VectorXd;

namespace {

// splitmix64 + Box-Muller: seeded N(0,1) entries, independent of any library
struct Normal {
  uint64_t s;
  explicit Normal(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
  double operator()() {
    const double u1 = uniform(), u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
};

MatrixXd normal(Index r, Index c, uint64_t seed) {
  Normal g(seed);
  MatrixXd m(r, c);
  for (Index j = 0; j < c; ++j)
    for (Index i = 0; i < r; ++i) m(i, j) = g();
  return m;
}

std::string dir;

template <typename T>
void dump(const char* name, const T* p, size_t count) {
  FILE* f = std::fopen((dir + "/" + name).c_str(), "wb");
  if (!f) throw std::runtime_error(std::string("cannot write ") + name);
  std::fwrite(p, sizeof(T), count, f);
  std::fclose(f);
}
void dump(const char* name, const MatrixXd& m) { dump(name, m.data(), (size_t)m.size()); }

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s OUTDIR\n", argv[0]);
    return 2;
  }
  dir = argv[1];
  try {
    // PartialPivLU<MatrixXd>: factor, permutation, solve (C1 shape, smaller n)
    const Index n = 300;
    MatrixXd A = normal(n, n, 11), b = normal(n, 2, 12);
    Eigen::PartialPivLU<MatrixXd> lu(A);
    MatrixXd x = lu.solve(b);
    Eigen::VectorXi perm = lu.permutationP().indices();
    dump("lu_A.f64", A);
    dump("lu_b.f64", b);
    dump("lu_LU.f64", lu.matrixLU());
    dump("lu_perm.i32", perm.data(), (size_t)n);
    dump("lu_x.f64", x);
    double det = lu.determinant();
    dump("lu_det.f64", &det, 1);

    // LLT<MatrixXd>, lower and upper
    MatrixXd G = normal(n, n, 13);
    MatrixXd S = G * G.transpose() / double(n) + MatrixXd::Identity(n, n);
    MatrixXd bs = normal(n, 3, 14);
    Eigen::LLT<MatrixXd> llt(S);
    MatrixXd xs = llt.solve(bs);
    MatrixXd L = llt.matrixL();
    Eigen::LLT<MatrixXd, Eigen::Upper> lltu(S);
    MatrixXd U = lltu.matrixU();
    int info[2] = {llt.info() == Eigen::Success ? 0 : 1, lltu.info() == Eigen::Success ? 0 : 1};
    dump("llt_S.f64", S);
    dump("llt_b.f64", bs);
    dump("llt_L.f64", L);
    dump("llt_U.f64", U);
    dump("llt_x.f64", xs);
    dump("llt_info.i32", info, 2);

    // a non-SPD matrix: the failing column is reported as NumericalIssue
    MatrixXd Sbad = S;
    Sbad(37, 37) = -5.0;
    Eigen::LLT<MatrixXd> bad(Sbad);
    int bad_info = bad.info() == Eigen::NumericalIssue ? 1 : 0;
    dump("llt_bad_info.i32", &bad_info, 1);

    // GEMM: C += A*B and C = A^T * B
    MatrixXd Ga = normal(211, 173, 15), Gb = normal(173, 157, 16), Gc = normal(211, 157, 17);
    MatrixXd Gc0 = Gc;
    Gc.noalias() += Ga * Gb;
    MatrixXd Gt = Ga.transpose() * normal(211, 64, 18);
    dump("gemm_A.f64", Ga);
    dump("gemm_B.f64", Gb);
    dump("gemm_C0.f64", Gc0);
    dump("gemm_C.f64", Gc);
    dump("gemm_T.f64", Gt);

    // TRSM: lower-left, upper-right (unit) solves in place
    MatrixXd T = normal(180, 180, 19);
    T.diagonal().array() += 20.0;
    MatrixXd X1 = normal(180, 40, 20), X2 = normal(55, 180, 21);
    MatrixXd X10 = X1, X20 = X2;
    // a unit triangle with N(0,1) entries is exponentially ill-conditioned:
    // the unit-upper case uses entries scaled by 1/20
    MatrixXd Tu = T / 20.0;
    T.triangularView<Eigen::Lower>().solveInPlace(X1);
    Tu.triangularView<Eigen::UnitUpper>().solveInPlace<Eigen::OnTheRight>(X2);
    dump("trsm_T.f64", T);
    dump("trsm_Tu.f64", Tu);
    dump("trsm_X10.f64", X10);
    dump("trsm_X20.f64", X20);
    dump("trsm_X1.f64", X1);
    dump("trsm_X2.f64", X2);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  std::printf("eigen_plugin_check: ok\n");
  return 0;
}
