// eigen_plugin_check.cpp -- unmodified Eigen user code, compiled with
// include/eigen_amd/eigen_plugin.h so that PartialPivLU<MatrixXd>,
// LLT<MatrixXd>, C.noalias() += A*B, selfadjointView().rankUpdate(A) and
// TriangularView::solveInPlace (matrix and vector right-hand sides) run
// through the eamd C ABI on the GPU.  Writes the inputs, the results and the
// plug points' call counts per section as raw little-endian arrays into the
// directory given as argv[1]; the Python test compares them with the oracle.
// Exit status 0 when every call returned, 1 on an exception from the GPU
// path.
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
void dump(const char* name, const Eigen::MatrixXf& m) { dump(name, m.data(), (size_t)m.size()); }

// the plug points' call counts: getrf, potrf, gemm, syrk, trsm, trsv
void dump_counts(const char* name) {
  const auto& c = eigen_amd_plugin::call_counts();
  const long v[6] = {c.getrf, c.potrf, c.gemm, c.syrk, c.trsm, c.trsv};
  dump(name, v, 6);
}

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
    dump_counts("cnt_lu.i64");

    // one right-hand side as a vector (C1's shape): triangular_solve_vector
    VectorXd bv = normal(n, 1, 22).col(0);
    VectorXd xv = lu.solve(bv);
    dump("luv_b.f64", bv.data(), (size_t)n);
    dump("luv_x.f64", xv.data(), (size_t)n);
    dump_counts("cnt_luv.i64");

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
    VectorXd bsv = bs.col(0);
    VectorXd xsv = llt.solve(bsv), xsu = lltu.solve(bsv);
    dump("lltv_x.f64", xsv.data(), (size_t)n);
    dump("lltv_xu.f64", xsu.data(), (size_t)n);
    dump_counts("cnt_llt.i64");

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
    dump_counts("cnt_trsm.i64");

    // SYRK through selfadjointView().rankUpdate: C += alpha A A^T (lower),
    // C += alpha B^T B (upper, B read transposed)
    MatrixXd Sa = normal(150, 70, 23), Sb = normal(60, 150, 24), Sc0 = normal(150, 150, 25);
    MatrixXd Sl = Sc0, Su = Sc0;
    Sl.selfadjointView<Eigen::Lower>().rankUpdate(Sa, 0.5);
    Su.selfadjointView<Eigen::Upper>().rankUpdate(Sb.transpose(), -2.0);
    dump("syrk_A.f64", Sa);
    dump("syrk_B.f64", Sb);
    dump("syrk_C0.f64", Sc0);
    dump("syrk_L.f64", Sl);
    dump("syrk_U.f64", Su);
    dump_counts("cnt_syrk.i64");

    // FP32: PartialPivLU<MatrixXf> and LLT<MatrixXf> (lower, upper) through
    // the float plug points; their solves stay in Eigen's FP32 substitution
    {
      const Index nf = 257;
      Eigen::MatrixXf Af = normal(nf, nf, 31).cast<float>(), bf = normal(nf, 2, 32).cast<float>();
      Eigen::PartialPivLU<Eigen::MatrixXf> luf(Af);
      Eigen::MatrixXf xf = luf.solve(bf);
      Eigen::VectorXi permf = luf.permutationP().indices();
      dump("luf_A.f32", Af);
      dump("luf_b.f32", bf);
      dump("luf_LU.f32", luf.matrixLU());
      dump("luf_perm.i32", permf.data(), (size_t)nf);
      dump("luf_x.f32", xf);
      Eigen::MatrixXf Gf = normal(nf, nf, 33).cast<float>();
      Eigen::MatrixXf Sf = Gf * Gf.transpose() / float(nf) + Eigen::MatrixXf::Identity(nf, nf);
      Eigen::MatrixXf bsf = normal(nf, 3, 34).cast<float>();
      Eigen::LLT<Eigen::MatrixXf> lltf(Sf);
      Eigen::MatrixXf xsf = lltf.solve(bsf);
      Eigen::MatrixXf Lf = lltf.matrixL();
      Eigen::LLT<Eigen::MatrixXf, Eigen::Upper> lltfu(Sf);
      Eigen::MatrixXf Uf = lltfu.matrixU();
      Eigen::MatrixXf Sfbad = Sf;
      Sfbad(20, 20) = -1.0f;
      Eigen::LLT<Eigen::MatrixXf> badf(Sfbad);
      int infof[3] = {lltf.info() == Eigen::Success ? 0 : 1,
                      lltfu.info() == Eigen::Success ? 0 : 1,
                      badf.info() == Eigen::NumericalIssue ? 1 : 0};
      dump("lltf_S.f32", Sf);
      dump("lltf_b.f32", bsf);
      dump("lltf_L.f32", Lf);
      dump("lltf_U.f32", Uf);
      dump("lltf_x.f32", xsf);
      dump("lltf_info.i32", infof, 3);
      dump_counts("cnt_f32.i64");
    }

    // empty matrices: Success, no GPU call needed
    MatrixXd E(0, 0);
    Eigen::PartialPivLU<MatrixXd> elu(E);
    Eigen::LLT<MatrixXd> ellt(E);
    int empty_ok = ellt.info() == Eigen::Success && elu.matrixLU().size() == 0 ? 1 : 0;
    dump("empty_ok.i32", &empty_ok, 1);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  std::printf("eigen_plugin_check: ok\n");
  return 0;
}
