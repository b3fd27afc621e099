// Tests of the C++ host mirror (include/rest_tensors.hpp), written after the
// reference's own examples: the doc examples of MatrixFull / _dgemm
// (src/matrix/mod.rs:437-452, src/matrix/matrix_blas_lapack.rs:76-119), the
// ao2mo benchmark shape (benches/bench_tensors.rs:5-16), a Davidson solve
// judged by its residual, and the _dsyev doc example (matrix_blas_lapack.rs:286-318).  `--cpu-only` runs the checks that touch no GPU.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "rest_tensors.hpp"

using rest::MatrixFull;
using rest::MatrixUpper;
using rest::Range;
using rest::RIFull;

static int failures = 0;
#define CHECK(name, cond)                                   \
  do {                                                      \
    if (cond) {                                             \
      std::printf("ok %s\n", name);                         \
    } else {                                                \
      std::printf("FAIL %s (%s:%d)\n", name, __FILE__, __LINE__); \
      ++failures;                                           \
    }                                                       \
  } while (0)

static std::vector<double> seq(int a, int b) {  // a..b-1 as doubles
  std::vector<double> v;
  for (int i = a; i < b; ++i) v.push_back(i);
  return v;
}

static void cpu_checks() {
  // RIFull layout (ri.rs:18-32): element (i,j,k) at i + n0 j + n0 n1 k
  RIFull<double> r({3, 4, 5}, 0.0);
  CHECK("rifull_indicing", r.indicing[0] == 1 && r.indicing[1] == 3 && r.indicing[2] == 12);
  r(2, 1, 3) = 7.0;
  CHECK("rifull_index", r.data[2 + 3 * 1 + 12 * 3] == 7.0);
  CHECK("rifull_reducing_matrix", r.get_reducing_matrix(3) == r.data.data() + 36 &&
                                      r.get_reducing_matrix(5) == nullptr);
  bool panicked = false;
  try {
    (void)RIFull<double>::from_vec({2, 2, 2}, std::vector<double>(7, 0.0));
  } catch (const rest::Panic&) {
    panicked = true;
  }
  CHECK("rifull_from_vec_short_panics", panicked);
  CHECK("rifull_from_vec_exact", RIFull<double>::from_vec({2, 2, 2}, seq(0, 8)).has_value());
  // MatrixFull (matrixfull.rs:196-240) column-major
  auto m = *MatrixFull<double>::from_vec({3, 3}, seq(1, 10));
  CHECK("matrixfull_index", m(1, 2) == 8.0 && m(2, 0) == 3.0);
  // MatrixUpper::size() of a packed length (matrixupper.rs:289-292)
  auto up = MatrixUpper<double>::from_vec(10, std::vector<double>(10, 0.0));
  CHECK("matrixupper_dims", up.dims()[0] == 4);
  CHECK("matrixupper_not_triangular",
        MatrixUpper<double>::from_vec(11, std::vector<double>(11, 0.0)).dims()[0] == 0);
}

static void gpu_checks() {
  // iter_matrixupper doc order (matrix/mod.rs:437-452): 4x4 of 1..16
  auto m4 = *MatrixFull<double>::from_vec({4, 4}, seq(1, 17));
  auto up = m4.to_matrixupper();
  const std::vector<double> order = {1, 5, 6, 9, 10, 11, 13, 14, 15, 16};
  CHECK("to_matrixupper_doc_order", up.size == 10 && up.data == order);
  auto full = up.to_matrixfull();
  bool sym = full.has_value();
  for (int j = 0; sym && j < 4; ++j)
    for (int i = 0; i <= j; ++i)
      sym = sym && (*full)(i, j) == m4(i, j) && (*full)(j, i) == m4(i, j);
  CHECK("to_matrixfull_mirrors", sym);
  CHECK("to_matrixfull_not_triangular",
        !MatrixUpper<double>::from_vec(11, std::vector<double>(11, 1.0)).to_matrixfull());
  CHECK("to_matrixfull_empty", MatrixUpper<double>::empty().to_matrixfull()->data.empty());

  // _dgemm doc examples (matrix_blas_lapack.rs:76-119)
  auto A = *MatrixFull<double>::from_vec({3, 3}, seq(1, 10));
  auto B = *MatrixFull<double>::from_vec({3, 3}, seq(6, 15));
  MatrixFull<double> C({3, 3}, 2.0);
  rest::_dgemm(A, {Range{1, 3}, Range{1, 3}}, 'N', B, {Range{0, 2}, Range{0, 2}}, 'N', C,
               {Range{1, 3}, Range{0, 2}}, 1.0, 1.0);
  CHECK("dgemm_doc_nn", C(1, 0) == 88.0 && C(2, 0) == 101.0 && C(1, 1) == 127.0 &&
                            C(2, 1) == 146.0 && C(0, 0) == 2.0 && C(2, 2) == 2.0);
  MatrixFull<double> C2({3, 3}, 2.0);
  rest::_dgemm(A, {Range{1, 3}, Range{1, 2}}, 'T', B, {Range{0, 2}, Range{0, 2}}, 'N', C2,
               {Range{1, 2}, Range{0, 2}}, 1.0, 1.0);
  CHECK("dgemm_doc_tn", C2(1, 0) == 74.0 && C2(1, 1) == 107.0 && C2(0, 0) == 2.0);
  MatrixFull<double> C3({3, 3}, 2.0);
  rest::_dgemm(A, {Range{1, 3}, Range{0, 2}}, 'N', B, {Range{0, 1}, Range{0, 2}}, 'T', C3,
               {Range{0, 2}, Range{0, 1}}, 1.0, 1.0);
  CHECK("dgemm_doc_nt", C3(0, 0) == 59.0 && C3(1, 0) == 74.0 && C3(2, 0) == 2.0);
  bool panicked = false;
  try {
    rest::_dgemm(A, {Range{0, 3}, Range{0, 3}}, 'N', B, {Range{0, 2}, Range{0, 3}}, 'N', C,
                 {Range{0, 3}, Range{0, 3}}, 1.0, 0.0);
  } catch (const rest::Panic&) {
    panicked = true;
  }
  CHECK("dgemm_shape_panics", panicked);

  // ao2mo on the benchmark's shape (benches/bench_tensors.rs:5-10): ri3fn = 2.0
  // [10,10,20], C = 1.0 [10,10] -> every ri3mo element = 2 * 10 * 10 = 200
  RIFull<double> ri3fn({10, 10, 20}, 2.0);
  MatrixFull<double> eig({10, 10}, 1.0);
  auto ri3mo = ri3fn.ao2mo(eig);
  bool all200 = ri3mo.size[0] == 20 && ri3mo.size[1] == 10 && ri3mo.size[2] == 10;
  for (double v : ri3mo.data) all200 = all200 && v == 200.0;
  CHECK("ao2mo_bench_closed_form", all200);
  // ao2mo with C = I is a layout permutation: ri3mo(P,i,j) = ri3fn(i,j,P)
  RIFull<double> ri({6, 6, 5}, 0.0);
  for (size_t k = 0; k < ri.data.size(); ++k) ri.data[k] = (double)(k % 97) - 40.0;
  for (int P = 0; P < 5; ++P)  // symmetric in (i,j), as ri3fn is
    for (int j = 0; j < 6; ++j)
      for (int i = 0; i < j; ++i) ri(j, i, P) = ri(i, j, P);
  MatrixFull<double> eye({6, 6}, 0.0);
  for (int i = 0; i < 6; ++i) eye(i, i) = 1.0;
  auto perm = ri.ao2mo(eye);
  bool ok = true;
  for (int P = 0; P < 5; ++P)
    for (int j = 0; j < 6; ++j)
      for (int i = 0; i < 6; ++i) ok = ok && perm(P, i, j) == ri(i, j, P);
  CHECK("ao2mo_identity_permutation", ok);
  // rifull_to_matfull_symm (ri.rs:297-306): column P = packed upper of slice P
  auto packed = ri.rifull_to_matfull_symm();
  ok = packed.size[0] == 21 && packed.size[1] == 5;
  for (int P = 0; P < 5 && ok; ++P)
    for (int j = 0; j < 6; ++j)
      for (int i = 0; i <= j; ++i) ok = ok && packed(j * (j + 1) / 2 + i, P) == ri(i, j, P);
  CHECK("rifull_to_matfull_symm", ok);

  // davidson_solve (davidson.rs:48-53): A = diag(1..n) + 0.01 u u^T, closure a_x
  const size_t n = 120;
  std::vector<double> u(n);
  for (size_t i = 0; i < n; ++i) u[i] = std::sin(0.37 * (double)i + 0.1);
  auto a_x = [&](const std::vector<double>& x) {
    double ux = 0.0;
    for (size_t i = 0; i < n; ++i) ux += u[i] * x[i];
    std::vector<double> y(n);
    for (size_t i = 0; i < n; ++i) y[i] = (double)(i + 1) * x[i] + 0.01 * u[i] * ux;
    return y;
  };
  std::vector<double> hdiag(n), x0(n, 0.0);
  for (size_t i = 0; i < n; ++i) hdiag[i] = (double)(i + 1) + 0.01 * u[i] * u[i];
  x0[0] = 1.0;
  rest::DavidsonParams params;
  params.tol = 1e-12;
  params.nroots = 3;
  params.maxcyc = 200;
  auto [conv, e, xs] = rest::davidson_solve(a_x, x0, hdiag, params, 0);
  double worst = 0.0;
  for (size_t k = 0; k < 3; ++k) {
    auto ax = a_x(xs[k]);
    double r2 = 0.0;
    for (size_t i = 0; i < n; ++i) r2 += (ax[i] - e[k] * xs[k][i]) * (ax[i] - e[k] * xs[k][i]);
    worst = std::max(worst, std::sqrt(r2));
  }
  CHECK("davidson_converged", conv[0] && conv[1] && conv[2]);
  CHECK("davidson_residual", worst < 1e-6);
  // a positive rank-one update interlaces: e_k lies in [d_k, d_{k+1}] = [k+1, k+2]
  CHECK("davidson_interlacing", e[0] >= 1.0 && e[0] <= 2.0 && e[1] >= 2.0 && e[1] <= 3.0 &&
                                    e[2] >= 3.0 && e[2] <= 4.0);

  // transposes on the reference's test tensor [3,2,2] of 0..11 (ri.rs:437-449)
  {
    auto ri = *RIFull<double>::from_vec({3, 2, 2}, seq(0, 12));
    auto a = ri.transpose_jik(), b = ri.transpose_jki(), c = ri.transpose_kji(),
         d = ri.transpose_ikj();
    bool ok = a.size == std::array<size_t, 3>{2, 3, 2} && b.size == std::array<size_t, 3>{2, 2, 3} &&
              c.size == std::array<size_t, 3>{2, 2, 3} && d.size == std::array<size_t, 3>{3, 2, 2};
    for (size_t x = 0; x < 3; ++x)
      for (size_t y = 0; y < 2; ++y)
        for (size_t z = 0; z < 2; ++z) {
          const double v = ri(x, y, z);
          ok = ok && a(y, x, z) == v && b(y, z, x) == v && c(z, y, x) == v && d(x, z, y) == v;
        }
    CHECK("rifull_transposes", ok);
  }

  // _dspgvx with B = diag(2, 2, 2): eigenvalues of A / 2, Z^T B Z = I
  {
    auto A = MatrixUpper<double>::from_vec(6, {4.0, 1.0, 6.0, 0.0, 1.0, 8.0});
    auto Bm = MatrixUpper<double>::from_vec(6, {2.0, 0.0, 2.0, 0.0, 0.0, 2.0});
    auto res = rest::_dspgvx(A, Bm, 2);
    auto [vec_a, w_a, n_a] = rest::lapack_dspevx(A);
    bool ok = res.has_value() && n_a == 3 && std::fabs(res->second[0] - 0.5 * w_a[0]) < 1e-13 &&
              std::fabs(res->second[1] - 0.5 * w_a[1]) < 1e-13 && res->second[2] == 0.0;
    double g = 0.0;  // (Z^T B Z)(0,0) with B = 2 I
    for (int r = 0; r < 3; ++r) g += 2.0 * res->first(r, 0) * res->first(r, 0);
    CHECK("dspgvx_scaled_overlap", ok && std::fabs(g - 1.0) < 1e-13);
  }

  // _dsyev doc example (matrix_blas_lapack.rs:286-318), eigenvectors sign-sensitive
  {
    auto up = MatrixUpper<double>::from_vec(6, seq(1, 7));
    auto a = *up.to_matrixfull();
    auto [vec, w, nd] = rest::_dsyev(a, 'V');
    const double wref[3] = {-1.5066326307865059, -0.05739624271478554, 11.564028873501286};
    const double vref[9] = {-0.6827362941552275, -0.38559063640162244, 0.6206375864887483,
                            0.6202872696512856,  -0.7547821848190948,  0.21341873532627673,
                            -0.3861539275363389, -0.5306823104260265,  -0.7544941548144386};
    double dv = 0.0, dw = 0.0;
    for (int k = 0; k < 9; ++k) dv += (vec->data[k] - vref[k]) * (vec->data[k] - vref[k]);
    for (int k = 0; k < 3; ++k) dw += (w[k] - wref[k]) * (w[k] - wref[k]);
    CHECK("dsyev_doc_example", nd == 3 && dv < 1e-6 && dw < 1e-6);
    auto [none, w2, nd2] = rest::_dsyev(a, 'N');
    CHECK("dsyev_doc_example_values_only", !none.has_value() && std::fabs(w2[2] - wref[2]) < 1e-12);
    // _power(-0.5): X A X = I (A has eigenvalues of both signs: the negative
    // ones fall below the threshold and are dropped, so only the projector
    // onto the positive eigenvector remains)
    auto x = rest::_power(a, -0.5, 1e-10);
    double xax = 0.0;  // u^T (X A X) u for the kept eigenvector u = column 2
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s2 = 0.0;
        for (int k = 0; k < 3; ++k)
          for (int l = 0; l < 3; ++l) s2 += (*x)(i, k) * a(k, l) * (*x)(l, j);
        xax += vec->data[6 + i] * s2 * vec->data[6 + j];
      }
    CHECK("power_drops_negative_eigenvalues", std::fabs(xax - 1.0) < 1e-12);
  }
}

int main(int argc, char** argv) {
  const bool cpu_only = argc > 1 && std::strcmp(argv[1], "--cpu-only") == 0;
  try {
    cpu_checks();
    if (!cpu_only) gpu_checks();
  } catch (const std::exception& ex) {
    std::printf("FAIL exception: %s\n", ex.what());
    ++failures;
  }
  std::printf("%s (%d failures)\n", failures ? "FAILED" : "ALL OK", failures);
  return failures ? 1 : 0;
}
