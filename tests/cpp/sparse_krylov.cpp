// sparse_krylov.cpp -- the drop-in C++ API on the HIP domain, exercised the way the reference's
// tests/src/sparse.cpp and examples/tutorial/iterative.cpp use it: STL matrices copied into
// compressed_matrix / sliced_ell_matrix, prod() against the serial STL product, vector
// expressions against host loops, and solve() with cg_tag / bicgstab_tag / gmres_tag, Jacobi and a
// user-written preconditioner, checked by the FP64 true residual.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <vector>

#include "viennacl/compressed_matrix.hpp"
#include "viennacl/coordinate_matrix.hpp"
#include "viennacl/ell_matrix.hpp"
#include "viennacl/hyb_matrix.hpp"
#include "viennacl/io/matrix_market.hpp"
#include "viennacl/linalg/bicgstab.hpp"
#include "viennacl/linalg/cg.hpp"
#include "viennacl/linalg/gmres.hpp"
#include "viennacl/linalg/inner_prod.hpp"
#include "viennacl/linalg/jacobi_precond.hpp"
#include "viennacl/linalg/mixed_precision_cg.hpp"
#include "viennacl/linalg/norm_1.hpp"
#include "viennacl/linalg/norm_2.hpp"
#include "viennacl/linalg/norm_inf.hpp"
#include "viennacl/linalg/prod.hpp"
#include "viennacl/sliced_ell_matrix.hpp"
#include "viennacl/tools/matrix_generation.hpp"
#include "viennacl/vector.hpp"

static int failures = 0;
#define CHECK(cond, ...)                               \
  do {                                                 \
    if (!(cond)) {                                     \
      std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
      std::printf(__VA_ARGS__);                        \
      std::printf("\n");                               \
      ++failures;                                      \
    }                                                  \
  } while (0)

template <typename T>
using StlMatrix = std::vector<std::map<unsigned int, T> >;

template <typename T>
std::vector<T> stl_prod(StlMatrix<T> const& A, std::vector<T> const& x)
{
  std::vector<T> y(A.size(), T(0));
  for (size_t i = 0; i < A.size(); ++i)
    for (auto const& kv : A[i]) y[i] += kv.second * x[kv.first];
  return y;
}

template <typename T>
double rel_inf(std::vector<T> const& a, std::vector<T> const& b)
{
  double num = 0, den = 0;
  for (size_t i = 0; i < a.size(); ++i) {
    num = std::max(num, std::fabs(double(a[i]) - double(b[i])));
    den = std::max(den, std::fabs(double(b[i])));
  }
  return den > 0 ? num / den : num;
}

/** Random nonsymmetric, diagonally dominant matrix with ragged rows (some empty off-diagonal sets,
 *  some long rows) so that every kernel path of the CSR planner is reached. */
template <typename T>
StlMatrix<T> random_matrix(unsigned n, std::mt19937_64& g)
{
  StlMatrix<T> A(n);
  std::uniform_real_distribution<double> u(-1, 1);
  for (unsigned i = 0; i < n; ++i) {
    unsigned len = (i % 97 == 0) ? 3000 : (i % 13 == 0 ? 0 : unsigned(g() % 9));
    double s = 0;
    for (unsigned k = 0; k < len; ++k) {
      unsigned j = unsigned(g() % n);
      if (j == i) continue;
      T v = T(u(g));
      A[i][j] += v;
    }
    for (auto const& kv : A[i]) s += std::fabs(double(kv.second));
    A[i][i] = T(1.5 * s + 1.0);
  }
  return A;
}

template <typename T>
std::vector<T> to_host(viennacl::vector<T> const& v)
{
  std::vector<T> h(v.size());
  viennacl::copy(v, h);
  return h;
}

template <typename T>
double true_relres(StlMatrix<T> const& A, std::vector<T> const& b, std::vector<T> const& x)
{
  std::vector<double> r(b.size());
  double nr = 0, nb = 0;
  for (size_t i = 0; i < A.size(); ++i) {
    double s = 0;
    for (auto const& kv : A[i]) s += double(kv.second) * double(x[kv.first]);
    double d = double(b[i]) - s;
    nr += d * d;
    nb += double(b[i]) * double(b[i]);
  }
  return std::sqrt(nr / nb);
}

/** A user preconditioner unknown to the library (diagonal scaling written by hand), so that
 *  solve() takes its generic path built on the HIP vector operations. */
template <typename T>
struct scaled_identity
{
  T s;
  void apply(viennacl::vector<T>& v) const { v = s * v; }
};

template <typename T>
void test_spmv_and_vectors(double eps)
{
  std::mt19937_64 g(7);
  const unsigned n = 5000;
  StlMatrix<T> A = random_matrix<T>(n, g);
  std::vector<T> x(n), y0(n);
  std::uniform_real_distribution<double> u(-1, 1);
  for (unsigned i = 0; i < n; ++i) { x[i] = T(u(g)); y0[i] = T(u(g)); }

  viennacl::compressed_matrix<T> vA;
  viennacl::copy(A, vA);
  viennacl::sliced_ell_matrix<T> sA;
  viennacl::copy(A, sA);
  CHECK(vA.size1() == n && vA.size2() == n, "sizes");

  viennacl::vector<T> vx(n), vy(n);
  viennacl::copy(x, vx);
  viennacl::copy(y0, vy);
  std::vector<T> ref = stl_prod(A, x);

  vy = viennacl::linalg::prod(vA, vx);
  CHECK(rel_inf(to_host(vy), ref) < eps, "csr prod %g", rel_inf(to_host(vy), ref));
  vy = viennacl::linalg::prod(sA, vx);
  CHECK(rel_inf(to_host(vy), ref) < eps, "sell prod %g", rel_inf(to_host(vy), ref));

  // y += A x and y -= A x (compressed_matrix.hpp op_executor inplace_add / inplace_sub)
  viennacl::copy(y0, vy);
  vy += viennacl::linalg::prod(vA, vx);
  std::vector<T> r2(n);
  for (unsigned i = 0; i < n; ++i) r2[i] = y0[i] + ref[i];
  CHECK(rel_inf(to_host(vy), r2) < eps, "csr prod += %g", rel_inf(to_host(vy), r2));
  viennacl::copy(y0, vy);
  vy -= viennacl::linalg::prod(sA, vx);
  for (unsigned i = 0; i < n; ++i) r2[i] = y0[i] - ref[i];
  CHECK(rel_inf(to_host(vy), r2) < eps, "sell prod -= %g", rel_inf(to_host(vy), r2));

  // aliasing: x = A x goes through a temporary
  viennacl::vector<T> vz = vx;
  vz = viennacl::linalg::prod(vA, vz);
  CHECK(rel_inf(to_host(vz), ref) < eps, "aliased prod %g", rel_inf(to_host(vz), ref));

  // round trip of the CSR matrix back to STL
  StlMatrix<T> back(n);
  viennacl::copy(vA, back);
  bool same = true;
  for (unsigned i = 0; i < n && same; ++i) same = (back[i] == A[i]);
  CHECK(same, "compressed_matrix -> STL round trip");

  // vector expressions against host loops (tests/src/vector_float_double.cpp style)
  T a = T(0.75), b = T(-1.25);
  viennacl::vector<T> w(n);
  w = a * vx + b * vy;
  std::vector<T> hy = to_host(vy), hw(n);
  for (unsigned i = 0; i < n; ++i) hw[i] = a * x[i] + b * hy[i];
  CHECK(rel_inf(to_host(w), hw) < eps, "avbv %g", rel_inf(to_host(w), hw));
  w += vx;
  for (unsigned i = 0; i < n; ++i) hw[i] += x[i];
  CHECK(rel_inf(to_host(w), hw) < eps, "w += x");
  w -= a * vx;
  for (unsigned i = 0; i < n; ++i) hw[i] -= a * x[i];
  CHECK(rel_inf(to_host(w), hw) < eps, "w -= a x");

  double ip = 0, nn = 0;
  for (unsigned i = 0; i < n; ++i) { ip += double(x[i]) * double(hw[i]); nn += double(x[i]) * double(x[i]); }
  T dip = viennacl::linalg::inner_prod(vx, w);
  T dn = viennacl::linalg::norm_2(vx);
  CHECK(std::fabs(double(dip) - ip) <= eps * std::fabs(ip) * 10, "inner_prod %g vs %g", double(dip), ip);
  CHECK(std::fabs(double(dn) - std::sqrt(nn)) <= eps * std::sqrt(nn) * 10, "norm_2");

  viennacl::vector<T> q = viennacl::linalg::element_div(vx, w);
  std::vector<T> hq(n);
  for (unsigned i = 0; i < n; ++i) hq[i] = x[i] / hw[i];
  CHECK(rel_inf(to_host(q), hq) < eps, "element_div");

  // strided view: every other entry of a 2n vector
  viennacl::vector<T> big(2 * n);
  std::vector<T> hb(2 * n, T(0));
  for (unsigned i = 0; i < n; ++i) hb[2 * i] = x[i];
  viennacl::copy(hb, big);
  viennacl::vector_base<T> view(big.handle(), n, 0, 2);
  T dv = viennacl::linalg::inner_prod(view, vx);
  CHECK(std::fabs(double(dv) - nn) <= eps * nn * 10, "strided inner_prod");
  viennacl::vector<T> vy2(n);
  vy2 = viennacl::linalg::prod(vA, view);
  CHECK(rel_inf(to_host(vy2), ref) < eps, "prod with strided x");

  // Jacobi diagonal (row_info SPARSE_ROW_DIAGONAL)
  viennacl::linalg::jacobi_precond<viennacl::compressed_matrix<T> > J(vA, viennacl::linalg::jacobi_tag());
  std::vector<T> hd = to_host(viennacl::vector<T>(J.diagonal()));
  bool diag_ok = true;
  for (unsigned i = 0; i < n; ++i) diag_ok = diag_ok && hd[i] == A[i].at(i);
  CHECK(diag_ok, "jacobi diagonal");
}

template <typename T>
void test_solvers(double tol_solver, double tol_check)
{
  // 2-D Laplacian from tools::generate_fdm_laplace, b ~ U[-1,1]
  viennacl::compressed_matrix<T> A;
  viennacl::tools::generate_fdm_laplace(A, 48, 40);
  const unsigned n = 48 * 40;
  StlMatrix<T> hA(n);
  viennacl::copy(A, hA);
  viennacl::sliced_ell_matrix<T> S;
  viennacl::copy(hA, S);

  std::mt19937_64 g(0);
  std::uniform_real_distribution<double> u(-1, 1);
  std::vector<T> b(n);
  for (auto& v : b) v = T(u(g));
  viennacl::vector<T> vb(n);
  viennacl::copy(b, vb);

  using namespace viennacl::linalg;
  {
    cg_tag tag(tol_solver, 1000);
    viennacl::vector<T> x = solve(A, vb, tag);
    double rr = true_relres(hA, b, to_host(x));
    CHECK(rr < tol_check && tag.iters() > 10 && tag.iters() < 1000, "cg csr: iters %u relres %g", tag.iters(), rr);
    cg_tag tag2(tol_solver, 1000);
    viennacl::vector<T> x2 = solve(S, vb, tag2);
    CHECK(tag2.iters() == tag.iters(), "cg sell iters %u vs csr %u", tag2.iters(), tag.iters());
    CHECK(true_relres(hA, b, to_host(x2)) < tol_check, "cg sell residual");
  }
  {
    jacobi_precond<viennacl::compressed_matrix<T> > J(A, jacobi_tag());
    cg_tag tag(tol_solver, 1000);
    viennacl::vector<T> x = solve(A, vb, tag, J);
    CHECK(true_relres(hA, b, to_host(x)) < tol_check, "jacobi cg residual");
    // a preconditioner the library does not know: generic PCG on the HIP vector ops
    cg_tag tag2(tol_solver, 1000);
    viennacl::vector<T> x2 = solve(A, vb, tag2, scaled_identity<T>{T(0.25)});
    CHECK(true_relres(hA, b, to_host(x2)) < tol_check, "generic pcg residual");
    CHECK(std::abs(int(tag2.iters()) - int(tag.iters())) <= 2, "generic pcg iters %u vs %u", tag2.iters(), tag.iters());
  }
  {
    bicgstab_tag tag(tol_solver, 1000, 200);
    viennacl::vector<T> x = solve(A, vb, tag);
    CHECK(true_relres(hA, b, to_host(x)) < tol_check, "bicgstab residual (iters %lu)", (unsigned long)tag.iters());
    jacobi_precond<viennacl::compressed_matrix<T> > J(A, jacobi_tag());
    bicgstab_tag tag2(tol_solver, 1000, 200);
    viennacl::vector<T> x2 = solve(A, vb, tag2, J);
    CHECK(true_relres(hA, b, to_host(x2)) < tol_check * 10, "jacobi bicgstab residual");
    bicgstab_tag tag3(tol_solver, 1000, 200);
    viennacl::vector<T> x3 = solve(S, vb, tag3, scaled_identity<T>{T(0.25)});
    CHECK(true_relres(hA, b, to_host(x3)) < tol_check * 10, "generic bicgstab residual");
  }
  {
    // solver objects: initial guess and a monitor that stops early
    cg_solver<viennacl::vector<T> > solver(cg_tag(tol_solver, 1000));
    viennacl::vector<T> guess(n);
    std::vector<T> hg(n, T(0.1));
    viennacl::copy(hg, guess);
    solver.set_initial_guess(guess);
    viennacl::vector<T> x = solver(A, vb);
    CHECK(true_relres(hA, b, to_host(x)) < tol_check, "cg_solver with guess");

    struct Count { int calls; };
    Count c = {0};
    cg_solver<viennacl::vector<T> > s2(cg_tag(tol_solver, 1000));
    s2.set_monitor([](viennacl::vector<T> const& xi, T, void* d) {
      Count* cc = static_cast<Count*>(d);
      return ++cc->calls >= 5 && xi.size() > 0;
    }, &c);
    s2(A, vb);
    CHECK(c.calls == 5 && s2.tag().iters() == 5, "monitor stop: calls %d iters %u", c.calls, s2.tag().iters());
  }
  {
    // STL overload (cg.hpp:349-363)
    std::vector<T> x = solve(hA, b, cg_tag(tol_solver, 1000), no_precond());
    CHECK(true_relres(hA, b, x) < tol_check, "stl solve");
  }
}

template <typename T>
void test_gmres(double tol_solver, double tol_check)
{
  // 2-D convection-diffusion, upwind (nonsymmetric), b ~ U[-1,1]
  const unsigned px = 50, py = 40, n = px * py;
  StlMatrix<T> hA(n);
  for (unsigned j = 0; j < py; ++j)
    for (unsigned i = 0; i < px; ++i) {
      const unsigned r = j * px + i;
      hA[r][r] = T(4.8);
      if (i > 0) hA[r][r - 1] = T(-1.5);
      if (i + 1 < px) hA[r][r + 1] = T(-1);
      if (j > 0) hA[r][r - px] = T(-1.3);
      if (j + 1 < py) hA[r][r + px] = T(-1);
    }
  viennacl::compressed_matrix<T> A;
  viennacl::copy(hA, A);
  viennacl::sliced_ell_matrix<T> S;
  viennacl::copy(hA, S);
  std::mt19937_64 g(3);
  std::uniform_real_distribution<double> u(-1, 1);
  std::vector<T> b(n);
  for (auto& v : b) v = T(u(g));
  viennacl::vector<T> vb(n);
  viennacl::copy(b, vb);

  using namespace viennacl::linalg;
  {
    gmres_tag tag(tol_solver, 2000, 20);
    viennacl::vector<T> x = solve(A, vb, tag);
    double rr = true_relres(hA, b, to_host(x));
    CHECK(rr < tol_check && tag.iters() >= 20 && tag.iters() < 2000, "gmres csr: iters %u relres %g", tag.iters(), rr);
    gmres_tag tag2(tol_solver, 2000, 20);
    viennacl::vector<T> x2 = solve(S, vb, tag2);
    CHECK(tag2.iters() == tag.iters(), "gmres sell iters %u vs csr %u", tag2.iters(), tag.iters());
    CHECK(true_relres(hA, b, to_host(x2)) < tol_check, "gmres sell residual");
  }
  {
    // preconditioned: the Householder GMRES on the HIP vector operations
    jacobi_precond<viennacl::compressed_matrix<T> > J(A, jacobi_tag());
    gmres_tag tag(tol_solver, 2000, 20);
    viennacl::vector<T> x = solve(A, vb, tag, J);
    double rr = true_relres(hA, b, to_host(x));
    CHECK(rr < 10 * tol_check, "jacobi gmres: iters %u relres %g", tag.iters(), rr);
    gmres_tag tag2(tol_solver, 2000, 20);
    viennacl::vector<T> x2 = solve(A, vb, tag2, scaled_identity<T>{T(0.25)});
    CHECK(true_relres(hA, b, to_host(x2)) < 10 * tol_check, "generic gmres residual");
  }
  {
    // per-cycle monitor that stops after two cycles, and the initial guess
    struct Count { int calls; };
    Count c = {0};
    gmres_solver<viennacl::vector<T> > s(gmres_tag(tol_solver, 2000, 15));
    s.set_monitor([](viennacl::vector<T> const& xi, T, void* d) {
      Count* cc = static_cast<Count*>(d);
      return ++cc->calls >= 2 && xi.size() > 0;
    }, &c);
    s(A, vb);
    CHECK(c.calls == 2 && s.tag().iters() == 30, "gmres monitor stop: calls %d iters %u", c.calls, s.tag().iters());
    gmres_solver<viennacl::vector<T> > s2(gmres_tag(tol_solver, 2000, 20));
    viennacl::vector<T> guess(n);
    std::vector<T> hg(n, T(0.1));
    viennacl::copy(hg, guess);
    s2.set_initial_guess(guess);
    CHECK(true_relres(hA, b, to_host(s2(A, vb))) < tol_check, "gmres_solver with guess");
  }
  {
    std::vector<T> x = solve(hA, b, gmres_tag(tol_solver, 2000, 20), no_precond());
    CHECK(true_relres(hA, b, x) < tol_check, "stl gmres solve");
  }
}

template <typename T, typename MatrixT>
void check_format(const char* name, StlMatrix<T> const& hA, std::vector<T> const& b, viennacl::vector<T> const& vb,
                  double eps, double tol_solver, double tol_check)
{
  using namespace viennacl::linalg;
  const unsigned n = unsigned(hA.size());
  MatrixT M;
  viennacl::copy(hA, M);
  CHECK(M.size1() == n && M.size2() == n, "%s sizes", name);
  // prod against the STL product, y = A x and y += A x
  std::vector<T> hx(n);
  for (unsigned i = 0; i < n; ++i) hx[i] = T(std::sin(0.37 * i));
  viennacl::vector<T> vx(n), vy(n);
  viennacl::copy(hx, vx);
  vy = viennacl::linalg::prod(M, vx);
  std::vector<T> ref = stl_prod(hA, hx);
  CHECK(rel_inf(to_host(vy), ref) < eps, "%s prod %g", name, rel_inf(to_host(vy), ref));
  vy += viennacl::linalg::prod(M, vx);
  std::vector<T> r2(ref);
  for (auto& v : r2) v *= T(2);
  CHECK(rel_inf(to_host(vy), r2) < eps, "%s prod += %g", name, rel_inf(to_host(vy), r2));
  // back to the STL matrix
  StlMatrix<T> back(n);
  viennacl::copy(M, back);
  bool same = true;
  for (unsigned i = 0; i < n; ++i) same = same && back[i] == hA[i];
  CHECK(same, "%s -> STL round trip", name);
  // solvers
  bicgstab_tag tag(tol_solver, 2000);
  viennacl::vector<T> x = solve(M, vb, tag);
  CHECK(true_relres(hA, b, to_host(x)) < tol_check, "%s bicgstab (iters %lu)", name, (unsigned long)tag.iters());
  gmres_tag gtag(tol_solver, 2000, 20);
  viennacl::vector<T> x2 = solve(M, vb, gtag);
  CHECK(true_relres(hA, b, to_host(x2)) < tol_check, "%s gmres (iters %u)", name, gtag.iters());
}

template <typename T>
void test_formats(double eps, double tol_solver, double tol_check)
{
  // nonsymmetric convection-diffusion with a few long rows, so that HYB has a CSR part
  const unsigned px = 40, py = 30, n = px * py;
  StlMatrix<T> hA(n);
  for (unsigned j = 0; j < py; ++j)
    for (unsigned i = 0; i < px; ++i) {
      const unsigned r = j * px + i;
      hA[r][r] = T(4.8);
      if (i > 0) hA[r][r - 1] = T(-1.5);
      if (i + 1 < px) hA[r][r + 1] = T(-1);
      if (j > 0) hA[r][r - px] = T(-1.3);
      if (j + 1 < py) hA[r][r + px] = T(-1);
    }
  for (unsigned r = 0; r < n; r += 97)
    for (unsigned k = 1; k <= 12; ++k) hA[r][(r + 37 * k) % n] += T(-0.01);
  std::mt19937_64 g(5);
  std::uniform_real_distribution<double> u(-1, 1);
  std::vector<T> b(n);
  for (auto& v : b) v = T(u(g));
  viennacl::vector<T> vb(n);
  viennacl::copy(b, vb);
  check_format<T, viennacl::coordinate_matrix<T> >("coordinate_matrix", hA, b, vb, eps, tol_solver, tol_check);
  check_format<T, viennacl::ell_matrix<T> >("ell_matrix", hA, b, vb, eps, tol_solver, tol_check);
  check_format<T, viennacl::hyb_matrix<T> >("hyb_matrix", hA, b, vb, eps, tol_solver, tol_check);
  // Jacobi on coordinate_matrix (row_info of the COO diagonal, jacobi_precond.hpp:104-130)
  viennacl::coordinate_matrix<T> C;
  viennacl::copy(hA, C);
  viennacl::linalg::jacobi_precond<viennacl::coordinate_matrix<T> > J(C, viennacl::linalg::jacobi_tag());
  viennacl::linalg::bicgstab_tag tag(tol_solver, 2000);
  viennacl::vector<T> x = viennacl::linalg::solve(C, vb, tag, J);
  CHECK(true_relres(hA, b, to_host(x)) < 10 * tol_check, "jacobi bicgstab on coordinate_matrix");
  viennacl::hyb_matrix<T> H;
  viennacl::copy(hA, H);
  CHECK(H.csr_nnz() > 1 && H.internal_ellnnz() == 5, "hyb split: ell width %lu, csr part %lu",
        (unsigned long)H.internal_ellnnz(), (unsigned long)H.csr_nnz());
}

void test_mixed_precision()
{
  // 2-D variable-coefficient problem: diagonal entries vary, so Jacobi changes the iteration
  viennacl::compressed_matrix<double> A;
  viennacl::tools::generate_fdm_laplace(A, 60, 50);
  const unsigned n = 60 * 50;
  StlMatrix<double> hA(n);
  viennacl::copy(A, hA);
  for (unsigned i = 0; i < n; ++i)
    for (auto& kv : hA[i]) kv.second *= (1.0 + 0.5 * std::sin(0.01 * i)) * (1.0 + 0.5 * std::sin(0.01 * kv.first));
  for (unsigned i = 0; i < n; ++i) {
    double s = 0;
    for (auto const& kv : hA[i]) if (kv.first != i) s += std::fabs(kv.second);
    hA[i][i] = s + 0.01;
  }
  viennacl::copy(hA, A);
  viennacl::sliced_ell_matrix<double> S;
  viennacl::copy(hA, S);
  std::mt19937_64 g(0);
  std::uniform_real_distribution<double> u(-1, 1);
  std::vector<double> b(n);
  for (auto& v : b) v = u(g);
  viennacl::vector<double> vb(n);
  viennacl::copy(b, vb);

  using namespace viennacl::linalg;
  mixed_precision_cg_tag tag(1e-10, 5000);
  viennacl::vector<double> x = solve(A, vb, tag);
  double rr = true_relres(hA, b, to_host(x));
  CHECK(rr < 1e-10 && tag.error() < 1e-10 && tag.corrections() >= 2, "mixed cg: relres %g iters %u corr %u", rr,
        tag.iters(), tag.corrections());
  mixed_precision_cg_tag tag1(1e-10, 5000);
  tag1.reliable_updates(true);
  viennacl::vector<double> x1 = solve(S, vb, tag1);
  double rr1 = true_relres(hA, b, to_host(x1));
  CHECK(rr1 < 1e-10 && tag1.iters() < tag.iters(), "reliable mixed cg: relres %g iters %u vs %u", rr1, tag1.iters(),
        tag.iters());
  jacobi_precond<viennacl::compressed_matrix<double> > J(A, jacobi_tag());
  mixed_precision_cg_tag tag2(1e-10, 5000);
  tag2.reliable_updates(true);
  viennacl::vector<double> x2 = solve(A, vb, tag2, J);
  CHECK(true_relres(hA, b, to_host(x2)) < 1e-10, "jacobi mixed cg");
  bool threw = false;
  try {
    solve(A, vb, tag, J);
  } catch (std::invalid_argument const&) {
    threw = true;
  }
  CHECK(threw, "reference schedule with a preconditioner must throw");
  cg_tag t64(1e-12, 5000);
  viennacl::vector<double> x64 = solve(A, vb, t64);
  std::vector<double> h64 = to_host(x64), h1 = to_host(x1);
  double num = 0, den = 0;
  for (unsigned i = 0; i < n; ++i) { num += (h1[i] - h64[i]) * (h1[i] - h64[i]); den += h64[i] * h64[i]; }
  CHECK(std::sqrt(num / den) < 1e-7, "mixed vs FP64 solution %g", std::sqrt(num / den));
}

void test_matrix_market()
{
  // write an STL matrix, read it back into the STL type and into compressed_matrix
  // (io/matrix_market.hpp; the writer keeps 6 significant digits, as the reference's does)
  std::mt19937_64 g(11);
  StlMatrix<double> A = random_matrix<double>(700, g);
  const char* file = "/tmp/vclhip_sparse_krylov_test.mtx";
  viennacl::io::write_matrix_market_file(A, file);
  StlMatrix<double> B;
  long lines = viennacl::io::read_matrix_market_file(B, file);
  size_t nnz = 0;
  for (auto const& r : A) nnz += r.size();
  CHECK(lines == long(nnz) + 2 && B.size() == A.size(), "mm read lines %ld", lines);
  bool same_pattern = true;
  double max_rel = 0;
  for (size_t i = 0; i < A.size() && same_pattern; ++i) {
    same_pattern = A[i].size() == B[i].size();
    for (auto const& kv : A[i]) {
      auto it = B[i].find(kv.first);
      if (it == B[i].end()) { same_pattern = false; break; }
      max_rel = std::max(max_rel, std::fabs(it->second - kv.second) / std::fabs(kv.second));
    }
  }
  CHECK(same_pattern && max_rel < 5e-6, "mm round trip: pattern %d, max rel %g", int(same_pattern), max_rel);
  viennacl::compressed_matrix<double> vA;
  CHECK(viennacl::io::read_matrix_market_file(vA, file) > 0 && vA.nnz() == nnz, "mm into compressed_matrix");
  std::vector<double> x(700, 1.0);
  viennacl::vector<double> vx(700), vy(700);
  viennacl::copy(x, vx);
  vy = viennacl::linalg::prod(vA, vx);
  CHECK(rel_inf(to_host(vy), stl_prod(B, x)) < 1e-13, "SpMV of the file's matrix");
  CHECK(viennacl::io::read_matrix_market_file(B, "does_not_exist.mtx") == EXIT_FAILURE, "missing file status");
  std::remove(file);
}

/* norm_1, norm_inf, index_norm_inf, plane_rotation and swap against host loops
 * (tests/src/vector_float_double.cpp checks the same operations against uBLAS). */
template <typename T>
void test_vector_extras(double eps)
{
  const size_t n = 10007;
  std::mt19937_64 rng(42);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  std::vector<T> hx(n), hy(n);
  for (size_t i = 0; i < n; ++i) {
    hx[i] = T(u(rng));
    hy[i] = T(u(rng));
  }
  hx[777] = T(-3);  // the largest |x_i|, first of two
  hx[9000] = T(3);
  viennacl::vector<T> x(n), y(n);
  viennacl::copy(hx, x);
  viennacl::copy(hy, y);
  double n1 = 0;
  for (size_t i = 0; i < n; ++i) n1 += std::fabs(double(hx[i]));
  CHECK(std::fabs(double(viennacl::linalg::norm_1(x)) - n1) <= eps * n1 * 10, "norm_1");
  CHECK(viennacl::linalg::index_norm_inf(x) == 777, "index_norm_inf %lu", (unsigned long)viennacl::linalg::index_norm_inf(x));
  CHECK(double(viennacl::linalg::norm_inf(x)) == 3.0, "norm_inf");
  viennacl::vector<T> z(n);
  z.clear();
  CHECK(viennacl::linalg::norm_inf(z) == T(0) && viennacl::linalg::index_norm_inf(z) == 0, "norm_inf of zeros");
  const T c = T(0.6), d = T(0.8);
  viennacl::linalg::plane_rotation(x, y, c, d);
  std::vector<T> rx(n), ry(n);
  for (size_t i = 0; i < n; ++i) {
    rx[i] = c * hx[i] + d * hy[i];
    ry[i] = c * hy[i] - d * hx[i];
  }
  CHECK(rel_inf(to_host(x), rx) < eps && rel_inf(to_host(y), ry) < eps, "plane_rotation");
  const std::vector<T> sx = to_host(x), sy = to_host(y);
  viennacl::swap(x, y);
  CHECK(to_host(x) == sy && to_host(y) == sx, "swap");
}

int main()
{
  int ndev = 0;
  if (vclhip_device_count(&ndev) != 0 || ndev == 0) {
    std::printf("no HIP device\n");
    return 2;
  }
  test_spmv_and_vectors<double>(1e-13);
  test_spmv_and_vectors<float>(1e-5);
  test_solvers<double>(1e-10, 1e-9);
  test_solvers<float>(1e-5, 1e-4);
  test_gmres<double>(1e-10, 1e-9);
  test_gmres<float>(1e-5, 1e-4);
  test_formats<double>(1e-13, 1e-10, 1e-9);
  test_formats<float>(1e-5, 1e-5, 1e-4);
  test_vector_extras<double>(1e-13);
  test_vector_extras<float>(1e-6);
  test_mixed_precision();
  test_matrix_market();
  if (failures) {
    std::printf("%d FAILED\n", failures);
    return 1;
  }
  std::printf("ALL PASSED\n");
  return 0;
}
