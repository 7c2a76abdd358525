// reference_call_sites.cpp -- call sites written the way the reference's own programs write them,
// compiled unchanged against the drop-in headers and run on the HIP domain:
//   * iterator-range copies, viennacl::copy(cpu.begin(), cpu.end(), gpu.begin()) and back
//     (examples/tutorial/iterative.cpp:128-129, tests/src/sparse.cpp:65-70, :366);
//   * vector views: range, slice, project(), and products into and out of them
//     (tests/src/sparse.cpp:158-207, viennacl/vector_proxy.hpp);
//   * the STL product, the tests' ground truth (linalg/prod.hpp:115-131; tests/src/sparse.cpp:370-410);
//   * compressed_matrix resize/reserve/operator()(i, j) (tests/src/sparse.cpp:213-283,
//     compressed_matrix.hpp:946-1076);
//   * a matrix-free operator with apply(x, y) and size1() in prod() and in solve() with cg_tag,
//     bicgstab_tag and gmres_tag (examples/tutorial/matrix-free.cpp:58-82, vector.hpp:3315-3320).
// Each failed check prints one FAIL line; the program ends with ALL PASSED or a nonzero status.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <list>
#include <map>
#include <random>
#include <vector>

#include "viennacl/compressed_matrix.hpp"
#include "viennacl/coordinate_matrix.hpp"
#include "viennacl/ell_matrix.hpp"
#include "viennacl/hyb_matrix.hpp"
#include "viennacl/linalg/bicgstab.hpp"
#include "viennacl/linalg/cg.hpp"
#include "viennacl/linalg/gmres.hpp"
#include "viennacl/linalg/inner_prod.hpp"
#include "viennacl/linalg/jacobi_precond.hpp"
#include "viennacl/linalg/norm_2.hpp"
#include "viennacl/linalg/prod.hpp"
#include "viennacl/sliced_ell_matrix.hpp"
#include "viennacl/tools/matrix_generation.hpp"
#include "viennacl/vector.hpp"
#include "viennacl/vector_proxy.hpp"

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

// The reference test's relative difference between a host vector and a device vector
// (tests/src/sparse.cpp:64-97), read back through the iterator-range copy.
template <typename T>
T diff(std::vector<T> const& v1, viennacl::vector<T>& v2)
{
  std::vector<T> v2_cpu(v2.size());
  viennacl::backend::finish();
  viennacl::copy(v2.begin(), v2.end(), v2_cpu.begin());
  T worst = 0;
  for (std::size_t i = 0; i < v1.size(); ++i) {
    const T m = std::max(std::fabs(v2_cpu[i]), std::fabs(v1[i]));
    const T d = m > 0 ? std::fabs(v2_cpu[i] - v1[i]) / m : T(0);
    worst = std::max(worst, d);
  }
  return worst;
}

// The reference test's entry-wise difference of two sparse matrices (tests/src/sparse.cpp:100-148).
template <typename T, typename MatrixT>
T diff(std::vector<std::map<unsigned int, T> >& cpu_A, MatrixT& vcl_A)
{
  std::vector<std::map<unsigned int, T> > from_gpu(vcl_A.size1());
  viennacl::backend::finish();
  viennacl::copy(vcl_A, from_gpu);
  T error = 0;
  for (std::size_t i = 0; i < cpu_A.size(); ++i)
    for (auto const& kv : cpu_A[i]) {
      const T g = from_gpu.size() > i && from_gpu[i].count(kv.first) ? from_gpu[i][kv.first] : T(0);
      const T m = std::max(std::fabs(kv.second), std::fabs(g));
      if (m > 0) error = std::max(error, std::fabs(kv.second - g) / m);
    }
  for (std::size_t i = 0; i < from_gpu.size(); ++i)
    for (auto const& kv : from_gpu[i]) {
      const T c = i < cpu_A.size() && cpu_A[i].count(kv.first) ? cpu_A[i].at(kv.first) : T(0);
      const T m = std::max(std::fabs(kv.second), std::fabs(c));
      if (m > 0) error = std::max(error, std::fabs(kv.second - c) / m);
    }
  return error;
}

// ------------------------------------------------------------------ iterator-range copies
template <typename T>
void test_iterator_copies(T eps)
{
  const std::size_t n = 1000;
  std::vector<T> host(n);
  for (std::size_t i = 0; i < n; ++i) host[i] = T(0.5) + T(i % 37) / T(8);
  viennacl::vector<T> vcl(n);
  viennacl::copy(host.begin(), host.end(), vcl.begin());  // examples/tutorial/iterative.cpp:128
  CHECK(diff(host, vcl) == 0, "copy(cpu.begin(), cpu.end(), gpu.begin())");

  std::vector<T> back(n, T(-1));
  viennacl::copy(vcl.begin(), vcl.end(), back.begin());
  CHECK(back == host, "copy(gpu.begin(), gpu.end(), cpu.begin())");

  // a shorter range into the middle of the device vector (copy(std_vec.begin(), ..., vcl.begin() + k))
  std::vector<T> part(100, T(7));
  viennacl::copy(part.begin(), part.end(), vcl.begin() + 250);
  for (std::size_t i = 250; i < 350; ++i) host[i] = T(7);
  CHECK(diff(host, vcl) == 0, "copy into gpu.begin() + 250");

  // a std::list (not one linear piece of memory) in both directions
  std::list<T> lst(host.begin(), host.begin() + 64);
  viennacl::vector<T> small(64);
  viennacl::copy(lst.begin(), lst.end(), small.begin());
  std::list<T> lst_back(64, T(0));
  viennacl::copy(small.begin(), small.end(), lst_back.begin());
  CHECK(lst_back == lst, "copy through std::list iterators");

  // fast_copy on linear host memory, and entry access through the iterators
  std::vector<T> fast(n);
  viennacl::fast_copy(vcl.begin(), vcl.end(), fast.begin());
  CHECK(fast == host, "fast_copy(gpu.begin(), gpu.end(), cpu.begin())");
  auto it = vcl.begin() + 5;
  *it = T(3.25);
  host[5] = T(3.25);
  CHECK(T(*(vcl.begin() + 5)) == T(3.25) && T(vcl[5]) == T(3.25), "*iterator = value");
  viennacl::vector<T> other(n);
  viennacl::copy(vcl.begin(), vcl.end(), other.begin());  // device range -> device range
  CHECK(diff(host, other) == 0, "copy(gpu.begin(), gpu.end(), gpu2.begin())");
  (void)eps;
}

// ------------------------------------------------------------------ vector proxies
template <typename T>
void test_vector_proxies(T eps)
{
  const std::size_t n = 40;
  std::vector<T> hx(n), hy(n);
  for (std::size_t i = 0; i < n; ++i) {
    hx[i] = T(1) + T(i) / T(4);
    hy[i] = T(-2) + T(i % 7);
  }
  viennacl::vector<T> x(n), y(n);
  viennacl::copy(hx.begin(), hx.end(), x.begin());
  viennacl::copy(hy.begin(), hy.end(), y.begin());

  // y(range(10, 20)) = x(range(0, 10)) + 2 * x(range(20, 30))
  viennacl::range r_dst(10, 20), r_a(0, 10), r_b(20, 30);
  viennacl::project(y, r_dst) = viennacl::project(x, r_a) + T(2) * viennacl::project(x, r_b);
  for (std::size_t i = 0; i < 10; ++i) hy[10 + i] = hx[i] + T(2) * hx[20 + i];
  CHECK(diff(hy, y) <= eps, "range expression");

  // y(slice(1, 3, 8)) -= x(slice(0, 5, 8))
  viennacl::slice s_dst(1, 3, 8), s_src(0, 5, 8);
  viennacl::vector_slice<viennacl::vector<T> > ys = viennacl::project(y, s_dst);
  ys -= viennacl::project(x, s_src);
  for (std::size_t i = 0; i < 8; ++i) hy[1 + 3 * i] -= hx[5 * i];
  CHECK(diff(hy, y) <= eps, "slice -=");

  // nested proxies: a slice of a range and a range of a slice
  viennacl::vector_range<viennacl::vector<T> > xr = viennacl::project(x, viennacl::range(4, 36));
  viennacl::vector_slice<viennacl::vector<T> > xrs = viennacl::project(xr, viennacl::slice(2, 3, 9));  // x[6 + 3i]
  viennacl::vector_slice<viennacl::vector<T> > xrsr = viennacl::project(xrs, viennacl::range(1, 5));  // x[9 + 3i]
  std::vector<T> got(4);
  viennacl::copy(xrsr.begin(), xrsr.end(), got.begin());
  bool ok = true;
  for (std::size_t i = 0; i < 4; ++i) ok = ok && got[i] == hx[9 + 3 * i];
  CHECK(ok, "project(project(project(x, range), slice), range)");
  T dot_ref = 0;
  for (std::size_t i = 0; i < 9; ++i) dot_ref += hx[6 + 3 * i] * hx[6 + 3 * i];
  const T dot = viennacl::linalg::inner_prod(xrs, xrs);
  CHECK(std::fabs(dot - dot_ref) <= eps * dot_ref, "inner_prod on a slice: %g vs %g", double(dot), double(dot_ref));
  CHECK(std::fabs(viennacl::linalg::norm_2(xrs) - std::sqrt(dot_ref)) <= eps * std::sqrt(dot_ref), "norm_2 on a slice");

  // writing a host vector into a range and reading it back
  std::vector<T> block(6, T(9.5));
  viennacl::vector_range<viennacl::vector<T> > yr = viennacl::project(y, viennacl::range(30, 36));
  viennacl::copy(block, yr);
  for (std::size_t i = 30; i < 36; ++i) hy[i] = T(9.5);
  CHECK(diff(hy, y) == 0, "copy(host, vector_range)");
  std::vector<T> yr_back;
  viennacl::copy(yr, yr_back);
  CHECK(yr_back == block, "copy(vector_range, host)");

  // entry proxies on a view write into the parent
  ys(2) = T(-4);
  hy[1 + 3 * 2] = T(-4);
  CHECK(diff(hy, y) == 0, "vector_slice(i) = value");
}

// ------------------------------------------------------------------ products with strided vectors
// tests/src/sparse.cpp:151-207: project(result, slice(1, 3, 5)) = prod(A, project(rhs, slice(3, 2, 4))).
template <typename T, typename MatrixT>
void strided_product(const char* name, T eps)
{
  std::vector<std::map<unsigned int, T> > std_A(5);
  std_A[0][0] = T(2.0); std_A[0][2] = T(-1.0);
  std_A[1][0] = T(3.0); std_A[1][2] = T(-5.0);
  std_A[2][1] = T(5.0); std_A[2][2] = T(-2.0);
  std_A[3][2] = T(1.0); std_A[3][3] = T(-6.0);
  std_A[4][1] = T(7.0); std_A[4][2] = T(-5.0);

  const std::size_t n = 16;
  std::vector<T> rhs(n), result(n);
  for (std::size_t i = 0; i < n; ++i) {
    rhs[i] = T(1) + T(i) / T(3);
    result[i] = T(i % 5) - T(1);
  }
  std::vector<T> expect = result;
  for (std::size_t i = 0; i < 5; ++i) {
    T v = 0;
    for (auto const& kv : std_A[i]) v += kv.second * rhs[3 + 2 * kv.first];
    expect[1 + 3 * i] = v;
  }

  MatrixT A2;
  viennacl::copy(std_A, A2);
  viennacl::vector<T> vcl_rhs(n), vcl_result(n);
  viennacl::copy(rhs.begin(), rhs.end(), vcl_rhs.begin());
  viennacl::copy(result.begin(), result.end(), vcl_result.begin());
  viennacl::project(vcl_result, viennacl::slice(1, 3, 5)) =
      viennacl::linalg::prod(A2, viennacl::project(vcl_rhs, viennacl::slice(3, 2, 4)));
  CHECK(diff(expect, vcl_result) <= eps, "%s: strided product, part 1", name);

  // part 2: a contiguous right-hand side gathered entry by entry
  viennacl::vector<T> vec(4);
  vec(0) = rhs[3];
  vec(1) = rhs[5];
  vec(2) = rhs[7];
  vec(3) = rhs[9];
  vcl_result(1) = T(1.0);
  vcl_result(4) = T(1.0);
  vcl_result(7) = T(1.0);
  vcl_result(10) = T(1.0);
  vcl_result(13) = T(1.0);
  viennacl::project(vcl_result, viennacl::slice(1, 3, 5)) = viennacl::linalg::prod(A2, vec);
  CHECK(diff(expect, vcl_result) <= eps, "%s: strided product, part 2", name);
}

// ------------------------------------------------------------------ STL product and += / -=
template <typename T>
std::vector<std::map<unsigned int, T> > random_stl(unsigned n, std::mt19937_64& g)
{
  std::vector<std::map<unsigned int, T> > A(n);
  std::uniform_real_distribution<double> u(-1, 1);
  for (unsigned i = 0; i < n; ++i) {
    const unsigned len = (i % 101 == 0) ? 300 : unsigned(g() % 12);
    for (unsigned k = 0; k < len; ++k) A[i][unsigned(g() % n)] = T(u(g));
    A[i][i] = T(0.5);  // tests/src/sparse.cpp:351: row sums away from zero
  }
  return A;
}

template <typename T, typename MatrixT>
void products_against_stl(const char* name, std::vector<std::map<unsigned int, T> > const& std_matrix,
                          std::vector<T> const& rhs, T eps)
{
  MatrixT A;
  viennacl::copy(std_matrix, A);
  viennacl::vector<T> vcl_rhs(rhs.size()), vcl_result(rhs.size());
  viennacl::copy(rhs.begin(), rhs.end(), vcl_rhs.begin());

  std::vector<T> result = viennacl::linalg::prod(std_matrix, rhs);  // linalg/prod.hpp:115-131
  vcl_result = viennacl::linalg::prod(A, vcl_rhs);
  CHECK(diff(result, vcl_result) <= eps, "%s: y = prod(A, x): %g", name, double(diff(result, vcl_result)));

  for (std::size_t i = 0; i < result.size(); ++i) result[i] += rhs[i];
  vcl_result = vcl_rhs;
  vcl_result += viennacl::linalg::prod(A, vcl_rhs);
  CHECK(diff(result, vcl_result) <= eps, "%s: y += prod(A, x)", name);

  result = viennacl::linalg::prod(std_matrix, rhs);
  for (std::size_t i = 0; i < result.size(); ++i) result[i] = rhs[i] - result[i];
  vcl_result = vcl_rhs;
  vcl_result -= viennacl::linalg::prod(A, vcl_rhs);
  CHECK(diff(result, vcl_result) <= eps, "%s: y -= prod(A, x)", name);
}

// ------------------------------------------------------------------ resize / reserve / operator()
// tests/src/sparse.cpp:213-283
template <typename T>
void test_resize(T eps)
{
  typedef viennacl::compressed_matrix<T> MatrixT;
  std::vector<std::map<unsigned int, T> > std_A(5);
  for (unsigned i = 0; i < 5; ++i)
    for (unsigned j = 0; j < 5; ++j) std_A[i][j] = T(i) + T(j) / T(10) + (i == 0 && j == 0 ? T(10) : T(0));
  MatrixT vcl_matrix;
  viennacl::copy(std_A, vcl_matrix);
  std::vector<std::map<unsigned int, T> > std_B(std_A.size());
  viennacl::copy(vcl_matrix, std_B);
  CHECK(diff(std_A, vcl_matrix) <= eps, "equality after copy");

  std_A.resize(10);
  vcl_matrix.resize(10, 10, true);
  CHECK(vcl_matrix.size1() == 10 && vcl_matrix.size2() == 10, "resize to larger: dimensions");
  CHECK(diff(std_A, vcl_matrix) <= eps, "resize to larger: entries");

  for (unsigned i = 5; i < 10; ++i)
    for (unsigned j = 5; j < 10; ++j) std_A[i][j] = T(i) + T(j) / T(10);
  viennacl::copy(std_A, vcl_matrix);

  // resize to smaller with preserve: rows 7.. and columns 7.. are dropped
  std::vector<std::map<unsigned int, T> > expect(7);
  for (unsigned i = 0; i < 7; ++i)
    for (auto const& kv : std_A[i])
      if (kv.first < 7) expect[i][kv.first] = kv.second;
  vcl_matrix.resize(7, 7);
  CHECK(vcl_matrix.size1() == 7 && vcl_matrix.size2() == 7, "resize to smaller: dimensions");
  CHECK(diff(expect, vcl_matrix) <= eps, "resize to smaller: entries");

  // the product on the shrunk matrix (its row-block plan was rebuilt)
  std::vector<T> x(7);
  for (unsigned i = 0; i < 7; ++i) x[i] = T(1) + T(i);
  viennacl::vector<T> vx(7), vy(7);
  viennacl::copy(x.begin(), x.end(), vx.begin());
  vy = viennacl::linalg::prod(vcl_matrix, vx);
  CHECK(diff(viennacl::linalg::prod(expect, x), vy) <= eps, "prod after resize to smaller");

  // operator()(i, j): read a stored entry, write it, and insert one outside the pattern
  const T a23 = vcl_matrix(2, 3);
  CHECK(a23 == expect[2][3], "operator()(i, j) read: %g vs %g", double(a23), double(expect[2][3]));
  vcl_matrix(2, 3) = T(-8);
  expect[2][3] = T(-8);
  vcl_matrix(0, 6) = T(4.5);  // row 0 holds columns 0..4 only: inserted
  expect[0][6] = T(4.5);
  CHECK(diff(expect, vcl_matrix) <= eps, "operator()(i, j) write and insert");
  vy = viennacl::linalg::prod(vcl_matrix, vx);
  CHECK(diff(viennacl::linalg::prod(expect, x), vy) <= eps, "prod after operator() inserts");

  // reserve keeps the entries and the product
  const std::size_t stored = vcl_matrix.nnz();
  vcl_matrix.reserve(stored + 50);
  CHECK(vcl_matrix.nnz() == stored + 50, "reserve: nnz() reports the new capacity");
  CHECK(diff(expect, vcl_matrix) <= eps, "reserve keeps the entries");
  vy = viennacl::linalg::prod(vcl_matrix, vx);
  CHECK(diff(viennacl::linalg::prod(expect, x), vy) <= eps, "prod after reserve");

  // resize without preserve leaves no entries
  vcl_matrix.resize(9, 9, false);
  std::vector<std::map<unsigned int, T> > empty(9);
  CHECK(vcl_matrix.size1() == 9 && diff(empty, vcl_matrix) == 0, "resize without preserve");
}

// ------------------------------------------------------------------ matrix-free operator
// examples/tutorial/matrix-free.cpp: the N x N 5-point Poisson operator applied without a stored
// matrix.  apply() reads x, applies the stencil and writes y through the vector API.
template <typename NumericT>
class MyOperator
{
public:
  explicit MyOperator(std::size_t N) : N_(N) {}

  void apply(viennacl::vector_base<NumericT> const& x, viennacl::vector_base<NumericT>& y) const
  {
    if (viennacl::traits::active_handle_id(x) == viennacl::HIP_MEMORY) apply_hip(x, y);
  }
  std::size_t size1() const { return N_ * N_; }
  mutable std::size_t calls = 0;

private:
  void apply_hip(viennacl::vector_base<NumericT> const& x, viennacl::vector_base<NumericT>& y) const
  {
    ++calls;
    std::vector<NumericT> hx(x.size()), hy(x.size());
    viennacl::copy(x.begin(), x.end(), hx.begin());
    for (std::size_t i = 0; i < N_; ++i)
      for (std::size_t j = 0; j < N_; ++j) {
        const std::size_t r = i * N_ + j;
        NumericT v = NumericT(4) * hx[r];
        if (i > 0) v -= hx[r - N_];
        if (i + 1 < N_) v -= hx[r + N_];
        if (j > 0) v -= hx[r - 1];
        if (j + 1 < N_) v -= hx[r + 1];
        hy[r] = v;
      }
    viennacl::copy(hy.begin(), hy.end(), y.begin());
  }
  std::size_t N_;
};

template <typename T>
void test_matrix_free(T tol_solver, T tol_check)
{
  const std::size_t N = 24;
  MyOperator<T> op(N);
  viennacl::vector<T> rhs = viennacl::scalar_vector<T>(N * N, T(-1));  // matrix-free.cpp:93

  // the same operator stored: generate_fdm_laplace numbers row i + j * N, symmetric in i and j
  viennacl::compressed_matrix<T> A;
  viennacl::tools::generate_fdm_laplace(A, N, N);

  viennacl::vector<T> y_free = viennacl::linalg::prod(op, rhs);
  viennacl::vector<T> y_csr = viennacl::linalg::prod(A, rhs);
  std::vector<T> h_free(N * N), h_csr(N * N);
  viennacl::copy(y_free, h_free);
  viennacl::copy(y_csr, h_csr);
  CHECK(h_free == h_csr, "prod(user operator, x) equals the stored Laplacian's product");

  // y += prod(op, x) goes through the operator as well
  viennacl::vector<T> acc = rhs;
  acc += viennacl::linalg::prod(op, rhs);
  std::vector<T> h_acc(N * N);
  viennacl::copy(acc, h_acc);
  bool ok = true;
  for (std::size_t i = 0; i < h_acc.size(); ++i) ok = ok && h_acc[i] == T(-1) + h_csr[i];
  CHECK(ok, "y += prod(user operator, x)");

  auto true_res = [&](viennacl::vector<T> const& x) {
    viennacl::vector<T> r = viennacl::linalg::prod(A, x);
    r = rhs - r;
    return viennacl::linalg::norm_2(r) / viennacl::linalg::norm_2(rhs);
  };
  viennacl::vector<T> x_cg = viennacl::linalg::solve(op, rhs, viennacl::linalg::cg_tag(tol_solver, 500));
  CHECK(true_res(x_cg) <= tol_check, "solve(op, b, cg_tag): true residual %g", double(true_res(x_cg)));
  viennacl::linalg::bicgstab_tag btag(tol_solver, 500);
  viennacl::vector<T> x_bi = viennacl::linalg::solve(op, rhs, btag);
  CHECK(true_res(x_bi) <= tol_check, "solve(op, b, bicgstab_tag): true residual %g", double(true_res(x_bi)));
  viennacl::linalg::gmres_tag gtag(tol_solver, 500, 30);
  viennacl::vector<T> x_gm = viennacl::linalg::solve(op, rhs, gtag);
  // In single precision the Householder GMRES of gmres.hpp:436-626 stops on a degenerate estimate:
  // once r[k] / rho rounds to 1, rho = rho sin(acos(1)) = 0.  The reference build does so on this
  // very problem (24 x 24, tol 1e-5: estimate 0 after 27 steps, true residual 4.6e-2), so only the
  // double-precision solve is held to its tolerance.
  if (sizeof(T) == 8)
    CHECK(true_res(x_gm) <= tol_check, "solve(op, b, gmres_tag): true residual %g", double(true_res(x_gm)));
  CHECK(op.calls > 0, "the operator's apply() ran");

  // the stored matrix gives the same CG solution to solver accuracy
  viennacl::vector<T> x_csr = viennacl::linalg::solve(A, rhs, viennacl::linalg::cg_tag(tol_solver, 500));
  viennacl::vector<T> d = x_csr - x_cg;
  CHECK(viennacl::linalg::norm_2(d) <= T(100) * tol_check * viennacl::linalg::norm_2(x_csr), "matrix-free CG vs stored CG");
}

// ------------------------------------------------------------------ the iterative tutorial's calls
// examples/tutorial/iterative.cpp:120-200 with STL containers in place of uBLAS.
template <typename T>
void test_iterative_tutorial(T tol_check)
{
  const std::size_t N = 30;
  viennacl::compressed_matrix<T> vcl_compressed_matrix;
  viennacl::tools::generate_fdm_laplace(vcl_compressed_matrix, N, N);
  std::vector<std::map<unsigned int, T> > stl_matrix(N * N);
  viennacl::copy(vcl_compressed_matrix, stl_matrix);
  viennacl::coordinate_matrix<T> vcl_coordinate_matrix;
  viennacl::copy(stl_matrix, vcl_coordinate_matrix);
  std::vector<std::map<unsigned int, T> > stl_back(N * N);
  viennacl::copy(vcl_coordinate_matrix, stl_back);
  CHECK(stl_back == stl_matrix, "coordinate_matrix round trip");

  std::vector<T> rhs(N * N);
  for (std::size_t i = 0; i < rhs.size(); ++i) rhs[i] = T(1) + T(i % 11) / T(10);
  viennacl::vector<T> vcl_rhs(rhs.size()), vcl_result(rhs.size());
  viennacl::copy(rhs.begin(), rhs.end(), vcl_rhs.begin());

  viennacl::linalg::jacobi_precond<viennacl::compressed_matrix<T> > vcl_jacobi(vcl_compressed_matrix, viennacl::linalg::jacobi_tag());
  vcl_result = viennacl::linalg::solve(vcl_compressed_matrix, vcl_rhs, viennacl::linalg::cg_tag());
  vcl_result = viennacl::linalg::solve(vcl_compressed_matrix, vcl_rhs, viennacl::linalg::cg_tag(1e-6, 200), vcl_jacobi);
  std::vector<T> result(rhs.size());
  viennacl::copy(vcl_result.begin(), vcl_result.end(), result.begin());
  std::vector<T> Ax = viennacl::linalg::prod(stl_matrix, result);
  double num = 0, den = 0;
  for (std::size_t i = 0; i < rhs.size(); ++i) {
    num += double(rhs[i] - Ax[i]) * double(rhs[i] - Ax[i]);
    den += double(rhs[i]) * double(rhs[i]);
  }
  // the tutorial's tags ask for 1e-6 (Jacobi: the M^-1-norm, the 2-norm ratio for this constant diagonal)
  const double tut_check = std::max<double>(1.05e-6, tol_check);
  CHECK(std::sqrt(num / den) <= tut_check, "tutorial Jacobi CG: residual %g", std::sqrt(num / den));
  vcl_result = viennacl::linalg::solve(vcl_coordinate_matrix, vcl_rhs, viennacl::linalg::bicgstab_tag(1e-6, 200));
  viennacl::copy(vcl_result.begin(), vcl_result.end(), result.begin());
  Ax = viennacl::linalg::prod(stl_matrix, result);
  num = 0;
  for (std::size_t i = 0; i < rhs.size(); ++i) num += double(rhs[i] - Ax[i]) * double(rhs[i] - Ax[i]);
  CHECK(std::sqrt(num / den) <= tut_check, "tutorial BiCGStab (COO): residual %g", std::sqrt(num / den));
}

template <typename T>
void run_all(T eps, T tol_solver, T tol_check)
{
  test_iterator_copies<T>(eps);
  test_vector_proxies<T>(eps);
  strided_product<T, viennacl::compressed_matrix<T> >("compressed_matrix", eps);
  strided_product<T, viennacl::coordinate_matrix<T> >("coordinate_matrix", eps);
  strided_product<T, viennacl::ell_matrix<T> >("ell_matrix", eps);
  strided_product<T, viennacl::sliced_ell_matrix<T> >("sliced_ell_matrix", eps);
  strided_product<T, viennacl::hyb_matrix<T> >("hyb_matrix", eps);
  std::mt19937_64 g(7);
  const unsigned n = 3000;
  std::vector<std::map<unsigned int, T> > std_matrix = random_stl<T>(n, g);
  std::vector<T> rhs(n);
  std::uniform_real_distribution<double> u(0, 1);
  for (unsigned i = 0; i < n; ++i) rhs[i] = T(1) + T(u(g));
  products_against_stl<T, viennacl::compressed_matrix<T> >("compressed_matrix", std_matrix, rhs, eps);
  products_against_stl<T, viennacl::coordinate_matrix<T> >("coordinate_matrix", std_matrix, rhs, eps);
  products_against_stl<T, viennacl::sliced_ell_matrix<T> >("sliced_ell_matrix", std_matrix, rhs, eps);
  products_against_stl<T, viennacl::hyb_matrix<T> >("hyb_matrix", std_matrix, rhs, eps);
  test_resize<T>(eps);
  test_matrix_free<T>(tol_solver, tol_check);
  test_iterative_tutorial<T>(tol_check);
}

int main()
{
  try {
    run_all<double>(1e-12, 1e-10, 1e-9);  // tests/src/sparse.cpp: epsilon 1e-12 for double
    run_all<float>(1e-4f, 1e-5f, 5e-4f);  // ... and 1e-4 for float
  } catch (std::exception const& e) {
    std::printf("FAIL exception: %s\n", e.what());
    ++failures;
  }
  if (failures) {
    std::printf("%d check(s) FAILED\n", failures);
    return 1;
  }
  std::printf("ALL PASSED\n");
  return 0;
}
