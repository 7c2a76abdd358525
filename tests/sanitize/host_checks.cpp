// host_checks.cpp -- the host-only code of libvclhip (generators, CSR row-block plans, SELL / COO /
// ELL / HYB conversions, the MatrixMarket reader and writer) and the CPU oracle, run under
// AddressSanitizer and UndefinedBehaviorSanitizer (tests/sanitize/Makefile builds this file with
// those sources and -fsanitize=address,undefined).  Every check also asserts the invariants of its
// output, so a wrong size or an index out of range fails here even where it stays inside a buffer.
// The MatrixMarket reader parses files it cannot trust: a deterministic set of mutated files
// (bytes replaced, inserted, deleted, lines cut) must each be read or refused cleanly.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include <unistd.h>

#include "vclhip.h"
#include "vcl_oracle.h"

namespace {

int g_failures = 0;

#define CHECK(cond)                                                          \
  do {                                                                       \
    if (!(cond)) {                                                           \
      std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #cond); \
      ++g_failures;                                                          \
    }                                                                        \
  } while (0)

#define OK(expr) CHECK((expr) == VCLHIP_SUCCESS)

struct Csr {
  uint32_t n = 0, n_cols = 0;
  std::vector<uint32_t> rp, col;
  std::vector<double> val;
  uint32_t nnz() const { return rp.empty() ? 0 : rp[n]; }
};

bool valid_csr(const Csr& A)
{
  if (A.rp.size() != A.n + 1 || A.rp[0] != 0) return false;
  for (uint32_t r = 0; r < A.n; ++r) {
    if (A.rp[r + 1] < A.rp[r]) return false;
    for (uint32_t k = A.rp[r]; k < A.rp[r + 1]; ++k) {
      if (A.col[k] >= A.n_cols) return false;
      if (k > A.rp[r] && A.col[k] <= A.col[k - 1]) return false;
    }
  }
  return A.col.size() == A.nnz() && A.val.size() == A.nnz();
}

Csr laplace2d(uint32_t px, uint32_t py)
{
  Csr A;
  A.n = A.n_cols = px * py;
  A.rp.resize(A.n + 1);
  A.col.resize(5ull * A.n);
  A.val.resize(5ull * A.n);
  uint32_t nnz = 0;
  OK(vclhip_gen_fdm_laplace_2d(px, py, A.rp.data(), A.col.data(), A.val.data(), &nnz));
  A.col.resize(nnz);
  A.val.resize(nnz);
  return A;
}

Csr laplace3d(uint32_t nx, uint32_t ny, uint32_t nz)
{
  Csr A;
  A.n = A.n_cols = nx * ny * nz;
  A.rp.resize(A.n + 1);
  A.col.resize(7ull * A.n);
  A.val.resize(7ull * A.n);
  uint32_t nnz = 0;
  OK(vclhip_gen_laplace_3d(nx, ny, nz, A.rp.data(), A.col.data(), A.val.data(), &nnz));
  A.col.resize(nnz);
  A.val.resize(nnz);
  return A;
}

Csr diffusion(uint32_t nx, uint32_t ny, uint32_t nz, uint64_t seed)
{
  Csr A;
  A.n = A.n_cols = nx * ny * nz;
  std::vector<double> g(A.n), k(A.n);
  OK(vclhip_random_normal(seed, A.n, 0.0, 1.0, g.data()));
  for (uint32_t i = 0; i < A.n; ++i) k[i] = std::exp(g[i]);
  A.rp.resize(A.n + 1);
  A.col.resize(7ull * A.n);
  A.val.resize(7ull * A.n);
  uint32_t nnz = 0;
  OK(vclhip_gen_diffusion(nx, ny, nz, k.data(), A.rp.data(), A.col.data(), A.val.data(), &nnz));
  A.col.resize(nnz);
  A.val.resize(nnz);
  return A;
}

Csr powerlaw(uint32_t n, uint32_t max_len, uint64_t seed)
{
  Csr A;
  A.n = A.n_cols = n;
  uint32_t nnz = 0;
  OK(vclhip_gen_powerlaw(n, max_len, 2.0, 1.1, seed, nullptr, nullptr, nullptr, &nnz));
  A.rp.resize(n + 1);
  A.col.resize(nnz);
  A.val.resize(nnz);
  uint32_t nnz2 = 0;
  OK(vclhip_gen_powerlaw(n, max_len, 2.0, 1.1, seed, A.rp.data(), A.col.data(), A.val.data(), &nnz2));
  CHECK(nnz2 == nnz);
  return A;
}

// Rows of random lengths, some empty, one longer than a row block's staging capacity.
Csr ragged(uint32_t n, uint64_t seed)
{
  std::mt19937_64 g(seed);
  Csr A;
  A.n = A.n_cols = n;
  A.rp.assign(n + 1, 0);
  for (uint32_t r = 0; r < n; ++r) {
    uint32_t len = (g() % 4 == 0) ? 0 : static_cast<uint32_t>(g() % 40);
    if (r == n / 2) len = std::min<uint32_t>(n, 5000);
    if (len > n) len = n;
    std::vector<uint32_t> cs;
    for (uint32_t c = 0; c < n && cs.size() < len; ++c)
      if (g() % n < len || n - c <= len - cs.size()) cs.push_back(c);
    A.rp[r + 1] = A.rp[r] + static_cast<uint32_t>(cs.size());
    for (uint32_t c : cs) {
      A.col.push_back(c);
      A.val.push_back(static_cast<double>(g() % 1000) / 100.0 - 5.0);
    }
  }
  return A;
}

void check_plans(const Csr& A)
{
  std::vector<uint32_t> blocks(A.n + 1);
  uint32_t nb = 0;
  OK(vclhip_csr_plan_blocks(A.rp.data(), A.n, blocks.data(), &nb));
  CHECK(nb <= A.n + (A.n == 0 ? 1 : 0));
  std::vector<uint32_t> rb(A.n + 2);
  uint32_t nrb = 0;
  OK(vclhip_csr_reference_row_blocks(A.rp.data(), A.n, rb.data(), &nrb));
  CHECK(nrb <= A.n + 1);
  if (A.n) {
    std::vector<uint32_t> orb(A.n + 2);
    const uint32_t onb = orc_row_blocks(A.n, A.rp.data(), orb.data());
    CHECK(onb == nrb);
  }
}

void check_sell(const Csr& A, uint32_t C)
{
  uint32_t ns = 0;
  uint64_t stored = 0;
  if (A.n == 0) {  // an empty matrix has no SELL form: refused
    CHECK(vclhip_sell_from_csr_f64(0, C, A.rp.data(), nullptr, nullptr, &ns, &stored, nullptr, nullptr, nullptr,
                                   nullptr) == VCLHIP_ERROR_INVALID_VALUE);
    return;
  }
  OK(vclhip_sell_from_csr_f64(A.n, C, A.rp.data(), A.col.data(), A.val.data(), &ns, &stored, nullptr, nullptr,
                              nullptr, nullptr));
  std::vector<uint32_t> cpb(ns), bs(ns), ci(stored);
  std::vector<double> el(stored);
  OK(vclhip_sell_from_csr_f64(A.n, C, A.rp.data(), A.col.data(), A.val.data(), &ns, &stored, cpb.data(), ci.data(),
                              bs.data(), el.data()));
  CHECK(ns == (A.n + C - 1) / C);
  for (uint32_t s = 0; s < ns; ++s) CHECK(static_cast<uint64_t>(bs[s]) + uint64_t(cpb[s]) * C <= stored);
  // the oracle's SELL product equals its CSR product
  std::vector<double> x(A.n), y1(A.n, 0.0), y2(A.n, 0.0);
  OK(vclhip_random_uniform(7, A.n, -1, 1, x.data()));
  orc_csr_spmv_f64(A.n, A.rp.data(), A.col.data(), A.val.data(), x.data(), 0, 1, 1.0, y1.data(), 0, 1, 0.0);
  orc_sell_spmv_f64(A.n, C, cpb.data(), ci.data(), bs.data(), el.data(), x.data(), 0, 1, 1.0, y2.data(), 0, 1, 0.0);
  for (uint32_t i = 0; i < A.n; ++i) CHECK(std::fabs(y1[i] - y2[i]) <= 1e-12 * (1 + std::fabs(y1[i])));
}

void check_formats(const Csr& A)
{
  if (A.n == 0) {  // the conversions refuse an empty matrix
    uint32_t w = 0, ew = 0, cn = 0, nb = 0;
    std::vector<uint32_t> rp(1), blocks(1);
    CHECK(vclhip_coo_plan(nullptr, 0, 0, rp.data(), blocks.data(), &nb) == VCLHIP_ERROR_INVALID_VALUE);
    CHECK(vclhip_ell_from_csr_f64(0, 32, A.rp.data(), nullptr, nullptr, &w, nullptr, nullptr) ==
          VCLHIP_ERROR_INVALID_VALUE);
    CHECK(vclhip_hyb_from_csr_f64(0, 0.5, 32, A.rp.data(), nullptr, nullptr, &ew, &cn, nullptr, nullptr, nullptr,
                                  nullptr, nullptr) == VCLHIP_ERROR_INVALID_VALUE);
    return;
  }
  // COO from the CSR pattern
  std::vector<uint32_t> coords(2ull * A.nnz());
  for (uint32_t r = 0; r < A.n; ++r)
    for (uint32_t k = A.rp[r]; k < A.rp[r + 1]; ++k) {
      coords[2ull * k] = r;
      coords[2ull * k + 1] = A.col[k];
    }
  std::vector<uint32_t> rp(A.n + 1), blocks(A.n + 1);
  uint32_t nb = 0;
  OK(vclhip_coo_plan(coords.data(), A.nnz(), A.n, rp.data(), blocks.data(), &nb));
  CHECK(rp == A.rp);
  for (uint32_t align : {1u, 32u, 64u}) {
    uint32_t w = 0;
    OK(vclhip_ell_from_csr_f64(A.n, align, A.rp.data(), A.col.data(), A.val.data(), &w, nullptr, nullptr));
    const size_t ld = (A.n + align - 1) / align * align;
    std::vector<uint32_t> ec(ld * w);
    std::vector<double> ev(ld * w);
    OK(vclhip_ell_from_csr_f64(A.n, align, A.rp.data(), A.col.data(), A.val.data(), &w, ec.data(), ev.data()));
    for (uint32_t c : ec) CHECK(c < std::max<uint32_t>(A.n, 1));
  }
  for (double thr : {0.0, 0.5, 0.8, 1.0}) {
    uint32_t ew = 0, cn = 0;
    OK(vclhip_hyb_from_csr_f64(A.n, thr, 32, A.rp.data(), A.col.data(), A.val.data(), &ew, &cn, nullptr, nullptr,
                               nullptr, nullptr, nullptr));
    const size_t ld = (A.n + 31) / 32 * 32;
    std::vector<uint32_t> ec(ld * ew), cr(A.n + 1), cc(cn);
    std::vector<double> ev(ld * ew), cv(cn);
    OK(vclhip_hyb_from_csr_f64(A.n, thr, 32, A.rp.data(), A.col.data(), A.val.data(), &ew, &cn, ec.data(), ev.data(),
                               cr.data(), cc.data(), cv.data()));
    CHECK(cr[A.n] == cn || (cr[A.n] == 0 && cn == 1));  // an empty CSR part stores one zero
  }
}

std::string tmp_path(const char* name)
{
  const char* d = std::getenv("TMPDIR");
  return std::string(d && *d ? d : "/tmp") + "/vclhip_host_checks_" + std::to_string(::getpid()) + "_" + name;
}

bool read_mm(const std::string& path, Csr* out, int64_t base = 1)
{
  vclhip_mm_matrix* m = nullptr;
  uint32_t nr = 0, nc = 0, nz = 0;
  int sym = 0;
  int64_t lines = 0;
  if (vclhip_mm_read(path.c_str(), base, &m, &nr, &nc, &nz, &sym, &lines) != VCLHIP_SUCCESS) {
    CHECK(m == nullptr);
    CHECK(std::strlen(vclhip_last_error()) > 0);
    return false;
  }
  out->n = nr;
  out->n_cols = nc;
  out->rp.assign(nr + 1, 0);
  out->col.assign(nz, 0);
  out->val.assign(nz, 0.0);
  OK(vclhip_mm_csr(m, out->rp.data(), out->col.data(), out->val.data()));
  OK(vclhip_mm_free(m));
  return true;
}

void check_mm_round_trip(const Csr& A)
{
  const std::string f = tmp_path("rt.mtx");
  OK(vclhip_mm_write(f.c_str(), 1, A.n, A.n_cols, A.rp.data(), A.col.data(), A.val.data()));
  Csr B;
  CHECK(read_mm(f, &B));
  CHECK(valid_csr(B) && B.rp == A.rp && B.col == A.col);
  for (size_t k = 0; k < B.val.size(); ++k) CHECK(std::fabs(B.val[k] - A.val[k]) <= 1e-5 * (1 + std::fabs(A.val[k])));
  std::remove(f.c_str());
}

void fuzz_mm(int rounds)
{
  const char* seeds[] = {
      "%%MatrixMarket matrix coordinate real general\n% c\n\n 4 3 5\n1 1 1.5\n4 3 -2\n1 1 7.25\n2 2 3e-3\n3 1 1\n",
      "%%MatrixMarket matrix coordinate real symmetric\n4 4 4\n1 1 2\n2 1 -1\n3 2 -1.5\n4 4 3\n",
      "%%MatrixMarket matrix coordinate pattern general\n3 3 3\n2 1\n3 3\n1 3\n",
      "%%MatrixMarket matrix coordinate complex general\n2 2 2\n1 2 3.5 -1\n2 1 -0.5 2\n",
      "3 3 2\n1 1 4\n3 2 1\n",
  };
  const char alphabet[] = "0123456789 -+.eE%\n\t\rMmatrixcoordinatesymmetricpattern";
  std::mt19937_64 g(12345);
  const std::string f = tmp_path("fuzz.mtx");
  int accepted = 0;
  for (int it = 0; it < rounds; ++it) {
    std::string s = seeds[it % (sizeof(seeds) / sizeof(seeds[0]))];
    const int edits = 1 + static_cast<int>(g() % 4);
    for (int e = 0; e < edits && !s.empty(); ++e) {
      const size_t pos = g() % s.size();
      const char ch = alphabet[g() % (sizeof(alphabet) - 1)];
      switch (g() % 5) {
        case 0: s[pos] = ch; break;
        case 1: s.insert(s.begin() + pos, ch); break;
        case 2: s.erase(pos, 1 + g() % 3); break;
        case 3: s.resize(pos); break;
        default: s.insert(pos, std::to_string(static_cast<long long>(g() % 100000) - 50000)); break;
      }
    }
    if (g() % 8 == 0) s += std::string(5000 + g() % 80000, ' ') + "1\n";  // a line longer than any buffer
    FILE* fp = std::fopen(f.c_str(), "wb");
    CHECK(fp != nullptr);
    if (!fp) return;
    std::fwrite(s.data(), 1, s.size(), fp);
    std::fclose(fp);
    Csr A;
    if (read_mm(f, &A, (it & 1) ? 1 : 0)) {
      ++accepted;
      CHECK(valid_csr(A));
    }
  }
  std::remove(f.c_str());
  std::printf("MatrixMarket: %d mutated files, %d read, the rest refused\n", rounds, accepted);
  CHECK(accepted > 0 && accepted < rounds);
}

void check_oracle_solves(const Csr& A)
{
  const uint32_t n = A.n;
  std::vector<double> b(n), x(n, 0.0), hist(4000), diag(n);
  OK(vclhip_random_uniform(0, n, -1, 1, b.data()));
  orc_csr64 d = {n, A.rp.data(), A.col.data(), A.val.data()};
  orc_result r = orc_cg_pipelined_csr(&d, b.data(), x.data(), 1e-10, 0.0, 4000, hist.data());
  CHECK(r.iters > 0 && r.error < 1e-9);
  orc_csr_row_info_f64(n, A.rp.data(), A.col.data(), A.val.data(), diag.data(), 3);
  std::fill(x.begin(), x.end(), 0.0);
  r = orc_cg_classic_csr(&d, b.data(), x.data(), diag.data(), 1e-10, 0.0, 4000, hist.data());
  CHECK(r.error < 1e-9);
  std::fill(x.begin(), x.end(), 0.0);
  r = orc_bicgstab_pipelined_csr(&d, b.data(), x.data(), 1e-10, 0.0, 4000, hist.data());
  CHECK(r.error < 1e-9);
  std::fill(x.begin(), x.end(), 0.0);
  r = orc_bicgstab_classic_csr(&d, b.data(), x.data(), diag.data(), 1e-10, 0.0, 4000, 200, hist.data());
  CHECK(r.error < 1e-9);
  std::fill(x.begin(), x.end(), 0.0);
  r = orc_gmres_pipelined_csr(&d, b.data(), x.data(), 1e-10, 0.0, 4000, 20, 1, hist.data());
  CHECK(r.error < 1e-9);
  std::fill(x.begin(), x.end(), 0.0);
  uint32_t restarts = 0;
  r = orc_mixed_cg_csr(&d, b.data(), x.data(), 1e-10, 4000, 1e-2f, &restarts);
  CHECK(r.error < 1e-9);
  std::fill(x.begin(), x.end(), 0.0);
  uint32_t updates = 0;
  r = orc_mixed_cg2_csr(&d, b.data(), x.data(), diag.data(), 1e-10, 4000, 1e-2f, 1, 1, &updates, hist.data());
  CHECK(r.error < 1e-9);
  std::vector<double> lam(n), dA(A.nnz());
  OK(vclhip_random_uniform(3, n, -1, 1, lam.data()));
  orc_adjoint_grad_csr(n, A.rp.data(), A.col.data(), lam.data(), x.data(), dA.data());
  // SELL-64 solve of the same system
  std::vector<uint32_t> cpb, ci, bs;
  std::vector<double> el;
  uint32_t ns = 0;
  uint64_t stored = 0;
  orc_sell_sizes(n, 64, A.rp.data(), &ns, &stored);
  cpb.resize(ns);
  bs.resize(ns);
  ci.resize(stored);
  el.resize(stored);
  orc_sell_from_csr_f64(n, 64, A.rp.data(), A.col.data(), A.val.data(), cpb.data(), ci.data(), bs.data(), el.data());
  orc_sell64 s = {n, 64, cpb.data(), ci.data(), bs.data(), el.data()};
  std::fill(x.begin(), x.end(), 0.0);
  r = orc_cg_pipelined_sell(&s, b.data(), x.data(), 1e-10, 0.0, 4000, hist.data());
  CHECK(r.error < 1e-9);
}

void check_batch()
{
  uint32_t rows = 0, nnz = 0;
  OK(vclhip_gen_graph_laplacian_batch(40, 64, 300, 8.0, 0.1, 1.0, 0.01, 9, nullptr, nullptr, nullptr, nullptr, &rows,
                                      &nnz));
  Csr A;
  A.n = A.n_cols = rows;
  std::vector<uint32_t> sys(41);
  A.rp.resize(rows + 1);
  A.col.resize(nnz);
  A.val.resize(nnz);
  OK(vclhip_gen_graph_laplacian_batch(40, 64, 300, 8.0, 0.1, 1.0, 0.01, 9, sys.data(), A.rp.data(), A.col.data(),
                                      A.val.data(), &rows, &nnz));
  CHECK(valid_csr(A) && sys[40] == rows);
  std::vector<double> b(rows), x(rows, 0.0), err(40);
  std::vector<uint32_t> its(40);
  OK(vclhip_random_uniform(1, rows, -1, 1, b.data()));
  orc_csr64 d = {rows, A.rp.data(), A.col.data(), A.val.data()};
  orc_batched_cg_csr(40, sys.data(), &d, b.data(), x.data(), 1e-8, 5000, its.data(), err.data());
  for (double e : err) CHECK(e < 1e-7);
  std::fill(x.begin(), x.end(), 0.0);
  orc_batched_cg2_csr(40, sys.data(), &d, b.data(), x.data(), 1e-5, 5000, 1, its.data(), err.data());
  check_plans(A);
}

void check_errors()
{
  uint32_t nnz = 0, nb = 0;
  std::vector<uint32_t> bad_rp = {0, 3, 2, 4};
  std::vector<uint32_t> blocks(4);
  CHECK(vclhip_csr_plan_blocks(bad_rp.data(), 3, blocks.data(), &nb) == VCLHIP_ERROR_INVALID_VALUE);
  CHECK(vclhip_gen_powerlaw(1, 4, 2.0, 1.1, 0, nullptr, nullptr, nullptr, &nnz) == VCLHIP_ERROR_INVALID_VALUE);
  CHECK(vclhip_csr_plan_blocks(nullptr, 3, blocks.data(), &nb) == VCLHIP_ERROR_INVALID_VALUE);
  vclhip_mm_matrix* m = nullptr;
  CHECK(vclhip_mm_read(tmp_path("absent.mtx").c_str(), 1, &m, nullptr, nullptr, nullptr, nullptr, nullptr) != 0);
  CHECK(m == nullptr);
}

}  // namespace

int main()
{
  orc_set_num_threads(2);
  std::vector<Csr> ms;
  ms.push_back(laplace2d(1, 1));
  ms.push_back(laplace2d(7, 5));
  ms.push_back(laplace2d(30, 20));
  ms.push_back(laplace3d(1, 1, 1));
  ms.push_back(laplace3d(6, 5, 4));
  ms.push_back(diffusion(9, 8, 1, 2));
  ms.push_back(diffusion(7, 6, 5, 3));
  ms.push_back(powerlaw(2, 1, 1));
  ms.push_back(powerlaw(3000, 512, 4));
  ms.push_back(ragged(0, 5));
  ms.push_back(ragged(1, 5));
  ms.push_back(ragged(777, 6));
  ms.push_back(ragged(6000, 7));
  for (const Csr& A : ms) {
    CHECK(valid_csr(A));
    check_plans(A);
    for (uint32_t C : {1u, 2u, 3u, 32u, 64u, 256u}) check_sell(A, C);
    check_formats(A);
    check_mm_round_trip(A);
  }
  check_oracle_solves(laplace2d(30, 20));
  check_oracle_solves(diffusion(7, 6, 5, 3));
  check_batch();
  check_errors();
  fuzz_mm(3000);
  if (g_failures) {
    std::fprintf(stderr, "%d checks failed\n", g_failures);
    return 1;
  }
  std::printf("host checks passed under the sanitizers\n");
  return 0;
}
