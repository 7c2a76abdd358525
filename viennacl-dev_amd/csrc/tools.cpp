// tools.cpp -- host-side tools of libvclhip: the project RNG, the matrix generators used by the
// tests and benchmarks, the CSR row-block planner and the CSR -> SELL-C converter.
// None of this touches the GPU.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <unordered_set>
#include <vector>

#include "errors.hpp"
#include "plan.hpp"

namespace {

// Project RNG: mt19937_64, 53-bit uniforms in [0,1).
struct rng {
  std::mt19937_64 eng;
  explicit rng(uint64_t seed) : eng(seed) {}
  double uniform() { return static_cast<double>(eng() >> 11) * (1.0 / 9007199254740992.0); }
  // Box-Muller on a pair of uniforms; returns two normals.
  void normal_pair(double& z0, double& z1)
  {
    const double two_pi = 6.283185307179586476925286766559;
    double u1 = uniform();
    double u2 = uniform();
    double r = std::sqrt(-2.0 * std::log(1.0 - u1));
    z0 = r * std::cos(two_pi * u2);
    z1 = r * std::sin(two_pi * u2);
  }
};

// Streams normals one at a time from consecutive Box-Muller pairs.
struct normal_stream {
  rng& g;
  bool have = false;
  double spare = 0;
  explicit normal_stream(rng& gen) : g(gen) {}
  double next()
  {
    if (have) { have = false; return spare; }
    double z0, z1;
    g.normal_pair(z0, z1);
    spare = z1;
    have = true;
    return z0;
  }
};

}  // namespace

extern "C" {

int vclhip_random_uniform(uint64_t seed, uint64_t n, double a, double b, double* out)
{
  if (n && !out) return vclhip::invalid("out is NULL");
  rng g(seed);
  for (uint64_t i = 0; i < n; ++i) out[i] = a + (b - a) * g.uniform();
  return VCLHIP_SUCCESS;
}

int vclhip_random_normal(uint64_t seed, uint64_t n, double mean, double sigma, double* out)
{
  if (n && !out) return vclhip::invalid("out is NULL");
  rng g(seed);
  normal_stream ns(g);
  for (uint64_t i = 0; i < n; ++i) out[i] = mean + sigma * ns.next();
  return VCLHIP_SUCCESS;
}

// tools/matrix_generation.hpp:47-86 (same entries; columns ascending as a std::map row yields them)
int vclhip_gen_fdm_laplace_2d(uint32_t px, uint32_t py, uint32_t* row_ptr, uint32_t* col, double* val,
                              uint32_t* nnz)
{
  if (!row_ptr || !col || !val || !nnz) return vclhip::invalid("null output");
  uint64_t n = static_cast<uint64_t>(px) * py;
  if (n == 0 || n * 5 > 0xFFFFFFFFull) return vclhip::invalid("grid size");
  uint32_t k = 0;
  row_ptr[0] = 0;
  for (uint32_t j = 0; j < py; ++j) {
    for (uint32_t i = 0; i < px; ++i) {
      uint32_t row = i + j * px;
      if (j > 0) { col[k] = row - px; val[k++] = -1.0; }
      if (i > 0) { col[k] = row - 1; val[k++] = -1.0; }
      col[k] = row; val[k++] = 4.0;
      if (i + 1 < px) { col[k] = row + 1; val[k++] = -1.0; }
      if (j + 1 < py) { col[k] = row + px; val[k++] = -1.0; }
      row_ptr[row + 1] = k;
    }
  }
  *nnz = k;
  return VCLHIP_SUCCESS;
}

int vclhip_gen_laplace_3d(uint32_t nx, uint32_t ny, uint32_t nz, uint32_t* row_ptr, uint32_t* col, double* val,
                          uint32_t* nnz)
{
  if (!row_ptr || !col || !val || !nnz) return vclhip::invalid("null output");
  uint64_t n = static_cast<uint64_t>(nx) * ny * nz;
  if (n == 0 || n * 7 > 0xFFFFFFFFull) return vclhip::invalid("grid size");
  const uint32_t sy = nx, sz = nx * ny;
  uint32_t k = 0;
  row_ptr[0] = 0;
  for (uint32_t z = 0; z < nz; ++z)
    for (uint32_t y = 0; y < ny; ++y)
      for (uint32_t x = 0; x < nx; ++x) {
        uint32_t row = x + sy * y + sz * z;
        if (z > 0) { col[k] = row - sz; val[k++] = -1.0; }
        if (y > 0) { col[k] = row - sy; val[k++] = -1.0; }
        if (x > 0) { col[k] = row - 1; val[k++] = -1.0; }
        col[k] = row; val[k++] = 6.0;
        if (x + 1 < nx) { col[k] = row + 1; val[k++] = -1.0; }
        if (y + 1 < ny) { col[k] = row + sy; val[k++] = -1.0; }
        if (z + 1 < nz) { col[k] = row + sz; val[k++] = -1.0; }
        row_ptr[row + 1] = k;
      }
  *nnz = k;
  return VCLHIP_SUCCESS;
}

int vclhip_gen_diffusion(uint32_t nx, uint32_t ny, uint32_t nz, const double* kc, uint32_t* row_ptr, uint32_t* col,
                         double* val, uint32_t* nnz)
{
  if (!kc || !row_ptr || !col || !val || !nnz) return vclhip::invalid("null argument");
  uint64_t n = static_cast<uint64_t>(nx) * ny * nz;
  if (n == 0 || n * 7 > 0xFFFFFFFFull) return vclhip::invalid("grid size");
  const uint32_t sy = nx, sz = nx * ny;
  auto face = [](double a, double b) { return 2.0 * a * b / (a + b); };
  uint32_t k = 0;
  row_ptr[0] = 0;
  for (uint32_t z = 0; z < nz; ++z)
    for (uint32_t y = 0; y < ny; ++y)
      for (uint32_t x = 0; x < nx; ++x) {
        uint32_t row = x + sy * y + sz * z;
        double kr = kc[row];
        // Six faces (four in 2-D): interior faces couple to the neighbour, boundary faces add k to the diagonal.
        double diag = 0.0;
        uint32_t dpos;
        if (nz > 1) {
          if (z > 0) { double f = face(kr, kc[row - sz]); col[k] = row - sz; val[k++] = -f; diag += f; }
          else diag += kr;
        }
        if (y > 0) { double f = face(kr, kc[row - sy]); col[k] = row - sy; val[k++] = -f; diag += f; }
        else diag += kr;
        if (x > 0) { double f = face(kr, kc[row - 1]); col[k] = row - 1; val[k++] = -f; diag += f; }
        else diag += kr;
        dpos = k;
        col[k] = row; val[k++] = 0.0;
        if (x + 1 < nx) { double f = face(kr, kc[row + 1]); col[k] = row + 1; val[k++] = -f; diag += f; }
        else diag += kr;
        if (y + 1 < ny) { double f = face(kr, kc[row + sy]); col[k] = row + sy; val[k++] = -f; diag += f; }
        else diag += kr;
        if (nz > 1) {
          if (z + 1 < nz) { double f = face(kr, kc[row + sz]); col[k] = row + sz; val[k++] = -f; diag += f; }
          else diag += kr;
        }
        val[dpos] = diag;
        row_ptr[row + 1] = k;
      }
  *nnz = k;
  return VCLHIP_SUCCESS;
}

int vclhip_gen_powerlaw(uint32_t n, uint32_t max_len, double alpha, double diag_factor, uint64_t seed,
                        uint32_t* row_ptr, uint32_t* col, double* val, uint32_t* nnz)
{
  if (!nnz || n < 2 || max_len == 0) return vclhip::invalid("powerlaw arguments");
  if (max_len >= n) max_len = n - 1;
  std::vector<double> cdf(max_len);
  double acc = 0;
  for (uint32_t k = 1; k <= max_len; ++k) { acc += std::pow(static_cast<double>(k), -alpha); cdf[k - 1] = acc; }
  for (auto& c : cdf) c /= acc;
  const bool fill = row_ptr && col && val;
  rng g(seed);
  normal_stream ns(g);
  uint64_t total = 0;
  std::vector<uint32_t> cols;
  std::vector<double> vals;
  std::unordered_set<uint32_t> seen;
  if (fill) row_ptr[0] = 0;
  for (uint32_t row = 0; row < n; ++row) {
    double u = g.uniform();
    uint32_t L = static_cast<uint32_t>(std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) + 1;
    if (L > max_len) L = max_len;
    cols.clear();
    seen.clear();
    while (cols.size() < L) {
      uint32_t c = static_cast<uint32_t>(g.uniform() * n);
      if (c >= n) c = n - 1;
      if (c == row || !seen.insert(c).second) continue;
      cols.push_back(c);
    }
    vals.resize(L);
    double rowsum = 0;
    for (uint32_t t = 0; t < L; ++t) { vals[t] = ns.next(); rowsum += std::fabs(vals[t]); }
    if (fill) {
      // sorted columns with the diagonal inserted in place
      std::vector<std::pair<uint32_t, double> > ent(L);
      for (uint32_t t = 0; t < L; ++t) ent[t] = std::make_pair(cols[t], vals[t]);
      ent.push_back(std::make_pair(row, diag_factor * rowsum));
      std::sort(ent.begin(), ent.end());
      for (size_t t = 0; t < ent.size(); ++t) { col[total + t] = ent[t].first; val[total + t] = ent[t].second; }
    }
    total += L + 1;
    if (total > 0xFFFFFFFFull) return vclhip::invalid("nnz exceeds 32-bit index range");
    if (fill) row_ptr[row + 1] = static_cast<uint32_t>(total);
  }
  *nnz = static_cast<uint32_t>(total);
  return VCLHIP_SUCCESS;
}

int vclhip_gen_graph_laplacian_batch(uint32_t num_systems, uint32_t n_min, uint32_t n_max, double avg_degree,
                                     double w_lo, double w_hi, double shift, uint64_t seed, uint32_t* sys_ptr,
                                     uint32_t* row_ptr, uint32_t* col, double* val, uint32_t* total_rows,
                                     uint32_t* nnz)
{
  if (!total_rows || !nnz || num_systems == 0 || n_min < 2 || n_max < n_min) return vclhip::invalid("batch arguments");
  const bool fill = sys_ptr && row_ptr && col && val;
  rng g(seed);
  uint64_t rows = 0, entries = 0;
  if (fill) { sys_ptr[0] = 0; row_ptr[0] = 0; }
  std::vector<std::vector<std::pair<uint32_t, double> > > adj;
  for (uint32_t s = 0; s < num_systems; ++s) {
    uint32_t ni = n_min + static_cast<uint32_t>(g.uniform() * (n_max - n_min + 1));
    if (ni > n_max) ni = n_max;
    uint64_t edges = static_cast<uint64_t>(std::llround(ni * avg_degree / 2.0));
    adj.assign(ni, std::vector<std::pair<uint32_t, double> >());
    for (uint64_t e = 0; e < edges; ++e) {
      uint32_t a, b;
      do {
        a = static_cast<uint32_t>(g.uniform() * ni);
        b = static_cast<uint32_t>(g.uniform() * ni);
        if (a >= ni) a = ni - 1;
        if (b >= ni) b = ni - 1;
      } while (a == b);
      double w = w_lo + (w_hi - w_lo) * g.uniform();
      adj[a].push_back(std::make_pair(b, w));
      adj[b].push_back(std::make_pair(a, w));
    }
    for (uint32_t r = 0; r < ni; ++r) {
      auto& lst = adj[r];
      std::sort(lst.begin(), lst.end());
      // merge repeated edges, insert the diagonal
      std::vector<std::pair<uint32_t, double> > merged;
      double deg = 0;
      for (size_t t = 0; t < lst.size(); ++t) {
        deg += lst[t].second;
        if (!merged.empty() && merged.back().first == lst[t].first) merged.back().second += lst[t].second;
        else merged.push_back(lst[t]);
      }
      std::vector<std::pair<uint32_t, double> > ent;
      ent.reserve(merged.size() + 1);
      bool placed = false;
      for (auto& m : merged) {
        if (!placed && m.first > r) { ent.push_back(std::make_pair(r, deg + shift)); placed = true; }
        ent.push_back(std::make_pair(m.first, -m.second));
      }
      if (!placed) ent.push_back(std::make_pair(r, deg + shift));
      if (fill) {
        for (size_t t = 0; t < ent.size(); ++t) {
          col[entries + t] = static_cast<uint32_t>(rows + ent[t].first);
          val[entries + t] = ent[t].second;
        }
      }
      entries += ent.size();
      if (entries > 0xFFFFFFFFull) return vclhip::invalid("nnz exceeds 32-bit index range");
      if (fill) row_ptr[rows + r + 1] = static_cast<uint32_t>(entries);
    }
    rows += ni;
    if (rows >= (1u << 28)) return vclhip::invalid("too many rows");
    if (fill) sys_ptr[s + 1] = static_cast<uint32_t>(rows);
  }
  *total_rows = static_cast<uint32_t>(rows);
  *nnz = static_cast<uint32_t>(entries);
  return VCLHIP_SUCCESS;
}

int vclhip_csr_plan_blocks(const uint32_t* row_ptr, uint32_t n_rows, uint32_t* blocks, uint32_t* num_blocks)
{
  if (!row_ptr || !blocks || !num_blocks) return vclhip::invalid("null argument");
  if (n_rows >= vclhip::kPlanRowLimit) return vclhip::invalid("n_rows must be below 2^28");
  for (uint32_t r = 0; r < n_rows; ++r)
    if (row_ptr[r + 1] < row_ptr[r]) {
      vclhip::set_error("invalid argument: row_ptr must be nondecreasing (row %u)", r);
      return VCLHIP_ERROR_INVALID_VALUE;
    }
  *num_blocks = vclhip::plan_csr_blocks(row_ptr, n_rows, blocks);
  return VCLHIP_SUCCESS;
}

// compressed_matrix.hpp:1152-1188
int vclhip_csr_reference_row_blocks(const uint32_t* row_ptr, uint32_t n_rows, uint32_t* row_blocks,
                                    uint32_t* num_blocks)
{
  if (!row_ptr || !row_blocks || !num_blocks) return vclhip::invalid("null argument");
  const uint64_t cap = 1024;
  uint64_t in_batch = 0;
  uint32_t nb = 0;
  row_blocks[0] = 0;
  for (uint64_t i = 0; i < n_rows; ++i) {
    in_batch += static_cast<uint64_t>(row_ptr[i + 1]) - row_ptr[i];
    if (in_batch > cap) {
      if (i - row_blocks[nb] > 0) {
        row_blocks[++nb] = static_cast<uint32_t>(i);
        --i;  // the current row opens the next block
      } else {
        row_blocks[++nb] = static_cast<uint32_t>(i + 1);
      }
      in_batch = 0;
    }
  }
  if (in_batch > 0) row_blocks[++nb] = n_rows;
  *num_blocks = nb;
  return VCLHIP_SUCCESS;
}

// sliced_ell_matrix.hpp:140-214
int vclhip_sell_from_csr_f64(uint32_t n_rows, uint32_t C, const uint32_t* row_ptr, const uint32_t* col,
                             const double* val, uint32_t* num_slices, uint64_t* stored, uint32_t* cpb,
                             uint32_t* ci, uint32_t* bs, double* el)
{
  if (!row_ptr || !num_slices || !stored || C == 0 || n_rows == 0) return vclhip::invalid("sell arguments");
  uint32_t ns = (n_rows + C - 1) / C;
  uint64_t offset = 0;
  const bool fill = cpb && ci && bs && el && col && val;
  for (uint32_t s = 0; s < ns; ++s) {
    uint32_t r_end = std::min<uint64_t>(static_cast<uint64_t>(s + 1) * C, n_rows);
    uint32_t m = 0;
    for (uint32_t r = s * C; r < r_end; ++r) m = std::max(m, row_ptr[r + 1] - row_ptr[r]);
    if (fill) {
      cpb[s] = m;
      bs[s] = static_cast<uint32_t>(offset);
      std::fill(ci + offset, ci + offset + static_cast<uint64_t>(m) * C, 0u);
      std::fill(el + offset, el + offset + static_cast<uint64_t>(m) * C, 0.0);
      for (uint32_t r = s * C; r < r_end; ++r) {
        uint32_t e = 0;
        for (uint32_t i = row_ptr[r]; i < row_ptr[r + 1]; ++i, ++e) {
          uint64_t idx = offset + static_cast<uint64_t>(e) * C + (r - s * C);
          ci[idx] = col[i];
          el[idx] = val[i];
        }
      }
    }
    offset += static_cast<uint64_t>(m) * C;
    if (offset > 0xFFFFFFFFull) return vclhip::invalid("SELL storage exceeds 32-bit offsets");
  }
  *num_slices = ns;
  *stored = offset;
  return VCLHIP_SUCCESS;
}

int vclhip_coo_plan(const uint32_t* coords, uint32_t nnz, uint32_t n_rows, uint32_t* row_ptr, uint32_t* blocks,
                    uint32_t* num_blocks)
{
  if ((nnz && !coords) || !row_ptr || !blocks || !num_blocks || n_rows == 0) return vclhip::invalid("coo plan arguments");
  if (n_rows >= vclhip::kPlanRowLimit) return vclhip::invalid("n_rows must be below 2^28");
  std::fill(row_ptr, row_ptr + n_rows + 1, 0u);
  uint32_t prev = 0;
  for (uint32_t e = 0; e < nnz; ++e) {
    const uint32_t r = coords[2 * static_cast<size_t>(e)];
    if (r >= n_rows) return vclhip::invalid("COO row index out of range");
    if (r < prev) return vclhip::invalid("COO entries must be ordered by row");
    prev = r;
    ++row_ptr[r + 1];
  }
  for (uint32_t r = 0; r < n_rows; ++r) row_ptr[r + 1] += row_ptr[r];
  *num_blocks = vclhip::plan_csr_blocks(row_ptr, n_rows, blocks);
  return VCLHIP_SUCCESS;
}

int vclhip_ell_from_csr_f64(uint32_t n_rows, uint32_t align, const uint32_t* row_ptr, const uint32_t* col,
                            const double* val, uint32_t* internal_maxnnz, uint32_t* coords, double* elements)
{
  if (!row_ptr || !internal_maxnnz || align == 0 || n_rows == 0) return vclhip::invalid("ell arguments");
  uint32_t m = 0;
  for (uint32_t r = 0; r < n_rows; ++r) m = std::max(m, row_ptr[r + 1] - row_ptr[r]);
  const uint64_t ld = (static_cast<uint64_t>(n_rows) + align - 1) / align * align;
  const uint64_t width = (static_cast<uint64_t>(m) + align - 1) / align * align;
  if (ld * width > 0xFFFFFFFFull || ld > 0xFFFFFFFFull) return vclhip::invalid("ELL storage exceeds 32-bit sizes");
  *internal_maxnnz = static_cast<uint32_t>(width);
  if (!coords || !elements) return VCLHIP_SUCCESS;
  if (row_ptr[n_rows] && (!col || !val)) return vclhip::invalid("ell arguments");
  std::fill(coords, coords + ld * width, 0u);
  std::fill(elements, elements + ld * width, 0.0);
  for (uint32_t r = 0; r < n_rows; ++r)
    for (uint32_t i = row_ptr[r], k = 0; i < row_ptr[r + 1]; ++i, ++k) {
      coords[ld * k + r] = col[i];
      elements[ld * k + r] = val[i];
    }
  return VCLHIP_SUCCESS;
}

int vclhip_hyb_from_csr_f64(uint32_t n_rows, double csr_threshold, uint32_t align, const uint32_t* row_ptr,
                            const uint32_t* col, const double* val, uint32_t* internal_ellnnz, uint32_t* csr_nnz,
                            uint32_t* ell_coords, double* ell_elements, uint32_t* csr_rows, uint32_t* csr_cols,
                            double* csr_elements)
{
  if (!row_ptr || !internal_ellnnz || !csr_nnz || align == 0 || n_rows == 0) return vclhip::invalid("hyb arguments");
  // ELL width: the first row length whose cumulative row count reaches csr_threshold * n_rows
  uint32_t m = 0;
  for (uint32_t r = 0; r < n_rows; ++r) m = std::max(m, row_ptr[r + 1] - row_ptr[r]);
  std::vector<uint64_t> hist(static_cast<size_t>(m) + 1, 0);
  for (uint32_t r = 0; r < n_rows; ++r) ++hist[row_ptr[r + 1] - row_ptr[r]];
  uint32_t width = m;
  uint64_t sum = 0;
  for (uint32_t len = 0; len <= m; ++len) {
    sum += hist[len];
    if (static_cast<double>(sum) >= csr_threshold * static_cast<double>(n_rows)) {
      width = len;
      break;
    }
  }
  const uint64_t ld = (static_cast<uint64_t>(n_rows) + align - 1) / align * align;
  const uint64_t iw = (static_cast<uint64_t>(width) + align - 1) / align * align;
  if (ld * iw > 0xFFFFFFFFull || ld > 0xFFFFFFFFull) return vclhip::invalid("HYB storage exceeds 32-bit sizes");
  uint64_t rest = 0;
  for (uint32_t r = 0; r < n_rows; ++r) {
    const uint32_t len = row_ptr[r + 1] - row_ptr[r];
    if (len > width) rest += len - width;
  }
  *internal_ellnnz = static_cast<uint32_t>(iw);
  *csr_nnz = static_cast<uint32_t>(rest ? rest : 1);  // an empty CSR part keeps one zero entry
  // sizes only when the outputs are not given (the ELL arrays may be null when the ELL part is empty)
  if (!csr_rows || !csr_cols || !csr_elements || (iw && (!ell_coords || !ell_elements))) return VCLHIP_SUCCESS;
  if (row_ptr[n_rows] && (!col || !val)) return vclhip::invalid("hyb arguments");
  if (iw) {
    std::fill(ell_coords, ell_coords + ld * iw, 0u);
    std::fill(ell_elements, ell_elements + ld * iw, 0.0);
  }
  uint32_t c = 0;
  for (uint32_t r = 0; r < n_rows; ++r) {
    csr_rows[r] = c;
    for (uint32_t i = row_ptr[r], k = 0; i < row_ptr[r + 1]; ++i, ++k) {
      if (k < width) {
        ell_coords[ld * k + r] = col[i];
        ell_elements[ld * k + r] = val[i];
      } else {
        csr_cols[c] = col[i];
        csr_elements[c] = val[i];
        ++c;
      }
    }
  }
  csr_rows[n_rows] = c;
  if (c == 0) {
    csr_cols[0] = 0;
    csr_elements[0] = 0.0;
  }
  return VCLHIP_SUCCESS;
}

}  // extern "C"
