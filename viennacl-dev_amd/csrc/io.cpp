// io.cpp -- MatrixMarket files (reference io/matrix_market.hpp:102-433) to and from CSR host arrays.
//
// Reading follows the reference reader's rules: an optional "%%MatrixMarket matrix coordinate"
// header with a real, pattern or complex field (pattern entries are 1; of a complex entry the real part is kept)
// and a general or symmetric structure; '%' comment lines and blank lines are skipped, leading
// blanks are ignored (blanks only: a line led by a tab is a data line, as there); the size line gives
// rows, columns and the entry count; reading stops after that many entries, and a file that ends
// before them is an error; indices are shifted by index_base and checked against the size; a
// symmetric file also sets (col, row); an entry given twice keeps the last value (the reference
// assigns mat(row, col) = value).  Two cases differ from the reference on purpose: a size line that
// declares 0 entries gives an empty matrix (the reference reads on and reports a parse error at the
// end of the file), and a file that ends early without a final newline is an error too (the
// reference then returns what it read).  Writing produces the reference writer's output: a general coordinate
// header, the size line, then one "row col value" line per entry in row order with the value in
// the default iostream format (6 significant digits).
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "errors.hpp"

struct vclhip_mm_matrix {
  uint32_t n_rows = 0, n_cols = 0;
  bool symmetric = false;
  std::vector<uint32_t> row_ptr, col;
  std::vector<double> val;
};

namespace {

struct Entry {
  uint32_t row, col;
  uint64_t seq;  // order of assignment: the last one wins
  double val;
};

std::string lower(std::string s)
{
  for (auto& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  return s;
}

// Next whitespace-separated token of `p` (advanced past it); empty at the end of the line.
std::string token(const char*& p)
{
  while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
  const char* b = p;
  while (*p && *p != ' ' && *p != '\t' && *p != '\r' && *p != '\n') ++p;
  return std::string(b, p);
}

bool parse_long(const char*& p, long long& v)
{
  while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
  if (!*p || *p == '\n') return false;
  char* end = nullptr;
  errno = 0;
  v = std::strtoll(p, &end, 10);
  if (end == p || errno) return false;
  p = end;
  return true;
}

bool parse_double(const char*& p, double& v)
{
  while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
  if (!*p || *p == '\n') return false;
  char* end = nullptr;
  v = std::strtod(p, &end);
  if (end == p) return false;
  p = end;
  return true;
}

int fail(long line, const char* what)
{
  vclhip::set_error("MatrixMarket: line %ld: %s", line, what);
  return VCLHIP_ERROR_INVALID_VALUE;
}

}  // namespace

extern "C" {

int vclhip_mm_read(const char* path, int64_t index_base, vclhip_mm_matrix** out, uint32_t* n_rows, uint32_t* n_cols,
                   uint32_t* nnz, int* symmetric, int64_t* lines_read)
{
  if (!path || !out) return vclhip::invalid("null argument");
  *out = nullptr;
  FILE* f = std::fopen(path, "r");
  if (!f) {
    vclhip::set_error("MatrixMarket: cannot open file %s", path);
    return VCLHIP_ERROR_INVALID_VALUE;
  }
  std::vector<char> buf(1 << 16);
  std::string line;
  long linenum = 0;
  bool size_seen = false, pattern = false, sym = false;
  long long rows = 0, cols = 0, count = 0, valid = 0;
  std::vector<Entry> entries;
  int status = VCLHIP_SUCCESS;
  while (status == VCLHIP_SUCCESS) {
    // one line of any length
    line.clear();
    bool got = false;
    while (std::fgets(buf.data(), static_cast<int>(buf.size()), f)) {
      got = true;
      line += buf.data();
      if (!line.empty() && line.back() == '\n') break;
    }
    if (!got) break;
    ++linenum;
    // only blanks are trimmed (io/matrix_market.hpp:44-75): a line that starts with a tab or holds
    // only a tab or a carriage return is a data line, and not one that parses
    const char* p = line.c_str();
    while (*p == ' ') ++p;
    if (*p == '\n' || *p == 0) continue;
    if (*p == '%') {
      if (p[1] != '%') continue;
      const char* q = p + 2;
      if (lower(token(q)) != "matrixmarket") { status = fail(linenum, "expected 'MatrixMarket'"); break; }
      if (lower(token(q)) != "matrix") { status = fail(linenum, "expected 'matrix'"); break; }
      const std::string fmt = lower(token(q));
      if (fmt == "array") { status = fail(linenum, "'array' files are not supported"); break; }
      if (fmt != "coordinate") { status = fail(linenum, "expected 'coordinate'"); break; }
      const std::string field = lower(token(q));
      if (field == "pattern") pattern = true;
      else if (field != "real" && field != "complex") { status = fail(linenum, "only real, complex or pattern"); break; }
      const std::string shape = lower(token(q));
      if (shape == "symmetric") sym = true;
      else if (shape != "general") { status = fail(linenum, "only general or symmetric matrices"); break; }
      continue;
    }
    if (!size_seen) {
      if (!parse_long(p, rows) || !parse_long(p, cols) || !parse_long(p, count)) {
        status = fail(linenum, "could not read the matrix size");
        break;
      }
      if (rows < 0 || cols < 0 || count < 0 || rows >= (1ll << 31) || cols >= (1ll << 31)) {
        status = fail(linenum, "matrix size out of range");
        break;
      }
      size_seen = true;
      entries.reserve(static_cast<size_t>(std::min<long long>(count, 1ll << 26)) * (sym ? 2 : 1));
      if (count == 0) break;
      continue;
    }
    long long r = 0, c = 0;
    double v = 1.0;
    if (!parse_long(p, r)) { status = fail(linenum, "could not read the row index"); break; }
    if (!parse_long(p, c)) { status = fail(linenum, "could not read the column index"); break; }
    if (!pattern && !parse_double(p, v)) { status = fail(linenum, "could not read the value"); break; }
    r -= index_base;
    c -= index_base;
    if (r < 0 || r >= rows) { status = fail(linenum, "row index out of bounds"); break; }
    if (c < 0 || c >= cols) { status = fail(linenum, "column index out of bounds"); break; }
    entries.push_back({static_cast<uint32_t>(r), static_cast<uint32_t>(c), entries.size(), v});
    if (sym) entries.push_back({static_cast<uint32_t>(c), static_cast<uint32_t>(r), entries.size(), v});
    if (++valid == count) break;
  }
  std::fclose(f);
  if (status) return status;
  if (!size_seen) return fail(linenum, "missing size line");
  if (valid < count) return fail(linenum, "the file ends before the entries its size line declares");
  // rows in order, columns ascending; of repeated (row, col) the last assignment stays
  std::sort(entries.begin(), entries.end(), [](const Entry& a, const Entry& b) {
    if (a.row != b.row) return a.row < b.row;
    if (a.col != b.col) return a.col < b.col;
    return a.seq < b.seq;
  });
  vclhip_mm_matrix* m = new vclhip_mm_matrix();
  m->n_rows = static_cast<uint32_t>(rows);
  m->n_cols = static_cast<uint32_t>(cols);
  m->symmetric = sym;
  m->row_ptr.assign(static_cast<size_t>(rows) + 1, 0);
  for (size_t i = 0; i < entries.size(); ++i) {
    if (i + 1 < entries.size() && entries[i + 1].row == entries[i].row && entries[i + 1].col == entries[i].col)
      continue;  // a later assignment of the same entry follows
    m->col.push_back(entries[i].col);
    m->val.push_back(entries[i].val);
    ++m->row_ptr[entries[i].row + 1];
  }
  if (m->col.size() > 0xFFFFFFFFull) {
    delete m;
    return fail(linenum, "more than 2^32 - 1 entries");
  }
  for (size_t i = 0; i < static_cast<size_t>(rows); ++i) m->row_ptr[i + 1] += m->row_ptr[i];
  *out = m;
  if (n_rows) *n_rows = m->n_rows;
  if (n_cols) *n_cols = m->n_cols;
  if (nnz) *nnz = static_cast<uint32_t>(m->col.size());
  if (symmetric) *symmetric = sym ? 1 : 0;
  if (lines_read) *lines_read = linenum;
  return VCLHIP_SUCCESS;
}

int vclhip_mm_csr(const vclhip_mm_matrix* m, uint32_t* row_ptr, uint32_t* col, double* val)
{
  if (!m || !row_ptr || (!m->col.empty() && (!col || !val))) return vclhip::invalid("null argument");
  std::copy(m->row_ptr.begin(), m->row_ptr.end(), row_ptr);
  std::copy(m->col.begin(), m->col.end(), col);
  std::copy(m->val.begin(), m->val.end(), val);
  return VCLHIP_SUCCESS;
}

int vclhip_mm_free(vclhip_mm_matrix* m)
{
  delete m;
  return VCLHIP_SUCCESS;
}

int vclhip_mm_write(const char* path, int64_t index_base, uint32_t n_rows, uint32_t n_cols, const uint32_t* row_ptr,
                    const uint32_t* col, const double* val)
{
  if (!path || !row_ptr || (row_ptr[n_rows] && (!col || !val))) return vclhip::invalid("null argument");
  FILE* f = std::fopen(path, "w");
  if (!f) {
    vclhip::set_error("MatrixMarket: cannot open file %s for writing", path);
    return VCLHIP_ERROR_INVALID_VALUE;
  }
  std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n");
  std::fprintf(f, "%u %u %u\n", n_rows, n_cols, row_ptr[n_rows]);
  for (uint32_t r = 0; r < n_rows; ++r)
    for (uint32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
      std::fprintf(f, "%lld %lld %g\n", static_cast<long long>(r) + index_base,
                   static_cast<long long>(col[k]) + index_base, val[k]);
  const bool ok = std::ferror(f) == 0;
  if (std::fclose(f) != 0 || !ok) {
    vclhip::set_error("MatrixMarket: write error on %s", path);
    return VCLHIP_ERROR_INVALID_VALUE;
  }
  return VCLHIP_SUCCESS;
}

}  // extern "C"
