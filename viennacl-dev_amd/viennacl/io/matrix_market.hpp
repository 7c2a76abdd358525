#ifndef VIENNACL_IO_MATRIX_MARKET_HPP
#define VIENNACL_IO_MATRIX_MARKET_HPP
// matrix_market.hpp -- io::read_matrix_market_file / write_matrix_market_file (reference
// io/matrix_market.hpp:102-433) for std::vector<std::map<unsigned int, T> >, compressed_matrix and
// sliced_ell_matrix.  Parsing and formatting are done by libvclhip (vclhip_mm_read /
// vclhip_mm_write); the return values follow the reference: the number of lines read on success,
// 0 on a format error, EXIT_FAILURE when the file cannot be opened.

#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "viennacl/compressed_matrix.hpp"
#include "viennacl/sliced_ell_matrix.hpp"

namespace viennacl
{
namespace io
{
namespace detail
{
struct mm_csr
{
  uint32_t n_rows = 0, n_cols = 0;
  std::vector<uint32_t> row_ptr, col;
  std::vector<double> val;
};

/** Reads `file` into CSR host arrays; returns the reference reader's status. */
inline long mm_read(const char* file, long index_base, mm_csr& out)
{
  {
    std::FILE* f = std::fopen(file, "r");
    if (!f) {
      std::cerr << "ViennaCL: Matrix Market Reader: Cannot open file " << file << std::endl;
      return EXIT_FAILURE;
    }
    std::fclose(f);
  }
  vclhip_mm_matrix* m = nullptr;
  uint32_t nnz = 0;
  int sym = 0;
  int64_t lines = 0;
  if (vclhip_mm_read(file, index_base, &m, &out.n_rows, &out.n_cols, &nnz, &sym, &lines) != 0) {
    std::cerr << "Error in file " << file << ": " << vclhip_last_error() << std::endl;
    return 0;
  }
  out.row_ptr.resize(out.n_rows + 1);
  out.col.resize(nnz);
  out.val.resize(nnz);
  const int st = vclhip_mm_csr(m, out.row_ptr.data(), out.col.data(), out.val.data());
  vclhip_mm_free(m);
  return st == 0 ? static_cast<long>(lines) : 0;
}
}  // namespace detail

template <typename ScalarT>
long read_matrix_market_file(std::vector<std::map<unsigned int, ScalarT> >& mat, const char* file,
                             long index_base = 1)
{
  detail::mm_csr c;
  const long st = detail::mm_read(file, index_base, c);
  if (st == 0 || (st == EXIT_FAILURE && c.row_ptr.empty())) return st;
  mat.clear();
  mat.resize(c.n_rows);
  for (uint32_t r = 0; r < c.n_rows; ++r)
    for (uint32_t k = c.row_ptr[r]; k < c.row_ptr[r + 1]; ++k) mat[r][c.col[k]] = static_cast<ScalarT>(c.val[k]);
  return st;
}

template <typename ScalarT>
long read_matrix_market_file(std::vector<std::map<unsigned int, ScalarT> >& mat, std::string const& file,
                             long index_base = 1)
{
  return read_matrix_market_file(mat, file.c_str(), index_base);
}

template <typename NumericT, unsigned int AlignmentV>
long read_matrix_market_file(compressed_matrix<NumericT, AlignmentV>& mat, const char* file, long index_base = 1)
{
  detail::mm_csr c;
  const long st = detail::mm_read(file, index_base, c);
  if (st == 0 || (st == EXIT_FAILURE && c.row_ptr.empty())) return st;
  std::vector<NumericT> v(c.val.begin(), c.val.end());
  mat.set(c.row_ptr.data(), c.col.data(), v.data(), c.n_rows, c.n_cols, c.col.size());
  return st;
}

template <typename NumericT, unsigned int AlignmentV>
long read_matrix_market_file(compressed_matrix<NumericT, AlignmentV>& mat, std::string const& file,
                             long index_base = 1)
{
  return read_matrix_market_file(mat, file.c_str(), index_base);
}

template <typename NumericT, typename IndexT>
long read_matrix_market_file(sliced_ell_matrix<NumericT, IndexT>& mat, const char* file, long index_base = 1)
{
  std::vector<std::map<unsigned int, NumericT> > stl;
  const long st = read_matrix_market_file(stl, file, index_base);
  if (st != 0 && !(st == EXIT_FAILURE && stl.empty())) viennacl::copy(stl, mat);
  return st;
}

template <typename ScalarT>
void write_matrix_market_file(std::vector<std::map<unsigned int, ScalarT> > const& mat, const char* file,
                              long index_base = 1)
{
  // the reference's adapter reports a square matrix of mat.size() rows (tools/adapter.hpp:193-194)
  std::vector<uint32_t> rp(mat.size() + 1, 0), col;
  std::vector<double> val;
  for (size_t r = 0; r < mat.size(); ++r) {
    for (auto const& kv : mat[r]) {
      col.push_back(kv.first);
      val.push_back(static_cast<double>(kv.second));
    }
    rp[r + 1] = static_cast<uint32_t>(col.size());
  }
  backend::hip::check(vclhip_mm_write(file, index_base, static_cast<uint32_t>(mat.size()),
                                      static_cast<uint32_t>(mat.size()), rp.data(), col.data(), val.data()),
                      "write_matrix_market_file");
}

template <typename ScalarT>
void write_matrix_market_file(std::vector<std::map<unsigned int, ScalarT> > const& mat, std::string const& file,
                              long index_base = 1)
{
  write_matrix_market_file(mat, file.c_str(), index_base);
}

template <typename NumericT, unsigned int AlignmentV>
void write_matrix_market_file(compressed_matrix<NumericT, AlignmentV> const& mat, std::string const& file,
                              long index_base = 1)
{
  std::vector<std::map<unsigned int, NumericT> > stl(mat.size1());
  viennacl::copy(mat, stl);
  std::vector<uint32_t> rp(mat.size1() + 1, 0), col;
  std::vector<double> val;
  for (size_t r = 0; r < stl.size(); ++r) {
    for (auto const& kv : stl[r]) {
      col.push_back(kv.first);
      val.push_back(static_cast<double>(kv.second));
    }
    rp[r + 1] = static_cast<uint32_t>(col.size());
  }
  backend::hip::check(vclhip_mm_write(file.c_str(), index_base, static_cast<uint32_t>(mat.size1()),
                                      static_cast<uint32_t>(mat.size2()), rp.data(), col.data(), val.data()),
                      "write_matrix_market_file");
}

}  // namespace io
}  // namespace viennacl

#endif
