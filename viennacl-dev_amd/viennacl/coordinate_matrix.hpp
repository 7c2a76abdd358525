#ifndef VIENNACL_COORDINATE_MATRIX_HPP
#define VIENNACL_COORDINATE_MATRIX_HPP
// coordinate_matrix.hpp -- COO matrix of the drop-in API (reference coordinate_matrix.hpp:46-260).
// handle12 = (row, col) pairs (uint32), handle = values, handle3 = the 64 group boundaries of the
// reference's device kernels (coordinate_matrix.hpp:68-91).  The entries are kept in row order, as
// copy() writes them; the HIP SpMV runs the CSR kernel over row pointers and a row-block plan of
// that order, built on the host with the entries (row_ptr_handle(), blocks_handle()).

#include <map>
#include <vector>

#include "viennacl/backend/memory.hpp"
#include "viennacl/vector.hpp"

namespace viennacl
{

template <typename NumericT, unsigned int AlignmentV>
class coordinate_matrix
{
public:
  typedef NumericT value_type;
  typedef backend::mem_handle handle_type;
  typedef vcl_size_t size_type;

  coordinate_matrix() : rows_(0), cols_(0), nonzeros_(0), num_blocks_(0) {}
  coordinate_matrix(size_type rows, size_type cols) : rows_(rows), cols_(cols), nonzeros_(0), num_blocks_(0) {}
  explicit coordinate_matrix(viennacl::context) : rows_(0), cols_(0), nonzeros_(0), num_blocks_(0) {}

  size_type size1() const { return rows_; }
  size_type size2() const { return cols_; }
  size_type nnz() const { return nonzeros_; }
  size_type internal_nnz() const { return tools::align_to_multiple<size_type>(nonzeros_, AlignmentV); }

  handle_type const& handle12() const { return coords_; }
  handle_type const& handle() const { return elements_; }
  handle_type const& handle3() const { return group_boundaries_; }
  handle_type const& row_ptr_handle() const { return row_ptr_; }
  handle_type const& blocks_handle() const { return blocks_; }

  vclhip_coo descriptor() const
  {
    vclhip_coo d;
    d.n_rows = uint32_t(rows_);
    d.n_cols = uint32_t(cols_);
    d.nnz = uint32_t(nonzeros_);
    d.num_blocks = num_blocks_;
    d.coords = coords_.template hip_ptr<uint32_t>();
    d.elements = elements_.hip_handle();
    d.row_ptr = row_ptr_.template hip_ptr<uint32_t>();
    d.blocks = blocks_.template hip_ptr<uint32_t>();
    return d;
  }

  /** Sets the matrix from host (row, col) pairs ordered by row and their values. */
  void set(std::vector<uint32_t> const& coords, std::vector<NumericT> const& values, size_type rows, size_type cols)
  {
    rows_ = rows;
    cols_ = cols;
    nonzeros_ = values.size();
    if (coords.size() != 2 * nonzeros_) throw std::invalid_argument("ViennaCL (HIP): coordinate_matrix::set sizes");
    std::vector<uint32_t> c(2 * internal_nnz(), 0u);
    std::copy(coords.begin(), coords.end(), c.begin());
    std::vector<NumericT> v(internal_nnz(), NumericT(0));
    std::copy(values.begin(), values.end(), v.begin());
    // Group boundaries as the reference computes them (coordinate_matrix.hpp:75-91).  Its threshold
    // truncates (fraction + 1) / 64 to an integer before scaling by the entry count, so groups 1..63
    // all end after the first nonempty row and the last group holds the rest.
    const size_type group_num = 64;
    std::vector<uint32_t> gb(group_num + 1, 0u);
    size_type fraction = 0;
    for (size_type r = 0, e = 0; r < rows_; ++r) {
      while (e < nonzeros_ && c[2 * e] == r) ++e;
      while (fraction < group_num && e > size_type(double(fraction + 1) / double(group_num)) * nonzeros_)
        gb[++fraction] = uint32_t(e);
    }
    gb[group_num] = uint32_t(nonzeros_);
    std::vector<uint32_t> rp(rows_ + 1), blocks(rows_ + 1);
    if (rows_)
      backend::hip::check(vclhip_coo_plan(c.data(), uint32_t(nonzeros_), uint32_t(rows_), rp.data(), blocks.data(),
                                          &num_blocks_), "coordinate_matrix plan");
    viennacl::context ctx(HIP_MEMORY);
    backend::memory_create(coords_, sizeof(uint32_t) * (c.empty() ? 2 : c.size()), ctx, c.empty() ? nullptr : c.data());
    backend::memory_create(elements_, sizeof(NumericT) * (v.empty() ? 1 : v.size()), ctx, v.empty() ? nullptr : v.data());
    backend::memory_create(group_boundaries_, sizeof(uint32_t) * gb.size(), ctx, gb.data());
    backend::memory_create(row_ptr_, sizeof(uint32_t) * rp.size(), ctx, rp.data());
    backend::memory_create(blocks_, sizeof(uint32_t) * (num_blocks_ + 1), ctx, blocks.data());
  }

private:
  size_type rows_;
  size_type cols_;
  size_type nonzeros_;
  uint32_t num_blocks_;
  handle_type coords_;
  handle_type elements_;
  handle_type group_boundaries_;
  handle_type row_ptr_;
  handle_type blocks_;
};

/** std::vector<std::map<index, value>> -> COO in row order (coordinate_matrix.hpp:46-130). */
template <typename IndexT, typename CpuT, typename NumericT, unsigned int AlignmentV>
void copy(std::vector<std::map<IndexT, CpuT> > const& cpu, coordinate_matrix<NumericT, AlignmentV>& gpu)
{
  std::vector<uint32_t> coords;
  std::vector<NumericT> val;
  vcl_size_t max_col = 0;
  for (vcl_size_t i = 0; i < cpu.size(); ++i)
    for (typename std::map<IndexT, CpuT>::const_iterator it = cpu[i].begin(); it != cpu[i].end(); ++it) {
      coords.push_back(uint32_t(i));
      coords.push_back(uint32_t(it->first));
      val.push_back(NumericT(it->second));
      if (vcl_size_t(it->first) + 1 > max_col) max_col = vcl_size_t(it->first) + 1;
    }
  gpu.set(coords, val, cpu.size(), gpu.size2() > max_col ? gpu.size2() : max_col);
}

/** COO -> std::vector<std::map<index, value>> (coordinate_matrix.hpp:133-181). */
template <typename IndexT, typename CpuT, typename NumericT, unsigned int AlignmentV>
void copy(coordinate_matrix<NumericT, AlignmentV> const& gpu, std::vector<std::map<IndexT, CpuT> >& cpu)
{
  cpu.assign(gpu.size1(), std::map<IndexT, CpuT>());
  if (gpu.nnz() == 0) return;
  std::vector<uint32_t> c(2 * gpu.nnz());
  std::vector<NumericT> v(gpu.nnz());
  backend::memory_read(gpu.handle12(), 0, sizeof(uint32_t) * c.size(), c.data());
  backend::memory_read(gpu.handle(), 0, sizeof(NumericT) * v.size(), v.data());
  for (vcl_size_t e = 0; e < gpu.nnz(); ++e) cpu[c[2 * e]][IndexT(c[2 * e + 1])] = CpuT(v[e]);
}

namespace traits
{
template <typename T, unsigned A> vcl_size_t size1(coordinate_matrix<T, A> const& m) { return m.size1(); }
template <typename T, unsigned A> vcl_size_t size2(coordinate_matrix<T, A> const& m) { return m.size2(); }
}  // namespace traits

}  // namespace viennacl

#include "viennacl/linalg/sparse_matrix_operations.hpp"

#endif
