#ifndef VIENNACL_SLICED_ELL_MATRIX_HPP
#define VIENNACL_SLICED_ELL_MATRIX_HPP
// sliced_ell_matrix.hpp -- SELL-C-sigma (sigma = 1) matrix of the drop-in API (reference
// sliced_ell_matrix.hpp:46-375).  handle1 = columns per slice, handle2 = column indices,
// handle3 = slice start offsets, handle = values; slice s, entry k, row-in-slice i lives at
// block_start[s] + k*C + i.  With C = 64 one slice is exactly one wavefront of the HIP kernel.

#include <map>
#include <vector>

#include "viennacl/backend/memory.hpp"
#include "viennacl/vector.hpp"

namespace viennacl
{

template <typename ScalarT, typename IndexT>
class sliced_ell_matrix
{
  static_assert(sizeof(IndexT) == 4, "the HIP backend uses 32-bit indices");

public:
  typedef ScalarT value_type;
  typedef backend::mem_handle handle_type;
  typedef vcl_size_t size_type;

  /** Rows per slice when none is given: one 64-lane wavefront (the reference defaults to 32,
   *  sliced_ell_matrix.hpp:146-147, sized for 32-wide warps). */
  static const size_type default_rows_per_block = 64;

  sliced_ell_matrix() : rows_(0), cols_(0), rows_per_block_(0) {}
  sliced_ell_matrix(size_type rows, size_type cols, size_type rows_per_block = 0)
    : rows_(rows), cols_(cols), rows_per_block_(rows_per_block) {}
  explicit sliced_ell_matrix(viennacl::context) : rows_(0), cols_(0), rows_per_block_(0) {}

  size_type size1() const { return rows_; }
  size_type size2() const { return cols_; }
  size_type rows_per_block() const { return rows_per_block_; }
  size_type internal_size1() const { return tools::align_to_multiple<size_type>(rows_, rows_per_block_); }
  size_type internal_size2() const { return cols_; }

  handle_type const& handle1() const { return columns_per_block_; }
  handle_type const& handle2() const { return column_indices_; }
  handle_type const& handle3() const { return block_start_; }
  handle_type const& handle() const { return elements_; }
  handle_type& handle1() { return columns_per_block_; }
  handle_type& handle2() { return column_indices_; }
  handle_type& handle3() { return block_start_; }
  handle_type& handle() { return elements_; }

  vclhip_sell descriptor() const
  {
    vclhip_sell d;
    d.n_rows = uint32_t(rows_);
    d.n_cols = uint32_t(cols_);
    d.C = uint32_t(rows_per_block_);
    d.num_slices = rows_per_block_ ? uint32_t((rows_ + rows_per_block_ - 1) / rows_per_block_) : 0;
    d.columns_per_block = columns_per_block_.template hip_ptr<uint32_t>();
    d.column_indices = column_indices_.template hip_ptr<uint32_t>();
    d.block_start = block_start_.template hip_ptr<uint32_t>();
    d.elements = elements_.hip_handle();
    return d;
  }

  /** Builds the SELL arrays from host CSR arrays (the construction of sliced_ell_matrix.hpp:140-214). */
  void set_from_csr(std::vector<uint32_t> const& rp, std::vector<uint32_t> const& col, std::vector<ScalarT> const& val,
                    size_type cols)
  {
    if (rows_per_block_ == 0) rows_per_block_ = default_rows_per_block;
    rows_ = rp.size() - 1;
    cols_ = cols;
    std::vector<double> dval(val.begin(), val.end());
    uint32_t ns = 0;
    uint64_t stored = 0;
    backend::hip::check(vclhip_sell_from_csr_f64(uint32_t(rows_), uint32_t(rows_per_block_), rp.data(), col.data(),
                                                 dval.data(), &ns, &stored, nullptr, nullptr, nullptr, nullptr),
                        "sliced_ell sizes");
    std::vector<uint32_t> cpb(ns), bs(ns), ci(stored ? stored : 1);
    std::vector<double> el(stored ? stored : 1);
    backend::hip::check(vclhip_sell_from_csr_f64(uint32_t(rows_), uint32_t(rows_per_block_), rp.data(), col.data(),
                                                 dval.data(), &ns, &stored, cpb.data(), ci.data(), bs.data(), el.data()),
                        "sliced_ell build");
    std::vector<ScalarT> elc(el.begin(), el.end());
    viennacl::context ctx(HIP_MEMORY);
    backend::memory_create(columns_per_block_, sizeof(uint32_t) * ns, ctx, cpb.data());
    backend::memory_create(block_start_, sizeof(uint32_t) * ns, ctx, bs.data());
    backend::memory_create(column_indices_, sizeof(uint32_t) * ci.size(), ctx, ci.data());
    backend::memory_create(elements_, sizeof(ScalarT) * elc.size(), ctx, elc.data());
  }

private:
  size_type rows_;
  size_type cols_;
  size_type rows_per_block_;
  handle_type columns_per_block_;
  handle_type column_indices_;
  handle_type block_start_;
  handle_type elements_;
};

/** std::vector<std::map<index, value>> -> SELL-C (sliced_ell_matrix.hpp:223-235). */
template <typename IndexT, typename CpuT, typename ScalarT, typename IndexT2>
void copy(std::vector<std::map<IndexT, CpuT> > const& cpu, sliced_ell_matrix<ScalarT, IndexT2>& gpu)
{
  std::vector<uint32_t> rp(cpu.size() + 1, 0), col;
  std::vector<ScalarT> val;
  vcl_size_t max_col = 0;
  for (vcl_size_t i = 0; i < cpu.size(); ++i) {
    for (typename std::map<IndexT, CpuT>::const_iterator it = cpu[i].begin(); it != cpu[i].end(); ++it) {
      col.push_back(uint32_t(it->first));
      val.push_back(ScalarT(it->second));
      if (vcl_size_t(it->first) + 1 > max_col) max_col = vcl_size_t(it->first) + 1;
    }
    rp[i + 1] = uint32_t(col.size());
  }
  gpu.set_from_csr(rp, col, val, gpu.size2() > max_col ? gpu.size2() : max_col);
}

namespace traits
{
template <typename T, typename I> vcl_size_t size1(sliced_ell_matrix<T, I> const& m) { return m.size1(); }
template <typename T, typename I> vcl_size_t size2(sliced_ell_matrix<T, I> const& m) { return m.size2(); }
}  // namespace traits

}  // namespace viennacl

#include "viennacl/linalg/sparse_matrix_operations.hpp"

#endif
