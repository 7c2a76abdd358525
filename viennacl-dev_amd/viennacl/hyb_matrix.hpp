#ifndef VIENNACL_HYB_MATRIX_HPP
#define VIENNACL_HYB_MATRIX_HPP
// hyb_matrix.hpp -- hybrid ELL + CSR matrix of the drop-in API (reference hyb_matrix.hpp:38-420).
// handle2 / handle = the ELL part (column-major, internal_ellnnz() columns of internal_size1()
// rows), handle3 / handle4 / handle5 = the CSR part (row pointers, columns, values) that holds the
// entries of each row beyond the ELL width.  The ELL width is the smallest row length that at least
// csr_threshold() of the rows do not exceed (default 0.8, hyb_matrix.hpp:44, :150-169), so a few
// long rows do not widen the ELL part.  When some row's CSR part is longer than 32 entries,
// blocks_handle() holds the row-block plan of the CSR part, with which the product sums such a
// part by a wavefront; otherwise one lane sums each row, ELL part then CSR part.

#include <algorithm>
#include <map>
#include <vector>

#include "viennacl/backend/memory.hpp"
#include "viennacl/ell_matrix.hpp"
#include "viennacl/vector.hpp"

namespace viennacl
{

template <typename NumericT, unsigned int AlignmentV>
class hyb_matrix
{
public:
  typedef NumericT value_type;
  typedef backend::mem_handle handle_type;
  typedef vcl_size_t size_type;

  hyb_matrix() : csr_threshold_(NumericT(0.8)), rows_(0), cols_(0), ellnnz_(0), csrnnz_(0), num_blocks_(0) {}
  explicit hyb_matrix(viennacl::context)
    : csr_threshold_(NumericT(0.8)), rows_(0), cols_(0), ellnnz_(0), csrnnz_(0), num_blocks_(0) {}
  hyb_matrix(size_type rows, size_type cols)
    : csr_threshold_(NumericT(0.8)), rows_(rows), cols_(cols), ellnnz_(0), csrnnz_(0), num_blocks_(0) {}

  NumericT csr_threshold() const { return csr_threshold_; }
  void csr_threshold(NumericT thr) { csr_threshold_ = thr; }

  size_type size1() const { return rows_; }
  size_type size2() const { return cols_; }
  size_type internal_size1() const { return tools::align_to_multiple<size_type>(rows_, AlignmentV); }
  size_type internal_size2() const { return tools::align_to_multiple<size_type>(cols_, AlignmentV); }
  size_type ell_nnz() const { return ellnnz_; }
  size_type internal_ellnnz() const { return tools::align_to_multiple<size_type>(ellnnz_, AlignmentV); }
  size_type csr_nnz() const { return csrnnz_; }

  handle_type const& handle() const { return ell_elements_; }
  handle_type const& handle2() const { return ell_coords_; }
  handle_type const& handle3() const { return csr_rows_; }
  handle_type const& handle4() const { return csr_cols_; }
  handle_type const& handle5() const { return csr_elements_; }
  handle_type const& blocks_handle() const { return blocks_; }

  vclhip_hyb descriptor() const
  {
    vclhip_hyb d;
    d.n_rows = uint32_t(rows_);
    d.n_cols = uint32_t(cols_);
    d.internal_size1 = uint32_t(internal_size1());
    d.internal_ellnnz = uint32_t(internal_ellnnz());
    d.ell_coords = ell_coords_.template hip_ptr<uint32_t>();
    d.ell_elements = ell_elements_.hip_handle();
    d.csr_rows = csr_rows_.template hip_ptr<uint32_t>();
    d.csr_cols = csr_cols_.template hip_ptr<uint32_t>();
    d.csr_elements = csr_elements_.hip_handle();
    d.num_blocks = num_blocks_;
    d.blocks = num_blocks_ ? blocks_.template hip_ptr<uint32_t>() : nullptr;
    return d;
  }

  /** Builds both parts from host CSR arrays (the split of hyb_matrix.hpp:136-240). */
  void set_from_csr(std::vector<uint32_t> const& rp, std::vector<uint32_t> const& col, std::vector<NumericT> const& val,
                    size_type cols)
  {
    rows_ = rp.size() - 1;
    cols_ = cols;
    if (rows_ == 0) return;
    std::vector<double> dval(val.begin(), val.end());
    uint32_t width = 0, cnnz = 0;
    backend::hip::check(vclhip_hyb_from_csr_f64(uint32_t(rows_), double(csr_threshold_), AlignmentV, rp.data(),
                                                col.data(), dval.data(), &width, &cnnz, nullptr, nullptr, nullptr,
                                                nullptr, nullptr), "hyb_matrix sizes");
    ellnnz_ = width;  // already a multiple of AlignmentV
    csrnnz_ = cnnz;
    const size_type ell = internal_size1() * width;
    std::vector<uint32_t> ec(ell ? ell : 1), cr(rows_ + 1), cc(cnnz);
    std::vector<double> ee(ell ? ell : 1), ce(cnnz);
    backend::hip::check(vclhip_hyb_from_csr_f64(uint32_t(rows_), double(csr_threshold_), AlignmentV, rp.data(),
                                                col.data(), dval.data(), &width, &cnnz, ec.data(), ee.data(),
                                                cr.data(), cc.data(), ce.data()), "hyb_matrix build");
    std::vector<NumericT> eev(ee.begin(), ee.end()), cev(ce.begin(), ce.end());
    uint32_t longest = 0;
    for (size_type r = 0; r < rows_; ++r) longest = std::max(longest, cr[r + 1] - cr[r]);
    std::vector<uint32_t> blocks(rows_ + 1);
    num_blocks_ = 0;
    if (longest > 32)
      backend::hip::check(vclhip_csr_plan_blocks(cr.data(), uint32_t(rows_), blocks.data(), &num_blocks_),
                          "hyb_matrix plan");
    viennacl::context ctx(HIP_MEMORY);
    backend::memory_create(ell_coords_, sizeof(uint32_t) * ec.size(), ctx, ec.data());
    backend::memory_create(ell_elements_, sizeof(NumericT) * eev.size(), ctx, eev.data());
    backend::memory_create(csr_rows_, sizeof(uint32_t) * cr.size(), ctx, cr.data());
    backend::memory_create(csr_cols_, sizeof(uint32_t) * cc.size(), ctx, cc.data());
    backend::memory_create(csr_elements_, sizeof(NumericT) * cev.size(), ctx, cev.data());
    if (num_blocks_) backend::memory_create(blocks_, sizeof(uint32_t) * (num_blocks_ + 1), ctx, blocks.data());
  }

private:
  NumericT csr_threshold_;
  size_type rows_;
  size_type cols_;
  size_type ellnnz_;
  size_type csrnnz_;
  uint32_t num_blocks_;
  handle_type ell_coords_;
  handle_type ell_elements_;
  handle_type csr_rows_;
  handle_type csr_cols_;
  handle_type csr_elements_;
  handle_type blocks_;
};

/** std::vector<std::map<index, value>> -> HYB (hyb_matrix.hpp:136-255). */
template <typename IndexT, typename CpuT, typename NumericT, unsigned int AlignmentV>
void copy(std::vector<std::map<IndexT, CpuT> > const& cpu, hyb_matrix<NumericT, AlignmentV>& gpu)
{
  std::vector<uint32_t> rp, col;
  std::vector<NumericT> val;
  const vcl_size_t max_col = detail::stl_to_csr(cpu, rp, col, val);
  gpu.set_from_csr(rp, col, val, gpu.size2() > max_col ? gpu.size2() : max_col);
}

/** HYB -> std::vector<std::map<index, value>> (hyb_matrix.hpp:258-320). */
template <typename IndexT, typename CpuT, typename NumericT, unsigned int AlignmentV>
void copy(hyb_matrix<NumericT, AlignmentV> const& gpu, std::vector<std::map<IndexT, CpuT> >& cpu)
{
  cpu.assign(gpu.size1(), std::map<IndexT, CpuT>());
  if (gpu.size1() == 0) return;
  const vcl_size_t ell = gpu.internal_size1() * gpu.internal_ellnnz();
  std::vector<uint32_t> ec(ell), cr(gpu.size1() + 1), cc(gpu.csr_nnz());
  std::vector<NumericT> ee(ell), ce(gpu.csr_nnz());
  if (ell) {
    backend::memory_read(gpu.handle2(), 0, sizeof(uint32_t) * ell, ec.data());
    backend::memory_read(gpu.handle(), 0, sizeof(NumericT) * ell, ee.data());
  }
  backend::memory_read(gpu.handle3(), 0, sizeof(uint32_t) * cr.size(), cr.data());
  backend::memory_read(gpu.handle4(), 0, sizeof(uint32_t) * cc.size(), cc.data());
  backend::memory_read(gpu.handle5(), 0, sizeof(NumericT) * ce.size(), ce.data());
  for (vcl_size_t r = 0; r < gpu.size1(); ++r) {
    for (vcl_size_t k = 0; k < gpu.internal_ellnnz(); ++k) {
      const vcl_size_t off = k * gpu.internal_size1() + r;
      if (ee[off] < 0 || ee[off] > 0) cpu[r][IndexT(ec[off])] = CpuT(ee[off]);
    }
    for (uint32_t e = cr[r]; e < cr[r + 1]; ++e) cpu[r][IndexT(cc[e])] = CpuT(ce[e]);
  }
}

namespace traits
{
template <typename T, unsigned A> vcl_size_t size1(hyb_matrix<T, A> const& m) { return m.size1(); }
template <typename T, unsigned A> vcl_size_t size2(hyb_matrix<T, A> const& m) { return m.size2(); }
}  // namespace traits

}  // namespace viennacl

#include "viennacl/linalg/sparse_matrix_operations.hpp"

#endif
