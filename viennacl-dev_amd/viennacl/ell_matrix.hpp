#ifndef VIENNACL_ELL_MATRIX_HPP
#define VIENNACL_ELL_MATRIX_HPP
// ell_matrix.hpp -- ELL matrix of the drop-in API (reference ell_matrix.hpp:46-240).  handle2 =
// column indices, handle = values, both column-major with internal_size1() rows: entry k of row i
// is at k * internal_size1() + i, k < internal_maxnnz(); padding entries hold 0.  The HIP SpMV
// runs one lane per row, so a wavefront reads 64 consecutive entries of each column.

#include <map>
#include <vector>

#include "viennacl/backend/memory.hpp"
#include "viennacl/vector.hpp"

namespace viennacl
{

template <typename NumericT, unsigned int AlignmentV>
class ell_matrix
{
public:
  typedef NumericT value_type;
  typedef backend::mem_handle handle_type;
  typedef vcl_size_t size_type;

  ell_matrix() : rows_(0), cols_(0), maxnnz_(0) {}
  ell_matrix(size_type rows, size_type cols) : rows_(rows), cols_(cols), maxnnz_(0) {}
  explicit ell_matrix(viennacl::context) : rows_(0), cols_(0), maxnnz_(0) {}

  size_type size1() const { return rows_; }
  size_type size2() const { return cols_; }
  size_type internal_size1() const { return tools::align_to_multiple<size_type>(rows_, AlignmentV); }
  size_type internal_size2() const { return tools::align_to_multiple<size_type>(cols_, AlignmentV); }
  size_type maxnnz() const { return maxnnz_; }
  size_type internal_maxnnz() const { return tools::align_to_multiple<size_type>(maxnnz_, AlignmentV); }
  size_type nnz() const { return rows_ * maxnnz_; }
  size_type internal_nnz() const { return internal_size1() * internal_maxnnz(); }

  handle_type const& handle2() const { return coords_; }
  handle_type const& handle() const { return elements_; }

  vclhip_ell descriptor() const
  {
    vclhip_ell d;
    d.n_rows = uint32_t(rows_);
    d.n_cols = uint32_t(cols_);
    d.internal_size1 = uint32_t(internal_size1());
    d.internal_maxnnz = uint32_t(internal_maxnnz());
    d.coords = coords_.template hip_ptr<uint32_t>();
    d.elements = elements_.hip_handle();
    return d;
  }

  /** Builds the ELL arrays from host CSR arrays (the layout of ell_matrix.hpp:123-171). */
  void set_from_csr(std::vector<uint32_t> const& rp, std::vector<uint32_t> const& col, std::vector<NumericT> const& val,
                    size_type cols)
  {
    rows_ = rp.size() - 1;
    cols_ = cols;
    maxnnz_ = 0;
    for (size_type r = 0; r < rows_; ++r) maxnnz_ = std::max<size_type>(maxnnz_, rp[r + 1] - rp[r]);
    std::vector<double> dval(val.begin(), val.end());
    uint32_t width = 0;
    if (rows_ == 0) return;
    backend::hip::check(vclhip_ell_from_csr_f64(uint32_t(rows_), AlignmentV, rp.data(), col.data(), dval.data(),
                                                &width, nullptr, nullptr), "ell_matrix sizes");
    const size_type total = internal_size1() * width;
    std::vector<uint32_t> c(total ? total : 1);
    std::vector<double> e(total ? total : 1);
    backend::hip::check(vclhip_ell_from_csr_f64(uint32_t(rows_), AlignmentV, rp.data(), col.data(), dval.data(),
                                                &width, c.data(), e.data()), "ell_matrix build");
    std::vector<NumericT> ev(e.begin(), e.end());
    viennacl::context ctx(HIP_MEMORY);
    backend::memory_create(coords_, sizeof(uint32_t) * c.size(), ctx, c.data());
    backend::memory_create(elements_, sizeof(NumericT) * ev.size(), ctx, ev.data());
  }

private:
  size_type rows_;
  size_type cols_;
  size_type maxnnz_;
  handle_type coords_;
  handle_type elements_;
};

namespace detail
{
/** Host CSR arrays of an STL matrix, and the column count it implies. */
template <typename IndexT, typename CpuT, typename NumericT>
vcl_size_t stl_to_csr(std::vector<std::map<IndexT, CpuT> > const& cpu, std::vector<uint32_t>& rp,
                      std::vector<uint32_t>& col, std::vector<NumericT>& val)
{
  rp.assign(cpu.size() + 1, 0);
  col.clear();
  val.clear();
  vcl_size_t max_col = 0;
  for (vcl_size_t i = 0; i < cpu.size(); ++i) {
    for (typename std::map<IndexT, CpuT>::const_iterator it = cpu[i].begin(); it != cpu[i].end(); ++it) {
      col.push_back(uint32_t(it->first));
      val.push_back(NumericT(it->second));
      if (vcl_size_t(it->first) + 1 > max_col) max_col = vcl_size_t(it->first) + 1;
    }
    rp[i + 1] = uint32_t(col.size());
  }
  return max_col;
}
}  // namespace detail

/** std::vector<std::map<index, value>> -> ELL (ell_matrix.hpp:123-197). */
template <typename IndexT, typename CpuT, typename NumericT, unsigned int AlignmentV>
void copy(std::vector<std::map<IndexT, CpuT> > const& cpu, ell_matrix<NumericT, AlignmentV>& gpu)
{
  std::vector<uint32_t> rp, col;
  std::vector<NumericT> val;
  const vcl_size_t max_col = detail::stl_to_csr(cpu, rp, col, val);
  gpu.set_from_csr(rp, col, val, gpu.size2() > max_col ? gpu.size2() : max_col);
}

/** ELL -> std::vector<std::map<index, value>>, nonzero entries (ell_matrix.hpp:200-240). */
template <typename IndexT, typename CpuT, typename NumericT, unsigned int AlignmentV>
void copy(ell_matrix<NumericT, AlignmentV> const& gpu, std::vector<std::map<IndexT, CpuT> >& cpu)
{
  cpu.assign(gpu.size1(), std::map<IndexT, CpuT>());
  const vcl_size_t total = gpu.internal_nnz();
  if (total == 0) return;
  std::vector<uint32_t> c(total);
  std::vector<NumericT> v(total);
  backend::memory_read(gpu.handle2(), 0, sizeof(uint32_t) * total, c.data());
  backend::memory_read(gpu.handle(), 0, sizeof(NumericT) * total, v.data());
  for (vcl_size_t k = 0; k < gpu.internal_maxnnz(); ++k)
    for (vcl_size_t r = 0; r < gpu.size1(); ++r) {
      const vcl_size_t off = k * gpu.internal_size1() + r;
      if (v[off] < 0 || v[off] > 0) cpu[r][IndexT(c[off])] = CpuT(v[off]);
    }
}

namespace traits
{
template <typename T, unsigned A> vcl_size_t size1(ell_matrix<T, A> const& m) { return m.size1(); }
template <typename T, unsigned A> vcl_size_t size2(ell_matrix<T, A> const& m) { return m.size2(); }
}  // namespace traits

}  // namespace viennacl

#include "viennacl/linalg/sparse_matrix_operations.hpp"

#endif
