#ifndef VIENNACL_COMPRESSED_MATRIX_HPP
#define VIENNACL_COMPRESSED_MATRIX_HPP
// compressed_matrix.hpp -- CSR matrix of the drop-in API (reference compressed_matrix.hpp:628-1318).
// handle1 = row pointers (uint32), handle2 = column indices (uint32), handle = values,
// handle3 = the row-block plan of the HIP SpMV kernels (built on the host whenever the
// structure is set, like the reference's generate_row_block_information).

#include <map>
#include <memory>
#include <vector>

#include "viennacl/backend/memory.hpp"
#include "viennacl/vector.hpp"

namespace viennacl
{

template <typename NumericT, unsigned int AlignmentV>
class compressed_matrix
{
public:
  typedef NumericT value_type;
  typedef backend::mem_handle handle_type;
  typedef vcl_size_t size_type;

  compressed_matrix() : rows_(0), cols_(0), nonzeros_(0), row_block_num_(0), wrapped_(false) {}

  /** Allocates (zeroed) storage for a rows x cols matrix with nonzeros entries (compressed_matrix.hpp:642-676). */
  compressed_matrix(size_type rows, size_type cols, size_type nonzeros = 0, viennacl::context ctx = viennacl::context())
    : rows_(rows), cols_(cols), nonzeros_(nonzeros), row_block_num_(0), wrapped_(false)
  {
    std::vector<uint32_t> rp(rows + 1, 0);
    backend::memory_create(row_buffer_, sizeof(uint32_t) * (rows + 1), ctx, rp.data());
    if (nonzeros) {
      backend::memory_create(col_buffer_, sizeof(uint32_t) * nonzeros, ctx);
      backend::memory_create(elements_, sizeof(NumericT) * nonzeros, ctx);
    }
    generate_row_block_information();
  }

  explicit compressed_matrix(viennacl::context) : rows_(0), cols_(0), nonzeros_(0), row_block_num_(0), wrapped_(false) {}

  /** Wraps caller-owned device CSR arrays (compressed_matrix.hpp:740-781); the caller keeps ownership
   *  and the arrays must outlive the matrix.  The row-block plan is built from a host copy of row_ptr. */
  compressed_matrix(const uint32_t* row_ptr_dev, const uint32_t* col_dev, const NumericT* values_dev,
                    memory_types mem_type, size_type rows, size_type cols, size_type nonzeros)
    : rows_(rows), cols_(cols), nonzeros_(nonzeros), row_block_num_(0), wrapped_(false)
  {
    if (mem_type != HIP_MEMORY) throw memory_exception("only HIP_MEMORY arrays can be wrapped");
    wrapped_ = true;
    row_buffer_.wrap_hip(const_cast<uint32_t*>(row_ptr_dev), sizeof(uint32_t) * (rows + 1));
    col_buffer_.wrap_hip(const_cast<uint32_t*>(col_dev), sizeof(uint32_t) * nonzeros);
    elements_.wrap_hip(const_cast<NumericT*>(values_dev), sizeof(NumericT) * nonzeros);
    generate_row_block_information();
  }

  compressed_matrix(compressed_matrix const& other)
    : rows_(0), cols_(0), nonzeros_(0), row_block_num_(0), wrapped_(false)
  {
    *this = other;
  }

  compressed_matrix& operator=(compressed_matrix const& other)
  {
    if (this == &other) return *this;
    dia_.reset();
    wrapped_ = false;
    rows_ = other.rows_;
    cols_ = other.cols_;
    nonzeros_ = other.nonzeros_;
    row_block_num_ = other.row_block_num_;
    viennacl::context ctx(HIP_MEMORY);
    backend::memory_create(row_buffer_, other.row_buffer_.raw_size(), ctx);
    backend::memory_copy(other.row_buffer_, row_buffer_, 0, 0, other.row_buffer_.raw_size());
    backend::memory_create(col_buffer_, other.col_buffer_.raw_size(), ctx);
    backend::memory_copy(other.col_buffer_, col_buffer_, 0, 0, other.col_buffer_.raw_size());
    backend::memory_create(elements_, other.elements_.raw_size(), ctx);
    backend::memory_copy(other.elements_, elements_, 0, 0, other.elements_.raw_size());
    backend::memory_create(row_blocks_, other.row_blocks_.raw_size(), ctx);
    backend::memory_copy(other.row_blocks_, row_blocks_, 0, 0, other.row_blocks_.raw_size());
    return *this;
  }

  /** Sets the matrix from host CSR arrays (compressed_matrix.hpp:917-944). */
  void set(const void* row_jumper, const void* col_buffer, const NumericT* elements, size_type rows, size_type cols,
           size_type nonzeros)
  {
    dia_.reset();
    wrapped_ = false;
    viennacl::context ctx(HIP_MEMORY);
    backend::memory_create(row_buffer_, sizeof(uint32_t) * (rows + 1), ctx, row_jumper);
    backend::memory_create(col_buffer_, sizeof(uint32_t) * nonzeros, ctx, col_buffer);
    backend::memory_create(elements_, sizeof(NumericT) * nonzeros, ctx, elements);
    nonzeros_ = nonzeros;
    rows_ = rows;
    cols_ = cols;
    generate_row_block_information(static_cast<const uint32_t*>(row_jumper));
  }

  /** Builds the row-block plan of the SpMV kernels (role of compressed_matrix.hpp:1152-1188). */
  void generate_row_block_information(const uint32_t* host_row_ptr = nullptr)
  {
    std::vector<uint32_t> rp;
    if (!host_row_ptr) {
      rp.resize(rows_ + 1, 0);
      if (row_buffer_.raw_size()) backend::memory_read(row_buffer_, 0, sizeof(uint32_t) * (rows_ + 1), rp.data());
      host_row_ptr = rp.data();
    }
    std::vector<uint32_t> blocks(rows_ + 1, 0);
    uint32_t nb = 0;
    if (rows_ > 0) backend::hip::check(vclhip_csr_plan_blocks(host_row_ptr, uint32_t(rows_), blocks.data(), &nb), "row blocks");
    row_block_num_ = nb;
    backend::memory_create(row_blocks_, sizeof(uint32_t) * (nb + 1), viennacl::context(HIP_MEMORY), blocks.data());
  }

  /** Room for new_nonzeros entries (compressed_matrix.hpp:946-973): the column and value arrays grow,
   *  keeping their entries when `preserve`; nnz() then reports the new capacity, as in the reference. */
  void reserve(size_type new_nonzeros, bool preserve = true)
  {
    dia_.reset();
    if (new_nonzeros <= nonzeros_) return;
    viennacl::context ctx(HIP_MEMORY);
    handle_type col_new, val_new;
    backend::memory_create(col_new, sizeof(uint32_t) * new_nonzeros, ctx);
    backend::memory_create(val_new, sizeof(NumericT) * new_nonzeros, ctx);
    if (preserve && nonzeros_ > 0) {
      backend::memory_copy(col_buffer_, col_new, 0, 0, sizeof(uint32_t) * nonzeros_);
      backend::memory_copy(elements_, val_new, 0, 0, sizeof(NumericT) * nonzeros_);
    }
    col_buffer_.swap(col_new);
    elements_.swap(val_new);
    nonzeros_ = new_nonzeros;
  }

  /** New dimensions (compressed_matrix.hpp:982-1040).  With `preserve` the entries inside the new
   *  dimensions are kept and those in removed rows or columns are dropped (growing a matrix that had
   *  no rows stores an explicit zero at (0, 0), as the reference does); without it the matrix is
   *  left with no entries. */
  void resize(size_type new_rows, size_type new_cols, bool preserve = true)
  {
    if (new_rows == 0 || new_cols == 0) throw std::invalid_argument("ViennaCL (HIP): cannot resize a sparse matrix to zero size");
    if (new_rows == rows_ && new_cols == cols_) return;
    if (!preserve) {
      *this = compressed_matrix(new_rows, new_cols, 0);
      return;
    }
    std::vector<std::map<uint32_t, NumericT> > stl(rows_);
    if (rows_ > 0) {
      copy_to_stl(stl);
    } else {
      stl.resize(new_rows);
      stl[0][0] = NumericT(0);
    }
    stl.resize(new_rows);
    if (new_cols < cols_)
      for (vcl_size_t i = 0; i < stl.size(); ++i) stl[i].erase(stl[i].lower_bound(uint32_t(new_cols)), stl[i].end());
    set_from_stl(stl, new_cols);
  }

  /** Removes every entry, keeping the dimensions (compressed_matrix.hpp:1043-1055). */
  void clear() { *this = compressed_matrix(rows_, cols_, 0); }

  /** Entry (i, j) as a host-side proxy (compressed_matrix.hpp:1057-1076); an entry outside the
   *  sparsity pattern is inserted first (with the value 0), which rebuilds the matrix. */
  entry_proxy<NumericT> operator()(size_type i, size_type j)
  {
    dia_.reset();
    if (i >= rows_ || j >= cols_) throw std::out_of_range("ViennaCL (HIP): compressed_matrix index out of range");
    size_type index = element_index(i, j);
    if (index < nonzeros_) return entry_proxy<NumericT>(index, elements_);
    std::vector<std::map<uint32_t, NumericT> > stl(rows_);
    copy_to_stl(stl);
    stl[i][uint32_t(j)] = NumericT(0);
    set_from_stl(stl, cols_);
    return entry_proxy<NumericT>(element_index(i, j), elements_);
  }

  size_type size1() const { return rows_; }
  size_type size2() const { return cols_; }
  size_type nnz() const { return nonzeros_; }
  size_type blocks1() const { return row_block_num_; }

  handle_type const& handle1() const { return row_buffer_; }
  handle_type const& handle2() const { return col_buffer_; }
  handle_type const& handle3() const { return row_blocks_; }
  handle_type const& handle() const { return elements_; }
  // Writable access may change the matrix: it drops the solvers' DIA copy.
  handle_type& handle1() { dia_.reset(); return row_buffer_; }
  handle_type& handle2() { dia_.reset(); return col_buffer_; }
  handle_type& handle3() { dia_.reset(); return row_blocks_; }
  handle_type& handle() { dia_.reset(); return elements_; }

  /** The solvers' DIA detection and copy of this matrix (vclhip_dia_create_*), made on the first
   *  solve and kept until the matrix can have changed: set(), resize(), reserve(), operator()(i, j)
   *  and writable handle access drop it.  A matrix wrapping caller-owned arrays keeps none (the
   *  caller may change them at any time), and neither does an empty one. */
  vclhip_dia const* dia_cache() const
  {
    if (wrapped_ || rows_ == 0 || rows_ != cols_ || nonzeros_ == 0) return nullptr;
    if (!dia_) {
      vclhip_csr d = descriptor();
      vclhip_dia* p = nullptr;
      const int st = std::is_same<NumericT, float>::value ? vclhip_dia_create_f32(&d, &p, backend::hip::current_stream())
                                                          : vclhip_dia_create_f64(&d, &p, backend::hip::current_stream());
      backend::hip::check(st, "DIA operator");
      dia_.reset(p, [](vclhip_dia* q) { (void)vclhip_dia_destroy(q); });
    }
    return dia_.get();
  }

  /** The descriptor passed to the C ABI. */
  vclhip_csr descriptor() const
  {
    vclhip_csr d;
    d.n_rows = uint32_t(rows_);
    d.n_cols = uint32_t(cols_);
    d.nnz = uint32_t(nonzeros_);
    d.num_blocks = uint32_t(row_block_num_);
    d.row_ptr = row_buffer_.template hip_ptr<uint32_t>();
    d.col_idx = col_buffer_.template hip_ptr<uint32_t>();
    d.values = elements_.hip_handle();
    d.blocks = row_blocks_.template hip_ptr<uint32_t>();
    return d;
  }

  void switch_memory_context(viennacl::context ctx)
  {
    if (ctx.memory_type() != HIP_MEMORY) throw memory_exception("this build keeps sparse matrices in HIP_MEMORY");
  }

private:
  /** Position of entry (i, j) in the column and value arrays, or nonzeros_ when it is not stored. */
  size_type element_index(size_type i, size_type j) const
  {
    uint32_t rr[2];
    backend::memory_read(row_buffer_, sizeof(uint32_t) * i, sizeof(rr), rr);
    if (rr[1] <= rr[0]) return nonzeros_;
    std::vector<uint32_t> cols(rr[1] - rr[0]);
    backend::memory_read(col_buffer_, sizeof(uint32_t) * rr[0], sizeof(uint32_t) * cols.size(), cols.data());
    for (size_type k = 0; k < cols.size(); ++k)
      if (cols[k] == j) return rr[0] + k;
    return nonzeros_;
  }

  void copy_to_stl(std::vector<std::map<uint32_t, NumericT> >& stl) const
  {
    std::vector<uint32_t> rp(rows_ + 1, 0);
    if (row_buffer_.raw_size()) backend::memory_read(row_buffer_, 0, sizeof(uint32_t) * rp.size(), rp.data());
    const uint32_t stored = rp[rows_];
    std::vector<uint32_t> col(stored);
    std::vector<NumericT> val(stored);
    if (stored) {
      backend::memory_read(col_buffer_, 0, sizeof(uint32_t) * stored, col.data());
      backend::memory_read(elements_, 0, sizeof(NumericT) * stored, val.data());
    }
    for (size_type i = 0; i < rows_ && i < stl.size(); ++i)
      for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) stl[i][col[k]] = val[k];
  }

  void set_from_stl(std::vector<std::map<uint32_t, NumericT> > const& stl, size_type cols)
  {
    std::vector<uint32_t> rp(stl.size() + 1, 0), col;
    std::vector<NumericT> val;
    for (size_type i = 0; i < stl.size(); ++i) {
      for (typename std::map<uint32_t, NumericT>::const_iterator it = stl[i].begin(); it != stl[i].end(); ++it) {
        col.push_back(it->first);
        val.push_back(it->second);
      }
      rp[i + 1] = uint32_t(col.size());
    }
    set(rp.data(), col.data(), val.data(), stl.size(), cols, val.size());
  }

  size_type rows_;
  size_type cols_;
  size_type nonzeros_;
  size_type row_block_num_;
  bool wrapped_;
  mutable std::shared_ptr<vclhip_dia> dia_;
  handle_type row_buffer_;
  handle_type row_blocks_;
  handle_type col_buffer_;
  handle_type elements_;
};

/** Raw host CSR arrays -> device (compressed_matrix.hpp:107-146). */
template <typename NumericT, unsigned int AlignmentV>
void copy(std::vector<uint32_t> const& row_ptr, std::vector<uint32_t> const& col, std::vector<NumericT> const& val,
          vcl_size_t cols, compressed_matrix<NumericT, AlignmentV>& gpu)
{
  gpu.set(row_ptr.data(), col.data(), val.data(), row_ptr.size() - 1, cols, val.size());
}

/** std::vector<std::map<index, value>> -> device (compressed_matrix.hpp:201-218). */
template <typename IndexT, typename CpuT, typename NumericT, unsigned int AlignmentV>
void copy(std::vector<std::map<IndexT, CpuT> > const& cpu, compressed_matrix<NumericT, AlignmentV>& gpu)
{
  std::vector<uint32_t> rp(cpu.size() + 1, 0), col;
  std::vector<NumericT> val;
  vcl_size_t max_col = 0;
  for (vcl_size_t i = 0; i < cpu.size(); ++i) {
    for (typename std::map<IndexT, CpuT>::const_iterator it = cpu[i].begin(); it != cpu[i].end(); ++it) {
      col.push_back(uint32_t(it->first));
      val.push_back(NumericT(it->second));
      if (vcl_size_t(it->first) + 1 > max_col) max_col = vcl_size_t(it->first) + 1;
    }
    rp[i + 1] = uint32_t(col.size());
  }
  // columns: one past the largest column index (the reference's max_col + 1), or the matrix's own
  // column count when that is larger
  const vcl_size_t cols = gpu.size2() > max_col ? gpu.size2() : max_col;
  gpu.set(rp.data(), col.data(), val.data(), cpu.size(), cols, val.size());
}

/** device -> std::vector<std::map<index, value>> (compressed_matrix.hpp copy back). */
template <typename IndexT, typename CpuT, typename NumericT, unsigned int AlignmentV>
void copy(compressed_matrix<NumericT, AlignmentV> const& gpu, std::vector<std::map<IndexT, CpuT> >& cpu)
{
  std::vector<uint32_t> rp(gpu.size1() + 1), col(gpu.nnz());
  std::vector<NumericT> val(gpu.nnz());
  if (gpu.size1()) backend::memory_read(gpu.handle1(), 0, sizeof(uint32_t) * rp.size(), rp.data());
  if (gpu.nnz()) {
    backend::memory_read(gpu.handle2(), 0, sizeof(uint32_t) * col.size(), col.data());
    backend::memory_read(gpu.handle(), 0, sizeof(NumericT) * val.size(), val.data());
  }
  cpu.assign(gpu.size1(), std::map<IndexT, CpuT>());
  for (vcl_size_t i = 0; i < gpu.size1(); ++i)
    for (uint32_t k = rp[i]; k < rp[i + 1]; ++k) cpu[i][IndexT(col[k])] = CpuT(val[k]);
}

namespace traits
{
template <typename T, unsigned A> vcl_size_t size1(compressed_matrix<T, A> const& m) { return m.size1(); }
template <typename T, unsigned A> vcl_size_t size2(compressed_matrix<T, A> const& m) { return m.size2(); }
template <typename T, unsigned A> viennacl::context context(compressed_matrix<T, A> const&) { return viennacl::context(HIP_MEMORY); }
}  // namespace traits

}  // namespace viennacl

#include "viennacl/linalg/sparse_matrix_operations.hpp"

#endif
