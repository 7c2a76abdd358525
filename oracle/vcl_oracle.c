/*
 * vcl_oracle.c -- CPU restatement of ViennaCL 1.7.1's host_based backend and
 * Krylov solvers for CSR / SELL-C matrices.  TEST INFRASTRUCTURE ONLY: the
 * checker for the HIP path and the CPU baseline; never linked by the product.
 *
 * Each function names the reference lines it follows.  Expressions keep the
 * reference's operand order; build with -ffp-contract=off (oracle/Makefile).
 * OpenMP reductions follow the reference's `#pragma omp parallel for
 * reduction(+: ...)`; with one thread the summation order is sequential.
 */
#include "vcl_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void orc_set_num_threads(int n)
{
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int orc_get_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------ SpMV */

/* host_based/sparse_matrix_operations.hpp:145-186.  The alpha=1/beta=0 shortcut at
 * :152-159 never fires (it tests result.stride()==0) and computes the same values. */
#define ORC_CSR_SPMV(NAME, T)                                                              \
  void NAME(uint32_t n_rows, const uint32_t* row_ptr, const uint32_t* col, const T* val,   \
            const T* x, size_t x_start, size_t x_stride, T alpha,                          \
            T* y, size_t y_start, size_t y_stride, T beta)                                 \
  {                                                                                        \
    long row;                                                                              \
    _Pragma("omp parallel for")                                                            \
    for (row = 0; row < (long)n_rows; ++row) {                                             \
      T dot_prod = 0;                                                                      \
      size_t row_end = row_ptr[row + 1];                                                   \
      for (size_t i = row_ptr[row]; i < row_end; ++i)                                      \
        dot_prod += val[i] * x[(size_t)col[i] * x_stride + x_start];                       \
      size_t index = (size_t)row * y_stride + y_start;                                     \
      if (beta < 0 || beta > 0)                                                            \
        y[index] = alpha * dot_prod + beta * y[index];                                     \
      else                                                                                 \
        y[index] = alpha * dot_prod;                                                       \
    }                                                                                      \
  }

ORC_CSR_SPMV(orc_csr_spmv_f64, double)
ORC_CSR_SPMV(orc_csr_spmv_f32, float)

/* host_based/sparse_matrix_operations.hpp:1796-1868 */
#define ORC_SELL_SPMV(NAME, T)                                                             \
  void NAME(uint32_t n_rows, uint32_t C, const uint32_t* columns_per_block,                \
            const uint32_t* column_indices, const uint32_t* block_start, const T* elements,\
            const T* x, size_t x_start, size_t x_stride, T alpha,                          \
            T* y, size_t y_start, size_t y_stride, T beta)                                 \
  {                                                                                        \
    long num_blocks = (long)((n_rows + C - 1) / C);                                        \
    long b;                                                                                \
    _Pragma("omp parallel for")                                                            \
    for (b = 0; b < num_blocks; ++b) {                                                     \
      T* result_values = (T*)calloc(C, sizeof(T));                                         \
      size_t current_columns = columns_per_block[b];                                       \
      for (size_t k = 0; k < current_columns; ++k) {                                       \
        size_t stride_start = block_start[b] + k * C;                                      \
        for (uint32_t r = 0; r < C; ++r) {                                                 \
          T v = elements[stride_start + r];                                                \
          result_values[r] += (v > 0 || v < 0)                                             \
              ? x[(size_t)column_indices[stride_start + r] * x_stride + x_start] * v : 0;  \
        }                                                                                  \
      }                                                                                    \
      size_t first_row = (size_t)b * C;                                                    \
      for (uint32_t r = 0; r < C; ++r) {                                                   \
        if (first_row + r < n_rows) {                                                      \
          size_t index = (first_row + r) * y_stride + y_start;                             \
          if (beta < 0 || beta > 0)                                                        \
            y[index] = alpha * result_values[r] + beta * y[index];                         \
          else                                                                             \
            y[index] = alpha * result_values[r];                                           \
        }                                                                                  \
      }                                                                                    \
      free(result_values);                                                                 \
    }                                                                                      \
  }

ORC_SELL_SPMV(orc_sell_spmv_f64, double)
ORC_SELL_SPMV(orc_sell_spmv_f32, float)

/* sliced_ell_matrix.hpp:149-171: per-slice max row length. */
void orc_sell_sizes(uint32_t n_rows, uint32_t C, const uint32_t* row_ptr, uint32_t* num_slices, uint64_t* stored)
{
  uint32_t ns = (n_rows + C - 1) / C;
  uint64_t total = 0;
  for (uint32_t s = 0; s < ns; ++s) {
    uint32_t m = 0;
    for (uint32_t r = s * C; r < n_rows && r < (s + 1) * C; ++r) {
      uint32_t len = row_ptr[r + 1] - row_ptr[r];
      if (len > m) m = len;
    }
    total += (uint64_t)m * C;
  }
  *num_slices = ns;
  *stored = total;
}

/* sliced_ell_matrix.hpp:140-214: slice-major, column-major inside a slice, padding 0 / column 0. */
void orc_sell_from_csr_f64(uint32_t n_rows, uint32_t C, const uint32_t* row_ptr, const uint32_t* col, const double* val,
                           uint32_t* columns_per_block, uint32_t* column_indices, uint32_t* block_start, double* elements)
{
  uint32_t ns = (n_rows + C - 1) / C;
  uint64_t offset = 0;
  for (uint32_t s = 0; s < ns; ++s) {
    uint32_t m = 0;
    for (uint32_t r = s * C; r < n_rows && r < (s + 1) * C; ++r) {
      uint32_t len = row_ptr[r + 1] - row_ptr[r];
      if (len > m) m = len;
    }
    columns_per_block[s] = m;
    block_start[s] = (uint32_t)offset;
    for (uint64_t k = 0; k < (uint64_t)m * C; ++k) {
      column_indices[offset + k] = 0;
      elements[offset + k] = 0.0;
    }
    for (uint32_t r = s * C; r < n_rows && r < (s + 1) * C; ++r) {
      uint32_t e = 0;
      for (uint32_t i = row_ptr[r]; i < row_ptr[r + 1]; ++i, ++e) {
        uint64_t idx = offset + (uint64_t)e * C + (r - s * C);
        column_indices[idx] = col[i];
        elements[idx] = val[i];
      }
    }
    offset += (uint64_t)m * C;
  }
}

/* compressed_matrix.hpp:1152-1188 */
uint32_t orc_row_blocks(uint32_t n_rows, const uint32_t* row_ptr, uint32_t* row_blocks)
{
  const size_t shared_mem_size = 1024;
  size_t num_entries = 0;
  uint32_t nb = 0;
  row_blocks[0] = 0;
  for (size_t i = 0; i < n_rows; ++i) {
    size_t entries_in_row = (size_t)row_ptr[i + 1] - (size_t)row_ptr[i];
    num_entries += entries_in_row;
    if (num_entries > shared_mem_size) {
      size_t rows_in_batch = i - row_blocks[nb];
      if (rows_in_batch > 0)
        row_blocks[++nb] = (uint32_t)(i--);
      else
        row_blocks[++nb] = (uint32_t)(i + 1);
      num_entries = 0;
    }
  }
  if (num_entries > 0)
    row_blocks[++nb] = n_rows;
  return nb;
}

/* host_based/sparse_matrix_operations.hpp:52-98 */
void orc_csr_row_info_f64(uint32_t n_rows, const uint32_t* row_ptr, const uint32_t* col, const double* val,
                          double* out, int info)
{
  for (size_t row = 0; row < n_rows; ++row) {
    double value = 0;
    uint32_t row_end = row_ptr[row + 1];
    switch (info) {
      case 0:
        for (uint32_t i = row_ptr[row]; i < row_end; ++i)
          value = fmax(value, fabs(val[i]));
        break;
      case 1:
        for (uint32_t i = row_ptr[row]; i < row_end; ++i)
          value += fabs(val[i]);
        break;
      case 2:
        for (uint32_t i = row_ptr[row]; i < row_end; ++i)
          value += val[i] * val[i];
        value = sqrt(value);
        break;
      default:
        for (uint32_t i = row_ptr[row]; i < row_end; ++i)
          if (col[i] == row) { value = val[i]; break; }
        break;
    }
    out[row] = value;
  }
}

/* ------------------------------------------------------------ vector ops */

/* Reduction order.  0 (default): the reference's own loops (sequential with one OpenMP thread).
 * k > 0: every inner product is summed as k contiguous blocks, each sequential, the block sums added
 * in order -- the order of the reference's `omp parallel for reduction(+:)` with k threads and a
 * static schedule, applied regardless of the reference's 5000-element OpenMP threshold.  Used to
 * measure how far a solver's iteration count moves when only the summation order changes. */
static int g_red_blocks = 0;

void orc_set_reduction_blocks(int k) { g_red_blocks = k > 0 ? k : 0; }

static size_t blk_lo(size_t n, int b) { return (size_t)(((unsigned long long)n * (unsigned)b) / (unsigned)g_red_blocks); }

/* host_based/vector_operations.hpp:465-544 */
double orc_dot_f64(size_t n, const double* x, const double* y)
{
  if (g_red_blocks) {
    double tot = 0;
    for (int b = 0; b < g_red_blocks; ++b) {
      double t = 0;
      for (size_t i = blk_lo(n, b); i < blk_lo(n, b + 1); ++i) t += x[i] * y[i];
      tot += t;
    }
    return tot;
  }
  double temp = 0;
  long i;
  #pragma omp parallel for reduction(+: temp) if (n > 5000)
  for (i = 0; i < (long)n; ++i)
    temp += x[i] * y[i];
  return temp;
}

float orc_dot_f32(size_t n, const float* x, const float* y)
{
  if (g_red_blocks) {
    float tot = 0;
    for (int b = 0; b < g_red_blocks; ++b) {
      float t = 0;
      for (size_t i = blk_lo(n, b); i < blk_lo(n, b + 1); ++i) t += x[i] * y[i];
      tot += t;
    }
    return tot;
  }
  float temp = 0;
  long i;
  #pragma omp parallel for reduction(+: temp) if (n > 5000)
  for (i = 0; i < (long)n; ++i)
    temp += x[i] * y[i];
  return temp;
}

/* host_based/vector_operations.hpp:727-833 */
double orc_norm2_f64(size_t n, const double* x)
{
  if (g_red_blocks) return sqrt(orc_dot_f64(n, x, x));
  double temp = 0;
  long i;
  #pragma omp parallel for reduction(+: temp) if (n > 5000)
  for (i = 0; i < (long)n; ++i) {
    double d = x[i];
    temp += d * d;
  }
  return sqrt(temp);
}

float orc_norm2_f32(size_t n, const float* x)
{
  if (g_red_blocks) return sqrtf(orc_dot_f32(n, x, x));
  float temp = 0;
  long i;
  #pragma omp parallel for reduction(+: temp) if (n > 5000)
  for (i = 0; i < (long)n; ++i) {
    float d = x[i];
    temp += d * d;
  }
  return sqrtf(temp);
}

/* ------------------------------------------------- pipelined operations */

/* host_based/iterative_operations.hpp:58-103 */
void orc_pipelined_prod_csr_f64(uint32_t n_rows, const uint32_t* row_ptr, const uint32_t* col, const double* val,
                                const double* p, double* Ap, const double* r0star,
                                double* buf, size_t chunk, size_t offset)
{
  if (g_red_blocks) {
    double tA = 0, tP = 0, tR = 0;
    for (int b = 0; b < g_red_blocks; ++b) {
      double a = 0, q = 0, c = 0;
      for (size_t row = blk_lo(n_rows, b); row < blk_lo(n_rows, b + 1); ++row) {
        double dot_prod = 0;
        double vpd = p[row];
        for (size_t i = row_ptr[row]; i < row_ptr[row + 1]; ++i) dot_prod += val[i] * p[col[i]];
        Ap[row] = dot_prod;
        a += dot_prod * dot_prod;
        q += vpd * dot_prod;
        c += r0star ? dot_prod * r0star[row] : 0.0;
      }
      tA += a;
      tP += q;
      tR += c;
    }
    buf[chunk] = tA;
    buf[2 * chunk] = tP;
    if (r0star) buf[offset] = tR;
    return;
  }
  double ApAp = 0, pAp = 0, Apr0 = 0;
  long row;
  #pragma omp parallel for reduction(+: ApAp, pAp, Apr0)
  for (row = 0; row < (long)n_rows; ++row) {
    double dot_prod = 0;
    double val_p_diag = p[row];
    size_t row_end = row_ptr[row + 1];
    for (size_t i = row_ptr[row]; i < row_end; ++i)
      dot_prod += val[i] * p[col[i]];
    Ap[row] = dot_prod;
    ApAp += dot_prod * dot_prod;
    pAp += val_p_diag * dot_prod;
    Apr0 += r0star ? dot_prod * r0star[row] : 0.0;
  }
  buf[chunk] = ApAp;
  buf[2 * chunk] = pAp;
  if (r0star)
    buf[offset] = Apr0;
}

/* host_based/iterative_operations.hpp:223-292 (ceil(n/C) slices) */
void orc_pipelined_prod_sell_f64(uint32_t n_rows, uint32_t C, const uint32_t* columns_per_block,
                                 const uint32_t* column_indices, const uint32_t* block_start, const double* elements,
                                 const double* p, double* Ap, const double* r0star,
                                 double* buf, size_t chunk, size_t offset)
{
  if (g_red_blocks) {
    orc_sell_spmv_f64(n_rows, C, columns_per_block, column_indices, block_start, elements, p, 0, 1, 1.0, Ap, 0, 1, 0.0);
    buf[chunk] = orc_dot_f64(n_rows, Ap, Ap);
    buf[2 * chunk] = orc_dot_f64(n_rows, p, Ap);
    if (r0star) buf[offset] = orc_dot_f64(n_rows, Ap, r0star);
    return;
  }
  long num_blocks = (long)((n_rows + C - 1) / C);
  double ApAp = 0, pAp = 0, Apr0 = 0;
  long b;
  #pragma omp parallel for reduction(+: ApAp, pAp, Apr0)
  for (b = 0; b < num_blocks; ++b) {
    double* result_values = (double*)calloc(C, sizeof(double));
    size_t current_columns = columns_per_block[b];
    for (size_t k = 0; k < current_columns; ++k) {
      size_t stride_start = block_start[b] + k * C;
      for (uint32_t r = 0; r < C; ++r) {
        double v = elements[stride_start + r];
        result_values[r] += v ? p[column_indices[stride_start + r]] * v : 0;
      }
    }
    size_t first_row = (size_t)b * C;
    for (uint32_t r = 0; r < C; ++r) {
      size_t row = first_row + r;
      if (row < n_rows) {
        double row_result = result_values[r];
        Ap[row] = row_result;
        ApAp += row_result * row_result;
        pAp += p[row] * row_result;
        Apr0 += r0star ? row_result * r0star[row] : 0.0;
      }
    }
    free(result_values);
  }
  buf[chunk] = ApAp;
  buf[2 * chunk] = pAp;
  if (r0star)
    buf[offset] = Apr0;
}

/* host_based/iterative_operations.hpp:378-418 */
void orc_pipelined_cg_vector_update_f64(size_t n, double* x, double alpha, double* p, double* r,
                                        const double* Ap, double beta, double* buf)
{
  if (g_red_blocks) {
    for (size_t k = 0; k < n; ++k) {
      double value_p = p[k];
      double value_r = r[k];
      x[k] += alpha * value_p;
      value_r -= alpha * Ap[k];
      p[k] = value_r + beta * value_p;
      r[k] = value_r;
    }
    buf[0] = orc_dot_f64(n, r, r);
    return;
  }
  double inner_prod_r = 0;
  long i;
  #pragma omp parallel for reduction(+: inner_prod_r)
  for (i = 0; i < (long)n; ++i) {
    double value_p = p[i];
    double value_r = r[i];
    x[i] += alpha * value_p;
    value_r -= alpha * Ap[i];
    value_p = value_r + beta * value_p;
    inner_prod_r += value_r * value_r;
    p[i] = value_p;
    r[i] = value_r;
  }
  buf[0] = inner_prod_r;
}

/* host_based/iterative_operations.hpp:518-563 */
void orc_pipelined_bicgstab_update_s_f64(size_t n, double* s, const double* r, const double* Ap,
                                         double* buf, size_t chunk, size_t offset)
{
  double r_in_r0 = 0, Ap_in_r0 = 0;
  for (size_t i = 0; i < chunk; ++i) {
    r_in_r0 += buf[i];
    Ap_in_r0 += buf[i + 3 * chunk];
  }
  double alpha = r_in_r0 / Ap_in_r0;
  if (g_red_blocks) {
    for (size_t k = 0; k < n; ++k) s[k] = r[k] - alpha * Ap[k];
    buf[offset] = orc_dot_f64(n, s, s);
    return;
  }
  double inner_prod_s = 0;
  long i;
  #pragma omp parallel for reduction(+: inner_prod_s)
  for (i = 0; i < (long)n; ++i) {
    double value_s = r[i] - alpha * Ap[i];
    inner_prod_s += value_s * value_s;
    s[i] = value_s;
  }
  buf[offset] = inner_prod_s;
}

/* host_based/iterative_operations.hpp:572-621 */
void orc_pipelined_bicgstab_vector_update_f64(size_t n, double* x, double alpha, double* p, double omega,
                                              const double* s, double* r, const double* As, double beta,
                                              const double* Ap, const double* r0star, double* buf)
{
  if (g_red_blocks) {
    for (size_t k = 0; k < n; ++k) {
      double value_p = p[k];
      double value_s = s[k];
      x[k] += alpha * value_p + omega * value_s;
      double value_residual = value_s - omega * As[k];
      p[k] = value_residual + beta * (value_p - omega * Ap[k]);
      r[k] = value_residual;
    }
    buf[0] = orc_dot_f64(n, r, r0star);
    return;
  }
  double inner_prod_r_r0star = 0;
  long i;
  #pragma omp parallel for reduction(+: inner_prod_r_r0star)
  for (i = 0; i < (long)n; ++i) {
    double value_result = x[i];
    double value_p = p[i];
    double value_s = s[i];
    double value_As = As[i];
    double value_Ap = Ap[i];
    double value_r0star = r0star[i];
    value_result += alpha * value_p + omega * value_s;
    double value_residual = value_s - omega * value_As;
    value_p = value_residual + beta * (value_p - omega * value_Ap);
    inner_prod_r_r0star += value_residual * value_r0star;
    x[i] = value_result;
    r[i] = value_residual;
    p[i] = value_p;
  }
  buf[0] = inner_prod_r_r0star;
}

/* ------------------------------------------- coordinate / ELL / HYB matrices */

/* host_based/sparse_matrix_operations.hpp:1221-1246 (COO), :1502-1537 (ELL), :1872-1929 (HYB);
 * contiguous vectors. */
void orc_fmt_spmv_f64(const orc_fmt64* A, const double* x, double alpha, double* y, double beta)
{
  const size_t n = A->n_rows;
  if (A->format == ORC_FORMAT_COO) {
    for (size_t i = 0; i < n; ++i) {
      if (beta < 0 || beta > 0)
        y[i] *= beta;
      else
        y[i] = 0;
    }
    for (size_t i = 0; i < A->nnz; ++i)
      y[A->coords[2 * i]] += alpha * A->elements[i] * x[A->coords[2 * i + 1]];
    return;
  }
  for (size_t row = 0; row < n; ++row) {
    double sum = 0;
    for (size_t k = 0; k < A->width; ++k) {
      const size_t off = row + k * A->ld;
      const double v = A->elements[off];
      if (v > 0 || v < 0) sum += x[A->coords[off]] * v;
    }
    if (A->format == ORC_FORMAT_HYB)
      for (size_t e = A->csr_rows[row]; e < A->csr_rows[row + 1]; ++e) sum += x[A->csr_cols[e]] * A->csr_elements[e];
    if (beta < 0 || beta > 0)
      y[row] = alpha * sum + beta * y[row];
    else
      y[row] = alpha * sum;
  }
}

/* pipelined_prod_impl for coordinate_matrix / ell_matrix / hyb_matrix
 * (host_based/iterative_operations.hpp:113-156, :166-210, :302-365), single-threaded order. */
void orc_fmt_pipelined_prod_f64(const orc_fmt64* A, const double* p, double* Ap, const double* r0star,
                                double* buf, size_t chunk, size_t offset)
{
  const size_t n = A->n_rows;
  double ApAp = 0, pAp = 0, Apr0 = 0;
  if (A->format == ORC_FORMAT_COO) {
    for (size_t i = 0; i < n; ++i) Ap[i] = 0;
    for (size_t i = 0; i < A->nnz; ++i) Ap[A->coords[2 * i]] += A->elements[i] * p[A->coords[2 * i + 1]];
    for (size_t i = 0; i < n; ++i) {
      const double a = Ap[i], q = p[i];
      ApAp += a * a;
      pAp += a * q;
      Apr0 += r0star ? a * r0star[i] : 0.0;
    }
  } else {
    for (size_t row = 0; row < n; ++row) {
      const double pd = p[row];
      double sum = 0;
      for (size_t k = 0; k < A->width; ++k) {
        const size_t off = row + k * A->ld;
        const double v = A->elements[off];
        if (v != 0) sum += p[A->coords[off]] * v;
      }
      if (A->format == ORC_FORMAT_HYB)
        for (size_t e = A->csr_rows[row]; e < A->csr_rows[row + 1]; ++e) sum += p[A->csr_cols[e]] * A->csr_elements[e];
      Ap[row] = sum;
      ApAp += sum * sum;
      pAp += pd * sum;
      Apr0 += r0star ? sum * r0star[row] : 0.0;
    }
  }
  buf[chunk] = ApAp;
  buf[2 * chunk] = pAp;
  if (r0star) buf[offset] = Apr0;
}

/* --------------------------------------------------------------- solvers */

typedef struct orc_op {
  int is_sell;  /* 0 CSR, 1 SELL, 2 COO / ELL / HYB */
  const orc_csr64* csr;
  const orc_sell64* sell;
  const orc_fmt64* fmt;
} orc_op;

static uint32_t op_rows(const orc_op* A)
{
  return A->is_sell == 2 ? A->fmt->n_rows : A->is_sell ? A->sell->n_rows : A->csr->n_rows;
}

static void op_prod(const orc_op* A, const double* x, double* y)
{
  if (A->is_sell == 2)
    orc_fmt_spmv_f64(A->fmt, x, 1.0, y, 0.0);
  else if (A->is_sell)
    orc_sell_spmv_f64(A->sell->n_rows, A->sell->C, A->sell->columns_per_block, A->sell->column_indices,
                      A->sell->block_start, A->sell->elements, x, 0, 1, 1.0, y, 0, 1, 0.0);
  else
    orc_csr_spmv_f64(A->csr->n_rows, A->csr->row_ptr, A->csr->col, A->csr->val, x, 0, 1, 1.0, y, 0, 1, 0.0);
}

static void op_pipelined_prod(const orc_op* A, const double* p, double* Ap, const double* r0star,
                              double* buf, size_t chunk, size_t offset)
{
  if (A->is_sell == 2)
    orc_fmt_pipelined_prod_f64(A->fmt, p, Ap, r0star, buf, chunk, offset);
  else if (A->is_sell)
    orc_pipelined_prod_sell_f64(A->sell->n_rows, A->sell->C, A->sell->columns_per_block, A->sell->column_indices,
                                A->sell->block_start, A->sell->elements, p, Ap, r0star, buf, chunk, offset);
  else
    orc_pipelined_prod_csr_f64(A->csr->n_rows, A->csr->row_ptr, A->csr->col, A->csr->val, p, Ap, r0star,
                               buf, chunk, offset);
}

static double chunk_sum(const double* buf, size_t begin, size_t end)
{
  double s = 0; /* std::accumulate order */
  for (size_t i = begin; i < end; ++i) s += buf[i];
  return s;
}

/* cg.hpp:128-187 */
static orc_result cg_pipelined(const orc_op* A, const double* b, double* x, double tol, double abs_tol,
                               uint32_t max_iters, double* history)
{
  orc_result res = {0, 0.0};
  size_t n = op_rows(A);
  double* r = (double*)malloc(n * sizeof(double));
  double* p = (double*)malloc(n * sizeof(double));
  double* Ap = (double*)malloc(n * sizeof(double));
  double* buf = (double*)calloc(3 * 256, sizeof(double));
  const size_t C = 256;
  memset(x, 0, n * sizeof(double));
  memcpy(r, b, n * sizeof(double));
  memcpy(p, b, n * sizeof(double));
  op_prod(A, p, Ap);

  double norm_rhs_squared = orc_norm2_f64(n, r);
  norm_rhs_squared *= norm_rhs_squared;
  if (norm_rhs_squared <= abs_tol * abs_tol) {
    free(r); free(p); free(Ap); free(buf);
    return res;
  }
  double inner_prod_rr = norm_rhs_squared;
  double alpha = inner_prod_rr / orc_dot_f64(n, p, Ap);
  double beta = orc_norm2_f64(n, Ap);
  beta = (alpha * alpha * beta * beta - inner_prod_rr) / inner_prod_rr;
  double inner_prod_ApAp = 0, inner_prod_pAp = 0;

  for (uint32_t i = 0; i < max_iters; ++i) {
    res.iters = i + 1;
    orc_pipelined_cg_vector_update_f64(n, x, alpha, p, r, Ap, beta, buf);
    op_pipelined_prod(A, p, Ap, NULL, buf, C, 0);
    inner_prod_rr = chunk_sum(buf, 0, C);
    inner_prod_ApAp = chunk_sum(buf, C, 2 * C);
    inner_prod_pAp = chunk_sum(buf, 2 * C, 3 * C);
    if (history) history[i] = inner_prod_rr;
    if (fabs(inner_prod_rr / norm_rhs_squared) < tol * tol || fabs(inner_prod_rr) < abs_tol * abs_tol)
      break;
    alpha = inner_prod_rr / inner_prod_pAp;
    beta = (alpha * alpha * inner_prod_ApAp - inner_prod_rr) / inner_prod_rr;
  }
  res.error = sqrt(fabs(inner_prod_rr) / norm_rhs_squared);
  free(r); free(p); free(Ap); free(buf);
  return res;
}

orc_result orc_cg_pipelined_csr(const orc_csr64* A, const double* b, double* x, double tol, double abs_tol,
                                uint32_t max_iters, double* history)
{
  orc_op op = {0, A, NULL, NULL};
  return cg_pipelined(&op, b, x, tol, abs_tol, max_iters, history);
}

orc_result orc_cg_pipelined_sell(const orc_sell64* A, const double* b, double* x, double tol, double abs_tol,
                                 uint32_t max_iters, double* history)
{
  orc_op op = {1, NULL, A, NULL};
  return cg_pipelined(&op, b, x, tol, abs_tol, max_iters, history);
}

/* cg.hpp:257-322; jacobi apply = element_div (jacobi_precond.hpp:121-126).
 * With diag == NULL the z handler aliases z to the residual (cg.hpp:102-109) and
 * the new <r,z> is pow(norm_2(r), 2) (cg.hpp:301-302). */
orc_result orc_cg_classic_csr(const orc_csr64* A, const double* b, double* x, const double* diag,
                              double tol, double abs_tol, uint32_t max_iters, double* history)
{
  orc_result res = {0, 0.0};
  orc_op op = {0, A, NULL, NULL};
  size_t n = A->n_rows;
  double* r = (double*)malloc(n * sizeof(double));
  double* tmp = (double*)malloc(n * sizeof(double));
  double* zbuf = (double*)malloc(n * sizeof(double));
  double* p = (double*)malloc(n * sizeof(double));
  double* z = diag ? zbuf : r;
  memset(x, 0, n * sizeof(double));
  memcpy(r, b, n * sizeof(double));
  if (diag) {
    for (size_t i = 0; i < n; ++i) z[i] = r[i] / diag[i];
  }
  memcpy(p, z, n * sizeof(double));
  double ip_rr = orc_dot_f64(n, r, z);
  double new_ip_rr = 0;
  double norm_rhs_squared = ip_rr;
  if (fabs(norm_rhs_squared) <= abs_tol * abs_tol) {
    free(r); free(tmp); free(zbuf); free(p);
    return res;
  }
  for (uint32_t i = 0; i < max_iters; ++i) {
    res.iters = i + 1;
    op_prod(&op, p, tmp);
    double alpha = ip_rr / orc_dot_f64(n, tmp, p);
    for (size_t k = 0; k < n; ++k) x[k] = x[k] + p[k] * alpha;
    for (size_t k = 0; k < n; ++k) r[k] = r[k] + tmp[k] * (-alpha);
    if (diag) {
      for (size_t k = 0; k < n; ++k) z[k] = r[k] / diag[k];
      new_ip_rr = orc_dot_f64(n, r, z);
    } else {
      double nrm = orc_norm2_f64(n, r);
      new_ip_rr = pow(nrm, 2);
    }
    if (history) history[i] = new_ip_rr;
    double ratio = new_ip_rr / norm_rhs_squared;
    if (fabs(ratio) < tol * tol || fabs(new_ip_rr) < abs_tol * abs_tol)
      break;
    double beta = new_ip_rr / ip_rr;
    ip_rr = new_ip_rr;
    for (size_t k = 0; k < n; ++k) p[k] = z[k] + p[k] * beta;
  }
  res.error = sqrt(fabs(new_ip_rr / norm_rhs_squared));
  free(r); free(tmp); free(zbuf); free(p);
  return res;
}

/* bicgstab.hpp:97-215 */
static orc_result bicgstab_pipelined(const orc_op* A, const double* b, double* x, double tol, double abs_tol,
                                     uint32_t max_iters, double* history)
{
  orc_result res = {0, 0.0};
  size_t n = op_rows(A);
  const size_t C = 256;
  double* r = (double*)malloc(n * sizeof(double));
  double* p = (double*)malloc(n * sizeof(double));
  double* r0star = (double*)malloc(n * sizeof(double));
  double* Ap = (double*)malloc(n * sizeof(double));
  double* s = (double*)malloc(n * sizeof(double));
  double* As = (double*)malloc(n * sizeof(double));
  double* buf = (double*)calloc(6 * C, sizeof(double));
  memset(x, 0, n * sizeof(double));
  memcpy(r, b, n * sizeof(double));
  memcpy(p, b, n * sizeof(double));
  memcpy(r0star, b, n * sizeof(double));
  memcpy(Ap, b, n * sizeof(double));
  memcpy(s, b, n * sizeof(double));
  memcpy(As, b, n * sizeof(double));

  double norm_rhs_host = orc_norm2_f64(n, r);
  double residual_norm = norm_rhs_host;
  buf[0] = norm_rhs_host * norm_rhs_host;
  if (norm_rhs_host <= abs_tol) {
    free(r); free(p); free(r0star); free(Ap); free(s); free(As); free(buf);
    return res;
  }
  for (uint32_t i = 0; i < max_iters; ++i) {
    res.iters = i + 1;
    op_pipelined_prod(A, p, Ap, r0star, buf, C, 3 * C);
    orc_pipelined_bicgstab_update_s_f64(n, s, r, Ap, buf, C, 5 * C);
    op_pipelined_prod(A, s, As, r0star, buf, C, 4 * C);
    double r_dot_r0 = chunk_sum(buf, 0, C);
    double As_dot_As = chunk_sum(buf, C, 2 * C);
    double As_dot_s = chunk_sum(buf, 2 * C, 3 * C);
    double Ap_dot_r0 = chunk_sum(buf, 3 * C, 4 * C);
    double As_dot_r0 = chunk_sum(buf, 4 * C, 5 * C);
    double s_dot_s = chunk_sum(buf, 5 * C, 6 * C);
    double alpha = r_dot_r0 / Ap_dot_r0;
    double beta = -As_dot_r0 / Ap_dot_r0;
    double omega = As_dot_s / As_dot_As;
    residual_norm = sqrt(s_dot_s - 2.0 * omega * As_dot_s + omega * omega * As_dot_As);
    if (history) history[i] = residual_norm;
    if (fabs(residual_norm / norm_rhs_host) < tol || residual_norm < abs_tol)
      break;
    orc_pipelined_bicgstab_vector_update_f64(n, x, alpha, p, omega, s, r, As, beta, Ap, r0star, buf);
  }
  res.error = residual_norm / norm_rhs_host;
  free(r); free(p); free(r0star); free(Ap); free(s); free(As); free(buf);
  return res;
}

orc_result orc_bicgstab_pipelined_csr(const orc_csr64* A, const double* b, double* x, double tol, double abs_tol,
                                      uint32_t max_iters, double* history)
{
  orc_op op = {0, A, NULL, NULL};
  return bicgstab_pipelined(&op, b, x, tol, abs_tol, max_iters, history);
}

orc_result orc_bicgstab_pipelined_sell(const orc_sell64* A, const double* b, double* x, double tol, double abs_tol,
                                       uint32_t max_iters, double* history)
{
  orc_op op = {1, NULL, A, NULL};
  return bicgstab_pipelined(&op, b, x, tol, abs_tol, max_iters, history);
}

orc_result orc_cg_pipelined_fmt(const orc_fmt64* A, const double* b, double* x, double tol, double abs_tol,
                                uint32_t max_iters, double* history)
{
  orc_op op = {2, NULL, NULL, A};
  return cg_pipelined(&op, b, x, tol, abs_tol, max_iters, history);
}

orc_result orc_bicgstab_pipelined_fmt(const orc_fmt64* A, const double* b, double* x, double tol, double abs_tol,
                                      uint32_t max_iters, double* history)
{
  orc_op op = {2, NULL, NULL, A};
  return bicgstab_pipelined(&op, b, x, tol, abs_tol, max_iters, history);
}

/* bicgstab.hpp:398-489 (diag != NULL) and :295-383 (diag == NULL) */
orc_result orc_bicgstab_classic_csr(const orc_csr64* A, const double* b, double* x, const double* diag,
                                    double tol, double abs_tol, uint32_t max_iters, uint32_t max_iters_before_restart,
                                    double* history)
{
  orc_result res = {0, 0.0};
  orc_op op = {0, A, NULL, NULL};
  size_t n = A->n_rows;
  double* r = (double*)malloc(n * sizeof(double));
  double* r0star = (double*)malloc(n * sizeof(double));
  double* tmp0 = (double*)malloc(n * sizeof(double));
  double* tmp1 = (double*)malloc(n * sizeof(double));
  double* s = (double*)malloc(n * sizeof(double));
  double* p = (double*)malloc(n * sizeof(double));
  memset(x, 0, n * sizeof(double));
  memcpy(r, b, n * sizeof(double));
  memcpy(r0star, b, n * sizeof(double));
  memcpy(p, b, n * sizeof(double));
  double ip_rr0star;
  double norm_rhs_host = orc_norm2_f64(n, r);
  if (diag) {
    ip_rr0star = orc_norm2_f64(n, r); /* bicgstab.hpp:419 (overwritten at the first restart) */
  } else {
    ip_rr0star = norm_rhs_host * norm_rhs_host;
  }
  double new_ip_rr0star = 0;
  double residual_norm = norm_rhs_host;
  if (norm_rhs_host <= abs_tol) {
    free(r); free(r0star); free(tmp0); free(tmp1); free(s); free(p);
    return res;
  }
  int restart_flag = 1;
  size_t last_restart = 0;
  for (size_t i = 0; i < max_iters; ++i) {
    if (restart_flag) {
      op_prod(&op, x, r);
      for (size_t k = 0; k < n; ++k) r[k] = b[k] - r[k];
      if (diag)
        for (size_t k = 0; k < n; ++k) r[k] = r[k] / diag[k];
      memcpy(p, r, n * sizeof(double));
      memcpy(r0star, r, n * sizeof(double));
      ip_rr0star = orc_norm2_f64(n, r);
      ip_rr0star *= ip_rr0star;
      restart_flag = 0;
      last_restart = i;
    }
    res.iters = (uint32_t)(i + 1);
    op_prod(&op, p, tmp0);
    if (diag)
      for (size_t k = 0; k < n; ++k) tmp0[k] = tmp0[k] / diag[k];
    double alpha = ip_rr0star / orc_dot_f64(n, tmp0, r0star);
    for (size_t k = 0; k < n; ++k) s[k] = r[k] - tmp0[k] * alpha;
    op_prod(&op, s, tmp1);
    if (diag)
      for (size_t k = 0; k < n; ++k) tmp1[k] = tmp1[k] / diag[k];
    double norm_tmp1 = orc_norm2_f64(n, tmp1);
    double omega = orc_dot_f64(n, tmp1, s) / (norm_tmp1 * norm_tmp1);
    /* result += alpha * p + omega * s  (vector.hpp avbv_v: x + (p*alpha + s*omega)) */
    for (size_t k = 0; k < n; ++k) x[k] = x[k] + (p[k] * alpha + s[k] * omega);
    for (size_t k = 0; k < n; ++k) r[k] = s[k] - tmp1[k] * omega;
    if (diag) {
      residual_norm = orc_norm2_f64(n, r);
      if (history) history[i] = residual_norm;
      if (residual_norm / norm_rhs_host < tol || residual_norm < abs_tol)
        break;
      new_ip_rr0star = orc_dot_f64(n, r, r0star);
    } else {
      new_ip_rr0star = orc_dot_f64(n, r, r0star);
      residual_norm = orc_norm2_f64(n, r);
      if (history) history[i] = residual_norm;
      if (fabs(residual_norm / norm_rhs_host) < tol || residual_norm < abs_tol)
        break;
    }
    double beta = new_ip_rr0star / ip_rr0star * alpha / omega;
    ip_rr0star = new_ip_rr0star;
    if ((ip_rr0star <= 0 && ip_rr0star >= 0) || (omega <= 0 && omega >= 0) ||
        (i - last_restart > max_iters_before_restart))
      restart_flag = 1;
    for (size_t k = 0; k < n; ++k) p[k] = p[k] - tmp0[k] * omega;
    for (size_t k = 0; k < n; ++k) p[k] = r[k] + p[k] * beta;
  }
  res.error = residual_norm / norm_rhs_host;
  free(r); free(r0star); free(tmp0); free(tmp1); free(s); free(p);
  return res;
}

/* mixed_precision_cg.hpp:95-187 */
/* Inner product of FP32 vectors accumulated in double (sequential). */
static double dot32_acc64(size_t n, const float* x, const float* y)
{
  double t = 0;
  for (size_t i = 0; i < n; ++i) t += (double)(x[i] * y[i]);
  return t;
}

static float dot32(size_t n, const float* x, const float* y, int acc64)
{
  return acc64 ? (float)dot32_acc64(n, x, y) : orc_dot_f32(n, x, y);
}

/* mixed_precision_cg.hpp:95-187 (mode 0) and the reliable-update variant (mode 1). */
orc_result orc_mixed_cg2_csr(const orc_csr64* A, const double* b, double* x, const double* diag, double tol,
                             uint32_t max_iters, float inner_tol, int mode, int acc64, uint32_t* updates,
                             double* history)
{
  orc_result res = {0, 0.0};
  size_t n = A->n_rows;
  uint32_t nnz = A->row_ptr[n];
  uint32_t nrest = 0;
  double* residual = (double*)malloc(n * sizeof(double));
  memset(x, 0, n * sizeof(double));
  memcpy(residual, b, n * sizeof(double));
  double ip_rr = orc_dot_f64(n, b, b);
  double new_ip_rr = 0;
  double norm_rhs_squared = ip_rr;
  if (updates) *updates = 0;
  if (norm_rhs_squared <= 0) {
    free(residual);
    return res;
  }
  float* r_lo = (float*)malloc(n * sizeof(float));
  float* x_lo = (float*)calloc(n, sizeof(float));
  float* p_lo = (float*)malloc(n * sizeof(float));
  float* t_lo = (float*)malloc(n * sizeof(float));
  float* d_lo = diag ? (float*)malloc(n * sizeof(float)) : NULL;
  float* val_lo = (float*)malloc((size_t)nnz * sizeof(float));
  for (size_t k = 0; k < nnz; ++k) val_lo[k] = (float)A->val[k];
  if (d_lo)
    for (size_t k = 0; k < n; ++k) d_lo[k] = (float)diag[k];

  if (mode == 0) {
    float inner_ip_rr = (float)ip_rr;
    float new_inner_ip_rr = 0;
    float initial_inner_rhs_norm_squared = (float)ip_rr;
    for (size_t k = 0; k < n; ++k) p_lo[k] = (float)b[k];
    memcpy(r_lo, p_lo, n * sizeof(float));
    for (uint32_t i = 0; i < max_iters; ++i) {
      res.iters = i + 1;
      orc_csr_spmv_f32((uint32_t)n, A->row_ptr, A->col, val_lo, p_lo, 0, 1, 1.0f, t_lo, 0, 1, 0.0f);
      float alpha = inner_ip_rr / dot32(n, t_lo, p_lo, acc64);
      for (size_t k = 0; k < n; ++k) x_lo[k] = x_lo[k] + p_lo[k] * alpha;
      for (size_t k = 0; k < n; ++k) r_lo[k] = r_lo[k] + t_lo[k] * (-alpha);
      new_inner_ip_rr = dot32(n, r_lo, r_lo, acc64);
      if (history) history[i] = new_inner_ip_rr;
      float beta = new_inner_ip_rr / inner_ip_rr;
      inner_ip_rr = new_inner_ip_rr;
      for (size_t k = 0; k < n; ++k) p_lo[k] = r_lo[k] + p_lo[k] * beta;

      if (new_inner_ip_rr < inner_tol * initial_inner_rhs_norm_squared || i == max_iters - 1) {
        ++nrest;
        for (size_t k = 0; k < n; ++k) residual[k] = (double)x_lo[k];
        for (size_t k = 0; k < n; ++k) x[k] = x[k] + residual[k];
        orc_csr_spmv_f64((uint32_t)n, A->row_ptr, A->col, A->val, x, 0, 1, 1.0, residual, 0, 1, 0.0);
        for (size_t k = 0; k < n; ++k) residual[k] = b[k] - residual[k];
        new_ip_rr = orc_dot_f64(n, residual, residual);
        if (new_ip_rr / norm_rhs_squared < tol * tol)
          break;
        for (size_t k = 0; k < n; ++k) p_lo[k] = (float)residual[k];
        memset(x_lo, 0, n * sizeof(float));
        memcpy(r_lo, p_lo, n * sizeof(float));
        initial_inner_rhs_norm_squared = (float)new_ip_rr;
        inner_ip_rr = (float)new_ip_rr;
      }
    }
  } else {
    /* Pipelined CG in FP32 (the order of the HIP kernels: update, product, scalars). */
    const float* dd = d_lo;
    for (size_t k = 0; k < n; ++k) {
      r_lo[k] = (float)b[k];
      p_lo[k] = dd ? r_lo[k] / dd[k] : r_lo[k];
    }
    double rz = 0, rr = 0;
    for (size_t k = 0; k < n; ++k) {
      rz += (double)(r_lo[k] * p_lo[k]);
      rr += (double)(r_lo[k] * r_lo[k]);
    }
    double rr_ref = rr;
    orc_csr_spmv_f32((uint32_t)n, A->row_ptr, A->col, val_lo, p_lo, 0, 1, 1.0f, t_lo, 0, 1, 0.0f);
    double ApAp = 0, pAp = 0;
    for (size_t k = 0; k < n; ++k) {
      float s = t_lo[k];
      ApAp += (double)(dd ? s * (s / dd[k]) : s * s);
      pAp += (double)(p_lo[k] * s);
    }
    double alpha = rz / pAp;
    double beta = (alpha * alpha * ApAp - rz) / rz;
    for (uint32_t i = 0; i < max_iters; ++i) {
      res.iters = i + 1;
      const float fa = (float)alpha, fb = (float)beta;
      rz = 0;
      rr = 0;
      for (size_t k = 0; k < n; ++k) {
        x_lo[k] += fa * p_lo[k];
        r_lo[k] -= fa * t_lo[k];
        float z = dd ? r_lo[k] / dd[k] : r_lo[k];
        p_lo[k] = z + fb * p_lo[k];
        rz += (double)(r_lo[k] * z);
        rr += (double)(r_lo[k] * r_lo[k]);
      }
      orc_csr_spmv_f32((uint32_t)n, A->row_ptr, A->col, val_lo, p_lo, 0, 1, 1.0f, t_lo, 0, 1, 0.0f);
      ApAp = 0;
      pAp = 0;
      for (size_t k = 0; k < n; ++k) {
        float s = t_lo[k];
        ApAp += (double)(dd ? s * (s / dd[k]) : s * s);
        pAp += (double)(p_lo[k] * s);
      }
      if (history) history[i] = rr;
      if (rr < (double)inner_tol * rr_ref || rr / norm_rhs_squared < tol * tol || i == max_iters - 1) {
        ++nrest;
        for (size_t k = 0; k < n; ++k) {
          x[k] = x[k] + (double)x_lo[k];
          x_lo[k] = 0.0f;
        }
        orc_csr_spmv_f64((uint32_t)n, A->row_ptr, A->col, A->val, x, 0, 1, 1.0, residual, 0, 1, 0.0);
        new_ip_rr = 0;
        for (size_t k = 0; k < n; ++k) {
          residual[k] = b[k] - residual[k];
          new_ip_rr += residual[k] * residual[k];
        }
        if (new_ip_rr / norm_rhs_squared < tol * tol || i == max_iters - 1)
          break;
        rz = 0;
        rr = 0;
        for (size_t k = 0; k < n; ++k) {
          float rn = (float)residual[k];
          float zo = dd ? r_lo[k] / dd[k] : r_lo[k];
          float zn = dd ? rn / dd[k] : rn;
          p_lo[k] = p_lo[k] + (zn - zo);
          r_lo[k] = rn;
          rz += (double)(rn * zn);
          rr += (double)(rn * rn);
        }
        rr_ref = rr;
        orc_csr_spmv_f32((uint32_t)n, A->row_ptr, A->col, val_lo, p_lo, 0, 1, 1.0f, t_lo, 0, 1, 0.0f);
        ApAp = 0;
        pAp = 0;
        for (size_t k = 0; k < n; ++k) {
          float s = t_lo[k];
          ApAp += (double)(dd ? s * (s / dd[k]) : s * s);
          pAp += (double)(p_lo[k] * s);
        }
      }
      alpha = rz / pAp;
      beta = (alpha * alpha * ApAp - rz) / rz;
    }
  }
  res.error = sqrt(new_ip_rr / norm_rhs_squared);
  if (updates) *updates = nrest;
  free(residual); free(r_lo); free(x_lo); free(p_lo); free(t_lo); free(val_lo); free(d_lo);
  return res;
}

orc_result orc_mixed_cg_csr(const orc_csr64* A, const double* b, double* x, double tol, uint32_t max_iters,
                            float inner_tol, uint32_t* restarts)
{
  return orc_mixed_cg2_csr(A, b, x, NULL, tol, max_iters, inner_tol, 0, 0, restarts, NULL);
}

/* Independent classic CG per diagonal block (cg.hpp:257-322 with no preconditioner). */
/* Pipelined GMRES (gmres.hpp:182-370, no preconditioner) with the host kernels of
 * host_based/iterative_operations.hpp:734-937.  gs_skip_tail != 0 reproduces the reference's Gram-
 * Schmidt stage 1 with one thread, whose range [0, (n-1)/T*T) leaves out the last entry
 * (:820-822); 0 computes the full inner products.  iters counts the Krylov steps that entered the
 * error estimator, as tag.iters() does. */
orc_result orc_gmres_pipelined_csr(const orc_csr64* A, const double* b, double* x, double tol, double abs_tol,
                                   uint32_t max_iters, uint32_t krylov_dim, int gs_skip_tail, double* history)
{
  orc_result res = {0, 0.0};
  const size_t n = A->n_rows;
  const size_t K = krylov_dim;
  const size_t C = 128; /* buffer_size_per_vector */
  uint32_t max_restarts = max_iters / krylov_dim;
  if (max_restarts > 0 && max_restarts * krylov_dim == max_iters) max_restarts -= 1;
  double* r = (double*)malloc(n * sizeof(double));
  double* V = (double*)calloc(K * n, sizeof(double));
  double* R = (double*)calloc(K * K, sizeof(double));
  double* buf = (double*)calloc(3 * C, sizeof(double));
  double* r_dot_vk = (double*)calloc(C * K, sizeof(double));
  double* vi_in_vk = (double*)calloc(C * K, sizeof(double));
  double* xi = (double*)calloc(K, sizeof(double));
  double* eta = (double*)calloc(K, sizeof(double));
  double* coef = (double*)calloc(K, sizeof(double));
  double* h = (double*)calloc(K, sizeof(double));
  memcpy(r, b, n * sizeof(double));
  memset(x, 0, n * sizeof(double));
  const double norm_rhs = orc_norm2_f64(n, r);
  double rho_0 = norm_rhs, rho = 1.0;
  const size_t gs_end = gs_skip_tail ? (n > 0 ? n - 1 : 0) : n;

  for (uint32_t restart = 0; restart <= max_restarts; ++restart) {
    if (restart > 0) {
      orc_csr_spmv_f64((uint32_t)n, A->row_ptr, A->col, A->val, x, 0, 1, 1.0, r, 0, 1, 0.0);
      for (size_t i = 0; i < n; ++i) r[i] = b[i] - r[i];
      rho_0 = orc_norm2_f64(n, r);
    }
    if (rho_0 <= abs_tol) break;
    for (size_t i = 0; i < n; ++i) r[i] = r[i] / rho_0;
    rho = 1.0;
    if (rho_0 / norm_rhs < tol || rho_0 < abs_tol) break;

    size_t k;
    for (k = 0; k < K; ++k) {
      double* vk = V + k * n;
      /* v_k = A v_{k-1} (v_{-1} = r); <v_k, v_k> into buf[C] (pipelined_gmres_prod = pipelined_cg_prod) */
      orc_pipelined_prod_csr_f64((uint32_t)n, A->row_ptr, A->col, A->val, k == 0 ? r : V + (k - 1) * n, vk, NULL,
                                 buf, C, 0);
      if (k > 0) {
        /* stage 1: <v_i, v_k>, i < k */
        for (size_t j = 0; j < k; ++j) h[j] = 0.0;
        for (size_t i = 0; i < gs_end; ++i) {
          const double vki = vk[i];
          for (size_t j = 0; j < k; ++j) h[j] += V[j * n + i] * vki;
        }
        for (size_t j = 0; j < k; ++j) vi_in_vk[j * C] = h[j];
        /* stage 2: reduce, v_k -= sum_j <v_j,v_k> v_j, ||v_k||^2 */
        for (size_t j = 0; j < k; ++j) {
          double t = 0.0;
          for (size_t c = 0; c < C; ++c) t += vi_in_vk[j * C + c];
          h[j] = t;
        }
        double nsq = 0.0;
        for (size_t i = 0; i < n; ++i) {
          double v = vk[i];
          for (size_t j = 0; j < k; ++j) v -= h[j] * V[j * n + i];
          nsq += v * v;
          vk[i] = v;
        }
        for (size_t j = 0; j < k; ++j) R[j + k * K] = h[j];
        buf[C] = nsq;
      }
      /* normalize: ||v_k|| from buf[C..2C), v_k /= ||v_k||, <r, v_k> */
      double nv = 0.0;
      for (size_t c = 0; c < C; ++c) nv += buf[c + C];
      nv = sqrt(nv);
      R[k * K + k] = nv;
      double rv = 0.0;
      for (size_t i = 0; i < n; ++i) {
        const double v = vk[i] / nv;
        rv += r[i] * v;
        vk[i] = v;
      }
      r_dot_vk[k * C] = rv;
    }
    for (size_t i = 0; i < k; ++i) {
      double t = 0.0;
      for (size_t c = 0; c < C; ++c) t += r_dot_vk[i * C + c];
      xi[i] = t;
    }
    const size_t full = k;
    for (size_t i = 0; i < k; ++i)
      if (fabs(R[i + i * k]) < tol * R[0]) {
        k = i;
        break;
      }
    for (size_t i = 0; i < k; ++i) {
      res.iters += 1;
      if (xi[i] >= rho || xi[i] <= -rho) {
        if (history && res.iters <= max_iters) history[res.iters - 1] = fabs(rho * rho_0 / norm_rhs);
        k = i;
        break;
      }
      rho *= sin(acos(xi[i] / rho));
      if (history && res.iters <= max_iters) history[res.iters - 1] = fabs(rho * rho_0 / norm_rhs);
    }
    for (size_t i = 0; i < k; ++i) eta[i] = xi[i];
    for (long i2 = (long)k - 1; i2 > -1; --i2) {
      const size_t i = (size_t)i2;
      for (size_t j = i + 1; j < k; ++j) eta[i] -= R[i + j * full] * eta[j];
      eta[i] /= R[i + i * full];
    }
    for (size_t i = 0; i < k; ++i) coef[i] = rho_0 * eta[i];
    for (size_t i = 0; i < n; ++i) {
      double v = x[i];
      v += coef[0] * r[i];
      for (size_t j = 1; j < k; ++j) v += coef[j] * V[(j - 1) * n + i];
      x[i] = v;
    }
    res.error = fabs(rho * rho_0 / norm_rhs);
  }
  free(r); free(V); free(R); free(buf); free(r_dot_vk); free(vi_in_vk); free(xi); free(eta); free(coef); free(h);
  return res;
}

void orc_gmres_normalize_vk(size_t n, double* v_k, const double* residual, double* R, size_t offset_in_R,
                            const double* buf, double* r_dot_vk, size_t chunk, size_t chunk_offset)
{
  double norm_vk = 0;
  for (size_t i = 0; i < chunk; ++i) norm_vk += buf[i + chunk];
  norm_vk = sqrt(norm_vk);
  R[offset_in_R] = norm_vk;
  double rv = 0;
  for (size_t i = 0; i < n; ++i) {
    const double v = v_k[i] / norm_vk;
    rv += residual[i] * v;
    v_k[i] = v;
  }
  r_dot_vk[chunk_offset] = rv;
}

void orc_gmres_stage1(const double* basis, size_t n, size_t ld, size_t k, double* vi_in_vk, size_t chunk,
                      int skip_tail)
{
  const size_t end = skip_tail ? (n > 0 ? n - 1 : 0) : n;
  for (size_t j = 0; j < k; ++j) {
    double t = 0;
    for (size_t i = 0; i < end; ++i) t += basis[i + j * ld] * basis[i + k * ld];
    vi_in_vk[j * chunk] = t;
  }
}

void orc_gmres_stage2(double* basis, size_t n, size_t ld, size_t k, const double* vi_in_vk, double* R,
                      size_t krylov_dim, double* buf, size_t chunk)
{
  double* h = (double*)calloc(k ? k : 1, sizeof(double));
  for (size_t i = 0; i < k; ++i)
    for (size_t j = 0; j < chunk; ++j) h[i] += vi_in_vk[i * chunk + j];
  double nsq = 0;
  for (size_t i = 0; i < n; ++i) {
    double v = basis[i + k * ld];
    for (size_t j = 0; j < k; ++j) v -= h[j] * basis[i + j * ld];
    nsq += v * v;
    basis[i + k * ld] = v;
  }
  for (size_t i = 0; i < k; ++i) R[i + k * krylov_dim] = h[i];
  buf[chunk] = nsq;
  free(h);
}

void orc_gmres_update_result(double* x, const double* residual, const double* basis, size_t n, size_t ld,
                             const double* coef, size_t k)
{
  for (size_t i = 0; i < n; ++i) {
    double v = x[i];
    v += coef[0] * residual[i];
    for (size_t j = 1; j < k; ++j) v += coef[j] * basis[i + (j - 1) * ld];
    x[i] = v;
  }
}

/* Pipelined CG (cg.hpp:128-187) on FP32 vectors, inner products accumulated in double. */
static orc_result cg_pipelined_f32(uint32_t n, const uint32_t* rp, const uint32_t* cl, const float* val,
                                   const float* b, float* x, double tol, uint32_t max_iters)
{
  orc_result res = {0, 0.0};
  float* r = (float*)malloc((n ? n : 1) * sizeof(float));
  float* p = (float*)malloc((n ? n : 1) * sizeof(float));
  float* Ap = (float*)malloc((n ? n : 1) * sizeof(float));
  double bb = 0, ApAp = 0, pAp = 0;
  for (uint32_t i = 0; i < n; ++i) {
    x[i] = 0.0f;
    r[i] = b[i];
    p[i] = b[i];
    bb += (double)(b[i] * b[i]);
  }
  orc_csr_spmv_f32(n, rp, cl, val, p, 0, 1, 1.0f, Ap, 0, 1, 0.0f);
  for (uint32_t i = 0; i < n; ++i) {
    ApAp += (double)(Ap[i] * Ap[i]);
    pAp += (double)(p[i] * Ap[i]);
  }
  double nrm = sqrt(bb);
  double nrs = nrm * nrm;
  double rho = nrs;
  if (nrs > 0) {
    double alpha = nrs / pAp;
    double nap = sqrt(ApAp);
    double beta = (alpha * alpha * nap * nap - nrs) / nrs;
    for (uint32_t it = 0; it < max_iters; ++it) {
      float ta = (float)alpha, tb = (float)beta;
      double rr = 0;
      for (uint32_t i = 0; i < n; ++i) {
        x[i] += ta * p[i];
        r[i] -= ta * Ap[i];
        p[i] = r[i] + tb * p[i];
        rr += (double)(r[i] * r[i]);
      }
      orc_csr_spmv_f32(n, rp, cl, val, p, 0, 1, 1.0f, Ap, 0, 1, 0.0f);
      ApAp = 0;
      pAp = 0;
      for (uint32_t i = 0; i < n; ++i) {
        ApAp += (double)(Ap[i] * Ap[i]);
        pAp += (double)(p[i] * Ap[i]);
      }
      res.iters = it + 1;
      rho = rr;
      if (fabs(rho / nrs) < tol * tol) break;
      alpha = rho / pAp;
      beta = (alpha * alpha * ApAp - rho) / rho;
    }
  }
  res.error = nrs > 0 ? sqrt(fabs(rho) / nrs) : 0.0;
  free(r); free(p); free(Ap);
  return res;
}

/* Independent pipelined CG (cg.hpp:128-187) per diagonal block; single != 0 runs it on FP32 copies
 * of the values and b (inner products accumulated in double) and returns x rounded back. */
void orc_batched_cg_csr(uint32_t num_systems, const uint32_t* sys_ptr, const orc_csr64* A, const double* b,
                        double* x, double tol, uint32_t max_iters, uint32_t* iters, double* errors)
{
  orc_batched_cg2_csr(num_systems, sys_ptr, A, b, x, tol, max_iters, 0, iters, errors);
}

void orc_batched_cg2_csr(uint32_t num_systems, const uint32_t* sys_ptr, const orc_csr64* A, const double* b,
                         double* x, double tol, uint32_t max_iters, int single, uint32_t* iters, double* errors)
{
  for (uint32_t s = 0; s < num_systems; ++s) {
    uint32_t r0 = sys_ptr[s], r1 = sys_ptr[s + 1];
    uint32_t n = r1 - r0;
    uint32_t base = A->row_ptr[r0];
    uint32_t* rp = (uint32_t*)malloc((n + 1) * sizeof(uint32_t));
    uint32_t nnz = A->row_ptr[r1] - base;
    uint32_t* cl = (uint32_t*)malloc((nnz ? nnz : 1) * sizeof(uint32_t));
    for (uint32_t k = 0; k <= n; ++k) rp[k] = A->row_ptr[r0 + k] - base;
    for (uint32_t k = 0; k < nnz; ++k) cl[k] = A->col[base + k] - r0;
    orc_result res;
    if (single) {
      float* v32 = (float*)malloc((nnz ? nnz : 1) * sizeof(float));
      float* b32 = (float*)malloc((n ? n : 1) * sizeof(float));
      float* x32 = (float*)malloc((n ? n : 1) * sizeof(float));
      for (uint32_t k = 0; k < nnz; ++k) v32[k] = (float)A->val[base + k];
      for (uint32_t k = 0; k < n; ++k) b32[k] = (float)b[r0 + k];
      res = cg_pipelined_f32(n, rp, cl, v32, b32, x32, tol, max_iters);
      for (uint32_t k = 0; k < n; ++k) x[r0 + k] = (double)x32[k];
      free(v32); free(b32); free(x32);
    } else {
      orc_csr64 sub = {n, rp, cl, A->val + base};
      res = orc_cg_pipelined_csr(&sub, b + r0, x + r0, tol, 0.0, max_iters, NULL);
    }
    iters[s] = res.iters;
    errors[s] = res.error;
    free(rp);
    free(cl);
  }
}

void orc_adjoint_grad_csr(uint32_t n_rows, const uint32_t* row_ptr, const uint32_t* col,
                          const double* lambda, const double* x, double* dA)
{
  for (uint32_t row = 0; row < n_rows; ++row)
    for (uint32_t k = row_ptr[row]; k < row_ptr[row + 1]; ++k)
      dA[k] = -lambda[row] * x[col[k]];
}
