// libviennacl.cpp -- the libviennacl-style C API of include/viennacl_hip.h on top of the vclhip_*
// entry points.  A backend holds the device and the stream; a CSR handle holds the descriptor and
// the row-block plan built from the caller's row pointers.
#include <cstdint>
#include <string>
#include <vector>

#include "internal.hpp"
#include "viennacl_hip.h"

struct ViennaCLHIPBackend_impl {
  int device = 0;
  void* stream = nullptr;
};

struct ViennaCLHIPCsr_impl {
  ViennaCLHIPPrecision precision = ViennaCLHIPInvalidPrecision;
  vclhip_csr desc{};
  uint32_t* blocks_dev = nullptr;
};

namespace {

ViennaCLHIPStatus status(int st) { return st == VCLHIP_SUCCESS ? ViennaCLHIPSuccess : ViennaCLHIPGenericFailure; }

ViennaCLHIPStatus fail(const char* what)
{
  vclhip::set_error("%s", what);
  return ViennaCLHIPGenericFailure;
}

bool vec_ok(ViennaCLInt n, ViennaCLInt off, ViennaCLInt inc) { return n >= 0 && off >= 0 && inc > 0; }

// Makes the backend's device current.
int enter(ViennaCLHIPBackend b)
{
  if (!b) {
    vclhip::set_error("backend is NULL");
    return VCLHIP_ERROR_INVALID_VALUE;
  }
  return vclhip_set_device(b->device);
}

vclhip_solver_params params(double tol, ViennaCLInt max_iters)
{
  vclhip_solver_params P;
  vclhip_solver_params_default(&P);
  P.tol = tol;
  P.max_iters = static_cast<uint32_t>(max_iters);
  return P;
}

ViennaCLHIPStatus report(int st, const vclhip_solver_result& R, ViennaCLInt* iters, double* error)
{
  if (st) return ViennaCLHIPGenericFailure;
  if (iters) *iters = static_cast<ViennaCLInt>(R.iters);
  if (error) *error = R.error;
  return ViennaCLHIPSuccess;
}

bool csr_ok(ViennaCLHIPCsr A, ViennaCLHIPPrecision p)
{
  if (!A) return vclhip::set_error("matrix is NULL"), false;
  if (A->precision != p) return vclhip::set_error("matrix precision does not match the function"), false;
  return true;
}

}  // namespace

extern "C" {

ViennaCLHIPStatus ViennaCLHIPBackendCreate(ViennaCLHIPBackend* backend)
{
  if (!backend) return fail("backend is NULL");
  ViennaCLHIPBackend b = new ViennaCLHIPBackend_impl();
  if (vclhip_get_device(&b->device) != VCLHIP_SUCCESS) {
    delete b;
    *backend = nullptr;
    return ViennaCLHIPGenericFailure;
  }
  *backend = b;
  return ViennaCLHIPSuccess;
}

ViennaCLHIPStatus ViennaCLHIPBackendSetStream(ViennaCLHIPBackend backend, void* stream)
{
  if (!backend) return fail("backend is NULL");
  backend->stream = stream;
  return ViennaCLHIPSuccess;
}

ViennaCLHIPStatus ViennaCLHIPBackendDestroy(ViennaCLHIPBackend* backend)
{
  if (!backend) return fail("backend is NULL");
  delete *backend;
  *backend = nullptr;
  return ViennaCLHIPSuccess;
}

const char* ViennaCLHIPLastError(void) { return vclhip_last_error(); }

// ---- BLAS 1

#define VCLHIP_VEC1(b, n, off, inc)                              \
  do {                                                           \
    if (!vec_ok(n, off, inc)) return fail("invalid vector size"); \
    if (enter(b)) return ViennaCLHIPGenericFailure;              \
  } while (0)

#define VCLHIP_VEC2(b, n, ox, ix, oy, iy)                                           \
  do {                                                                              \
    if (!vec_ok(n, ox, ix) || !vec_ok(n, oy, iy)) return fail("invalid vector size"); \
    if (enter(b)) return ViennaCLHIPGenericFailure;                                 \
  } while (0)

ViennaCLHIPStatus ViennaCLHIPiSamax(ViennaCLHIPBackend b, ViennaCLInt n, ViennaCLInt* index, float* x, ViennaCLInt offx,
                                    ViennaCLInt incx)
{
  VCLHIP_VEC1(b, n, offx, incx);
  if (!index) return fail("index is NULL");
  uint32_t i = 0;
  const int st = vclhip_iamax_f32(uint32_t(n), x, uint32_t(offx), uint32_t(incx), &i, b->stream);
  if (!st) *index = static_cast<ViennaCLInt>(i);
  return status(st);
}

ViennaCLHIPStatus ViennaCLHIPiDamax(ViennaCLHIPBackend b, ViennaCLInt n, ViennaCLInt* index, double* x,
                                    ViennaCLInt offx, ViennaCLInt incx)
{
  VCLHIP_VEC1(b, n, offx, incx);
  if (!index) return fail("index is NULL");
  uint32_t i = 0;
  const int st = vclhip_iamax_f64(uint32_t(n), x, uint32_t(offx), uint32_t(incx), &i, b->stream);
  if (!st) *index = static_cast<ViennaCLInt>(i);
  return status(st);
}

ViennaCLHIPStatus ViennaCLHIPSasum(ViennaCLHIPBackend b, ViennaCLInt n, float* alpha, float* x, ViennaCLInt offx,
                                   ViennaCLInt incx)
{
  VCLHIP_VEC1(b, n, offx, incx);
  return status(vclhip_asum_f32(uint32_t(n), x, uint32_t(offx), uint32_t(incx), alpha, b->stream));
}

ViennaCLHIPStatus ViennaCLHIPDasum(ViennaCLHIPBackend b, ViennaCLInt n, double* alpha, double* x, ViennaCLInt offx,
                                   ViennaCLInt incx)
{
  VCLHIP_VEC1(b, n, offx, incx);
  return status(vclhip_asum_f64(uint32_t(n), x, uint32_t(offx), uint32_t(incx), alpha, b->stream));
}

ViennaCLHIPStatus ViennaCLHIPSaxpy(ViennaCLHIPBackend b, ViennaCLInt n, float alpha, float* x, ViennaCLInt offx,
                                   ViennaCLInt incx, float* y, ViennaCLInt offy, ViennaCLInt incy)
{
  VCLHIP_VEC2(b, n, offx, incx, offy, incy);
  return status(vclhip_avbv_f32(uint32_t(n), y, uint32_t(offy), uint32_t(incy), y, uint32_t(offy), uint32_t(incy), 1.0f,
                                x, uint32_t(offx), uint32_t(incx), alpha, b->stream));
}

ViennaCLHIPStatus ViennaCLHIPDaxpy(ViennaCLHIPBackend b, ViennaCLInt n, double alpha, double* x, ViennaCLInt offx,
                                   ViennaCLInt incx, double* y, ViennaCLInt offy, ViennaCLInt incy)
{
  VCLHIP_VEC2(b, n, offx, incx, offy, incy);
  return status(vclhip_avbv_f64(uint32_t(n), y, uint32_t(offy), uint32_t(incy), y, uint32_t(offy), uint32_t(incy), 1.0,
                                x, uint32_t(offx), uint32_t(incx), alpha, b->stream));
}

ViennaCLHIPStatus ViennaCLHIPScopy(ViennaCLHIPBackend b, ViennaCLInt n, float* x, ViennaCLInt offx, ViennaCLInt incx,
                                   float* y, ViennaCLInt offy, ViennaCLInt incy)
{
  VCLHIP_VEC2(b, n, offx, incx, offy, incy);
  return status(
      vclhip_copy_f32(uint32_t(n), x, uint32_t(offx), uint32_t(incx), y, uint32_t(offy), uint32_t(incy), b->stream));
}

ViennaCLHIPStatus ViennaCLHIPDcopy(ViennaCLHIPBackend b, ViennaCLInt n, double* x, ViennaCLInt offx, ViennaCLInt incx,
                                   double* y, ViennaCLInt offy, ViennaCLInt incy)
{
  VCLHIP_VEC2(b, n, offx, incx, offy, incy);
  return status(
      vclhip_copy_f64(uint32_t(n), x, uint32_t(offx), uint32_t(incx), y, uint32_t(offy), uint32_t(incy), b->stream));
}

ViennaCLHIPStatus ViennaCLHIPSdot(ViennaCLHIPBackend b, ViennaCLInt n, float* alpha, float* x, ViennaCLInt offx,
                                  ViennaCLInt incx, float* y, ViennaCLInt offy, ViennaCLInt incy)
{
  VCLHIP_VEC2(b, n, offx, incx, offy, incy);
  return status(vclhip_dot_f32(uint32_t(n), x, uint32_t(offx), uint32_t(incx), y, uint32_t(offy), uint32_t(incy), alpha,
                               b->stream));
}

ViennaCLHIPStatus ViennaCLHIPDdot(ViennaCLHIPBackend b, ViennaCLInt n, double* alpha, double* x, ViennaCLInt offx,
                                  ViennaCLInt incx, double* y, ViennaCLInt offy, ViennaCLInt incy)
{
  VCLHIP_VEC2(b, n, offx, incx, offy, incy);
  return status(vclhip_dot_f64(uint32_t(n), x, uint32_t(offx), uint32_t(incx), y, uint32_t(offy), uint32_t(incy), alpha,
                               b->stream));
}

ViennaCLHIPStatus ViennaCLHIPSnrm2(ViennaCLHIPBackend b, ViennaCLInt n, float* alpha, float* x, ViennaCLInt offx,
                                   ViennaCLInt incx)
{
  VCLHIP_VEC1(b, n, offx, incx);
  return status(vclhip_norm2_f32(uint32_t(n), x, uint32_t(offx), uint32_t(incx), alpha, b->stream));
}

ViennaCLHIPStatus ViennaCLHIPDnrm2(ViennaCLHIPBackend b, ViennaCLInt n, double* alpha, double* x, ViennaCLInt offx,
                                   ViennaCLInt incx)
{
  VCLHIP_VEC1(b, n, offx, incx);
  return status(vclhip_norm2_f64(uint32_t(n), x, uint32_t(offx), uint32_t(incx), alpha, b->stream));
}

ViennaCLHIPStatus ViennaCLHIPSrot(ViennaCLHIPBackend b, ViennaCLInt n, float* x, ViennaCLInt offx, ViennaCLInt incx,
                                  float* y, ViennaCLInt offy, ViennaCLInt incy, float c, float s)
{
  VCLHIP_VEC2(b, n, offx, incx, offy, incy);
  return status(vclhip_rot_f32(uint32_t(n), x, uint32_t(offx), uint32_t(incx), y, uint32_t(offy), uint32_t(incy), c, s,
                               b->stream));
}

ViennaCLHIPStatus ViennaCLHIPDrot(ViennaCLHIPBackend b, ViennaCLInt n, double* x, ViennaCLInt offx, ViennaCLInt incx,
                                  double* y, ViennaCLInt offy, ViennaCLInt incy, double c, double s)
{
  VCLHIP_VEC2(b, n, offx, incx, offy, incy);
  return status(vclhip_rot_f64(uint32_t(n), x, uint32_t(offx), uint32_t(incx), y, uint32_t(offy), uint32_t(incy), c, s,
                               b->stream));
}

ViennaCLHIPStatus ViennaCLHIPSscal(ViennaCLHIPBackend b, ViennaCLInt n, float alpha, float* x, ViennaCLInt offx,
                                   ViennaCLInt incx)
{
  VCLHIP_VEC1(b, n, offx, incx);
  return status(vclhip_scal_f32(uint32_t(n), x, uint32_t(offx), uint32_t(incx), alpha, b->stream));
}

ViennaCLHIPStatus ViennaCLHIPDscal(ViennaCLHIPBackend b, ViennaCLInt n, double alpha, double* x, ViennaCLInt offx,
                                   ViennaCLInt incx)
{
  VCLHIP_VEC1(b, n, offx, incx);
  return status(vclhip_scal_f64(uint32_t(n), x, uint32_t(offx), uint32_t(incx), alpha, b->stream));
}

ViennaCLHIPStatus ViennaCLHIPSswap(ViennaCLHIPBackend b, ViennaCLInt n, float* x, ViennaCLInt offx, ViennaCLInt incx,
                                   float* y, ViennaCLInt offy, ViennaCLInt incy)
{
  VCLHIP_VEC2(b, n, offx, incx, offy, incy);
  return status(
      vclhip_swap_f32(uint32_t(n), x, uint32_t(offx), uint32_t(incx), y, uint32_t(offy), uint32_t(incy), b->stream));
}

ViennaCLHIPStatus ViennaCLHIPDswap(ViennaCLHIPBackend b, ViennaCLInt n, double* x, ViennaCLInt offx, ViennaCLInt incx,
                                   double* y, ViennaCLInt offy, ViennaCLInt incy)
{
  VCLHIP_VEC2(b, n, offx, incx, offy, incy);
  return status(
      vclhip_swap_f64(uint32_t(n), x, uint32_t(offx), uint32_t(incx), y, uint32_t(offy), uint32_t(incy), b->stream));
}

// ---- sparse

ViennaCLHIPStatus ViennaCLHIPCsrCreate(ViennaCLHIPBackend b, ViennaCLHIPPrecision precision, ViennaCLInt rows,
                                       ViennaCLInt cols, ViennaCLInt nnz, const void* row_ptr, const void* col,
                                       const void* values, ViennaCLHIPCsr* matrix)
{
  if (!matrix) return fail("matrix is NULL");
  *matrix = nullptr;
  if (precision != ViennaCLHIPFloat && precision != ViennaCLHIPDouble) return fail("invalid precision");
  if (rows <= 0 || cols <= 0 || nnz < 0 || !row_ptr || (nnz && (!col || !values))) return fail("invalid CSR arrays");
  if (enter(b)) return ViennaCLHIPGenericFailure;
  std::vector<uint32_t> rp(static_cast<size_t>(rows) + 1), blocks(static_cast<size_t>(rows) + 1);
  int st = vclhip_memcpy_d2h(rp.data(), row_ptr, rp.size() * sizeof(uint32_t), b->stream);
  if (!st) st = vclhip_stream_synchronize(b->stream);
  if (st) return ViennaCLHIPGenericFailure;
  if (rp[0] != 0 || rp[static_cast<size_t>(rows)] != static_cast<uint32_t>(nnz)) return fail("row_ptr does not match nnz");
  uint32_t nb = 0;
  if (vclhip_csr_plan_blocks(rp.data(), uint32_t(rows), blocks.data(), &nb)) return ViennaCLHIPGenericFailure;
  ViennaCLHIPCsr A = new ViennaCLHIPCsr_impl();
  void* dev = nullptr;
  st = vclhip_malloc(&dev, (nb + 1) * sizeof(uint32_t));
  if (!st) st = vclhip_memcpy_h2d(dev, blocks.data(), (nb + 1) * sizeof(uint32_t), b->stream);
  if (!st) st = vclhip_stream_synchronize(b->stream);
  if (st) {
    if (dev) vclhip_free(dev);
    delete A;
    return ViennaCLHIPGenericFailure;
  }
  A->precision = precision;
  A->blocks_dev = static_cast<uint32_t*>(dev);
  A->desc.n_rows = uint32_t(rows);
  A->desc.n_cols = uint32_t(cols);
  A->desc.nnz = uint32_t(nnz);
  A->desc.num_blocks = nb;
  A->desc.row_ptr = static_cast<const uint32_t*>(row_ptr);
  A->desc.col_idx = static_cast<const uint32_t*>(col);
  A->desc.values = values;
  A->desc.blocks = A->blocks_dev;
  *matrix = A;
  return ViennaCLHIPSuccess;
}

ViennaCLHIPStatus ViennaCLHIPCsrDestroy(ViennaCLHIPCsr* matrix)
{
  if (!matrix) return fail("matrix is NULL");
  if (*matrix) {
    if ((*matrix)->blocks_dev) vclhip_free((*matrix)->blocks_dev);
    delete *matrix;
  }
  *matrix = nullptr;
  return ViennaCLHIPSuccess;
}

ViennaCLHIPStatus ViennaCLHIPScsrmv(ViennaCLHIPBackend b, ViennaCLHIPCsr A, float alpha, float* x, ViennaCLInt offx,
                                    ViennaCLInt incx, float beta, float* y, ViennaCLInt offy, ViennaCLInt incy)
{
  if (!csr_ok(A, ViennaCLHIPFloat)) return ViennaCLHIPGenericFailure;
  VCLHIP_VEC2(b, 1, offx, incx, offy, incy);
  return status(vclhip_csr_spmv_f32(&A->desc, x, uint32_t(offx), uint32_t(incx), alpha, y, uint32_t(offy),
                                    uint32_t(incy), beta, b->stream));
}

ViennaCLHIPStatus ViennaCLHIPDcsrmv(ViennaCLHIPBackend b, ViennaCLHIPCsr A, double alpha, double* x, ViennaCLInt offx,
                                    ViennaCLInt incx, double beta, double* y, ViennaCLInt offy, ViennaCLInt incy)
{
  if (!csr_ok(A, ViennaCLHIPDouble)) return ViennaCLHIPGenericFailure;
  VCLHIP_VEC2(b, 1, offx, incx, offy, incy);
  return status(vclhip_csr_spmv_f64(&A->desc, x, uint32_t(offx), uint32_t(incx), alpha, y, uint32_t(offy),
                                    uint32_t(incy), beta, b->stream));
}

ViennaCLHIPStatus ViennaCLHIPScg(ViennaCLHIPBackend b, ViennaCLHIPCsr A, float* rhs, float* x, double tol,
                                 ViennaCLInt max_iters, ViennaCLInt jacobi, ViennaCLInt* iters, double* error)
{
  if (!csr_ok(A, ViennaCLHIPFloat) || max_iters < 0) return ViennaCLHIPGenericFailure;
  if (enter(b)) return ViennaCLHIPGenericFailure;
  vclhip_solver_params P = params(tol, max_iters);
  vclhip_solver_result R;
  const int st = vclhip_cg_solve_csr_f32(&A->desc, rhs, x, jacobi ? VCLHIP_PRECOND_JACOBI : VCLHIP_PRECOND_NONE, &P,
                                         &R, b->stream);
  return report(st, R, iters, error);
}

ViennaCLHIPStatus ViennaCLHIPDcg(ViennaCLHIPBackend b, ViennaCLHIPCsr A, double* rhs, double* x, double tol,
                                 ViennaCLInt max_iters, ViennaCLInt jacobi, ViennaCLInt* iters, double* error)
{
  if (!csr_ok(A, ViennaCLHIPDouble) || max_iters < 0) return ViennaCLHIPGenericFailure;
  if (enter(b)) return ViennaCLHIPGenericFailure;
  vclhip_solver_params P = params(tol, max_iters);
  vclhip_solver_result R;
  const int st = vclhip_cg_solve_csr_f64(&A->desc, rhs, x, jacobi ? VCLHIP_PRECOND_JACOBI : VCLHIP_PRECOND_NONE, &P,
                                         &R, b->stream);
  return report(st, R, iters, error);
}

ViennaCLHIPStatus ViennaCLHIPSbicgstab(ViennaCLHIPBackend b, ViennaCLHIPCsr A, float* rhs, float* x, double tol,
                                       ViennaCLInt max_iters, ViennaCLInt jacobi, ViennaCLInt* iters, double* error)
{
  if (!csr_ok(A, ViennaCLHIPFloat) || max_iters < 0) return ViennaCLHIPGenericFailure;
  if (enter(b)) return ViennaCLHIPGenericFailure;
  vclhip_solver_params P = params(tol, max_iters);
  vclhip_solver_result R;
  const int st = vclhip_bicgstab_solve_csr_f32(&A->desc, rhs, x,
                                               jacobi ? VCLHIP_PRECOND_JACOBI : VCLHIP_PRECOND_NONE, &P, &R,
                                               b->stream);
  return report(st, R, iters, error);
}

ViennaCLHIPStatus ViennaCLHIPDbicgstab(ViennaCLHIPBackend b, ViennaCLHIPCsr A, double* rhs, double* x, double tol,
                                       ViennaCLInt max_iters, ViennaCLInt jacobi, ViennaCLInt* iters, double* error)
{
  if (!csr_ok(A, ViennaCLHIPDouble) || max_iters < 0) return ViennaCLHIPGenericFailure;
  if (enter(b)) return ViennaCLHIPGenericFailure;
  vclhip_solver_params P = params(tol, max_iters);
  vclhip_solver_result R;
  const int st = vclhip_bicgstab_solve_csr_f64(&A->desc, rhs, x,
                                               jacobi ? VCLHIP_PRECOND_JACOBI : VCLHIP_PRECOND_NONE, &P, &R,
                                               b->stream);
  return report(st, R, iters, error);
}

ViennaCLHIPStatus ViennaCLHIPSgmres(ViennaCLHIPBackend b, ViennaCLHIPCsr A, float* rhs, float* x, double tol,
                                    ViennaCLInt max_iters, ViennaCLInt krylov_dim, ViennaCLInt* iters, double* error)
{
  if (!csr_ok(A, ViennaCLHIPFloat) || max_iters < 0 || krylov_dim <= 0) return ViennaCLHIPGenericFailure;
  if (enter(b)) return ViennaCLHIPGenericFailure;
  vclhip_solver_params P = params(tol, max_iters);
  P.krylov_dim = static_cast<uint32_t>(krylov_dim);
  vclhip_solver_result R;
  return report(vclhip_gmres_solve_csr_f32(&A->desc, rhs, x, &P, &R, b->stream), R, iters, error);
}

ViennaCLHIPStatus ViennaCLHIPDgmres(ViennaCLHIPBackend b, ViennaCLHIPCsr A, double* rhs, double* x, double tol,
                                    ViennaCLInt max_iters, ViennaCLInt krylov_dim, ViennaCLInt* iters, double* error)
{
  if (!csr_ok(A, ViennaCLHIPDouble) || max_iters < 0 || krylov_dim <= 0) return ViennaCLHIPGenericFailure;
  if (enter(b)) return ViennaCLHIPGenericFailure;
  vclhip_solver_params P = params(tol, max_iters);
  P.krylov_dim = static_cast<uint32_t>(krylov_dim);
  vclhip_solver_result R;
  return report(vclhip_gmres_solve_csr_f64(&A->desc, rhs, x, &P, &R, b->stream), R, iters, error);
}

}  // extern "C"
