/*
 * viennacl_hip.h -- a libviennacl-style C API for the HIP path (the pattern of the reference's
 * libviennacl/include/viennacl.hpp:49-618: a backend handle, BLAS-1 functions per precision that
 * take (pointer, offset, increment) vectors, status codes), extended with CSR sparse products and
 * the Krylov solvers so that C, MATLAB MEX or any FFI caller reaches the HIP kernels without C++.
 *
 * Vectors and matrix arrays are device pointers (vclhip_malloc or any HIP allocation).  Every
 * function returns ViennaCLHIPSuccess (0) or ViennaCLHIPGenericFailure (1), the values of the
 * reference's ViennaCLStatus (viennacl.hpp:98-102); ViennaCLHIPLastError() gives the reason.
 * Work is ordered on the backend's stream (the null stream unless set); scalar results are
 * returned to the host synchronously.
 */
#ifndef VIENNACL_HIP_H
#define VIENNACL_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#ifndef VIENNACL_VIENNACL_HPP
typedef int ViennaCLInt; /* viennacl.hpp:49 */
#endif

typedef enum { ViennaCLHIPSuccess = 0, ViennaCLHIPGenericFailure } ViennaCLHIPStatus;
typedef enum { ViennaCLHIPInvalidPrecision, ViennaCLHIPFloat, ViennaCLHIPDouble } ViennaCLHIPPrecision;

typedef struct ViennaCLHIPBackend_impl* ViennaCLHIPBackend;
typedef struct ViennaCLHIPCsr_impl* ViennaCLHIPCsr;

/* ---- backend (ViennaCLBackendCreate / ViennaCLBackendDestroy, viennacl.hpp:111-113) */
ViennaCLHIPStatus ViennaCLHIPBackendCreate(ViennaCLHIPBackend* backend);
ViennaCLHIPStatus ViennaCLHIPBackendSetStream(ViennaCLHIPBackend backend, void* stream);
ViennaCLHIPStatus ViennaCLHIPBackendDestroy(ViennaCLHIPBackend* backend);
const char* ViennaCLHIPLastError(void);

/* ---- BLAS 1 (the ViennaCLCUDA* family of viennacl.hpp:134-390 on HIP device pointers) */
/* index of the first largest |x_i| (ViennaCLCUDAiSamax, viennacl.hpp:136-141) */
ViennaCLHIPStatus ViennaCLHIPiSamax(ViennaCLHIPBackend backend, ViennaCLInt n, ViennaCLInt* index, float* x,
                                    ViennaCLInt offx, ViennaCLInt incx);
ViennaCLHIPStatus ViennaCLHIPiDamax(ViennaCLHIPBackend backend, ViennaCLInt n, ViennaCLInt* index, double* x,
                                    ViennaCLInt offx, ViennaCLInt incx);
/* sum |x_i| (ViennaCLCUDASasum, viennacl.hpp:164-169) */
ViennaCLHIPStatus ViennaCLHIPSasum(ViennaCLHIPBackend backend, ViennaCLInt n, float* alpha, float* x,
                                   ViennaCLInt offx, ViennaCLInt incx);
ViennaCLHIPStatus ViennaCLHIPDasum(ViennaCLHIPBackend backend, ViennaCLInt n, double* alpha, double* x,
                                   ViennaCLInt offx, ViennaCLInt incx);
/* y += alpha x (ViennaCLCUDASaxpy, viennacl.hpp:193-200) */
ViennaCLHIPStatus ViennaCLHIPSaxpy(ViennaCLHIPBackend backend, ViennaCLInt n, float alpha, float* x, ViennaCLInt offx,
                                   ViennaCLInt incx, float* y, ViennaCLInt offy, ViennaCLInt incy);
ViennaCLHIPStatus ViennaCLHIPDaxpy(ViennaCLHIPBackend backend, ViennaCLInt n, double alpha, double* x,
                                   ViennaCLInt offx, ViennaCLInt incx, double* y, ViennaCLInt offy, ViennaCLInt incy);
/* y = x (ViennaCLCUDAScopy, viennacl.hpp:227-233) */
ViennaCLHIPStatus ViennaCLHIPScopy(ViennaCLHIPBackend backend, ViennaCLInt n, float* x, ViennaCLInt offx,
                                   ViennaCLInt incx, float* y, ViennaCLInt offy, ViennaCLInt incy);
ViennaCLHIPStatus ViennaCLHIPDcopy(ViennaCLHIPBackend backend, ViennaCLInt n, double* x, ViennaCLInt offx,
                                   ViennaCLInt incx, double* y, ViennaCLInt offy, ViennaCLInt incy);
/* x . y (ViennaCLCUDASdot, viennacl.hpp:254-262) */
ViennaCLHIPStatus ViennaCLHIPSdot(ViennaCLHIPBackend backend, ViennaCLInt n, float* alpha, float* x, ViennaCLInt offx,
                                  ViennaCLInt incx, float* y, ViennaCLInt offy, ViennaCLInt incy);
ViennaCLHIPStatus ViennaCLHIPDdot(ViennaCLHIPBackend backend, ViennaCLInt n, double* alpha, double* x,
                                  ViennaCLInt offx, ViennaCLInt incx, double* y, ViennaCLInt offy, ViennaCLInt incy);
/* ||x||_2 (ViennaCLCUDASnrm2, viennacl.hpp:287-292) */
ViennaCLHIPStatus ViennaCLHIPSnrm2(ViennaCLHIPBackend backend, ViennaCLInt n, float* alpha, float* x,
                                   ViennaCLInt offx, ViennaCLInt incx);
ViennaCLHIPStatus ViennaCLHIPDnrm2(ViennaCLHIPBackend backend, ViennaCLInt n, double* alpha, double* x,
                                   ViennaCLInt offx, ViennaCLInt incx);
/* x' = c x + s y, y' = c y - s x (ViennaCLCUDASrot, viennacl.hpp:316-325) */
ViennaCLHIPStatus ViennaCLHIPSrot(ViennaCLHIPBackend backend, ViennaCLInt n, float* x, ViennaCLInt offx,
                                  ViennaCLInt incx, float* y, ViennaCLInt offy, ViennaCLInt incy, float c, float s);
ViennaCLHIPStatus ViennaCLHIPDrot(ViennaCLHIPBackend backend, ViennaCLInt n, double* x, ViennaCLInt offx,
                                  ViennaCLInt incx, double* y, ViennaCLInt offy, ViennaCLInt incy, double c,
                                  double s);
/* x = alpha x (ViennaCLCUDASscal, viennacl.hpp:351-356) */
ViennaCLHIPStatus ViennaCLHIPSscal(ViennaCLHIPBackend backend, ViennaCLInt n, float alpha, float* x, ViennaCLInt offx,
                                   ViennaCLInt incx);
ViennaCLHIPStatus ViennaCLHIPDscal(ViennaCLHIPBackend backend, ViennaCLInt n, double alpha, double* x,
                                   ViennaCLInt offx, ViennaCLInt incx);
/* x <-> y (ViennaCLCUDASswap, viennacl.hpp:379-386) */
ViennaCLHIPStatus ViennaCLHIPSswap(ViennaCLHIPBackend backend, ViennaCLInt n, float* x, ViennaCLInt offx,
                                   ViennaCLInt incx, float* y, ViennaCLInt offy, ViennaCLInt incy);
ViennaCLHIPStatus ViennaCLHIPDswap(ViennaCLHIPBackend backend, ViennaCLInt n, double* x, ViennaCLInt offx,
                                   ViennaCLInt incx, double* y, ViennaCLInt offy, ViennaCLInt incy);

/* ---- sparse (no counterpart in libviennacl; the C view of compressed_matrix, prod and solve) */
/* Wraps caller-owned device CSR arrays (row_ptr: rows + 1, col and values: nnz; 32-bit indices) of
 * the given precision and builds the kernels' row-block plan (compressed_matrix.hpp:740-781,
 * :1152-1188).  The arrays must outlive the handle. */
ViennaCLHIPStatus ViennaCLHIPCsrCreate(ViennaCLHIPBackend backend, ViennaCLHIPPrecision precision, ViennaCLInt rows,
                                       ViennaCLInt cols, ViennaCLInt nnz, const void* row_ptr, const void* col,
                                       const void* values, ViennaCLHIPCsr* matrix);
ViennaCLHIPStatus ViennaCLHIPCsrDestroy(ViennaCLHIPCsr* matrix);
/* y = alpha A x + beta y (prod_impl, linalg/sparse_matrix_operations.hpp:90-121) */
ViennaCLHIPStatus ViennaCLHIPScsrmv(ViennaCLHIPBackend backend, ViennaCLHIPCsr A, float alpha, float* x,
                                    ViennaCLInt offx, ViennaCLInt incx, float beta, float* y, ViennaCLInt offy,
                                    ViennaCLInt incy);
ViennaCLHIPStatus ViennaCLHIPDcsrmv(ViennaCLHIPBackend backend, ViennaCLHIPCsr A, double alpha, double* x,
                                    ViennaCLInt offx, ViennaCLInt incx, double beta, double* y, ViennaCLInt offy,
                                    ViennaCLInt incy);
/* x = solve(A, b, cg_tag(tol, max_iters) [, jacobi_precond]) (cg.hpp:338-375); jacobi != 0 selects
 * the Jacobi preconditioner; *iters and *error receive tag.iters() and tag.error(). */
ViennaCLHIPStatus ViennaCLHIPScg(ViennaCLHIPBackend backend, ViennaCLHIPCsr A, float* b, float* x, double tol,
                                 ViennaCLInt max_iters, ViennaCLInt jacobi, ViennaCLInt* iters, double* error);
ViennaCLHIPStatus ViennaCLHIPDcg(ViennaCLHIPBackend backend, ViennaCLHIPCsr A, double* b, double* x, double tol,
                                 ViennaCLInt max_iters, ViennaCLInt jacobi, ViennaCLInt* iters, double* error);
/* x = solve(A, b, bicgstab_tag(tol, max_iters) [, jacobi_precond]) (bicgstab.hpp:493-529) */
ViennaCLHIPStatus ViennaCLHIPSbicgstab(ViennaCLHIPBackend backend, ViennaCLHIPCsr A, float* b, float* x, double tol,
                                       ViennaCLInt max_iters, ViennaCLInt jacobi, ViennaCLInt* iters, double* error);
ViennaCLHIPStatus ViennaCLHIPDbicgstab(ViennaCLHIPBackend backend, ViennaCLHIPCsr A, double* b, double* x, double tol,
                                       ViennaCLInt max_iters, ViennaCLInt jacobi, ViennaCLInt* iters, double* error);
/* x = solve(A, b, gmres_tag(tol, max_iters, krylov_dim)) (gmres.hpp:374-383) */
ViennaCLHIPStatus ViennaCLHIPSgmres(ViennaCLHIPBackend backend, ViennaCLHIPCsr A, float* b, float* x, double tol,
                                    ViennaCLInt max_iters, ViennaCLInt krylov_dim, ViennaCLInt* iters,
                                    double* error);
ViennaCLHIPStatus ViennaCLHIPDgmres(ViennaCLHIPBackend backend, ViennaCLHIPCsr A, double* b, double* x, double tol,
                                    ViennaCLInt max_iters, ViennaCLInt krylov_dim, ViennaCLInt* iters,
                                    double* error);

#ifdef __cplusplus
}
#endif

#endif /* VIENNACL_HIP_H */
