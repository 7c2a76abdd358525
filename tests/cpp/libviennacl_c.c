/* libviennacl_c.c -- the libviennacl-style C API (include/viennacl_hip.h) used from plain C99, the
 * way the reference's libviennacl tests call its C functions: BLAS 1 on offset/strided device
 * vectors against host loops, and CSR products and solvers on a 2-D Laplacian checked by the host
 * residual.  Prints one FAIL line per failed check and ALL PASSED at the end. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "vclhip.h"
#include "viennacl_hip.h"

static int failures = 0;
#define CHECK(cond, ...)                                  \
  do {                                                    \
    if (!(cond)) {                                        \
      printf("FAIL %s:%d: ", __FILE__, __LINE__);         \
      printf(__VA_ARGS__);                                \
      printf("\n");                                       \
      ++failures;                                         \
    }                                                     \
  } while (0)

static void* to_dev(const void* host, size_t bytes)
{
  void* d = NULL;
  if (vclhip_malloc(&d, bytes) || vclhip_memcpy_h2d(d, host, bytes, NULL) || vclhip_stream_synchronize(NULL)) {
    printf("device allocation failed: %s\n", vclhip_last_error());
    exit(3);
  }
  return d;
}

static void from_dev(void* host, const void* dev, size_t bytes)
{
  if (vclhip_memcpy_d2h(host, dev, bytes, NULL) || vclhip_stream_synchronize(NULL)) {
    printf("copy failed: %s\n", vclhip_last_error());
    exit(3);
  }
}

static double maxdiff(const double* a, const double* b, int n)
{
  double m = 0;
  for (int i = 0; i < n; ++i) m = fmax(m, fabs(a[i] - b[i]));
  return m;
}

static void test_blas1(ViennaCLHIPBackend be)
{
  enum { N = 1000, OFFX = 3, INCX = 2, OFFY = 5, INCY = 3 };
  const int lx = OFFX + N * INCX, ly = OFFY + N * INCY;
  double* hx = malloc(sizeof(double) * lx);
  double* hy = malloc(sizeof(double) * ly);
  double* tx = malloc(sizeof(double) * lx);
  double* ty = malloc(sizeof(double) * ly);
  for (int i = 0; i < lx; ++i) hx[i] = sin(0.7 * i) * (i % 7 == 3 ? 3.0 : 1.0);
  for (int i = 0; i < ly; ++i) hy[i] = cos(0.3 * i);
  double* dx = to_dev(hx, sizeof(double) * lx);
  double* dy = to_dev(hy, sizeof(double) * ly);

  /* amax: the first index of the largest |x_i| */
  ViennaCLInt idx = -1;
  CHECK(ViennaCLHIPiDamax(be, N, &idx, dx, OFFX, INCX) == ViennaCLHIPSuccess, "iDamax");
  int best = 0;
  for (int i = 1; i < N; ++i)
    if (fabs(hx[OFFX + i * INCX]) > fabs(hx[OFFX + best * INCX])) best = i;
  CHECK(idx == best, "iDamax %d vs %d", idx, best);
  /* asum, dot, nrm2 */
  double asum = 0, dot = 0, nrm = 0, r = 0;
  for (int i = 0; i < N; ++i) {
    asum += fabs(hx[OFFX + i * INCX]);
    dot += hx[OFFX + i * INCX] * hy[OFFY + i * INCY];
    nrm += hx[OFFX + i * INCX] * hx[OFFX + i * INCX];
  }
  nrm = sqrt(nrm);
  CHECK(ViennaCLHIPDasum(be, N, &r, dx, OFFX, INCX) == 0 && fabs(r - asum) <= 1e-12 * asum, "Dasum %g %g", r, asum);
  CHECK(ViennaCLHIPDdot(be, N, &r, dx, OFFX, INCX, dy, OFFY, INCY) == 0 && fabs(r - dot) <= 1e-12 * N, "Ddot");
  CHECK(ViennaCLHIPDnrm2(be, N, &r, dx, OFFX, INCX) == 0 && fabs(r - nrm) <= 1e-12 * nrm, "Dnrm2");
  /* axpy, scal, rot, swap, copy: compare the whole arrays (untouched entries included) */
  memcpy(tx, hx, sizeof(double) * lx);
  memcpy(ty, hy, sizeof(double) * ly);
  for (int i = 0; i < N; ++i) ty[OFFY + i * INCY] += 0.25 * tx[OFFX + i * INCX];
  CHECK(ViennaCLHIPDaxpy(be, N, 0.25, dx, OFFX, INCX, dy, OFFY, INCY) == 0, "Daxpy");
  for (int i = 0; i < N; ++i) tx[OFFX + i * INCX] *= -1.5;
  CHECK(ViennaCLHIPDscal(be, N, -1.5, dx, OFFX, INCX) == 0, "Dscal");
  for (int i = 0; i < N; ++i) {
    const double a = tx[OFFX + i * INCX], b = ty[OFFY + i * INCY];
    tx[OFFX + i * INCX] = 0.6 * a + 0.8 * b;
    ty[OFFY + i * INCY] = 0.6 * b - 0.8 * a;
  }
  CHECK(ViennaCLHIPDrot(be, N, dx, OFFX, INCX, dy, OFFY, INCY, 0.6, 0.8) == 0, "Drot");
  for (int i = 0; i < N; ++i) {
    const double a = tx[OFFX + i * INCX];
    tx[OFFX + i * INCX] = ty[OFFY + i * INCY];
    ty[OFFY + i * INCY] = a;
  }
  CHECK(ViennaCLHIPDswap(be, N, dx, OFFX, INCX, dy, OFFY, INCY) == 0, "Dswap");
  from_dev(hx, dx, sizeof(double) * lx);
  from_dev(hy, dy, sizeof(double) * ly);
  CHECK(maxdiff(hx, tx, lx) == 0 && maxdiff(hy, ty, ly) == 0, "axpy/scal/rot/swap %g %g", maxdiff(hx, tx, lx),
        maxdiff(hy, ty, ly));
  for (int i = 0; i < N; ++i) ty[OFFY + i * INCY] = tx[OFFX + i * INCX];
  CHECK(ViennaCLHIPDcopy(be, N, dx, OFFX, INCX, dy, OFFY, INCY) == 0, "Dcopy");
  from_dev(hy, dy, sizeof(double) * ly);
  CHECK(maxdiff(hy, ty, ly) == 0, "Dcopy values");

  /* single precision: amax and dot on contiguous vectors */
  float fx[N], fy[N];
  for (int i = 0; i < N; ++i) {
    fx[i] = (float)(i % 17) - 8.0f;
    fy[i] = 0.5f;
  }
  fx[611] = -40.0f;
  float* dfx = to_dev(fx, sizeof(fx));
  float* dfy = to_dev(fy, sizeof(fy));
  CHECK(ViennaCLHIPiSamax(be, N, &idx, dfx, 0, 1) == 0 && idx == 611, "iSamax %d", idx);
  float fr = 0;
  CHECK(ViennaCLHIPSasum(be, N, &fr, dfx, 0, 1) == 0 && fr > 0, "Sasum");
  CHECK(ViennaCLHIPSaxpy(be, N, 2.0f, dfx, 0, 1, dfy, 0, 1) == 0, "Saxpy");
  from_dev(fy, dfy, sizeof(fy));
  CHECK(fy[611] == -79.5f && fy[0] == -15.5f, "Saxpy values %g %g", fy[611], fy[0]);
  /* all zeros: the reference returns the start offset */
  memset(fx, 0, sizeof(fx));
  vclhip_memcpy_h2d(dfx, fx, sizeof(fx), NULL);
  CHECK(ViennaCLHIPiSamax(be, N - 10, &idx, dfx, 7, 1) == 0 && idx == 7, "iSamax of zeros %d", idx);
  vclhip_free(dfx);
  vclhip_free(dfy);
  vclhip_free(dx);
  vclhip_free(dy);
  free(hx);
  free(hy);
  free(tx);
  free(ty);
}

static void test_sparse(ViennaCLHIPBackend be)
{
  /* 2-D Laplacian, 60 x 50, CSR built on the host */
  enum { PX = 60, PY = 50, N = PX * PY };
  unsigned* rp = malloc(sizeof(unsigned) * (N + 1));
  unsigned* col = malloc(sizeof(unsigned) * 5 * N);
  double* val = malloc(sizeof(double) * 5 * N);
  float* valf = malloc(sizeof(float) * 5 * N);
  unsigned nnz = 0;
  for (int j = 0; j < PY; ++j)
    for (int i = 0; i < PX; ++i) {
      const int r = j * PX + i;
      rp[r] = nnz;
      if (j > 0) { col[nnz] = r - PX; val[nnz++] = -1; }
      if (i > 0) { col[nnz] = r - 1; val[nnz++] = -1; }
      col[nnz] = r; val[nnz++] = 4;
      if (i + 1 < PX) { col[nnz] = r + 1; val[nnz++] = -1; }
      if (j + 1 < PY) { col[nnz] = r + PX; val[nnz++] = -1; }
    }
  rp[N] = nnz;
  for (unsigned k = 0; k < nnz; ++k) valf[k] = (float)val[k];
  double* b = malloc(sizeof(double) * N);
  double* x = malloc(sizeof(double) * N);
  double* y = malloc(sizeof(double) * N);
  for (int i = 0; i < N; ++i) b[i] = sin(0.01 * i * i);
  void* drp = to_dev(rp, sizeof(unsigned) * (N + 1));
  void* dcol = to_dev(col, sizeof(unsigned) * nnz);
  void* dval = to_dev(val, sizeof(double) * nnz);
  void* dvalf = to_dev(valf, sizeof(float) * nnz);
  double* db = to_dev(b, sizeof(double) * N);
  double* dx = to_dev(b, sizeof(double) * N);
  ViennaCLHIPCsr A = NULL, Af = NULL;
  CHECK(ViennaCLHIPCsrCreate(be, ViennaCLHIPDouble, N, N, (ViennaCLInt)nnz, drp, dcol, dval, &A) == 0, "CsrCreate %s",
        ViennaCLHIPLastError());
  CHECK(ViennaCLHIPCsrCreate(be, ViennaCLHIPFloat, N, N, (ViennaCLInt)nnz, drp, dcol, dvalf, &Af) == 0, "CsrCreate f");
  /* y = 2 A b - y */
  double* dy = to_dev(b, sizeof(double) * N);
  CHECK(ViennaCLHIPDcsrmv(be, A, 2.0, db, 0, 1, -1.0, dy, 0, 1) == 0, "Dcsrmv");
  from_dev(y, dy, sizeof(double) * N);
  double err = 0;
  for (int r = 0; r < N; ++r) {
    double s = 0;
    for (unsigned k = rp[r]; k < rp[r + 1]; ++k) s += val[k] * b[col[k]];
    err = fmax(err, fabs(y[r] - (2 * s - b[r])));
  }
  CHECK(err <= 1e-13, "Dcsrmv values %g", err);
  /* the wrong precision is refused */
  CHECK(ViennaCLHIPScsrmv(be, A, 1.0f, (float*)db, 0, 1, 0.0f, (float*)dy, 0, 1) == ViennaCLHIPGenericFailure &&
        strlen(ViennaCLHIPLastError()) > 0, "precision mismatch");
  /* solvers: true residual on the host */
  ViennaCLInt iters = 0;
  double est = 0;
  for (int kind = 0; kind < 3; ++kind) {
    ViennaCLHIPStatus st = kind == 0 ? ViennaCLHIPDcg(be, A, db, dx, 1e-10, 2000, 0, &iters, &est)
                         : kind == 1 ? ViennaCLHIPDbicgstab(be, A, db, dx, 1e-10, 2000, 1, &iters, &est)
                                     : ViennaCLHIPDgmres(be, A, db, dx, 1e-10, 3000, 30, &iters, &est);
    from_dev(x, dx, sizeof(double) * N);
    double rn = 0, bn = 0;
    for (int r = 0; r < N; ++r) {
      double s = 0;
      for (unsigned k = rp[r]; k < rp[r + 1]; ++k) s += val[k] * x[col[k]];
      rn += (b[r] - s) * (b[r] - s);
      bn += b[r] * b[r];
    }
    CHECK(st == 0 && sqrt(rn / bn) < 1e-8 && iters > 0, "solver %d: status %d, %d iterations, residual %g", kind,
          (int)st, iters, sqrt(rn / bn));
  }
  /* single precision CG */
  float* fb = malloc(sizeof(float) * N);
  for (int i = 0; i < N; ++i) fb[i] = (float)b[i];
  float* dfb = to_dev(fb, sizeof(float) * N);
  float* dfx = to_dev(fb, sizeof(float) * N);
  CHECK(ViennaCLHIPScg(be, Af, dfb, dfx, 1e-5, 2000, 0, &iters, &est) == 0 && est < 1e-5, "Scg %g", est);
  CHECK(ViennaCLHIPSgmres(be, Af, dfb, dfx, 1e-5, 2000, 20, &iters, &est) == 0 && est < 1e-4, "Sgmres %g", est);
  CHECK(ViennaCLHIPCsrDestroy(&A) == 0 && A == NULL && ViennaCLHIPCsrDestroy(&Af) == 0, "CsrDestroy");
  vclhip_free(drp);
  vclhip_free(dcol);
  vclhip_free(dval);
  vclhip_free(dvalf);
  vclhip_free(db);
  vclhip_free(dx);
  vclhip_free(dy);
  vclhip_free(dfb);
  vclhip_free(dfx);
  free(rp);
  free(col);
  free(val);
  free(valf);
  free(b);
  free(x);
  free(y);
  free(fb);
}

int main(void)
{
  int ndev = 0;
  if (vclhip_device_count(&ndev) != 0 || ndev == 0) {
    printf("no HIP device\n");
    return 2;
  }
  ViennaCLHIPBackend be = NULL;
  CHECK(ViennaCLHIPBackendCreate(&be) == ViennaCLHIPSuccess && be, "BackendCreate");
  test_blas1(be);
  test_sparse(be);
  CHECK(ViennaCLHIPBackendDestroy(&be) == 0 && be == NULL, "BackendDestroy");
  if (failures) {
    printf("%d FAILED\n", failures);
    return 1;
  }
  printf("ALL PASSED\n");
  return 0;
}
