"""Pins the oracle against what the mathematics fixes, independently of the reference:
closed-form Laplacian eigenpairs and row sums, dense numpy products and solves, finite
termination of CG, the Chronopoulos-Gear recurrence, and central finite differences.
Each check is chosen so that a dropped term, a wrong sign or index, or a transposed operand in
the oracle fails it; test_planted_mistakes_are_caught demonstrates this on the CPU.
"""
import numpy as np
import pytest


def dense(rp, col, val, n=None):
    n = rp.size - 1 if n is None else n
    A = np.zeros((rp.size - 1, n))
    for i in range(rp.size - 1):
        for k in range(rp[i], rp[i + 1]):
            A[i, col[k]] += val[k]
    return A


def test_2d_laplacian_eigenpair(oracle, vh):
    px, py = 23, 17
    rp, col, val = vh.fdm_laplace_2d(px, py)
    j, k = 3, 5
    i_idx = np.arange(px)[None, :]
    j_idx = np.arange(py)[:, None]
    v = (np.sin((i_idx + 1) * j * np.pi / (px + 1)) * np.sin((j_idx + 1) * k * np.pi / (py + 1))).ravel()
    lam = 4 - 2 * np.cos(j * np.pi / (px + 1)) - 2 * np.cos(k * np.pi / (py + 1))
    y = oracle.csr_spmv(rp, col, val, v)
    assert np.max(np.abs(y - lam * v)) < 1e-13


def test_3d_laplacian_eigenpair_and_row_sums(oracle, vh):
    nx, ny, nz = 9, 7, 6
    rp, col, val = vh.laplace_3d(nx, ny, nz)
    a, b, c = 2, 3, 1
    X, Y, Z = np.meshgrid(np.arange(nx) + 1, np.arange(ny) + 1, np.arange(nz) + 1, indexing="ij")
    v = (np.sin(X * a * np.pi / (nx + 1)) * np.sin(Y * b * np.pi / (ny + 1)) * np.sin(Z * c * np.pi / (nz + 1)))
    v = v.transpose(2, 1, 0).ravel()  # row = x + nx (y + ny z)
    lam = sum(2 - 2 * np.cos(m * np.pi / (N + 1)) for m, N in ((a, nx), (b, ny), (c, nz)))
    assert np.max(np.abs(oracle.csr_spmv(rp, col, val, v) - lam * v)) < 1e-13
    # A * 1 = number of boundary faces of each cell
    ones = oracle.csr_spmv(rp, col, val, np.ones(rp.size - 1))
    bnd = ((X == 1).astype(int) + (X == nx) + (Y == 1) + (Y == ny) + (Z == 1) + (Z == nz)).transpose(2, 1, 0).ravel()
    assert np.array_equal(ones, bnd.astype(float))


def test_spmv_alpha_beta_strides_against_dense(oracle, vh):
    rp, col, val = vh.powerlaw(300, max_len=40, seed=9)
    A = dense(rp, col, val)
    x = vh.random_uniform(1, 300)
    y0 = vh.random_uniform(2, 300)
    y = oracle.csr_spmv(rp, col, val, x, -0.75, y0, 2.5)
    assert np.allclose(y, -0.75 * A @ x + 2.5 * y0, rtol=1e-13, atol=1e-13)
    # strided operands: x read at 3i+1, y written at 2i+5
    xs = np.zeros(3 * 300 + 1)
    xs[1::3] = x
    ys = np.full(2 * 300 + 6, 7.0)
    ys = oracle.csr_spmv(rp, col, val, xs, 1.0, ys, 0.0, xs=1, xi=3, ys=5, yi=2)
    assert np.allclose(ys[5::2][:300], A @ x, rtol=1e-13, atol=1e-13)
    assert np.all(ys[:5] == 7.0) and np.all(ys[6::2] == 7.0)


def test_sell_equals_csr_bitwise(oracle, vh):
    for C_ in (32, 64, 128):
        rp, col, val = vh.powerlaw(777, max_len=60, seed=C_)
        S = oracle.SellHost(rp, col, val, C_)
        x = vh.random_uniform(5, 777)
        assert np.array_equal(oracle.sell_spmv(S, x), oracle.csr_spmv(rp, col, val, x))


def test_cg_finite_termination_on_diagonal(oracle):
    # a diagonal SPD matrix with 5 distinct eigenvalues: CG converges in at most 5 steps
    n = 200
    d = np.repeat([1.0, 2.0, 5.0, 11.0, 30.0], n // 5)
    rp = np.arange(n + 1, dtype=np.uint32)
    A = oracle.CsrHost(rp, np.arange(n, dtype=np.uint32), d)
    b = np.linspace(-1, 1, n) + 0.3
    x, it, err, _ = oracle.cg_pipelined(A, b, 1e-10, 50)
    assert it <= 6 and err < 1e-10
    assert np.allclose(x, b / d, rtol=1e-9)


@pytest.mark.parametrize("variant", ["pipelined", "classic", "jacobi"])
def test_cg_solves_against_numpy(oracle, vh, variant):
    k = np.exp(vh.random_normal(3, 15 * 14, 0.0, 1.0))
    rp, col, val = vh.diffusion(15, 14, 1, k)
    A = dense(rp, col, val)
    b = vh.random_uniform(0, rp.size - 1)
    H = oracle.CsrHost(rp, col, val)
    if variant == "pipelined":
        x, it, err, hist = oracle.cg_pipelined(H, b, 1e-12, 2000, history=True)
    elif variant == "classic":
        x, it, err, hist = oracle.cg_classic(H, b, 1e-12, 2000, history=True)
    else:
        x, it, err, hist = oracle.cg_classic(H, b, 1e-12, 2000, diag=np.diag(A).copy(), history=True)
    xs = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-9
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-10
    assert err < 1e-12 and hist.size == it


def test_chronopoulos_gear_recurrence(oracle, vh):
    """rr_{k+1} predicted by beta = (alpha^2 <Ap,Ap> - rr)/rr equals the explicit <r,r> (cg.hpp:179-180)."""
    rp, col, val = vh.fdm_laplace_2d(20, 20)
    b = vh.random_uniform(0, 400)
    H = oracle.CsrHost(rp, col, val)
    _, it, _, hist = oracle.cg_pipelined(H, b, 1e-12, 500, history=True)
    _, itc, _, histc = oracle.cg_classic(H, b, 1e-12, 500, history=True)
    m = min(it, itc, 30)
    assert np.allclose(hist[:m], histc[:m], rtol=1e-6)
    assert abs(it - itc) <= 2


@pytest.mark.parametrize("jacobi", [False, True])
def test_bicgstab_against_numpy(oracle, vh, jacobi):
    rp, col, val = vh.powerlaw(400, max_len=30, diag_factor=1.5, seed=11)
    A = dense(rp, col, val)
    b = vh.random_uniform(0, 400)
    H = oracle.CsrHost(rp, col, val)
    if jacobi:
        x, it, err, _ = oracle.bicgstab_classic(H, b, 1e-11, 500, diag=np.diag(A).copy())
    else:
        x, it, err, _ = oracle.bicgstab_pipelined(H, b, 1e-11, 500)
    xs = np.linalg.solve(A, b)
    assert np.linalg.norm(x - xs) / np.linalg.norm(xs) < 1e-8
    assert err < 1e-11


def test_mixed_cg_true_residual(oracle, vh):
    rp, col, val = vh.laplace_3d(10, 9, 8)
    A = dense(rp, col, val)
    b = vh.random_uniform(0, rp.size - 1)
    x, it, err, restarts = oracle.mixed_cg(oracle.CsrHost(rp, col, val), b, 1e-10, 3000)
    rel = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert rel < 1e-10 and abs(rel - err) < 1e-12 and restarts >= 2


@pytest.mark.parametrize("jacobi", [False, True])
def test_mixed_reliable_update_pins(oracle, vh, jacobi):
    """Mixed CG with reliable updates (no reference counterpart): the FP64 true residual is below the
    tolerance and is the error it reports, x matches a dense solve, the FP32 iteration count stays
    within 15% of FP64 CG's, and the history is the FP32 <r,r> that the updates reset."""
    k = np.exp(vh.random_normal(3, 9 * 8 * 7))
    rp, col, val = vh.diffusion(9, 8, 7, k)
    A = dense(rp, col, val)
    b = vh.random_uniform(0, rp.size - 1)
    diag = oracle.row_info(rp, col, val, 3) if jacobi else None
    x, it, err, upd, hist = oracle.mixed_cg2(oracle.CsrHost(rp, col, val), b, 1e-10, 3000, mode=1, diag=diag,
                                             history=True)
    rel = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert rel < 1e-10 and abs(rel - err) < 1e-12 and upd >= 2
    assert np.linalg.norm(x - np.linalg.solve(A, b)) / np.linalg.norm(x) < 1e-8
    _, it64, _, _ = oracle.cg_pipelined(oracle.CsrHost(rp, col, val), b, 1e-10, 3000)
    if not jacobi:
        assert it <= 1.15 * it64
    assert hist.size == it and (hist > 0).all()


def test_mixed_reference_schedule_accumulation(oracle, vh):
    """mode 0 with FP32-accumulated inner products is the reference restatement itself; FP64
    accumulation (the HIP kernels') moves the iteration count only slightly."""
    rp, col, val = vh.laplace_3d(12, 11, 10)
    b = vh.random_uniform(0, rp.size - 1)
    A = oracle.CsrHost(rp, col, val)
    x0, it0, err0, _ = oracle.mixed_cg(A, b, 1e-10, 3000)
    x1, it1, err1, _, _ = oracle.mixed_cg2(A, b, 1e-10, 3000, mode=0, acc64=False)
    assert it0 == it1 and err0 == err1 and np.array_equal(x0, x1)
    _, it2, err2, _, _ = oracle.mixed_cg2(A, b, 1e-10, 3000, mode=0, acc64=True)
    assert abs(it2 - it0) <= max(2, 0.02 * it0) and err2 < 1e-10


def test_batched_fp32_against_dense(oracle, vh):
    sys_ptr, rp, col, val = vh.graph_laplacian_batch(4, 20, 60, seed=5)
    b = vh.random_uniform(0, rp.size - 1)
    x, iters, errs = oracle.batched_cg(sys_ptr, oracle.CsrHost(rp, col, val), b, 1e-5, 500, single=True)
    A = dense(rp, col, val)
    for s in range(4):
        r0, r1 = sys_ptr[s], sys_ptr[s + 1]
        xs = np.linalg.solve(A[r0:r1, r0:r1], b[r0:r1])
        assert np.linalg.norm(x[r0:r1] - xs) / np.linalg.norm(xs) < 1e-4
        assert errs[s] < 1e-5


def test_batched_oracle_per_system(oracle, vh):
    sys_ptr, rp, col, val = vh.graph_laplacian_batch(5, 10, 40, seed=2)
    b = vh.random_uniform(0, rp.size - 1)
    x, iters, errs = oracle.batched_cg(sys_ptr, oracle.CsrHost(rp, col, val), b, 1e-10, 500)
    A = dense(rp, col, val)
    for s in range(5):
        r0, r1 = sys_ptr[s], sys_ptr[s + 1]
        As = A[r0:r1, r0:r1]
        assert np.allclose(A[r0:r1, :r0], 0) and np.allclose(A[r0:r1, r1:], 0)
        xs = np.linalg.solve(As, b[r0:r1])
        assert np.linalg.norm(x[r0:r1] - xs) / np.linalg.norm(xs) < 1e-7
        assert errs[s] < 1e-10 and iters[s] >= 1


def test_adjoint_gradient_central_differences(oracle, vh):
    """dL/dA_ij = -lambda_i x_j with A^T lambda = dL/dx, for L = 0.5 ||x - t||^2 (finite differences)."""
    k = np.exp(vh.random_normal(4, 8 * 7, 0.0, 0.5))
    rp, col, val = vh.diffusion(8, 7, 1, k)
    n = rp.size - 1
    b = vh.random_uniform(0, n)
    t = vh.random_uniform(2, n)

    def loss(v):
        x = np.linalg.solve(dense(rp, col, v), b)
        return 0.5 * np.sum((x - t) ** 2)

    A = dense(rp, col, val)
    x = np.linalg.solve(A, b)
    lam = np.linalg.solve(A.T, x - t)
    g = oracle.adjoint_grad(rp, col, lam, x)
    h = 1e-6
    for kk in np.random.default_rng(0).choice(col.size, 12, replace=False):
        vp, vm = val.copy(), val.copy()
        vp[kk] += h
        vm[kk] -= h
        fd = (loss(vp) - loss(vm)) / (2 * h)
        assert abs(fd - g[kk]) <= 1e-6 * max(1.0, abs(fd))


def test_planted_mistakes_are_caught(oracle, vh):
    """Plants typical mistakes in a copy of the oracle's SpMV and CG (on the CPU) and checks that the
    closed-form / dense pins above would reject them."""
    rp, col, val = vh.laplace_3d(6, 5, 4)
    A = dense(rp, col, val)
    x = vh.random_uniform(1, rp.size - 1)
    good = oracle.csr_spmv(rp, col, val, x)
    assert np.allclose(good, A @ x, rtol=1e-13, atol=1e-13)
    # transposed operand: A^T x differs from A x for a nonsymmetric matrix
    rp2, col2, val2 = vh.powerlaw(200, max_len=20, seed=3)
    A2 = dense(rp2, col2, val2)
    x2 = vh.random_uniform(4, 200)
    assert not np.allclose(A2.T @ x2, oracle.csr_spmv(rp2, col2, val2, x2), rtol=1e-6)
    # dropped last entry of each row
    bad = np.array([sum(val[k] * x[col[k]] for k in range(rp[i], rp[i + 1] - 1)) for i in range(rp.size - 1)])
    assert not np.allclose(bad, A @ x, rtol=1e-6)
    # wrong beta sign in CG: the iteration no longer solves the system
    b = vh.random_uniform(0, rp.size - 1)
    xx, r, pp = np.zeros_like(b), b.copy(), b.copy()
    rr = r @ r
    for _ in range(200):
        Ap = A @ pp
        alpha = rr / (pp @ Ap)
        xx += alpha * pp
        r -= alpha * Ap
        rr_new = r @ r
        pp = r - (rr_new / rr) * pp  # planted sign error
        rr = rr_new
    assert np.linalg.norm(b - A @ xx) / np.linalg.norm(b) > 1e-6


def krylov_minimizer(A, b, K):
    """argmin of ||b - A x|| over the Krylov space span{b, A b, ..., A^(K-1) b} (dense, via an
    orthonormal Arnoldi basis)."""
    n = b.size
    Q = np.zeros((n, K))
    Q[:, 0] = b / np.linalg.norm(b)
    for j in range(1, K):
        w = A @ Q[:, j - 1]
        for _ in range(2):
            w -= Q[:, :j] @ (Q[:, :j].T @ w)
        Q[:, j] = w / np.linalg.norm(w)
    y = np.linalg.lstsq(A @ Q, b, rcond=None)[0]
    return Q @ y


def test_gmres_cycle_is_the_krylov_minimizer(oracle, vh):
    """One cycle of krylov_dim steps from x = 0 minimizes ||b - A x|| over K_krylov_dim(A, b), and
    the cycle's estimate |rho rho_0| / ||b|| equals the true residual.  The Gram-Schmidt range that
    leaves out the last entry (the reference's OpenMP kernel) misses the minimizer: the pin
    distinguishes the two."""
    rp, col, val = vh.convection_diffusion(9, 8, 3, beta=(0.9, 0.4, 0.2))
    A = dense(rp, col, val)
    b = vh.random_uniform(3, rp.size - 1)
    Ah = oracle.CsrHost(rp, col, val)
    for K in (1, 4, 8):
        xs = krylov_minimizer(A, b, K)
        x, it, err, _ = oracle.gmres(Ah, b, 1e-15, K, K)
        assert it == K
        assert np.linalg.norm(x - xs) <= 1e-9 * np.linalg.norm(xs)
        true = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        assert abs(err - true) <= 1e-8 * true
    xk, _, _, _ = oracle.gmres(Ah, b, 1e-15, 8, 8, gs_skip_tail=True)
    assert np.linalg.norm(xk - krylov_minimizer(A, b, 8)) > 1e-7 * np.linalg.norm(xk)


def test_gmres_full_space_and_restarts(oracle, vh):
    """krylov_dim >= n: the premature-convergence check stops the basis where it becomes dependent,
    after at most n + 1 steps, and a restarted solve reaches the dense solution; restarted cycles
    never increase the true residual."""
    rng = np.random.default_rng(4)
    n = 12
    A = rng.standard_normal((n, n)) + 6 * np.eye(n)
    rp = np.arange(0, n * n + 1, n, dtype=np.uint32)
    col = np.tile(np.arange(n, dtype=np.uint32), n)
    b = rng.standard_normal(n)
    xs = np.linalg.solve(A, b)
    x, it, err, _ = oracle.gmres(oracle.CsrHost(rp, col, A.ravel()), b, 1e-13, 20, 20)
    assert it <= n + 1 and np.linalg.norm(x - xs) <= 1e-4 * np.linalg.norm(xs)
    # classical Gram-Schmidt loses orthogonality over n steps; the restart recovers it
    x, it, err, _ = oracle.gmres(oracle.CsrHost(rp, col, A.ravel()), b, 1e-13, 60, 20)
    assert np.linalg.norm(x - xs) <= 1e-11 * np.linalg.norm(xs)
    rp, col, val = vh.convection_diffusion(10, 9, 2)
    Ad = dense(rp, col, val)
    b = vh.random_uniform(6, rp.size - 1)
    Ah = oracle.CsrHost(rp, col, val)
    last = np.inf
    for cycles in range(1, 6):
        x, it, err, _ = oracle.gmres(Ah, b, 1e-15, 6 * cycles, 6)
        assert it == 6 * cycles
        true = np.linalg.norm(b - Ad @ x) / np.linalg.norm(b)
        assert true <= last * (1 + 1e-12) and abs(err - true) <= 1e-6 * true
        last = true


def test_gmres_ops_against_numpy(oracle):
    """The operations with all entries: stage 1 is the dense product V^T v_k, stage 2 the
    projection v_k - V h, update x + c0 r + V c."""
    rng = np.random.default_rng(8)
    n, ld, K, k, chunk = 301, 320, 5, 4, 16
    basis = rng.standard_normal(K * ld)
    V = basis.reshape(K, ld)[:, :n]
    out = np.zeros(K * chunk)
    oracle.gmres_stage1(basis, n, ld, k, out, chunk)
    assert np.allclose(out[::chunk][:k], V[:k] @ V[k], rtol=1e-13, atol=1e-12)
    vi = np.zeros(K * chunk)
    vi[::chunk][:k] = h = rng.standard_normal(k)
    b2, R, buf = basis.copy(), np.zeros(K * K), np.zeros(3 * chunk)
    oracle.gmres_stage2(b2, n, ld, k, vi, R, K, buf, chunk)
    want = V[k] - h @ V[:k]
    assert np.allclose(b2.reshape(K, ld)[k, :n], want, rtol=1e-13, atol=1e-13)
    assert np.allclose(R.reshape(K, K)[k, :k], h) and np.isclose(buf[chunk], want @ want, rtol=1e-13)
    x0, r, c = rng.standard_normal(n), rng.standard_normal(n), rng.standard_normal(K)
    x = x0.copy()
    oracle.gmres_update_result(x, r, basis, n, ld, c, K)
    assert np.allclose(x, x0 + c[0] * r + c[1:] @ V[:K - 1], rtol=1e-13, atol=1e-13)
