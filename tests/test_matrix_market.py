"""MatrixMarket reader/writer (vclhip_mm_*, io/matrix_market.hpp:102-433) against the reference's
own reader compiled in oracle/_ref (read_matrix_market_file into std::vector<std::map>), against
scipy.io, an independent implementation of the format, and against the reference reader's rules:
comments, blank lines and leading blanks, symmetric and pattern files, the entry count of the size
line, repeated entries (last value kept), index base, and the errors the reference reports."""
import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp


def csr_of(vh, path, base=1):
    rp, col, val, ncols, sym, lines = vh.read_matrix_market(path, base)
    return sp.csr_matrix((val, col, rp), shape=(rp.size - 1, ncols)), sym, lines


def test_scipy_written_files_read_back_exactly(vh, tmp_path):
    rng = np.random.default_rng(3)
    A = sp.random(300, 250, density=0.03, format="csr", random_state=4)
    A.data = rng.standard_normal(A.nnz)
    f = tmp_path / "a.mtx"
    scipy.io.mmwrite(str(f), A, precision=17)
    B, sym, lines = csr_of(vh, f)
    assert not sym and B.shape == A.shape and lines >= A.nnz + 2
    assert (B != A).nnz == 0
    # symmetric storage: scipy writes one triangle, the reader mirrors it
    S = sp.random(200, 200, density=0.05, format="csr", random_state=5)
    S = (S + S.T).tocsr()
    S.setdiag(4.0)
    fs = tmp_path / "s.mtx"
    scipy.io.mmwrite(str(fs), S, symmetry="symmetric", precision=17)
    assert "symmetric" in fs.read_text().splitlines()[0]
    T, sym, _ = csr_of(vh, fs)
    assert sym and (T != S).nnz == 0


def test_writer_matches_the_reference_format(vh, tmp_path):
    rp, col, val = vh.fdm_laplace_2d(9, 7)
    val = val * np.pi
    f = tmp_path / "w.mtx"
    vh.write_matrix_market(f, rp, col, val)
    lines = f.read_text().splitlines()
    assert lines[0] == "%%MatrixMarket matrix coordinate real general"
    assert lines[1] == f"{rp.size - 1} {rp.size - 1} {rp[-1]}"
    # row order, 1-based, default iostream formatting (6 significant digits)
    r0 = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    assert lines[2] == f"{r0[0] + 1} {col[0] + 1} {val[0]:g}"
    M = scipy.io.mmread(str(f)).tocsr()
    ref = sp.csr_matrix((val, col, rp))
    assert np.allclose(M.toarray(), ref.toarray(), rtol=1e-5, atol=0)
    # zero-based
    vh.write_matrix_market(f, rp, col, val, index_base=0)
    assert f.read_text().splitlines()[2].startswith("0 ")
    B, _, _ = csr_of(vh, f, base=0)
    assert B.nnz == rp[-1]


def test_reader_rules(vh, tmp_path):
    f = tmp_path / "r.mtx"
    f.write_text("%%MatrixMarket matrix coordinate real general\n"
                 "% a comment\n"
                 "\n"
                 "   4 3 5\n"
                 "1 1 1.5\n"
                 "  % indented comment\n"
                 "4 3 -2\n"
                 "1 1 7.25\n"            # repeated entry: the last value stays
                 "2 2 3e-3\n"
                 "3 1 1\n"
                 "4 1 99\n")             # beyond the 5 entries of the size line: not read
    A, sym, lines = csr_of(vh, f)
    D = A.toarray()
    assert D.shape == (4, 3) and not sym
    assert D[0, 0] == 7.25 and D[3, 2] == -2 and D[1, 1] == 3e-3 and D[2, 0] == 1 and D[3, 0] == 0
    assert lines == 10
    # pattern and complex fields
    f.write_text("%%MatrixMarket matrix coordinate pattern symmetric\n3 3 2\n2 1\n3 3\n")
    A, sym, _ = csr_of(vh, f)
    assert sym and np.array_equal(A.toarray(), [[0, 1, 0], [1, 0, 0], [0, 0, 1]])
    f.write_text("%%MatrixMarket matrix coordinate complex general\n2 2 1\n1 2 3.5 -1\n")
    A, _, _ = csr_of(vh, f)
    assert A.toarray()[0, 1] == 3.5


@pytest.mark.parametrize("text", [
    "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",
    "%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n2 1 1\n",
    "%%MatrixMarket matrix coordinate integer general\n2 2 1\n2 1 1\n",
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n",
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 0 1\n",
    "%%MatrixMarket vector coordinate real general\n2 2 1\n1 1 1\n",
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1\n",
    "%%MatrixMarket matrix coordinate real general\n% no size line\n",
])
def test_reader_errors(vh, tmp_path, text):
    f = tmp_path / "e.mtx"
    f.write_text(text)
    with pytest.raises(vh.VclHipError):
        vh.read_matrix_market(f)


def test_header_is_optional(vh, tmp_path):
    # the reference reader parses a file without the %%MatrixMarket line as real general
    f = tmp_path / "h.mtx"
    f.write_text("2 2 1\n2 1 4.5\n")
    A, sym, _ = csr_of(vh, f)
    assert not sym and A.toarray()[1, 0] == 4.5


def test_missing_file(vh, tmp_path):
    with pytest.raises(vh.VclHipError):
        vh.read_matrix_market(tmp_path / "absent.mtx")


# Rule cases read by both readers: the same lines-read count and the same matrix, or an error from
# both (the reference returns 0 on a parse error).
REFERENCE_CASES = {
    "rules": "%%MatrixMarket matrix coordinate real general\n% a comment\n\n   4 3 5\n1 1 1.5\n"
             "  % indented comment\n4 3 -2\n1 1 7.25\n2 2 3e-3\n3 1 1\n4 1 99\n",
    "symmetric": "%%MatrixMarket matrix coordinate real symmetric\n4 4 5\n1 1 2\n2 1 -1\n3 2 -1.5\n4 4 3\n"
                 "4 2 0.25\n",
    "symmetric_repeated": "%%MatrixMarket matrix coordinate real symmetric\n3 3 3\n2 1 1\n1 2 5\n2 1 7\n",
    "pattern_symmetric": "%%MatrixMarket matrix coordinate pattern symmetric\n3 3 2\n2 1\n3 3\n",
    "pattern_general": "%%MatrixMarket matrix coordinate pattern general\n3 3 3\n2 1\n3 3\n1 3\n",
    "complex": "%%MatrixMarket matrix coordinate complex general\n2 2 2\n1 2 3.5 -1\n2 1 -0.5 2\n",
    "upper_case_banner": "%%MatrixMarket MATRIX Coordinate REAL Symmetric\n3 3 2\n2 1 4\n3 3 1\n",
    "no_banner": "2 2 1\n2 1 4.5\n",
    "stored_zeros": "%%MatrixMarket matrix coordinate real general\n3 3 3\n1 1 0\n2 2 0.0\n3 1 -0\n",
    "extra_tokens": "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.0 junk\n2 2 2.0 7 8\n",
    "no_final_newline": "%%MatrixMarket matrix coordinate real general\n2 2 1\n2 1 4.5",
    "crlf": "%%MatrixMarket matrix coordinate real general\r\n2 2 1\r\n2 1 4.5\r\n",
    "tabs_between_fields": "%%MatrixMarket matrix coordinate real general\n2\t2\t1\n2\t1\t4.5\n",
    "blank_lines": "%%MatrixMarket matrix coordinate real general\n2 2 1\n\n\n2 1 4.5\n",
    "exponents": "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.25E+02\n2 2 -7.5e-300\n",
    # errors
    "fewer_entries": "%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1\n",
    "tab_led_comment": "%%MatrixMarket matrix coordinate real general\n2 2 1\n\t% c\n2 1 4.5\n",
    "array": "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",
    "skew_symmetric": "%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n2 1 1\n",
    "row_out_of_range": "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n",
    "index_below_base": "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 0 1\n",
    "missing_value": "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1\n",
    "no_size_line": "%%MatrixMarket matrix coordinate real general\n% no size line\n",
    "not_matrixmarket": "%%MatrixMarkets matrix coordinate real general\n2 2 1\n1 1 1\n",
}


def read_both(vh, oracle, path, base=1):
    r, rp, col, val = oracle.ref_read_matrix_market(path, base)
    try:
        ours = vh.read_matrix_market(path, base)
    except vh.VclHipError:
        ours = None
    return r, (rp, col, val), ours


needs_ref = pytest.mark.skipif(not __import__("pyoracle").ref_available(),
                               reason="reference build oracle/_ref not present")


@needs_ref
@pytest.mark.parametrize("case", sorted(REFERENCE_CASES))
def test_reader_equals_reference_reader(vh, oracle, tmp_path, case):
    f = tmp_path / "c.mtx"
    f.write_bytes(REFERENCE_CASES[case].encode())
    r, (rp, col, val), ours = read_both(vh, oracle, f)
    if r == 0:
        assert ours is None, f"{case}: the reference rejects the file, vclhip_mm_read accepted it"
        return
    assert ours is not None, f"{case}: the reference reads the file ({r} lines), vclhip_mm_read rejected it"
    orp, ocol, oval, _, _, lines = ours
    assert lines == r
    assert np.array_equal(orp, rp) and np.array_equal(ocol, col)
    assert np.array_equal(oval.view(np.uint64), val.view(np.uint64))  # bit for bit, signed zeros too


@needs_ref
@pytest.mark.parametrize("base", [0, 1, 3])
def test_reader_equals_reference_reader_large(vh, oracle, tmp_path, base):
    """A 3000-row power-law matrix with 17-digit values written by scipy (comments, general), and a
    symmetric 2D Laplacian, read with index bases 0, 1 and 3 by both readers."""
    rng = np.random.default_rng(11 + base)
    n = 3000
    lens = np.minimum((rng.pareto(1.0, n) + 1).astype(np.int64), 200)
    rows = np.repeat(np.arange(n), lens)
    cols = rng.integers(0, n, rows.size)
    A = sp.csr_matrix((rng.standard_normal(rows.size), (rows, cols)), shape=(n, n))
    A.sum_duplicates()
    f = tmp_path / "p.mtx"
    scipy.io.mmwrite(str(f), A, precision=17, comment="power-law rows")
    text = f.read_text().splitlines()
    k = next(i for i, t in enumerate(text) if not t.startswith("%"))  # the size line
    body = [" ".join(str(int(v) - 1 + base) if j < 2 else v for j, v in enumerate(t.split()))
            for t in text[k + 1:]]
    f.write_text("\n".join(text[: k + 1] + body) + "\n")
    L = sp.diags([-1, -1, 4, -1, -1], [-40, -1, 0, 1, 40], shape=(1600, 1600), format="csr")
    fs = tmp_path / "s.mtx"
    scipy.io.mmwrite(str(fs), L, symmetry="symmetric", precision=17)
    for path, b in ((f, base), (fs, 1)):
        r, (rp, col, val), ours = read_both(vh, oracle, path, b)
        assert r > 0 and ours is not None
        orp, ocol, oval, _, _, lines = ours
        assert lines == r and np.array_equal(orp, rp) and np.array_equal(ocol, col) and np.array_equal(oval, val)


@needs_ref
def test_zero_entry_file_is_an_empty_matrix(vh, oracle, tmp_path):
    """The one deliberate difference: a size line declaring 0 entries reads as an empty matrix here;
    the reference reader goes on to the end of the file and reports a parse error there."""
    f = tmp_path / "z.mtx"
    f.write_text("%%MatrixMarket matrix coordinate real general\n3 2 0\n")
    r, _, ours = read_both(vh, oracle, f)
    assert r == 0
    rp, col, val, ncols, _, _ = ours
    assert rp.tolist() == [0, 0, 0, 0] and col.size == 0 and ncols == 2


@pytest.mark.gpu
def test_file_read_into_device_and_host_gives_the_same_product(vh, dev, oracle, tmp_path):
    """One file, two domains: vclhip_mm_read -> compressed_matrix on the device -> SpMV, and the
    reference reader -> the oracle's host CSR product: the same arrays, and the same products bit for
    bit on rows of at most 32 entries (one lane sums them in the host's order); longer rows are
    summed by a wavefront and agree to 1e-13 relative (inf-norm), the SpMV parity bar."""
    rng = np.random.default_rng(5)
    n = 20000
    lens = np.minimum((rng.pareto(1.0, n) + 1).astype(np.int64), 512)
    rows = np.repeat(np.arange(n), lens)
    A = sp.csr_matrix((rng.standard_normal(rows.size), (rows, rng.integers(0, n, rows.size))), shape=(n, n))
    A.sum_duplicates()
    f = tmp_path / "d.mtx"
    scipy.io.mmwrite(str(f), A, precision=17)
    rp, col, val, _, _, _ = vh.read_matrix_market(f)
    if oracle.ref_available():
        r, (hrp, hcol, hval), _ = read_both(vh, oracle, f)
        assert r > 0 and np.array_equal(hrp, rp) and np.array_equal(hcol, col) and np.array_equal(hval, val)
    else:  # the reference build did not travel: scipy's reader gives the host arrays
        H = scipy.io.mmread(str(f)).tocsr()
        H.sort_indices()
        hrp, hcol, hval = H.indptr.astype(np.uint32), H.indices.astype(np.uint32), H.data
    x = rng.uniform(-1, 1, n)
    dA = vh.CsrMatrix(rp, col, val)
    dx = vh.DeviceArray.from_numpy(x)
    dy = vh.DeviceArray.zeros(n, np.float64)
    dA.spmv(dx, dy)
    y = dy.to_numpy()
    y_host = oracle.csr_spmv(hrp, hcol, hval, x)
    short = np.diff(rp.astype(np.int64)) <= 32
    assert short.sum() > n // 2 and (~short).any()
    assert np.array_equal(y[short], y_host[short])
    assert np.max(np.abs(y - y_host)) <= 1e-13 * np.max(np.abs(y_host))
