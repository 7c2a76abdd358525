"""MatrixMarket reader/writer (vclhip_mm_*, io/matrix_market.hpp:102-433) against scipy.io, an
independent implementation of the format, and against the reference reader's rules: comments,
blank lines and leading blanks, symmetric and pattern files, the entry count of the size line,
repeated entries (last value kept), index base, and the errors the reference reports."""
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
