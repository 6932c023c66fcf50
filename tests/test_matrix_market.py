"""Matrix Market reader (host side; no GPU): hand-written files of every
accepted layout, files written by scipy's writer, and the refusals."""
import gzip

import numpy as np
import pytest
import scipy.io
import scipy.sparse

from amfls_solver_amd import matrix_market as MM


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return p


def test_coordinate_general_canonical_form(tmp_path):
    # repeated (1,1) added, stored zero dropped, rows given out of order
    p = _write(tmp_path, "a.mtx", "%%MatrixMarket matrix coordinate real general\n% comment\n"
               "3 4 6\n3 1 4.0\n1 3 1.0\n1 1 2.0\n2 2 3.0\n1 1 0.5\n3 4 0\n")
    A = MM.load_sparse(p)
    assert A.shape == (3, 4) and A.nnz == 4
    assert np.array_equal(A.toarray(), [[2.5, 0, 1, 0], [0, 3, 0, 0], [4, 0, 0, 0]])
    assert all(np.all(np.diff(A.indices[A.indptr[i]:A.indptr[i + 1]]) > 0) for i in range(3))
    gz = tmp_path / "a.mtx.gz"
    gz.write_bytes(gzip.compress(p.read_bytes()))
    assert (MM.load_sparse(gz) != A).nnz == 0


def test_symmetric_skew_pattern_and_array_layouts(tmp_path):
    sym = _write(tmp_path, "s.mtx", "%%MatrixMarket matrix coordinate integer symmetric\n"
                 "3 3 3\n1 1 5\n3 1 -2\n3 2 7\n")
    assert np.array_equal(MM.load_sparse(sym).toarray(), [[5, 0, -2], [0, 0, 7], [-2, 7, 0]])
    skew = _write(tmp_path, "k.mtx", "%%MatrixMarket matrix coordinate real skew-symmetric\n"
                  "2 2 1\n2 1 1.5\n")
    assert np.array_equal(MM.load_sparse(skew).toarray(), [[0, -1.5], [1.5, 0]])
    pat = _write(tmp_path, "p.mtx", "%%MatrixMarket matrix coordinate pattern general\n2 3 2\n1 3\n2 1\n")
    assert np.array_equal(MM.load_sparse(pat).toarray(), [[0, 0, 1], [1, 0, 0]])
    # dense arrays are column-major; symmetric arrays store the lower triangle by column
    arr = _write(tmp_path, "d.mtx", "%%MatrixMarket matrix array real general\n2 3\n1\n2\n3\n4\n5\n6\n")
    assert np.array_equal(MM.load_sparse(arr).toarray(), [[1, 3, 5], [2, 4, 6]])
    sarr = _write(tmp_path, "e.mtx", "%%MatrixMarket matrix array real symmetric\n3 3\n1\n2\n3\n4\n5\n6\n")
    assert np.array_equal(MM.load_sparse(sarr).toarray(), [[1, 2, 3], [2, 4, 5], [3, 5, 6]])


def test_agrees_with_scipy_writer(tmp_path):
    rng = np.random.default_rng(3)
    A = scipy.sparse.random(60, 45, density=0.1, random_state=5, format="coo")
    scipy.io.mmwrite(str(tmp_path / "r.mtx"), A)
    assert np.array_equal(MM.load_sparse(tmp_path / "r.mtx").toarray(), A.toarray())
    D = rng.standard_normal((7, 5))
    scipy.io.mmwrite(str(tmp_path / "dd.mtx"), D)
    assert np.array_equal(MM.load_sparse(tmp_path / "dd.mtx").toarray(), D)
    S = scipy.sparse.random(30, 30, density=0.2, random_state=9)
    S = (S + S.T).tocoo()
    scipy.io.mmwrite(str(tmp_path / "ss.mtx"), S, symmetry="symmetric")
    assert np.allclose(MM.load_sparse(tmp_path / "ss.mtx").toarray(), S.toarray(), rtol=0, atol=0)


def test_vectors(tmp_path):
    col = _write(tmp_path, "b.mtx", "%%MatrixMarket matrix array real general\n3 1\n3\n0\n4\n")
    assert np.allclose(MM.load_vector(col, 3), [0.6, 0.0, 0.8])
    assert np.array_equal(MM.load_vector(col, 3, unit=False), [3.0, 0.0, 4.0])
    row = _write(tmp_path, "r.mtx", "%%MatrixMarket matrix coordinate real general\n1 4 2\n1 2 -2\n1 4 1\n")
    assert np.array_equal(MM.load_vector(row, 4, unit=False), [0, -2, 0, 1])


@pytest.mark.parametrize("text", [
    "%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 2\n",      # complex field
    "%%MatrixMarket vector coordinate real general\n1 1 1\n1 1 1\n",           # not a matrix
    "%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1\n2 2 1\n",    # fewer entries than declared
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n",           # row out of range
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 nan\n",         # non-finite value
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 0\n",           # only a stored zero
    "%%MatrixMarket matrix coordinate real symmetric\n2 3 1\n1 1 1\n",         # symmetric but not square
    "%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n1 1 1\n",    # skew diagonal
    "%%MatrixMarket matrix array pattern general\n1 1\n1\n",                   # pattern array
])
def test_refuses_malformed_matrices(tmp_path, text):
    with pytest.raises(MM.MatrixMarketError):
        MM.load_sparse(_write(tmp_path, "bad.mtx", text))


def test_refuses_bad_vectors_and_missing_files(tmp_path):
    with pytest.raises(FileNotFoundError):
        MM.load_sparse(tmp_path / "missing.mtx")
    z = _write(tmp_path, "z.mtx", "%%MatrixMarket matrix array real general\n2 1\n0\n0\n")
    with pytest.raises(MM.MatrixMarketError, match="zero"):
        MM.load_vector(z, 2)
    with pytest.raises(MM.MatrixMarketError, match="not a vector"):
        MM.load_vector(z, 3)
    assert issubclass(MM.MatrixMarketError, ValueError)
