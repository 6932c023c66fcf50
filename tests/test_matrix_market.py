"""Matrix Market ingestion (host side; no GPU)."""
import gzip

import numpy as np
import pytest
import scipy.io
import scipy.sparse

from amfls_solver_amd import matrix_market as MM


def test_reads_canonical_csr_and_unit_rhs(tmp_path):
    # duplicates summed, explicit zeros dropped, indices sorted
    rows = np.array([0, 0, 1, 2, 2, 0])
    cols = np.array([2, 0, 1, 0, 0, 1])
    vals = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 0.0])
    coo = scipy.sparse.coo_matrix((vals, (rows, cols)), shape=(3, 3))
    p = tmp_path / "a.mtx"
    scipy.io.mmwrite(str(p), coo)
    A = MM.read_matrix_market(p)
    assert A.shape == (3, 3) and A.nnz == 4
    assert np.array_equal(A.toarray(), np.array([[2.0, 0, 1.0], [0, 3.0, 0], [9.0, 0, 0]]))
    assert all(np.all(np.diff(A.indices[A.indptr[i]:A.indptr[i + 1]]) > 0) for i in range(3))
    gz = tmp_path / "a.mtx.gz"
    gz.write_bytes(gzip.compress(p.read_bytes()))
    assert (MM.read_matrix_market(gz) != A).nnz == 0
    bp = tmp_path / "b.mtx"
    scipy.io.mmwrite(str(bp), np.array([[3.0], [0.0], [4.0]]))
    b = MM.read_rhs(bp, 3)
    assert np.allclose(b, [0.6, 0.0, 0.8])


def test_rejects_bad_inputs(tmp_path):
    with pytest.raises(FileNotFoundError):
        MM.read_matrix_market(tmp_path / "missing.mtx")
    c = tmp_path / "c.mtx"
    scipy.io.mmwrite(str(c), scipy.sparse.coo_matrix(np.array([[1 + 1j, 0], [0, 1]])))
    with pytest.raises(ValueError):
        MM.read_matrix_market(c)
    z = tmp_path / "z.mtx"
    scipy.io.mmwrite(str(z), np.zeros((2, 1)))
    with pytest.raises(ValueError):
        MM.read_rhs(z, 2)
    with pytest.raises(ValueError):
        MM.read_rhs(z, 3)
