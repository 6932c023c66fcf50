"""The allocation-free product entry point (amfls_operator_apply_rows_noalloc)
against the ordinary one (amfls_operator_apply_rows), through the C ABI.

Two calls in a row per operator, the second on a smaller input, each in both
directions and block precisions: the results agree with the ordinary entry
point's to FP64 rounding (FP64 blocks) or FP32 rounding (FP32 blocks) -- the
Aᵀ products may combine fewer partial sums to fit the fixed array, so the
summation order, not the result, can differ -- and no device memory is
allocated by the calls.  A product too large for the array is refused with
the length error.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ERR_LENGTH = 3


def _product(torch, M, fn, transpose, X, s, Y, dtype_code):
    lib, h = M.context._lib, M.context.handle
    return getattr(lib, fn)(h, M.handle, int(transpose), ctypes.c_void_p(X.data_ptr()), s, s,
                            ctypes.c_void_p(Y.data_ptr()), s, dtype_code)


def _compare(torch, M, rows, cols, s, seed):
    from amfls_solver_amd import _lib
    g = torch.Generator(device="cuda").manual_seed(seed)
    for dt, code, tol in ((torch.float64, _lib.F64, 1e-12), (torch.float32, _lib.F32, 2e-5)):
        for transpose in (0, 1):
            n_in, n_out = (rows, cols) if transpose else (cols, rows)
            X = torch.randn(n_in, s, device="cuda", dtype=dt, generator=g)
            want = torch.full((n_out, s), float("nan"), device="cuda", dtype=dt)
            got = torch.full((n_out, s), float("nan"), device="cuda", dtype=dt)
            torch.cuda.synchronize()
            assert _product(torch, M, "amfls_operator_apply_rows", transpose, X, s, want, code) == 0
            M.context.synchronize()
            free0 = torch.cuda.mem_get_info()[0]
            rc = _product(torch, M, "amfls_operator_apply_rows_noalloc", transpose, X, s, got, code)
            assert rc == 0, M.context._lib.amfls_last_error().decode()
            M.context.synchronize()
            assert torch.cuda.mem_get_info()[0] == free0, "the allocation-free product allocated device memory"
            w, o = want.double().cpu().numpy(), got.double().cpu().numpy()
            assert np.all(np.isfinite(o))
            err = np.linalg.norm(o - w) / np.linalg.norm(w)
            assert err <= tol, (dt, transpose, s, err)


def test_noalloc_dense_matches_apply_rows(gpu):
    import torch
    import amfls_solver_amd as am
    from amfls_solver_amd import problems as P
    for rows, cols, s in ((300_000, 1000, 32), (40_000, 600, 5)):   # second call: smaller input
        a, _, _ = P.config2_device(rows, cols)
        M = am.DeviceMatrix(a, value_dtype=torch.bfloat16)
        print(f"dense {rows}x{cols} width {s}", flush=True)
        _compare(torch, M, rows, cols, s, seed=rows)


def test_noalloc_csr_matches_apply_rows(gpu):
    import torch
    import amfls_solver_amd as am
    from amfls_solver_amd import problems as P
    for rows, cols, per, s in ((200_000, 20_000, 12, 16), (30_000, 4000, 7, 1)):
        p = P.csr_rows(rows, cols, per, seed=5, regularization=0.0)
        A = torch.sparse_csr_tensor(torch.from_numpy(p.row_ptr), torch.from_numpy(p.col_idx.astype(np.int64)),
                                    torch.from_numpy(p.values), size=(p.rows, p.cols))
        M = am.DeviceMatrix(A, value_dtype=torch.float32)
        print(f"csr {rows}x{cols} width {s}", flush=True)
        _compare(torch, M, rows, cols, s, seed=rows)


def test_noalloc_refuses_products_beyond_its_array(gpu):
    import torch
    import amfls_solver_amd as am
    from amfls_solver_amd import _lib
    # the packed three-plane copy of a 100000 x 64 block exceeds 16 MiB
    a = torch.randn(64, 100_000, device="cuda").to(torch.bfloat16)
    M = am.DeviceMatrix(a, value_dtype=torch.bfloat16)
    X = torch.randn(100_000, 64, device="cuda", dtype=torch.float32)
    Y = torch.empty(64, 64, device="cuda", dtype=torch.float32)
    rc = _product(torch, M, "amfls_operator_apply_rows_noalloc", 0, X, 64, Y, _lib.F32)
    assert rc == ERR_LENGTH, rc
    assert b"scratch" in M.context._lib.amfls_last_error()
