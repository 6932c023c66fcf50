"""Rank reveal of search blocks (CholeskyQR passes + SVD of the triangular
factor, the engine's replacement for the reference's thin SVD of the block,
krylov_workspace.cpp:263-309) against numpy's SVD of the same block.

Blocks W = U diag(sigma) V^T with orthonormal U, V and singular values
spread over up to ten decades (condition numbers to 1e10), plus exactly
repeated columns.  The engine's cutoff must be the reference's
64 eps max(m, n, k, w) ||W||_F for FP64 blocks, its rank the number of
numpy singular values above that cutoff, and its retained singular values
those of numpy to 1e-12 of sigma_max (FP64).  FP32 blocks use the FP32
rule 64 eps_32 sqrt(w) ||W||_F and a 1e-5 tolerance.  Singular values are
kept at least a factor 20 away from the cutoff so the decision is not a
rounding coin flip."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env(gpu):
    import torch
    import amfls_solver_amd as am
    ctx = am.Context()
    return torch, ctx


def _block(rng, n, sigma):
    w = sigma.size
    u, _ = np.linalg.qr(rng.standard_normal((n, w)))
    v, _ = np.linalg.qr(rng.standard_normal((w, w)))
    return (u * sigma) @ v.T


def _reveal(torch, ctx, W, dtype, factor=0.0):
    from amfls_solver_amd import _lib
    t = torch.from_numpy(W.astype(dtype)).contiguous().cuda()
    n, w = W.shape
    rank, cut, disc = ctypes.c_int32(), ctypes.c_double(), ctypes.c_double()
    sv = np.zeros(w)
    code = _lib.F64 if dtype == np.float64 else _lib.F32
    torch.cuda.synchronize()
    _lib.check(ctx._lib.amfls_block_rank_reveal(ctx.handle, n, w, ctypes.c_void_p(t.data_ptr()), w, code, w,
                                                float(factor), ctypes.byref(rank), ctypes.byref(cut),
                                                ctypes.byref(disc), sv.ctypes.data_as(ctypes.c_void_p)))
    q = t.double().cpu().numpy()
    return rank.value, cut.value, disc.value, sv, q


CASES64 = [(8, 0.0), (32, 4.0), (64, 8.0), (32, 10.0), (64, 14.0), (48, 20.0)]


@pytest.mark.parametrize("w,decades", CASES64)
def test_fp64_rank_reveal_matches_numpy_svd(env, w, decades):
    torch, ctx = env
    rng = np.random.default_rng(100 * w + int(decades))
    n = 20000
    sigma = np.logspace(0, -decades, w)
    W = _block(rng, n, sigma)
    s_np = np.linalg.svd(W, compute_uv=False)
    cutoff_ref = 64 * np.finfo(np.float64).eps * max(n, w) * np.linalg.norm(W)
    keep = s_np > cutoff_ref
    # stay clear of the cutoff: no singular value within a factor 20 of it
    if np.any((s_np > cutoff_ref / 20) & (s_np < cutoff_ref * 20)):
        sigma[(sigma > cutoff_ref / 20) & (sigma < cutoff_ref * 20)] *= 100.0
        W = _block(rng, n, sigma)
        s_np = np.linalg.svd(W, compute_uv=False)
        cutoff_ref = 64 * np.finfo(np.float64).eps * max(n, w) * np.linalg.norm(W)
        keep = s_np > cutoff_ref
    rank, cut, disc, sv, q = _reveal(torch, ctx, W, np.float64)
    assert abs(cut - cutoff_ref) <= 1e-12 * cutoff_ref
    assert rank == int(keep.sum()), (rank, s_np[:3], s_np[-3:], cutoff_ref)
    assert np.all(np.abs(sv[:rank] - s_np[:rank]) <= 1e-12 * s_np[0]), np.max(np.abs(sv[:rank] - s_np[:rank]))
    assert disc <= np.sqrt(w) * cutoff_ref
    # the retained columns are orthonormal and span the dominant singular subspace
    Q = q[:, :rank]
    assert np.max(np.abs(Q.T @ Q - np.eye(rank))) < 1e-12
    if rank == w:
        assert np.linalg.norm(W - Q @ (Q.T @ W)) <= 1e-12 * np.linalg.norm(W)


def test_fp64_repeated_columns_lose_exactly_their_rank(env):
    torch, ctx = env
    rng = np.random.default_rng(5)
    base = rng.standard_normal((5000, 12))
    W = np.concatenate([base, base[:, :4], 3.0 * base[:, 5:7]], axis=1)   # 18 columns, rank 12
    rank, cut, disc, sv, q = _reveal(torch, ctx, W, np.float64)
    s_np = np.linalg.svd(W, compute_uv=False)
    assert rank == 12
    assert np.all(np.abs(sv[:12] - s_np[:12]) <= 1e-12 * s_np[0])


@pytest.mark.parametrize("w,decades", [(16, 3.0), (32, 5.0), (64, 7.0)])
def test_fp32_rank_reveal_matches_numpy_svd(env, w, decades):
    torch, ctx = env
    rng = np.random.default_rng(7 * w)
    n = 20000
    sigma = np.logspace(0, -decades, w)
    W = _block(rng, n, sigma).astype(np.float32).astype(np.float64)      # the stored FP32 block
    s_np = np.linalg.svd(W, compute_uv=False)
    cutoff_ref = 64 * np.finfo(np.float32).eps * np.sqrt(w) * np.linalg.norm(W)
    near = (s_np > cutoff_ref / 20) & (s_np < cutoff_ref * 20)
    rank, cut, disc, sv, q = _reveal(torch, ctx, W, np.float32)
    assert abs(cut - cutoff_ref) <= 1e-6 * cutoff_ref
    if not near.any():
        assert rank == int((s_np > cutoff_ref).sum())
    else:  # the decision can go either way inside the band, never outside it
        assert int((s_np > 20 * cutoff_ref).sum()) <= rank <= int((s_np > cutoff_ref / 20).sum())
    k = min(rank, int((s_np > 20 * cutoff_ref).sum()))
    assert np.all(np.abs(sv[:k] - s_np[:k]) <= 1e-5 * s_np[0])
