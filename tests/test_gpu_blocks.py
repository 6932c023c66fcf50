"""The engine's FP64 and FP32 block kernels (tall-skinny products, the row-local
product and the norm statistics) through their C entry points, against
float64 references computed on the host: every width class the dispatch
distinguishes (the matrix-core kernels for K, N <= 16 and the tiled kernels
beyond), ragged row counts, padded row strides, in-place row products and
statistics over values spanning the double exponent range."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env(gpu):
    import torch
    import amfls_solver_amd as am
    ctx = am.Context()
    return torch, ctx, ctx._lib


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _padded(torch, rng, m, k, ld):
    a = torch.zeros(m, ld, dtype=torch.float64)
    a[:, :k] = torch.from_numpy(rng.standard_normal((m, k)))
    return a.cuda()


SHAPES = [(1, 1, 1), (37, 5, 3), (1000, 16, 16), (1000, 17, 20), (4099, 32, 32), (777, 33, 8), (513, 8, 40),
          (2000, 64, 64), (300, 1, 32), (300, 32, 1)]


@pytest.mark.parametrize("m,k,n", SHAPES)
def test_block_gemm_nn(env, m, k, n):
    torch, ctx, lib = env
    from amfls_solver_amd import _lib
    rng = np.random.default_rng(m + 7 * k + 13 * n)
    for lda, ldy, alpha, beta in ((k, n, 1.0, 0.0), (k + 3, n + 5, -0.5, 2.0)):
        a = _padded(torch, rng, m, k, lda)
        c = torch.from_numpy(rng.standard_normal((k, n))).cuda()
        y = _padded(torch, rng, m, n, ldy)
        want = alpha * (a[:, :k].cpu() @ c.cpu()) + beta * y[:, :n].cpu()
        _lib.check(lib.amfls_block_gemm_nn(ctx.handle, m, n, k, _p(a), lda, _p(c), n, _p(y), ldy, alpha, beta))
        got = y[:, :n].cpu()
        scale = (abs(alpha) * (a[:, :k].abs().cpu() @ c.abs().cpu()) + abs(beta)).max().item()
        assert (got - want).abs().max().item() <= 1e-14 * scale * (k + 2), (lda, ldy)
        # padding columns untouched
        if ldy > n:
            assert torch.isfinite(y[:, n:]).all()


@pytest.mark.parametrize("m,k,n", SHAPES + [(100003, 16, 16), (65537, 32, 20)])
def test_block_gemm_tn(env, m, k, n):
    torch, ctx, lib = env
    from amfls_solver_amd import _lib
    rng = np.random.default_rng(3 * m + k + n)
    for lda, ldb, alpha in ((k, n, 1.0), (k + 2, n + 1, 0.25)):
        a = _padded(torch, rng, m, k, lda)
        b = _padded(torch, rng, m, n, ldb)
        c = torch.full((k, n + 2), 7.0, dtype=torch.float64, device="cuda")
        _lib.check(lib.amfls_block_gemm_tn(ctx.handle, m, k, n, _p(a), lda, _p(b), ldb, _p(c), n + 2, alpha))
        want = alpha * (a[:, :k].cpu().T @ b[:, :n].cpu())
        scale = abs(alpha) * (a[:, :k].abs().cpu().T @ b[:, :n].abs().cpu()).max().item()
        assert (c[:, :n].cpu() - want).abs().max().item() <= 1e-14 * scale * np.sqrt(m) * 4, (lda, ldb)
        assert (c[:, n:] == 7.0).all()


@pytest.mark.parametrize("rows,w,r", [(1, 1, 1), (1000, 16, 16), (4097, 32, 32), (999, 20, 7), (500, 40, 33)])
def test_block_rowmul_in_place(env, rows, w, r):
    torch, ctx, lib = env
    from amfls_solver_amd import _lib
    rng = np.random.default_rng(rows + w + r)
    ld = max(w, r) + 1
    x = _padded(torch, rng, rows, w, ld)
    mat = torch.from_numpy(rng.standard_normal((w, r))).cuda()
    want = x[:, :w].cpu() @ mat.cpu()
    _lib.check(lib.amfls_block_rowmul(ctx.handle, rows, w, r, _p(x), ld, _p(mat), _p(x), ld))
    scale = want.abs().max().item() + 1.0
    assert (x[:, :r].cpu() - want).abs().max().item() <= 1e-14 * scale * (w + 1)


def _stats_ref(v):
    v = np.abs(v.ravel())
    fin = np.isfinite(v)
    bad = float((~fin).sum())
    v = v[fin]
    mx = float(v.max()) if v.size else 0.0
    if mx == 0.0:
        return mx, 0.0, bad
    e = np.frexp(mx)[1]
    return mx, float(np.sum(np.ldexp(v, -e) ** 2)), bad


@pytest.mark.parametrize("rows,cols,ld", [(1, 1, 1), (1000, 1, 1), (200001, 1, 3), (5000, 16, 16), (4099, 7, 9),
                                          (1, 40, 40)])
def test_block_stats(env, rows, cols, ld):
    torch, ctx, lib = env
    from amfls_solver_amd import _lib
    rng = np.random.default_rng(rows + cols)
    for case in ("plain", "wide", "special"):
        v = rng.standard_normal((rows, cols))
        if case == "wide":  # magnitudes over six hundred decades
            v *= 10.0 ** rng.uniform(-300, 300, size=v.shape)
        if case == "special" and v.size > 3:
            v.ravel()[0] = np.inf
            v.ravel()[1] = np.nan
            v.ravel()[2] = 0.0
        x = torch.zeros(rows, ld, dtype=torch.float64)
        x[:, cols:] = float("nan")  # padding must be ignored
        x[:, :cols] = torch.from_numpy(v)
        x = x.cuda()
        out = (ctypes.c_double * 3)()
        _lib.check(lib.amfls_block_stats(ctx.handle, rows, cols, _p(x), ld, out))
        mx, ss, bad = _stats_ref(v)
        assert out[0] == mx and out[2] == bad, case
        assert abs(out[1] - ss) <= 1e-13 * max(ss, 1e-300) * np.log2(v.size + 2), case


# FP32 blocks (the mixed-precision search's local products): the matrix-core
# tall kernels for K, N <= 64, against float64 host products of the same
# FP32 inputs.  Products with FP32 outputs are checked to FP32 rounding of
# the absolute-value product; A^T B accumulates exact products in FP64.
SHAPES_F32 = [(1, 1, 1), (37, 5, 3), (1000, 16, 16), (1000, 17, 20), (4099, 32, 32), (777, 33, 8), (513, 8, 40),
              (20001, 64, 64), (300, 1, 64), (300, 64, 1), (100003, 48, 64), (2000, 300, 32), (777, 130, 17),
              (3001, 2000, 64)]


def _padded32(torch, rng, m, k, ld):
    a = torch.zeros(m, ld, dtype=torch.float32)
    a[:, :k] = torch.from_numpy(rng.standard_normal((m, k)).astype(np.float32))
    return a.cuda()


@pytest.mark.parametrize("m,k,n", SHAPES_F32)
def test_block_gemm_nn_f32(env, m, k, n):
    torch, ctx, lib = env
    from amfls_solver_amd import _lib
    rng = np.random.default_rng(m + 5 * k + 11 * n)
    for lda, ldy, alpha, beta in ((k, n, 1.0, 0.0), (k + 3, n + 5, -0.5, 2.0)):
        a = _padded32(torch, rng, m, k, lda)
        c = torch.from_numpy(rng.standard_normal((k, n))).cuda()
        y = _padded32(torch, rng, m, n, ldy)
        a64, y64 = a[:, :k].double().cpu(), y[:, :n].double().cpu()
        want = alpha * (a64 @ c.cpu()) + beta * y64
        _lib.check(lib.amfls_block_gemm_nn_f32(ctx.handle, m, n, k, _p(a), lda, _p(c), n, _p(y), ldy, alpha, beta))
        got = y[:, :n].double().cpu()
        scale = abs(alpha) * (a64.abs() @ c.abs().cpu()) + abs(beta) * y64.abs()
        assert ((got - want).abs() <= 4e-7 * (k + 2) * scale + 1e-30).all(), (lda, ldy)
        if ldy > n:
            assert torch.isfinite(y[:, n:]).all()


@pytest.mark.parametrize("m,k,n", SHAPES_F32)
def test_block_gemm_tn_f32(env, m, k, n):
    torch, ctx, lib = env
    from amfls_solver_amd import _lib
    rng = np.random.default_rng(7 * m + k + 3 * n)
    for lda, ldb, alpha in ((k, n, 1.0), (k + 2, n + 1, 0.25)):
        a = _padded32(torch, rng, m, k, lda)
        b = _padded32(torch, rng, m, n, ldb)
        c = torch.full((k, n + 2), 7.0, dtype=torch.float64, device="cuda")
        _lib.check(lib.amfls_block_gemm_tn_f32(ctx.handle, m, k, n, _p(a), lda, _p(b), ldb, _p(c), n + 2, alpha))
        a64, b64 = a[:, :k].double().cpu(), b[:, :n].double().cpu()
        want = alpha * (a64.T @ b64)
        scale = abs(alpha) * (a64.abs().T @ b64.abs()).max().item()
        assert (c[:, :n].cpu() - want).abs().max().item() <= 1e-14 * scale * np.sqrt(m) * 4, (lda, ldb)
        assert (c[:, n:] == 7.0).all()


@pytest.mark.parametrize("rows,w,r", [(1, 1, 1), (1000, 16, 16), (4097, 32, 32), (999, 20, 7), (500, 40, 33),
                                      (100001, 64, 64), (700, 100, 50), (300, 128, 64)])
def test_block_rowmul_f32_in_place(env, rows, w, r):
    torch, ctx, lib = env
    from amfls_solver_amd import _lib
    rng = np.random.default_rng(rows + 3 * w + r)
    ld = max(w, r) + 1
    x = _padded32(torch, rng, rows, w, ld)
    mat = torch.from_numpy(rng.standard_normal((w, r))).cuda()
    x64 = x[:, :w].double().cpu()
    want = x64 @ mat.cpu()
    _lib.check(lib.amfls_block_rowmul_f32(ctx.handle, rows, w, r, _p(x), ld, _p(mat), _p(x), ld))
    scale = x64.abs() @ mat.abs().cpu()
    assert ((x[:, :r].double().cpu() - want).abs() <= 4e-7 * (w + 1) * scale + 1e-30).all()
