"""GPU parity of the HIP engine against the CPU oracle (tests/golden fixtures
and direct oracle runs).  Adaptive runs use the synthetic clock so both sides
make the same width decisions; the trajectories (depth, rank, width per
candidate) must then be identical and the solutions agree to within FP64
rounding (relative 1e-8 here)."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(__file__), "golden", "fixtures.npz")


@pytest.fixture(scope="module")
def fx():
    return np.load(FIX)


@pytest.fixture(scope="module")
def torch_mod(gpu):
    import torch
    return torch


def csr_tensor(torch, rp, ci, val, shape):
    return torch.sparse_csr_tensor(torch.from_numpy(np.asarray(rp, np.int64)),
                                   torch.from_numpy(np.asarray(ci, np.int64)),
                                   torch.from_numpy(np.asarray(val)), size=shape)


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_gaussian_stream_matches_oracle(torch_mod, fx):
    import ctypes
    import amfls_solver_amd as am
    from amfls_solver_amd import _lib
    torch = torch_mod
    ctx = am.binding.default_context()
    y = torch.empty((64, 4), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    _lib.check(ctx._lib.amfls_gaussian_fill(ctx.handle, ctypes.c_void_p(y.data_ptr()), 64, 4, 4, 0, 7, 3, 2))
    g = y.cpu().numpy()
    # splitmix64 stages are exact; device log/cos may differ from libm by ulps
    assert np.allclose(g, fx["gauss_block"], rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("vt", ["float32", "float64"])
def test_csr_products_all_variants(torch_mod, fx, vt):
    import amfls_solver_amd as am
    torch = torch_mod
    rp, ci, val = fx["csr_rp"], fx["csr_ci"], fx["csr_val"].astype(vt)
    shape = (3000, 700)
    import scipy.sparse as sp
    A = sp.csr_matrix((val.astype(np.float64), ci, rp), shape=shape)
    M = am.DeviceMatrix(csr_tensor(torch, rp, ci, val, shape), value_dtype=getattr(torch, vt))
    lib = M.context._lib
    nvar = lib.amfls_spmm_variant_count()
    rng = np.random.default_rng(0)
    for s in (1, 2, 3, 4, 8, 16, 17, 33, 64, 100, 257):
        X = rng.standard_normal((700, s))
        Y = A @ X
        Z = A.T @ Y
        for v in [-1] + [i for i in range(nvar) if lib.amfls_spmm_variant_supports(i, s)]:
            M.set_variant(s, False, v)
            M.set_variant(s, True, v)
            y = M.apply(torch.from_numpy(X)).cpu().numpy()
            z = M.apply(torch.from_numpy(Y), transpose=True).cpu().numpy()
            assert np.abs(y - Y).max() <= 1e-13 * np.abs(Y).max(), (s, v)
            assert np.abs(z - Z).max() <= 1e-13 * np.abs(Z).max(), (s, v)
        y32 = M.apply(torch.from_numpy(X).float()).cpu().numpy()
        assert np.abs(y32 - Y).max() <= 1e-5 * np.abs(Y).max()


def _slab_matrix(rng, rows, cols, mean):
    """Random CSR with empty rows (every 97th) and one full row."""
    lens = rng.integers(0, 2 * mean, size=rows)
    lens[::97] = 0
    full = rows // 3
    cis = []
    for i, L in enumerate(lens):
        cis.append(np.arange(cols) if i == full else np.unique(rng.integers(0, cols, size=L)))
    rp = np.concatenate([[0], np.cumsum([len(c) for c in cis])]).astype(np.int64)
    return rp, np.concatenate(cis).astype(np.int32)


def _variant_index(lib, name):
    return [lib.amfls_spmm_variant_name(v).decode() for v in range(lib.amfls_spmm_variant_count())].index(name)


@pytest.mark.parametrize("vt", ["float32", "float64"])
def test_slab_kernel_products(torch_mod, vt):
    """The width-one slab kernel (x staged in LDS one column slab at a time)
    against scipy: three and four slabs of 8192 columns (A, A^T), many
    panels, empty rows, and a full row whose tiles span several chunks
    (more than the two loaded ahead); FP64 and FP32 blocks; the same bits
    on a second launch."""
    import scipy.sparse as sp
    import amfls_solver_amd as am
    torch = torch_mod
    rng = np.random.default_rng(7)
    rows, cols = 30000, 20000
    rp, ci = _slab_matrix(rng, rows, cols, 12)
    val = rng.standard_normal(len(ci)).astype(vt)
    A = sp.csr_matrix((val.astype(np.float64), ci, rp), shape=(rows, cols))
    M = am.DeviceMatrix(csr_tensor(torch, rp, ci, val, (rows, cols)), value_dtype=getattr(torch, vt))
    slab = _variant_index(M.context._lib, "slab")
    M.set_variant(1, False, slab)
    M.set_variant(1, True, slab)
    x = rng.standard_normal(cols)
    y_ref = A @ x
    z_ref = A.T @ y_ref
    y = M.apply(torch.from_numpy(x)).cpu().numpy()[:, 0]
    z = M.apply(torch.from_numpy(y_ref), transpose=True).cpu().numpy()[:, 0]
    assert np.abs(y - y_ref).max() <= 1e-13 * np.abs(y_ref).max()
    assert np.abs(z - z_ref).max() <= 1e-13 * np.abs(z_ref).max()
    assert np.array_equal(y, M.apply(torch.from_numpy(x)).cpu().numpy()[:, 0])
    assert np.array_equal(z, M.apply(torch.from_numpy(y_ref), transpose=True).cpu().numpy()[:, 0])
    assert np.all(y[::97] == 0.0)  # empty rows
    y32 = M.apply(torch.from_numpy(x).float()).cpu().numpy()[:, 0]
    z32 = M.apply(torch.from_numpy(y_ref).float(), transpose=True).cpu().numpy()[:, 0]
    assert np.abs(y32 - y_ref).max() <= 1e-5 * np.abs(y_ref).max()
    assert np.abs(z32 - z_ref).max() <= 1e-5 * np.abs(z_ref).max()


def test_slab_kernel_in_ridge_solves(torch_mod):
    """Ridge and minimum-norm solves with the slab kernel for every width-one
    product (A^T with the augmented add) certify and agree with the default
    kernels' solves."""
    import amfls_solver_amd as am
    torch = torch_mod
    rng = np.random.default_rng(8)
    rows, cols = 20000, 9000
    rp, ci = _slab_matrix(rng, rows, cols, 10)
    val = rng.standard_normal(len(ci))
    b = torch.from_numpy(rng.standard_normal(rows)).cuda()
    M = am.DeviceMatrix(csr_tensor(torch, rp, ci, val, (rows, cols)))
    slab = _variant_index(M.context._lib, "slab")
    for lam in (1e-2, 1e-4, 0.0):
        ref = am.run_amfls(M, b, regularization=lam, tolerance=1e-8, clock="synthetic")
        M.set_variant(1, False, slab)
        M.set_variant(1, True, slab)
        r = am.run_amfls(M, b, regularization=lam, tolerance=1e-8, clock="synthetic")
        M.set_variant(1, False, -1)
        M.set_variant(1, True, -1)
        assert r.status == ref.status, (lam, r.status, ref.status)
        if lam > 0:
            assert r.status == 0 and r.relative_energy_error_upper_bound <= 1e-8, lam
        assert rel(r.solution.cpu(), ref.solution.cpu()) <= 1e-6, lam


@pytest.mark.parametrize("layout", ["col", "row"])
@pytest.mark.parametrize("vt", ["float64", "float32", "bfloat16"])
def test_dense_products(torch_mod, fx, layout, vt):
    import amfls_solver_amd as am
    torch = torch_mod
    a = torch.from_numpy(fx["dense_a"]).to(getattr(torch, vt))
    stored = a.double().numpy()
    t = a.t().contiguous().t() if layout == "col" else a.contiguous()
    M = am.DeviceMatrix(t, value_dtype=getattr(torch, vt))
    rng = np.random.default_rng(1)
    for s in (1, 3, 8, 9, 64, 70):
        X = rng.standard_normal((80, s))
        Y = stored @ X
        y = M.apply(torch.from_numpy(X)).cpu().numpy()
        z = M.apply(torch.from_numpy(Y), transpose=True).cpu().numpy()
        assert np.abs(y - Y).max() <= 1e-13 * np.abs(Y).max()
        assert np.abs(z - stored.T @ Y).max() <= 1e-13 * np.abs(stored.T @ Y).max()


def test_validation_models_match_oracle(torch_mod, fx):
    import amfls_solver_amd as am
    from oracle import operators as O
    torch = torch_mod
    M = am.DeviceMatrix(csr_tensor(torch, fx["csr_rp"], fx["csr_ci"], fx["csr_val"], (3000, 700)),
                        value_dtype=torch.float32)
    up, lo, sa, st = M.validation_model()
    assert (sa, st) == (int(fx["csr_model"][2]), int(fx["csr_model"][3]))
    assert np.isclose(up, fx["csr_model"][0], rtol=1e-13) and np.isclose(lo, fx["csr_model"][1], rtol=1e-13)
    a = fx["dense_a"]
    D = am.DeviceMatrix(torch.from_numpy(a))
    mo = O.DenseOperator(a).validation_error_model()
    up, lo, sa, st = D.validation_model()
    assert (sa, st) == (mo[2], mo[3])
    assert np.isclose(up, mo[0], rtol=1e-13) and np.isclose(lo, mo[1], rtol=1e-13)


def _check_trajectory(res, trace, status):
    got = res.trace_float.numpy()
    # depth, basis rank, auxiliary width of every evaluated candidate
    msg = f"engine trace\n{got[:, [1, 3, 2, 13, 14, 15]]}\noracle trace\n{trace[:, [0, 1, 2, 3]]}"
    assert got.shape[0] == trace.shape[0], msg
    assert np.array_equal(got[:, [1, 3, 2]], trace[:, [0, 1, 2]]), msg
    assert res.status == int(status[0]) and res.stop_reason == int(status[1])
    assert res.depth == int(status[2]) and res.basis_rank == int(status[3])


def test_fixed_rbgk_parity(torch_mod, fx):
    import amfls_solver_amd as am
    torch = torch_mod
    r = am.run_fixed_rbgk(torch.from_numpy(fx["dense_a"]), torch.from_numpy(fx["dense_b"]), auxiliary_width=3,
                          maximum_depth=40, tolerance=1e-10, seed=5, clock="synthetic")
    st = fx["rbgk_status"]
    assert (r.status, r.stop_reason, r.depth, r.basis_rank) == tuple(int(v) for v in st)
    assert rel(r.solution.numpy(), fx["rbgk_solution"]) < 1e-8
    got = r.trace_float.numpy()
    assert np.array_equal(got[:, [1, 3]], fx["rbgk_trace"][:, [0, 1]])


@pytest.mark.parametrize("tag,lam", [("l0", 0.0), ("l3", 1e-3)])
def test_amfls_dense_trajectory_parity(torch_mod, fx, tag, lam):
    import amfls_solver_amd as am
    torch = torch_mod
    r = am.run_amfls(torch.from_numpy(fx["dense_a"]), torch.from_numpy(fx["dense_b"]), regularization=lam,
                     tolerance=1e-10, seed=11, clock="synthetic")
    _check_trajectory(r, fx[f"amfls_dense_{tag}_trace"], fx[f"amfls_dense_{tag}_status"])
    assert r.auxiliary_width == int(fx[f"amfls_dense_{tag}_status"][4])
    assert rel(r.solution.numpy(), fx[f"amfls_dense_{tag}_solution"]) < 1e-8


@pytest.mark.parametrize("tag,lam", [("l0", 0.0), ("l3", 1e-3)])
def test_amfls_csr_trajectory_parity(torch_mod, fx, tag, lam):
    import amfls_solver_amd as am
    torch = torch_mod
    A = csr_tensor(torch, fx["csr_rp"], fx["csr_ci"], fx["csr_val"], (3000, 700))
    r = am.run_amfls(A, torch.from_numpy(fx["csr_b"]), regularization=lam, tolerance=1e-9, seed=1,
                     clock="synthetic", value_dtype=torch.float32)
    _check_trajectory(r, fx[f"amfls_csr_{tag}_trace"], fx[f"amfls_csr_{tag}_status"])
    assert rel(r.solution.numpy(), fx[f"amfls_csr_{tag}_solution"]) < 1e-8
    # the certificate bounds the true least-squares backward error of the
    # returned x (recomputed in FP64 with the exact operator norm)
    import scipy.sparse as sp
    As = sp.csr_matrix((fx["csr_val"].astype(np.float64), fx["csr_ci"], fx["csr_rp"]), shape=(3000, 700))
    x = r.solution.numpy()
    res = As @ x - fx["csr_b"]
    g = As.T @ res + lam * x
    k_norm = np.hypot(np.linalg.svd(As.toarray(), compute_uv=False)[0], np.sqrt(lam))
    true_ls = np.linalg.norm(g) / (k_norm * (1 + 1e-12) * np.hypot(np.linalg.norm(res), np.sqrt(lam) * np.linalg.norm(x)))
    assert true_ls <= r.least_squares_backward_error_upper_bound


def test_lsqr_parity(torch_mod, fx):
    """LSQR (reference lsqr.cpp) against the oracle on the CSR fixture.  The
    tolerance sits well above the certificate at the depth-256 checkpoint
    (1.3e-9 in the oracle), so both stop there; LSQR without
    reorthogonalization amplifies rounding differences, so certificates are
    compared to a relative 50 percent and the solution to 1e-7."""
    import amfls_solver_amd as am
    from oracle import operators as O, solvers as S
    torch = torch_mod
    A = csr_tensor(torch, fx["csr_rp"], fx["csr_ci"], fx["csr_val"], (3000, 700))
    r = am.run_lsqr(A, torch.from_numpy(fx["csr_b"]), regularization=1e-3, tolerance=2e-9,
                    value_dtype=torch.float32)
    cop = O.CsrOperator(3000, 700, fx["csr_rp"], fx["csr_ci"], fx["csr_val"].astype(np.float64))
    ro = S.solve_lsqr(cop, fx["csr_b"], regularization=1e-3, tolerance=2e-9)
    assert r.status == int(ro.status) == 0
    assert r.depth == ro.depth == 256
    got = r.trace_float.numpy()
    want = np.array([[t.depth, t.relative_energy_error_upper_bound] for t in ro.trace])
    assert np.array_equal(got[:, 1], want[:, 0])
    fin = np.isfinite(want[:, 1])
    assert np.array_equal(np.isfinite(got[:, 17]), fin)
    assert np.all(np.abs(got[fin, 17] / want[fin, 1] - 1.0) < 0.5)
    assert rel(r.solution.numpy(), ro.solution) < 1e-7
    # the oracle's stopping depth at tolerance 1e-9 (fixture written by the oracle)
    assert int(fx["lsqr_csr_status"][1]) == 260


@pytest.mark.parametrize("lam", [0.0, 1e-3])
def test_mixed_precision_reaches_fp64_tolerance(torch_mod, lam):
    """FP32 search + FP64 refinement certifies the same tolerance and agrees
    with the FP64 oracle's certified solution."""
    import amfls_solver_amd as am
    from amfls_solver_amd import problems as P
    from oracle import operators as O, solvers as S
    torch = torch_mod
    p = P.csr_rows(20000, 2000, 12, seed=3, regularization=lam)
    A = csr_tensor(torch, p.row_ptr, p.col_idx, p.values, (p.rows, p.cols))
    tol = 1e-8
    r = am.run_amfls(A, torch.from_numpy(p.b), regularization=lam, tolerance=tol, precision="mixed",
                     value_dtype=torch.float32)
    assert r.success, (r.status, r.stop_reason, r.refinement_steps)
    cert = r.relative_energy_error_upper_bound if lam > 0 else r.backward_error_upper_bound
    assert cert <= tol
    op = O.CsrOperator(p.rows, p.cols, p.row_ptr, p.col_idx, p.values.astype(np.float64))
    ro = S.solve_amfls(op, p.b, regularization=lam, tolerance=1e-12)
    assert rel(r.solution.numpy(), ro.solution) < 1e-6


def test_mixed_forward_target_falls_back_to_fp64(torch_mod):
    """A mixed solve whose FP64 refinement cannot reach its forward-error
    target within its iteration budget hands the problem to the FP64 engine:
    the result is flagged and is the FP64 engine's own (bitwise, synthetic
    clock), with both runs' checkpoints in its trace; with room to refine,
    the same problem reaches the target on the mixed path (no fallback) and
    agrees with a dense least-squares solve."""
    import amfls_solver_amd as am
    from amfls_solver_amd import problems as P
    torch = torch_mod
    a, b, _ = P.dense_gd(4000, 300, 4.0, seed=12, noise=1e-8, relative_noise=True)
    A = torch.from_numpy(np.ascontiguousarray(a))
    bt = torch.from_numpy(b)
    x_ls = np.linalg.lstsq(a, b, rcond=None)[0]
    kw = dict(tolerance=1e-8, clock="synthetic")
    ok = am.run_amfls(A, bt, precision="mixed", forward_tolerance=1e-9, **kw)
    assert not ok.fp64_fallback
    assert ok.forward_error_estimate <= 1e-9
    assert rel(ok.solution.numpy(), x_ls) < 1e-8
    # a forward target no refinement reaches, and a two-iteration budget
    fb = am.run_amfls(A, bt, precision="mixed", forward_tolerance=1e-30, maximum_refinement_iterations=2, **kw)
    ref = am.run_amfls(A, bt, precision="fp64", **kw)
    assert fb.fp64_fallback and not ref.fp64_fallback
    assert fb.refinement_steps == 2
    assert (fb.status, fb.stop_reason, fb.depth, fb.basis_rank) == (ref.status, ref.stop_reason, ref.depth,
                                                                    ref.basis_rank)
    assert torch.equal(fb.solution, ref.solution)
    assert fb.trace_float.shape[0] > ref.trace_float.shape[0]  # the mixed run's checkpoints too


def test_edge_cases(torch_mod):
    import amfls_solver_amd as am
    from amfls_solver_amd import problems as P
    torch = torch_mod
    a, b, _ = P.dense_gd(60, 20, 2.0, seed=4, noise=1e-3)
    A = torch.from_numpy(a)
    # zero right-hand side: x = 0 is exact
    r = am.run_amfls(A, torch.zeros(60, dtype=torch.float64))
    assert r.success and float(r.solution.abs().max()) == 0.0
    # work limits forced from the host
    r = am.run_amfls(A, torch.from_numpy(b), tolerance=1e-14, maximum_depth=2, clock="synthetic")
    assert r.status == 1 and r.stop_reason == 3 and r.depth == 2
    r = am.run_amfls(A, torch.from_numpy(b), tolerance=1e-14, maximum_basis_size=5, clock="synthetic")
    assert r.status == 2 and r.basis_rank == 5
    # invalid inputs raise, never reach a kernel
    with pytest.raises(ValueError):
        am.run_amfls(A, torch.full((60,), float("nan"), dtype=torch.float64))
    with pytest.raises(ValueError):
        am.run_amfls(A, torch.zeros(59, dtype=torch.float64))
    with pytest.raises(ValueError):
        am.run_amfls(A, torch.from_numpy(b), tolerance=1.5)
    dup = csr_tensor(torch, [0, 2, 3], [1, 1, 0], [1.0, 2.0, 3.0], (2, 2))
    with pytest.raises(ValueError):
        am.DeviceMatrix(dup)
    # empty rows and a single nonzero
    e = csr_tensor(torch, [0, 0, 1, 1], [2], [4.0], (3, 3))
    r = am.run_amfls(e, torch.tensor([0.0, 8.0, 1.0], dtype=torch.float64))
    assert r.success and np.allclose(r.solution.numpy(), [0.0, 0.0, 2.0])


def test_readme_diagonal_example(torch_mod):
    import amfls_solver_amd as am
    torch = torch_mod
    A = torch.diag(torch.tensor([100.0, 3.0, 2.0, 1.0], dtype=torch.float64))
    r = am.run_amfls(A, torch.tensor([100.0, 6.0, 6.0, 4.0], dtype=torch.float64), tolerance=1e-10, seed=7)
    assert r.success
    assert np.allclose(r.solution.numpy(), [1, 2, 3, 4], atol=1e-12)


def test_torch_ops_entry_points(torch_mod, fx):
    """torch.ops.amfls.* on dense host tensors and on device CSR tensors give
    the reference's 8-tuple and the same answers as the Python mirror."""
    import amfls_solver_amd as am
    torch = torch_mod
    A = torch.diag(torch.tensor([100.0, 3.0, 2.0, 1.0], dtype=torch.float64))
    b = torch.tensor([100.0, 6.0, 6.0, 4.0], dtype=torch.float64)
    out = torch.ops.amfls.run_amfls(A, b, 0.0, 1e-10, 1e-6, seed=7)
    assert len(out) == 8 and out[6] == 0
    assert out[1].shape == (21,) and out[2].shape == (36,) and out[3].shape == (9,)
    assert out[4].shape[1] == 29 and out[5].shape[1] == 4
    assert np.allclose(out[0].numpy(), [1, 2, 3, 4], atol=1e-12)
    Ad = csr_tensor(torch, fx["csr_rp"], fx["csr_ci"], fx["csr_val"], (3000, 700)).to("cuda")
    bd = torch.from_numpy(fx["csr_b"]).cuda()
    o1 = torch.ops.amfls.run_lsqr(Ad, bd, 1e-3, 2e-9)
    r1 = am.run_lsqr(Ad, bd, regularization=1e-3, tolerance=2e-9)
    assert o1[6] == r1.status == 0 and np.allclose(o1[0].numpy(), r1.solution.numpy(), rtol=0, atol=1e-12)
    o2 = torch.ops.amfls.run_fixed_rbgk(Ad, bd, 1e-3, 1e-9, 1e-6, 2, 6, 0, 3, 0)
    assert o2[2][9] == 6  # depth counter


def test_torch_ops_every_dispatch_key(torch_mod, fx):
    """run_amfls through each registered key on one ridge problem: host CSR
    (the reference's SparseCsrCPU path), device CSR, host and device dense,
    and a non-contiguous dense view (converted like the reference's
    .contiguous()).  Every call certifies, and its true relative energy error
    against a dense FP64 ridge solve is within the tolerance."""
    torch = torch_mod
    lam, tol = 1e-3, 1e-8
    A_host = csr_tensor(torch, fx["csr_rp"], fx["csr_ci"], fx["csr_val"], (3000, 700))
    b_host = torch.from_numpy(fx["csr_b"])
    D = A_host.to_dense().to(torch.float64)
    Dn = D.numpy()
    H = Dn.T @ Dn + lam * np.eye(Dn.shape[1])
    xs = np.linalg.solve(H, Dn.T @ fx["csr_b"])
    denom = np.sqrt(xs @ H @ xs)
    cases = {
        "csr_cpu": (A_host, b_host),
        "csr_cuda": (A_host.to("cuda"), b_host.cuda()),
        "dense_cpu": (D, b_host),
        "dense_cuda": (D.cuda(), b_host.cuda()),
        "dense_view_cuda": (D.T.contiguous().cuda().T, b_host.cuda()),
    }
    assert not cases["dense_view_cuda"][0].is_contiguous()
    for name, (A, b) in cases.items():
        out = torch.ops.amfls.run_amfls(A, b, lam, tol, 1e-6, seed=5)
        assert out[6] == 0, name
        assert out[0].device.type == "cpu" and out[0].shape == (700,), name
        d = out[0].numpy() - xs
        assert np.sqrt(d @ H @ d) / denom <= tol, name


@pytest.mark.parametrize("vt", ["bfloat16", "float32", "float64"])
def test_dense_matrix_core_products(torch_mod, vt):
    """Row-major dense products: widths up to 8 on both the vector-ALU
    kernels (FP64 accumulation) and the matrix cores, wider blocks on the matrix
    cores; fp64 blocks are fp64-accurate (A upconverted exactly), fp32
    blocks fp32-accurate (split-bf16 or fp32 matrix cores); ragged sizes
    and widths up to 64 (70 falls back to the tall-skinny kernels)."""
    import amfls_solver_amd as am
    torch = torch_mod
    rng = np.random.default_rng(7)
    # (1531, 203, 5) is a view with a padded leading dimension (208), so the
    # 16-byte-aligned A^T kernels run on a column count that is not a
    # multiple of 8 and a last row chunk that is partial; (3001, 520) and
    # (100003, 264) take the direct-to-LDS A^T kernel (FP32 blocks, columns a
    # multiple of 8) with a last strip of one or a few 8-column pieces, a
    # last 64-row trip past the matrix, and chunks of more than six trips
    for m, n, pad in ((1000, 203, 0), (600, 517, 0), (37, 64, 0), (4099, 1, 0), (1531, 203, 5), (3001, 520, 0),
                      (100003, 264, 0)):
        full = torch.from_numpy(rng.standard_normal((m, n + pad))).to(getattr(torch, vt))
        a = full.cuda()[:, :n] if pad else full
        stored = a.double().cpu().numpy()
        M = am.DeviceMatrix(a if pad else a.contiguous(), value_dtype=getattr(torch, vt))
        for s in (1, 5, 16, 17, 32, 40, 64, 70):
            X = rng.standard_normal((n, s))
            Y = stored @ X
            Yg = rng.standard_normal((m, s))
            Z = stored.T @ Yg
            y = M.apply(torch.from_numpy(X)).cpu().numpy()
            z = M.apply(torch.from_numpy(Yg), transpose=True).cpu().numpy()
            assert np.abs(y - Y).max() <= 1e-13 * max(1.0, np.abs(Y).max()) * np.sqrt(n)
            assert np.abs(z - Z).max() <= 1e-13 * max(1.0, np.abs(Z).max()) * np.sqrt(m)
            y32 = M.apply(torch.from_numpy(X).float()).cpu().double().numpy()
            z32 = M.apply(torch.from_numpy(Yg).float(), transpose=True).cpu().double().numpy()
            assert np.abs(y32 - Y).max() <= 1e-5 * np.abs(np.abs(stored) @ np.abs(X)).max()
            assert np.abs(z32 - Z).max() <= 1e-5 * np.abs(np.abs(stored).T @ np.abs(Yg)).max()
            if s <= 8:  # the other kernel family at widths <= 8
                for tr in (False, True):
                    M.set_variant(s, tr, 1 if s == 1 else 0)
                y = M.apply(torch.from_numpy(X)).cpu().numpy()
                z = M.apply(torch.from_numpy(Yg), transpose=True).cpu().numpy()
                assert np.abs(y - Y).max() <= 1e-13 * max(1.0, np.abs(Y).max()) * np.sqrt(n)
                assert np.abs(z - Z).max() <= 1e-13 * max(1.0, np.abs(Z).max()) * np.sqrt(m)
                for tr in (False, True):
                    M.set_variant(s, tr, -1)


def _mlp_data(n, seed):
    r = np.random.default_rng(seed)
    X = (r.standard_normal((n, 12)) * 2.0 + 0.5).astype(np.float32)
    y = (r.random(n) < 0.4).astype(np.float32)
    norm = np.concatenate([np.full(12, 0.5), np.full(12, 0.5)]).astype(np.float32)
    return X, y, norm


def test_width_mlp_kernels_match_oracle(torch_mod):
    """Forward logits and the gradient of the mean BCE from the HIP kernels
    against the float64 CPU oracle (oracle/mlp.py)."""
    from amfls_solver_amd.width_model import WidthModel, init_params
    from oracle import mlp as OM
    torch = torch_mod
    X, y, norm = _mlp_data(300, 4)
    p = init_params(9)
    p[768:832] = 0.03
    m = WidthModel(p, norm)
    z = m.logits(X)
    zo, _ = OM.forward(p.astype(np.float64), norm, X)
    assert np.abs(z - zo).max() <= 1e-4 * max(1.0, np.abs(zo).max())
    loss, g = m.gradient(torch.from_numpy(X), torch.from_numpy(y))
    go = OM.gradient(p.astype(np.float64), norm, X, y)
    assert abs(loss - OM.loss(p.astype(np.float64), norm, X, y)) <= 1e-5
    assert np.abs(g - go).max() <= 1e-5 + 1e-4 * np.abs(go).max()


def test_width_mlp_training_reduces_loss(torch_mod):
    """Adam on the GPU learns a separable rule (label = x0 + x1 > 0)."""
    from amfls_solver_amd.width_model import WidthModel, init_params
    torch = torch_mod
    r = np.random.default_rng(1)
    X = r.standard_normal((512, 12)).astype(np.float32)
    y = (X[:, 0] + X[:, 1] > 0).astype(np.float32)
    norm = np.concatenate([np.zeros(12), np.ones(12)]).astype(np.float32)
    m = WidthModel(init_params(2), norm)
    Xd, yd = torch.from_numpy(X).cuda(), torch.from_numpy(y).cuda()
    first = m.train_step(Xd, yd, 1e-2)
    for _ in range(300):
        last = m.train_step(Xd, yd, 1e-2)
    assert first > 0.3 and last < 0.1
    assert np.mean((m.logits(X) > 0) == (y > 0.5)) > 0.97


def test_width_policies_record_and_learned_certify(torch_mod):
    """The record policy returns one 13-value row per widening decision and
    follows the heuristic; the learned policy (trained model shipped with the
    package, else an untrained one) still certifies: the FP64 certificate is
    the only termination test."""
    import amfls_solver_amd as am
    from amfls_solver_amd import width_model as W
    torch = torch_mod
    A, b, lam, _ = W.training_problem(3)
    M = am.DeviceMatrix(A, value_dtype=torch.float32 if A.layout == torch.sparse_csr else torch.float64)
    rec = W.record(M, b.cuda(), lam, 1e-8)
    assert rec.ndim == 2 and rec.shape[1] == 13 and set(np.unique(rec[:, 12])) <= {0.0, 1.0}
    model = W.WidthModel.load() if W.DEFAULT_PATH.exists() else W.WidthModel(W.init_params(0), np.concatenate(
        [np.zeros(12), np.ones(12)]).astype(np.float32))
    model.attach()
    M2 = am.DeviceMatrix(A, value_dtype=M.value_dtype, context=model.context)
    r = am.run_amfls(M2, b.cuda(), regularization=lam, tolerance=1e-8, width_policy=1, context=model.context)
    assert r.status == 0
    with pytest.raises(ValueError):
        ctx2 = am.Context()
        M3 = am.DeviceMatrix(A, value_dtype=M.value_dtype, context=ctx2)
        am.run_amfls(M3, b.cuda(), regularization=lam, tolerance=1e-8, width_policy=1, context=ctx2)


def test_learned_width_policy_gains_over_heuristic(torch_mod):
    """What the trained network gains: over the held-out problems of its
    training-time evaluation (drawn like the training set, never trained
    on) the learned policy is not slower than the heuristic controller in
    geometric mean, and on the held-out problems where it widened and won,
    it still widens and certifies sooner when run again here; both policies
    certify the same tolerance."""
    from amfls_solver_amd import width_model as W
    model = W.WidthModel.load()
    assert model.meta["heldout_speedup"] >= 1.0
    cases = model.meta["gain_cases"][:3]
    assert len(cases) >= 2, model.meta
    gm, rows = W.evaluate(model, cases, repeats=3, log=print)
    for r in rows:
        assert r["heuristic"][1] == 0 and r["learned"][1] == 0
        assert r["learned"][2] > r["heuristic"][2]  # the model widened
    assert gm > 1.1, rows


def test_learned_width_policy_falls_back_outside_its_domain(torch_mod):
    """Solver states outside the box the model was trained on go to the
    heuristic: with a domain no decision can fall in (active widths of
    2^10..2^11) and the synthetic clock, the learned policy takes no decision
    itself and reproduces the heuristic trajectory exactly."""
    import amfls_solver_amd as am
    from amfls_solver_amd import problems as P
    from amfls_solver_amd import width_model as W
    torch = torch_mod
    model = W.WidthModel.load()
    assert model.domain is not None
    lo, hi = model.domain
    lo, hi = lo.copy(), hi.copy()
    lo[0], hi[0] = 10.0, 11.0
    model.set_domain(lo, hi)
    a, b, _ = P.dense_gd(3000, 300, 4.0, seed=3, noise=1e-6)
    M = am.DeviceMatrix(torch.from_numpy(np.ascontiguousarray(a)), context=model.context)
    model.attach()
    bd = torch.from_numpy(b).cuda()
    kw = dict(tolerance=1e-10, clock="synthetic", context=model.context)
    r_h = am.run_amfls(M, bd, width_policy=0, **kw)
    d0 = model.decisions()
    r_l = am.run_amfls(M, bd, width_policy=1, **kw)
    d1 = model.decisions()
    assert d1[0] == d0[0] and d1[1] > d0[1]  # every decision went to the heuristic
    assert r_l.status == r_h.status and r_l.stop_reason == r_h.stop_reason
    assert (r_l.depth, r_l.auxiliary_width, r_l.basis_rank) == (r_h.depth, r_h.auxiliary_width, r_h.basis_rank)
    assert torch.equal(r_l.solution, r_h.solution)


def test_width_script_replays_and_forks_decisions(torch_mod):
    """The counterfactual replays behind the learned controller's labels:
    scripting a recorded run's own choices reproduces its trajectory and
    solution exactly (synthetic clock), and flipping the first decision
    changes the widths the solve takes."""
    import ctypes
    import amfls_solver_amd as am
    from amfls_solver_amd import _lib
    from amfls_solver_amd import width_model as W
    from amfls_solver_amd.binding import _decode, _result_tuple
    torch = torch_mod
    A, b, lam, _ = W.training_problem(3)
    M = am.DeviceMatrix(A, value_dtype=torch.float32 if A.layout == torch.sparse_csr else torch.float64)
    bd = b.cuda()

    def solve(policy, script=None):
        o = _lib.options(regularization=float(lam), tolerance=1e-8, width_policy=_lib.WIDTH_POLICY[policy],
                         clock_mode=_lib.CLOCK["synthetic"])
        arr = None
        if script is not None:
            arr = np.ascontiguousarray(script, dtype=np.int8)
            o.width_script, o.width_script_length = arr.ctypes.data, arr.size
        res = ctypes.c_void_p()
        torch.cuda.synchronize()
        _lib.check(M.context._lib.amfls_solve(M.context.handle, M.handle, bd.data_ptr(), 1, ctypes.byref(o),
                                              ctypes.byref(res)))
        rec = None
        if policy == "record":
            rows = ctypes.c_int64()
            _lib.check(M.context._lib.amfls_result_width_records(res, None, ctypes.byref(rows)))
            rec = np.empty((rows.value, 13), dtype=np.float32)
            _lib.check(M.context._lib.amfls_result_width_records(res, rec.ctypes.data_as(ctypes.c_void_p),
                                                                 ctypes.byref(rows)))
        return _decode(*_result_tuple(M.context, res)), rec

    base, rec = solve("record")
    assert rec.shape[0] >= 1
    chosen = rec[:, 12].astype(np.int8)
    again, _ = solve("heuristic", chosen)
    assert (again.depth, again.auxiliary_width, again.basis_rank) == (base.depth, base.auxiliary_width,
                                                                        base.basis_rank)
    assert torch.equal(again.solution, base.solution)
    flipped, _ = solve("heuristic", [1 - chosen[0]])
    assert flipped.status == 0
    assert flipped.auxiliary_width != base.auxiliary_width or flipped.depth != base.depth


def _edge_problems():
    """Small shapes at the edges of the engine's dispatch: a wide matrix
    (minimum-norm and regularized), one column, one row, CSR with empty rows
    and columns, and a CSR matrix with no entries."""
    from oracle import operators as O

    def dense(m, n, seed):
        r = np.random.default_rng(seed)
        return r.standard_normal((m, n)), r.standard_normal(m)
    out = []
    a, b = dense(30, 80, 1)
    out += [("wide_l0", a, None, b, 0.0, O.RowMajorDenseOperator(a)),
            ("wide_l2", a, None, b, 1e-2, O.RowMajorDenseOperator(a))]
    a, b = dense(50, 1, 2)
    out.append(("onecol", a, None, b, 0.0, O.RowMajorDenseOperator(a)))
    a, b = dense(1, 20, 3)
    out.append(("onerow", a, None, b, 1e-3, O.RowMajorDenseOperator(a)))
    rp, ci, va = np.array([0, 0, 2, 2, 3, 3]), np.array([0, 3, 3]), np.array([2.0, -1.0, 4.0])
    out.append(("csr_gaps", None, (rp, ci, va, (5, 6)), np.array([1.0, 2.0, -1.0, 3.0, 0.5]), 1e-2,
                O.CsrOperator(5, 6, rp, ci, va)))
    rp0 = np.zeros(5, np.int64)
    out.append(("csr_zero", None, (rp0, np.zeros(0, np.int64), np.zeros(0), (4, 3)), np.ones(4), 0.0,
                O.CsrOperator(4, 3, rp0, np.zeros(0, np.int64), np.zeros(0))))
    return out


@pytest.mark.parametrize("case", range(6))
def test_edge_shapes_match_oracle(torch_mod, case):
    """Status, stop reason, depth, basis rank and solution of the engine equal
    the oracle's on small edge shapes (synthetic clock, so both take the same
    width decisions)."""
    import amfls_solver_amd as am
    from oracle import solvers as S
    from oracle.clock import SyntheticClock
    torch = torch_mod
    name, a, csr, b, lam, op = _edge_problems()[case]
    ro = S.solve_amfls(op, b, regularization=lam, tolerance=1e-10, seed=3, clock=SyntheticClock())
    if a is not None:
        A = torch.from_numpy(np.ascontiguousarray(a))
    else:
        rp, ci, va, shape = csr
        A = csr_tensor(torch, rp, ci, va, shape)
    r = am.run_amfls(A, torch.from_numpy(b), regularization=lam, tolerance=1e-10, seed=3, clock="synthetic")
    assert (r.status, r.stop_reason, r.depth, r.basis_rank) == \
        (int(ro.status), int(ro.stop_reason), ro.depth, ro.basis_rank), name
    if np.abs(ro.solution).max() == 0.0:
        assert float(r.solution.abs().max()) == 0.0, name
    else:
        # certified results agree to rounding; an uncertified one (the
        # regularized wide case stops at the basis limit) only to what its
        # certificate bounds, since rounding differs once the basis spans A
        tol = 1e-8 if ro.status == 0 else max(1e-8, 10.0 * ro.relative_energy_error_upper_bound)
        assert rel(r.solution.numpy(), ro.solution) < tol, name
