"""GPU runs of the benchmark configurations at their full sizes.

* 2000 x 500 dense G D (four decades, lambda 0): the FP64 engine follows the
  oracle's trajectory and certificates on the synthetic clock, and the mixed
  engine with a forward-error target reaches the FP64 least-squares solution
  (numpy lstsq).
* 1e6 x 2000 dense G D stored as bf16 (six decades, lambda 0): the FP64 and
  the mixed-precision engines both reach a forward relative error of at most
  1e-8 against the FP64 oracle solution in tests/golden/c2_solution.npz
  (TSQR of the stored matrix, computed by tests/golden/make_c2_fixture.py);
  the matrix built on the GPU is first shown to be the fixture's bit for bit.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

C2_FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "c2_solution.npz")


def rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def torch_mod(gpu):
    import torch
    return torch


def bf16_checksums(torch, a):
    """The three bit checksums of make_c2_fixture.py over a row-major bf16 A."""
    rows, cols = a.shape
    tot = np.zeros(3, dtype=np.int64)
    wcol = torch.arange(1, cols + 1, device=a.device, dtype=torch.int64)[None, :]
    for i0 in range(0, rows, 50_000):
        i1 = min(rows, i0 + 50_000)
        bits = a[i0:i1].view(torch.int16).to(torch.int64) & 0xFFFF
        wrow = (torch.arange(i0, i1, device=a.device, dtype=torch.int64) % 1009 + 1)[:, None]
        tot += np.array([int(bits.sum()), int((bits * wcol).sum()), int((bits * wrow).sum())], dtype=np.int64)
    return tot


def test_config0_dense_2000x500(torch_mod):
    """C0: trajectory and certificate parity with the oracle (synthetic
    clock), and the mixed engine's forward error against lstsq."""
    import amfls_solver_amd as am
    from amfls_solver_amd import problems as P
    from oracle import operators as O, solvers as S
    from oracle.clock import SyntheticClock
    torch = torch_mod
    a, b, _ = P.config0()
    x_ls = np.linalg.lstsq(a, b, rcond=None)[0]
    A = torch.from_numpy(a)
    ro = S.solve_amfls(O.DenseOperator(a), b, tolerance=1e-8, clock=SyntheticClock())
    r = am.run_amfls(A, torch.from_numpy(b), tolerance=1e-8, clock="synthetic")
    got = r.trace_float.numpy()
    want = np.array([[t.depth, t.basis_rank, t.auxiliary_width] for t in ro.trace], dtype=float)
    assert np.array_equal(got[:, [1, 3, 2]], want), (got[:, [1, 3, 2]], want)
    # same outcome as the oracle: on this workload the reference's
    # certificate does not reach 1e-8 before the basis spans R^500
    assert (r.status, r.stop_reason, r.depth, r.basis_rank) == (ro.status, ro.stop_reason, ro.depth, ro.basis_rank)
    assert rel(r.solution.numpy(), ro.solution) < 1e-8
    # certificates of every evaluated candidate within 1e-6 of the oracle's
    want_be = np.array([t.backward_error_upper_bound for t in ro.trace])
    assert np.allclose(got[:, 15], want_be, rtol=1e-6), (got[:, 15], want_be)
    # the returned certificate bounds the backward error recomputed on the host
    x = r.solution.numpy()
    res = b - a @ x
    be = min(np.linalg.norm(res) / (np.linalg.norm(a, 2) * np.linalg.norm(x) + np.linalg.norm(b)),
             np.linalg.norm(a.T @ res) / (np.linalg.norm(a, 2) * np.linalg.norm(res)))
    assert be <= r.backward_error_upper_bound * (1 + 1e-6), (be, r.backward_error_upper_bound)
    # mixed precision with a forward-error target reaches the FP64 solution;
    # its certificate ends at the same rounding floor (the exact solution's
    # own backward error is 2.9e-8 here)
    rm = am.run_amfls(A, torch.from_numpy(b), tolerance=1e-8, precision="mixed", forward_tolerance=1e-9)
    assert rm.backward_error_upper_bound <= 5e-8, rm.backward_error_upper_bound
    assert rm.refinement_steps <= 64
    fe = rel(rm.solution.numpy(), x_ls)
    assert fe <= 1e-8, fe
    assert rm.forward_error_estimate <= 1e-9
    # the estimate tracks the true error to within a factor 10 (or both tiny)
    assert fe <= max(10 * rm.forward_error_estimate, 1e-12)


@pytest.fixture(scope="module")
def config2(torch_mod):
    from amfls_solver_amd import problems as P
    a, b, xs = P.config2_device()
    fx = np.load(C2_FIXTURE)
    return a, b, fx


def test_config2_matrix_is_the_fixture(torch_mod, config2):
    a, b, fx = config2
    assert tuple(a.shape) == (int(fx["rows"]), int(fx["cols"]))
    assert np.array_equal(bf16_checksums(torch_mod, a), fx["bits_checksums"])
    assert np.isclose(float(torch_mod.linalg.vector_norm(b)), float(fx["b_norm"]), rtol=1e-12)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_config2_forward_error(torch_mod, config2, precision):
    """C2: the FP64 engine at tolerance 1e-8 and the mixed engine with a
    forward-error target both land within 1e-8 of the FP64 oracle solution."""
    import amfls_solver_amd as am
    torch = torch_mod
    a, b, fx = config2
    M = am.DeviceMatrix(a, value_dtype=torch.bfloat16)
    kw = {"forward_tolerance": 1e-9} if precision == "mixed" else {}
    r = am.run_amfls(M, b, tolerance=1e-8, precision=precision, **kw)
    assert r.success, (r.status, r.stop_reason)
    fe = rel(r.solution.numpy(), fx["x_ls"])
    print(f"{precision}: forward error {fe:.3e} estimate {r.forward_error_estimate:.3e} depth {r.depth} "
          f"rank {r.basis_rank} refinement {r.refinement_steps}")
    assert fe <= 1e-8, fe
    if precision == "mixed":
        assert r.forward_error_estimate <= 1e-9
        assert fe <= 10 * r.forward_error_estimate + 1e-12
        assert 0 < r.refinement_steps <= 64
