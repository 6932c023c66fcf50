"""GPU tests of graph-captured width-one level batches (context option
set_level_graphs): with the batches captured into HIP graphs, solves equal
the directly enqueued solves bitwise on the synthetic clock, and certify with
the same trajectory on the wall clock, where the CSR products stamp their
own times inside the captured kernels.  Graphs and direct solves alternate
on one context, with host synchronization between them, and every solve is
compared with the direct one."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(torch, rows, cols, seed, context):
    import amfls_solver_amd as am
    from amfls_solver_amd.binding import gaussian_block, generate_csr
    # rows scaled by log-uniform factors in [1e-3, 1]: about a hundred levels, so
    # the solve runs at least one full 64-level batch
    rp, ci, val = generate_csr(rows, cols, (4, 32), seed=seed, scale_range=(1e-3, 1.0), value_dtype=torch.float32,
                               context=context)
    M = am.DeviceMatrix.from_csr_arrays(rows, cols, rp, ci, val, context=context)
    b = gaussian_block(rows, 1, 500 + seed, 0, context=context).reshape(-1)
    return M, b


@pytest.mark.parametrize("lam", [1e-3, 1e-5])
def test_level_graphs_match_direct_enqueue(gpu, lam):
    import torch
    import amfls_solver_amd as am
    from amfls_solver_amd.binding import Context
    ctx = Context()
    M, b = _problem(torch, 60000, 6000, 21, ctx)
    kw = dict(regularization=lam, tolerance=1e-8, clock="synthetic", context=ctx)
    ctx.set_level_graphs(False)
    ref = am.run_amfls(M, b, **kw)
    assert ref.status == 0 and ref.depth > 64  # at least two full level batches
    for rep in range(3):
        ctx.set_level_graphs(True)
        g = am.run_amfls(M, b, **kw)
        ctx.set_level_graphs(False)
        d = am.run_amfls(M, b, **kw)
        ctx.synchronize()
        for r in (g, d):
            assert (r.status, r.depth, r.basis_rank, r.auxiliary_width) == \
                (ref.status, ref.depth, ref.basis_rank, ref.auxiliary_width), rep
            assert torch.equal(r.solution, ref.solution), rep
            # every certificate column of every checkpoint, bitwise (NaN marks fields a
            # checkpoint did not compute)
            torch.testing.assert_close(r.trace_float, ref.trace_float, rtol=0, atol=0, equal_nan=True)


def test_level_graphs_wall_clock_certify(gpu):
    import torch
    import amfls_solver_amd as am
    from amfls_solver_amd.binding import Context
    ctx = Context()
    M, b = _problem(torch, 80000, 8000, 5, ctx)
    kw = dict(regularization=1e-3, tolerance=1e-8, context=ctx)
    ctx.set_level_graphs(False)
    ref = am.run_amfls(M, b, **kw)
    ctx.set_level_graphs(True)
    for _ in range(2):
        g = am.run_amfls(M, b, **kw)
        assert g.status == 0
        assert g.relative_energy_error_upper_bound <= 1e-8
        # products stamped inside the graphs are charged: positive operator seconds
        assert g.a_seconds > 0.0 and g.at_seconds > 0.0
        # both certified to a relative energy error of 1e-8 (the width
        # decisions of a wall-clock solve may differ between runs)
        rel = float(torch.linalg.norm(g.solution - ref.solution) / torch.linalg.norm(ref.solution))
        assert rel <= 1e-6, rel
    ctx.set_level_graphs(False)
