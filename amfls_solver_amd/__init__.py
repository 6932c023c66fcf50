"""MI355X-native AMFLS adaptive-width block Golub-Kahan least-squares engine.

The solver runs in libamfls_hip.so (hand-written gfx950 HIP kernels and a
C++20 host engine behind a C ABI, include/amfls_hip.h); this package is the
Python mirror of the reference's binding (run_amfls, run_fixed_rbgk,
run_lsqr, SolveResult) plus the torch.ops.amfls registration, and the
engine's extensions: regularization sweeps (run_amfls_sweep), concurrent
batches of CSR problems (run_amfls_batch) and the SpMM autotuner
(DeviceMatrix.autotune) and the certified validation of a given
candidate (validate_candidate).
"""
from .binding import (  # noqa: F401
    Context,
    DeviceMatrix,
    HostOperator,
    SolveResult,
    run_amfls,
    run_amfls_batch,
    run_amfls_sweep,
    run_fixed_rbgk,
    run_lsqr,
    validate_candidate,
)

from . import torch_ops as _torch_ops

_torch_ops.register()

__all__ = ["Context", "DeviceMatrix", "HostOperator", "SolveResult", "run_amfls", "run_amfls_batch", "run_amfls_sweep",
           "run_fixed_rbgk", "run_lsqr", "validate_candidate"]
