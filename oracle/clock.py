"""Wall and synthetic clocks for the oracle solvers.

The reference reads std::chrono::steady_clock around operator calls and
search intervals (krylov_workspace.cpp:383-405, amfls.cpp:500-529), which makes
the adaptive width schedule timing dependent.  The synthetic clock advances
only by deterministic charges so two implementations make identical width
decisions:

* operator call of s columns: ``op_unit * relative_block_product_cost(s)``
* orthogonalization of an N x w block against k basis columns:
  ``local_unit * w * (k + w)``

The HIP engine implements the same charges (its ``clock_mode`` option).
"""
from __future__ import annotations

import time

OP_UNIT = 1.0e-3
LOCAL_UNIT = 1.0e-6


class WallClock:
    synthetic = False

    def now(self) -> float:
        return time.perf_counter()

    def charge_operator(self, cost: float) -> None:
        pass

    def charge_local(self, block_cols: int, basis_cols: int) -> None:
        pass


class SyntheticClock:
    synthetic = True

    def __init__(self, op_unit: float = OP_UNIT, local_unit: float = LOCAL_UNIT):
        self.t = 0.0
        self.op_unit = op_unit
        self.local_unit = local_unit

    def now(self) -> float:
        return self.t

    def charge_operator(self, cost: float) -> None:
        self.t += self.op_unit * cost

    def charge_local(self, block_cols: int, basis_cols: int) -> None:
        self.t += self.local_unit * block_cols * (basis_cols + block_cols)
