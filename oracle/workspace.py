"""Variable-width block Golub-Kahan workspace (oracle).

Follows src/algorithms/mathematics/krylov/krylov_workspace.cpp:
rank_revealing_cutoff (:63-80), orthogonalize_and_factor (:82-310),
stable_ridge_filter (:312-331), the workspace constructor (:335-381),
initialize (:470-503), append_left_seed (:505-587), advance_level (:589-831),
evaluate (:833-925) and the step-audit predicates (:965-979).

Thin SVD is numpy.linalg.svd(full_matrices=False) (LAPACK dgesdd, as
svd.cpp:23-60); the lambda = 0 projected solve is scipy's dgelsy with the
reference's rcond (svd.cpp:90-153).

Pinned by tests/test_oracle_solvers.py: V^T V = I, nested decrease of the
projected objective, the candidate equal to the least-squares solution over
span(V) computed directly from A (test_workspace_krylov_invariants), and
on a rank-deficient operator the basis stops at the rank with the
minimum-norm solution (test_rank_deficient_operator_stops_at_rank).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
import scipy.linalg as sla

from . import results as R
from .validation import Counted, validate_original_candidate

EPS = float(np.finfo(np.float64).eps)
INF = math.inf
LD = np.longdouble


@dataclass
class StepAudit:
    requested_injected_columns: int = 0
    retained_injected_columns: int = 0
    left_input_width: int = 0
    retained_right_width: int = 0
    retained_next_left_width: int = 0
    basis_rank_before: int = 0
    basis_rank_after: int = 0
    right_rank_cutoff: float = 0.0
    right_discarded_residual_norm: float = 0.0
    right_active_discarded_residual_norm: float = 0.0
    left_rank_cutoff: float = 0.0
    left_discarded_residual_norm: float = 0.0
    right_rank_loss: bool = False
    left_rank_loss: bool = False
    right_positive_finite_below_cutoff: bool = False
    left_positive_finite_below_cutoff: bool = False
    right_nonfinite_residual: bool = False
    left_nonfinite_residual: bool = False
    right_basis_cap_truncation: bool = False
    left_basis_cap_truncation: bool = False


def step_has_numerical_rank_truncation(s: StepAudit) -> bool:
    return (s.right_rank_loss and s.right_positive_finite_below_cutoff
            and s.right_active_discarded_residual_norm > 0.0) or \
        (s.left_rank_loss and s.left_positive_finite_below_cutoff)


def step_has_nonfinite_breakdown(s: StepAudit) -> bool:
    return s.right_nonfinite_residual or s.left_nonfinite_residual


def rank_revealing_cutoff(scale, m, n, k, w):
    if not (scale >= 0.0) or not math.isfinite(scale):
        return INF
    return 64.0 * EPS * max(1, m, n, k, w) * scale


@dataclass
class Factor:
    coefficients: np.ndarray
    q: np.ndarray
    factor: np.ndarray
    discarded: float = 0.0
    discarded_tracked: float = 0.0
    cutoff: float = 0.0
    positive_finite_below_cutoff: bool = False
    nonfinite_residual: bool = False


def _norm(v) -> float:
    return float(np.linalg.norm(v))


def orthogonalize_and_factor(block, basis, max_rank, m, n, scale, tracked, retain, clock, stats):
    """krylov_workspace.cpp:82-310."""
    N, w = block.shape
    k = basis.shape[1]
    start = clock.now()
    clock.charge_local(w, k)
    block = block.copy()
    retain_scalar = w != 1 or retain
    coeffs = np.zeros((k if retain_scalar else 0, w))
    coeffs_finite = True
    if w == 1 and k > 0:
        before = _norm(block)
        acc = np.zeros(k)

        def projection_pass():
            c = basis.T @ block[:, 0]
            if retain_scalar:
                coeffs[:, 0] += c
            else:
                acc[:] += c
            block[:, 0] -= basis @ c
        projection_pass()
        resid = _norm(block)
        if math.isfinite(before) and math.isfinite(resid) and resid < math.sqrt(0.5) * before:
            projection_pass()
            resid = _norm(block)
        if not retain_scalar:
            coeffs_finite = bool(np.all(np.isfinite(acc)))
    else:
        for _ in range(2 if k > 0 else 0):
            c = basis.T @ block
            coeffs += c
            block -= basis @ c
        resid = _norm(block)
    cutoff = rank_revealing_cutoff(scale, m, n, k, w)
    finite = math.isfinite(scale) and math.isfinite(resid) and coeffs_finite and \
        bool(np.all(np.isfinite(coeffs)))
    f = Factor(coeffs, np.zeros((N, 0)), np.zeros((0, w)), resid, 0.0, cutoff, False, not finite)
    if not finite:
        f.discarded = INF
        f.discarded_tracked = INF if tracked > 0 else 0.0
        stats.orthogonalization_seconds += clock.now() - start
        return f
    if max_rank == 0 or w == 0:
        t = 0.0
        for c in range(tracked):
            t = math.hypot(t, _norm(block[:, c]))
        f.discarded_tracked = t
        stats.orthogonalization_seconds += clock.now() - start
        return f
    if w == 1:
        if math.isfinite(resid) and resid > cutoff:
            f.q = block * (1.0 / resid)
            f.factor = np.array([[resid]])
            f.discarded = 0.0
            f.discarded_tracked = 0.0
        elif math.isfinite(resid) and 0.0 < resid <= cutoff and max_rank > 0:
            f.positive_finite_below_cutoff = True
            if tracked == 1:
                f.discarded_tracked = resid
        stats.orthogonalization_seconds += clock.now() - start
        return f
    u, s, vt = np.linalg.svd(block, full_matrices=False)
    rank = 0
    while rank < s.size and rank < max_rank and math.isfinite(s[rank]) and s[rank] > cutoff:
        rank += 1
    if rank < s.size and rank < max_rank:
        fd = float(s[rank])
        f.positive_finite_below_cutoff = math.isfinite(fd) and 0.0 < fd <= cutoff
    if rank > 0:
        f.q = u[:, :rank].copy()
        f.factor = s[:rank, None] * vt[:rank, :]
    f.discarded = 0.0
    for i in range(rank, s.size):
        f.discarded = math.hypot(f.discarded, float(s[i]))
        tw = 0.0
        for c in range(tracked):
            tw = math.hypot(tw, float(vt[i, c]))
        f.discarded_tracked = math.hypot(f.discarded_tracked, float(s[i]) * tw)
    stats.orthogonalization_seconds += clock.now() - start
    return f


def stable_ridge_filter(sigma: float, lam: float) -> float:
    """krylov_workspace.cpp:312-331."""
    if not (sigma > 0.0) or not (lam > 0.0):
        return 0.0
    root = math.sqrt(lam)
    if not (root > 0.0) or not math.isfinite(root):
        return sigma / (sigma * sigma + lam)
    if sigma <= root:
        ratio = sigma / root
        return (ratio / (1.0 + ratio * ratio)) / root
    inv = root / sigma
    return (1.0 / sigma) / (1.0 + inv * inv)


class Workspace:
    def __init__(self, op, b, lam, maximum_basis_size, stats: R.RunStatistics, clock):
        if op.rows <= 0 or op.cols <= 0:
            raise ValueError("block Golub-Kahan operator dimensions must be positive")
        if not math.isfinite(lam) or lam < 0.0:
            raise ValueError("block Golub-Kahan regularization must be finite and nonnegative")
        if maximum_basis_size < 0:
            raise ValueError("block Golub-Kahan basis limit must be nonnegative")
        b = np.asarray(b, dtype=np.float64)
        if not np.all(np.isfinite(b)):
            raise ValueError("block Golub-Kahan right-hand side must be finite")
        self.op, self.m, self.n = op, op.rows, op.cols
        self.b = b.copy()
        self.bnorm = _norm(b)
        self.lam = lam
        mn = min(self.m, self.n)
        self.basis_limit = min(maximum_basis_size, mn) if maximum_basis_size > 0 else mn
        self.stats, self.clock = stats, clock
        self.counted = Counted(op, stats, clock)
        self.U = np.zeros((self.m, 0))
        self.pending = np.zeros((self.m, 0))
        self.V = np.zeros((self.n, 0))
        self.T = np.zeros((min(self.basis_limit + 1, 8), 0))
        self.g: list = []
        self.active_first = self.active_width = 0
        self.prev_first = self.prev_width = 0
        self.coupling = np.zeros((0, 0))
        self.pending_requested = self.pending_retained = 0
        self.depth = 0
        self.initialized = False
        self.reliable = True
        self.audit = StepAudit()

    # --- operator search calls (krylov_workspace.cpp:383-405) ---
    def _apply(self, x, transpose: bool):
        s, c = self.stats, self.clock
        start = c.now()
        c.charge_operator(self.op.relative_block_product_cost(x.shape[1]))
        y = (self.op.apply_transpose if transpose else self.op.apply)(x)
        y = np.asarray(y, dtype=np.float64).reshape(self.n if transpose else self.m, x.shape[1])
        sec = c.now() - start
        if transpose:
            s.at_seconds += sec
            s.at_columns += x.shape[1]
            s.search_at_columns += x.shape[1]
            s.at_block_calls += 1
            s.search_at_block_calls += 1
        else:
            s.a_seconds += sec
            s.a_columns += x.shape[1]
            s.search_a_columns += x.shape[1]
            s.a_block_calls += 1
            s.search_a_block_calls += 1
        return y

    def can_advance(self) -> bool:
        return self.initialized and self.reliable and \
            self.U.shape[1] + self.pending.shape[1] > 0 and self.V.shape[1] < self.basis_limit

    @property
    def basis_rank(self) -> int:
        return self.V.shape[1]

    def _ensure_rows(self, required):
        maxrows = self.basis_limit + 1
        if required > maxrows:
            raise RuntimeError("block Golub-Kahan projected row limit was exceeded")
        if required <= self.T.shape[0]:
            return
        grown = max(1, self.T.shape[0])
        while grown < required:
            grown = maxrows if grown > maxrows // 2 else grown * 2
        t = np.zeros((grown, self.T.shape[1]))
        t[:self.T.shape[0]] = self.T
        self.T = t

    def projected_operator(self):
        return self.T[:len(self.g), :self.V.shape[1]].copy()

    def initialize(self, aux=None):
        if self.initialized:
            raise RuntimeError("block Golub-Kahan workspace is already initialized")
        self.initialized = True
        self.audit = StepAudit()
        start = self.clock.now()
        if math.isfinite(self.bnorm) and self.bnorm > 0.0 and self.basis_limit > 0:
            self.U = (self.b * (1.0 / self.bnorm)).reshape(-1, 1)
            self.g.append(self.bnorm)
            self.active_first, self.active_width = 0, 1
        self.stats.orthogonalization_seconds += self.clock.now() - start
        if self.bnorm > 0.0 and aux is not None and aux.shape[1] > 0:
            self.append_left_seed(aux)

    def append_left_seed(self, seed) -> int:
        """krylov_workspace.cpp:505-587."""
        if not self.initialized:
            raise RuntimeError("block Golub-Kahan workspace is not initialized")
        if not self.reliable:
            raise RuntimeError("cannot continue block Golub-Kahan after numerical truncation")
        if self.pending_requested == 0 and self.pending.shape[1] == 0:
            self.audit = StepAudit()
        self.pending_requested += seed.shape[1]
        self.audit.requested_injected_columns = self.pending_requested
        self.audit.retained_injected_columns = self.pending_retained
        if seed.shape[1] == 0:
            return 0
        avail = max(0, self.basis_limit - self.V.shape[1] - self.U.shape[1] - self.pending.shape[1])
        max_rank = min(avail, self.m - self.U.shape[1] - self.pending.shape[1], seed.shape[1])
        if max_rank <= 0:
            return 0
        local = np.hstack([self.U, self.pending]) if self.pending.shape[1] else self.U
        scale = _norm(seed)
        f = orthogonalize_and_factor(seed, local, max_rank, self.m, self.n, scale, 0, False,
                                     self.clock, self.stats)
        kept = f.q.shape[1]
        if kept <= 0:
            return 0
        self.pending = np.hstack([self.pending, f.q]) if self.pending.shape[1] else f.q
        self.pending_retained += kept
        self.audit.retained_injected_columns = self.pending_retained
        return kept

    def advance_level(self) -> bool:
        """krylov_workspace.cpp:589-831."""
        req, ret = self.pending_requested, self.pending_retained
        a = StepAudit(requested_injected_columns=req, retained_injected_columns=ret,
                      basis_rank_before=self.V.shape[1], basis_rank_after=self.V.shape[1],
                      left_input_width=self.U.shape[1] + self.pending.shape[1])
        self.audit = a
        if not self.can_advance():
            return False
        if self.U.shape[1] != self.active_width or self.active_first + self.active_width != len(self.g):
            raise RuntimeError("block Golub-Kahan active left frontier is inconsistent")
        old_right = self.V.shape[1]
        old_left = len(self.g)
        aw = self.U.shape[1]
        combined = np.hstack([self.U, self.pending]) if self.pending.shape[1] else self.U
        R_ = self._apply(combined, True)
        self.pending = np.zeros((self.m, 0))
        self.pending_requested = self.pending_retained = 0
        rscale = _norm(R_)
        if self.prev_width > 0 and aw > 0:
            if self.coupling.shape != (aw, self.prev_width):
                raise RuntimeError("block Golub-Kahan frontier coupling dimensions are invalid")
            cc = self.coupling
            if combined.shape[1] != aw:
                cc = np.zeros((combined.shape[1], self.prev_width))
                cc[:aw] = self.coupling
            Vp = self.V[:, self.prev_first:self.prev_first + self.prev_width]
            R_ = R_ - Vp @ cc.T
        max_right = min(combined.shape[1], self.basis_limit - self.V.shape[1])
        rf = orthogonalize_and_factor(R_, self.V, max_right, self.m, self.n, rscale, aw, False,
                                      self.clock, self.stats)
        rw = rf.q.shape[1]
        a.retained_right_width = rw
        a.right_rank_cutoff = rf.cutoff
        a.right_discarded_residual_norm = rf.discarded
        a.right_active_discarded_residual_norm = rf.discarded_tracked
        a.right_positive_finite_below_cutoff = rf.positive_finite_below_cutoff
        a.right_nonfinite_residual = rf.nonfinite_residual
        a.right_rank_loss = rw < max_right
        a.right_basis_cap_truncation = max_right < combined.shape[1]
        if step_has_numerical_rank_truncation(a) or step_has_nonfinite_breakdown(a):
            self.reliable = False
        if rw == 0:
            self.U = np.zeros((self.m, 0))
            self.active_first, self.active_width = old_left, 0
            self.prev_first, self.prev_width = self.V.shape[1], 0
            self.coupling = np.zeros((0, 0))
            return False
        newV = rf.q
        L = self._apply(newV, False)
        lscale = _norm(L)
        coup = rf.factor[:, :aw]
        if aw > 0:
            L = L - self.U @ coup.T
        max_left = min(rw, self.m - old_left, self.basis_limit + 1 - old_left)
        lf = orthogonalize_and_factor(L, self.U, max_left, self.m, self.n, lscale, 0, True,
                                      self.clock, self.stats)
        lw = lf.q.shape[1]
        a.retained_next_left_width = lw
        a.left_rank_cutoff = lf.cutoff
        a.left_discarded_residual_norm = lf.discarded
        a.left_positive_finite_below_cutoff = lf.positive_finite_below_cutoff
        a.left_nonfinite_residual = lf.nonfinite_residual
        a.left_rank_loss = lw < max_left
        a.left_basis_cap_truncation = max_left < rw
        if step_has_numerical_rank_truncation(a) or step_has_nonfinite_breakdown(a):
            self.reliable = False
        if lf.nonfinite_residual:
            return False
        self.V = np.hstack([self.V, newV]) if self.V.shape[1] else newV
        self._ensure_rows(old_left + lw)
        newcols = np.zeros((self.T.shape[0], rw))
        for c in range(rw):
            for r in range(aw):
                newcols[self.active_first + r, c] = coup[c, r] + lf.coefficients[r, c]
            for r in range(lw):
                newcols[old_left + r, c] = lf.factor[r, c]
        self.T = np.hstack([self.T, newcols])
        self.U = lf.q
        for c in range(lw):
            self.g.append(float(np.dot(self.U[:, c], self.b)))
        self.active_first, self.active_width = old_left, lw
        self.prev_first, self.prev_width = old_right, rw
        self.coupling = lf.factor
        self.depth += 1
        a.basis_rank_after = self.V.shape[1]
        return True

    def projected_solution(self):
        """The projected solve of krylov_workspace.cpp:855-908 (returns x)."""
        T = self.projected_operator()
        g = np.array(self.g, dtype=np.float64)
        k = self.V.shape[1]
        if k == 0:
            return np.zeros(self.n)
        rel = min(0.5, 64.0 * EPS * max(T.shape))
        if self.lam == 0.0:
            y = sla.lstsq(T, g, cond=rel, lapack_driver="gelsy", check_finite=False)[0]
        else:
            u, s, vt = np.linalg.svd(T, full_matrices=False)
            spec = np.array([stable_ridge_filter(float(s[i]), self.lam) * float(np.dot(u[:, i], g))
                             for i in range(s.size)])
            y = (vt.T.astype(LD) @ spec.astype(LD)).astype(np.float64)
        return self.V @ y

    def evaluate(self, tol, epoch, aux_width) -> R.LeastSquaresResult:
        """krylov_workspace.cpp:833-925."""
        if not self.initialized:
            raise RuntimeError("block Golub-Kahan workspace is not initialized")
        if not math.isfinite(tol) or tol <= 0.0 or tol >= 1.0:
            raise ValueError("block Golub-Kahan tolerance must lie strictly between zero and one")
        start = self.clock.now()
        x = self.projected_solution()
        self.stats.projected_solve_seconds += self.clock.now() - start
        res = validate_original_candidate(self.counted, self.b, x, self.lam, tol)
        res.depth = self.depth
        res.auxiliary_width = aux_width
        res.basis_rank = self.V.shape[1]
        res.trace.append(R.make_iteration_record(res, epoch, self.depth, aux_width))
        return res
