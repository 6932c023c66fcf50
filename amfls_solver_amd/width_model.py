"""Learned width controller: data, GPU training and evaluation.

The engine's widening decision (reference src/algorithms/amfls/amfls.cpp:
322-358, cost gate amfls_internal.hpp:318-324) can be taken by a 12 -> 64 ->
64 -> 1 ReLU MLP (csrc/device/mlp.hip) instead of the heuristic.  Its inputs
are what AMFLS records about itself at each decision: the active width, the
last four log10 certificate decreases, the last four per-level interval
times, and log10 of the operator's rows, columns and nonzeros.

Labels.  Each decision is labelled by a counterfactual: the solve is
replayed from the start with every earlier decision scripted to what the
recorded run chose (the engine's width_script option; products and
reductions are deterministic, so the replay reaches the same state), the
decision itself forced to widen and then to hold, and the heuristic taking
every later decision; the decision is labelled "widen" when the widened
replay certified sooner (wall clock on the GPU, best of two runs for short
solves) and dropped when the two are within 3 percent.  Decisions are taken
along the heuristic's path and along the always-widen path, so the states
the model meets after its own widenings are represented too.  The model is
trained on the GPU with the engine's own forward/backward kernels and Adam;
the FP64 certificate still decides termination, so a poor model costs time,
never correctness.

Problems are drawn like the benchmark configurations (CSR like the 200000 x
20000 and batched ones, dense G D up to 40000 rows on the host, and dense
G D of 2e5..1e6 rows stored as bf16 and solved by the mixed-precision engine
like the 1e6 x 2000 one); see training_problem.

CLI (GPU):  python -m amfls_solver_amd.width_model train --first 0 --problems 1000 --collect-only --data part0.npz
            python -m amfls_solver_amd.width_model train --from-data part0.npz part1.npz --wide --c2
            python -m amfls_solver_amd.width_model evaluate --wide --c2
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import time
from pathlib import Path

import numpy as np
import torch

from . import _lib
from .binding import DeviceMatrix, default_context, run_amfls

N_IN, N_HID, N_PARAMS = _lib.WIDTH_INPUTS, 64, _lib.WIDTH_PARAMS
DEFAULT_PATH = Path(__file__).resolve().parent / "models" / "width_mlp.npz"


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights and zero biases in the engine's parameter layout."""
    r = np.random.default_rng(seed)
    W1 = r.standard_normal((N_HID, N_IN)) * math.sqrt(2.0 / N_IN)
    W2 = r.standard_normal((N_HID, N_HID)) * math.sqrt(2.0 / N_HID)
    W3 = r.standard_normal(N_HID) * math.sqrt(1.0 / N_HID)
    return np.concatenate([W1.ravel(), np.zeros(N_HID), W2.ravel(), np.zeros(N_HID), W3, [0.0]]).astype(np.float32)


class WidthModel:
    """Device-resident width MLP (C ABI amfls_width_model_*)."""

    def __init__(self, params: np.ndarray, norm: np.ndarray, context=None):
        self.context = context or default_context()
        self._lib = self.context._lib
        p = np.ascontiguousarray(params, dtype=np.float32)
        nm = np.ascontiguousarray(norm, dtype=np.float32)
        if p.size != N_PARAMS or nm.size != 2 * N_IN:
            raise ValueError("width model needs 5057 parameters and 24 normalization values")
        h = ctypes.c_void_p()
        _lib.check(self._lib.amfls_width_model_create(self.context.handle, p.ctypes.data_as(ctypes.c_void_p),
                                                      nm.ctypes.data_as(ctypes.c_void_p), ctypes.byref(h)))
        self.handle = h
        self.norm = nm.copy()
        self.domain = None

    def set_domain(self, lo: np.ndarray, hi: np.ndarray) -> None:
        """Per-input box of the training data; decisions outside it go to the heuristic."""
        lo = np.ascontiguousarray(lo, dtype=np.float32)
        hi = np.ascontiguousarray(hi, dtype=np.float32)
        if lo.size != N_IN or hi.size != N_IN:
            raise ValueError("width model domain needs 12 lower and 12 upper bounds")
        _lib.check(self._lib.amfls_width_model_set_domain(self.handle, lo.ctypes.data_as(ctypes.c_void_p),
                                                          hi.ctypes.data_as(ctypes.c_void_p)))
        self.domain = (lo.copy(), hi.copy())

    def decisions(self) -> tuple[int, int]:
        """(decisions taken by the model, decisions handed to the heuristic) so far."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(self._lib.amfls_width_model_decisions(self.handle, ctypes.byref(a), ctypes.byref(b)))
        return int(a.value), int(b.value)

    def params(self) -> np.ndarray:
        out = np.empty(N_PARAMS, dtype=np.float32)
        _lib.check(self._lib.amfls_width_model_get_params(self.handle, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def logits(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, N_IN)
        out = np.empty(X.shape[0], dtype=np.float32)
        _lib.check(self._lib.amfls_width_model_logits(self.handle, X.ctypes.data_as(ctypes.c_void_p), X.shape[0],
                                                      out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def gradient(self, X: torch.Tensor, y: torch.Tensor):
        """(mean BCE, gradient) for device tensors X (n x 12) and y (n)."""
        Xd = X.to(device="cuda", dtype=torch.float32).contiguous()
        yd = y.to(device="cuda", dtype=torch.float32).contiguous()
        g = np.empty(N_PARAMS, dtype=np.float32)
        loss = ctypes.c_float()
        torch.cuda.synchronize()
        _lib.check(self._lib.amfls_width_model_gradient(self.handle, Xd.data_ptr(), yd.data_ptr(), Xd.shape[0],
                                                        g.ctypes.data_as(ctypes.c_void_p), ctypes.byref(loss)))
        return float(loss.value), g

    def train_step(self, X: torch.Tensor, y: torch.Tensor, lr: float) -> float:
        loss = ctypes.c_float()
        _lib.check(self._lib.amfls_width_model_train_step(self.handle, X.data_ptr(), y.data_ptr(), X.shape[0],
                                                          float(lr), ctypes.byref(loss)))
        return float(loss.value)

    def attach(self) -> None:
        """Use this model for width_policy="learned" solves on its context."""
        _lib.check(self._lib.amfls_context_set_width_model(self.context.handle, self.handle))

    def save(self, path: Path, **meta) -> None:
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        extra = {} if self.domain is None else {"domain_lo": self.domain[0], "domain_hi": self.domain[1]}
        np.savez(path, params=self.params(), norm=self.norm, meta=np.array(json.dumps(meta)), **extra)

    @classmethod
    def load(cls, path: Path = DEFAULT_PATH, context=None) -> "WidthModel":
        with np.load(Path(path), allow_pickle=False) as z:
            m = cls(z["params"], z["norm"], context)
            m.meta = json.loads(str(z["meta"])) if "meta" in z else {}
            if "domain_lo" in z:
                m.set_domain(z["domain_lo"], z["domain_hi"])
        return m

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            self._lib.amfls_context_set_width_model(self.context.handle, None)
            self._lib.amfls_width_model_destroy(h)
            self.handle = None


# ----------------------------------------------------------------- problems
def training_problem(index: int, seed: int = 11):
    """Problem `index` of the training/evaluation distribution, drawn like the
    benchmark configurations: CSR as the 200000 x 20000 and batched
    configurations (m log-uniform in [5000, 200000], n log-uniform in
    [500, min(20000, m / 2)], distinct uniform columns, N(0,1) values, either 12
    or U[4, 32] nonzeros per row, rows optionally scaled by log-uniform
    factors in [1e-3, 1]) or dense G D as the tall dense configuration
    (m log-uniform in [1000, 40000], n in [100, min(1000, m)], D log-spaced
    over 1-6 decades, b = A x* + 1e-8-relative noise); lambda = 0 or one of
    logspace(-6, -1, 8); one problem in fifty is a tall dense G D of 2e5..1e6
    rows and 500..2000 columns built on the GPU in bf16 (2..6 decades,
    lambda 0), solved by the mixed-precision engine with a 1e-9 forward-error
    target.  Returns (matrix tensor, b, lambda, description)."""
    from . import problems as P
    r = np.random.default_rng([seed, index])
    if index % 50 == 7:  # tall dense bf16 like the 1e6 x 2000 configuration, mixed precision
        m = int(round(10.0 ** r.uniform(np.log10(200_000), 6.0)))
        n = int(r.integers(500, 2001))
        dec = float(r.uniform(2.0, 6.0))
        a, b, _ = P.config2_device(m, n, decades=dec, seed=int(r.integers(0, 2**31)))
        return a, b, 0.0, f"dense-bf16 {m}x{n} dec {dec:.1f} lam 0 (mixed)"
    lam = 0.0 if r.random() < 0.4 else float(10.0 ** (-6.0 + 5.0 * int(r.integers(0, 8)) / 7.0))
    if r.random() < 0.65:
        m = int(round(10.0 ** r.uniform(np.log10(5000), np.log10(200_000))))
        n = int(round(10.0 ** r.uniform(np.log10(500), np.log10(min(20_000, m // 2)))))
        per = np.full(m, 12) if r.random() < 0.3 else r.integers(4, 33, size=m)
        per = np.minimum(per, n)
        scaled = r.random() < 0.5
        p = P.csr_rows(m, n, per, seed=int(r.integers(0, 2**31)), scale_range=(1e-3, 1.0) if scaled else (1.0, 1.0),
                       regularization=lam)
        A = torch.sparse_csr_tensor(torch.from_numpy(p.row_ptr), torch.from_numpy(p.col_idx.astype(np.int64)),
                                    torch.from_numpy(p.values), size=(m, n))
        return A, torch.from_numpy(p.b), lam, f"csr {m}x{n} nnz {p.nnz} lam {lam:.1e}"
    m = int(round(10.0 ** r.uniform(3.0, np.log10(40_000))))
    n = int(r.integers(100, min(1000, m) + 1))
    dec = float(r.uniform(1.0, 6.0))
    a, b, _ = P.dense_gd(m, n, dec, seed=int(r.integers(0, 2**31)), noise=1e-8, relative_noise=True)
    return torch.from_numpy(np.ascontiguousarray(a)), torch.from_numpy(b), lam, f"dense {m}x{n} dec {dec:.1f} lam {lam:.1e}"


def problem_kind(index: int, seed: int = 11) -> str:
    """"dense-bf16", "csr" or "dense": the kind of training_problem(index,
    seed), found from the same random draws without building the matrix."""
    if index % 50 == 7:
        return "dense-bf16"
    r = np.random.default_rng([seed, index])
    if r.random() >= 0.4:
        r.integers(0, 8)
    return "csr" if r.random() < 0.65 else "dense"


def wide_problem(rows: int = 50_000, cols: int = 500_000, per_row: int = 8, seed: int = 5):
    """The wide minimum-norm evaluation case: CSR rows x cols with `per_row`
    distinct uniform columns per row, N(0,1) values, b ~ N(0,1), lambda = 0."""
    from . import problems as P
    p = P.csr_rows(rows, cols, per_row, seed=seed, scale_range=(1.0, 1.0), regularization=0.0)
    A = torch.sparse_csr_tensor(torch.from_numpy(p.row_ptr), torch.from_numpy(p.col_idx.astype(np.int64)),
                                torch.from_numpy(p.values), size=(rows, cols))
    return A, torch.from_numpy(p.b), 0.0, f"csr {rows}x{cols} nnz {p.nnz} lam 0 (wide)"


def _matrix(A, context=None):
    if A.is_cuda and A.dtype == torch.bfloat16:
        return DeviceMatrix(A, value_dtype=torch.bfloat16, context=context)
    return DeviceMatrix(A, value_dtype=torch.float32 if A.layout == torch.sparse_csr else torch.float64,
                        context=context)


def _mode(desc: str) -> dict:
    """Engine options of a problem: the mixed-precision engine with a
    forward-error target for the dense bf16 ones, FP64 otherwise."""
    return {"precision": "mixed", "forward_tolerance": 1e-9} if "(mixed)" in desc else {"precision": "fp64"}


def _solve(M, b, lam, tolerance, mode, policy="heuristic", script=None):
    """(wall seconds to certification or inf, status, auxiliary width,
    decision records rows x 13 or None) of one solve."""
    from .binding import _decode, _result_tuple
    ctx = M.context
    o = _lib.options(regularization=float(lam), tolerance=float(tolerance), width_policy=_lib.WIDTH_POLICY[policy],
                     precision=_lib.PRECISION[mode["precision"]],
                     forward_tolerance=float(mode.get("forward_tolerance", 0.0)))
    arr = None
    if script is not None and len(script) > 0:
        arr = np.ascontiguousarray(script, dtype=np.int8)
        o.width_script = arr.ctypes.data
        o.width_script_length = arr.size
    res = ctypes.c_void_p()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _lib.check(ctx._lib.amfls_solve(ctx.handle, M.handle, b.data_ptr(), 1, ctypes.byref(o), ctypes.byref(res)))
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    rec = None
    try:
        if policy.startswith("record"):
            rows = ctypes.c_int64()
            _lib.check(ctx._lib.amfls_result_width_records(res, None, ctypes.byref(rows)))
            rec = np.empty((rows.value, N_IN + 1), dtype=np.float32)
            _lib.check(ctx._lib.amfls_result_width_records(res, rec.ctypes.data_as(ctypes.c_void_p),
                                                           ctypes.byref(rows)))
    except Exception:
        ctx._lib.amfls_result_destroy(res)
        raise
    r = _decode(*_result_tuple(ctx, res))  # releases the result
    return (dt if r.status == 0 else math.inf), int(r.status), int(r.auxiliary_width), rec


def counterfactual_labels(M, b, lam, tolerance, mode, *, heuristic_decisions=3, widen_decisions=1,
                          first_decision=0, margin=0.03):
    """Decision records with counterfactual labels for one problem.

    Records the decisions of the heuristic's run and of the always-widen
    run; for decisions first_decision .. first_decision + count - 1 of each,
    replays the solve with the recorded choices before d, d forced to widen
    or to hold, and the heuristic after d.  Returns (X rows x 12, y,
    summary)."""
    _solve(M, b, lam, tolerance, mode)  # warm
    X, y, out = [], [], {"heuristic_path": [], "widen_path": []}
    for path, policy, limit in (("heuristic_path", "record", heuristic_decisions),
                                ("widen_path", "record_widen", widen_decisions)):
        t_path, _, _, rec = _solve(M, b, lam, tolerance, mode, policy)
        if rec is None or rec.shape[0] == 0:
            continue
        chosen = rec[:, N_IN].astype(np.int8) if policy == "record" else np.ones(rec.shape[0], np.int8)
        for d in range(first_decision, min(first_decision + limit, rec.shape[0])):
            times = []
            for c in (1, 0):
                script = np.concatenate([chosen[:d], [c]]).astype(np.int8)
                t = _solve(M, b, lam, tolerance, mode, "heuristic", script)[0]
                if t < 0.05 or not math.isfinite(t):
                    t = min(t, _solve(M, b, lam, tolerance, mode, "heuristic", script)[0])
                times.append(t)
            t_widen, t_hold = times
            entry = {"decision": d, "widen_seconds": t_widen, "hold_seconds": t_hold, "heuristic": int(rec[d, N_IN])}
            if math.isfinite(t_widen) or math.isfinite(t_hold):
                if abs(t_widen - t_hold) > margin * min(t_widen, t_hold):
                    X.append(rec[d, :N_IN])
                    y.append(1.0 if t_widen < t_hold else 0.0)
                    entry["label"] = int(y[-1])
            out[path].append(entry)
    return np.array(X, dtype=np.float32).reshape(-1, N_IN), np.array(y, dtype=np.float32), out


def collect(indices, *, seed: int = 11, tolerance: float = 1e-8, budget_seconds: float = math.inf, log=print,
            kinds=None, first_decision: int = 0, heuristic_decisions: int = 3, widen_decisions: int = 1):
    """Counterfactually labelled decision records for the given problem
    indices (stops starting new problems once `budget_seconds` have passed);
    `kinds` restricts them to problem kinds (problem_kind), and the decision
    window is first_decision .. + heuristic_decisions on the heuristic's
    path (+ widen_decisions on the always-widen path)."""
    X, y, info = [], [], []
    t_start = time.perf_counter()
    for i in indices:
        if kinds is not None and problem_kind(i, seed) not in kinds:
            continue
        if time.perf_counter() - t_start > budget_seconds:
            log(f"  collection budget of {budget_seconds:.0f}s reached after {len(info)} problems")
            break
        A, b, lam, desc = training_problem(i, seed)
        M = _matrix(A)
        bd = b.cuda()
        Xi, yi, summary = counterfactual_labels(M, bd, lam, tolerance, _mode(desc), first_decision=first_decision,
                                                heuristic_decisions=heuristic_decisions,
                                                widen_decisions=widen_decisions)
        X.extend(Xi)
        y.extend(yi)
        info.append({"index": int(i), "problem": desc, "labelled": int(len(yi)), **summary})
        log(f"  [{i}] {desc}: {len(yi)} labelled decisions ({int(np.sum(yi))} widen)")
        del M, A, bd
    return np.array(X, dtype=np.float32).reshape(-1, N_IN), np.array(y, dtype=np.float32), info


def record(M, b, lam, tolerance, widen: bool = False):
    """Model inputs and heuristic choices (rows x 13) of every widening
    decision met while following the heuristic, or while always widening."""
    raw, _, _ = _record_raw(M, b, lam, tolerance, "record_widen" if widen else "record")
    return raw


def _record_raw(M, b, lam, tolerance, policy):
    """(decision records rows x 13, status, auxiliary width) of one solve."""
    from .binding import _decode, _result_tuple
    ctx = M.context
    o = _lib.options(regularization=float(lam), tolerance=float(tolerance), width_policy=_lib.WIDTH_POLICY[policy])
    res = ctypes.c_void_p()
    torch.cuda.synchronize()
    _lib.check(ctx._lib.amfls_solve(ctx.handle, M.handle, b.data_ptr(), 1, ctypes.byref(o), ctypes.byref(res)))
    try:
        rows = ctypes.c_int64()
        _lib.check(ctx._lib.amfls_result_width_records(res, None, ctypes.byref(rows)))
        out = np.empty((rows.value, N_IN + 1), dtype=np.float32)
        _lib.check(ctx._lib.amfls_result_width_records(res, out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(rows)))
    except Exception:
        ctx._lib.amfls_result_destroy(res)
        raise
    r = _decode(*_result_tuple(ctx, res))  # releases the result
    return out, int(r.status), int(r.auxiliary_width)


def feature_domain(X: np.ndarray, margin: float = 0.05):
    """Box of the training inputs, widened by `margin` of each range, for the
    solver-state inputs (width, certificate decreases, level times); the
    operator-size inputs (log rows, columns, nonzeros) are left open, so the
    model also decides for operators larger or wider than it was trained
    on, from the state they put the solver in."""
    lo, hi = X.min(axis=0), X.max(axis=0)
    pad = margin * np.maximum(hi - lo, 1e-3)
    lo, hi = (lo - pad).astype(np.float32), (hi + pad).astype(np.float32)
    lo[9:] = -np.inf
    hi[9:] = np.inf
    return lo, hi


def train(X: np.ndarray, y: np.ndarray, *, epochs: int = 400, lr: float = 3e-3, batch: int = 256, seed: int = 0,
          log=print) -> WidthModel:
    """Adam on the GPU with the engine's kernels; features normalized by the
    training set's mean and standard deviation; the model's domain is the
    box of the training inputs."""
    mean = X.mean(axis=0)
    std = X.std(axis=0)
    norm = np.concatenate([mean, 1.0 / np.maximum(std, 1e-3)]).astype(np.float32)
    model = WidthModel(init_params(seed), norm)
    model.set_domain(*feature_domain(X))
    Xd = torch.from_numpy(X).cuda()
    yd = torch.from_numpy(y).cuda()
    gen = torch.Generator(device="cpu").manual_seed(seed)
    n = X.shape[0]
    for ep in range(epochs):
        perm = torch.randperm(n, generator=gen).cuda()
        tot = 0.0
        for s0 in range(0, n, batch):
            idx = perm[s0:s0 + batch]
            xb, yb = Xd[idx].contiguous(), yd[idx].contiguous()
            tot += model.train_step(xb, yb, lr) * idx.numel()
        if ep % 50 == 0 or ep == epochs - 1:
            log(f"  epoch {ep}: mean loss {tot / n:.4f}")
    return model


def evaluate(model: WidthModel, problems, *, seed: int = 11, tolerance: float = 1e-8, repeats: int = 2, log=print):
    """Time-to-tolerance of the heuristic and the learned policy per problem
    (median of `repeats` runs each, after a warm run).  `problems` holds
    distribution indices or (matrix, b, lambda, description) tuples; dense
    bf16 problems run on the mixed-precision engine."""
    model.attach()
    rows = []
    for item in problems:
        A, b, lam, desc = training_problem(item, seed) if isinstance(item, (int, np.integer)) else item
        M = _matrix(A, model.context)
        bd = b.cuda()
        mode = _mode(desc)
        out = {}
        for pol in ("heuristic", "learned"):
            _solve(M, bd, lam, tolerance, mode, pol)
            d0 = model.decisions()
            runs = [_solve(M, bd, lam, tolerance, mode, pol) for _ in range(repeats)]
            d1 = model.decisions()
            out[pol] = (float(np.median([r[0] for r in runs])), runs[-1][1], runs[-1][2],
                        (d1[0] - d0[0]) // repeats, (d1[1] - d0[1]) // repeats)
        rows.append({"index": int(item) if isinstance(item, (int, np.integer)) else -1, "problem": desc,
                     "heuristic": out["heuristic"][:3], "learned": out["learned"][:3],
                     "model_decisions": out["learned"][3], "heuristic_fallbacks": out["learned"][4]})
        log(f"  [{rows[-1]['index']}] {desc}: heuristic {out['heuristic'][0]:.4f}s (aux {out['heuristic'][2]}), "
            f"learned {out['learned'][0]:.4f}s (aux {out['learned'][2]}; {out['learned'][3]} model decisions, "
            f"{out['learned'][4]} outside the domain)")
        del M
    ratios = [r["heuristic"][0] / r["learned"][0] for r in rows
              if math.isfinite(r["heuristic"][0]) and math.isfinite(r["learned"][0])]
    gm = float(np.exp(np.mean(np.log(ratios)))) if ratios else float("nan")
    return gm, rows


def c2_problem():
    """The 1e6 x 2000 bf16 dense configuration (mixed precision, forward-error target)."""
    from . import problems as P
    a, b, _ = P.config2_device()
    return a, b, 0.0, "dense-bf16 1000000x2000 dec 6.0 lam 0 (mixed)"


def gain_cases(rows, factor: float = 1.2, limit: int = 4):
    """Held-out problems on which the learned policy widened beyond the
    heuristic and certified at least `factor` times sooner."""
    out = []
    for r in rows:
        h, l = r["heuristic"], r["learned"]
        if h[1] == 0 and l[1] == 0 and l[2] > h[2] and h[0] > factor * l[0]:
            out.append(r["index"])
    return out[:limit]


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m amfls_solver_amd.width_model")
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--problems", type=int, default=2000)
    t.add_argument("--budget-seconds", type=float, default=math.inf,
                   help="stop starting new training problems after this many seconds")
    t.add_argument("--heldout", type=int, default=40)
    t.add_argument("--epochs", type=int, default=400)
    t.add_argument("--out", type=Path, default=DEFAULT_PATH)
    t.add_argument("--data", type=Path, default=None, help="also save the decision records (npz) here")
    t.add_argument("--wide", action="store_true", help="also evaluate the wide 50000 x 500000 minimum-norm CSR case")
    t.add_argument("--c2", action="store_true", help="also evaluate the 1e6 x 2000 bf16 dense case (mixed)")
    t.add_argument("--first", type=int, default=0, help="first training problem index")
    t.add_argument("--collect-only", action="store_true", help="only collect decision records (to --data)")
    t.add_argument("--kinds", default=None, help="comma-separated problem kinds to collect (csr, dense, dense-bf16)")
    t.add_argument("--first-decision", type=int, default=0, help="first labelled decision of each path")
    t.add_argument("--decisions", type=int, default=3, help="labelled decisions on the heuristic's path")
    t.add_argument("--widen-decisions", type=int, default=1, help="labelled decisions on the always-widen path")
    t.add_argument("--from-data", type=Path, nargs="*", default=None,
                   help="train on these collected record files instead of collecting")
    e = sub.add_parser("evaluate")
    e.add_argument("--model", type=Path, default=DEFAULT_PATH)
    e.add_argument("--heldout", type=int, default=40)
    e.add_argument("--update", type=Path, default=None, help="write the model with the held-out result here")
    e.add_argument("--wide", action="store_true", help="also the wide 50000 x 500000 minimum-norm CSR case")
    e.add_argument("--c2", action="store_true", help="also the 1e6 x 2000 bf16 dense case (mixed)")
    a = ap.parse_args(argv)
    if a.cmd == "train":
        if a.from_data:
            Xs, ys, info = [], [], []
            for f in a.from_data:
                with np.load(f, allow_pickle=False) as z:
                    Xs.append(z["X"])
                    ys.append(z["y"])
                    info.extend(json.loads(str(z["info"])))
            X, y = np.concatenate(Xs), np.concatenate(ys)
        else:
            print(f"collecting problems {a.first}..{a.first + a.problems - 1}", flush=True)
            X, y, info = collect(range(a.first, a.first + a.problems), budget_seconds=a.budget_seconds,
                                 log=lambda s: print(s, flush=True),
                                 kinds=None if a.kinds is None else set(a.kinds.split(",")),
                                 first_decision=a.first_decision, heuristic_decisions=a.decisions,
                                 widen_decisions=a.widen_decisions)
        print(f"{X.shape[0]} labelled decisions from {len(info)} problems, {y.mean():.3f} labelled widen", flush=True)
        if a.data:
            a.data.parent.mkdir(parents=True, exist_ok=True)
            np.savez(a.data, X=X, y=y, info=np.array(json.dumps(info)))
        if a.collect_only:
            return 0
        model = train(X, y, epochs=a.epochs, log=lambda s: print(s, flush=True))
        z = model.logits(X)
        acc = float(np.mean((z > 0) == (y > 0.5)))
        print(f"training accuracy {acc:.3f}", flush=True)
        gm, rows = evaluate(model, range(100000, 100000 + a.heldout), log=lambda s: print(s, flush=True))
        print(f"held-out geometric-mean speedup of learned over heuristic: {gm:.3f}", flush=True)
        meta = dict(problems=len(info), decisions=int(X.shape[0]), widen_fraction=float(y.mean()),
                    labels="per-decision counterfactual replays", train_accuracy=acc, heldout=a.heldout,
                    heldout_speedup=gm, epochs=a.epochs, gain_cases=gain_cases(rows))
        for flag, name, make in ((a.wide, "wide_case", wide_problem), (a.c2, "c2_case", c2_problem)):
            if flag:
                _, wr = evaluate(model, [make()], repeats=2, log=lambda s: print(s, flush=True))
                meta[name] = {k: wr[0][k] for k in ("problem", "heuristic", "learned", "model_decisions",
                                                     "heuristic_fallbacks")}
        model.save(a.out, **meta)
        print(f"saved {a.out}", flush=True)
    else:
        model = WidthModel.load(a.model)
        gm, rows = evaluate(model, range(100000, 100000 + a.heldout), log=lambda s: print(s, flush=True))
        print(f"held-out geometric-mean speedup of learned over heuristic: {gm:.3f}", flush=True)
        meta = dict(model.meta)
        meta.update(heldout=a.heldout, heldout_speedup=gm, gain_cases=gain_cases(rows))
        for flag, name, make in ((a.wide, "wide_case", wide_problem), (a.c2, "c2_case", c2_problem)):
            if flag:
                _, wr = evaluate(model, [make()], repeats=2, log=lambda s: print(s, flush=True))
                meta[name] = {k: wr[0][k] for k in ("problem", "heuristic", "learned", "model_decisions",
                                                     "heuristic_fallbacks")}
        if a.update:
            model.save(a.update, **meta)
            print(f"wrote {a.update}", flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
