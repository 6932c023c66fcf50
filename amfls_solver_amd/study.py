"""Study-runner parity for the dense comparative suites.

Draws the reference's synthetic dense problems
(experiments/runners/run_comparative_study.py:193-384), measures a solution
with the same external error definitions (:577-655) and writes the CSV
schema (:38-107), so that the reference's
study configurations run on the HIP engine and produce rows in the same
layout.  The generator draws from a torch CPU generator in the reference's
order (left basis, right basis, coefficients, then signs or noise), so an
instance seed gives the same problem as in the reference's result tables;
the reference objective of each table row pins that.
"""
from __future__ import annotations

import math
import time

import numpy as np
import torch

# CSV columns of the reference's study tables (run_comparative_study.py:38-107)
FIELDS = [
    "suite", "case", "instance_seed", "algorithm_seed", "method", "m", "n", "rank", "coherence", "spectrum", "rhs",
    "regularization", "regularization_ratio", "tolerance", "success", "external_success", "precision_status_case",
    "precision_status_correct", "status", "stop_reason", "objective", "reference_objective", "relative_objective_gap",
    "residual_norm", "normal_residual_norm", "compatible_backward_error_external",
    "least_squares_backward_error_external", "backward_error_external", "relative_energy_error_external",
    "reported_accuracy_upper", "relative_solution_error", "relative_null_component", "a_columns", "at_columns",
    "total_vector_work", "a_block_calls", "at_block_calls", "total_block_calls", "search_vector_work",
    "validation_vector_work", "search_block_calls", "validation_block_calls", "sketch_a_columns",
    "sketch_at_columns", "sketch_a_block_calls", "sketch_at_block_calls", "iterative_a_columns",
    "iterative_at_columns", "iterative_a_block_calls", "iterative_at_block_calls", "gaussian_random_columns",
    "gaussian_random_values", "gaussian_random_block_requests", "iterations", "depth", "auxiliary_width",
    "basis_rank", "kernel_seconds", "a_seconds", "at_seconds", "orthogonalization_seconds",
    "projected_solve_seconds", "validation_seconds", "other_seconds", "wall_seconds", "peak_resident_memory_mib",
    "additional_resident_memory_mib", "error",
]


def _orthonormal(rows: int, cols: int, gen: torch.Generator, coherence: str) -> torch.Tensor:
    if coherence == "coherent":  # the leading coordinate vectors
        out = torch.zeros((rows, cols), dtype=torch.float64)
        idx = torch.arange(cols)
        out[idx, idx] = 1.0
        return out
    if coherence != "haar":
        raise ValueError(f"coherence {coherence!r} is neither 'haar' nor 'coherent'")
    g = torch.randn((rows, cols), generator=gen, dtype=torch.float64)
    q, r = torch.linalg.qr(g, mode="reduced")
    d = torch.sign(torch.diagonal(r))
    d[d == 0.0] = 1.0
    return q * d  # Haar-distributed: Q with the signs of diag(R) taken out


def _spectrum(case: dict, rank: int) -> torch.Tensor:
    kind = str(case["spectrum"])
    i = torch.arange(rank, dtype=torch.float64)
    one = torch.ones(rank, dtype=torch.float64)
    if kind == "flat":
        return one
    if kind in ("outliers", "separated_outliers"):
        cnt = min(int(case.get("outlier_count", 1)), rank)
        big = float(case.get("outlier_ratio", 1.0e4))
        v = torch.linspace(1.0, float(case.get("tail_minimum", 0.2)), rank, dtype=torch.float64)
        if kind == "outliers" or cnt == 1:
            v[:cnt] = big
        else:
            v[:cnt] = torch.logspace(math.log10(big), math.log10(float(case.get("smallest_outlier_ratio", 10.0))),
                                     cnt, dtype=torch.float64)
        return v
    if kind == "clusters":
        v = torch.full((rank,), 0.1, dtype=torch.float64)
        a = min(rank, max(1, rank // 16))
        b = min(rank, max(a + 1, rank // 2))
        v[:a] = 100.0
        v[a:b] = 1.0
        return v
    if kind == "polynomial":
        return (i + 1.0).pow(-float(case.get("alpha", 1.0)))
    if kind in ("exponential", "numerical"):
        end = float(case.get("endpoint", 1.0e-8)) if kind == "exponential" else float(case["endpoint"])
        return one if rank == 1 else torch.logspace(0.0, math.log10(end), rank, dtype=torch.float64)
    if kind == "smooth":
        return one if rank == 1 else torch.exp(-4.0 * (i / float(rank - 1)).pow(1.5))
    if kind == "precision_two_mode":
        v = one.clone()
        v[-1] = float(case["two_mode_small_singular_value"])
        return v
    raise ValueError(f"no singular-value profile is named {kind!r}")


def make_problem(case: dict, seed: int) -> dict:
    """A = U diag(sigma) V^T (or a coordinate embedding for coherent cases),
    b built from coefficients in the left basis (plus an orthogonal noise
    component for inconsistent cases), scaled to unit norm (a power of two
    for exact coordinate inconsistent cases), and the exact solution of the
    lambda = ratio * sigma_1^2 problem.  Returns matrix, b, reference,
    row_basis, operator_norm, regularization, reference_objective."""
    m, n, rank = int(case["m"]), int(case["n"]), int(case["rank"])
    if rank <= 0 or rank > min(m, n):
        raise ValueError(f"a {m} x {n} case cannot have rank {rank}")
    gen = torch.Generator(device="cpu").manual_seed(seed)
    coherence = str(case.get("coherence", "haar"))
    U = _orthonormal(m, rank, gen, coherence)
    V = _orthonormal(n, rank, gen, coherence)
    sigma = _spectrum(case, rank)
    if coherence == "coherent":
        A = torch.zeros((m, n), dtype=torch.float64)
        idx = torch.arange(rank)
        A[idx, idx] = sigma
    else:
        A = ((U * sigma) @ V.T).contiguous()
    c = torch.randn(rank, generator=gen, dtype=torch.float64)
    rhs = str(case["rhs"])
    ratio = float(case.get("regularization_ratio", 0.0))
    exact_inconsistent = coherence == "coherent" and rhs == "inconsistent" and rank < min(m, n) and ratio == 0.0
    if exact_inconsistent:
        c = torch.sign(c)
        c[c == 0.0] = 1.0
    if rhs == "top_loaded":
        c[max(1, rank // 8):] = 0.0
    elif rhs == "tail_loaded":
        c[:max(0, rank - max(1, rank // 8))] = 0.0
    elif rhs == "outlier_orthogonal":
        c[:min(int(case.get("outlier_count", 1)), rank)] = 0.0
    elif rhs == "energy_balanced":
        lam0 = ratio * float(sigma[0]) ** 2
        c = (sigma.square() + lam0).rsqrt()
        sg = torch.sign(torch.randn(rank, generator=gen, dtype=torch.float64))
        sg[sg == 0.0] = 1.0
        c = c * sg
    elif rhs not in ("random", "consistent", "inconsistent"):
        raise ValueError(f"no right-hand-side recipe is named {rhs!r}")
    signal = U @ (sigma * c)
    b = signal.clone()
    if rhs == "inconsistent":
        if rank >= m:
            raise ValueError(f"rank {rank} leaves no room in R^{m} for a residual component")
        if exact_inconsistent:
            noise = torch.zeros(m, dtype=torch.float64)
            noise[rank:] = torch.randn(m - rank, generator=gen, dtype=torch.float64)
        else:
            noise = torch.randn(m, generator=gen, dtype=torch.float64)
            noise -= U @ (U.T @ noise)
        nn_ = float(torch.linalg.vector_norm(noise))
        if nn_ == 0.0:
            raise RuntimeError("the residual component drawn for this seed vanished")
        b += float(case.get("noise_ratio", 0.05)) * float(torch.linalg.vector_norm(signal)) * noise / nn_
    lam = ratio * float(sigma[0]) ** 2
    x = V @ ((sigma.square() / (sigma.square() + lam)) * c)
    bn = float(torch.linalg.vector_norm(b))
    if bn == 0.0:
        raise ValueError("this case and seed give b = 0, which has only the trivial solution")
    if exact_inconsistent:  # exact power-of-two scaling keeps the coordinates exact
        _, e = math.frexp(bn)
        scale = math.ldexp(1.0, -e)
        b *= scale
        x *= scale
    else:
        b /= bn
        x /= bn
    r = A @ x - b
    return {"matrix": A, "b": b.contiguous(), "reference": x.contiguous(), "row_basis": V,
            "operator_norm": float(sigma[0]), "regularization": lam,
            "reference_objective": float(torch.dot(r, r) + lam * torch.dot(x, x))}


def _quotient(num: float, den: float) -> float:
    """num / den for error ratios: 0 for a zero numerator, +inf for a zero
    denominator."""
    if num == 0.0:
        return 0.0
    return math.inf if den == 0.0 else num / den


# A row passes externally when its primary error is within the tolerance up
# to this allowance for the rounding of the check itself (the reference's
# study uses the same allowance, so success columns compare row for row).
def _allowance(tolerance: float) -> float:
    return 256.0 * np.finfo(np.float64).eps + 1.0e-7 * tolerance


def diagnostics(problem: dict, solution: torch.Tensor, tolerance: float) -> dict:
    """Externally computed accuracy of `solution` on `problem`.

    Everything is measured on the augmented system K = [A; sqrt(lam) I],
    whose residual is [A x - b; sqrt(lam) x] and whose normal residual is
    g = A^T (A x - b) + lam x:

    * backward errors: compatible ||r_K|| / (||K|| ||x|| + ||b||), least
      squares ||g|| / (||K|| ||r_K||), and the smaller of the two, with
      ||K|| = sqrt(sigma_1^2 + lam);
    * the relative error in the energy norm ||K e|| / ||K x_ref||, and the
      relative 2-norm error of x;
    * for rank-deficient unregularized cases, the share of x outside the
      row space (minimum-norm check).

    external_success uses the energy error when lam > 0 and the backward
    error otherwise, plus the minimum-norm check where it applies."""
    A = problem["matrix"].numpy()
    b = problem["b"].numpy()
    x_ref = problem["reference"].numpy()
    V = problem["row_basis"].numpy()
    lam = float(problem["regularization"])
    x = solution.to(dtype=torch.float64, device="cpu").numpy()
    root = math.sqrt(lam)
    norm = np.linalg.norm

    def augmented(v):  # K v
        return np.concatenate([A @ v, root * v])

    r = A @ x - b
    r_aug = np.concatenate([r, root * x])
    g = A.T @ r + lam * x
    k_norm = math.hypot(float(problem["operator_norm"]), root)
    x_norm, b_norm, r_aug_norm = norm(x), norm(b), norm(r_aug)
    compatible = _quotient(r_aug_norm, k_norm * x_norm + b_norm)
    least_squares = _quotient(norm(g), k_norm * r_aug_norm)
    err = x - x_ref
    energy = _quotient(norm(augmented(err)), norm(augmented(x_ref)))
    backward = min(compatible, least_squares)
    tol = tolerance + _allowance(tolerance)
    full_row_space = lam > 0.0 or V.shape[1] >= A.shape[1]
    outside = 0.0 if full_row_space else _quotient(norm(x - V @ (V.T @ x)), norm(x_ref))
    objective = float(r_aug @ r_aug)
    best = float(problem["reference_objective"])
    primary = energy if lam > 0.0 else backward
    return {
        "external_success": int(primary <= tol and outside <= tol),
        "objective": objective,
        "relative_objective_gap": (objective - best) / max(best, np.finfo(np.float64).tiny),
        "residual_norm": float(norm(r)),
        "normal_residual_norm": float(norm(g)),
        "compatible_backward_error_external": compatible,
        "least_squares_backward_error_external": least_squares,
        "backward_error_external": backward,
        "relative_energy_error_external": energy,
        "relative_solution_error": _quotient(norm(err), norm(x_ref)),
        "relative_null_component": outside,
    }


def run_case(suite: str, options: dict, case: dict, instance_seed: int, *, precision: str = "fp64",
             context=None) -> dict:
    """One AMFLS row of a study table on the HIP engine: the problem of
    `instance_seed`, the suite's limits, algorithm seed = instance seed +
    the suite's offset; every FIELDS key is present ('' where the engine has
    no such counter)."""
    from . import DeviceMatrix, run_amfls
    p = make_problem(case, instance_seed)
    seed = instance_seed + int(options["algorithm_seed_offset"])
    M = DeviceMatrix(p["matrix"], context=context)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = run_amfls(M, p["b"], regularization=p["regularization"], tolerance=float(options["tolerance"]),
                  failure_probability=float(options["failure_probability"]),
                  maximum_epochs=int(options["maximum_epochs"]), maximum_depth=int(options["maximum_depth"]),
                  maximum_auxiliary_width=int(options["maximum_auxiliary_width"]),
                  maximum_basis_size=int(options["maximum_basis_size"]), seed=seed, precision=precision,
                  context=context)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    row = {k: "" for k in FIELDS}
    row.update(suite=suite, case=case["name"], instance_seed=instance_seed, algorithm_seed=seed, method="amfls",
               m=case["m"], n=case["n"], rank=case["rank"], coherence=case.get("coherence", "haar"),
               spectrum=case["spectrum"], rhs=case["rhs"], regularization=p["regularization"],
               regularization_ratio=case.get("regularization_ratio", 0.0), tolerance=options["tolerance"],
               success=int(r.status == 0), status=r.status, stop_reason=r.stop_reason,
               reference_objective=p["reference_objective"], wall_seconds=wall)
    row.update(diagnostics(p, r.solution, float(options["tolerance"])))
    row["reported_accuracy_upper"] = (r.relative_energy_error_upper_bound if p["regularization"] > 0.0
                                      else r.backward_error_upper_bound)
    for k in ("a_columns", "at_columns", "a_block_calls", "at_block_calls", "gaussian_random_columns",
              "gaussian_random_values", "gaussian_random_block_requests", "iterations", "depth", "auxiliary_width",
              "basis_rank", "a_seconds", "at_seconds", "orthogonalization_seconds", "projected_solve_seconds",
              "validation_seconds"):
        if hasattr(r, k):
            row[k] = getattr(r, k)
    row["total_block_calls"] = r.a_block_calls + r.at_block_calls
    row["total_vector_work"] = r.a_columns + r.at_columns
    return row


def write_csv(rows, path) -> None:
    """Rows in the reference's column order (extra keys are refused)."""
    import csv
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for r in rows:
            w.writerow(r)


def main(argv=None) -> int:
    """python -m amfls_solver_amd.study CONFIG.json [--seeds K] [--out rows.csv]
    runs the AMFLS rows of a dense study configuration in the reference's
    format (suite, tolerance, limits, algorithm_seed_offset, instance_seeds,
    cases) on the HIP engine."""
    import argparse
    import json
    ap = argparse.ArgumentParser(prog="python -m amfls_solver_amd.study")
    ap.add_argument("config")
    ap.add_argument("--seeds", type=int, default=1, help="first K instance seeds of the configuration")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    with open(a.config) as f:
        cfg = json.load(f)
    rows = []
    for case in cfg["cases"]:
        for seed in cfg["instance_seeds"][:a.seeds]:
            row = run_case(cfg["suite"], cfg, case, int(seed))
            rows.append(row)
            print(f"{case['name']} seed {seed}: status {row['status']} external_success {row['external_success']} "
                  f"depth {row['depth']} width {row['auxiliary_width']} {row['wall_seconds']:.3f}s", flush=True)
    if a.out:
        write_csv(rows, a.out)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
