"""Study-runner parity for the dense comparative suites.

Restates the reference's synthetic dense problem generator
(experiments/runners/run_comparative_study.py:193-384), its external
diagnostics (:577-655) and its CSV schema (:38-107), so that the reference's
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
        raise ValueError(f"unknown coherence: {coherence}")
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
    raise ValueError(f"unknown spectrum: {kind}")


def make_problem(case: dict, seed: int) -> dict:
    """A = U diag(sigma) V^T (or a coordinate embedding for coherent cases),
    b built from coefficients in the left basis (plus an orthogonal noise
    component for inconsistent cases), scaled to unit norm (a power of two
    for exact coordinate inconsistent cases), and the exact solution of the
    lambda = ratio * sigma_1^2 problem.  Returns matrix, b, reference,
    row_basis, operator_norm, regularization, reference_objective."""
    m, n, rank = int(case["m"]), int(case["n"]), int(case["rank"])
    if rank <= 0 or rank > min(m, n):
        raise ValueError(f"invalid rank {rank} for shape ({m},{n})")
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
        raise ValueError(f"unknown right-hand side: {rhs}")
    signal = U @ (sigma * c)
    b = signal.clone()
    if rhs == "inconsistent":
        if rank >= m:
            raise ValueError("an inconsistent case requires rank < m")
        if exact_inconsistent:
            noise = torch.zeros(m, dtype=torch.float64)
            noise[rank:] = torch.randn(m - rank, generator=gen, dtype=torch.float64)
        else:
            noise = torch.randn(m, generator=gen, dtype=torch.float64)
            noise -= U @ (U.T @ noise)
        nn_ = float(torch.linalg.vector_norm(noise))
        if nn_ == 0.0:
            raise RuntimeError("failed to construct an inconsistent component")
        b += float(case.get("noise_ratio", 0.05)) * float(torch.linalg.vector_norm(signal)) * noise / nn_
    lam = ratio * float(sigma[0]) ** 2
    x = V @ ((sigma.square() / (sigma.square() + lam)) * c)
    bn = float(torch.linalg.vector_norm(b))
    if bn == 0.0:
        raise ValueError("right-hand side cannot be zero")
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


def _ratio(num: float, den: float) -> float:
    if num == 0.0:
        return 0.0
    return math.inf if den == 0.0 else num / den


def diagnostics(problem: dict, solution: torch.Tensor, tolerance: float) -> dict:
    """External accuracy of a solution against the problem's exact answer:
    backward errors (compatible, least squares, their minimum), relative
    energy and solution errors, the null-space component of minimum-norm
    problems, and external_success (the primary error, energy for lambda > 0
    else backward, within tolerance plus a small slack)."""
    A, b, x_ref, V = problem["matrix"], problem["b"], problem["reference"], problem["row_basis"]
    lam, anorm = float(problem["regularization"]), float(problem["operator_norm"])
    x = solution.to(dtype=torch.float64, device="cpu")
    r = A @ x - b
    g = A.T @ r + lam * x
    rn, gn, xn, bn = (float(torch.linalg.vector_norm(v)) for v in (r, g, x, b))
    aug_r = math.sqrt(rn * rn + lam * xn * xn)
    aug_a = math.sqrt(anorm * anorm + lam)
    compatible = _ratio(aug_r, aug_a * xn + bn)
    least_squares = _ratio(gn, aug_a * aug_r)
    backward = min(compatible, least_squares)
    e = x - x_ref
    en, refn = float(torch.linalg.vector_norm(e)), float(torch.linalg.vector_norm(x_ref))
    ea, refa = float(torch.linalg.vector_norm(A @ e)), float(torch.linalg.vector_norm(A @ x_ref))
    energy = _ratio(math.sqrt(ea * ea + lam * en * en), math.sqrt(refa * refa + lam * refn * refn))
    primary = energy if lam > 0.0 else backward
    eps = float(torch.finfo(torch.float64).eps)
    slack = 256.0 * eps + 1.0e-7 * tolerance
    obj = float(torch.dot(r, r) + lam * torch.dot(x, x))
    ref_obj = float(problem["reference_objective"])
    if lam > 0.0 or V.shape[1] >= A.shape[1]:
        null_rel, min_norm_ok = 0.0, True
    else:
        null_rel = _ratio(float(torch.linalg.vector_norm(x - V @ (V.T @ x))), refn)
        min_norm_ok = null_rel <= tolerance + slack
    return {
        "external_success": int(primary <= tolerance + slack and min_norm_ok),
        "objective": obj,
        "relative_objective_gap": (obj - ref_obj) / max(ref_obj, float(torch.finfo(torch.float64).tiny)),
        "residual_norm": rn,
        "normal_residual_norm": gn,
        "compatible_backward_error_external": compatible,
        "least_squares_backward_error_external": least_squares,
        "backward_error_external": backward,
        "relative_energy_error_external": energy,
        "relative_solution_error": _ratio(en, refn),
        "relative_null_component": null_rel,
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
