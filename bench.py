#!/usr/bin/env python
"""Time-to-tolerance benchmark of the AMFLS HIP engine on one MI355X.

Headline workload: the 200000 x 20000 CSR least-squares problem (exactly 12
distinct uniform columns per row, N(0,1) values, rows scaled by log-uniform
factors in [1e-3, 1], values stored as float32, b ~ N(0,1), lambda = 1e-3,
certified relative energy error 1e-8).  A step is one complete certified
solve with the matrix and b already resident in HBM.  Reported: seconds per
solve of the plain FP64 engine (headline), of mixed precision (FP32 search +
FP64 refinement/certificate) and of the learned width policy; the CSR SpMM
kernel's achieved algorithmic GB/s measured with HIP events on the engine
stream over the timed solves; an SpMM bandwidth sweep on a CSR matrix larger
than the 256 MB last-level cache; the 1,000,000 x 2000 dense BF16 problem
(G D, D log-spaced over six decades, lambda = 0) with its matrix-core product
rates; and the CPU restatement (oracle) timed on this host with one BLAS
thread.

python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def measured_traffic():
    """HBM bytes per launch of the headline SpMM kernels from the committed
    PMC summary (FETCH_SIZE and WRITE_SIZE passes scaled by a calibration
    copy), or None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "csr_spmm_c1_pmc_traffic.txt")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def csr_tensor(torch, np, p):
    return torch.sparse_csr_tensor(torch.from_numpy(p.row_ptr), torch.from_numpy(p.col_idx.astype(np.int64)),
                                   torch.from_numpy(p.values), size=(p.rows, p.cols))


def spmm_sweep(torch, np, am, P, widths=(1, 4, 16, 64), reps=20):
    """Achieved GB/s of the CSR products on a matrix beyond the last-level cache."""
    p = P.csr_rows(2_000_000, 200_000, 12, seed=9, regularization=0.0)
    M = am.DeviceMatrix(csr_tensor(torch, np, p), value_dtype=torch.float32)
    out = {"matrix": f"csr_{p.rows}x{p.cols}_nnz{p.nnz}_f32", "widths": {}}
    lib, h = M.context._lib, M.context.handle
    import ctypes
    for s in widths:
        X = torch.randn(p.cols, s, device="cuda", dtype=torch.float32)
        Y = torch.empty(p.rows, s, device="cuda", dtype=torch.float32)
        Z = torch.empty(p.cols, s, device="cuda", dtype=torch.float32)
        row = {}
        for tr, (src, dst, ldi, ldo) in ((0, (X, Y, s, s)), (1, (Y, Z, s, s))):
            kinds = [("", -1)] + ([("_stream", lib.amfls_spmm_variant_count() - 1)] if s == 1 else [])
            for tag, var in kinds:
                M.set_variant(s, bool(tr), var)
                torch.cuda.synchronize()
                for _ in range(3):
                    lib.amfls_operator_apply_rows(h, M.handle, tr, ctypes.c_void_p(src.data_ptr()), s, ldi,
                                                  ctypes.c_void_p(dst.data_ptr()), ldo, 1)
                M.context.synchronize()
                M.profile(True)
                for _ in range(reps):
                    lib.amfls_operator_apply_rows(h, M.handle, tr, ctypes.c_void_p(src.data_ptr()), s, ldi,
                                                  ctypes.c_void_p(dst.data_ptr()), ldo, 1)
                M.context.synchronize()
                n, sec, by = M.kernel_stats()
                M.profile(False)
                row[("AT" if tr else "A") + tag] = {"us": 1e6 * sec / n, "GBs": by / sec / 1e9,
                                                    "frac": by / sec / 1e9 / HBM_PEAK_GBS}
            M.set_variant(s, bool(tr), -1)
        out["widths"][str(s)] = row
    del M
    return out


def dense_c2(torch, am, P, rows=1_000_000, cols=2000):
    """The 1e6 x 2000 BF16 dense problem: FP64 certified solve (A upconverted
    exactly on the matrix cores) and the product kernels' rates."""
    import ctypes
    a, b, xs = P.config2_device(rows, cols)
    M = am.DeviceMatrix(a, value_dtype=torch.bfloat16)
    out = {"matrix": f"dense_{rows}x{cols}_bf16_GD_6decades_lambda0"}
    lib, h = M.context._lib, M.context.handle
    prod = {}
    for s in (1, 8, 16, 64):
        for dt, code, name in ((torch.float64, 0, "f64"), (torch.float32, 1, "f32")):
            X = torch.randn(cols, s, device="cuda", dtype=dt)
            Y = torch.empty(rows, s, device="cuda", dtype=dt)
            Z = torch.empty(cols, s, device="cuda", dtype=dt)
            for tr, (src, dst) in ((0, (X, Y)), (1, (Y, Z))):
                # widths <= 8: the default kernel and the other family (vector
                # ALUs by default for one column, matrix cores above)
                alt = ((1, "_mfma") if s == 1 else (0, "_valu"))
                for var, tag in (((-1, ""), alt) if s <= 8 else ((-1, ""),)):
                    M.set_variant(s, bool(tr), var)
                    lib.amfls_operator_apply_rows(h, M.handle, tr, ctypes.c_void_p(src.data_ptr()), s, s,
                                                  ctypes.c_void_p(dst.data_ptr()), s, code)
                    M.context.synchronize()
                    M.profile(True)
                    for _ in range(5):
                        lib.amfls_operator_apply_rows(h, M.handle, tr, ctypes.c_void_p(src.data_ptr()), s, s,
                                                      ctypes.c_void_p(dst.data_ptr()), s, code)
                    M.context.synchronize()
                    n, sec, by = M.kernel_stats()
                    M.profile(False)
                    prod[f"{'AT' if tr else 'A'}_s{s}_{name}{tag}"] = {
                        "ms": 1e3 * sec / n, "GBs": by / sec / 1e9, "TFs": 2.0 * rows * cols * s / (sec / n) / 1e12}
                M.set_variant(s, bool(tr), -1)
    out["products"] = prod
    for mode in ("fp64", "mixed"):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = am.run_amfls(M, b, tolerance=1e-8, precision=mode)
        torch.cuda.synchronize()
        out[mode] = {"seconds": time.perf_counter() - t0, "status": int(r.status), "depth": r.depth,
                     "auxiliary_width": r.auxiliary_width, "basis_rank": r.basis_rank,
                     "backward_error": r.backward_error_upper_bound, "refinement_steps": r.refinement_steps}
        log(f"C2 {mode}: {out[mode]}")
    del M, a
    return out


def batch_c3(torch, np, am, problems=256, lambdas=8, workers=6, seed=3, tolerance=1e-8):
    """The batched configuration: `problems` independent CSR problems
    (m ~ U[5000, 200000], n ~ U[500, min(10000, m)], row lengths ~ U[4, 32],
    N(0,1) values stored as float32, b ~ N(0,1)), each with `lambdas` values
    logspace(-6, -1); built on the device by the counter-addressed generator,
    autotuned at width 1, then solved concurrently by `workers` host threads
    with a stream each, one regularization sweep per problem (FP64 blocks,
    FP64 certificates)."""
    from amfls_solver_amd.binding import generate_csr
    rng = np.random.default_rng(seed)
    dims = []
    for _ in range(problems):
        m = int(rng.integers(5000, 200_001))
        n = int(rng.integers(500, min(10_000, m) + 1))
        dims.append((m, n))
    t0 = time.perf_counter()
    mats, bs = [], []
    for i, (m, n) in enumerate(dims):
        rp, ci, val = generate_csr(m, n, (4, 32), seed=1000 + i, value_dtype=torch.float32)
        mats.append(am.DeviceMatrix.from_csr_arrays(m, n, rp, ci, val))
        bs.append(am.binding.gaussian_block(m, 1, 2000 + i, 0).reshape(-1))
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    for M in mats:
        M.autotune(widths=(1,), dtype=torch.float64, reps=3)
    t_tune = time.perf_counter() - t0
    lam = np.tile(np.logspace(-6.0, -1.0, lambdas), (problems, 1))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = am.run_amfls_batch(mats, bs, lam, workers=workers, tolerance=tolerance)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ok = sum(1 for row in res for r in row if r.status == 0)
    statuses = {}
    for row in res:
        for r in row:
            statuses[str(r.status)] = statuses.get(str(r.status), 0) + 1
    deep = sorted(range(problems), key=lambda i: -max(r.depth for r in res[i]))[:3]
    prods = sum(row[0].a_block_calls + row[0].at_block_calls for row in res)
    nnz = sum(M.nnz for M in mats)
    out = {"problems": problems, "lambdas_per_problem": lambdas, "workers": workers, "total_nnz": nnz,
           "seconds": dt, "certified": ok, "solves": problems * lambdas, "solves_per_s": problems * lambdas / dt,
           "block_products": prods, "block_products_per_s": prods / dt, "build_seconds": t_build,
           "autotune_seconds": t_tune, "max_depth": max(r.depth for row in res for r in row),
           "statuses": statuses,
           "deepest": [{"m": dims[i][0], "n": dims[i][1], "depth": [r.depth for r in res[i]],
                        "status": [r.status for r in res[i]]} for i in deep]}
    # the same first problems solved one lambda at a time, for the sharing gain
    k = min(8, problems)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(k):
        for l in lam[i]:
            am.run_amfls(mats[i], bs[i], regularization=float(l), tolerance=tolerance)
    torch.cuda.synchronize()
    t_sep = time.perf_counter() - t0
    t0 = time.perf_counter()
    for i in range(k):
        am.run_amfls_sweep(mats[i], bs[i], lam[i], tolerance=tolerance)
    torch.cuda.synchronize()
    t_sw = time.perf_counter() - t0
    out["first_problems_serial"] = {"problems": k, "separate_solves_seconds": t_sep, "sweep_seconds": t_sw}
    del mats, bs
    return out


def cpu_baseline(np, P, seconds_budget=25.0):
    """Oracle (CPU restatement, numpy/scipy FP64) on the same workload, 1 BLAS thread."""
    from threadpoolctl import threadpool_limits
    from oracle import operators as O
    from oracle import solvers as S
    p = P.config1()
    times = []
    res = None
    with threadpool_limits(limits=1):
        op = O.CsrOperator(p.rows, p.cols, p.row_ptr, p.col_idx, p.values.astype(np.float64))
        t_start = time.time()
        while len(times) < 3 and (time.time() - t_start) < seconds_budget:
            t0 = time.perf_counter()
            res = S.solve_amfls(op, p.b, regularization=p.regularization, tolerance=1e-8)
            times.append(time.perf_counter() - t0)
    return float(np.median(times)), len(times), res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--headline", choices=["mixed", "fp64"], default="fp64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-dense", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-problems", type=int, default=64,
                    help="problems of the batched configuration to run (the full configuration has 256)")
    ap.add_argument("--batch-workers", type=int, default=6,
                    help="host threads of the batch, a stream each (six fill the GPU best on the batched workload)")
    ap.add_argument("--pmc-calibrate", action="store_true",
                    help="start with a 1 GiB device copy (reference kernel for PMC byte counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {a.gpus}); multi-GPU runs are not part of this build",
              file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import amfls_solver_amd as am
    from amfls_solver_amd import problems as P

    if a.pmc_calibrate:
        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda")  # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log("pmc calibration copy: 1 GiB read + 1 GiB written")
        del src, dst

    log("generating 200000x20000 CSR (12 nnz/row, lambda 1e-3)")
    p = P.config1()
    M = am.DeviceMatrix(csr_tensor(torch, np, p), value_dtype=torch.float32)
    b = torch.from_numpy(p.b).cuda()
    # width-one kernel variants: the measured default table (an autotuned
    # choice varies from run to run between variants within noise of each
    # other, which would make the PMC traffic runs incomparable)
    lib = M.context._lib
    mean_a, mean_at = M.nnz / M.rows, M.nnz / M.cols
    variants = {"A": lib.amfls_spmm_variant_name(lib.amfls_spmm_default_variant(1, mean_a)).decode(),
                "AT": lib.amfls_spmm_variant_name(lib.amfls_spmm_default_variant(1, mean_at)).decode()}
    modes = ["fp64", "mixed"]
    results = {}
    for mode in modes:
        for _ in range(a.warmup):
            am.run_amfls(M, b, regularization=p.regularization, tolerance=1e-8, precision=mode)
        log(f"warm {mode}")
    for mode in modes:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        last = None
        for _ in range(a.steps):
            last = am.run_amfls(M, b, regularization=p.regularization, tolerance=1e-8, precision=mode)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / a.steps
        if mode == a.headline:
            # the same K solves again with every product's kernel timed by its
            # dispatch timestamps (kept out of the headline timing: recording
            # and resolving the event pairs costs host time)
            M.profile(True)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for _ in range(a.steps):
                am.run_amfls(M, b, regularization=p.regularization, tolerance=1e-8, precision=mode)
            torch.cuda.synchronize()
            dtp = (time.perf_counter() - t1) / a.steps
            n, sec, by = M.kernel_stats()
            M.profile(False)
            results["kernel"] = (n, sec, by, dtp)
        results[mode] = (dt, last)
        log(f"{mode}: {dt:.4f} s/solve status {last.status} cert {last.relative_energy_error_upper_bound:.3e} "
            f"depth {last.depth} width {last.auxiliary_width} rank {last.basis_rank} refine {last.refinement_steps}")

    # learned width policy (shipped model) on the same workload, FP64
    from amfls_solver_amd.width_model import WidthModel
    wm = WidthModel.load(context=M.context)
    wm.attach()
    am.run_amfls(M, b, regularization=p.regularization, tolerance=1e-8, width_policy=1, context=M.context)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        rl = am.run_amfls(M, b, regularization=p.regularization, tolerance=1e-8, width_policy=1, context=M.context)
    torch.cuda.synchronize()
    learned_s = (time.perf_counter() - t0) / a.steps
    log(f"learned width policy: {learned_s:.4f} s/solve status {rl.status} width {rl.auxiliary_width}")

    dt, r = results[a.headline]
    other = "mixed" if a.headline == "fp64" else "fp64"
    n, sec, by, dt_profiled = results["kernel"]
    achieved = by / sec / 1e9 if sec > 0 else 0.0
    sol = {m: results[m][1].solution.numpy() for m in modes}

    cpu = None
    if not a.no_cpu_baseline:
        log("cpu baseline (oracle, 1 thread)")
        cpu_s, nrun, ro = cpu_baseline(np, P)
        rel = {m: float(np.linalg.norm(sol[m] - ro.solution) / np.linalg.norm(ro.solution)) for m in modes}
        cpu = {"value": cpu_s, "unit": "s", "cores": 1, "kind": "port",
               "sample": f"{nrun} full certified solves of the same 200000x20000 CSR workload "
                         f"(numpy/scipy FP64 restatement, BLAS limited to 1 thread); median",
               "status": int(ro.status), "depth": int(ro.depth)}
    else:
        rel = {}

    dense = None
    if not a.no_dense:
        log("dense 1000000x2000 bf16 (matrix cores)")
        dense = dense_c2(torch, am, P)

    sweep = None
    if not a.no_sweep:
        log("SpMM sweep on 2000000x200000 CSR (beyond the 256 MB cache)")
        sweep = spmm_sweep(torch, np, am, P)

    batch = None
    if not a.no_batch and a.batch_problems > 0:
        log(f"batched CSR problems: {a.batch_problems} x 8 lambdas")
        batch = batch_c3(torch, np, am, problems=a.batch_problems, workers=a.batch_workers)
        log(f"batch: {batch}")

    line = {
        "metric": "time-to-rel-err 1e-8 (s): mixed/learned vs FP64 AMFLS, 1 GPU; SpMM GB/s",
        "value": dt,
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * dt,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.headline == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": "csr_200000x20000_nnz12_f32values_lambda1e-3_relenergy1e-8",
            "mode": a.headline,
            "certified_status": int(r.status),
            "certificate": r.relative_energy_error_upper_bound,
            "depth": r.depth, "auxiliary_width": r.auxiliary_width, "basis_rank": r.basis_rank,
            "block_products": r.a_block_calls + r.at_block_calls,
            "block_products_per_s": (r.a_block_calls + r.at_block_calls) / dt,
            f"{other}_seconds": results[other][0],
            f"{other}_status": int(results[other][1].status),
            f"{other}_certificate": results[other][1].relative_energy_error_upper_bound,
            "mixed_refinement_steps": results["mixed"][1].refinement_steps,
            "learned_width_seconds": learned_s,
            "learned_width_status": int(rl.status),
            "learned_model_heldout_speedup": wm.meta.get("heldout_speedup"),
            "rel_diff_vs_oracle": rel,
            "spmm_variants_width1": variants,
            "parallelism": "single_gpu",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(),
            "kernel": "csr_spmm (A and A^T products of the timed solves)",
            "launches": n,
            "avg_us": 1e6 * sec / max(n, 1),
            "timed_pass": f"{a.steps} further solves of the workload with per-product dispatch-timestamp events "
                          f"({1e3 * dt_profiled:.2f} ms per solve with the events)",
            "note": "working set fits the 256 MB last-level cache; see spmm_sweep for an HBM-resident matrix",
        },
        "cpu_baseline": cpu,
        "spmm_sweep": sweep,
        "dense_1e6x2000_bf16": dense,
        "batched_csr": batch,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
