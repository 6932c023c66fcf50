#!/usr/bin/env python
"""Time-to-relative-error benchmark of the AMFLS HIP engine on one MI355X.

Headline workload: the 1,000,000 x 2000 dense least-squares problem A = G D
(G i.i.d. N(0,1) from the counter-addressed Gaussian stream, D log-spaced
from 1 to 1e-6, A stored as bfloat16, b = A x* + 1e-8-relative noise,
lambda = 0).  A step is one complete solve with A and b already resident in
HBM.  `value` is the seconds per solve of mixed-precision AMFLS with the
learned width controller (FP32 block Golub-Kahan search on the split-bf16
matrix cores whose widening decisions the shipped MLP takes, FP64 LSQR
refinement preconditioned by the search's factorization, FP64 certificate)
run until its forward-error estimate is at most 1e-9; every timed solution's
forward relative error against the FP64 oracle solution
(tests/golden/c2_solution.npz, TSQR of the stored matrix) must be at most
1e-8, with no fallback to the FP64 engine, or the line reports "valid":
false.  Beside it: the same mixed engine with the reference's heuristic
width controller, plain FP64 AMFLS on the same workload (its time to the
same forward error) and the ratios; the product kernels' achieved HBM
bandwidth in the timed solves (HIP events on the engine stream); the 200000
x 20000 CSR solve (lambda 1e-3) with its SpMM
bandwidth; an SpMM sweep beyond the 256 MB last-level cache; the batched CSR
configuration; and the FP64 CPU restatement (oracle) timed on a sample of the
dense workload on this host's cores.

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


def measured_traffic(name="csr_spmm_c1_pmc_traffic.txt", kernel=None):
    """HBM bytes per launch from a committed PMC summary under profiles/
    (FETCH_SIZE and WRITE_SIZE passes scaled by a calibration copy): the
    "traffic_bytes_per_launch:" line, or with `kernel` the
    "traffic_bytes_per_launch[<kernel>]:" line; None when absent."""
    path = os.path.join(ROOT, "profiles", name)
    key = "traffic_bytes_per_launch:" if kernel is None else f"traffic_bytes_per_launch[{kernel}]:"
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


DENSE_PMC_PROFILE = "dense_c2_split3_pmc_traffic.txt"


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
            names = [lib.amfls_spmm_variant_name(v).decode() for v in range(lib.amfls_spmm_variant_count())]
            kinds = [("", -1)] + ([("_stream", names.index("stream")), ("_stream8k", names.index("stream8k")),
                                   ("_slab", names.index("slab"))] if s == 1 else [])
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


C2_FIXTURE = os.path.join(ROOT, "tests", "golden", "c2_solution.npz")


def bf16_checksums(torch, np, a):
    """Bit checksums of a row-major bf16 matrix (as tests/golden/make_c2_fixture.py computes them)."""
    rows, cols = a.shape
    tot = np.zeros(3, dtype=np.int64)
    wcol = torch.arange(1, cols + 1, device=a.device, dtype=torch.int64)[None, :]
    for i0 in range(0, rows, 50_000):
        i1 = min(rows, i0 + 50_000)
        bits = a[i0:i1].view(torch.int16).to(torch.int64) & 0xFFFF
        wrow = (torch.arange(i0, i1, device=a.device, dtype=torch.int64) % 1009 + 1)[:, None]
        tot += np.array([int(bits.sum()), int((bits * wcol).sum()), int((bits * wrow).sum())], dtype=np.int64)
    return tot


def dense_products(torch, M, rows, cols, widths=(1, 16, 32, 64), reps=5):
    """Per-launch time and achieved bandwidth of the dense products (FP32 and FP64 blocks)."""
    import ctypes
    lib, h = M.context._lib, M.context.handle
    prod = {}
    for s_ in widths:
        for dt, code, name in ((torch.float64, 0, "f64"), (torch.float32, 1, "f32")):
            X = torch.randn(cols, s_, device="cuda", dtype=dt)
            Y = torch.empty(rows, s_, device="cuda", dtype=dt)
            Z = torch.empty(cols, s_, device="cuda", dtype=dt)
            for tr, (src, dst) in ((0, (X, Y)), (1, (Y, Z))):
                lib.amfls_operator_apply_rows(h, M.handle, tr, ctypes.c_void_p(src.data_ptr()), s_, s_,
                                              ctypes.c_void_p(dst.data_ptr()), s_, code)
                M.context.synchronize()
                M.profile(True)
                for _ in range(reps):
                    lib.amfls_operator_apply_rows(h, M.handle, tr, ctypes.c_void_p(src.data_ptr()), s_, s_,
                                                  ctypes.c_void_p(dst.data_ptr()), s_, code)
                M.context.synchronize()
                n, sec, by = M.kernel_stats()
                M.profile(False)
                prod[f"{'AT' if tr else 'A'}_s{s_}_{name}"] = {
                    "ms": 1e3 * sec / n, "GBs": by / sec / 1e9, "TFs": 2.0 * rows * cols * s_ / (sec / n) / 1e12}
    return prod


def dense_kernel_name(c):
    """Product kernel of one class of bf16 dense products (the events around
    the product also cover its small pack or reduce launch)."""
    lo, hi = c["widths"]
    if hi == 1:
        return f"narrow_{'tn' if c['transpose'] else 'nn'}_kernel<bf16, {'float' if c['block'] == 'f32' else 'double'}, 1>"
    nt = (hi + 15) // 16
    if c["block"] == "f32":
        if c["transpose"]:
            # products of at most 32 columns on the 2000-column matrix: the direct-to-LDS kernels
            return f"tn_lds_kernel<float, {nt}>" if nt <= 2 else f"tn_split3_kernel<float, {nt}>"
        return f"nn_lds_kernel<float, {nt}>" if nt <= 2 else f"nn_split3_kernel<float, {nt}>"
    return f"dense_{'tn' if c['transpose'] else 'nn_f64'}_kernel<bf16, double, ..., {nt}>"


def c2_headline(torch, np, am, P, steps, warmup, forward_target=1e-8):
    """The 1e6 x 2000 bf16 dense workload: K timed solves per engine, forward
    errors against the FP64 oracle, kernel bandwidth of the mixed solves."""
    a, b, _ = P.config2_device()
    rows, cols = a.shape
    fx = np.load(C2_FIXTURE) if os.path.exists(C2_FIXTURE) else None
    same = fx is not None and bool(np.array_equal(bf16_checksums(torch, np, a), fx["bits_checksums"]))
    x_ref = fx["x_ls"] if same else None
    M = am.DeviceMatrix(a, value_dtype=torch.bfloat16)
    # the shipped learned width controller, attached to the operator's context
    from amfls_solver_amd.width_model import WidthModel
    wm = WidthModel.load(context=M.context)
    wm.attach()
    engines = {
        "mixed_learned": dict(precision="mixed", forward_tolerance=0.1 * forward_target, width_policy=1,
                              context=M.context),
        "mixed": dict(precision="mixed", forward_tolerance=0.1 * forward_target),
        "fp64": dict(precision="fp64"),
    }
    out = {"matrix": f"dense_{rows}x{cols}_bf16_GD_6decades_lambda0", "matrix_matches_fixture": same}
    for name, kw in engines.items():
        for _ in range(warmup):
            am.run_amfls(M, b, tolerance=1e-8, **kw)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = []
        for _ in range(steps):
            res.append(am.run_amfls(M, b, tolerance=1e-8, **kw))
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        fe = [float(np.linalg.norm(r.solution.numpy() - x_ref) / np.linalg.norm(x_ref)) if x_ref is not None
              else float("nan") for r in res]
        r = res[-1]
        out[name] = {"seconds": dt, "status": [int(q.status) for q in res], "forward_error": fe,
                     "fp64_fallback": [bool(q.fp64_fallback) for q in res],
                     "forward_error_max": max(fe), "forward_error_estimate": r.forward_error_estimate,
                     "depth": r.depth, "auxiliary_width": r.auxiliary_width, "basis_rank": r.basis_rank,
                     "refinement_iterations": r.refinement_steps,
                     "backward_error": r.backward_error_upper_bound,
                     "block_products": r.a_block_calls + r.at_block_calls,
                     "block_products_per_s": (r.a_block_calls + r.at_block_calls) / dt,
                     "operator_seconds": r.a_seconds + r.at_seconds}
        log(f"C2 {name}: {dt:.4f} s/solve status {out[name]['status']} forward error max {max(fe):.3e} "
            f"depth {r.depth} width {r.auxiliary_width + 1} refine {r.refinement_steps}")
    d0 = wm.decisions()
    am.run_amfls(M, b, tolerance=1e-8, **engines["mixed_learned"])
    d1 = wm.decisions()
    out["mixed_learned"]["model_decisions"] = d1[0] - d0[0]
    out["mixed_learned"]["heuristic_fallbacks"] = d1[1] - d0[1]
    # the same headline solves again with every product's kernel timed by HIP
    # events on the engine stream (kept out of the headline timing)
    M.profile(True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        am.run_amfls(M, b, tolerance=1e-8, **engines["mixed_learned"])
    torch.cuda.synchronize()
    out["profiled_seconds_per_solve"] = (time.perf_counter() - t1) / steps
    # per-class view: each (direction, width range, block type) class runs one
    # product kernel; the roofline reports the class with the most time
    classes = []
    for tr in (False, True):
        for lo, hi in ((1, 1), (2, 16), (17, 32), (33, 48), (49, 64)):
            for dt in (torch.float32, torch.float64):
                cn, cs, cb = M.kernel_stats_class(tr, lo, hi, dt)
                if cn:
                    classes.append({"transpose": tr, "widths": [lo, hi], "block": "f32" if dt == torch.float32
                                    else "f64", "launches": cn, "seconds": cs, "bytes": cb,
                                    "avg_us": 1e6 * cs / cn, "GBs": cb / cs / 1e9})
    n, sec, by = M.kernel_stats()
    M.profile(False)
    classes.sort(key=lambda c: -c["seconds"])
    out["kernel"] = {"launches": n, "seconds": sec, "bytes": by, "classes": classes}
    out["products"] = dense_products(torch, M, rows, cols)
    del M, a, b
    torch.cuda.empty_cache()
    return out


def c1_section(torch, np, am, P, steps, warmup):
    """The 200000 x 20000 CSR workload (lambda 1e-3): FP64 and mixed solves
    and the SpMM kernels' bandwidth in the timed FP64 solves."""
    p = P.config1()
    M = am.DeviceMatrix(csr_tensor(torch, np, p), value_dtype=torch.float32)
    b = torch.from_numpy(p.b).cuda()
    lib = M.context._lib
    mean_a, mean_at = M.nnz / M.rows, M.nnz / M.cols
    variants = {"A": lib.amfls_spmm_variant_name(lib.amfls_spmm_default_variant(1, mean_a)).decode(),
                "AT": lib.amfls_spmm_variant_name(lib.amfls_spmm_default_variant(1, mean_at)).decode()}
    out = {"matrix": "csr_200000x20000_nnz12_f32values_lambda1e-3", "spmm_variants_width1": variants}
    sols = {}
    for mode in ("fp64", "mixed"):
        for _ in range(warmup):
            am.run_amfls(M, b, regularization=p.regularization, tolerance=1e-8, precision=mode)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            r = am.run_amfls(M, b, regularization=p.regularization, tolerance=1e-8, precision=mode)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        sols[mode] = r.solution.numpy()
        out[mode] = {"seconds": dt, "status": int(r.status), "certificate": r.relative_energy_error_upper_bound,
                     "depth": r.depth, "auxiliary_width": r.auxiliary_width, "basis_rank": r.basis_rank,
                     "refinement_iterations": r.refinement_steps,
                     "block_products_per_s": (r.a_block_calls + r.at_block_calls) / dt}
        log(f"C1 {mode}: {dt:.4f} s/solve status {r.status} cert {r.relative_energy_error_upper_bound:.3e}")
    # the same FP64 solve with its width-one level batches captured into HIP graphs
    M.context.set_level_graphs(True)
    try:
        am.run_amfls(M, b, regularization=p.regularization, tolerance=1e-8, precision="fp64")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            r = am.run_amfls(M, b, regularization=p.regularization, tolerance=1e-8, precision="fp64")
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
    finally:
        M.context.set_level_graphs(False)
    out["fp64_level_graphs"] = {"seconds": dt, "status": int(r.status), "depth": r.depth,
                                "rel_diff_vs_direct": float(np.linalg.norm(r.solution.numpy() - sols["fp64"]) /
                                                            np.linalg.norm(sols["fp64"]))}
    log(f"C1 fp64 with level graphs: {dt:.4f} s/solve status {r.status}")
    M.profile(True)
    torch.cuda.synchronize()
    for _ in range(steps):
        am.run_amfls(M, b, regularization=p.regularization, tolerance=1e-8, precision="fp64")
    torch.cuda.synchronize()
    n, sec, by = M.kernel_stats()
    M.profile(False)
    achieved = by / sec / 1e9 if sec > 0 else 0.0
    out["spmm"] = {"launches": n, "avg_us": 1e6 * sec / max(n, 1), "achieved_GBs": achieved,
                   "frac": achieved / HBM_PEAK_GBS, "traffic_bytes_per_launch_committed": measured_traffic(),
                   "note": "working set fits the 256 MB last-level cache; spmm_sweep is HBM-resident"}
    out["mixed_rel_diff_vs_fp64"] = float(np.linalg.norm(sols["mixed"] - sols["fp64"]) /
                                          np.linalg.norm(sols["fp64"]))
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


def host_cpu():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count() or 1


def cpu_baseline(torch, np, P, rows=100_000, cols=2000, single_rows=20_000):
    """The FP64 CPU restatement (oracle AMFLS with numpy BLAS) on samples of
    the dense workload: its first `rows` rows with BLAS on every host core
    the GPU box grants (at most 16), and its first `single_rows` rows on one
    thread; same G D construction, same tolerance; times scaled by
    1e6 / rows to the full workload (products and left-block work are linear
    in the rows)."""
    from threadpoolctl import threadpool_limits
    from oracle import operators as O
    from oracle import solvers as S
    model, ncpu = host_cpu()
    cores = min(16, ncpu)
    runs = {}
    for label, nrows, threads in (("all", rows, cores), ("single", single_rows, 1)):
        a, b, _ = P.config2_device(nrows, cols)
        ah = a.double().cpu().numpy()
        bh = b.cpu().numpy()
        del a, b
        with threadpool_limits(limits=threads):
            op = O.RowMajorDenseOperator(ah)
            t0 = time.perf_counter()
            res = S.solve_amfls(op, bh, tolerance=1e-8)
            dt = time.perf_counter() - t0
        del ah, op
        runs[label] = (nrows, threads, dt, res)
    nrows, threads, dt, res = runs["all"]
    scale = 1_000_000 / nrows
    srows, _, sdt, sres = runs["single"]
    sscale = 1_000_000 / srows
    return {"value": dt * scale, "unit": "s", "cores": cores, "kind": "port",
            "sample": f"oracle FP64 AMFLS (numpy/scipy restatement, BLAS on {cores} threads) on the first {nrows} "
                      f"rows of the dense G D workload (x{scale:g} to 1e6 rows); measured {dt:.2f} s, "
                      f"status {int(res.status)}, depth {int(res.depth)}",
            "single_thread": {"value": sdt * sscale, "unit": "s", "cores": 1,
                              "sample": f"same restatement on one thread, first {srows} rows (x{sscale:g}); "
                                        f"measured {sdt:.2f} s, status {int(sres.status)}, depth {int(sres.depth)}"},
            "host_cpu": model, "host_nproc": ncpu}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--no-dense", action="store_true",
                    help="skip the dense headline section (profiling runs of the CSR sections; no headline value)")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-problems", type=int, default=256,
                    help="problems of the batched configuration (256 is the full configuration)")
    ap.add_argument("--batch-workers", type=int, default=6,
                    help="host threads of the batch, a stream each (six fill the GPU best on the batched workload)")
    ap.add_argument("--pmc-calibrate", action="store_true",
                    help="start with a 1 GiB device copy (reference kernel for PMC byte counters)")
    ap.add_argument("--headline-solves-only", action="store_true",
                    help="profiling runs: only the headline engine's dense solves (warmup + steps), one summary line")
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

    if a.headline_solves_only:
        return headline_solves_only(torch, np, am, P, a.steps, a.warmup)

    c2 = None
    if not a.no_dense:
        log("dense 1000000x2000 bf16 workload")
        c2 = c2_headline(torch, np, am, P, a.steps, a.warmup)

    c1 = None if a.no_c1 else c1_section(torch, np, am, P, a.steps, a.warmup)
    sweep = None
    if not a.no_sweep:
        log("SpMM sweep on 2000000x200000 CSR (beyond the 256 MB cache)")
        sweep = spmm_sweep(torch, np, am, P)
    batch = None
    if not a.no_batch and a.batch_problems > 0:
        log(f"batched CSR problems: {a.batch_problems} x 8 lambdas")
        batch = batch_c3(torch, np, am, problems=a.batch_problems, workers=a.batch_workers)
        log(f"batch: {batch}")
    cpu = None
    if not a.no_cpu_baseline:
        log("cpu baseline (oracle on a sample of the dense workload)")
        cpu = cpu_baseline(torch, np, P)
        log(f"cpu baseline: {cpu}")

    line = headline_fields(c2, a) if c2 is not None else {
        "value": None, "ms_per_step": None, "valid": False,
        "config": {"workload": "dense_1000000x2000_bf16_GD_6decades_lambda0_forward_rel_err_1e-8",
                   "note": "dense section skipped (--no-dense): no headline in this run"},
        "roofline": None}
    line = {
        "metric": "time-to-rel-err 1e-8 (s): mixed/learned vs FP64 AMFLS, 1 GPU; SpMM GB/s",
        "value": line["value"],
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": line["ms_per_step"],
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "valid": line["valid"],
        "config": line["config"],
        "roofline": line["roofline"],
        "cpu_baseline": cpu,
        "dense_1e6x2000_bf16": c2,
        "csr_200000x20000": c1,
        "spmm_sweep": sweep,
        "batched_csr": batch,
    }
    print(json.dumps(line), flush=True)
    return 0


def headline_solves_only(torch, np, am, P, steps, warmup):
    """The headline engine's solves of the dense workload alone (mixed
    precision, learned width controller), for a kernel trace of just them."""
    a, b, _ = P.config2_device()
    M = am.DeviceMatrix(a, value_dtype=torch.bfloat16)
    from amfls_solver_amd.width_model import WidthModel
    wm = WidthModel.load(context=M.context)
    wm.attach()
    kw = dict(precision="mixed", forward_tolerance=1e-9, width_policy=1, context=M.context, tolerance=1e-8)
    for _ in range(warmup):
        am.run_amfls(M, b, **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r = am.run_amfls(M, b, **kw)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    print(json.dumps({"headline_solves": steps, "warmup": warmup, "seconds_per_solve": dt, "status": int(r.status),
                      "depth": r.depth, "refinement_iterations": r.refinement_steps}), flush=True)
    return 0


def headline_fields(c2, a):
    """value, validity, config and roofline of the JSON line from the dense
    section: the headline engine is mixed precision with the learned width
    controller."""
    mixed, heur, fp64 = c2["mixed_learned"], c2["mixed"], c2["fp64"]
    valid = (c2["matrix_matches_fixture"] and mixed["forward_error_max"] <= 1e-8 and
             fp64["forward_error_max"] <= 1e-8 and all(st == 0 for st in mixed["status"]) and
             not any(mixed["fp64_fallback"]))
    top = c2["kernel"]["classes"][0]
    kn, ksec, kby = top["launches"], top["seconds"], top["bytes"]
    achieved = kby / ksec / 1e9 if ksec > 0 else 0.0
    return {
        "value": mixed["seconds"],
        "ms_per_step": 1e3 * mixed["seconds"],
        "valid": valid,
        "config": {
            "workload": "dense_1000000x2000_bf16_GD_6decades_lambda0_forward_rel_err_1e-8",
            "engine": "mixed/learned: FP32 block Golub-Kahan search (split-bf16 matrix cores) with the learned "
                      "width controller, FP64 LSQR refinement preconditioned by its factorization, FP64 certificate",
            "forward_error_max": mixed["forward_error_max"],
            "forward_error_estimate": mixed["forward_error_estimate"],
            "status": mixed["status"],
            "depth": mixed["depth"], "auxiliary_width": mixed["auxiliary_width"],
            "refinement_iterations": mixed["refinement_iterations"],
            "block_products_per_s": mixed["block_products_per_s"],
            "fp64_seconds": fp64["seconds"],
            "fp64_forward_error_max": fp64["forward_error_max"],
            "fp64_status": fp64["status"],
            "speedup_mixed_learned_vs_fp64": fp64["seconds"] / mixed["seconds"],
            "model_decisions_per_solve": mixed["model_decisions"],
            "mixed_heuristic_seconds": heur["seconds"],
            "mixed_heuristic_forward_error_max": heur["forward_error_max"],
            "mixed_heuristic_status": heur["status"],
            "speedup_mixed_heuristic_vs_fp64": fp64["seconds"] / heur["seconds"],
            "matrix_matches_oracle_fixture": c2["matrix_matches_fixture"],
            "parallelism": "single_gpu",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(DENSE_PMC_PROFILE, dense_kernel_name(top)),
            "traffic_source": f"profiles/{DENSE_PMC_PROFILE} (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE of the same "
                              f"kernel at its widest width in the class, scaled by a 1 GiB calibration copy; "
                              f"committed, not measured in this run)",
            "kernel": dense_kernel_name(top),
            "product_class": {k: top[k] for k in ("transpose", "widths", "block")},
            "bytes_per_launch": kby / max(kn, 1),
            "launches": kn,
            "avg_us": 1e6 * ksec / max(kn, 1),
            "all_products_GBs": c2["kernel"]["bytes"] / c2["kernel"]["seconds"] / 1e9
            if c2["kernel"]["seconds"] > 0 else None,
            "timed_pass": f"{a.steps} further mixed solves with HIP events around every product on the engine "
                          f"stream ({1e3 * c2['profiled_seconds_per_solve']:.1f} ms per solve with the events)",
        },
    }


if __name__ == "__main__":
    sys.exit(main())
