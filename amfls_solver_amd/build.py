"""Build libamfls_hip.so (HIP kernels + C++20 host engine + C ABI) for gfx950.

Usage: python -m amfls_solver_amd.build [--jobs N]

Every source is compiled by hipcc with the packed-FP32 instruction feature
disabled, then linked into one shared library next to this file.  The build
runs on a machine without a GPU (hipcc cross-compiles gfx950).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
LIB = HERE / "libamfls_hip.so"
ARCH = os.environ.get("AMFLS_OFFLOAD_ARCH", "gfx950")

SOURCES = [
    CSRC / "device" / "blas.hip",
    CSRC / "device" / "spmm.hip",
    CSRC / "device" / "rank_one.hip",
    CSRC / "device" / "level_chain.hip",
    CSRC / "device" / "dense_mfma.hip",
    CSRC / "device" / "mlp.hip",
    CSRC / "device" / "generate.hip",
    CSRC / "device" / "dense_narrow.hip",
    CSRC / "device" / "noalloc.hip",
    CSRC / "device" / "tall.hip",
    CSRC / "engine" / "operator.cpp",
    CSRC / "engine" / "small_la.cpp",
    CSRC / "engine" / "validation.cpp",
    CSRC / "engine" / "workspace.cpp",
    CSRC / "engine" / "solver.cpp",
    CSRC / "engine" / "refine.cpp",
    CSRC / "engine" / "width_model.cpp",
    CSRC / "engine" / "capi.cpp",
]

COMMON = [
    "-std=c++20", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: Path, obj: Path) -> tuple[Path, str]:
    lang = ["-x", "hip"] if src.suffix == ".hip" else []
    cmd = [hipcc(), *COMMON, *lang, "-c", str(src), "-o", str(obj)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{p.stdout}\n{p.stderr}")
    return obj, p.stderr


def build(jobs: int | None = None, verbose: bool = False) -> Path:
    out_dir = HERE / "_build"
    out_dir.mkdir(exist_ok=True)
    sources = [s for s in SOURCES if s.exists()]
    objs = [out_dir / (s.stem + ("_hip" if s.suffix == ".hip" else "") + ".o") for s in sources]
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_compile, s, o) for s, o in zip(sources, objs)]
        for f in futs:
            obj, err = f.result()
            if verbose and err.strip():
                print(err, file=sys.stderr)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{p.stdout}\n{p.stderr}")
    os.replace(tmp, LIB)
    return LIB


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))


if __name__ == "__main__":
    main()
