"""The C++ drop-in header (include/amfls_hip.hpp): a program written
against the reference's amfls::MatrixOperator / solve_amfls API compiles and
links unchanged (CPU), and on the GPU solves the README's diagonal example,
drives a custom operator's cost and long-double refinement hooks, and maps
errors to the reference's exceptions."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "tests" / "native" / "cpp_dropin_example.cpp"
LIBDIR = ROOT / "amfls_solver_amd"


def _build(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not available")
    if not (LIBDIR / "libamfls_hip.so").exists():
        pytest.skip("libamfls_hip.so is not built")
    exe = tmp_path / "dropin"
    subprocess.run([gxx, "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", f"-I{ROOT / 'include'}", str(SRC),
                    "-o", str(exe), f"-L{LIBDIR}", "-lamfls_hip", f"-Wl,-rpath,{LIBDIR}"], check=True)
    return exe


def test_header_compiles_and_links(tmp_path):
    assert _build(tmp_path).exists()


@pytest.mark.gpu
def test_dropin_program_solves_on_gpu(gpu, tmp_path):
    exe = _build(tmp_path)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True, timeout=120).stdout
    kv = dict(line.split("=", 1) for line in out.strip().splitlines())
    assert kv["diag_status"] == "0" and float(kv["diag_max_error"]) <= 1e-12
    assert int(kv["diag_trace_rows"]) >= 1 and float(kv["diag_backward"]) <= 1e-10
    # unreachable tolerance: the basis is exhausted (status 2, basis limit)
    assert kv["dense_status"] == "2" and kv["dense_basis_rank"] == "3"
    assert int(kv["dense_refinement_calls"]) >= 2 and int(kv["dense_refinement_calls"]) % 2 == 0
    assert int(kv["dense_cost_queries"]) >= 1
    assert float(kv["dense_ls_backward"]) < 1e-15
    A = np.array([[1, 2, 0], [0, 1, 1], [3, 0, 1], [1, 1, 1], [2, -1, 0], [0, 0, 4]], dtype=float)
    x = np.linalg.lstsq(A, np.array([1, -2, 3, 0, 5, 1], dtype=float), rcond=None)[0]
    got = np.array([float(kv[f"dense_x{j}"]) for j in range(3)])
    assert np.max(np.abs(got - x)) <= 1e-13 * np.max(np.abs(x))
    assert kv["throwing"] == "domain_error:operator refused"
    assert kv["bad_options"] == "invalid_argument"
