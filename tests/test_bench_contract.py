"""CPU checks of the benchmark's command-line contract and of the loader's
refusal to run without the built engine (no GPU calls)."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, cwd=ROOT)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)
    assert r.stdout.strip() == ""  # no JSON line for an unsupported configuration


def test_missing_engine_library_fails_loudly(tmp_path):
    from amfls_solver_amd import _lib
    with pytest.raises(RuntimeError, match="not built"):
        _lib.load(tmp_path / "libamfls_hip.so")


def test_bench_help_lists_contract_flags():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--help"], capture_output=True, text=True,
                       timeout=120, cwd=ROOT, env={**os.environ, "PYTHONWARNINGS": "ignore"})
    assert r.returncode == 0
    for flag in ("--gpus", "--steps", "--warmup"):
        assert flag in r.stdout
