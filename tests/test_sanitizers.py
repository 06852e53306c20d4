"""The product's host logic (tridiagonal / arrowhead eigensolve, group
closure and inverses, colex ranks of the spread of x) and the oracle's C
restatement, built with AddressSanitizer and UndefinedBehaviorSanitizer
(g++ / gcc, host code only) and run on small inputs (no GPU)."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("gcc") is None, reason="no host compiler")
def test_host_logic_and_oracle_sanitized(tmp_path):
    obj = str(tmp_path / "oracle.o")
    exe = str(tmp_path / "sanitize_host")
    subprocess.run(["gcc", "-std=c11", "-fopenmp", "-c", os.path.join(ROOT, "oracle", "oracle.c"), "-o", obj] + FLAGS,
                   check=True, capture_output=True, text=True)
    r = subprocess.run(["g++", "-std=c++17", "-fopenmp", os.path.join(ROOT, "tests", "native", "sanitize_host.cpp"),
                        obj, "-o", exe, "-lm"] + FLAGS, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert run.returncode == 0, (run.stdout[-2000:], run.stderr[-4000:])
    assert "sanitized host checks passed" in run.stdout
