"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer (host code only): the
oracle's own CPU pins (transposed sweeps, positivity, Riemann pins, ModelEvaluator, SOR,
second-order substitute, implicit and BDF steps, Roe-Jacobian blocks, viscous terms,
triangles and tetrahedra) are run again in a child process against `make -C oracle sanitize`,
with the sanitizer runtimes preloaded and every report fatal.  Skipped when the host compiler
has no sanitizer runtimes."""
from __future__ import annotations

import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
FILES = ["test_oracle_adjoint_sweep.py", "test_oracle_positivity.py", "test_oracle_riemann_pins.py",
         "test_oracle_model.py", "test_oracle_sor.py", "test_oracle_2nd.py", "test_oracle_implicit.py",
         "test_oracle_bdf.py", "test_oracle_roe_jacobian.py", "test_oracle_viscous.py", "test_oracle_simplex.py"]


def _runtime(name):
    try:
        p = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True, check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        return None
    return p if os.path.isabs(p) and os.path.exists(p) else None


def test_oracle_pins_clean_under_sanitizers():
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not (asan and ubsan):
        pytest.skip("the host compiler has no AddressSanitizer / UndefinedBehaviorSanitizer runtime")
    b = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], capture_output=True, text=True,
                       timeout=600)
    assert b.returncode == 0, b.stdout + b.stderr
    env = dict(os.environ, DNDSR_ORACLE_LIB=os.path.join(ROOT, "oracle", "_san", "liboracle.so"),
               LD_PRELOAD=f"{asan}:{ubsan}", ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", "-m", "not gpu",
                        *[os.path.join(HERE, f) for f in FILES]], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]
