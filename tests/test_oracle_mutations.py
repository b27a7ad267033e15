"""The oracle's pins reject plausible slips.

oracle/wsim_oracle.cpp carries compile-time mutation hooks: a build with
-DORC_MUTANT=k plants one mistake in one formula.  Each mutant is built
into a temporary directory and the pinning suites (test_oracle_pinning.py,
test_oracle_physics.py) run against it in a child process, which must
fail.  The unmutated oracle passes the same suites in the normal run.

  1  W's outer piece (2 - q)^2 instead of (2 - q)^3       (sph_kernels.cuh:20)
  2  dW/dq's inner piece 2.5 q^2 instead of 2.25 q^2     (sph_kernels.cuh:35)
  3  the stencil's +x column skipped                     (dfsph_kernels.cu:18-20)
  4  alpha's |sum m grad W|^2 term dropped               (dfsph_kernels.cu:166-168)
  5  rho-dot with v_j - v_i                              (dfsph_kernels.cu:209-211)
  6  kappa/rho in place of kappa/rho^2 in form 0         (dfsph_kernels.cu:257-258)
  7  the wall's mirror pressure term dropped             (dfsph_kernels.cu:291-294)
  8  XSPH without the -v_i term                          (dfsph_kernels.cu:333-337)
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
MUTANTS = range(1, 9)
SUITES = ["tests/test_oracle_pinning.py", "tests/test_oracle_physics.py"]


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    d = tmp_path_factory.mktemp("mutants")

    def build(k):
        out = d / f"liboracle_m{k}.so"
        r = subprocess.run(["g++", "-std=c++17", "-O1", "-fopenmp", "-ffp-contract=off",
                            "-fno-fast-math", "-fPIC", "-shared", f"-DORC_MUTANT={k}",
                            str(ROOT / "oracle" / "wsim_oracle.cpp"), "-o", str(out)],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-2000:]
        return k, out

    with ThreadPoolExecutor(max_workers=4) as ex:
        return dict(ex.map(build, MUTANTS))


@pytest.mark.parametrize("k", list(MUTANTS))
def test_pins_reject_mutant(k, mutant_libs):
    env = dict(os.environ, WSIM_ORACLE_LIB=str(mutant_libs[k]), OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        "-m", "not gpu", *SUITES], cwd=ROOT, env=env, capture_output=True,
                       text=True, timeout=600)
    assert r.returncode != 0, f"mutant {k} passed every pin:\n{r.stdout[-1500:]}"
    failed = [l for l in r.stdout.splitlines() if l.startswith("FAILED")]
    assert failed, r.stdout[-1500:]
    print(f"\nmutant {k}: rejected by {failed[0].split(' ')[1]}")
