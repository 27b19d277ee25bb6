"""The CPU suite again, against the host code and the oracle built with
AddressSanitizer and UndefinedBehaviorSanitizer (libfwi_hip_san.so,
oracle/liboracle_san_*.so; every finding aborts the run).  The sanitizer
runtime is preloaded into a child interpreter; the device code is not
instrumented and no GPU call is made."""
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "fwi2dacoustic-mpi-lbfgs_amd")
LIB_SAN = os.path.join(PKG, "libfwi_hip_san.so")
ORC_SAN = [os.path.join(ROOT, "oracle", f"liboracle_san_{p}.so") for p in ("f32", "f64")]


@pytest.mark.skipif(os.environ.get("FWI_SANITIZED_RUN") == "1", reason="already inside the sanitized run")
def test_cpu_suite_under_address_and_undefined_sanitizers():
    sys.path.insert(0, PKG)
    import build as B
    if not all(os.path.exists(p) for p in [LIB_SAN, *ORC_SAN]):
        pytest.skip("sanitized libraries not built (fwi2dacoustic-mpi-lbfgs_amd/build.py; make -C oracle san)")
    rt = B.sanitizer_runtime()
    if rt is None:
        pytest.skip("no shared sanitizer runtime")
    env = dict(os.environ)
    env.update(FWI_SANITIZED_RUN="1", FWI_HIP_LIBRARY=LIB_SAN, FWI_ORACLE_VARIANT="san",
               ASAN_OPTIONS="detect_leaks=0:halt_on_error=1", UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               LD_PRELOAD=rt + (":" + env["LD_PRELOAD"] if env.get("LD_PRELOAD") else ""))
    env.pop("PYTEST_XDIST_WORKER", None)
    cmd = [sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider", "-p", "no:xdist",
           os.path.join(ROOT, "tests")]
    p = subprocess.run(cmd, env=env, cwd=ROOT, capture_output=True, text=True, timeout=1500)
    out = p.stdout + p.stderr
    print(out[-3000:])
    assert "AddressSanitizer" not in out and "runtime error:" not in out
    assert p.returncode == 0, out[-3000:]
    m = re.search(r"(\d+) passed", p.stdout)
    assert m and int(m.group(1)) >= 200, p.stdout[-2000:]    # the suite ran, the library check below among it


@pytest.mark.skipif(os.environ.get("FWI_SANITIZED_RUN") != "1", reason="runs inside the sanitized run only")
def test_inside_the_sanitized_run_the_instrumented_libraries_are_loaded():
    import numpy as np
    from fwihip import lib as L
    from oracle import Oracle
    assert L.load()._name == LIB_SAN
    assert Oracle(np.float64).lib._name.endswith("liboracle_san_f64.so")
    with open("/proc/self/maps") as f:
        assert "libclang_rt.asan" in f.read()
