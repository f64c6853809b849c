"""AddressSanitizer and UndefinedBehaviorSanitizer on the CPU: the library's host
C++ (hex block meshes, polyMesh renumbering) and the oracle run under both
sanitizers (tests/sanitize/host_check.cpp), and the oracle's own tests run
against the oracle's sanitizer build.  Host code only: GPU code is not built
with sanitizers."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SAN_ENV = {"ASAN_OPTIONS": "halt_on_error=1:detect_leaks=1",
           "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1", "OMP_NUM_THREADS": "4"}


def _runtime(name):
    p = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True)
    path = p.stdout.strip()
    if p.returncode != 0 or not os.path.isabs(path):
        pytest.skip(f"{name} not available")
    return path


def test_host_code_and_oracle_under_sanitizers():
    _runtime("libasan.so")
    b = subprocess.run(["make", "-C", str(ROOT / "tests" / "sanitize")], capture_output=True,
                       text=True)
    assert b.returncode == 0, b.stderr[-2000:]
    r = subprocess.run([str(ROOT / "tests" / "sanitize" / "_build" / "host_check")],
                       capture_output=True, text=True, timeout=600,
                       env={**os.environ, **SAN_ENV})
    out = r.stdout + r.stderr
    assert r.returncode == 0 and "host_check: ok" in out, out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]


def test_oracle_tests_against_sanitized_oracle():
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    b = subprocess.run(["make", "-C", str(ROOT / "oracle"), "sanitize"], capture_output=True,
                       text=True)
    assert b.returncode == 0, b.stderr[-2000:]
    so = ROOT / "oracle" / "_build" / "liboracle_san.so"
    preload = ":".join(x for x in (asan, ubsan, os.environ.get("LD_PRELOAD", "")) if x)
    env = {**os.environ, **SAN_ENV, "ASAN_OPTIONS": "halt_on_error=1:detect_leaks=0",
           "LD_PRELOAD": preload, "FV_ORACLE_SO": str(so)}
    probe = ("import oracle_cpu as o, ctypes; "
             "assert o.ORACLE_SO.name == 'liboracle_san.so'; "
             "assert hasattr(ctypes.CDLL(None), '__asan_init'); print('sanitized oracle loaded')")
    r = subprocess.run([sys.executable, "-c", probe], capture_output=True, text=True, env=env,
                       cwd=str(ROOT / "tests"), timeout=300)
    assert r.returncode == 0 and "sanitized oracle loaded" in r.stdout, r.stderr[-2000:]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        "-m", "not gpu", str(ROOT / "tests" / "test_oracle_cpu.py")],
                       capture_output=True, text=True, env=env, cwd=str(ROOT), timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]
