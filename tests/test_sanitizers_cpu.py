"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the host code, run
on the CPU:

* the oracle (oracle/rpp_oracle.c) driven over every entry point with exactly
  sized operands, random windows, sparse matrices and ragged signature
  batches (tests/native/oracle_san_check.c, `make -C oracle san`);
* the library's host validation and Hall basis metadata code
  (csrc/*.cpp, instrumented on the host pass only) driven with NULL and
  random descriptors of every C-ABI entry point and with damaged metadata
  files (tests/native/capi_san_check.cpp, `make -C rough-path-primitives_amd/csrc san`).

Any sanitizer report aborts the driver with a non-zero status."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "_build", "san")
ENV = dict(os.environ, ASAN_OPTIONS="abort_on_error=0:detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")


def _make(directory):
    r = subprocess.run(["make", "-C", directory, "-j8", "san"], capture_output=True, text=True, timeout=1200)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]


def test_oracle_under_sanitizers():
    if shutil.which("gcc") is None:
        pytest.skip("no host C compiler")
    _make(os.path.join(ROOT, "oracle"))
    r = subprocess.run([os.path.join(SAN, "oracle_san_check"), "300"], capture_output=True, text=True,
                       timeout=600, env=ENV)
    assert r.returncode == 0 and "ORACLE SAN OK" in r.stdout, r.stdout[-3000:] + r.stderr[-6000:]


def test_capi_host_code_under_sanitizers(tmp_path):
    if not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("no hipcc")
    _make(os.path.join(ROOT, "rough-path-primitives_amd", "csrc"))
    # the HIP runtime keeps allocations until exit: leak reports would name it
    env = dict(ENV, ASAN_OPTIONS="abort_on_error=0:detect_leaks=0")
    r = subprocess.run([os.path.join(SAN, "capi_san_check"), str(tmp_path), "300"], capture_output=True, text=True,
                       timeout=900, env=env)
    if "CAPI SAN SKIPPED" in r.stdout:
        pytest.skip(r.stdout.strip())
    assert r.returncode == 0 and "CAPI SAN OK" in r.stdout, r.stdout[-3000:] + r.stderr[-6000:]
