"""The hip::block strategy header compiles against the reference's headers
(syntax check only; no GPU needed)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/include"


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference headers not present on this machine")
def test_strategy_header_compiles_against_reference():
    src = os.path.join(ROOT, "tests", "dropin", "dropin_check.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I" + REF, "-I" + os.path.join(ROOT, "include"),
                        "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", src],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
