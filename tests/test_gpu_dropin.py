"""Drop-in: the reference's own free functions (rpp::ops::ft_mul, ...)
instantiated with rpp::hip::strategies::BlockStrategy run on the GPU and
match the reference cpu::single_thread on identical inputs.  The check binary
is compiled against the reference headers (oracle/Makefile, target dropin)
where they exist; it reads nothing of the reference at run time."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "dropin_check")

pytestmark = pytest.mark.gpu


@pytest.mark.skipif(not os.path.exists(BIN), reason="drop-in check not built (needs the reference headers)")
def test_reference_free_functions_with_hip_strategy():
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "DROPIN OK" in r.stdout
