"""The C ABI's argument checks and host-side launch planning under
AddressSanitizer (CPU only): diffmpc_amd/csrc/asan/build_check.sh builds the
library with its host code instrumented and runs check_abi.c, which calls
every entry point with arguments that must be refused and the host-only
queries over every built model and a range of horizons; no call reaches the
GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCRIPT = os.path.join(ROOT, "diffmpc_amd", "csrc", "asan", "build_check.sh")


@pytest.mark.skipif(not shutil.which("/opt/rocm/bin/hipcc"), reason="hipcc not available")
def test_abi_argument_checks_are_clean_under_address_sanitizer(tmp_path):
    r = subprocess.run(["bash", SCRIPT, str(tmp_path)], capture_output=True, text=True, timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out, out[-4000:]
    assert "all argument checks returned the expected codes" in out
