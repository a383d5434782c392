"""The C ABI library loads and exports what include/dmpc.h declares (no GPU calls)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dmpc.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\**(dmpc_\w+)\s*\(", text, re.M)))


@pytest.fixture(scope="module")
def lib():
    from diffmpc_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        subprocess.run(["bash", os.path.join(ROOT, "diffmpc_amd", "csrc", "build.sh")], check=True)
    return _lib.load()


def test_header_declares_the_abi():
    names = _declared()
    assert "dmpc_sqp_solve" in names and "dmpc_sqp_backward" in names
    assert len(names) >= 10


def test_library_exports_every_declared_symbol(lib):
    from diffmpc_amd import _lib

    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (dmpc_\w+)", out))
    for name in _declared():
        assert name in exported, name
        assert name in _lib.exported_symbols(), name


def test_version_and_support_table(lib):
    assert lib.dmpc_version().startswith(b"dmpc")
    assert lib.dmpc_supported(0, 1, 13, 4, 50) == 1      # fused quadrotor N=50
    assert lib.dmpc_supported(0, 3, 7, 3, 100) == 1      # fused spacecraft7 N=100
    assert lib.dmpc_supported(1, 5, 48, 12, 256) == 1    # plain chain
    assert lib.dmpc_supported(0, 1, 12, 4, 50) == 0      # wrong nx for the model
    assert lib.dmpc_supported(7, 1, 13, 4, 50) == 0      # unknown mode


def test_argument_errors_return_codes_without_touching_the_gpu(lib):
    from diffmpc_amd import _lib

    assert lib.dmpc_sqp_solve(None, None, None, None, None, 0, None) == _lib.DMPC_ERR_ARG
    p = _lib.Problem(model=1, nx=13, nu=4, horizon=0, batch=1, scheme=2, n_params=10, penalize_control_reference=0,
                     dt=0.02)
    assert lib.dmpc_sqp_solve(ctypes.byref(p), None, None, None, None, 0, None) == _lib.DMPC_ERR_ARG
    p.horizon = 10
    p.model = 9
    assert lib.dmpc_sqp_solve(ctypes.byref(p), None, None, None, None, 0, None) == _lib.DMPC_ERR_UNSUPPORTED
    assert lib.dmpc_set_blocks_per_group(3) == _lib.DMPC_ERR_ARG
    assert lib.dmpc_set_blocks_per_group(0) == _lib.DMPC_OK
    assert lib.dmpc_pcg_solve(0, 1, 1, None, None, None, None, 1, 0.0, None, None, None) == _lib.DMPC_ERR_ARG
