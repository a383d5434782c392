"""User-defined models (CustomDynamics): tracing into C++, the host (torch)
and oracle (numpy) evaluations of the same function, the custom builds."""
import os

import numpy as np
import pytest
import torch

from diffmpc_amd import SpacecraftDynamics
from diffmpc_amd.dynamics import examples
from diffmpc_amd.dynamics.custom_dynamics import CustomDynamics, cpp_model_source
from oracle import dynamics as odyn
from tests.helpers import OracleCustomModel


def test_unicycle_host_and_oracle_evaluations_follow_the_formula():
    dyn = examples.unicycle()
    rng = np.random.default_rng(0)
    x, u, k = rng.normal(size=(5, 3)), rng.normal(size=(5, 2)), rng.uniform(0.5, 2, (5, 1))
    want = np.stack([u[:, 0] * np.cos(x[:, 2]), u[:, 0] * np.sin(x[:, 2]), k[:, 0] * u[:, 1]], -1)
    host = dyn.state_dot(torch.tensor(x), torch.tensor(u), {"turn_gain": torch.tensor(k)}).numpy()
    assert np.allclose(host, want, rtol=1e-15, atol=0)
    assert np.allclose(OracleCustomModel(dyn).state_dot(x, u, k), want, rtol=1e-15, atol=0)


def test_custom_rate_spacecraft_matches_the_built_in_model_on_the_host():
    dyn = examples.rate_spacecraft()
    ref = SpacecraftDynamics()
    rng = np.random.default_rng(1)
    w, tau, J = rng.normal(size=(7, 3)), rng.normal(size=(7, 3)), rng.uniform(1, 5, 3)
    a = dyn.state_dot(torch.tensor(w), torch.tensor(tau), {"inertia_vector": torch.tensor(J)}).numpy()
    b = ref.state_dot(torch.tensor(w), torch.tensor(tau), {"inertia_vector": torch.tensor(J)}).numpy()
    assert np.allclose(a, b, rtol=1e-14, atol=1e-15)
    o = odyn.make_model("spacecraft3").state_dot(w, tau, np.broadcast_to(J, (7, 3)))
    assert np.allclose(OracleCustomModel(dyn).state_dot(w, tau, np.broadcast_to(J, (7, 3))), o, rtol=1e-14)


def test_traced_source_shares_subexpressions_and_declares_every_output():
    def f(x, u, p, m):
        s = m.sin(x[0])
        return [s * s + 2.0, -s / (1.0 + p[0]) + x[1] ** 3]

    src = cpp_model_source(f, 2, 1, 1, "t")
    assert "static constexpr int NX = 2, NU = 1, NPARAM = 1;" in src
    assert src.count("dsin(x[0])") == 1  # the common subexpression is emitted once
    assert "xd[0] =" in src and "xd[1] =" in src
    assert "sc<T>(S(2.0))" in src and "const T P0 = sc<T>(p[0]);" in src


def test_tracer_rejects_what_the_device_code_cannot_express():
    with pytest.raises(TypeError):
        cpp_model_source(lambda x, u, p, m: [x[0] ** 1.5], 1, 1, 1)
    with pytest.raises(ValueError):
        cpp_model_source(lambda x, u, p, m: [x[0], x[0]], 1, 1, 1)
    with pytest.raises(TypeError):
        cpp_model_source(lambda x, u, p, m: ["a"], 1, 1, 1)


def test_models_without_parameters_get_a_placeholder_slot():
    dyn = CustomDynamics(1, 1, lambda x, u, p, m: [u[0] - x[0]], name="lag")
    assert dyn.num_params == 1 and dyn.param_keys == ()
    flat = dyn.flat_params({}, 3)
    assert flat.shape == (3, 1) and float(flat.abs().sum()) == 0.0


@pytest.mark.parametrize("make", examples.EXAMPLES, ids=lambda m: m.__name__)
def test_prebuilt_custom_library_exports_the_abi_and_reports_its_model(make):
    import ctypes

    from diffmpc_amd import _lib

    dyn = make()
    path = os.path.join(dyn.build_dir(), "libdmpc_custom.so")
    if not os.path.exists(path):
        pytest.skip("custom example library not built (run __graft_entry__.build())")
    lib = _lib.load(path)
    nx, nu, npar = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    assert lib.dmpc_custom_model_info(ctypes.byref(nx), ctypes.byref(nu), ctypes.byref(npar)) == 0
    assert (nx.value, nu.value, npar.value) == (dyn.num_states, dyn.num_controls, dyn.num_params)
    assert _lib.load().dmpc_custom_model_info(ctypes.byref(nx), ctypes.byref(nu), ctypes.byref(npar)) == -2
