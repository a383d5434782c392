"""GPU parity of gradients with respect to dynamics parameters passed as
weights (masses, inertias, spring constants, A/B/b of the linear model)
against the float64 oracle, through torch autograd and the C ABI
(dmpc_sqp_backward with the adjoint outputs, then
dmpc_dynamics_param_grads).  Tolerance: 1e-4 of the largest gradient entry
of the instance, as for the other gradients."""
import sys

import numpy as np
import pytest
import torch

from diffmpc_amd import workloads as W
from oracle import sqp as osqp
from tests.helpers import oracle_problem

pytestmark = pytest.mark.gpu

CASES = {
    "lti": lambda: W.lti(batch=4, seed=0),
    "quadrotor": lambda: W.quadrotor(batch=6, seed=1, horizon=30),
    "spacecraft7": lambda: W.spacecraft7(batch=4, seed=2, horizon=40),
    "spacecraft3": lambda: W.spacecraft3(batch=3, seed=5),
    "cartpole": lambda: W.cartpole(batch=3, seed=6),
}


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _np(t):
    return t.detach().cpu().numpy()


def _rel_rows(a, b):
    """max over instances of max |a - b| / max |b| (per instance)."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.max(np.abs(a - b), axis=1) / np.maximum(np.max(np.abs(b), axis=1), 1e-30)))


def _gpu_dyn_grads(name, prob, sp, cuda, mode="fused", horizons=None, shared=()):
    """Flat (B, P) d loss / d dynamics parameters from autograd, with the
    loss sum(cx * x) + sum(cu * u); keys in `shared` are passed unbatched
    (their gradient is the batch sum)."""
    solver, params, guess = W.make_solver(name, prob, sp, device=cuda, mode=mode, horizons=horizons)
    keys = solver.program.dynamics.param_keys
    leaves = {k: (params[k][0] if k in shared else params[k]).clone().requires_grad_(True) for k in keys}
    sol = solver.solve(guess, params, leaves)
    rng = np.random.default_rng(11)
    cx = rng.normal(size=tuple(sol.states.shape))
    cu = rng.normal(size=tuple(sol.controls.shape))
    if horizons is not None:
        for b, n in enumerate(horizons):
            cx[b, n + 1:] = 0.0
            cu[b, n + 1:] = 0.0
    loss = (sol.states * torch.tensor(cx, device=cuda)).sum() + (sol.controls * torch.tensor(cu, device=cuda)).sum()
    grads = torch.autograd.grad(loss, [leaves[k] for k in keys])
    torch.cuda.synchronize()
    B = sol.states.shape[0]
    flat = {k: _np(g).reshape(1 if k in shared else B, -1) for k, g in zip(keys, grads)}
    return flat, keys, cx, cu


def _oracle_flat(name, prob, sp, cx, cu):
    op = oracle_problem(name, prob)
    ref = osqp.sqp_solve(op, sp)
    g = osqp.sqp_backward(op, sp, ref, cx, cu)
    return osqp.dynamics_param_gradient(op, ref, g)


@pytest.mark.parametrize("variant", ["register", "streamed", "plain"])
@pytest.mark.parametrize("case", list(CASES))
def test_dynamics_parameter_gradients_match_oracle(cuda, case, variant):
    say(f"[dyn-grad] case={case} variant={variant}")
    name, prob, sp = CASES[case]()
    mode = "plain" if variant == "plain" else "fused"
    if mode == "fused":
        sp["fused_variant"] = variant
    flat, keys, cx, cu = _gpu_dyn_grads(name, prob, sp, cuda, mode=mode)
    got = np.concatenate([flat[k] for k in keys], axis=1)
    want = _oracle_flat(name, prob, sp, cx, cu)
    e = _rel_rows(got, want)
    say(f"  keys={keys} rel {e:.2e}")
    assert got.shape == want.shape
    assert e <= 1e-4


def test_chain_dynamics_parameter_gradients_match_oracle(cuda):
    say("[dyn-grad] chain N=20")
    name, prob, sp = W.chain(batch=2, seed=4, horizon=20, pcg_eps=1e-12)
    flat, keys, cx, cu = _gpu_dyn_grads(name, prob, sp, cuda)
    got = np.concatenate([flat[k] for k in keys], axis=1)
    e = _rel_rows(got, _oracle_flat(name, prob, sp, cx, cu))
    say(f"  rel {e:.2e}")
    assert e <= 1e-4


def test_shared_dynamics_parameters_get_batch_summed_gradient(cuda):
    say("[dyn-grad] quadrotor, shared mass and inertia")
    name, prob, sp = CASES["quadrotor"]()
    prob["dyn"] = np.repeat(prob["dyn"][:1], prob["dyn"].shape[0], 0)  # one vehicle for the batch
    flat, keys, cx, cu = _gpu_dyn_grads(name, prob, sp, cuda, shared=("mass", "inertia"))
    got = np.concatenate([flat[k] for k in keys], axis=1)
    want = _oracle_flat(name, prob, sp, cx, cu).sum(0, keepdims=True)
    e = _rel_rows(got, want)
    say(f"  rel {e:.2e}")
    assert e <= 1e-4


@pytest.mark.parametrize("variant", ["register", "streamed"])
def test_ragged_dynamics_parameter_gradients_match_oracle_per_instance(cuda, variant):
    say(f"[dyn-grad] ragged attitude variant={variant}")
    name, prob, sp = W.spacecraft7(batch=5, seed=8, horizon=24)
    h = [24, 7, 15, 1, 20]
    sp["fused_variant"] = variant
    flat, keys, cx, cu = _gpu_dyn_grads(name, prob, sp, cuda, horizons=h)
    got = np.concatenate([flat[k] for k in keys], axis=1)
    for b, n in enumerate(h):
        one = W.instance_problem(prob, b, n)
        want = _oracle_flat(name, one, sp, cx[b:b + 1, :n + 1], cu[b:b + 1, :n + 1])
        e = _rel_rows(got[b:b + 1], want)
        say(f"  b={b} N={n}: rel {e:.2e}")
        assert e <= 1e-4


def test_batch_params_solves_match_shared_parameter_solves(cuda):
    """SQPSolver.batch_params (shared values expanded once, the model's flat
    parameter block precomputed) gives the same solutions and gradients as
    passing the shared values to every solve, with and without a dynamics
    parameter among the weights."""
    say("[batch-params] quadrotor")
    name, prob, sp = CASES["quadrotor"]()
    solver, params, guess = W.make_solver(name, prob, sp, device=cuda)
    B = prob["x0"].shape[0]
    key = "weights_penalization_reference_state_trajectory"
    base = {k: v for k, v in params.items() if k != key}
    shared_w = params[key][0]
    bp = solver.batch_params({**base, key: shared_w}, B, cuda)
    assert "_flat_dynamics_parameters" in bp

    def run(p, with_mass, batched):
        guess.dual_backward_guess.zero_()  # each backward writes its adjoint warm start back into the guess
        w = shared_w.clone().requires_grad_(True)
        m = params["mass"].clone().requires_grad_(True)
        weights = {key: w}
        if batched:
            weights = solver.batch_params(weights, B, cuda)
        if with_mass:
            weights["mass"] = m
        sol = solver.solve(guess, p, weights)
        loss = (sol.states - params["reference_state_trajectory"]).square().sum()
        leaves = [w, m] if with_mass else [w]
        return sol.states.detach(), torch.autograd.grad(loss, leaves)

    for with_mass in (False, True):
        x1, g1 = run({**base, key: shared_w}, with_mass, False)
        x2, g2 = run(bp, with_mass, True)
        torch.cuda.synchronize()
        assert torch.equal(x1, x2)
        for a, b in zip(g1, g2):
            assert torch.allclose(a, b, rtol=1e-12, atol=0.0)
