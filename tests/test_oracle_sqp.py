"""Oracle SQP and adjoint: closed forms, finite differences and the values the
reference's notebooks recorded (tests/golden/notebook_golden.json)."""
import json
import os

import numpy as np
import pytest

from oracle import dynamics as odyn
from oracle import rollout, sqp
from diffmpc_amd import workloads as W
from tests.helpers import dense_kkt, notebook_spacecraft, oracle_problem

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "notebook_golden.json")))


def test_lti_single_iteration_is_exact_qp_solution():
    name, prob, sp = W.lti(batch=3, seed=0)
    prob = oracle_problem(name, prob)
    sp["pcg"]["tol_epsilon"] = 1e-26
    sp["pcg"]["max_iter"] = 1000
    sol = sqp.sqp_solve(prob, sp)
    sv = sol["saved"]
    for b in range(3):
        x, u, lam, _, _ = dense_kkt(sv["As_next"][b], sv["As"][b], sv["Bs"][b], sv["Qinv"][b], sv["qvec"][b],
                                    sv["Rinv"][b], sv["rvec"][b], sv["Cs"][b])
        assert np.allclose(sol["states"][b], x, atol=1e-9)
        assert np.allclose(sol["controls"][b, :-1], u, atol=1e-9)
        assert np.allclose(sol["dual"][b], lam, atol=1e-8)
    assert np.all(sol["num_iter"] == 1)
    # status: one iteration leaves err > tol_convergence (sqp.py:535-544)
    assert np.all(sol["status"] == sqp.ERROR)


def test_cost_terms_match_cost_derivatives():
    name, prob, sp = W.quadrotor(batch=2, seed=3, horizon=6)
    prob = oracle_problem(name, prob)
    x, u, _ = sqp.initial_guess(prob)
    x = x + np.random.default_rng(0).normal(size=x.shape)
    Qs, qs, Rs, rs = sqp.cost_terms(prob, x, u)
    # the quadratic model reproduces the cost up to a constant: c(z) - c(0) = 1/2 z^T H z + q^T z
    z = np.random.default_rng(1).normal(size=x.shape), np.random.default_rng(2).normal(size=u.shape)
    N = prob["horizon"]
    zx, zu = z
    zu[:, N] = 0.0
    quad = 0.5 * np.einsum("bti,btij,btj->b", zx, Qs, zx) + np.einsum("bti,bti->b", qs, zx)
    quad += 0.5 * np.einsum("bti,btij,btj->b", zu[:, :N], Rs, zu[:, :N]) + np.einsum("bti,bti->b", rs, zu[:, :N])
    c0 = sqp.cost(prob, np.zeros_like(x), np.zeros_like(u))
    assert np.allclose(sqp.cost(prob, zx, zu) - c0, quad, rtol=1e-10)


def test_select_alpha_follows_clamped_indexing():
    a = [0.01, 0.1, 0.3, 0.7, 1.0]
    assert sqp.select_alpha(np.array([1, 1, 1, 1, -1.0]), a) == 1.0
    assert sqp.select_alpha(np.array([1, 1, 1, -1, 1.0]), a) == 0.7
    assert sqp.select_alpha(np.array([1, -1, 1, 1, 1.0]), a) == 0.1
    assert sqp.select_alpha(np.array([-1, 1, 1, 1, 1.0]), a) == 0.01  # index 0 is never tested
    assert sqp.select_alpha(np.array([1, 1, 1, 1, 1.0]), a) == 0.01
    assert sqp.select_alpha(np.array([1.0]), [1.0]) == 1.0
    assert sqp.select_alpha(np.array([-1.0]), [1.0]) == 1.0


KEYS = ["w_x", "w_u", "w_f", "x0", "x_ref", "u_ref", "x_final"]


@pytest.mark.parametrize("penalize", [False, True])
def test_backward_matches_central_differences_on_linear_problem(penalize):
    """The adjoint of a linear-quadratic problem is exact: central FD of the solve."""
    name, prob, sp = W.lti(batch=2, seed=4, horizon=8)
    prob = oracle_problem(name, prob)
    rng = np.random.default_rng(0)
    prob["penalize_control_reference"] = penalize
    prob["x_ref"] = rng.normal(size=prob["x_ref"].shape)
    prob["u_ref"] = rng.normal(size=prob["u_ref"].shape)
    prob["x_final"] = rng.normal(size=prob["x_final"].shape)
    prob["w_f"] = rng.uniform(0.5, 2, size=prob["w_f"].shape)
    sp["pcg"]["tol_epsilon"] = 1e-28
    sp["pcg"]["max_iter"] = 1000
    cx = rng.normal(size=(2, 9, 4))
    cu = rng.normal(size=(2, 9, 2))

    def loss(p):
        s = sqp.sqp_solve(p, sp)
        return np.sum(cx * s["states"], axis=(1, 2)) + np.sum(cu * s["controls"], axis=(1, 2)), s

    _, sol = loss(prob)
    g = sqp.sqp_backward(prob, sp, sol, cx, cu)
    h = 1e-6
    for k in KEYS:
        arr = prob[k]
        num = np.zeros_like(arr)
        it = np.nditer(arr[0], flags=["multi_index"])
        for _ in it:
            idx = (slice(None),) + it.multi_index
            pp, pm = dict(prob), dict(prob)
            pp[k], pm[k] = arr.copy(), arr.copy()
            pp[k][idx] += h
            pm[k][idx] -= h
            num[idx] = (loss(pp)[0] - loss(pm)[0]) / (2 * h)
        assert np.allclose(g[k], num, rtol=1e-5, atol=1e-6), k


def test_notebook_spacecraft_first_loss_is_reproduced_exactly():
    inertia, x_init, prob, params = notebook_spacecraft()
    prob = oracle_problem("spacecraft3", prob)
    reward = lambda x, u: -(10 * np.sum(x ** 2, 1) + np.sum(u ** 2, 1))  # noqa: E731
    total, _ = rollout.rollout(prob, params, x_init, 20, 0, reward)
    assert abs(total.sum() - GOLDEN["spacecraft"]["losses"][0]) <= 1e-13 * abs(GOLDEN["spacecraft"]["losses"][0])


def test_notebook_spacecraft_training_trace():
    """notebooks/RL_spacecraft.ipynb cells 12 and 14: ten gradient steps, lr = 1.

    The loss of step 1 pins the forward; the weights after step 1 and the later
    losses pin the adjoint to the level of the adjoint PCG truncation
    (eps = 1e-12 absolute on eta): warm and cold adjoint starts already move
    the gradient by ~1e-5.
    """
    inertia, x_init, prob, params = notebook_spacecraft()
    prob = oracle_problem("spacecraft3", prob)
    reward = lambda x, u: -(10 * np.sum(x ** 2, 1) + np.sum(u ** 2, 1))  # noqa: E731
    reward_grad = lambda x, u: (-20 * x, -2 * u)  # noqa: E731
    wx, wu = np.ones(3), np.ones(3)
    ref = GOLDEN["spacecraft"]
    for step in range(10):
        prob["w_x"] = np.repeat(wx[None], 2, 0)
        prob["w_u"] = np.repeat(wu[None], 2, 0)
        total, rec = rollout.rollout(prob, params, x_init, 20, 0, reward)
        assert abs(total.sum() - ref["losses"][step]) <= 2e-6 * abs(ref["losses"][step]), step
        g = rollout.rollout_gradient(prob, params, rec, 0, reward_grad, keys=("w_x", "w_u"))
        wx = np.maximum(1e-10, wx + np.nan_to_num(g["w_x"]))
        wu = np.maximum(1e-10, wu + np.nan_to_num(g["w_u"]))
        if step == 0:
            assert np.allclose(wx, ref["w_x_after_step1"], atol=2e-5, rtol=0)
            assert np.allclose(wu, ref["w_u_after_step1"], atol=2e-5, rtol=0)
