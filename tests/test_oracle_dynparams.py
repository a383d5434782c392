"""Oracle gradients with respect to dynamics parameters passed as weights
(sqp.py:605-660, 714-737), pinned by central differences:

* on dynamics that are linear in (x, u) -- the LTI model (A, B, b) and the
  spring-damper chain (k, c) -- the KKT system of the QP is exact, so the
  implicit gradient is the derivative of the whole solve;
* on nonlinear models the reference's backward differentiates the KKT system
  of the last QP (cost Hessian, linearized dynamics), so its gradient is the
  derivative of that QP's solution with the linearization point held at the
  converged solution and the dynamics parameters entering A_t, B_t and C_t.
"""
import numpy as np
import pytest

from diffmpc_amd import workloads as W
from oracle import sqp
from tests.helpers import oracle_problem


def _fd_full_solve(prob, sp, loss, j_list, h):
    out = {}
    for j in j_list:
        hi, lo = dict(prob), dict(prob)
        hi["dyn"], lo["dyn"] = prob["dyn"].copy(), prob["dyn"].copy()
        hi["dyn"][:, j] += h
        lo["dyn"][:, j] -= h
        out[j] = (loss(hi) - loss(lo)) / (2 * h)
    return out


def test_lti_dynamics_parameter_gradient_is_the_derivative_of_the_solve():
    name, prob, sp = W.lti(batch=2, seed=4, horizon=8)
    prob = oracle_problem(name, prob)
    rng = np.random.default_rng(1)
    prob["x_ref"] = rng.normal(size=prob["x_ref"].shape)
    prob["x_final"] = rng.normal(size=prob["x_final"].shape)
    prob["dyn"] = prob["dyn"] + 0.05 * rng.normal(size=prob["dyn"].shape)  # b != 0, instance-specific A, B
    sp["pcg"]["tol_epsilon"] = 1e-28
    sp["pcg"]["max_iter"] = 1000
    cx = rng.normal(size=prob["x_ref"].shape)
    cu = rng.normal(size=prob["u_ref"].shape)

    def loss(p):
        s = sqp.sqp_solve(p, sp)
        return np.sum(cx * s["states"], axis=(1, 2)) + np.sum(cu * s["controls"], axis=(1, 2))

    sol = sqp.sqp_solve(prob, sp)
    g = sqp.sqp_backward(prob, sp, sol, cx, cu)
    gd = sqp.dynamics_param_gradient(prob, sol, g)
    P = prob["dyn"].shape[1]
    assert gd.shape == (2, P) == (2, 4 * 4 + 4 * 2 + 4)
    num = _fd_full_solve(prob, sp, loss, range(P), 1e-6)
    for j in range(P):
        assert np.allclose(gd[:, j], num[j], rtol=1e-6, atol=1e-7), (j, gd[:, j], num[j])


def test_chain_spring_and_damper_gradient_is_the_derivative_of_the_solve():
    name, prob, sp = W.chain(batch=1, seed=4, horizon=6, pcg_eps=1e-28)
    prob = oracle_problem(name, prob)
    sp["pcg"]["max_iter"] = 2000
    rng = np.random.default_rng(2)
    cx = rng.normal(size=prob["x_ref"].shape)

    def loss(p):
        return np.sum(cx * sqp.sqp_solve(p, sp)["states"], axis=(1, 2))

    sol = sqp.sqp_solve(prob, sp)
    g = sqp.sqp_backward(prob, sp, sol, cx, np.zeros_like(sol["controls"]))
    gd = sqp.dynamics_param_gradient(prob, sol, g)
    picks = [0, 5, 11, 12, 17, 23]  # springs and dampers along the chain
    num = _fd_full_solve(prob, sp, loss, picks, 1e-6)
    scale = np.max(np.abs(gd))
    for j in picks:
        assert abs(gd[0, j] - num[j][0]) <= 1e-6 * scale, (j, gd[0, j], num[j][0])


def _converged(name, prob, sp):
    prob = oracle_problem(name, prob)
    sp["num_scp_iteration_max"] = 60
    sp["tol_convergence"] = 1e-13
    sp["pcg"]["tol_epsilon"] = 1e-28
    sp["pcg"]["max_iter"] = 4000
    s = sqp.sqp_solve(prob, sp)
    assert np.all(s["convergence_error"] < 1e-10), s["convergence_error"]
    return prob, sp, s


NONLINEAR = [
    ("quadrotor", lambda: W.quadrotor(batch=2, seed=3, horizon=10), range(10)),
    ("spacecraft7", lambda: W.spacecraft7(batch=2, seed=3, horizon=12), range(3)),
    ("cartpole", lambda: W.cartpole(batch=2, seed=3, horizon=12), range(4)),
    ("spacecraft3", lambda: W.spacecraft3(batch=2, seed=3, horizon=12), range(3)),
]


@pytest.mark.parametrize("name,make,params", NONLINEAR, ids=[c[0] for c in NONLINEAR])
def test_nonlinear_dynamics_parameter_gradient_is_the_converged_qp_sensitivity(name, make, params):
    n, prob, sp = make()
    assert n == name
    prob, sp, s = _converged(name, prob, sp)
    xbar, ubar, lam = s["states"], s["controls"], s["dual"]
    rng = np.random.default_rng(7)
    cu = rng.normal(size=ubar.shape)
    cu[:, -1] = 0.0

    def loss_of(x, u):
        return 0.5 * np.sum((x - prob["x_ref"]) ** 2, axis=(1, 2)) + np.sum(cu * u, axis=(1, 2))

    g = sqp.sqp_backward(prob, sp, s, xbar - prob["x_ref"], cu)
    gd = sqp.dynamics_param_gradient(prob, s, g)

    def qp_loss(p):
        x, u, _, _, _, _ = sqp.solve_qp(sqp.qp_parameters(p, xbar, ubar), lam, sp["pcg"])
        return loss_of(x, u)

    for j in params:
        h = 1e-6 * max(1.0, float(np.max(np.abs(prob["dyn"][:, j]))))
        num = _fd_full_solve(prob, sp, qp_loss, [j], h)[j]
        scale = np.max(np.abs(gd))
        assert np.allclose(gd[:, j], num, rtol=0, atol=2e-6 * scale), (name, j, gd[:, j], num)
