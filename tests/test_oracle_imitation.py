"""Pins the oracle's imitation-learning loss and gradient
(oracle/imitation.py, restating train_imitation_diffmpc.py:55-81 and
utils.py:20-31) with checks independent of its formulas: the Euler rollout
of converged controls reproduces the converged Euler-scheme trajectory, the
gradient equals central differences of the loss of converged solves, and at
the expert's own weights loss and gradient vanish."""
import numpy as np
import pytest

from diffmpc_amd import workloads as W
from oracle import dynamics as odyn
from oracle import imitation as oil
from oracle import sqp as osqp
from tests.helpers import oracle_problem

TRUE_W = np.array([1.0, 2.0, 1.5, 1.0])  # cartpole.yaml


def _problem(B=2, N=20, seed=0):
    rng = np.random.default_rng(seed)
    x0 = rng.uniform(-0.3, 0.3, size=(B, 4))
    rep = lambda v: np.repeat(np.asarray(v, np.float64)[None], B, 0)  # noqa: E731
    return dict(model=odyn.make_model("cartpole"), horizon=N, dt=0.05, scheme=0, penalize_control_reference=False,
                x0=x0, x_final=np.zeros((B, 4)), x_ref=np.zeros((B, N + 1, 4)), u_ref=np.zeros((B, N + 1, 1)),
                w_x=rep(TRUE_W), w_u=rep([5e-2]), w_f=np.zeros((B, 4)), dyn=rep([1.0, 0.1, 0.5, 9.81]))


def _converged():
    return W.solver_params(80, 1e-12, 1e-24, linesearch=False)


@pytest.fixture(scope="module")
def expert():
    prob = _problem()
    sol = osqp.sqp_solve(prob, _converged())
    assert np.all(sol["convergence_error"] < 1e-10)
    return prob, sol


def test_euler_rollout_of_converged_controls_reproduces_the_trajectory(expert):
    prob, sol = expert
    xs = oil.euler_rollout(prob, sol["controls"])
    assert np.max(np.abs(xs - sol["states"])) < 1e-9


def test_gradient_matches_central_differences_of_the_loss_on_a_linear_problem():
    """On linear dynamics the QP of the last linearization is the problem
    itself, so the implicit gradient is the derivative of the loss of the
    converged solve (on the nonlinear cart-pole the reference's Gauss-Newton
    backward omits the dynamics curvature and central differences differ by
    design)."""
    name, lp, _ = W.lti(batch=3, seed=7, horizon=20)
    prob = oracle_problem(name, lp)
    params = W.solver_params(3, 1e-12, 1e-26, pcg_max=2000)
    us = osqp.sqp_solve(prob, params)["controls"]  # demonstrations at w_x = 1
    w = np.array([2.31993711, 2.10973335, 1.27643918, 2.40638906])  # first starting point of the benchmark
    loss, g = oil.il_loss_and_grad(prob, params, us, w)
    assert loss > 1e-6
    h = 1e-5
    fd = np.zeros(4)
    for i in range(4):
        e = np.zeros(4)
        e[i] = h
        fd[i] = (oil.il_loss(prob, params, us, w + e)[0] - oil.il_loss(prob, params, us, w - e)[0]) / (2 * h)
    assert np.max(np.abs(g - fd)) <= 1e-5 * np.max(np.abs(fd))


def test_loss_and_gradient_vanish_at_the_expert_weights(expert):
    prob, sol = expert
    loss, g = oil.il_loss_and_grad(prob, _converged(), sol["controls"], TRUE_W)
    assert loss < 1e-20
    assert np.max(np.abs(g)) < 1e-10
