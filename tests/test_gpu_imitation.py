"""GPU parity of the imitation-learning pipeline (diffmpc_amd.imitation,
dmpc_simulate) against the float64 oracle: the open-loop rollout kernel
against explicit integration steps, the imitation loss and its gradient on
the cart-pole benchmark's own settings (N = 100, <= 5 SQP iterations,
tolerance 0.1, PCG 1e-14, line search), the hipGraph-captured training step
against the eager one, and the training loop's bookkeeping.
Tolerances: states 1e-5 (relative to the instance's largest entry); loss and
gradient 1e-4 (relative), as for the other gradients."""
import sys

import numpy as np
import pytest
import torch

from diffmpc_amd import imitation as il
from diffmpc_amd import workloads as W
from oracle import dynamics as odyn
from oracle import imitation as oil
from tests.helpers import oracle_problem

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _np(t):
    return t.detach().cpu().numpy()


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    ax = tuple(range(1, a.ndim))
    return float(np.max(np.max(np.abs(a - b), axis=ax) / np.maximum(np.max(np.abs(b), axis=ax), 1e-30)))


def _oracle_rollout(prob, us, scheme, horizons=None):
    model = prob["model"]
    B, T = us.shape[0], us.shape[1]
    xs = np.zeros((B, T, model.nx))
    xs[:, 0] = prob["x0"]
    for t in range(T - 1):
        xs[:, t + 1] = odyn.predict_next_state(model, prob["dt"], scheme, prob["dyn"], xs[:, t], us[:, t])
    if horizons is not None:
        for b, n in enumerate(horizons):
            xs[b, n + 1:] = 0.0
    return xs


SIM_CASES = {
    "cartpole-euler": (lambda: W.cartpole(batch=5, seed=3, horizon=40), 0),
    "cartpole-midpoint": (lambda: W.cartpole(batch=5, seed=3, horizon=40), 1),
    "cartpole-rk4": (lambda: W.cartpole(batch=5, seed=3, horizon=40), 2),
    "quadrotor-rk4": (lambda: W.quadrotor(batch=7, seed=1, horizon=30), 2),
    "chain-euler": (lambda: W.chain(batch=3, seed=4, horizon=16), 0),
}


@pytest.mark.parametrize("case", list(SIM_CASES))
def test_simulate_matches_oracle_integration(cuda, case):
    say(f"[simulate] {case}")
    make, scheme = SIM_CASES[case]
    name, prob, sp = make()
    solver, params, _ = W.make_solver(name, prob, sp, device=cuda)
    rng = np.random.default_rng(5)
    us = 0.3 * rng.normal(size=tuple(prob["u_ref"].shape))
    us[:, -1] = 0.0
    got = _np(solver.program.simulate(torch.tensor(us, device=cuda), params, scheme=scheme))
    want = _oracle_rollout(oracle_problem(name, prob), us, scheme)
    e = rel(got, want)
    say(f"  rel {e:.2e}")
    assert e <= 1e-5


def test_simulate_zeroes_knots_past_ragged_horizons(cuda):
    say("[simulate] ragged cartpole")
    name, prob, sp = W.cartpole(batch=4, seed=2, horizon=25)
    h = [25, 3, 11, 1]
    solver, params, _ = W.make_solver(name, prob, sp, device=cuda, horizons=h)
    rng = np.random.default_rng(6)
    us = 0.3 * rng.normal(size=tuple(prob["u_ref"].shape))
    got = _np(solver.program.simulate(torch.tensor(us, device=cuda), params))
    want = _oracle_rollout(oracle_problem(name, prob), us, prob["scheme"], horizons=h)
    assert rel(got, want) <= 1e-5
    for b, n in enumerate(h):
        assert np.all(got[b, n + 1:] == 0.0)


@pytest.fixture(scope="module")
def demos(cuda):
    """Four expert demonstrations from the benchmark's initial-state
    distribution (seed 123), solved with 60 SQP iterations and PCG 1e-20."""
    say("[imitation] expert data, 4 trajectories")
    sp = il.expert_solver_params()
    sp["num_scp_iteration_max"] = 60
    sp["pcg"]["tol_epsilon"] = 1e-20
    x0, us, sol = il.collect_expert_data(num_trajs=4, seed=123, device=cuda, solver_params=sp)
    torch.cuda.synchronize()
    say(f"  expert convergence errors {_np(sol.convergence_error)} iterations {_np(sol.num_iter)}")
    assert np.allclose(_np(x0), il.sample_initial_states(4, 123))
    return _np(x0), _np(us)


def _trainer(cuda, x0, us):
    problem = il.cartpole_problem()
    solver = il.SQPSolver(program=problem, params=il.training_solver_params())
    return il.ImitationLearning(solver, x0, us, device=cuda)


@pytest.mark.parametrize("start", [0, 7])
def test_imitation_loss_and_gradient_match_oracle(cuda, demos, start):
    say(f"[imitation] loss and gradient, starting point {start}")
    x0, us = demos
    tr = _trainer(cuda, x0, us)
    w = torch.tensor(il.INITIAL_WEIGHTS[start], dtype=torch.float64, device=cuda, requires_grad=True)
    loss = tr.loss(w)
    (g,) = torch.autograd.grad(loss, [w])
    torch.cuda.synchronize()
    prob = oil.cartpole_problem(x0, tr.problem.params)
    want_loss, want_g = oil.il_loss_and_grad(prob, W.solver_params(5, 1e-1, 1e-14, linesearch=True), us,
                                             il.INITIAL_WEIGHTS[start])
    el = abs(float(loss) - want_loss) / abs(want_loss)
    eg = float(np.max(np.abs(_np(g) - want_g)) / np.max(np.abs(want_g)))
    say(f"  loss {float(loss):.8e} vs {want_loss:.8e} (rel {el:.2e}); grad rel {eg:.2e}")
    assert el <= 1e-4
    assert eg <= 1e-4


def test_demonstration_guess_is_the_euler_rollout(cuda, demos):
    say("[imitation] demonstration guess")
    x0, us = demos
    tr = _trainer(cuda, x0, us)
    got = _np(tr.demonstration_states())
    want = oil.euler_rollout(oil.cartpole_problem(x0, tr.problem.params), us)
    assert rel(got, want) <= 1e-5


def test_graphed_step_replays_match_the_eager_step(cuda, demos):
    say("[imitation] graphed step vs eager, 3 steps")
    x0, us = demos
    tr = _trainer(cuda, x0, us)
    w0 = torch.tensor(il.INITIAL_WEIGHTS[3], dtype=torch.float64, device=cuda)
    eager, w = [], w0
    for _ in range(3):
        w, loss = tr.step(w)
        eager.append((_np(w), float(loss)))
    step = tr.graphed_step()
    w = w0
    for k in range(3):
        w, loss = step(w)
        torch.cuda.synchronize()
        _ = tr.model_loss(w)  # eager work between replays
        ew, el = eager[k]
        say(f"  step {k}: loss {float(loss):.10e} eager {el:.10e}")
        assert np.max(np.abs(_np(w) - ew)) <= 1e-12 * np.max(np.abs(ew))
        assert abs(float(loss) - el) <= 1e-12 * abs(el)


def test_training_loop_bookkeeping(cuda, demos):
    say("[imitation] train, 2 starting points x 2 steps")
    x0, us = demos
    tr = _trainer(cuda, x0, us)
    out = tr.train(il.INITIAL_WEIGHTS[:2], num_steps=2, graphed=True)
    assert out["imitation_losses"].shape == (2, 3)
    assert out["model_losses"].shape == (2, 3)
    assert out["times"].shape == (2, 3) and np.all(np.diff(out["times"], axis=1) > 0)
    for i in range(2):
        w0 = torch.tensor(il.INITIAL_WEIGHTS[i], dtype=torch.float64, device=cuda)
        w1, l0 = tr.step(w0)
        _, l1 = tr.step(w1)
        true_w = np.array([1.0, 2.0, 1.5, 1.0])
        assert np.isclose(out["imitation_losses"][i, 0], float(l0), rtol=1e-12)
        assert np.isclose(out["imitation_losses"][i, 1], float(l0), rtol=1e-12)
        assert np.isclose(out["imitation_losses"][i, 2], float(l1), rtol=1e-12)
        assert np.isclose(out["model_losses"][i, 0], np.linalg.norm(il.INITIAL_WEIGHTS[i] - true_w), rtol=1e-12)
        assert np.isclose(out["model_losses"][i, 2], np.linalg.norm(_np(w1) - true_w), rtol=1e-12)
