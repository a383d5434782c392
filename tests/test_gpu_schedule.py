"""The register-resident forward of a batch of one horizon with more
instances than the device has CUs runs a whole solve as one launch: a
workgroup per (SQP iteration, instance) takes its position from a counter,
waits until the instance's previous iteration is done, linearizes the
instance and runs the iteration.  forward_schedule "per_iteration" runs one
linearization and one QP launch per SQP iteration instead.  The
per-instance computations are the same, so both give the same bits:
trajectories, multipliers, every history, the saved linearization the
backward reads (through the gradients).  The cases include instances that
meet the SQP stop rule at different iterations (their later positions are
skipped), instances whose state is not finite (they stop with the ERROR
status), the merit line search and two blocks per lane group, each with more
instances (300) than the device holds workgroups at once.
"""
import sys

import numpy as np
import pytest
import torch

from diffmpc_amd import workloads as W

pytestmark = pytest.mark.gpu

WKEY = "weights_penalization_reference_state_trajectory"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _np(t):
    return t.detach().cpu().numpy()


def _run(name, prob, sp, cuda, schedule, tuning):
    solver, params, guess = W.make_solver(name, prob, sp, device=cuda)
    solver.params.update(tuning, forward_schedule=schedule)
    wx = params[WKEY].clone().requires_grad_(True)
    x0 = params["initial_state"].clone().requires_grad_(True)
    sol = solver.solve(guess, params, {WKEY: wx, "initial_state": x0})
    loss = 0.5 * (sol.states - params["reference_state_trajectory"]).square().sum()
    gw, gx = torch.autograd.grad(loss, [wx, x0])
    torch.cuda.synchronize()
    out = {"states": sol.states, "controls": sol.controls, "dual": sol.dual, "num_iter": sol.num_iter,
           "status": sol.status, "pcg": sol.num_pcg_iters, "conv": sol.convergence_errors,
           "alphas": sol.linesearch_alphas, "grad_wx": gw, "grad_x0": gx}
    return {k: _np(v) for k, v in out.items()}


def _stop_rule(make):
    def f():
        name, prob, sp = make()
        return name, prob, dict(sp, tol_convergence=1e-3)
    return f


def _with_nan_instances():
    # three instances with a non-finite initial state: they stop after their
    # first SQP iteration (ERROR status), and their later positions are skipped
    name, prob, sp = W.quadrotor(batch=300, seed=5, horizon=16, sqp_iters=4)
    prob = dict(prob)
    prob["x0"] = np.array(prob["x0"], copy=True)
    prob["x0"][[0, 150, 299], 2] = np.nan
    return name, prob, sp


CASES = {
    "quadrotor-nan-instances-B300": (_with_nan_instances, {}),
    "quadrotor-N30-B300": (lambda: W.quadrotor(batch=300, seed=3, horizon=30, sqp_iters=4), {}),
    "quadrotor-N20-B300-K2": (lambda: W.quadrotor(batch=300, seed=4, horizon=20, sqp_iters=3),
                             {"fused_variant": "register", "blocks_per_group": 2}),
    "spacecraft7-stop-rule-B300": (_stop_rule(lambda: W.spacecraft7(batch=300, seed=2, horizon=12, sqp_iters=5,
                                                                    pcg_eps=1e-10)), {}),
    "cartpole-line-search-B300": (lambda: W.cartpole(batch=300, seed=6), {}),
}


@pytest.mark.parametrize("case", list(CASES))
def test_one_launch_solve_equals_per_iteration_launches(cuda, case):
    make, tuning = CASES[case]
    name, prob, sp = make()
    if case.startswith("cartpole"):
        sp = dict(sp, linesearch=True, num_scp_iteration_max=4, tol_convergence=1e-8)
    B, N = prob["x0"].shape[0], int(prob["horizon"])
    assert B > torch.cuda.get_device_properties(0).multi_processor_count  # the one-launch schedule runs
    say(f"[schedule] {case}: B={B} N={N}")
    one = _run(name, prob, sp, cuda, "auto", tuning)
    per_iter = _run(name, prob, sp, cuda, "per_iteration", tuning)
    if case.startswith("spacecraft7"):
        say("  SQP iterations", np.unique(one["num_iter"]).tolist())
        assert len(np.unique(one["num_iter"])) > 1  # instances stopped at different iterations
    if case.startswith("cartpole"):
        assert not np.all(one["alphas"] == 1.0)  # the line search shortened some steps
    if "nan" in case:
        assert (one["status"] != 0).sum() >= 3 and (one["status"][[0, 150, 299]] != 0).all()
    for k in one:
        assert np.array_equal(one[k], per_iter[k], equal_nan=True), k
    say("  identical; PCG iterations", int(one["pcg"].sum()))


def test_one_launch_solve_in_a_graph_matches_eager_on_every_replay(cuda):
    """The one-launch forward (300 quadrotor instances) with its backward,
    captured as one hipGraph: the schedule's counter is cleared by a kernel
    at the start of every replay, so each replay, with unrelated eager work
    and synchronisations in between, equals the eager step bit for bit; a
    replay after an in-place change of the initial states equals the eager
    step on the changed inputs."""
    from diffmpc_amd.graph import GraphedStep

    name, prob, sp = W.quadrotor(batch=300, seed=7, horizon=24, sqp_iters=3)
    assert prob["x0"].shape[0] > torch.cuda.get_device_properties(0).multi_processor_count
    say("[schedule, graph] quadrotor B=300 N=24")
    solver, params, guess = W.make_solver(name, prob, sp, device=cuda)
    w = params[WKEY].clone().requires_grad_(True)

    def step():
        guess.dual_backward_guess.zero_()
        sol = solver.solve(guess, params, {WKEY: w})
        loss = 0.5 * (sol.states - params["reference_state_trajectory"]).square().sum()
        (g,) = torch.autograd.grad(loss, [w])
        return sol.states, sol.num_pcg_iters, g

    def snap(outs):
        return [o.detach().clone() for o in outs]

    eager = snap(step())
    graphed = GraphedStep(step)
    junk = torch.randn(512, 512, device=cuda)
    for k in range(3):
        got = snap(graphed.replay())
        torch.cuda.synchronize()
        for e, g in zip(eager, got):
            assert torch.equal(e, g), f"replay {k}"
        junk = junk @ junk.T / 512.0  # eager work between replays
        torch.cuda.synchronize()
    params["initial_state"].mul_(0.5)
    got = snap(graphed.replay())
    eager2 = snap(step())
    for e, g in zip(eager2, got):
        assert torch.equal(e, g)
    assert not torch.equal(eager[0], eager2[0])
    say("  every replay identical")
