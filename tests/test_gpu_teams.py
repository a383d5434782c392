"""Several instances per workgroup ("teams" of the register-resident fused
kernels), through the C ABI.

A team computes what a one-instance workgroup computes: the same lane groups,
reduction rows and summation order, and the same PCG stop rule; a team whose
solve has stopped keeps passing the barriers with its state frozen.  So
without the line search the forward and adjoint results are the same bits as
with one instance per workgroup: trajectories, multipliers, PCG and SQP
counts, step norms and the gradient w.r.t. the state weights.  The cases
cover batches that leave teams of the last workgroup empty, ragged horizons
(teams of one workgroup with different PCG counts) and instances that meet
the SQP stop rule at different iterations.  With the line search the cost's
directional derivative is summed in another order, so those solves are
compared with the float64 oracle (step sizes exactly, as the one-instance
line-search test does).
"""
import sys

import numpy as np
import pytest
import torch

from diffmpc_amd import _lib
from diffmpc_amd import workloads as W
from oracle import sqp as osqp
from tests.helpers import oracle_problem

pytestmark = pytest.mark.gpu

WKEY = "weights_penalization_reference_state_trajectory"


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _np(t):
    return t.detach().cpu().numpy()


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def _lti_ragged():
    name, prob, sp = W.lti(batch=41, seed=12, horizon=24)
    hs = np.random.default_rng(5).integers(4, 25, size=41).astype(np.int32)
    return name, prob, sp, hs


def _attitude_stop_rule():
    # tol 1e-3: the oracle's instances stop after 3, 4 or 5 SQP iterations
    name, prob, sp = W.spacecraft7(batch=23, seed=2, horizon=12, sqp_iters=5, pcg_eps=1e-10)
    sp = dict(sp, tol_convergence=1e-3)
    return name, prob, sp, None


BITWISE_CASES = {
    "lti-N20-B37": lambda: W.lti(batch=37, seed=11, horizon=20) + (None,),
    "lti-ragged-N4..24-B41": _lti_ragged,
    "cartpole-N30-B29": lambda: W.cartpole(batch=29, seed=6, horizon=30, sqp_iters=3) + (None,),
    "spacecraft3-N15-B19": lambda: W.spacecraft3(batch=19, seed=5, horizon=15, sqp_iters=3) + (None,),
    "spacecraft7-stop-rule-B23": _attitude_stop_rule,
}


def _solve(name, prob, sp, hs, cuda, ipw):
    solver, params, guess = W.make_solver(name, prob, sp, device=cuda, horizons=hs)
    solver.params["instances_per_workgroup"] = ipw
    wx = params[WKEY].clone().requires_grad_(True)
    sol = solver.solve(guess, params, {WKEY: wx})
    loss = 0.5 * (sol.states - params["reference_state_trajectory"]).square().sum()
    (g,) = torch.autograd.grad(loss, [wx])
    torch.cuda.synchronize()
    return {"states": _np(sol.states), "controls": _np(sol.controls), "dual": _np(sol.dual),
            "num_iter": _np(sol.num_iter), "status": _np(sol.status), "pcg": _np(sol.num_pcg_iters),
            "conv": _np(sol.convergence_errors), "grad_wx": _np(g)}


@pytest.mark.parametrize("ipw", [2, 3, 4])
@pytest.mark.parametrize("case", list(BITWISE_CASES))
def test_instances_per_workgroup_give_the_same_bits(cuda, case, ipw):
    name, prob, sp, hs = BITWISE_CASES[case]()
    say(f"[teams] {case} instances per workgroup {ipw}")
    one = _solve(name, prob, sp, hs, cuda, 0)
    packed = _solve(name, prob, sp, hs, cuda, ipw)
    if case.startswith("spacecraft7"):
        say("  SQP iterations", np.unique(one["num_iter"]).tolist())
        assert len(np.unique(one["num_iter"])) > 1  # instances stopped at different iterations
    if hs is not None:
        assert len(np.unique(one["pcg"][:, 0])) > 1
    for k in one:
        assert np.array_equal(one[k], packed[k]), k
    say("  identical; PCG iterations", int(one["pcg"].sum()))


def test_instances_per_workgroup_with_the_line_search_match_oracle(cuda):
    """Cart-pole (4-lane groups) and attitude (8-lane groups) with the merit
    line search, 2 and 4 instances per workgroup (the last workgroup partly
    empty): step sizes equal the oracle's, trajectories within 1e-5."""
    cases = [("cartpole", W.cartpole(batch=3, seed=6), 4), ("spacecraft7", W.spacecraft7(batch=3, seed=2, horizon=30), 3)]
    for name, (_, prob, sp), maxit in cases:
        sp = dict(sp)
        sp.update(linesearch=True, num_scp_iteration_max=maxit, tol_convergence=1e-8)
        ref = osqp.sqp_solve(oracle_problem(name, prob), sp)
        for ipw in (2, 4):
            say(f"[teams, line search] {name} instances per workgroup {ipw}")
            solver, params, guess = W.make_solver(name, prob, sp, device=cuda)
            solver.params["instances_per_workgroup"] = ipw
            sol = solver.solve(guess, params)
            torch.cuda.synchronize()
            say("  alphas", _np(sol.linesearch_alphas).tolist())
            assert np.array_equal(_np(sol.linesearch_alphas), ref["linesearch_alphas"])
            assert rel(_np(sol.states), ref["states"]) <= 1e-5
            assert rel(_np(sol.controls), ref["controls"]) <= 1e-5
            assert rel(_np(sol.dual), ref["dual"]) <= 1e-5
            assert np.array_equal(_np(sol.num_iter), ref["num_iter"])


def test_instances_per_workgroup_outside_their_shape_are_rejected(cuda):
    """Teams exist for 4- and 8-lane groups with one block per group in the
    register-resident kernels: 16-lane models, the streamed variant, two
    blocks per group and teams that overflow a workgroup are unsupported."""
    def solve(name, prob, sp, **tuning):
        solver, params, guess = W.make_solver(name, prob, sp, device=cuda)
        solver.params.update(tuning)
        with torch.no_grad():
            solver.solve(guess, params)
        torch.cuda.synchronize()

    _, q, qsp = W.quadrotor(batch=2, seed=1, horizon=10, sqp_iters=1)
    _, c, csp = W.cartpole(batch=2, seed=6, horizon=30, sqp_iters=1)
    _, lc, lsp = W.cartpole(batch=2, seed=6, horizon=100, sqp_iters=1)  # 404 lanes -> 416 per team
    for name, prob, sp, tuning in [("quadrotor", q, qsp, dict(instances_per_workgroup=2)),
                                   ("cartpole", c, csp, dict(instances_per_workgroup=2, fused_variant="streamed")),
                                   ("cartpole", c, csp, dict(instances_per_workgroup=2, blocks_per_group=2)),
                                   ("cartpole", lc, lsp, dict(instances_per_workgroup=3))]:
        say(f"[teams, unsupported] {name} {tuning}")
        with pytest.raises(_lib.DmpcError, match="unsupported"):
            solve(name, prob, sp, **tuning)
    say("[teams] cartpole N=100, 2 per workgroup (832 threads) is supported")
    solve("cartpole", lc, lsp, instances_per_workgroup=2)
