"""The reference's own solver tests, run through this package's host API on
the GPU, each also checked against the float64 oracle:

* tests/test_optimal_control.py:30-41  spacecraft.yaml as loaded, PCG 1e-15
* tests/test_solver_sqp.py:30-100       spacecraft, horizons 10/20, batch 1/2,
                                         line search on/off, adjoint warm start on/off
* tests/test_solver_sqp.py:103-121      warm vs cold adjoint start
* tests/test_differentiable.py:21-92    gradient of ||x*|| + ||u*|| w.r.t. the
                                         state and control weights vs central differences

The problem set-ups follow the reference files (spacecraft.yaml, seeded
numpy draws); trajectories within 1e-5 and gradients within 1e-4 of the
oracle, besides the reference's own assertions."""
import os
import sys

import numpy as np
import pytest
import torch

from diffmpc_amd import OptimalControlProblem, SpacecraftDynamics, SQPSolver, load_problem_params, load_solver_params
from diffmpc_amd import _lib
from diffmpc_amd.solvers.sqp import SolverReturnStatus
from oracle import dynamics as odyn
from oracle import sqp as osqp

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _np(t):
    return t.detach().cpu().numpy()


def _oracle_problem(p, x0):
    """The oracle problem of a spacecraft.yaml-style parameter dict for
    initial states x0 (B, 3)."""
    x0 = np.atleast_2d(np.asarray(x0, np.float64))
    B, N = x0.shape[0], int(p["horizon"])
    a = lambda k: np.asarray(_np(torch.as_tensor(p[k])), np.float64)  # noqa: E731
    rep = lambda v: np.repeat(v.reshape(1, -1), B, 0)  # noqa: E731
    return dict(model=odyn.make_model("spacecraft3"), horizon=N, dt=float(p["discretization_resolution"]),
                scheme=int(p["discretization_scheme"]), penalize_control_reference=bool(p["penalize_control_reference"]),
                x0=x0, x_final=rep(a("final_state")), x_ref=np.repeat(a("reference_state_trajectory")[None], B, 0),
                u_ref=np.repeat(a("reference_control_trajectory")[None], B, 0),
                w_x=rep(a("weights_penalization_reference_state_trajectory")),
                w_u=rep(a("weights_penalization_control_squared")),
                w_f=rep(a("weights_penalization_final_state")),
                dyn=rep(a("inertia_vector")).astype(np.float32).astype(np.float64))


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def test_spacecraft_with_sqp_solver(cuda):
    """tests/test_optimal_control.py:30-41."""
    say("[reference tests] spacecraft.yaml with the SQP solver")
    problem_params = load_problem_params("spacecraft.yaml")
    solver_parameters = load_solver_params("sqp.yaml")
    solver_parameters["pcg"]["tol_epsilon"] = 1e-15
    problem = OptimalControlProblem(dynamics=SpacecraftDynamics(), params=problem_params)
    solver = SQPSolver(program=problem, params=solver_parameters)
    sol = solver.solve(solver.initial_guess(device=cuda), problem.params)
    torch.cuda.synchronize()
    assert float(sol.convergence_error) < 2e-5
    assert SolverReturnStatus(int(sol.status)) is SolverReturnStatus.SUCCESS
    ref = osqp.sqp_solve(_oracle_problem(problem_params, _np(problem_params["initial_state"])), solver_parameters)
    assert _rel(_np(sol.states), ref["states"][0]) <= 1e-5
    assert _rel(_np(sol.controls), ref["controls"][0]) <= 1e-5


def _sqp_solve_spacecraft(cuda, num_batch, horizon, use_linesearch, warm_start_backward):
    """sqp_solve_spacecraft of tests/test_solver_sqp.py:30-80: inertia and
    initial states from np.random.seed(0), zero references, no terminal
    weight, 20 SQP iterations to 1e-6, PCG 1e-12, alphas [1.0]; solves
    initial_states[0] from the straight-line guess."""
    np.random.seed(0)
    inertia_vector = 1.0 + np.random.rand(3) * 9.0
    initial_states = -0.1 + np.random.rand(num_batch, 3) * 0.2
    problem_params = load_problem_params("spacecraft.yaml")
    N = horizon - 1
    problem_params["horizon"] = N
    problem_params["inertia_vector"] = torch.tensor(inertia_vector)
    problem_params["reference_state_trajectory"] = torch.zeros(N + 1, 3, dtype=torch.float64)
    problem_params["reference_control_trajectory"] = torch.zeros(N + 1, 3, dtype=torch.float64)
    problem_params["weights_penalization_final_state"] = torch.zeros(3, dtype=torch.float64)
    solver_params = load_solver_params("sqp.yaml")
    solver_params["tol_convergence"] = 1.0e-6
    solver_params["num_scp_iteration_max"] = 20
    solver_params["pcg"]["tol_epsilon"] = 1.0e-12
    solver_params["linesearch"] = use_linesearch
    solver_params["warm_start_backward"] = warm_start_backward
    solver_params["linesearch_alphas"] = [1.0]
    dynamics = SpacecraftDynamics()
    problem = OptimalControlProblem(dynamics=dynamics, params=problem_params)
    solver = SQPSolver(program=problem, params=solver_params)
    sqp_guess = solver.initial_guess()
    weights = {k: problem_params[k] for k in ("weights_penalization_reference_state_trajectory",
                                              "weights_penalization_control_squared")}
    x0 = torch.tensor(initial_states[0])
    guess = solver.initial_guess({**problem_params, "initial_state": x0}, device=cuda)
    solution = solver.solve(guess, {**problem_params, "initial_state": x0}, weights)
    torch.cuda.synchronize()
    return solver, sqp_guess, solution, dynamics, problem_params, solver_params, initial_states[0]


@pytest.mark.parametrize("num_batch", [1, 2])
@pytest.mark.parametrize("horizon", [10, 20])
@pytest.mark.parametrize("use_linesearch", [True, False])
@pytest.mark.parametrize("warm_start_backward", [True, False])
def test_sqp_solve(cuda, num_batch, horizon, use_linesearch, warm_start_backward):
    """tests/test_solver_sqp.py:83-100 (plus the oracle's trajectories)."""
    say(f"[reference tests] test_sqp_solve B={num_batch} T={horizon} ls={use_linesearch} ws={warm_start_backward}")
    solver, sqp_guess, solution, dynamics, p, sp, x0 = _sqp_solve_spacecraft(
        cuda, num_batch, horizon, use_linesearch, warm_start_backward)
    assert solver.name == "SQPSolver"
    assert solver.pcg is not None
    assert sqp_guess.states.shape == (horizon, dynamics.num_states)
    assert sqp_guess.controls.shape == (horizon, dynamics.num_controls)
    assert sqp_guess.dual.shape == (horizon, dynamics.num_states)
    assert int(solution.status) == SolverReturnStatus.SUCCESS
    assert np.isfinite(float(solution.convergence_error))
    ref = osqp.sqp_solve(_oracle_problem(p, x0), sp)
    assert _rel(_np(solution.states), ref["states"][0]) <= 1e-5
    assert _rel(_np(solution.controls), ref["controls"][0]) <= 1e-5


def test_sqp_solve_warmstart(cuda):
    """tests/test_solver_sqp.py:103-121."""
    say("[reference tests] test_sqp_solve_warmstart")
    it = {}
    for warm_start_backward in (True, False):
        *_, solution, _, _, _, _ = _sqp_solve_spacecraft(cuda, 1, 20, True, warm_start_backward)
        assert int(solution.status) == SolverReturnStatus.SUCCESS
        assert np.isfinite(float(solution.convergence_error))
        it["warm" if warm_start_backward else "cold"] = int(solution.num_iter)
    assert it["warm"] <= it["cold"]


def test_sqp_solve_differentiable(cuda):
    """tests/test_differentiable.py:21-92: horizon 10, x0 from np.random.seed(0),
    30 SQP iterations to 1e-12, PCG 1e-24; the gradient of ||x*|| + ||u*||
    with respect to the shared state and control weights, by autograd through
    the implicit backward, against central differences of the oracle's
    objective (|fd - grad| <= 1e-3 as there, and 1e-4 relative)."""
    say("[reference tests] test_sqp_solve_differentiable")
    np.random.seed(0)
    horizon, N = 10, 9
    problem_params = load_problem_params("spacecraft.yaml")
    problem_params["initial_state"] = torch.tensor(-0.1 + np.random.rand(3) * 0.2)
    problem_params["horizon"] = N
    problem_params["reference_state_trajectory"] = torch.zeros(N + 1, 3, dtype=torch.float64)
    problem_params["reference_control_trajectory"] = torch.zeros(N + 1, 3, dtype=torch.float64)
    problem_params["weights_penalization_final_state"] = torch.zeros(3, dtype=torch.float64)
    solver_params = load_solver_params("sqp.yaml")
    solver_params["tol_convergence"] = 1.0e-12
    solver_params["num_scp_iteration_max"] = 30
    solver_params["pcg"]["tol_epsilon"] = 1.0e-24
    solver_params["linesearch"] = False
    solver_params["warm_start_backward"] = True
    problem = OptimalControlProblem(dynamics=SpacecraftDynamics(), params=problem_params)
    solver = SQPSolver(program=problem, params=solver_params)
    keys = ("weights_penalization_reference_state_trajectory", "weights_penalization_control_squared")
    weights = {k: problem_params[k].clone().to(cuda).requires_grad_(True) for k in keys}
    sol = solver.solve(solver.initial_guess(device=cuda), problem_params, weights)
    objective = torch.linalg.norm(sol.states) + torch.linalg.norm(sol.controls)
    grads = torch.autograd.grad(objective, [weights[k] for k in keys])
    torch.cuda.synchronize()

    base = _oracle_problem(problem_params, _np(problem_params["initial_state"]))

    def f(p):
        s = osqp.sqp_solve(p, solver_params)
        return np.linalg.norm(s["states"][0]) + np.linalg.norm(s["controls"][0])

    h = 1e-6
    for key, g in zip(("w_x", "w_u"), grads):
        fd = np.zeros(3)
        for i in range(3):
            hi, lo = dict(base), dict(base)
            hi[key], lo[key] = base[key].copy(), base[key].copy()
            hi[key][0, i] += h
            lo[key][0, i] -= h
            fd[i] = (f(hi) - f(lo)) / (2 * h)
        say(f"  {key}: grad {_np(g)} fd {fd}")
        assert np.allclose(fd - _np(g), 0.0, atol=1e-3, rtol=1e-3)
        assert _rel(_np(g), fd) <= 1e-4


# ---- tests/test_solver_pcg.py, one problem at a time as there ---------------
from diffmpc_amd.solvers.pcg import (PCGOptimalControl, QPCostParameters, QPDynamicsParameters,  # noqa: E402
                                     QPParameters)


def _t(a, cuda):
    return torch.as_tensor(np.asarray(a, np.float64), device=cuda)


def test_pcg_constructor_and_zero_shapes():
    """tests/test_solver_pcg.py:17-56 (no GPU work)."""
    solver_params = {"max_iter": 50, "tol_epsilon": 1e-5}
    oc = PCGOptimalControl(10, 4, 2, solver_params)
    assert (oc.name, oc.horizon, oc.num_states, oc.num_controls, oc.params) == (
        "PCGOptimalControl", 10, 4, 2, solver_params)
    qp = PCGOptimalControl(5, 3, 2).zero_qp_parameters()
    assert qp.dynamics.As_next.shape == (5, 3, 3) and qp.dynamics.Bs.shape == (5, 3, 2)
    assert qp.dynamics.Cs.shape == (6, 3) and qp.cost.Qinv.shape == (6, 3, 3) and qp.cost.rvec.shape == (5, 2)
    oc = PCGOptimalControl(4, 2, 1)
    assert oc.zero_dynamics_cost_premultiplied_matrices().Bs_x_Rinv.shape == (4, 2, 1)
    assert oc.zero_schur_complement_matrices().S.shape == (5, 2, 6)


def _identity_qp(T, nx, nu, cuda, As=None, As_next=None, Qinv=None):
    eye = np.eye(nx)
    return QPParameters(
        QPDynamicsParameters(As_next=_t(np.broadcast_to(eye if As_next is None else As_next, (T, nx, nx)), cuda),
                             As=_t(np.broadcast_to(eye if As is None else As, (T, nx, nx)), cuda),
                             Bs=_t(np.ones((T, nx, nu)), cuda), Cs=_t(np.ones((T, nx)), cuda)),
        QPCostParameters(Qinv=_t(np.tile(eye if Qinv is None else Qinv, (T + 1, 1, 1)), cuda),
                         qvec=_t(np.zeros((T + 1, nx)), cuda), Rinv=_t(np.tile(np.eye(nu), (T, 1, 1)), cuda),
                         rvec=_t(np.zeros((T, nu)), cuda)))


@pytest.mark.parametrize("T", [2, 5])
@pytest.mark.parametrize("nx", [1, 5])
@pytest.mark.parametrize("nu", [1, 5])
def test_pcg_layer_small_problems(cuda, T, nx, nu):
    """tests/test_solver_pcg.py:58-177: S/Phiinv of identity dynamics, gamma of
    zeros, recovery from zero multipliers (with the test's (T+1, 1) cost
    vectors, which broadcast), and the trivial KKT solve."""
    say(f"[reference tests] pcg layer T={T} nx={nx} nu={nu}")
    oc = PCGOptimalControl(T, nx, nu, {"max_iter": 10, "tol_epsilon": 1e-8})
    qp = _identity_qp(T, nx, nu, cuda)
    schur = oc.compute_S_Phiinv(qp)
    assert schur.S.shape == (T + 1, nx, 3 * nx) and schur.preconditioner_Phiinv.shape[0] == T + 1
    assert not torch.isnan(schur.S).any()
    dyn_cost = oc.zero_dynamics_cost_premultiplied_matrices(device=cuda)
    gamma = oc.compute_gamma(dyn_cost, _t(np.zeros((T + 1, nx)), cuda), _t(np.eye(nx), cuda),
                             _t(np.zeros((T + 1, nx)), cuda), _t(np.zeros((T, nu)), cuda))
    assert gamma.shape == (T + 1, nx) and torch.allclose(gamma, torch.zeros_like(gamma))
    z = oc.zero_qp_parameters(device=cuda)
    qpz = QPParameters(z.dynamics, QPCostParameters(Qinv=_t(np.tile(np.eye(nx), (T + 1, 1, 1)), cuda),
                                                    qvec=_t(np.zeros((T + 1, 1)), cuda),
                                                    Rinv=_t(np.tile(np.eye(nu), (T, 1, 1)), cuda),
                                                    rvec=_t(np.zeros((T, 1)), cuda)))
    sol = oc.get_states_controls_from_kkt_multipliers(_t(np.zeros((T + 1, nx)), cuda), qpz)
    assert sol.states.shape == (T + 1, nx) and sol.controls.shape == (T + 1, nu)
    assert not sol.states.abs().any() and not sol.controls.abs().any()
    if T <= 3 and nx <= 2 and nu <= 2:
        gammas = oc.compute_gamma(dyn_cost, _t(np.ones((T + 1, nx)), cuda), _t(np.eye(nx), cuda),
                                  _t(np.zeros((T + 1, nx)), cuda), _t(np.zeros((T, nu)), cuda))
        sol, debug = oc.solve_KKT_Schur(qp, schur, gammas, _t(np.zeros((T + 1, nx)), cuda))
        assert sol.states.shape == (T + 1, nx) and sol.controls.shape == (T + 1, nu)
        assert int(debug.num_iterations) <= 10 and float(debug.convergence_eta) >= 0.0


def test_pcg_layer_warmstart(cuda):
    """tests/test_solver_pcg.py:179-227: a converged guess is kept and needs
    fewer iterations."""
    say("[reference tests] pcg warm start")
    T, nx, nu = 3, 4, 5
    rng = np.random.default_rng(0)
    oc = PCGOptimalControl(T, nx, nu, {"max_iter": 20, "tol_epsilon": 1e-13})
    qp = _identity_qp(T, nx, nu, cuda, As=np.diag(rng.random(nx)), As_next=np.diag(rng.random(nx)),
                      Qinv=np.diag(rng.random(nx)))
    schur = oc.compute_S_Phiinv(qp)
    dyn_cost = oc.zero_dynamics_cost_premultiplied_matrices(device=cuda)
    gammas = oc.compute_gamma(dyn_cost, _t(np.ones((T + 1, nx)), cuda), _t(np.eye(nx), cuda),
                              _t(np.zeros((T + 1, nx)), cuda), _t(np.zeros((T, nu)), cuda))
    sol, debug = oc.solve_KKT_Schur(qp, schur, gammas, _t(np.zeros((T + 1, nx)), cuda))
    assert int(debug.num_iterations) <= 10 and float(debug.convergence_eta) >= 0.0
    sol_w, debug_w = oc.solve_KKT_Schur(qp, schur, gammas, sol.kkt_multipliers)
    assert torch.allclose(sol.states, sol_w.states) and torch.allclose(sol.controls, sol_w.controls)
    assert torch.allclose(sol.kkt_multipliers, sol_w.kkt_multipliers)
    assert int(debug_w.num_iterations) < int(debug.num_iterations)


def test_one_sqp_iteration_through_the_layer_matches_oracle(cuda):
    """SQPSolver.solve_one_scp_iteration (sqp.py:262-299: cost terms, GPU
    linearization, Qinv/Rinv, premultiplied products, S/Phiinv, gamma, PCG)
    on one quadrotor problem from quadrotor.yaml, against the oracle's QP of
    the same iteration, and the same for a batch of two."""
    say("[reference tests] one SQP iteration through the PCG layer")
    from diffmpc_amd import QuadrotorDynamics
    from diffmpc_amd.solvers.pcg import QPSolution

    p = load_problem_params("quadrotor.yaml")
    sp = load_solver_params("sqp.yaml")
    # quadrotor.yaml's weights (1e-4 .. 1e5) make this QP ill-conditioned: at
    # eps 1e-14 the two PCG runs (float32 and float64 A_t) stop ~1.5e-5 apart,
    # so both solve it to the end
    sp["pcg"]["tol_epsilon"] = 1e-24
    sp["pcg"]["max_iter"] = 20000
    solver = SQPSolver(OptimalControlProblem(QuadrotorDynamics(), p), params=sp)
    g = solver.initial_guess(device=cuda)
    rng = np.random.default_rng(2)
    x = g.states + 0.05 * torch.tensor(rng.normal(size=tuple(g.states.shape)), device=cuda)
    u = g.controls + 0.1 * torch.tensor(rng.normal(size=tuple(g.controls.shape)), device=cuda)
    sol, qp, pre, schur, iters = solver.solve_one_scp_iteration(QPSolution(x, u, g.dual), p)
    torch.cuda.synchronize()
    rep = lambda k: np.asarray(_np(torch.as_tensor(p[k])), np.float64)[None]  # noqa: E731
    op = dict(model=odyn.make_model("quadrotor"), horizon=int(p["horizon"]), dt=0.03, scheme=2,
              penalize_control_reference=True, x0=rep("initial_state"), x_final=rep("final_state"),
              x_ref=rep("reference_state_trajectory"), u_ref=rep("reference_control_trajectory"),
              w_x=rep("weights_penalization_reference_state_trajectory"),
              w_u=rep("weights_penalization_control_squared"), w_f=rep("weights_penalization_final_state"),
              dyn=np.concatenate([rep("mass"), rep("inertia")], axis=1).astype(np.float32).astype(np.float64))
    qpp = osqp.qp_parameters(op, _np(x)[None], _np(u)[None])
    xo, uo, lo, *_ = osqp.solve_qp(qpp, np.zeros((1, op["horizon"] + 1, 13)), sp["pcg"])
    say(f"  PCG iterations {int(iters)}; rel x {_rel(_np(sol.states), xo[0]):.2e} u {_rel(_np(sol.controls), uo[0]):.2e}")
    assert _rel(_np(sol.states), xo[0]) <= 1e-5 and _rel(_np(sol.controls), uo[0]) <= 1e-5
    assert _rel(_np(sol.kkt_multipliers), lo[0]) <= 1e-5
    xb, ub = torch.stack([x, x]), torch.stack([u, u])
    solb, *_ = solver.solve_one_scp_iteration(QPSolution(xb, ub, torch.stack([g.dual, g.dual])), p)
    torch.cuda.synchronize()
    assert torch.allclose(solb.states[1], sol.states, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("nx,nu,mode", [(6, 3, "auto"), (6, 3, "plain"), (2, 3, "auto"), (2, 3, "plain"),
                                         (2, 3, "fused")])
def test_linear_model_of_any_size(cuda, nx, nu, mode):
    """A linear model of a size the library itself is not built for solves as
    the reference solves any LinearDynamics: LinearDynamics runs it in a
    custom build of the library (prebuilt in the package by the build step).
    nx=6, nu=3 runs fused by default; nx=2, nu=3 (more controls than states)
    runs in the per-launch mode, and asking for the fused mode raises.
    Forward and backward (w_x, w_u, x0) are checked against the oracle."""
    say(f"[reference tests] linear nx={nx} nu={nu}, mode {mode}")
    from diffmpc_amd import LinearDynamics

    rng = np.random.default_rng({6: 9, 2: 11}[nx])
    N = 15
    A = 0.1 * rng.normal(size=(nx, nx))
    Bm = rng.normal(size=(nx, nu))
    p = dict(horizon=N, discretization_resolution=0.1, discretization_scheme=0, penalize_control_reference=False,
             initial_state=torch.tensor(rng.normal(size=nx)), final_state=torch.zeros(nx, dtype=torch.float64),
             reference_state_trajectory=torch.zeros(N + 1, nx, dtype=torch.float64),
             reference_control_trajectory=torch.zeros(N + 1, nu, dtype=torch.float64),
             weights_penalization_reference_state_trajectory=torch.ones(nx, dtype=torch.float64),
             weights_penalization_control_squared=0.1 * torch.ones(nu, dtype=torch.float64),
             weights_penalization_final_state=torch.zeros(nx, dtype=torch.float64),
             A=torch.tensor(A), B=torch.tensor(Bm), b=torch.zeros(nx, dtype=torch.float64))
    dyn = LinearDynamics({"num_states": nx, "num_controls": nu, "names_states": [f"x{i}" for i in range(nx)],
                          "names_controls": [f"u{i}" for i in range(nu)]})
    sp = load_solver_params("sqp.yaml")
    sp["pcg"]["tol_epsilon"] = 1e-14
    sp["mode"] = mode
    if not os.path.exists(os.path.join(dyn.prebuilt_dir(), "libdmpc_custom.so")):
        pytest.skip(f"linear{nx}x{nu} custom library not built (run __graft_entry__.build())")
    solver = SQPSolver(OptimalControlProblem(dyn, p), params=sp)
    wx = p["weights_penalization_reference_state_trajectory"].to(cuda).requires_grad_(True)
    wu = p["weights_penalization_control_squared"].to(cuda).requires_grad_(True)
    x0 = p["initial_state"].to(cuda).requires_grad_(True)
    weights = {"weights_penalization_reference_state_trajectory": wx,
               "weights_penalization_control_squared": wu, "initial_state": x0}
    if mode == "fused" and nu > nx:
        with pytest.raises(_lib.DmpcError):
            solver.solve(solver.initial_guess(device=cuda), p, weights)
        return
    sol = solver.solve(solver.initial_guess(device=cuda), p, weights)
    torch.cuda.synchronize()
    assert int(sol.status) == SolverReturnStatus.SUCCESS
    op = dict(model=odyn.make_model("linear", nx, nu), horizon=N, dt=0.1, scheme=0, penalize_control_reference=False,
              x0=_np(p["initial_state"])[None], x_final=np.zeros((1, nx)), x_ref=np.zeros((1, N + 1, nx)),
              u_ref=np.zeros((1, N + 1, nu)), w_x=np.ones((1, nx)), w_u=0.1 * np.ones((1, nu)), w_f=np.zeros((1, nx)),
              dyn=np.concatenate([A.ravel(), Bm.ravel(), np.zeros(nx)])[None].astype(np.float32).astype(np.float64))
    ref = osqp.sqp_solve(op, sp)
    assert _rel(_np(sol.states), ref["states"][0]) <= 1e-5
    assert _rel(_np(sol.controls), ref["controls"][0]) <= 1e-5
    # loss 1/2 |x|^2 + 1/2 |u|^2: every control row, including those past nx, feeds the gradient
    loss = 0.5 * sol.states.square().sum() + 0.5 * sol.controls.square().sum()
    gwx, gwu, gx0 = torch.autograd.grad(loss, [wx, wu, x0])
    gref = osqp.sqp_backward(op, sp, ref, ref["states"], ref["controls"])
    assert _rel(_np(gwx), gref["w_x"][0]) <= 1e-4
    assert _rel(_np(gwu), gref["w_u"][0]) <= 1e-4
    assert _rel(_np(gx0), gref["x0"][0]) <= 1e-4
