"""Host-side API: parameter files, problem class, guesses, parameter layouts."""
import numpy as np
import pytest
import torch

from diffmpc_amd import (CartpoleDynamics, ChainDynamics, LinearDynamics, OptimalControlProblem, QuadrotorDynamics,
                         SpacecraftAttitudeDynamics, SpacecraftDynamics, SQPSolver, load_problem_params,
                         load_solver_params)
from diffmpc_amd.dynamics.quadrotor_dynamics import S, get_rotation, q_conj, q_left
from oracle import dynamics as odyn
from oracle import sqp as osqp


def test_load_params_repeats_references():
    p = load_problem_params("spacecraft.yaml")
    assert p["horizon"] == 25
    assert p["reference_state_trajectory"].shape == (26, 3)
    assert p["reference_control_trajectory"].shape == (26, 3)
    assert p["initial_state"].dtype == torch.float64
    s = load_solver_params("sqp.yaml")
    assert s["num_scp_iteration_max"] == 15 and s["pcg"]["max_iter"] == 10000
    assert s["linesearch_alphas"] == [0.01, 0.1, 0.3, 0.7, 1.0]


def test_missing_required_key_raises():
    p = load_problem_params("spacecraft.yaml")
    del p["final_state"]
    with pytest.raises(KeyError):
        OptimalControlProblem(SpacecraftDynamics(), p)


def test_solver_rejects_other_programs_and_unknown_weight_keys():
    with pytest.raises(NotImplementedError):
        SQPSolver(program=object(), params=load_solver_params("sqp.yaml"))
    p = load_problem_params("spacecraft.yaml")
    solver = SQPSolver(OptimalControlProblem(SpacecraftDynamics(), p), params=load_solver_params("sqp.yaml"))
    # dynamics parameters (here inertia_vector) are differentiable keys; the
    # step size is not
    with pytest.raises(NotImplementedError):
        solver.solve(solver.initial_guess(), p, {"discretization_resolution": p["discretization_resolution"]})


def test_initial_guess_matches_oracle():
    p = load_problem_params("quadrotor.yaml")
    prob = OptimalControlProblem(QuadrotorDynamics(), p)
    B = 3
    x0 = torch.randn(B, 13, dtype=torch.float64)
    xs, us = prob.initial_guess({**p, "initial_state": x0})
    op = dict(horizon=prob.horizon, x0=x0.numpy(), x_final=np.repeat(p["final_state"].numpy()[None], B, 0),
              model=odyn.make_model("quadrotor"))
    ox, ou, _ = osqp.initial_guess(op)
    assert np.allclose(xs.numpy(), ox, atol=1e-15) and np.allclose(us.numpy(), ou)
    x1, u1 = prob.initial_guess()
    assert x1.shape == (26, 13) and u1.shape == (26, 4)


def test_flat_parameter_layouts_match_the_oracle_models():
    rng = np.random.default_rng(0)
    cases = [
        (QuadrotorDynamics(), {"mass": 0.7, "inertia": np.diag([0.01, 0.02, 0.03])}, "quadrotor"),
        (SpacecraftDynamics(), {"inertia_vector": [5.0, 2.0, 1.0]}, "spacecraft3"),
        (SpacecraftAttitudeDynamics(), {"inertia_vector": [0.5, 1.0, 1.5]}, "spacecraft7"),
        (CartpoleDynamics(), {"masscart": 1.0, "masspole": 0.1, "length": 0.5, "gravity": 9.81}, "cartpole"),
        (ChainDynamics(), {"spring_constants": np.linspace(0.5, 2, 12), "damping": np.linspace(0.05, 0.2, 12)},
         "chain"),
    ]
    for dyn, params, name in cases:
        flat = dyn.flat_params(params, 2).numpy().astype(np.float64)
        assert flat.shape == (2, dyn.num_params)
        m = odyn.make_model(name)
        x = rng.normal(size=(2, m.nx))
        u = rng.normal(size=(2, m.nu))
        if name in ("quadrotor",):
            x[:, 6:10] += [1, 0, 0, 0]
        if name == "spacecraft7":
            x[:, 0:4] += [1, 0, 0, 0]
        host = dyn.state_dot(torch.tensor(x), torch.tensor(u), params).numpy()
        assert np.allclose(host, m.state_dot(x, u, flat), rtol=1e-5, atol=1e-6), name
    lin = LinearDynamics()
    A, Bm, b = rng.normal(size=(4, 4)), rng.normal(size=(4, 2)), rng.normal(size=4)
    flat = lin.flat_params({"A": A, "B": Bm, "b": b}, 1).numpy()[0]
    assert np.allclose(flat, np.concatenate([A.ravel(), Bm.ravel(), b]), atol=1e-6)


def test_quadrotor_helpers():
    a, b = torch.randn(3, dtype=torch.float64), torch.randn(3, dtype=torch.float64)
    assert torch.allclose(S(a) @ b, torch.linalg.cross(a, b))
    q = torch.randn(4, dtype=torch.float64)
    q = q / q.norm()
    R = get_rotation(q)
    assert torch.allclose(R.T @ R, torch.eye(3, dtype=torch.float64), atol=1e-12)
    u = torch.randn(3, dtype=torch.float64)
    w = q_left(q_left(q) @ torch.cat([torch.zeros(1, dtype=torch.float64), u])) @ q_conj(q)
    assert torch.allclose(R @ u, w[1:], atol=1e-12)


def test_cost_and_constraints_match_oracle():
    p = load_problem_params("spacecraft.yaml")
    prob = OptimalControlProblem(SpacecraftDynamics(), p)
    N = prob.horizon
    rng = np.random.default_rng(3)
    x = rng.normal(size=(N + 1, 3))
    u = rng.normal(size=(N + 1, 3))
    c = prob.cost(torch.tensor(x), torch.tensor(u), p).item()
    g = prob.equality_constraints(torch.tensor(x), torch.tensor(u), p).numpy()
    op = dict(model=odyn.make_model("spacecraft3"), horizon=N, dt=0.1, scheme=0, penalize_control_reference=False,
              x0=p["initial_state"].numpy()[None], x_final=p["final_state"].numpy()[None],
              x_ref=p["reference_state_trajectory"].numpy()[None], u_ref=p["reference_control_trajectory"].numpy()[None],
              w_x=p["weights_penalization_reference_state_trajectory"].numpy()[None],
              w_u=p["weights_penalization_control_squared"].numpy()[None],
              w_f=p["weights_penalization_final_state"].numpy()[None], dyn=p["inertia_vector"].numpy()[None])
    assert np.isclose(c, osqp.cost(op, x[None], u[None])[0], rtol=1e-13)
    assert np.allclose(g, osqp.equality_constraints(op, x[None], u[None])[0], atol=1e-13)


def test_batch_params_expands_once_and_keeps_gradients():
    p = load_problem_params("quadrotor.yaml")
    solver = SQPSolver(OptimalControlProblem(QuadrotorDynamics(), p), params=load_solver_params("sqp.yaml"))
    w = p["weights_penalization_reference_state_trajectory"].clone().requires_grad_(True)
    out = solver.batch_params({**p, "weights_penalization_reference_state_trajectory": w}, 3)
    assert out["reference_state_trajectory"].shape == (3, 26, 13) and out["reference_state_trajectory"].is_contiguous()
    assert out["final_state"].shape == (3, 13) and out["initial_state"].shape == (3, 13)
    flat = out["_flat_dynamics_parameters"]
    assert flat.shape == (3, 10) and flat.dtype == torch.float32
    assert torch.equal(flat, QuadrotorDynamics().flat_params(p, 3).contiguous())
    out["weights_penalization_reference_state_trajectory"].sum().backward()
    assert torch.equal(w.grad, torch.full_like(w, 3.0))  # the expand's gradient sums over the batch
    # a weights dict without the model parameters gets no flat block
    ow = solver.batch_params({"weights_penalization_control_squared": p["weights_penalization_control_squared"]}, 3)
    assert "_flat_dynamics_parameters" not in ow and ow["weights_penalization_control_squared"].shape == (3, 4)


def test_jax_utils_counterparts():
    """value_and_jacfwd/jacrev against a closed-form Jacobian, and the PSD
    projection against the oracle's (diffmpc/utils/jax_utils.py:12-95;
    the reference's tests/test_utils.py:81-93 checks the projected spectrum)."""
    from diffmpc_amd.utils import (project_matrix_onto_positive_semidefinite_cone, value_and_jacfwd,
                                   value_and_jacrev)

    A = torch.tensor([[1.0, 2.0, 0.5], [0.0, -1.0, 3.0]], dtype=torch.float64)
    f = lambda x: A @ x + torch.sin(x[:2])  # noqa: E731
    x = torch.tensor([0.3, -0.7, 1.1], dtype=torch.float64)
    want = A.clone()
    want[0, 0] += torch.cos(x[0])
    want[1, 1] += torch.cos(x[1])
    for fn in (value_and_jacfwd, value_and_jacrev):
        y, J = fn(f, x)
        assert torch.allclose(y, f(x)) and J.shape == (2, 3) and torch.allclose(J, want, atol=1e-14)
    rng = np.random.default_rng(3)
    for n in (2, 5, 13):
        M = rng.normal(size=(n, n))
        M = 0.5 * (M + M.T)
        P = project_matrix_onto_positive_semidefinite_cone(torch.tensor(M), 1e-3)
        assert np.allclose(P.numpy(), osqp.project_psd(M, 1e-3), atol=1e-12)
        assert torch.linalg.eigvalsh(P).min() >= 1e-3 - 1e-12
        assert torch.equal(P, P.T)


def _quadrotor_oracle_problem(p, B=1):
    rep = lambda k: np.repeat(p[k].numpy()[None], B, 0)  # noqa: E731
    return dict(model=odyn.make_model("quadrotor"), horizon=int(p["horizon"]), dt=0.03, scheme=2,
                penalize_control_reference=True, x0=rep("initial_state"), x_final=rep("final_state"),
                x_ref=rep("reference_state_trajectory"), u_ref=rep("reference_control_trajectory"),
                w_x=rep("weights_penalization_reference_state_trajectory"),
                w_u=rep("weights_penalization_control_squared"), w_f=rep("weights_penalization_final_state"),
                dyn=np.repeat(np.concatenate([p["mass"].numpy(), p["inertia"].numpy()])[None], B, 0))


def test_cost_terms_match_oracle_and_the_cost():
    """SQPSolver.get_cost_terms / OptimalControlProblem.get_cost_linearized_matrices
    (sqp.py:160-197, ocp.py:249-323) against the oracle's cost_terms, with the
    control reference penalised (quadrotor.yaml), one problem and a batch; and
    step_cost + final_cost summing to cost (ocp.py:153-247)."""
    p = load_problem_params("quadrotor.yaml")
    prob = OptimalControlProblem(QuadrotorDynamics(), p)
    solver = SQPSolver(prob, params=load_solver_params("sqp.yaml"))
    N = prob.horizon
    rng = np.random.default_rng(4)
    x = rng.normal(size=(2, N + 1, 13))
    u = rng.normal(size=(2, N + 1, 4))
    op = _quadrotor_oracle_problem(p, 2)
    Qw, qw, Rw, rw = osqp.cost_terms(op, x, u)
    Qs, qs, Rs, rs = solver.get_cost_terms(torch.tensor(x), torch.tensor(u), p)
    for got, want in ((Qs, Qw), (qs, qw), (Rs, Rw), (rs, rw)):
        assert got.shape == want.shape and np.allclose(got.numpy(), want, rtol=1e-12, atol=1e-12)
    Q1, q1, R1, r1 = solver.get_cost_terms(torch.tensor(x[1]), torch.tensor(u[1]), p)
    assert np.allclose(Q1.numpy(), Qw[1]) and np.allclose(q1.numpy(), qw[1]) and np.allclose(r1.numpy(), rw[1])
    xt, ut = torch.tensor(x[0]), torch.tensor(u[0])
    steps = prob.step_cost(xt[:N], ut[:N], {**p, "reference_state_trajectory": p["reference_state_trajectory"][:N],
                                            "reference_control_trajectory": p["reference_control_trajectory"][:N]})
    total = steps.sum() + prob.final_cost(xt[N], p)
    assert np.isclose(total.item(), prob.cost(xt, ut, p).item(), rtol=1e-13)
    d = SQPSolver.state_control_difference(xt, ut, xt + 0.5, ut - 2.0)
    assert d.item() == 2.0


@pytest.mark.parametrize("field,shape", [
    ("dual", (1, 25, 13)), ("convergence_errors", (1, 7)), ("num_pcg_iters", (2, 5)),
    ("dual_backward_guess", (1, 26, 12)), ("states", (1, 26, 12)),
])
def test_solve_rejects_guess_arrays_of_another_shape(field, shape):
    """A guess, history or dual left over from a solver with another horizon,
    size or iteration cap raises ValueError before anything reaches the
    device (the kernels index them with fixed strides)."""
    p = load_problem_params("quadrotor.yaml")
    sp = load_solver_params("sqp.yaml")
    sp["num_scp_iteration_max"] = 5
    solver = SQPSolver(OptimalControlProblem(QuadrotorDynamics(), p), params=sp)
    g = solver.initial_guess(device="cpu")
    g = g._replace(**{field: torch.zeros(shape, dtype=torch.int32 if field == "num_pcg_iters" else torch.float64)})
    with pytest.raises(ValueError, match=field):
        solver.solve(g, p, device="cpu")


@pytest.mark.parametrize("key,shape", [
    ("reference_state_trajectory", (25, 13)), ("weights_penalization_control_squared", (3,)),
    ("initial_state", (2, 13)), ("final_state", (12,)),
])
def test_solve_rejects_parameters_of_another_shape(key, shape):
    p = load_problem_params("quadrotor.yaml")
    solver = SQPSolver(OptimalControlProblem(QuadrotorDynamics(), p), params=load_solver_params("sqp.yaml"))
    bad = dict(p)
    bad[key] = torch.zeros(shape, dtype=torch.float64)
    with pytest.raises(ValueError):
        solver.solve(solver.initial_guess(device="cpu"), bad, device="cpu")


def test_solve_requires_a_gpu_device():
    p = load_problem_params("quadrotor.yaml")
    solver = SQPSolver(OptimalControlProblem(QuadrotorDynamics(), p), params=load_solver_params("sqp.yaml"))
    with pytest.raises(RuntimeError, match="GPU"):
        solver.solve(solver.initial_guess(device="cpu"), p, device="cpu")
