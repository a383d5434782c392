"""Seeded synthetic workloads of the benchmark configurations.

Each generator returns ``(model_name, problem, solver_params)`` where
``problem`` is a dict of float64 numpy arrays with a leading batch axis and the
keys used throughout this package:

  horizon N, dt, scheme, penalize_control_reference,
  x0 (B,nx), x_final (B,nx), x_ref (B,N+1,nx), u_ref (B,N+1,nu),
  w_x (B,nx), w_u (B,nu), w_f (B,nx), dyn (B,P)

``solver_params`` uses the reference's solver keys (sqp.yaml).  The seed is
``config index + 1000 * repeat``.
"""
import numpy as np

QUAD_WX = np.array([10.0, 10, 10, 1, 1, 1, 1, 1, 1, 1, 0.1, 0.1, 0.1])


def solver_params(maxit, tol, pcg_eps, pcg_max=10000, linesearch=False, alphas=(0.01, 0.1, 0.3, 0.7, 1.0)):
    return dict(
        verbose=False, tol_convergence=tol, num_scp_iteration_max=maxit,
        linesearch=linesearch, linesearch_eta=0.4, linesearch_alphas=list(alphas),
        warm_start_backward=True,
        pcg=dict(verbose=False, max_iter=pcg_max, tol_epsilon=pcg_eps),
    )


def _rep(v, B):
    v = np.asarray(v, dtype=np.float64)
    return np.broadcast_to(v, (B,) + v.shape).copy()


def lti(batch=1, seed=0, horizon=20):
    """LTI MPC QP: nx=4, nu=2, A = I + 0.1 G, G ~ N(0, 0.1^2), B ~ 0.1 N(0, 1).

    Passed in the reference's continuous form with dt = 1 and Euler
    (A_c = A - I, B_c = B, b = 0).  w_x = 1, w_u = 0.1, terminal Qf = 10 I
    (w_f = 9), x0 ~ N(0, 1), zero reference.  One SQP iteration, PCG eps 1e-10,
    at most 200 PCG iterations.
    """
    rng = np.random.default_rng(seed)
    nx, nu, N = 4, 2, horizon
    G = rng.normal(0.0, 0.1, size=(batch, nx, nx))
    Ad = np.eye(nx) + 0.1 * G
    Bd = 0.1 * rng.normal(size=(batch, nx, nu))
    x0 = rng.normal(size=(batch, nx))
    Ac = Ad - np.eye(nx)
    dyn = np.concatenate([Ac.reshape(batch, -1), Bd.reshape(batch, -1), np.zeros((batch, nx))], axis=1)
    prob = dict(
        horizon=N, dt=1.0, scheme=0, penalize_control_reference=False,
        x0=x0, x_final=np.zeros((batch, nx)), x_ref=np.zeros((batch, N + 1, nx)),
        u_ref=np.zeros((batch, N + 1, nu)), w_x=_rep(np.ones(nx), batch),
        w_u=_rep(0.1 * np.ones(nu), batch), w_f=_rep(9.0 * np.ones(nx), batch), dyn=dyn,
    )
    return "linear", prob, solver_params(1, 1e-8, 1e-10, pcg_max=200)


def _axis_angle_quat(rng, batch, max_angle_rad):
    axis = rng.normal(size=(batch, 3))
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    ang = rng.uniform(0.0, max_angle_rad, size=batch)
    return np.concatenate([np.cos(ang / 2)[:, None], np.sin(ang / 2)[:, None] * axis], axis=1)


def quadrotor(batch=1024, seed=1, horizon=50, sqp_iters=5, pcg_eps=1e-6):
    """Quadrotor nonlinear MPC: nx=13, nu=4, dt=0.02, RK4, mass 1,
    J = diag(0.01, 0.01, 0.02).

    x0 = hover with position ~ U(-1,1)^3, velocity ~ N(0, 0.5^2) and a rotation
    of angle ~ U(0, 30 deg) about a random axis.  The target is a random
    waypoint ~ U(-1,1)^3 in hover; u_ref = [9.81, 0, 0, 0] is penalised.
    w_x = diag(10 I3, I3, I4, 0.1 I3), w_u = 0.01, w_f = 0.  Exactly
    ``sqp_iters`` SQP iterations (tol 0), PCG eps 1e-6.
    """
    rng = np.random.default_rng(seed)
    nx, nu, N = 13, 4, horizon
    x0 = np.zeros((batch, nx))
    x0[:, 0:3] = rng.uniform(-1, 1, size=(batch, 3))
    x0[:, 3:6] = rng.normal(0.0, 0.5, size=(batch, 3))
    x0[:, 6:10] = _axis_angle_quat(rng, batch, np.deg2rad(30.0))
    xt = np.zeros((batch, nx))
    xt[:, 0:3] = rng.uniform(-1, 1, size=(batch, 3))
    xt[:, 6] = 1.0
    J = np.diag([0.01, 0.01, 0.02])
    dyn = _rep(np.concatenate([[1.0], J.reshape(-1)]), batch)
    uref = np.array([9.81, 0.0, 0.0, 0.0])
    prob = dict(
        horizon=N, dt=0.02, scheme=2, penalize_control_reference=True,
        x0=x0, x_final=xt.copy(), x_ref=np.repeat(xt[:, None], N + 1, axis=1),
        u_ref=np.broadcast_to(uref, (batch, N + 1, nu)).copy(),
        w_x=_rep(QUAD_WX, batch), w_u=_rep(0.01 * np.ones(nu), batch),
        w_f=np.zeros((batch, nx)), dyn=dyn,
    )
    return "quadrotor", prob, solver_params(sqp_iters, 0.0, pcg_eps)


def spacecraft7(batch=4096, seed=2, horizon=100, sqp_iters=3, pcg_eps=1e-8):
    """Spacecraft attitude MPC: nx=7 (unit quaternion, rate), nu=3, dt=0.1,
    Euler, inertia diag ~ U(0.5, 2) per instance.

    x0 = random unit quaternion (scalar part made non-negative) with
    w ~ N(0, 0.1^2); the target is the identity attitude at rest.  The weights
    are log-parameterised, w = exp(theta), theta ~ U(-1, 1) per state / control
    component, shared by the batch; w_f = 0.  Returned with the thetas.
    """
    rng = np.random.default_rng(seed)
    nx, nu, N = 7, 3, horizon
    q = rng.normal(size=(batch, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    q *= np.where(q[:, :1] < 0, -1.0, 1.0)
    x0 = np.concatenate([q, rng.normal(0.0, 0.1, size=(batch, 3))], axis=1)
    xt = np.zeros((batch, nx))
    xt[:, 0] = 1.0
    theta_q = rng.uniform(-1, 1, size=nx)
    theta_r = rng.uniform(-1, 1, size=nu)
    dyn = rng.uniform(0.5, 2.0, size=(batch, 3))
    prob = dict(
        horizon=N, dt=0.1, scheme=0, penalize_control_reference=False,
        x0=x0, x_final=xt.copy(), x_ref=np.repeat(xt[:, None], N + 1, axis=1),
        u_ref=np.zeros((batch, N + 1, nu)), w_x=_rep(np.exp(theta_q), batch),
        w_u=_rep(np.exp(theta_r), batch), w_f=np.zeros((batch, nx)), dyn=dyn,
        theta_q=theta_q, theta_r=theta_r,
    )
    return "spacecraft7", prob, solver_params(sqp_iters, 0.0, pcg_eps)


def chain(batch=512, seed=4, horizon=256, pcg_eps=1e-8):
    """Planar 12-mass spring-damper chain: nx=48, nu=12, dt=0.05, RK4.

    k ~ U(0.5, 2), c ~ U(0.05, 0.2) per link and instance, x0 ~ N(0, 0.5^2),
    w_x = 1, w_u = 0.1, w_f = 0, zero reference.  The model is linear, so one
    SQP iteration solves it exactly.
    """
    rng = np.random.default_rng(seed)
    nx, nu, N = 48, 12, horizon
    k = rng.uniform(0.5, 2.0, size=(batch, 12))
    c = rng.uniform(0.05, 0.2, size=(batch, 12))
    x0 = rng.normal(0.0, 0.5, size=(batch, nx))
    prob = dict(
        horizon=N, dt=0.05, scheme=2, penalize_control_reference=False,
        x0=x0, x_final=np.zeros((batch, nx)), x_ref=np.zeros((batch, N + 1, nx)),
        u_ref=np.zeros((batch, N + 1, nu)), w_x=_rep(np.ones(nx), batch),
        w_u=_rep(0.1 * np.ones(nu), batch), w_f=np.zeros((batch, nx)),
        dyn=np.concatenate([k, c], axis=1),
    )
    return "chain", prob, solver_params(1, 0.0, pcg_eps)


def rl_lti(batch=64, seed=0, horizon=39, pcg_eps=1e-12):
    """The reference's reinforcement-learning benchmark problem
    (benchmarking/reinforcement-learning/utils.py generate_problem_data and
    benchmark_diffmpc.py:40-99): nx=8, nu=4, one shared discrete-time system
    A = I + 0.1 randn with its eigenvalues pulled inside the unit circle,
    B = randn, b = 0.01 randn, Q = I, R = I, x0 = 5 randn per instance (numpy's
    legacy seeded generator, so the data equal the reference's), passed in
    continuous form with dt = 1 and Euler; horizon 40 (N = 39), one SQP
    iteration, line search with alphas [1.0], PCG eps 1e-12.  Returns
    (name, prob, solver_params, (A, B, b) discrete)."""
    nx, nu, N = 8, 4, horizon
    rs = np.random.RandomState(seed)
    A = np.eye(nx) + 0.1 * rs.randn(nx, nx)
    lam, V = np.linalg.eig(A)
    mag = np.abs(lam)
    lam = np.where(mag < 1 - 1e-2, lam, lam / (mag + 1e-2))
    A = (V @ np.diag(lam) @ np.linalg.inv(V)).real
    Bm = rs.randn(nx, nu)
    b = 0.01 * rs.randn(nx)
    x0 = rs.randn(batch, nx) * 5.0
    dyn = np.concatenate([(A - np.eye(nx)).reshape(-1), Bm.reshape(-1), b])
    prob = dict(
        horizon=N, dt=1.0, scheme=0, penalize_control_reference=False,
        x0=x0, x_final=np.zeros((batch, nx)), x_ref=np.zeros((batch, N + 1, nx)),
        u_ref=np.zeros((batch, N + 1, nu)), w_x=_rep(np.ones(nx), batch),
        w_u=_rep(np.ones(nu), batch), w_f=np.zeros((batch, nx)), dyn=_rep(dyn, batch),
    )
    sp = solver_params(1, 1e-8, pcg_eps, linesearch=True, alphas=(1.0,))
    return "linear", prob, sp, (A, Bm, b)


def ragged(batch=8192, seed=3, tmin=10, tmax=200, sqp_iters=3):
    """Ragged mixed batch: horizons N ~ U{tmin..tmax} and systems drawn
    uniformly from {LTI nx=4/nu=2, spacecraft attitude nx=7/nu=3, quadrotor
    nx=13/nu=4}, each with the seeded per-instance x0 and weights of its own
    workload generator (lti, spacecraft7, quadrotor), `sqp_iters` SQP
    iterations.  Returns {name: (prob, solver_params, horizons)} with one
    group per system, padded to the group's longest horizon."""
    rng = np.random.default_rng(seed)
    kinds = rng.integers(0, 3, size=batch)
    hor = rng.integers(tmin, tmax + 1, size=batch)
    out = {}
    for k, gen in enumerate((lti, spacecraft7, quadrotor)):
        sel = np.nonzero(kinds == k)[0]
        if sel.size == 0:
            continue
        h = hor[sel].astype(np.int32)
        name, prob, sp = gen(batch=int(sel.size), seed=seed * 10 + k, horizon=int(h.max()))
        sp = dict(sp)
        sp["num_scp_iteration_max"] = sqp_iters
        sp["tol_convergence"] = 0.0
        out[name] = (prob, sp, h)
    return out


def split_by_horizon(prob, horizons, limit):
    """Two padded groups of a ragged problem: instances with horizon <= limit
    (padded to their own longest horizon) and the rest.  Returns a list of
    (prob, horizons) without empty groups."""
    out = []
    for sel in (np.nonzero(horizons <= limit)[0], np.nonzero(horizons > limit)[0]):
        if sel.size == 0:
            continue
        sub = slice_problem(prob, sel)
        n = int(horizons[sel].max())
        sub["horizon"] = n
        for k in ("x_ref", "u_ref"):
            sub[k] = sub[k][:, : n + 1]
        out.append((sub, horizons[sel]))
    return out


def instance_problem(prob, b, n):
    """Instance b of a padded (ragged) problem dict, cut to horizon n."""
    one = slice_problem(prob, np.array([b]))
    one["horizon"] = int(n)
    for k in ("x_ref", "u_ref"):
        one[k] = one[k][:, : n + 1]
    return one


def slice_problem(prob, idx):
    """Sub-batch of a problem dict (array entries with a batch axis)."""
    out = {}
    for k, v in prob.items():
        if isinstance(v, np.ndarray) and k not in ("theta_q", "theta_r"):
            out[k] = v[idx]
        else:
            out[k] = v
    return out


def dynamics_params(name, dyn, nx=4, nu=2):
    """Reference-style parameter entries from a flat (B, P) parameter block."""
    import torch

    d = torch.as_tensor(dyn, dtype=torch.float64)
    if name == "linear":
        return {"A": d[:, : nx * nx].reshape(-1, nx, nx), "B": d[:, nx * nx: nx * nx + nx * nu].reshape(-1, nx, nu),
                "b": d[:, nx * nx + nx * nu: nx * nx + nx * nu + nx]}
    if name == "quadrotor":
        return {"mass": d[:, 0], "inertia": d[:, 1:10]}
    if name in ("spacecraft3", "spacecraft7"):
        return {"inertia_vector": d[:, 0:3]}
    if name == "cartpole":
        return {"masscart": d[:, 0], "masspole": d[:, 1], "length": d[:, 2], "gravity": d[:, 3]}
    if name == "chain":
        return {"spring_constants": d[:, 0:12], "damping": d[:, 12:24]}
    raise ValueError(name)


def make_solver(name, prob, sp, device="cuda", mode="fused", horizons=None):
    """(solver, problem_params, guess) for a workload dict, on `device`.

    problem_params holds batched torch tensors under the reference's keys;
    `horizons` (B,) makes it a ragged batch padded to prob["horizon"].
    """
    import torch

    from .dynamics import (CartpoleDynamics, ChainDynamics, LinearDynamics, QuadrotorDynamics,
                           SpacecraftAttitudeDynamics, SpacecraftDynamics)
    from .problems import OptimalControlProblem
    from .solvers import SQPSolver

    nx, nu = prob["x0"].shape[1], prob["u_ref"].shape[2]
    if name == "linear":
        dyn = LinearDynamics({"num_states": nx, "num_controls": nu, "names_states": [f"x{i}" for i in range(nx)],
                              "names_controls": [f"u{i}" for i in range(nu)]})
    else:
        dyn = {"quadrotor": QuadrotorDynamics, "spacecraft3": SpacecraftDynamics,
               "spacecraft7": SpacecraftAttitudeDynamics, "cartpole": CartpoleDynamics,
               "chain": ChainDynamics}[name]()
    t = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.float64, device=device)  # noqa: E731
    params = dict(
        horizon=int(prob["horizon"]), discretization_resolution=float(prob["dt"]),
        discretization_scheme=int(prob["scheme"]), penalize_control_reference=bool(prob["penalize_control_reference"]),
        initial_state=t(prob["x0"]), final_state=t(prob["x_final"]),
        reference_state_trajectory=t(prob["x_ref"]), reference_control_trajectory=t(prob["u_ref"]),
        weights_penalization_reference_state_trajectory=t(prob["w_x"]),
        weights_penalization_control_squared=t(prob["w_u"]),
        weights_penalization_final_state=t(prob["w_f"]),
    )
    for k, v in dynamics_params(name, prob["dyn"], nx, nu).items():
        params[k] = v.to(device)
    if horizons is not None:
        params["horizons"] = torch.as_tensor(np.asarray(horizons), dtype=torch.int32, device=device)
    sp = dict(sp)
    sp["pcg"] = dict(sp["pcg"])
    sp["mode"] = mode
    problem = OptimalControlProblem(dyn, params)
    solver = SQPSolver(problem, params=sp)
    guess = solver.initial_guess(params, device=device)
    return solver, params, guess


def spacecraft3(batch=2, seed=5, horizon=39, sqp_iters=5, pcg_eps=1e-12):
    """Rate-only spacecraft (the reference's spacecraft.yaml weights), inertia
    1 + 9 U(0, 1), x0 ~ U(-0.1, 0.1)^3, zero reference, Euler, dt = 0.1."""
    rng = np.random.default_rng(seed)
    nx = nu = 3
    N = horizon
    prob = dict(
        horizon=N, dt=0.1, scheme=0, penalize_control_reference=False,
        x0=rng.uniform(-0.1, 0.1, size=(batch, nx)), x_final=np.zeros((batch, nx)),
        x_ref=np.zeros((batch, N + 1, nx)), u_ref=np.zeros((batch, N + 1, nu)), w_x=_rep(np.ones(nx), batch),
        w_u=_rep(np.ones(nu), batch), w_f=np.zeros((batch, nx)), dyn=1.0 + 9.0 * rng.random((batch, 3)),
    )
    return "spacecraft3", prob, solver_params(sqp_iters, 1e-8, pcg_eps)


def cartpole(batch=3, seed=6, horizon=30, sqp_iters=3, pcg_eps=1e-10):
    """Cart-pole swing-in: the reference's cartpole.yaml weights, RK4, dt = 0.05,
    pole angle ~ U(-0.3, 0.3), other states ~ N(0, 0.1^2)."""
    rng = np.random.default_rng(seed)
    nx, nu, N = 4, 1, horizon
    x0 = rng.normal(0.0, 0.1, size=(batch, nx))
    x0[:, 2] = rng.uniform(-0.3, 0.3, size=batch)
    prob = dict(
        horizon=N, dt=0.05, scheme=2, penalize_control_reference=False,
        x0=x0, x_final=np.zeros((batch, nx)), x_ref=np.zeros((batch, N + 1, nx)), u_ref=np.zeros((batch, N + 1, nu)),
        w_x=_rep([1.0, 2.0, 1.5, 1.0], batch), w_u=_rep([5e-2], batch), w_f=np.zeros((batch, nx)),
        dyn=_rep([1.0, 0.1, 0.5, 9.81], batch),
    )
    return "cartpole", prob, solver_params(sqp_iters, 1e-8, pcg_eps)
