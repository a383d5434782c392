"""Shared problem set-ups for the tests (oracle and GPU sides)."""
import numpy as np

from oracle import dynamics as odyn
from diffmpc_amd import workloads as W


def oracle_problem(name, prob):
    p = dict(prob)
    if name == "linear":
        p["model"] = odyn.make_model("linear", prob["x0"].shape[1], prob["u_ref"].shape[2])
    else:
        p["model"] = odyn.make_model(name)
    return p


def notebook_spacecraft():
    """notebooks/RL_spacecraft.ipynb cells 5, 7, 8: B=2, horizon 40 (N=39), 20 sim steps."""
    rs = np.random.RandomState(0)
    inertia = 1.0 + rs.rand(3) * 9.0
    x_init = -0.1 + rs.rand(2, 3) * 0.2
    B, N = 2, 39
    prob = dict(horizon=N, dt=0.1, scheme=0, penalize_control_reference=False,
                x0=x_init.copy(), x_final=np.zeros((B, 3)), x_ref=np.zeros((B, N + 1, 3)),
                u_ref=np.zeros((B, N + 1, 3)), w_x=np.ones((B, 3)), w_u=np.ones((B, 3)), w_f=np.zeros((B, 3)),
                dyn=np.repeat(inertia[None], B, 0))
    params = W.solver_params(1, 1e-8, 1e-12, linesearch=True, alphas=(1.0,))
    return inertia, x_init, prob, params


def notebook_quadrotor():
    """notebooks/RL_quadrotor.ipynb cells 5, 7, 8, 12: B=1, horizon 20 (N=19),
    10 RK4-simulated steps, quadrotor.yaml costs with mass 0.1 and inertia
    [[.1,.01,.01],[.01,.1,.01],[.01,.01,.1]], <= 50 SQP iterations with the
    default line search, PCG 1e-12.  Returns (x_init (1, 13), workload dict,
    solver params, guess function, reward, reward gradient)."""
    np.random.seed(0)  # generate_problem_data(batch_size=1, seed=0)
    x = np.array([0., 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0])[None].repeat(1, 0)
    x[:, :3] = -4.0 + np.random.rand(1, 3) * 8.0
    x[:, 3:6] = -0.1 + np.random.rand(1, 3) * 0.2
    delta = 0.5 * np.random.randn(4)
    q = np.array([1.0, 0, 0, 0]) + delta / np.linalg.norm(delta)
    x[:, 6:10] = q / np.linalg.norm(q)
    N, nx, nu = 19, 13, 4
    xr = np.array([0., 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0])
    ur = np.array([0.981, 0, 0, 0])
    rep = lambda v: np.asarray(v, np.float64)[None].copy()  # noqa: E731
    inertia = [0.1, 0.01, 0.01, 0.01, 0.1, 0.01, 0.01, 0.01, 0.1]
    prob = dict(horizon=N, dt=0.03, scheme=2, penalize_control_reference=True,
                x0=x.copy(), x_final=rep(xr), x_ref=np.tile(xr, (1, N + 1, 1)), u_ref=np.tile(ur, (1, N + 1, 1)),
                w_x=rep([1e-4, 1e-4, 1e-4, 1e-1, 1e-1, 1e-1, 1e-4, 1e-4, 1e-4, 1e-4, 1e-1, 1e-1, 1e-1]),
                w_u=rep([1e1, 1e-1, 1e-1, 1e-1]),
                w_f=rep([1e5, 1e5, 1e5, 1e3, 1e3, 1e3, 1e3, 1e3, 1e3, 1e3, 1e3, 1e3, 1e3]),
                dyn=rep([0.1] + inertia))
    params = W.solver_params(50, 1e-8, 1e-12, linesearch=True)

    def guess(p):
        """solver_initial_guess (cell 8): the straight line with normalised
        quaternions, controls [1, 0, 0, 0], zero multipliers."""
        t = np.arange(N + 1, dtype=np.float64)[None, :, None]
        xs = p["x0"][:, None, :] * (N - t) / N + p["x_final"][:, None, :] * t / N + 1e-6
        xs[:, :, 6:10] /= np.linalg.norm(xs[:, :, 6:10], axis=-1, keepdims=True)
        us = np.zeros((xs.shape[0], N + 1, nu))
        us[:, :, 0] = 1.0
        return xs, us, np.zeros((xs.shape[0], N + 1, nx))

    def reward(xn, u):
        return -(np.sum((xn - xr)[:, :3] ** 2, 1) + 0.1 * np.sum((u - ur) ** 2, 1))

    def reward_grad(xn, u):
        gx = np.zeros_like(xn)
        gx[:, :3] = -2.0 * (xn - xr)[:, :3]
        return gx, -0.2 * (u - ur)

    return x, prob, params, guess, reward, reward_grad


def small_random_qp(T=3, nx=4, nu=5, seed=0, B=2):
    """Random QP of the reference's PCG warm-start test (tests/test_solver_pcg.py:179-227), batched."""
    rng = np.random.default_rng(seed)
    As = np.stack([np.broadcast_to(np.diag(rng.random(nx)), (T, nx, nx)) for _ in range(B)])
    As_next = np.stack([np.broadcast_to(np.diag(rng.random(nx)), (T, nx, nx)) for _ in range(B)])
    Bs = np.ones((B, T, nx, nu))
    Qinv = np.stack([np.tile(np.diag(rng.random(nx)), (T + 1, 1, 1)) for _ in range(B)])
    Rinv = np.tile(np.eye(nu), (B, T, 1, 1))
    Cs = np.ones((B, T + 1, nx))
    qs = rng.normal(size=(B, T + 1, nx))
    rs = rng.normal(size=(B, T, nu))
    return [np.ascontiguousarray(a) for a in (As_next, As, Bs, Qinv, Rinv, Cs, qs, rs)]


def dense_kkt(As_next, As, Bs, Qinv, qvec, Rinv, rvec, Cs):
    """Dense KKT solve [H G^T; G 0][z; lam] = [-q; C] for one instance."""
    N, nx, nu = As.shape[0], As.shape[1], Bs.shape[2]
    nz = (N + 1) * (nx + nu)
    H = np.zeros((nz, nz))
    q = np.zeros(nz)
    G = np.zeros(((N + 1) * nx, nz))
    xi = lambda t: t * (nx + nu)  # noqa: E731
    ui = lambda t: t * (nx + nu) + nx  # noqa: E731
    for t in range(N + 1):
        H[xi(t):xi(t) + nx, xi(t):xi(t) + nx] = np.linalg.inv(Qinv[t])
        q[xi(t):xi(t) + nx] = qvec[t]
        if t < N:
            H[ui(t):ui(t) + nu, ui(t):ui(t) + nu] = np.linalg.inv(Rinv[t])
            q[ui(t):ui(t) + nu] = rvec[t]
        else:
            H[ui(t):ui(t) + nu, ui(t):ui(t) + nu] = np.eye(nu)
    G[0:nx, xi(0):xi(0) + nx] = np.eye(nx)
    for t in range(N):
        r = (t + 1) * nx
        G[r:r + nx, xi(t):xi(t) + nx] = As[t]
        G[r:r + nx, ui(t):ui(t) + nu] = Bs[t]
        G[r:r + nx, xi(t + 1):xi(t + 1) + nx] = As_next[t]
    K = np.block([[H, G.T], [G, np.zeros((G.shape[0],) * 2)]])
    s = np.linalg.solve(K, np.concatenate([-q, Cs.reshape(-1)]))
    z = s[:nz].reshape(N + 1, nx + nu)
    return z[:, :nx], z[:N, nx:], s[nz:].reshape(N + 1, nx), H, G


class OracleCustomModel(odyn.Model):
    """Oracle model of a user-defined model: the user's own state_dot function
    evaluated on numpy arrays (complex-step safe), float64."""

    def __init__(self, dyn):
        self.name = dyn.name
        self.nx = dyn.num_states
        self.nu = dyn.num_controls
        self.model_id = dyn.model_id
        self._fn = dyn._fn

    def state_dot(self, x, u, p):
        from diffmpc_amd.dynamics.custom_dynamics import LastAxis

        out = self._fn(LastAxis(x), LastAxis(u), LastAxis(p), np)
        shape = np.broadcast_shapes(x.shape[:-1], u.shape[:-1], p.shape[:-1])
        return np.stack([np.broadcast_to(np.asarray(o), shape) for o in out], -1)


def unicycle_problem(batch=4, seed=0, horizon=30):
    """Reach a goal pose with the example unicycle: RK4, dt = 0.1, per-instance
    turn gains; 10 SQP iterations to 1e-8, PCG 1e-10."""
    rng = np.random.default_rng(seed)
    B, N = batch, horizon
    x0 = np.concatenate([rng.uniform(-1, 1, (B, 2)), rng.uniform(-0.5, 0.5, (B, 1))], 1)
    goal = np.concatenate([rng.uniform(1, 2, (B, 2)), rng.uniform(-0.5, 0.5, (B, 1))], 1)
    prob = dict(horizon=N, dt=0.1, scheme=2, penalize_control_reference=False, x0=x0, x_final=goal,
                x_ref=np.repeat(goal[:, None], N + 1, 1), u_ref=np.zeros((B, N + 1, 2)),
                w_x=np.tile([1.0, 1.0, 0.1], (B, 1)), w_u=np.tile([0.1, 0.1], (B, 1)),
                w_f=np.tile([10.0, 10.0, 1.0], (B, 1)), dyn=rng.uniform(0.8, 1.2, (B, 1)))
    return prob, W.solver_params(10, 1e-8, 1e-10)


def custom_solver(dyn, prob, sp, device, mode="fused"):
    """SQPSolver for a CustomDynamics on a workload-style problem dict."""
    import torch

    from diffmpc_amd import OptimalControlProblem, SQPSolver

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
    for k, (o, n) in dyn.parameter_offsets().items():
        params[k] = t(prob["dyn"][:, o:o + n])
    sp = dict(sp)
    sp["pcg"] = dict(sp["pcg"])
    sp["mode"] = mode
    solver = SQPSolver(OptimalControlProblem(dyn, params), params=sp)
    return solver, params, solver.initial_guess(params, device=device)
