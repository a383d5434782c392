"""Physical invariants that pin the oracle's nonlinear models and integrators.

Each check is fixed by mechanics or by a closed form, not by the oracle's own
formulas, and each is shown to be sensitive: a planted mistake (a mutant
model, built here on the CPU from the oracle's function with one error put
in) must fail the same check.

* cart-pole (diffmpc/dynamics/cartpole_dynamics.py:36-80, the Gym equations):
  they are the Lagrangian equations of a cart of mass m_c carrying a uniform
  rod of mass m_p and half-length l (theta from the upright), whose energy
  E = 1/2 (m_c+m_p) xd^2 + m_p l cos(th) xd thd + 2/3 m_p l^2 thd^2 + m_p g l cos(th)
  changes at the rate of the applied power, dE/dt = F xd.
* chain (no reference counterpart): unit masses, zero-rest-length springs k_j
  and dampers c_j; E = 1/2 sum |v_j|^2 + 1/2 sum k_j |p_j - p_{j-1}|^2 obeys
  dE/dt = sum u_j v_jx - sum c_j |v_j - v_{j-1}|^2 (energy conserved at c = 0).
* attitude kinematics q' = 1/2 q (x) [0; w] (quadrotor_dynamics.py:129, shared
  by the attitude spacecraft): q . q' = 0, and the rotation matrix moves as
  R' = R S(w) (body-frame rates); Euler's equation J w' = tau - w x (J w)
  (spacecraft_dynamics.py:63, quadrotor_dynamics.py:130) gives
  d/dt (1/2 w.J w) = w . tau.
* integrators (diffmpc/dynamics/integrators.py:50-71) on xdot = A x: Euler
  I + hA, midpoint I + hA + (hA)^2/2, RK4 the degree-4 Taylor polynomial of
  exp(hA); with a control, midpoint adds dt (I + hA/2) B u.

Time derivatives of E along the flow are taken by complex step,
Im E(x + i h f(x)) / h, exact to rounding for these analytic functions.
"""
import numpy as np
import pytest

from oracle import dynamics as D

rng = np.random.default_rng(11)
H = 1e-30


def _rate(E, x, xdot, *args):
    """d/dt E(x(t)) = grad E . xdot by complex step."""
    return E(x.astype(np.complex128) + 1j * H * xdot, *args).imag / H


# ---------------------------------------------------------------- cart-pole
CARTPOLE_P = np.array([1.0, 0.1, 0.5, 9.81])


def cartpole_energy(x, p):
    xd, th, thd = x[..., 1], x[..., 2], x[..., 3]
    mc, mp, l, g = p[0], p[1], p[2], p[3]
    return (0.5 * (mc + mp) * xd**2 + mp * l * np.cos(th) * xd * thd
            + (2.0 / 3.0) * mp * l**2 * thd**2 + mp * g * l * np.cos(th))


def cartpole_power_residual(state_dot, p=CARTPOLE_P, n=64):
    """max |dE/dt - F xd| / scale over random states and forces."""
    x = rng.normal(size=(n, 4)) * np.array([1.0, 1.5, 2.0, 3.0])
    u = rng.normal(size=(n, 1)) * 5.0
    dE = _rate(cartpole_energy, x, state_dot(x, u, p), p)
    power = u[:, 0] * x[:, 1]
    scale = np.max(np.abs(power)) + np.max(np.abs(dE))
    return np.max(np.abs(dE - power)) / scale


def _cartpole_mutant(kind):
    def f(x, u, p):
        xd, th, thd = x[..., 1], x[..., 2], x[..., 3]
        F = u[..., 0]
        mc, mp, l, g = p[..., 0], p[..., 1], p[..., 2], p[..., 3]
        M, pml = mc + mp, mp * l
        s, c = np.sin(th), np.cos(th)
        temp = (F + pml * thd**2 * s) / M
        if kind == "no_centripetal":
            temp = F / M
        inertia = 4.0 / 3.0 if kind != "point_mass_pole" else 1.0
        th_acc = (g * s - c * temp) / (l * (inertia - (mp * c**2) / M))
        x_acc = temp - pml * th_acc * c / M
        if kind == "flipped_reaction":
            x_acc = temp + pml * th_acc * c / M
        return np.stack([xd, x_acc, thd, th_acc], -1)
    return f


def test_cartpole_energy_rate_is_applied_power():
    assert cartpole_power_residual(D.cartpole_state_dot) < 1e-13
    # F = 0: energy is conserved
    x = rng.normal(size=(32, 4))
    dE = _rate(cartpole_energy, x, D.cartpole_state_dot(x, np.zeros((32, 1)), CARTPOLE_P), CARTPOLE_P)
    assert np.max(np.abs(dE)) < 1e-12


@pytest.mark.parametrize("kind", ["flipped_reaction", "no_centripetal", "point_mass_pole"])
def test_cartpole_energy_check_catches_planted_mistakes(kind):
    assert cartpole_power_residual(_cartpole_mutant(kind)) > 1e-3


# ---------------------------------------------------------------- chain
M = D.CHAIN_MASSES


def chain_energy(x, p):
    pos = x[..., : 2 * M].reshape(x.shape[:-1] + (M, 2))
    vel = x[..., 2 * M:].reshape(x.shape[:-1] + (M, 2))
    k = p[:M]
    stretch = np.diff(pos, axis=-2, prepend=np.zeros_like(pos[..., :1, :]))
    return 0.5 * np.sum(vel * vel, axis=(-1, -2)) + 0.5 * np.sum(k * np.sum(stretch * stretch, -1), -1)


def chain_energy_residual(state_dot, p, n=32):
    """max |dE/dt - (input power - damper dissipation)| / scale."""
    x = rng.normal(size=(n, 4 * M))
    u = rng.normal(size=(n, M))
    dE = _rate(chain_energy, x, state_dot(x, u, p), p)
    vel = x[:, 2 * M:].reshape(n, M, 2)
    dvel = np.diff(vel, axis=-2, prepend=np.zeros_like(vel[:, :1, :]))
    expected = np.sum(u * vel[..., 0], -1) - np.sum(p[M:] * np.sum(dvel * dvel, -1), -1)
    scale = np.max(np.abs(expected)) + np.max(np.abs(dE))
    return np.max(np.abs(dE - expected)) / scale


def _chain_mutant(kind):
    def f(x, u, p):
        out = D.chain_state_dot(x, u, p)
        pos = x[..., : 2 * M].reshape(x.shape[:-1] + (M, 2))
        vel = x[..., 2 * M:].reshape(x.shape[:-1] + (M, 2))
        k, c = p[:M, None], p[M:, None]
        zero = np.zeros_like(pos[..., :1, :])
        link = -k * (pos - np.concatenate([zero, pos[..., :-1, :]], -2)) \
            - c * (vel - np.concatenate([zero, vel[..., :-1, :]], -2))
        if kind == "no_reaction":  # mass j does not feel link j+1
            acc = link
        elif kind == "damper_sign":
            link = link + 2 * c * (vel - np.concatenate([zero, vel[..., :-1, :]], -2))
            acc = link - np.concatenate([link[..., 1:, :], np.zeros_like(link[..., :1, :])], -2)
        else:  # control on y instead of x
            acc = link - np.concatenate([link[..., 1:, :], np.zeros_like(link[..., :1, :])], -2)
            acc = acc + np.stack([np.zeros_like(u), u], -1)
            return np.concatenate([out[..., : 2 * M], acc.reshape(x.shape[:-1] + (2 * M,))], -1)
        acc = acc + np.stack([u, np.zeros_like(u)], -1)
        return np.concatenate([out[..., : 2 * M], acc.reshape(x.shape[:-1] + (2 * M,))], -1)
    return f


CHAIN_P = np.concatenate([rng.uniform(0.5, 2.0, M), rng.uniform(0.05, 0.2, M)])


def test_chain_energy_conserved_without_damping_and_dissipates_with_it():
    p0 = CHAIN_P.copy()
    p0[M:] = 0.0
    assert chain_energy_residual(D.chain_state_dot, p0) < 1e-13
    x = rng.normal(size=(8, 4 * M))
    dE = _rate(chain_energy, x, D.chain_state_dot(x, np.zeros((8, M)), p0), p0)
    assert np.max(np.abs(dE)) < 1e-12
    assert chain_energy_residual(D.chain_state_dot, CHAIN_P) < 1e-13


@pytest.mark.parametrize("kind", ["no_reaction", "damper_sign", "control_axis"])
def test_chain_energy_check_catches_planted_mistakes(kind):
    assert chain_energy_residual(_chain_mutant(kind), CHAIN_P) > 1e-3


# ---------------------------------------------------------------- attitude
def _unit_quaternions(n):
    q = rng.normal(size=(n, 4))
    return q / np.linalg.norm(q, axis=-1, keepdims=True)


def attitude_residuals(qdot_fn, wdot_fn, Jd, n=32):
    """(q.q', |R' - R S(w)|, |d/dt 1/2 w.Jw - w.tau|, momentum balance), each relative.

    The last is the inertial angular momentum balance d/dt (R J w) = R tau,
    i.e. with R' = R S(w): S(w) J w + J w' = tau.
    """
    q = _unit_quaternions(n)
    w = rng.normal(size=(n, 3))
    tau = rng.normal(size=(n, 3))
    qd = qdot_fn(q, w)
    r_norm = np.max(np.abs(np.sum(q * qd, -1)))
    Rdot = D.rotation(q.astype(np.complex128) + 1j * H * qd).imag / H
    expect = np.einsum("nij,njk->nik", D.rotation(q), D.skew(w))
    r_frame = np.max(np.abs(Rdot - expect)) / np.max(np.abs(expect))
    wd = wdot_fn(w, tau)
    dE = np.sum(w * Jd * wd, -1)
    power = np.sum(w * tau, -1)
    r_energy = np.max(np.abs(dE - power)) / np.max(np.abs(power))
    bal = np.cross(w, Jd * w) + Jd * wd
    r_momentum = np.max(np.abs(bal - tau)) / np.max(np.abs(tau))
    return r_norm, r_frame, r_energy, r_momentum


SC7_J = np.array([0.7, 1.3, 1.9])


def _sc7_parts(state_dot):
    def qdot(q, w):
        return state_dot(np.concatenate([q, w], -1), np.zeros_like(w), SC7_J)[..., :4]

    def wdot(w, tau):
        q = _unit_quaternions(w.shape[0])
        return state_dot(np.concatenate([q, w], -1), tau, SC7_J)[..., 4:]
    return qdot, wdot


def test_attitude_kinematics_and_euler_equation():
    r_norm, r_frame, r_energy, r_momentum = attitude_residuals(*_sc7_parts(D.spacecraft7_state_dot), SC7_J)
    assert r_norm < 1e-14 and r_frame < 1e-13 and r_energy < 1e-13 and r_momentum < 1e-13
    # the rate part is the reference's rate-only spacecraft
    x = np.concatenate([_unit_quaternions(8), rng.normal(size=(8, 3))], -1)
    u = rng.normal(size=(8, 3))
    assert np.allclose(D.spacecraft7_state_dot(x, u, SC7_J)[:, 4:],
                       D.spacecraft3_state_dot(x[:, 4:], u, SC7_J), rtol=0, atol=1e-14)
    # the quaternion is normalised inside the derivative (quadrotor_dynamics.py:122)
    xs = x.copy()
    xs[:, :4] *= 3.0
    assert np.allclose(D.spacecraft7_state_dot(xs, u, SC7_J), D.spacecraft7_state_dot(x, u, SC7_J), atol=1e-14)


def _sc7_mutant(kind):
    def f(x, u, p):
        out = D.spacecraft7_state_dot(x, u, p)
        q, w = x[..., :4], x[..., 4:]
        if kind == "inertial_frame":  # q' = 1/2 [0; w] (x) q
            w4 = np.concatenate([np.zeros_like(w[..., :1]), w], -1)
            qd = 0.5 * np.einsum("...ij,...j->...i", D.q_left(w4), q)
            return np.concatenate([qd, out[..., 4:]], -1)
        if kind == "gyro_sign":
            wd = (u + np.cross(w, p * w)) / p
            return np.concatenate([out[..., :4], wd], -1)
        return np.concatenate([out[..., :4] * 2.0, out[..., 4:]], -1)  # missing 1/2
    return f


@pytest.mark.parametrize("kind", ["inertial_frame", "gyro_sign", "missing_half"])
def test_attitude_checks_catch_planted_mistakes(kind):
    r = attitude_residuals(*_sc7_parts(_sc7_mutant(kind)), SC7_J)
    assert max(r) > 1e-3


QUAD_P = np.concatenate([[1.3], np.array([[0.011, 0.001, 0.0], [0.001, 0.012, 0.0005],
                                          [0.0, 0.0005, 0.021]]).ravel()])


def test_quadrotor_translation_and_attitude():
    """p' = v, v' = (f/m) R e3 - g e3, and the attitude invariants with a full inertia."""
    n = 16
    q = _unit_quaternions(n)
    x = np.concatenate([rng.normal(size=(n, 6)), q, rng.normal(size=(n, 3))], -1)
    u = rng.normal(size=(n, 4))
    xd = D.quadrotor_state_dot(x, u, QUAD_P)
    assert np.allclose(xd[:, :3], x[:, 3:6], atol=0)
    # R e3 by the quaternion sandwich q (x) [0; e3] (x) q*
    e3 = np.array([0.0, 0.0, 0.0, 1.0])
    Re3 = np.stack([(D.q_left(D.q_left(qi) @ e3) @ D.q_conj(qi))[1:] for qi in q])
    assert np.allclose(xd[:, 3:6], Re3 * (u[:, :1] / QUAD_P[0]) - np.array([0, 0, 9.81]), atol=1e-13)
    J = QUAD_P[1:].reshape(3, 3)

    def qdot(qq, w):
        xx = np.concatenate([np.zeros((w.shape[0], 6)), qq, w], -1)
        return D.quadrotor_state_dot(xx, np.zeros((w.shape[0], 4)), QUAD_P)[:, 6:10]

    def wdot(w, tau):
        xx = np.concatenate([np.zeros((w.shape[0], 6)), _unit_quaternions(w.shape[0]), w], -1)
        return D.quadrotor_state_dot(xx, np.concatenate([np.zeros((w.shape[0], 1)), tau], -1), QUAD_P)[:, 10:]

    qq = _unit_quaternions(n)
    w = rng.normal(size=(n, 3))
    tau = rng.normal(size=(n, 3))
    assert np.max(np.abs(np.sum(qq * qdot(qq, w), -1))) < 1e-14
    Rdot = D.rotation(qq.astype(np.complex128) + 1j * H * qdot(qq, w)).imag / H
    assert np.allclose(Rdot, np.einsum("nij,njk->nik", D.rotation(qq), D.skew(w)), atol=1e-13)
    wd = wdot(w, tau)
    assert np.allclose(np.einsum("ni,ij,nj->n", w, J, wd), np.sum(w * tau, -1), rtol=1e-12, atol=1e-13)
    # torque-free: angular momentum J w is constant in the inertial frame,
    # d/dt (R J w) = R (S(w) J w + J w') = 0
    wd0 = wdot(w, np.zeros_like(tau))
    assert np.allclose(np.cross(w, w @ J.T) + wd0 @ J.T, 0.0, atol=1e-13)


# ---------------------------------------------------------------- integrators
def _linear_model(A, Bm):
    nx, nu = Bm.shape
    return D.make_model("linear", nx, nu), np.concatenate([A.ravel(), Bm.ravel(), np.zeros(nx)])


def integrator_residuals(step, A, Bm, dt):
    """|step - closed form| for Euler, midpoint, RK4 on xdot = A x + B u."""
    nx, nu = Bm.shape
    m, p = _linear_model(A, Bm)
    x = rng.normal(size=(6, nx))
    u = rng.normal(size=(6, nu))
    hA = dt * A
    I = np.eye(nx)
    P = {0: I + hA, 1: I + hA + hA @ hA / 2,
         2: I + hA + hA @ hA / 2 + hA @ hA @ hA / 6 + hA @ hA @ hA @ hA / 24}
    # control terms: Euler dt B u; midpoint dt (I + hA/2) B u;
    # RK4 dt (I + hA/2 + (hA)^2/6 + (hA)^3/24) B u
    Q = {0: I, 1: I + hA / 2, 2: I + hA / 2 + hA @ hA / 6 + hA @ hA @ hA / 24}
    out = {}
    for s in (0, 1, 2):
        want = x @ P[s].T + dt * (u @ Bm.T) @ Q[s].T
        got = step(m, dt, s, p, x, u)
        out[s] = np.max(np.abs(got - want)) / np.max(np.abs(want))
    return out


def test_integrators_match_closed_forms_on_linear_flow():
    # scalar xdot = a x: midpoint x1 = x0 (1 + a dt + (a dt)^2 / 2)
    m, p = _linear_model(np.array([[-0.7]]), np.zeros((1, 1)))
    x0 = np.array([[1.5]])
    h = -0.7 * 0.3
    assert np.isclose(D.predict_next_state(m, 0.3, 1, p, x0, np.zeros((1, 1)))[0, 0],
                      1.5 * (1 + h + h * h / 2), rtol=1e-15)
    A = rng.normal(size=(5, 5))
    Bm = rng.normal(size=(5, 2))
    r = integrator_residuals(D.predict_next_state, A, Bm, 0.13)
    assert max(r.values()) < 1e-14, r


def _integrator_mutant(kind):
    def step(model, dt, scheme, p, x, u):
        f = lambda z: model.state_dot(z, u, p)  # noqa: E731
        if scheme == 1:
            if kind == "midpoint_full_step":
                return x + dt * f(x + dt * f(x))
            if kind == "midpoint_at_start":
                return x + dt * f(x)
        if scheme == 2 and kind == "rk4_weights":
            k1 = f(x)
            k2 = f(x + 0.5 * dt * k1)
            k3 = f(x + 0.5 * dt * k2)
            k4 = f(x + dt * k3)
            return x + (k1 + k2 + 2 * k3 + 2 * k4) * dt / 6.0
        return D.predict_next_state(model, dt, scheme, p, x, u)
    return step


@pytest.mark.parametrize("kind,scheme", [("midpoint_full_step", 1), ("midpoint_at_start", 1), ("rk4_weights", 2)])
def test_integrator_check_catches_planted_mistakes(kind, scheme):
    A = rng.normal(size=(5, 5))
    Bm = rng.normal(size=(5, 2))
    assert integrator_residuals(_integrator_mutant(kind), A, Bm, 0.13)[scheme] > 1e-4
