"""Dynamics models, integrators and Jacobians of the oracle (float64 / complex128).

Every model takes ``x (..., nx)``, ``u (..., nu)`` and a flat parameter vector
``p (..., P)`` that broadcasts against the leading axes.  The flat layouts are
the ones the C ABI uses (``include/dmpc.h``):

  linear       P = nx*nx + nx*nu + nx : A (row-major), B (row-major), b
  quadrotor    P = 10                 : mass, inertia 3x3 (row-major)
  spacecraft3  P = 3                  : inertia_vector (diagonal inertia)
  spacecraft7  P = 3                  : inertia diagonal
  cartpole     P = 4                  : masscart, masspole, length, gravity
  chain        P = 24                 : spring k[12], damper c[12]; unit masses

All functions are written so that complex inputs propagate analytically
(no abs / conjugate), which is what complex-step differentiation needs.
"""
import numpy as np

GRAVITY = 9.81  # hard-coded in diffmpc/dynamics/quadrotor_dynamics.py:112

MODEL_IDS = {
    "linear": 0,
    "quadrotor": 1,
    "spacecraft3": 2,
    "spacecraft7": 3,
    "cartpole": 4,
    "chain": 5,
}


def skew(a):
    """Cross-product matrix S(a) with S(a) b = a x b.

    Follows diffmpc/dynamics/quadrotor_dynamics.py:35-36.
    """
    z = np.zeros_like(a[..., 0])
    return np.stack(
        [
            np.stack([z, -a[..., 2], a[..., 1]], -1),
            np.stack([a[..., 2], z, -a[..., 0]], -1),
            np.stack([-a[..., 1], a[..., 0], z], -1),
        ],
        -2,
    )


def q_left(q):
    """Left quaternion-product matrix L(q): q (x) p = L(q) p (scalar-first).

    Follows diffmpc/dynamics/quadrotor_dynamics.py:40-48.
    """
    qw, qx, qy, qz = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    return np.stack(
        [
            np.stack([qw, -qx, -qy, -qz], -1),
            np.stack([qx, qw, -qz, qy], -1),
            np.stack([qy, qz, qw, -qx], -1),
            np.stack([qz, -qy, qx, qw], -1),
        ],
        -2,
    )


def q_conj(q):
    """Quaternion conjugate (diffmpc/dynamics/quadrotor_dynamics.py:52-53)."""
    return np.concatenate([q[..., :1], -q[..., 1:]], -1)


def rotation(q):
    """Rotation matrix of a unit quaternion (quadrotor_dynamics.py:57-69)."""
    qw, qx, qy, qz = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    qw2, qx2, qy2, qz2 = qw * qw, qx * qx, qy * qy, qz * qz
    return np.stack(
        [
            np.stack([qw2 + qx2 - qy2 - qz2, 2 * (qx * qy - qw * qz), 2 * (qx * qz + qw * qy)], -1),
            np.stack([2 * (qx * qy + qw * qz), qw2 - qx2 + qy2 - qz2, 2 * (qy * qz - qw * qx)], -1),
            np.stack([2 * (qx * qz - qw * qy), 2 * (qy * qz + qw * qx), qw2 - qx2 - qy2 + qz2], -1),
        ],
        -2,
    )


def _mv(M, v):
    return np.einsum("...ij,...j->...i", M, v)


def linear_state_dot(x, u, p, nx, nu):
    """xdot = A x + B u + b (diffmpc/dynamics/linear_dynamics.py:46-70)."""
    A = p[..., : nx * nx].reshape(p.shape[:-1] + (nx, nx))
    Bm = p[..., nx * nx : nx * nx + nx * nu].reshape(p.shape[:-1] + (nx, nu))
    b = p[..., nx * nx + nx * nu : nx * nx + nx * nu + nx]
    return _mv(A, x) + _mv(Bm, u) + b


def quadrotor_state_dot(x, u, p):
    """Quadrotor, 13 states / 4 controls [thrust, tau_x, tau_y, tau_z].

    Follows diffmpc/dynamics/quadrotor_dynamics.py:84-133: the quaternion is
    normalised inside the derivative, v' = -g e3 + (f/m) R(q) e3,
    q' = 1/2 L(q) [0; w], w' = J^-1 (S(J w) w + tau).
    """
    m = p[..., 0]
    J = p[..., 1:10].reshape(p.shape[:-1] + (3, 3))
    Jinv = np.linalg.inv(J)  # complex J (complex-step in the inertia) stays complex
    v = x[..., 3:6]
    q = x[..., 6:10]
    w = x[..., 10:13]
    f = u[..., 0]
    tau = u[..., 1:4]
    qn = q / np.sqrt(np.sum(q * q, axis=-1, keepdims=True))
    R = rotation(qn)
    v_dot = R[..., :, 2] * (f / m)[..., None]
    v_dot = v_dot + np.array([0.0, 0.0, -GRAVITY])
    w4 = np.concatenate([np.zeros_like(w[..., :1]), w], -1)
    q_dot = 0.5 * _mv(q_left(qn), w4)
    Jw = _mv(J, w)
    w_dot = _mv(Jinv, np.cross(Jw, w) + tau)
    return np.concatenate([v, v_dot, q_dot, w_dot], -1)


def spacecraft3_state_dot(x, u, p):
    """w' = J^-1 (tau - w x (J w)), diagonal J (spacecraft_dynamics.py:31-66)."""
    Jd = p[..., 0:3]
    return (u - np.cross(x, Jd * x)) / Jd


def spacecraft7_state_dot(x, u, p):
    """Attitude spacecraft, 7 states [q (scalar-first), w], 3 torques.

    No reference counterpart (the reference spacecraft is rate-only): it joins
    the quaternion kinematics of quadrotor_dynamics.py:122,129 (normalised q,
    q' = 1/2 L(q) [0; w]) with Euler's equation of spacecraft_dynamics.py:63.
    """
    q = x[..., 0:4]
    w = x[..., 4:7]
    Jd = p[..., 0:3]
    qn = q / np.sqrt(np.sum(q * q, axis=-1, keepdims=True))
    w4 = np.concatenate([np.zeros_like(w[..., :1]), w], -1)
    q_dot = 0.5 * _mv(q_left(qn), w4)
    w_dot = (u - np.cross(w, Jd * w)) / Jd
    return np.concatenate([q_dot, w_dot], -1)


def cartpole_state_dot(x, u, p):
    """Cart-pole (cartpole_dynamics.py:36-80)."""
    xd, th, thd = x[..., 1], x[..., 2], x[..., 3]
    force = u[..., 0]
    mc, mp, length, g = p[..., 0], p[..., 1], p[..., 2], p[..., 3]
    total = mc + mp
    pml = mp * length
    s, c = np.sin(th), np.cos(th)
    temp = (force + pml * thd**2 * s) / total
    th_acc = (g * s - c * temp) / (length * (4.0 / 3.0 - (mp * c**2) / total))
    x_acc = temp - pml * th_acc * c / total
    return np.stack([xd, x_acc, thd, th_acc], -1)


CHAIN_MASSES = 12


def chain_state_dot(x, u, p):
    """Planar chain of 12 unit masses; no reference counterpart.

    State [p_1x, p_1y, ..., p_12x, p_12y, v_1x, ..., v_12y]; mass 1 hangs from a
    wall at the origin, link j joins mass j-1 and j with a zero-rest-length
    spring k_j and damper c_j; control u_j pushes mass j along x.
    """
    M = CHAIN_MASSES
    pos = x[..., : 2 * M].reshape(x.shape[:-1] + (M, 2))
    vel = x[..., 2 * M :].reshape(x.shape[:-1] + (M, 2))
    k = p[..., 0:M][..., None]
    c = p[..., M : 2 * M][..., None]
    zero = np.zeros_like(pos[..., :1, :])
    dpos = pos - np.concatenate([zero, pos[..., :-1, :]], -2)  # link j stretch
    dvel = vel - np.concatenate([zero, vel[..., :-1, :]], -2)
    link = -k * dpos - c * dvel  # force of link j on mass j
    acc = link - np.concatenate([link[..., 1:, :], np.zeros_like(link[..., :1, :])], -2)
    acc = acc + np.stack([u, np.zeros_like(u)], -1)
    return np.concatenate([vel.reshape(x.shape[:-1] + (2 * M,)), acc.reshape(x.shape[:-1] + (2 * M,))], -1)


class Model:
    """A dynamics model: name, sizes, state derivative."""

    def __init__(self, name, nx, nu):
        self.name = name
        self.nx = nx
        self.nu = nu
        self.model_id = MODEL_IDS[name]

    def state_dot(self, x, u, p):
        n = self.name
        if n == "linear":
            return linear_state_dot(x, u, p, self.nx, self.nu)
        if n == "quadrotor":
            return quadrotor_state_dot(x, u, p)
        if n == "spacecraft3":
            return spacecraft3_state_dot(x, u, p)
        if n == "spacecraft7":
            return spacecraft7_state_dot(x, u, p)
        if n == "cartpole":
            return cartpole_state_dot(x, u, p)
        if n == "chain":
            return chain_state_dot(x, u, p)
        raise ValueError(n)


def make_model(name, nx=None, nu=None):
    sizes = {"quadrotor": (13, 4), "spacecraft3": (3, 3), "spacecraft7": (7, 3),
             "cartpole": (4, 1), "chain": (4 * CHAIN_MASSES, CHAIN_MASSES)}
    if name == "linear":
        return Model(name, nx, nu)
    return Model(name, *sizes[name])


def predict_next_state(model, dt, scheme, p, x, u):
    """One integration step (diffmpc/dynamics/integrators.py:20-71).

    scheme 0 = Euler, 1 = midpoint, 2 = RK4.
    """
    if scheme == 0:
        return x + dt * model.state_dot(x, u, p)
    if scheme == 1:
        xm = x + 0.5 * dt * model.state_dot(x, u, p)
        return x + dt * model.state_dot(xm, u, p)
    if scheme == 2:
        k1 = model.state_dot(x, u, p)
        k2 = model.state_dot(x + 0.5 * dt * k1, u, p)
        k3 = model.state_dot(x + 0.5 * dt * k2, u, p)
        k4 = model.state_dot(x + dt * k3, u, p)
        return x + (1.0 / 6.0) * (k1 + 2 * k2 + 2 * k3 + k4) * dt
    raise ValueError(f"unknown discretization scheme {scheme}")


def next_state_and_jacobian(model, dt, scheme, p, x, u, h=1e-30):
    """f_d(x, u) and its Jacobian [A B] by complex-step differentiation.

    Mirrors what value_and_jacfwd(next_state, [x; u]) returns in
    optimal_control_problem.py:360-383 (value and d f_d / d[x; u]), using
    Im f(z + i h e_j) / h, exact to rounding for analytic f.
    Returns fd (..., nx) and J (..., nx, nx + nu).
    """
    nx, nu = model.nx, model.nu
    fd = predict_next_state(model, dt, scheme, p, x, u)
    z = np.concatenate([x, u], -1).astype(np.complex128)
    cols = []
    for j in range(nx + nu):
        zj = z.copy()
        zj[..., j] += 1j * h
        fj = predict_next_state(model, dt, scheme, p, zj[..., :nx], zj[..., nx:])
        cols.append(fj.imag / h)
    J = np.stack(cols, -1)
    return fd, J
