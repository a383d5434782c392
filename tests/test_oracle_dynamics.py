"""Oracle dynamics: the reference's invariants (tests/test_dynamics_quadrotor.py:26-108)
and complex-step Jacobians against central differences."""
import numpy as np
import pytest

from oracle import dynamics as D

rng = np.random.default_rng(0)


def test_skew_is_cross_product():
    a, b = rng.normal(size=3), rng.normal(size=3)
    assert np.allclose(D.skew(a) @ b, np.cross(a, b), rtol=1e-12, atol=1e-14)


def test_rotation_orthogonal_det_one():
    for _ in range(10):
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        R = D.rotation(q)
        assert np.allclose(R.T @ R, np.eye(3), atol=1e-12)
        assert np.isclose(np.linalg.det(R), 1.0, atol=1e-12)


def test_q_conj():
    q = rng.normal(size=4)
    qi = D.q_conj(q)
    assert qi[0] == q[0] and np.all(qi[1:] == -q[1:])


def test_rotation_matches_quaternion_sandwich():
    for _ in range(10):
        u = rng.normal(size=3)
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        b = D.q_left(D.q_left(q) @ np.concatenate([[0.0], u])) @ D.q_conj(q)
        assert np.allclose(D.rotation(q) @ u, b[1:], atol=1e-12)


MODELS = [
    ("quadrotor", 2, np.concatenate([[1.0], np.diag([0.01, 0.01, 0.02]).ravel()]), 0.02),
    ("spacecraft3", 0, np.array([5.0, 2.0, 1.0]), 0.1),
    ("spacecraft7", 0, np.array([0.7, 1.3, 1.9]), 0.1),
    ("cartpole", 1, np.array([1.0, 0.1, 0.5, 9.81]), 0.05),
    ("chain", 2, np.concatenate([np.linspace(0.5, 2, 12), np.linspace(0.05, 0.2, 12)]), 0.05),
]


@pytest.mark.parametrize("name,scheme,p,dt", MODELS)
def test_complex_step_matches_central_differences(name, scheme, p, dt):
    m = D.make_model(name)
    x = rng.normal(size=(3, m.nx)) * 0.5
    if name == "quadrotor":
        x[:, 6:10] += np.array([1.0, 0, 0, 0])
    if name == "spacecraft7":
        x[:, 0:4] += np.array([1.0, 0, 0, 0])
    u = rng.normal(size=(3, m.nu))
    fd, J = D.next_state_and_jacobian(m, dt, scheme, p, x, u)
    h = 1e-6
    z = np.concatenate([x, u], -1)
    for j in range(m.nx + m.nu):
        zp, zm = z.copy(), z.copy()
        zp[:, j] += h
        zm[:, j] -= h
        col = (D.predict_next_state(m, dt, scheme, p, zp[:, :m.nx], zp[:, m.nx:])
               - D.predict_next_state(m, dt, scheme, p, zm[:, :m.nx], zm[:, m.nx:])) / (2 * h)
        assert np.allclose(J[..., j], col, rtol=1e-6, atol=1e-7), (name, j)
    assert np.allclose(fd, D.predict_next_state(m, dt, scheme, p, x, u))


def test_linear_euler_jacobian_closed_form():
    nx, nu = 4, 2
    m = D.make_model("linear", nx, nu)
    Ac, Bc, b = rng.normal(size=(nx, nx)), rng.normal(size=(nx, nu)), rng.normal(size=nx)
    p = np.concatenate([Ac.ravel(), Bc.ravel(), b])
    x, u = rng.normal(size=(5, nx)), rng.normal(size=(5, nu))
    dt = 0.3
    fd, J = D.next_state_and_jacobian(m, dt, 0, p, x, u)
    assert np.allclose(J[..., :nx], np.eye(nx) + dt * Ac, atol=1e-14)
    assert np.allclose(J[..., nx:], dt * Bc, atol=1e-14)
    assert np.allclose(fd, x + dt * (x @ Ac.T + u @ Bc.T + b), atol=1e-13)


def test_chain_is_linear():
    m = D.make_model("chain")
    p = np.concatenate([np.linspace(0.5, 2, 12), np.linspace(0.05, 0.2, 12)])
    x1, x2 = rng.normal(size=(2, m.nx))
    u1, u2 = rng.normal(size=(2, m.nu))
    f = lambda x, u: m.state_dot(x, u, p)  # noqa: E731
    assert np.allclose(f(x1 + 2 * x2, u1 + 2 * u2), f(x1, u1) + 2 * f(x2, u2), atol=1e-12)
    assert np.allclose(f(np.zeros(m.nx), np.zeros(m.nu)), 0.0)
