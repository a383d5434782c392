"""Schur-complement PCG QP solver of the oracle (batched, float64).

Restates diffmpc/solvers/pcg.py.  Every array carries a leading batch axis B.
Shapes (N = horizon):
  As_next, As (B, N, nx, nx); Bs (B, N, nx, nu); Cs (B, N+1, nx)
  Qinv (B, N+1, nx, nx); qvec (B, N+1, nx); Rinv (B, N, nu, nu); rvec (B, N, nu)
  S, Phiinv (B, N+1, nx, 3 nx): row t holds the blocks of columns t-1, t, t+1.
"""
import numpy as np


def _T(M):
    return np.swapaxes(M, -1, -2)


def _mv(M, v):
    return np.einsum("...ij,...j->...i", M, v)


def compute_S_Phiinv(As_next, As, Bs, Qinv, Rinv):
    """Schur complement S = -G H^-1 G^T and its block-Jacobi-stair inverse.

    Follows diffmpc/solvers/pcg.py:160-243:
      theta_0 = Qinv_0,
      theta_{t+1} = A_t Qinv_t A_t^T + B_t Rinv_t B_t^T + An_t Qinv_{t+1} An_t^T,
      phi_{t+1} = A_t Qinv_t An_{t-1}^T (An_{-1} = I),
      S row t = -[phi_t, theta_t, phi_{t+1}^T]  (phi_0 = phi_{N+1} = 0),
      Phiinv row t = [-th_t^-1 phi_t th_{t-1}^-1, th_t^-1, -th_t^-1 phi_{t+1}^T th_{t+1}^-1].
    """
    Bsz, N, nx = As.shape[0], As.shape[1], As.shape[2]
    eye = np.broadcast_to(np.eye(nx, dtype=As.dtype), (Bsz, 1, nx, nx))
    An_prev = np.concatenate([eye, As_next[:, :-1]], axis=1)
    AQ = As @ Qinv[:, :-1]
    theta_rest = AQ @ _T(As) + (Bs @ Rinv) @ _T(Bs) + (As_next @ Qinv[:, 1:]) @ _T(As_next)
    phi = AQ @ _T(An_prev)
    theta = np.concatenate([Qinv[:, :1], theta_rest], axis=1)
    theta_inv = np.linalg.inv(theta)
    zeros = np.zeros((Bsz, 1, nx, nx), dtype=As.dtype)
    phis = np.concatenate([zeros, phi, zeros], axis=1)  # phis[k] = phi_k, k = 0..N+1
    S = -np.concatenate([phis[:, :-1], theta, _T(phis[:, 1:])], axis=-1)
    tinv = np.concatenate([zeros, theta_inv, zeros], axis=1)
    left = -(theta_inv @ phis[:, :-1]) @ tinv[:, :-2]
    right = -(theta_inv @ _T(phis[:, 1:])) @ tinv[:, 2:]
    Phiinv = np.concatenate([left, theta_inv, right], axis=-1)
    return S, Phiinv


def premultiplied(As_next, As, Bs, Qinv, Rinv):
    """A_t Qinv_t, An_t Qinv_{t+1}, B_t Rinv_t (diffmpc/solvers/sqp.py:247-251)."""
    return As @ Qinv[:, :-1], As_next @ Qinv[:, 1:], Bs @ Rinv


def compute_gamma(As_x_Qinv, Asnext_x_Qinvnext, Bs_x_Rinv, Cs, Q0inv, qs, rs):
    """gamma = C + G H^-1 q (diffmpc/solvers/pcg.py:245-288).

    gamma_0 = C_0 + Qinv_0 q_0;
    gamma_{t+1} = C_{t+1} + A_t Qinv_t q_t + B_t Rinv_t r_t + An_t Qinv_{t+1} q_{t+1}.
    """
    zeta = _mv(As_x_Qinv, qs[:, :-1]) + _mv(Bs_x_Rinv, rs) + _mv(Asnext_x_Qinvnext, qs[:, 1:])
    head = _mv(Q0inv, qs[:, 0])[:, None]
    return Cs + np.concatenate([head, zeta], axis=1)


def block_tridiag_matvec(M, v):
    """y_t = M_t [v_{t-1}; v_t; v_{t+1}] with zero padding (pcg.py:370-392, 405-415)."""
    z = np.zeros_like(v[:, :1])
    vp = np.concatenate([z, v, z], axis=1)
    vs = np.concatenate([vp[:, :-2], vp[:, 1:-1], vp[:, 2:]], axis=-1)
    return _mv(M, vs)


def solve_kkt_schur(S, Phiinv, gamma, lam0, tol_epsilon, max_iter):
    """Block-Jacobi-preconditioned CG on S lam = gamma (pcg.py:350-445).

    Each instance runs its own loop: at least one iteration, then continue
    while |eta| > eps and it <= max_iter - 1 (pcg.py:397-403); finished
    instances are frozen.  Returns lam (B, N+1, nx), iterations (B,), eta (B,).
    """
    Bsz = S.shape[0]
    lam = lam0.copy()
    r = gamma - block_tridiag_matvec(S, lam)
    rt = block_tridiag_matvec(Phiinv, r)
    p = rt.copy()
    eta = np.sum(r * rt, axis=(1, 2))
    it = np.zeros(Bsz, dtype=np.int64)

    def active_mask():
        cont = (np.abs(eta) > tol_epsilon) & (it <= max_iter - 1)
        return cont | (it < 1)

    act = active_mask()
    while np.any(act):
        a = act[:, None, None]
        ups = block_tridiag_matvec(S, p)
        v = np.sum(p * ups, axis=(1, 2))
        with np.errstate(divide="ignore", invalid="ignore"):
            alpha = np.where(act, eta / np.where(act, v, 1.0), 0.0)
        lam_n = lam + alpha[:, None, None] * p
        r_n = r - alpha[:, None, None] * ups
        rt_n = block_tridiag_matvec(Phiinv, r_n)
        eta_n = np.sum(r_n * rt_n, axis=(1, 2))
        with np.errstate(divide="ignore", invalid="ignore"):
            beta = np.where(act, eta_n / np.where(act, eta, 1.0), 0.0)
        p_n = rt_n + beta[:, None, None] * p
        lam = np.where(a, lam_n, lam)
        r = np.where(a, r_n, r)
        p = np.where(a, p_n, p)
        eta = np.where(act, eta_n, eta)
        it = it + act.astype(np.int64)
        act = active_mask()
    return lam, it, eta


def recover(lam, As_next, As, Bs, Qinv, qvec, Rinv, rvec):
    """Primal (x, u) from multipliers (pcg.py:290-348).

    x_t = -Qinv_t (q_t + An_{t-1}^T lam_t + A_t^T lam_{t+1}), t < N (An_{-1} = I);
    x_N = -Qinv_N (q_N + An_{N-1}^T lam_N); u_t = -Rinv_t (r_t + B_t^T lam_{t+1});
    the last control row is zero.
    """
    Bsz, N, nx = As.shape[0], As.shape[1], As.shape[2]
    nu = Bs.shape[-1]
    eye = np.broadcast_to(np.eye(nx, dtype=As.dtype), (Bsz, 1, nx, nx))
    An_prev = np.concatenate([eye, As_next[:, :-1]], axis=1)
    xs = -_mv(Qinv[:, :-1], qvec[:, :-1] + _mv(_T(An_prev), lam[:, :-1]) + _mv(_T(As), lam[:, 1:]))
    xN = -_mv(Qinv[:, -1], qvec[:, -1] + _mv(_T(As_next[:, -1]), lam[:, -1]))
    us = -_mv(Rinv, rvec + _mv(_T(Bs), lam[:, 1:]))
    states = np.concatenate([xs, xN[:, None]], axis=1)
    controls = np.concatenate([us, np.zeros((Bsz, 1, nu), dtype=us.dtype)], axis=1)
    return states, controls
