"""Oracle QP layer against dense linear algebra and the reference's PCG invariants."""
import numpy as np
import pytest

from oracle import qp
from tests.helpers import dense_kkt, small_random_qp


def _dense_S_gamma(As_next, As, Bs, Qinv, Rinv, Cs, qs, rs):
    N, nx, nu = As.shape[0], As.shape[1], Bs.shape[2]
    qv = qs
    x, u, lam, H, G = dense_kkt(As_next, As, Bs, Qinv, qv, Rinv, rs, Cs)
    Hinv = np.linalg.inv(H)
    S = -G @ Hinv @ G.T
    qfull = np.zeros(H.shape[0])
    for t in range(N + 1):
        qfull[t * (nx + nu): t * (nx + nu) + nx] = qs[t]
        if t < N:
            qfull[t * (nx + nu) + nx: (t + 1) * (nx + nu)] = rs[t]
    gamma = Cs.reshape(-1) + G @ Hinv @ qfull
    return S, gamma, x, u, lam


def _rows_to_dense(M):
    Np1, nx = M.shape[0], M.shape[1]
    D = np.zeros((Np1 * nx, Np1 * nx))
    for t in range(Np1):
        for k, tt in enumerate((t - 1, t, t + 1)):
            if 0 <= tt < Np1:
                D[t * nx:(t + 1) * nx, tt * nx:(tt + 1) * nx] = M[t, :, k * nx:(k + 1) * nx]
    return D


@pytest.mark.parametrize("T,nx,nu", [(2, 1, 1), (5, 3, 2), (3, 4, 5)])
def test_S_and_gamma_equal_dense_schur_complement(T, nx, nu):
    As_next, As, Bs, Qinv, Rinv, Cs, qs, rs = small_random_qp(T, nx, nu, seed=T + nx)
    rng = np.random.default_rng(1)
    As = As + 0.1 * rng.normal(size=As.shape)  # general, not only diagonal
    S, Phi = qp.compute_S_Phiinv(As_next, As, Bs, Qinv, Rinv)
    AQ, AnQ, BR = qp.premultiplied(As_next, As, Bs, Qinv, Rinv)
    gam = qp.compute_gamma(AQ, AnQ, BR, Cs, Qinv[:, 0], qs, rs)
    for b in range(As.shape[0]):
        Sd, gd, *_ = _dense_S_gamma(As_next[b], As[b], Bs[b], Qinv[b], Rinv[b], Cs[b], qs[b], rs[b])
        assert np.allclose(_rows_to_dense(S[b]), Sd, atol=1e-12)
        assert np.allclose(gam[b].reshape(-1), gd, atol=1e-12)
        assert np.allclose(_rows_to_dense(S[b]), _rows_to_dense(S[b]).T, atol=1e-13)
        assert np.allclose(_rows_to_dense(Phi[b]), _rows_to_dense(Phi[b]).T, atol=1e-12)


def test_preconditioner_is_exact_inverse_without_coupling():
    # A_t = 0 makes phi = 0: S block diagonal and Phiinv == (-S)^-1
    As_next, As, Bs, Qinv, Rinv, *_ = small_random_qp(4, 3, 2, seed=3)
    S, Phi = qp.compute_S_Phiinv(As_next, np.zeros_like(As), Bs, Qinv, Rinv)
    for b in range(S.shape[0]):
        assert np.allclose(_rows_to_dense(Phi[b]), np.linalg.inv(-_rows_to_dense(S[b])), atol=1e-10)


def test_pcg_solution_equals_dense_kkt_and_recovery_is_stationary():
    As_next, As, Bs, Qinv, Rinv, Cs, qs, rs = small_random_qp(6, 3, 2, seed=5)
    S, Phi = qp.compute_S_Phiinv(As_next, As, Bs, Qinv, Rinv)
    AQ, AnQ, BR = qp.premultiplied(As_next, As, Bs, Qinv, Rinv)
    gam = qp.compute_gamma(AQ, AnQ, BR, Cs, Qinv[:, 0], qs, rs)
    lam, it, eta = qp.solve_kkt_schur(S, Phi, gam, np.zeros_like(gam), 1e-26, 1000)
    x, u = qp.recover(lam, As_next, As, Bs, Qinv, qs, Rinv, rs)
    for b in range(S.shape[0]):
        xd, ud, lamd, H, G = dense_kkt(As_next[b], As[b], Bs[b], Qinv[b], qs[b], Rinv[b], rs[b], Cs[b])
        assert np.allclose(lam[b], lamd, atol=1e-9)
        assert np.allclose(x[b], xd, atol=1e-9)
        assert np.allclose(u[b, :-1], ud, atol=1e-9)
        assert np.all(u[b, -1] == 0.0)
    # stationarity H z + q + G^T lam = 0 holds for any lam (pcg.py:316-326)
    lam_any = np.random.default_rng(0).normal(size=lam.shape)
    x, u = qp.recover(lam_any, As_next, As, Bs, Qinv, qs, Rinv, rs)
    _, _, _, H, G = dense_kkt(As_next[0], As[0], Bs[0], Qinv[0], qs[0], Rinv[0], rs[0], Cs[0])
    N, nx, nu = As.shape[1], As.shape[2], Bs.shape[3]
    z = np.concatenate([x[0], u[0]], -1).reshape(-1)
    qf = np.concatenate([qs[0], np.concatenate([rs[0], np.zeros((1, nu))])], -1).reshape(-1)
    res = H @ z + qf + G.T @ lam_any[0].reshape(-1)
    mask = np.ones((N + 1, nx + nu), bool)
    mask[N, nx:] = False  # the padded last control is not a variable
    assert np.allclose(res.reshape(N + 1, nx + nu)[mask], 0.0, atol=1e-10)


def test_pcg_warm_start_invariant():
    """tests/test_solver_pcg.py:179-227: a converged guess is kept, with fewer iterations."""
    As_next, As, Bs, Qinv, Rinv, Cs, qs, rs = small_random_qp(3, 4, 5, seed=7)
    S, Phi = qp.compute_S_Phiinv(As_next, As, Bs, Qinv, Rinv)
    AQ, AnQ, BR = qp.premultiplied(As_next, As, Bs, Qinv, Rinv)
    gam = qp.compute_gamma(AQ, AnQ, BR, Cs, Qinv[:, 0], np.zeros_like(qs), np.zeros_like(rs))
    lam, it, eta = qp.solve_kkt_schur(S, Phi, gam, np.zeros_like(gam), 1e-13, 20)
    lam2, it2, eta2 = qp.solve_kkt_schur(S, Phi, gam, lam, 1e-13, 20)
    assert np.all(it <= 10) and np.all(eta >= 0)
    assert np.allclose(lam, lam2, atol=1e-8)
    assert np.all(it2 < it) and np.all(it2 >= 1)


def test_pcg_stop_rule_counts():
    """At least one iteration, at most max_iter (pcg.py:397-403)."""
    As_next, As, Bs, Qinv, Rinv, Cs, qs, rs = small_random_qp(8, 3, 2, seed=9)
    S, Phi = qp.compute_S_Phiinv(As_next, As, Bs, Qinv, Rinv)
    AQ, AnQ, BR = qp.premultiplied(As_next, As, Bs, Qinv, Rinv)
    gam = qp.compute_gamma(AQ, AnQ, BR, Cs, Qinv[:, 0], qs, rs)
    _, it, _ = qp.solve_kkt_schur(S, Phi, gam, np.zeros_like(gam), 1e300, 50)
    assert np.all(it == 1)
    _, it, _ = qp.solve_kkt_schur(S, Phi, gam, np.zeros_like(gam), 0.0, 3)
    assert np.all(it == 3)
