"""SQP driver and implicit-differentiation backward of the oracle (batched, float64).

Restates diffmpc/solvers/sqp.py and the parts of
diffmpc/problems/optimal_control_problem.py it calls.  A *problem* is a dict of
batched numpy arrays (leading axis B):

  model, horizon N, dt, scheme, penalize_control_reference,
  x0 (B,nx), x_final (B,nx), x_ref (B,N+1,nx), u_ref (B,N+1,nu),
  w_x (B,nx), w_u (B,nu), w_f (B,nx), dyn (B,P)

Solver parameters use the reference's keys (diffmpc/solvers/params/sqp.yaml).
"""
import numpy as np

from . import qp
from .dynamics import next_state_and_jacobian, predict_next_state

SUCCESS = 0
ERROR = -100
PSD_MIN_EIG = 1e-9  # sqp.py:191-196


def _diag(v):
    return v[..., :, None] * np.eye(v.shape[-1], dtype=v.dtype)


def project_psd(M, minimum_eigenvalue=0.0):
    """Eigenvalue clamp then symmetrise (diffmpc/utils/jax_utils.py:75-95)."""
    S, V = np.linalg.eigh(M)
    S = np.maximum(S, minimum_eigenvalue)
    P = V @ (S[..., :, None] * np.swapaxes(V, -1, -2))
    return 0.5 * (P + np.swapaxes(P, -1, -2))


def initial_guess(prob):
    """Straight line from x0 to x_final plus 1e-6; controls 1e-6 (ocp.py:82-102)."""
    N = prob["horizon"]
    t = np.arange(N + 1, dtype=np.float64)[None, :, None]
    a1 = (N - t) / N
    a2 = t / N
    x = prob["x0"][:, None, :] * a1 + prob["x_final"][:, None, :] * a2 + 1e-6
    B = x.shape[0]
    u = np.zeros((B, N + 1, prob["model"].nu)) + 1e-6
    lam = np.zeros((B, N + 1, prob["model"].nx))
    return x, u, lam


def _uref(prob):
    if prob["penalize_control_reference"]:
        return prob["u_ref"]
    return np.zeros_like(prob["u_ref"])


def cost(prob, states, controls):
    """Diagonal weighted tracking cost, no 1/2 (ocp.py:153-247)."""
    N = prob["horizon"]
    wx, wu, wf = prob["w_x"][:, None], prob["w_u"][:, None], prob["w_f"]
    c = np.sum(wx * (states[:, :N] - prob["x_ref"][:, :N]) ** 2, axis=(1, 2))
    c = c + np.sum(wu * (controls[:, :N] - _uref(prob)[:, :N]) ** 2, axis=(1, 2))
    c = c + np.sum((prob["w_x"] + wf) * (states[:, N] - prob["x_final"]) ** 2, axis=1)
    return c


def cost_gradient(prob, states, controls):
    """d cost / d(states, controls); last control row has zero gradient."""
    N = prob["horizon"]
    gx = np.zeros_like(states)
    gu = np.zeros_like(controls)
    gx[:, :N] = 2 * prob["w_x"][:, None] * (states[:, :N] - prob["x_ref"][:, :N])
    gx[:, N] = 2 * (prob["w_x"] + prob["w_f"]) * (states[:, N] - prob["x_final"])
    gu[:, :N] = 2 * prob["w_u"][:, None] * (controls[:, :N] - _uref(prob)[:, :N])
    return gx, gu


def equality_constraints(prob, states, controls):
    """g = [x_0 - x_init; f_d(x_t, u_t) - x_{t+1}] flattened (ocp.py:104-151)."""
    N = prob["horizon"]
    m = prob["model"]
    nxt = predict_next_state(m, prob["dt"], prob["scheme"], prob["dyn"][:, None], states[:, :N], controls[:, :N])
    g0 = states[:, 0] - prob["x0"]
    gd = nxt - states[:, 1:]
    return np.concatenate([g0, gd.reshape(gd.shape[0], -1)], axis=1)


def cost_terms(prob, states, controls):
    """Q, q, R, r of the quadratic model of the cost (sqp.py:160-197, ocp.py:249-323).

    Hessians are 2 diag(w) (terminal 2 diag(w_x + w_f)); q = grad - H xbar uses
    the raw Hessian, then Q and R are projected onto eigenvalues >= 1e-9.
    """
    N = prob["horizon"]
    gx, gu = cost_gradient(prob, states, controls)
    B = states.shape[0]
    wx = np.broadcast_to(prob["w_x"][:, None], (B, N, states.shape[-1]))
    Hx = _diag(2 * wx)
    HN = _diag(2 * (prob["w_x"] + prob["w_f"]))
    Hu = _diag(2 * np.broadcast_to(prob["w_u"][:, None], (B, N, controls.shape[-1])))
    q = gx[:, :N] - np.einsum("btj,btji->bti", states[:, :N], Hx)
    qN = gx[:, N] - np.einsum("bj,bji->bi", states[:, N], HN)
    r = gu[:, :N] - np.einsum("btj,btji->bti", controls[:, :N], Hu)
    Qs = np.concatenate([Hx, HN[:, None]], axis=1)
    qs = np.concatenate([q, qN[:, None]], axis=1)
    Qs = project_psd(Qs, PSD_MIN_EIG)
    Rs = project_psd(Hu, PSD_MIN_EIG)
    return Qs, qs, Rs, r


def linearize(prob, states, controls):
    """As_next = -I, A_t, B_t, C = [x0; -f_d + A xbar + B ubar] (ocp.py:325-398)."""
    N = prob["horizon"]
    m = prob["model"]
    fd, J = next_state_and_jacobian(m, prob["dt"], prob["scheme"], prob["dyn"][:, None], states[:, :N], controls[:, :N])
    As = J[..., : m.nx]
    Bs = J[..., m.nx :]
    C = -fd + qp._mv(As, states[:, :N]) + qp._mv(Bs, controls[:, :N])
    Cs = np.concatenate([prob["x0"][:, None], C], axis=1)
    B = states.shape[0]
    As_next = -np.broadcast_to(np.eye(m.nx), (B, N, m.nx, m.nx)).copy()
    return As_next, As, Bs, Cs


def qp_parameters(prob, states, controls):
    """Everything one SQP iteration needs (sqp.py:230-260)."""
    Qs, qs, Rs, rs = cost_terms(prob, states, controls)
    As_next, As, Bs, Cs = linearize(prob, states, controls)
    Qinv = np.linalg.inv(Qs)
    Rinv = np.linalg.inv(Rs)
    S, Phiinv = qp.compute_S_Phiinv(As_next, As, Bs, Qinv, Rinv)
    return dict(As_next=As_next, As=As, Bs=Bs, Cs=Cs, Qinv=Qinv, qvec=qs, Rinv=Rinv, rvec=rs, S=S, Phiinv=Phiinv)


def solve_qp(qpp, lam_guess, pcg_params):
    """gamma, PCG, recovery (sqp.py:262-299)."""
    AQ, AnQ, BR = qp.premultiplied(qpp["As_next"], qpp["As"], qpp["Bs"], qpp["Qinv"], qpp["Rinv"])
    gamma = qp.compute_gamma(AQ, AnQ, BR, qpp["Cs"], qpp["Qinv"][:, 0], qpp["qvec"], qpp["rvec"])
    lam, iters, eta = qp.solve_kkt_schur(qpp["S"], qpp["Phiinv"], gamma, lam_guess, pcg_params["tol_epsilon"], pcg_params["max_iter"])
    x, u = qp.recover(lam, qpp["As_next"], qpp["As"], qpp["Bs"], qpp["Qinv"], qpp["qvec"], qpp["Rinv"], qpp["rvec"])
    return x, u, lam, iters, eta, gamma


def linesearch(prob, params, states, controls, new_states, new_controls):
    """Merit-function backtracking of sqp.py:301-411 (Nocedal & Wright Alg. 18.3).

    Returns the step alpha per instance.  Candidates are scanned from the
    largest; out-of-range indices clamp as JAX gathers do (sqp.py:390-403).
    """
    eta = params["linesearch_eta"]
    alphas = np.asarray(params["linesearch_alphas"], dtype=np.float64)
    dx = new_states - states
    du = new_controls - controls
    c0 = cost(prob, states, controls)
    gx, gu = cost_gradient(prob, states, controls)
    dcdz = np.sum(gx * dx, axis=(1, 2)) + np.sum(gu * du, axis=(1, 2))
    g1 = np.sum(np.abs(equality_constraints(prob, states, controls)), axis=1)
    with np.errstate(divide="ignore", invalid="ignore"):
        mu = np.maximum(0.1, dcdz / (0.5 * g1))
    merit = c0 + mu * g1
    dmerit = dcdz - mu * g1
    dec = []
    for a in alphas:
        xs = states + a * dx
        us = controls + a * du
        m = cost(prob, xs, us) + mu * np.sum(np.abs(equality_constraints(prob, xs, us)), axis=1)
        dec.append(m - (merit + eta * a * dmerit))
    dec = np.stack(dec, axis=1)  # (B, n)
    return np.array([select_alpha(dec[b], alphas) for b in range(dec.shape[0])])


def select_alpha(decrease, alphas):
    """The step choice of sqp.py:389-403 for one instance.

    Start at index len+1 with alpha = max(alphas); while decrease[i] >= 0 and
    i != 0 step down, alpha = alphas[i-1].  Indices past the end clamp to the
    last entry, as JAX gathers do, so the last candidate is tested up to three
    times before alphas[-2]; when nothing passes, alphas[0] is returned.
    """
    n = len(alphas)
    i, alpha = n + 1, float(np.max(alphas))
    while not (decrease[min(i, n - 1)] < 0 or i == 0):
        alpha = float(alphas[min(i - 1, n - 1)])
        i = i - 1
    return alpha


def sqp_solve(prob, params, guess=None, histories=None):
    """The SQP loop of sqp.py:436-560 with per-instance stopping.

    guess = (x, u, lam) or None for the straight-line guess.  Returns a dict
    with the reference's SQPSolution fields plus the QP data of the last
    linearization point (what the backward uses, sqp.py:487-492, 512-518).
    """
    if guess is None:
        guess = initial_guess(prob)
    x, u, lam = (np.array(g, dtype=np.float64) for g in guess)
    B = x.shape[0]
    maxit = params["num_scp_iteration_max"]
    tol = params["tol_convergence"]
    if histories is None:
        conv_errs = np.zeros((B, maxit))
        ls_alphas = np.zeros((B, maxit))
        pcg_iters = np.zeros((B, maxit), dtype=np.int64)
        conv_err = np.zeros(B)
    else:
        conv_errs, ls_alphas, pcg_iters, conv_err = (np.array(h) for h in histories)
    num_iter = np.zeros(B, dtype=np.int64)
    saved = None
    while True:
        act = ((num_iter < maxit) & (conv_err > tol)) | (num_iter < 1)
        if not np.any(act):
            break
        qpp = qp_parameters(prob, x, u)
        xn, un, lamn, iters, _, _ = solve_qp(qpp, lam, params["pcg"])
        if params["linesearch"]:
            alpha = linesearch(prob, params, x, u, xn, un)
            xn = x + alpha[:, None, None] * (xn - x)
            un = u + alpha[:, None, None] * (un - u)
            ls_alphas[np.arange(B)[act], num_iter[act]] = alpha[act]
        err = np.maximum(np.max(np.abs(xn - x), axis=(1, 2)), np.max(np.abs(un - u), axis=(1, 2)))
        a3 = act[:, None, None]
        x = np.where(a3, xn, x)
        u = np.where(a3, un, u)
        lam = np.where(a3, lamn, lam)
        conv_err = np.where(act, err, conv_err)
        idx = np.arange(B)[act]
        conv_errs[idx, num_iter[act]] = err[act]
        pcg_iters[idx, num_iter[act]] = iters[act]
        if saved is None:
            saved = {k: v.copy() for k, v in qpp.items()}
        else:
            for k, v in qpp.items():
                sel = act.reshape((B,) + (1,) * (v.ndim - 1))
                saved[k] = np.where(sel, v, saved[k])
        num_iter = num_iter + act.astype(np.int64)
    bad = np.isnan(x).any(axis=(1, 2)) | np.isnan(u).any(axis=(1, 2)) | (conv_err > tol)
    status = np.where(bad, ERROR, SUCCESS)
    return dict(
        states=np.nan_to_num(x), controls=np.nan_to_num(u), dual=np.nan_to_num(lam),
        status=status, num_iter=num_iter, convergence_error=conv_err,
        convergence_errors=conv_errs, linesearch_alphas=ls_alphas, num_pcg_iters=pcg_iters,
        saved=saved,
    )


def sqp_backward(prob, params, sol, dl_dx, dl_du, lam_guess=None):
    """Adjoint backward of the custom VJP (sqp.py:605-798) in closed form.

    Solves the KKT system of the last linearization with q = -dl/dx,
    r = -dl/du[:-1], C = 0 by PCG on the saved S (sqp.py:668-708), recovers
    (x~, u~), and contracts with the mixed derivatives of the diagonal cost:
      dL/dw_x[i]  = -sum_{t<N} 2 (x_t - ref_t)_i x~_t,i - 2 (x_N - xf)_i x~_N,i
      dL/dw_f[i]  = -2 (x_N - xf)_i x~_N,i
      dL/dw_u[k]  = -sum_{t<N} 2 (u_t - uref_t)_k u~_t,k
      dL/dx_ref_t = 2 w_x x~_t (t < N; the terminal knot uses x_final)
      dL/dx_final = 2 (w_x + w_f) x~_N
      dL/du_ref_t = 2 w_u u~_t (t < N, only if the control reference is penalised)
      dL/dx0      = lam~_0   (g_0 = x_0 - x0)
    Returns a dict of per-instance gradients and lam~.
    """
    sv = sol["saved"]
    N = prob["horizon"]
    B = dl_dx.shape[0]
    qs = -dl_dx
    rs = -dl_du[:, :-1]
    AQ, AnQ, BR = qp.premultiplied(sv["As_next"], sv["As"], sv["Bs"], sv["Qinv"], sv["Rinv"])
    gamma = qp.compute_gamma(AQ, AnQ, BR, np.zeros_like(sv["Cs"]), sv["Qinv"][:, 0], qs, rs)
    if lam_guess is None or not params["warm_start_backward"]:
        lam_guess = np.zeros_like(sol["dual"])
    lamt, iters, _ = qp.solve_kkt_schur(sv["S"], sv["Phiinv"], gamma, lam_guess, params["pcg"]["tol_epsilon"], params["pcg"]["max_iter"])
    xt, ut = qp.recover(lamt, sv["As_next"], sv["As"], sv["Bs"], sv["Qinv"], qs, sv["Rinv"], rs)
    x, u = sol["states"], sol["controls"]
    ex = x[:, :N] - prob["x_ref"][:, :N]
    eN = x[:, N] - prob["x_final"]
    eu = u[:, :N] - _uref(prob)[:, :N]
    g = {}
    g["w_x"] = -(np.sum(2 * ex * xt[:, :N], axis=1) + 2 * eN * xt[:, N])
    g["w_f"] = -(2 * eN * xt[:, N])
    g["w_u"] = -np.sum(2 * eu * ut[:, :N], axis=1)
    gref = np.zeros_like(prob["x_ref"])
    gref[:, :N] = 2 * prob["w_x"][:, None] * xt[:, :N]
    g["x_ref"] = gref
    g["x_final"] = 2 * (prob["w_x"] + prob["w_f"]) * xt[:, N]
    guref = np.zeros_like(prob["u_ref"])
    if prob["penalize_control_reference"]:
        guref[:, :N] = 2 * prob["w_u"][:, None] * ut[:, :N]
    g["u_ref"] = guref
    g["x0"] = lamt[:, 0].copy()
    g["dual_backward"] = lamt
    g["pcg_iters"] = iters
    g["x_tilde"] = xt
    g["u_tilde"] = ut
    return g


def dynamics_param_gradient(prob, sol, g, rel_step=1e-5):
    """dL/d(dynamics parameters) (B, P) of the custom VJP for dynamics
    parameters among the weights (sqp.py:605-660 mixed derivatives, :714-737
    contraction).  The cost does not depend on them and of the constraints
    only the dynamics rows g_{t+1} = f_d(x_t, u_t; p) - x_{t+1} do
    (ocp.py:104-151), so

      dL/dp_j = -sum_{t<N} [ lam_{t+1} . d/dp_j (J_t z~_t) + lam~_{t+1} . d f_d/dp_j ]

    with J_t = d f_d / d(x_t, u_t) at the solution, lam its multipliers and
    (z~, lam~) the adjoint solution of sqp_backward.  d f_d/dp_j is taken by
    complex step (exact to rounding); d/dp_j (J_t z~_t) by central differences
    in p_j (step rel_step * max(1, |p_j|)) of the complex-step directional
    derivative J_t z~_t, accurate to about 1e-9 relative.
    """
    N = prob["horizon"]
    m = prob["model"]
    x, u, lam = sol["states"][:, :N], sol["controls"][:, :N], sol["dual"]
    xt, ut, lamt = g["x_tilde"][:, :N], g["u_tilde"][:, :N], g["dual_backward"]
    p = np.asarray(prob["dyn"], dtype=np.float64)
    B, P = p.shape
    h = 1e-30

    def f_d(pp, xx, uu):
        return predict_next_state(m, prob["dt"], prob["scheme"], pp[:, None], xx, uu)

    def jz(pp):
        return f_d(pp, x + 1j * h * xt, u + 1j * h * ut).imag / h

    out = np.zeros((B, P))
    for j in range(P):
        pc = p.astype(np.complex128)
        pc[:, j] += 1j * h
        dfdp = f_d(pc, x, u).imag / h
        d = rel_step * np.maximum(1.0, np.abs(p[:, j]))
        hi, lo = p.copy(), p.copy()
        hi[:, j] += d
        lo[:, j] -= d
        mixed = (jz(hi) - jz(lo)) / (2.0 * d[:, None, None])
        out[:, j] = -(np.sum(lam[:, 1:] * mixed, axis=(1, 2)) + np.sum(lamt[:, 1:] * dfdp, axis=(1, 2)))
    return out
