"""Closed-loop MPC rollout and its gradient in the oracle (float64).

Restates the rollout of the reference's RL examples
(notebooks/RL_spacecraft.ipynb cell 12, notebooks/RL_quadrotor.ipynb cell 12,
benchmarking/reinforcement-learning/benchmark_diffmpc.py:140-199): at every
simulation step the MPC problem is solved from the current state, warm-started
with the previous solution; the first control is applied to the simulator and
the reward accumulates.  The gradient of the summed reward with respect to the
cost weights follows jax reverse mode through that scan:

* the solve's initial_state and initial guess are stop-gradient (sqp.py:456-457),
  so d u* / d x0 is dropped and only the explicit simulator path carries state
  sensitivities;
* each solve contributes its custom-VJP weight gradient for the cotangent
  d L / d controls[0];
* the adjoint warm start of solve k is lam~ of solve k+1, the cotangent that
  reaches solve k's dual_backward_guess output (sqp.py:748-752, 794).
"""
import numpy as np

from .dynamics import next_state_and_jacobian, predict_next_state
from .sqp import initial_guess, sqp_backward, sqp_solve


def rollout(prob, params, x_init, num_sim_steps, sim_scheme, reward, guess_fn=None):
    """Forward rollout; returns (total reward per instance, per-step records)."""
    model = prob["model"]
    state = np.array(x_init, dtype=np.float64)
    p = dict(prob)
    p["x0"] = state.copy()
    x, u, lam = (guess_fn or initial_guess)(p)
    hist = None
    total = np.zeros(state.shape[0])
    records = []
    for _ in range(num_sim_steps):
        p = dict(prob)
        p["x0"] = state.copy()
        sol = sqp_solve(p, params, guess=(x, u, lam), histories=hist)
        u0 = sol["controls"][:, 0].copy()
        nxt = predict_next_state(model, prob["dt"], sim_scheme, prob["dyn"], state, u0)
        total = total + reward(nxt, u0)
        records.append((p, sol, state.copy(), u0, nxt))
        state = nxt
        x, u, lam = sol["states"], sol["controls"], sol["dual"]
        hist = (sol["convergence_errors"], sol["linesearch_alphas"], sol["num_pcg_iters"], sol["convergence_error"])
    return total, records


def rollout_gradient(prob, params, records, sim_scheme, reward_grad, keys=("w_x", "w_u", "w_f")):
    """d (sum of rewards over steps and instances) / d shared weights."""
    model = prob["model"]
    nx, nu = model.nx, model.nu
    B = records[0][2].shape[0]
    N = prob["horizon"]
    a_next = np.zeros((B, nx))
    lam_guess = np.zeros((B, N + 1, nx))
    grads = {k: 0.0 for k in keys}
    for p, sol, xk, uk, xk1 in reversed(records):
        dr_dx1, dr_du = reward_grad(xk1, uk)
        dl_dx1 = dr_dx1 + a_next
        _, J = next_state_and_jacobian(model, prob["dt"], sim_scheme, prob["dyn"], xk, uk)
        fx, fu = J[..., :nx], J[..., nx:]
        dl_du0 = dr_du + np.einsum("bi,bij->bj", dl_dx1, fu)
        a_next = np.einsum("bi,bij->bj", dl_dx1, fx)
        dl_dx = np.zeros((B, N + 1, nx))
        dl_du = np.zeros((B, N + 1, nu))
        dl_du[:, 0] = dl_du0
        g = sqp_backward(p, params, sol, dl_dx, dl_du, lam_guess)
        lam_guess = g["dual_backward"]
        for k in keys:
            grads[k] = grads[k] + np.sum(g[k], axis=0)
    return grads
