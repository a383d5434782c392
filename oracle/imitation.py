"""Imitation-learning loss and gradient of the oracle (float64) -- test
infrastructure, like the rest of this package.

Restates benchmarking/imitation-learning/train_imitation_diffmpc.py:55-81 and
utils.py:20-31 for a batch of demonstrations (x0_b, u_b):

  guess_b   = (Euler rollout of u_b from x0_b, u_b, 0)
  u*_b(w)   = controls of sqp_solve(problem with w_x = w, guess_b)
  loss(w)   = mean_b 0.5 mean_{t,k} (u*_b - u_b)^2
  dloss/dw  = sum_b (custom-VJP w_x gradient of instance b for the cotangent
              d loss / d u*_b = (u*_b - u_b) / (B (N+1) nu))

The state weights are shared by the batch (the reference vmaps with
in_axes=None), hence the sum over instances.  Pinned by
tests/test_oracle_imitation.py: the Euler rollout against explicit steps of
the model, and loss/gradient against central differences of the loss of
converged solves (where the implicit gradient is the true derivative).
"""
import numpy as np

from .dynamics import make_model, predict_next_state
from .sqp import sqp_backward, sqp_solve


def cartpole_problem(x0, problem_params):
    """The problem of cartpole.yaml (the YAML keys, as loaded; arrays or
    tensors) for initial states x0 (B, 4), as an oracle problem dict.  The
    dynamics parameters are rounded to float32, as the device kernels read
    them."""
    p = {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in problem_params.items()}
    x0 = np.asarray(x0, np.float64)
    B, N = x0.shape[0], int(p["horizon"])
    rep = lambda v: np.repeat(np.asarray(v, np.float64).reshape(1, -1), B, 0)  # noqa: E731
    dyn = rep([p["masscart"], p["masspole"], p["length"], p["gravity"]])
    return dict(model=make_model("cartpole"), horizon=N, dt=float(p["discretization_resolution"]),
                scheme=int(p["discretization_scheme"]), penalize_control_reference=bool(p["penalize_control_reference"]),
                x0=x0, x_final=rep(p["final_state"]),
                x_ref=np.repeat(np.asarray(p["reference_state_trajectory"], np.float64)[None], B, 0),
                u_ref=np.repeat(np.asarray(p["reference_control_trajectory"], np.float64)[None], B, 0),
                w_x=rep(p["weights_penalization_reference_state_trajectory"]),
                w_u=rep(p["weights_penalization_control_squared"]), w_f=rep(p["weights_penalization_final_state"]),
                dyn=dyn.astype(np.float32).astype(np.float64))


def euler_rollout(prob, us):
    """States (B, N+1, nx) of x_{t+1} = x_t + dt f(x_t, u_t) from prob["x0"]
    (utils.py:20-31: rolls all N+1 controls' first N)."""
    model = prob["model"]
    B, T = us.shape[0], us.shape[1]
    xs = np.zeros((B, T, model.nx))
    xs[:, 0] = prob["x0"]
    for t in range(T - 1):
        xs[:, t + 1] = predict_next_state(model, prob["dt"], 0, prob["dyn"], xs[:, t], us[:, t])
    return xs


def il_loss(prob, params, us, w):
    """(loss, solution) for shared state weights w (nx,)."""
    p = dict(prob)
    p["w_x"] = np.broadcast_to(np.asarray(w, np.float64), prob["x0"].shape).copy()
    guess = (euler_rollout(p, us), np.array(us, np.float64), np.zeros_like(prob["x_ref"]))
    sol = sqp_solve(p, params, guess=guess)
    loss = float(np.mean(0.5 * np.mean((sol["controls"] - us) ** 2, axis=(1, 2))))
    return loss, sol, p


def il_loss_and_grad(prob, params, us, w):
    """(loss, d loss / d w (nx,)) as jax.value_and_grad of batch_mean_loss."""
    loss, sol, p = il_loss(prob, params, us, w)
    B = us.shape[0]
    dl_du = (sol["controls"] - us) / (B * us.shape[1] * us.shape[2])
    g = sqp_backward(p, params, sol, np.zeros_like(sol["states"]), dl_du)
    return loss, g["w_x"].sum(axis=0)
