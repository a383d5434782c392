"""Closed-loop MPC rollout on the device.

The reference's reinforcement-learning benchmark and notebooks
(benchmarking/reinforcement-learning/benchmark_diffmpc.py:140-199,
notebooks/RL_spacecraft.ipynb cell 12) run a ``lax.scan`` over simulation
steps: solve the MPC problem from the current state, warm-started with the
previous solution; apply the first control to the simulator; accumulate the
reward; ``jax.grad`` of the summed reward with respect to the cost weights.
``Rollout`` is that loop over batched device solves with torch autograd.
Nothing in it reads back to the host, so a whole rollout -- forward, or
forward and gradient -- can be captured as one hipGraph
(``diffmpc_amd.graph.GraphedStep``).

Semantics kept from the reference:

* the solve's problem parameters, including ``initial_state`` (the simulated
  state), are stop-gradient (sqp.py:456-457): d u* / d x0 is dropped and state
  sensitivities flow only through the simulator;
* the carried solution is the next solve's initial guess and carries no
  gradient;
* ``adjoint_warm_start=True`` keeps the notebooks' adjoint warm start: the
  backward of solve k starts from the adjoint multipliers of solve k+1
  (sqp.py:748-752).  The benchmark script stop-gradients the carried solution,
  which cuts that channel; ``adjoint_warm_start=False`` reproduces it (each
  backward starts from zero).
"""
from typing import Callable, Dict, Optional

import torch

from .solvers.sqp import SQPSolution


def euler_simulator(dynamics, params: Dict, dt: float, scheme: int = 0) -> Callable:
    """x+ = one Euler (0), midpoint (1) or RK4 (2) step of the program's own
    dynamics, as the notebooks simulate (torch, differentiable)."""
    def f(x, u):
        return dynamics.state_dot(x, u, params)

    def step(x, u):
        k1 = f(x, u)
        if scheme == 0:
            return x + dt * k1
        if scheme == 1:
            return x + dt * f(x + 0.5 * dt * k1, u)
        k2 = f(x + 0.5 * dt * k1, u)
        k3 = f(x + 0.5 * dt * k2, u)
        k4 = f(x + dt * k3, u)
        return x + dt / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)

    return step


def linear_simulator(A: torch.Tensor, B: torch.Tensor, b: torch.Tensor) -> Callable:
    """x+ = A x + B u + b with shared discrete-time matrices
    (benchmark_diffmpc.py:154-157)."""
    def step(x, u):
        return x @ A.T + u @ B.T + b

    return step


class Rollout:
    """Batched closed-loop rollout of `num_sim_steps` MPC solves.

    solver, problem_params: the MPC problem (batched or shared parameters);
    simulate(x, u) -> x+; reward(x+, u) -> (B,) reward of one step;
    initial_guess(state) -> SQPSolution: the first solve's guess (default the
    solver's straight line; the quadrotor notebook, for instance, normalises
    the quaternions of that line and starts from hover thrust, cell 8).
    Calling returns (total reward (B,), states (T+1, B, nx), controls (T, B, nu)).
    """

    def __init__(self, solver, problem_params: Dict, simulate: Callable, reward: Callable, num_sim_steps: int,
                 warm_start: bool = True, adjoint_warm_start: bool = True,
                 initial_guess: Optional[Callable] = None):
        self.solver = solver
        self.problem_params = problem_params
        self.simulate = simulate
        self.reward = reward
        self.num_sim_steps = num_sim_steps
        self.warm_start = warm_start
        self.adjoint_warm_start = adjoint_warm_start
        self.initial_guess = initial_guess

    def _guess(self, state) -> SQPSolution:
        if self.initial_guess is not None:
            return self.initial_guess(state)
        return self.solver.initial_guess({**self.problem_params, "initial_state": state}, device=state.device)

    def __call__(self, x_init: torch.Tensor, weights: Optional[Dict] = None):
        weights = {} if weights is None else weights
        # expand shared parameters and weights to the batch once per rollout
        # rather than once per solve
        B, dev = x_init.shape[0], x_init.device
        problem_params = self.solver.batch_params(
            {k: v for k, v in self.problem_params.items() if k != "initial_state"}, B, dev)
        weights = self.solver.batch_params(weights, B, dev)
        state = x_init
        first = self._guess(x_init.detach())
        guess = first
        total = torch.zeros(x_init.shape[0], dtype=torch.float64, device=x_init.device)
        states, controls = [state], []
        for _ in range(self.num_sim_steps):
            if not self.adjoint_warm_start:
                guess = guess._replace(dual_backward_guess=torch.zeros_like(guess.dual))
            params = {**problem_params, "initial_state": state.detach()}
            sol = self.solver.solve(guess, params, weights)
            u0 = sol.controls[:, 0]
            state = self.simulate(state, u0)
            total = total + self.reward(state, u0)
            states.append(state)
            controls.append(u0)
            guess = sol if self.warm_start else first
        return total, torch.stack(states), torch.stack(controls)
