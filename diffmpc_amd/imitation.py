"""Cart-pole imitation learning of cost weights on the batched device solver.

The reference's imitation-learning benchmark (benchmarking/imitation-learning/)
is an end-to-end consumer of the differentiable solve:

* expert data (collect_expert_data.py:21-62): cart-pole problems of
  cartpole.yaml from random initial states, solved to convergence
  (<= 501 SQP iterations, PCG tolerance 1e-24, merit line search); the data
  are the initial states and the optimal control sequences;
* training (train_imitation_diffmpc.py:37-131): learn the state weights
  ``weights_penalization_reference_state_trajectory`` by gradient descent on
  the imitation loss mean_b 0.5 mean_t (u*_b(w) - u^demo_b)^2, where u*_b(w)
  is a short solve (<= 5 SQP iterations, tolerance 0.1, PCG 1e-14, line
  search) warm-started from the demonstration (controls = the demonstrated
  ones, states = their explicit-Euler rollout, utils.py:20-31); the step is
  w <- max(1e-10, w - lr nan_to_num(grad)) with lr = 1e-2, and the model loss
  is |w - w_true|.

Here the batch of 32 problems is one batched solve (the reference vmaps
them), the Euler rollout of the demonstrations is a device kernel
(dmpc_simulate), the gradient comes from torch autograd through the solver's
implicit-differentiation backward, and a training step -- rollout, solve,
loss, backward, update -- can be captured once and replayed as one hipGraph.
Demonstrations are stored as ``.npz`` (x0, ctrl_traj).
"""
import time
from typing import Any, Callable, Dict, Optional, Sequence

import numpy as np
import torch

from .dynamics.cartpole_dynamics import CartpoleDynamics
from .graph import GraphedStep
from .problems.optimal_control_problem import OptimalControlProblem
from .solvers.sqp import SQPSolver
from .utils.load_params import load_problem_params, load_solver_params

WEIGHT_KEY = "weights_penalization_reference_state_trajectory"
NUM_STEPS = 200
LEARNING_RATE = 1e-2

# The 20 starting points of the weight search, U(0.5, 2.5)^4 draws listed in
# benchmarking/imitation-learning/utils.py:34-58 (kept identical so that runs
# are comparable).
INITIAL_WEIGHTS = np.array([
    [2.31993711, 2.10973335, 1.27643918, 2.40638906],
    [2.16305628, 0.92978235, 0.8425473, 2.05522213],
    [0.54511473, 2.38979126, 1.82534442, 1.79607804],
    [1.52602309, 0.70958412, 2.28500543, 1.79144837],
    [1.47695067, 2.37700071, 1.10622824, 1.66791643],
    [1.77999985, 2.43335516, 2.13276163, 2.25013024],
    [1.08672352, 1.89562108, 2.44455541, 0.61613992],
    [0.91772376, 0.69279309, 1.3126017, 1.75713809],
    [2.35407733, 2.33764708, 0.71646967, 0.86020509],
    [1.98209716, 1.58648594, 1.59129719, 1.04037999],
    [1.6906858, 1.59305244, 2.1575869, 1.14175977],
    [2.25656044, 1.08790578, 1.8911752, 2.28738797],
    [1.71570255, 1.32924091, 0.6220067, 1.7408397],
    [0.95953363, 1.53826163, 1.24846342, 1.79400116],
    [1.12411295, 0.93086439, 2.40280909, 2.35137107],
    [1.99076212, 1.63976281, 1.6997142, 1.12886317],
    [2.20034932, 0.60009316, 0.79487412, 2.17139842],
    [2.07377772, 1.74829048, 1.48310797, 1.67964389],
    [0.82834224, 1.09670807, 1.86061252, 0.90516967],
    [0.75138763, 1.01916688, 0.80239714, 2.49789644],
])


def _device_params(params: Dict, device) -> Dict:
    return {k: (v.to(device) if isinstance(v, torch.Tensor) else v) for k, v in params.items()}


def sample_initial_states(num_trajs: int, seed: int = 123) -> np.ndarray:
    """(num_trajs, 4) cart-pole initial states drawn as collect_expert_data.py:35-39
    does after np.random.seed(seed) (its main uses 123): per trajectory, in
    this order, x ~ U(-0.5, 0.5), xdot ~ U(-0.5, 0.5), theta ~ U(-3.14, 3.14),
    thetadot ~ U(-1, 1)."""
    rs = np.random.RandomState(seed)
    out = np.empty((num_trajs, 4))
    for i in range(num_trajs):
        out[i, 0] = rs.uniform(low=-0.5, high=0.5)
        out[i, 1] = rs.uniform(low=-0.5, high=0.5)
        out[i, 2] = rs.uniform(low=-3.14, high=3.14)
        out[i, 3] = rs.uniform(low=-1.0, high=1.0)
    return out


def cartpole_problem(problem_params: Optional[Dict] = None) -> OptimalControlProblem:
    params = load_problem_params("cartpole.yaml") if problem_params is None else problem_params
    return OptimalControlProblem(dynamics=CartpoleDynamics(), params=params)


def expert_solver_params() -> Dict:
    """collect_expert_data.py:28-33: sqp.yaml with 501 SQP iterations, line
    search and PCG tolerance 1e-24."""
    sp = load_solver_params("sqp.yaml")
    sp["num_scp_iteration_max"] = 501
    sp["linesearch"] = True
    sp["verbose"] = sp["pcg"]["verbose"] = False
    sp["pcg"]["tol_epsilon"] = 1e-24
    return sp


def training_solver_params() -> Dict:
    """train_imitation_diffmpc.py:47-53: sqp.yaml with tolerance 0.1, 5 SQP
    iterations, PCG tolerance 1e-14 and line search."""
    sp = load_solver_params("sqp.yaml")
    sp["tol_convergence"] = 1e-1
    sp["num_scp_iteration_max"] = 5
    sp["pcg"]["tol_epsilon"] = 1e-14
    sp["linesearch"] = True
    sp["verbose"] = sp["pcg"]["verbose"] = False
    return sp


def collect_expert_data(num_trajs: int = 32, seed: int = 123, device=None, solver_params: Optional[Dict] = None,
                        initial_states: Optional[np.ndarray] = None):
    """Expert demonstrations (collect_expert_data.py:21-62) as one batched
    solve.  Returns (x0 (B, 4), controls (B, N+1, 1), solution) on the device."""
    device = torch.device("cuda") if device is None else torch.device(device)
    problem = cartpole_problem()
    sp = expert_solver_params() if solver_params is None else solver_params
    solver = SQPSolver(program=problem, params=sp)
    x0 = sample_initial_states(num_trajs, seed) if initial_states is None else np.asarray(initial_states)
    params = _device_params(problem.params, device)
    params["initial_state"] = torch.as_tensor(x0, dtype=torch.float64, device=device)
    guess = solver.initial_guess(params, device=device)
    sol = solver.solve(guess, params, {})
    return params["initial_state"], sol.controls, sol


def save_demonstrations(path: str, x0, controls) -> None:
    """Demonstrations as .npz: x0 (B, nx), ctrl_traj (B, N+1, nu)."""
    x0 = x0.detach().cpu().numpy() if isinstance(x0, torch.Tensor) else np.asarray(x0)
    u = controls.detach().cpu().numpy() if isinstance(controls, torch.Tensor) else np.asarray(controls)
    np.savez(path, x0=x0, ctrl_traj=u)


def load_dataset(path: str):
    """(x0 (B, nx), ctrl_traj (B, N+1, nu)) float64 arrays of a demonstrations
    file written by save_demonstrations (the .npz counterpart of
    utils.py:11-17)."""
    with np.load(path, allow_pickle=False) as z:
        return np.asarray(z["x0"], np.float64), np.asarray(z["ctrl_traj"], np.float64)


class ImitationLearning:
    """Weight learning from demonstrations (train_imitation_diffmpc.py:37-131).

    solver: an SQPSolver over the cart-pole problem (training_solver_params);
    initial_states (B, nx), demo_controls (B, N+1, nu): the demonstrations.
    ``step(w)`` is one gradient step of the shared state weights w (nx,)
    (grad_step_and_imitation_learning_loss) and returns (new w, loss);
    ``train`` repeats it from each starting point.
    """

    def __init__(self, solver: SQPSolver, initial_states, demo_controls, learning_rate: float = LEARNING_RATE,
                 device=None, weight_key: str = WEIGHT_KEY):
        self.solver = solver
        self.problem = solver.program
        self.device = torch.device("cuda") if device is None else torch.device(device)
        self.learning_rate = learning_rate
        self.weight_key = weight_key
        self.x0 = torch.as_tensor(initial_states, dtype=torch.float64, device=self.device).contiguous()
        self.demo_controls = torch.as_tensor(demo_controls, dtype=torch.float64, device=self.device).contiguous()
        self.params = _device_params(self.problem.params, self.device)
        self.params["initial_state"] = self.x0
        # the shared references and weights, expanded to the batch once
        self.params = solver.batch_params(self.params, self.x0.shape[0], self.device)
        self.true_weights = torch.as_tensor(self.problem.params[weight_key], dtype=torch.float64,
                                            device=self.device)
        self._graph = None

    @property
    def batch(self) -> int:
        return self.x0.shape[0]

    def demonstration_states(self) -> torch.Tensor:
        """Explicit-Euler rollout of the demonstrated controls (utils.py:20-31)."""
        return self.problem.simulate(self.demo_controls, self.params, scheme=0)

    def initial_guess(self):
        """initial_guess_from_data (train_imitation_diffmpc.py:55-62)."""
        guess = self.solver.initial_guess(self.params, device=self.device)
        return guess._replace(states=self.demonstration_states(), controls=self.demo_controls)

    def loss(self, w: torch.Tensor) -> torch.Tensor:
        """batch_mean_loss (train_imitation_diffmpc.py:64-74): the batch mean of
        0.5 mean((u* - u_demo)^2), differentiable in w."""
        sol = self.solver.solve(self.initial_guess(), self.params, {self.weight_key: w})
        per_sample = 0.5 * ((sol.controls - self.demo_controls) ** 2).mean(dim=(1, 2))
        return per_sample.mean()

    def step(self, w: torch.Tensor):
        """grad_step_and_imitation_learning_loss (train_imitation_diffmpc.py:76-81):
        (max(1e-10, w - lr nan_to_num(grad)), nan_to_num(loss))."""
        w = w.detach().to(device=self.device, dtype=torch.float64).requires_grad_(True)
        loss = self.loss(w)
        (grad,) = torch.autograd.grad(loss, [w])
        loss, grad = torch.nan_to_num(loss.detach()), torch.nan_to_num(grad)
        return torch.clamp(w.detach() - self.learning_rate * grad, min=1e-10), loss

    def model_loss(self, w: torch.Tensor) -> torch.Tensor:
        """|w - w_true| (train_imitation_diffmpc.py:83-85)."""
        return torch.linalg.norm(torch.as_tensor(w, dtype=torch.float64, device=self.device) - self.true_weights)

    def captured_step(self) -> Dict[str, Any]:
        """The hipGraph of one training step, created on first use: it reads
        the weights in `w`, writes the model loss of those weights, the
        step's imitation loss and the updated weights, and finally stores the
        updated weights back into `w`, so that consecutive replays are
        consecutive training steps with no host work in between."""
        if self._graph is None:
            nx = self.true_weights.numel()
            w = self.true_weights.clone().reshape(nx)
            w_out = torch.empty_like(w)
            loss_out = torch.empty((), dtype=torch.float64, device=self.device)
            model_out = torch.empty((), dtype=torch.float64, device=self.device)

            def fn():
                model_out.copy_(self.model_loss(w))
                w_new, loss = self.step(w)
                loss_out.copy_(loss)
                w_out.copy_(w_new)
                w.copy_(w_new)
                return w_out, loss_out, model_out

            self._graph = dict(graph=GraphedStep(fn), w=w, w_out=w_out, loss=loss_out, model_loss=model_out)
        return self._graph

    def graphed_step(self) -> Callable:
        """The step captured as one hipGraph.  Returns step(w) -> (new w, loss)
        with the same meaning as ``step``; the outputs are fresh tensors."""
        g = self.captured_step()

        def run(w):
            g["w"].copy_(torch.as_tensor(w, dtype=torch.float64, device=self.device).reshape(g["w"].shape))
            g["graph"].replay()
            return g["w_out"].clone(), g["loss"].clone()

        return run

    def train(self, initial_weights: Sequence = INITIAL_WEIGHTS, num_steps: int = NUM_STEPS, graphed: bool = True,
              log: Optional[Callable[[str], None]] = None) -> Dict[str, np.ndarray]:
        """The training loop of train_imitation_diffmpc.py:87-124.  For each
        starting point: entry 0 holds the loss and model loss at the starting
        weights and time 0, then one entry per step (model loss before the
        step, imitation loss of the step, seconds since the run's start).
        Returns model_losses, imitation_losses, times as (runs, num_steps + 1)
        arrays and the final weights (runs, nx).  With `graphed`, each step is
        one replay of the captured step (its first replay gives entry 0 and
        entry 1, which the reference computes twice)."""
        all_model, all_imit, all_times, finals = [], [], [], []
        for i, w0 in enumerate(initial_weights):
            w = torch.as_tensor(w0, dtype=torch.float64, device=self.device)
            times = [0.0]
            if graphed:
                g = self.captured_step()
                g["w"].copy_(w.reshape(g["w"].shape))
                torch.cuda.synchronize(self.device)
                s = time.monotonic()
                model, imit = [], []
                for k in range(num_steps):
                    g["graph"].replay()
                    model.append(g["model_loss"].clone())
                    imit.append(g["loss"].clone())
                    torch.cuda.synchronize(self.device)
                    times.append(time.monotonic() - s)
                    if log is not None:
                        log(f"run {i + 1}/{len(initial_weights)}, step {k}/{num_steps}: IL loss {float(imit[-1]):.6g}, "
                            f"model loss {float(model[-1]):.6g}, time {times[-1]:.3f}")
                model = [model[0]] + model
                imit = [imit[0]] + imit
                w = g["w"].clone()
            else:
                imit = [self.step(w)[1]]
                model = [self.model_loss(w)]
                torch.cuda.synchronize(self.device)
                s = time.monotonic()
                for k in range(num_steps):
                    model.append(self.model_loss(w))
                    w, loss = self.step(w)
                    imit.append(loss)
                    torch.cuda.synchronize(self.device)
                    times.append(time.monotonic() - s)
                    if log is not None:
                        log(f"run {i + 1}/{len(initial_weights)}, step {k}/{num_steps}: IL loss {float(loss):.6g}, "
                            f"model loss {float(model[-1]):.6g}, time {times[-1]:.3f}")
            all_model.append(torch.stack(model).cpu().numpy())
            all_imit.append(torch.stack(imit).cpu().numpy())
            all_times.append(np.array(times))
            finals.append(w.cpu().numpy())
        return dict(model_losses=np.array(all_model), imitation_losses=np.array(all_imit), times=np.array(all_times),
                    final_weights=np.array(finals))
