"""Optimal control problem (diffmpc/problems/optimal_control_problem.py:16-398).

Holds the dynamics and the parameter dict and offers the host-side pieces of
the reference class: sizes, the straight-line initial guess, the cost and the
equality constraints (torch, any device, batched over leading axes).  The
linearization itself runs on the GPU inside the solver; ``linearize`` exposes
it for parity checks.
"""
import copy
from typing import Any, Dict, Tuple

import torch

from .. import _lib
from ..dynamics.base_dynamics import Dynamics, batched
from ..dynamics.integrators import DiscretizationScheme, predict_next_state
from ..utils.load_params import check_parameters_dictionary_or_raise_errors

DYNAMICS_PARAM_KEYS = {
    "linear": ("A", "B", "b"),
}


def _as_tensor(v, device=None, dtype=torch.float64):
    return torch.as_tensor(v, dtype=dtype, device=device)


class OptimalControlProblem:
    """Optimal control problem class."""

    def __init__(self, dynamics: Dynamics, params: Dict[str, Any] = None, check_parameters_are_valid: bool = True):
        self._dynamics = dynamics
        if check_parameters_are_valid:
            check_parameters_dictionary_or_raise_errors(params)
        params = copy.deepcopy(params)
        self._params = params
        self._discretization_scheme = DiscretizationScheme(int(params["discretization_scheme"]))
        self._num_variables = (self.num_control_variables + self.num_state_variables) * (self.horizon + 1)
        self._num_constraints = self.num_state_variables * (self.horizon + 1)
        self.name = type(self).__name__

    @property
    def dynamics(self) -> Dynamics:
        return self._dynamics

    @property
    def params(self) -> Dict:
        return self._params

    @property
    def horizon(self) -> int:
        return int(self.params["horizon"])

    @property
    def discretization_scheme(self) -> DiscretizationScheme:
        return self._discretization_scheme

    @property
    def num_variables(self) -> int:
        return self._num_variables

    @property
    def num_constraints(self) -> int:
        return self._num_constraints

    @property
    def num_state_variables(self) -> int:
        return self.dynamics.num_states

    @property
    def num_control_variables(self) -> int:
        return self.dynamics.num_controls

    def initial_guess(self, params: Dict[str, Any] = None, device=None) -> Tuple[torch.Tensor, torch.Tensor]:
        """Straight line from initial to final state plus 1e-6; controls 1e-6 (ocp.py:82-102).

        A batched ``initial_state`` (B, nx) gives (B, N+1, nx) / (B, N+1, nu).
        """
        if params is None:
            params = self.params
        x0 = _as_tensor(params["initial_state"], device)
        xf = _as_tensor(params["final_state"], device)
        N = self.horizon
        if params.get("horizons") is not None:
            return self._ragged_initial_guess(x0, xf, params["horizons"])
        t = torch.arange(N + 1, dtype=torch.float64, device=x0.device).unsqueeze(-1)
        a1 = (N - t) / N
        a2 = t / N
        if x0.dim() == 2 or xf.dim() == 2:
            x0 = x0 if x0.dim() == 2 else x0.unsqueeze(0)
            xf = xf if xf.dim() == 2 else xf.unsqueeze(0)
            states = x0.unsqueeze(1) * a1 + xf.unsqueeze(1) * a2 + 1e-6
            controls = torch.zeros(states.shape[0], N + 1, self.num_control_variables, dtype=torch.float64,
                                   device=x0.device) + 1e-6
        else:
            states = x0 * a1 + xf * a2 + 1e-6
            controls = torch.zeros(N + 1, self.num_control_variables, dtype=torch.float64, device=x0.device) + 1e-6
        return states, controls

    def _ragged_initial_guess(self, x0, xf, horizons):
        """The straight-line guess over each instance's own horizon N_b (knots
        0..N_b); knots past N_b of the padded (B, N+1, .) arrays are zero."""
        N = self.horizon
        x0 = x0 if x0.dim() == 2 else x0.unsqueeze(0)
        xf = xf if xf.dim() == 2 else xf.unsqueeze(0)
        B = max(x0.shape[0], xf.shape[0])
        Nb = torch.as_tensor(horizons, device=x0.device).reshape(B, 1, 1).to(torch.float64)
        t = torch.arange(N + 1, dtype=torch.float64, device=x0.device).reshape(1, N + 1, 1)
        inside = (t <= Nb).to(torch.float64)
        states = (x0.unsqueeze(1) * ((Nb - t) / Nb) + xf.unsqueeze(1) * (t / Nb) + 1e-6) * inside
        controls = (torch.zeros(B, N + 1, self.num_control_variables, dtype=torch.float64, device=x0.device)
                    + 1e-6) * inside
        return states, controls

    def equality_constraints(self, states, controls, params: Dict[str, Any]) -> torch.Tensor:
        """g = [x_0 - x_init; f_d(x_t, u_t) - x_{t+1}] flattened (ocp.py:104-151)."""
        N = self.horizon
        x0 = _as_tensor(params["initial_state"], states.device)
        nxt = predict_next_state(self.dynamics, float(params["discretization_resolution"]), self.discretization_scheme,
                                 params, states[..., :N, :], controls[..., :N, :])
        g0 = states[..., 0, :] - x0
        gd = (nxt - states[..., 1:, :]).reshape(states.shape[:-2] + (-1,))
        return torch.cat([g0, gd], -1)

    def cost(self, states, controls, params: Dict[str, Any]) -> torch.Tensor:
        """Sum of w (z - ref)^2 over steps plus (w_x + w_f)(x_N - x_final)^2, no 1/2 (ocp.py:153-247)."""
        N = self.horizon
        dev = states.device
        xref = _as_tensor(params["reference_state_trajectory"], dev)
        uref = _as_tensor(params["reference_control_trajectory"], dev)
        if not self.params["penalize_control_reference"]:
            uref = torch.zeros_like(uref)
        wx = _as_tensor(params["weights_penalization_reference_state_trajectory"], dev)
        wu = _as_tensor(params["weights_penalization_control_squared"], dev)
        wf = _as_tensor(params["weights_penalization_final_state"], dev)
        xf = _as_tensor(params["final_state"], dev)
        c = (wx.unsqueeze(-2) * (states[..., :N, :] - xref[..., :N, :]) ** 2).sum((-1, -2))
        c = c + (wu.unsqueeze(-2) * (controls[..., :N, :] - uref[..., :N, :]) ** 2).sum((-1, -2))
        c = c + ((wx + wf) * (states[..., N, :] - xf) ** 2).sum(-1)
        return c

    def step_cost(self, state, control, params: Dict[str, Any]) -> torch.Tensor:
        """sum w (z - ref)^2 of one step over [x; u] (ocp.py:153-185); params
        hold that step's references.  Leading axes broadcast."""
        xr = _as_tensor(params["reference_state_trajectory"], state.device)
        ur = _as_tensor(params["reference_control_trajectory"], state.device)
        if not self.params["penalize_control_reference"]:
            ur = torch.zeros_like(ur)
        wx = _as_tensor(params["weights_penalization_reference_state_trajectory"], state.device)
        wu = _as_tensor(params["weights_penalization_control_squared"], state.device)
        return (wx * (state - xr) ** 2).sum(-1) + (wu * (control - ur) ** 2).sum(-1)

    def final_cost(self, state, params: Dict[str, Any]) -> torch.Tensor:
        """(w_x + w_f) (x_N - x_final)^2 summed (ocp.py:187-206)."""
        wx = _as_tensor(params["weights_penalization_reference_state_trajectory"], state.device)
        wf = _as_tensor(params["weights_penalization_final_state"], state.device)
        xf = _as_tensor(params["final_state"], state.device)
        return ((wx + wf) * (state - xf) ** 2).sum(-1)

    def get_cost_linearized_matrices(self, states, controls, params: Dict[str, Any] = None):
        """(QN, qN, Q, q, R, r) of the quadratic model of the cost at
        (states, controls) (ocp.py:249-323): c = 0.5 x'Qx + q'x + 0.5 u'Ru + r'u
        + const.  The reference takes them by autodiff of step_cost and
        final_cost; those are quadratic with diagonal weights, so here
        Q = 2 diag(w_x), q_t = grad - Q x_t = -2 w_x ref_t, R = 2 diag(w_u),
        r_t = -2 w_u uref_t (0 unless the control reference is penalised),
        QN = 2 diag(w_x + w_f), qN = -2 (w_x + w_f) x_final.  Leading batch axes
        are kept; Q (..., N, nx, nx), q (..., N, nx), R (..., N, nu, nu),
        r (..., N, nu)."""
        if params is None:
            params = self.params
        N = self.horizon
        dev = states.device
        wx = _as_tensor(params["weights_penalization_reference_state_trajectory"], dev)
        wu = _as_tensor(params["weights_penalization_control_squared"], dev)
        wf = _as_tensor(params["weights_penalization_final_state"], dev)
        xr = _as_tensor(params["reference_state_trajectory"], dev)[..., :N, :]
        ur = _as_tensor(params["reference_control_trajectory"], dev)[..., :N, :]
        xf = _as_tensor(params["final_state"], dev)
        lead = states.shape[:-2]
        wxt = wx.unsqueeze(-2).expand(lead + (N, wx.shape[-1]))
        wut = wu.unsqueeze(-2).expand(lead + (N, wu.shape[-1]))
        Q = torch.diag_embed(2.0 * wxt)
        q = (-2.0 * wxt * xr).expand(lead + (N, wx.shape[-1]))
        R = torch.diag_embed(2.0 * wut)
        r = (-2.0 * wut * ur if self.params["penalize_control_reference"] else torch.zeros_like(wut))
        wN = (wx + wf).expand(lead + (wx.shape[-1],))
        QN = torch.diag_embed(2.0 * wN)
        qN = -2.0 * wN * xf
        return QN, qN, Q, q, R, r

    def get_dynamics_linearized_matrices_with_states_controls(self, states, controls, params: Dict[str, Any] = None):
        """(As_next, As, Bs, Cs) at (states, controls) on the GPU (ocp.py:325-398);
        states (N+1, nx) as the reference passes them, or batched."""
        if states.dim() == 2:
            out = self.linearize(states.unsqueeze(0), controls.unsqueeze(0), params)
            return tuple(t[0] for t in out)
        return self.linearize(states, controls, params)

    def device_problem(self, params: Dict[str, Any], batch: int, device) -> Dict[str, torch.Tensor]:
        """Batched, contiguous device tensors of the C ABI's dmpc_problem_data."""
        N, nx, nu = self.horizon, self.num_state_variables, self.num_control_variables

        def b(key, nd):
            return batched(params[key], batch, nd, torch.float64, device)

        return dict(
            x0=b("initial_state", 1), x_final=b("final_state", 1),
            x_ref=b("reference_state_trajectory", 2), u_ref=b("reference_control_trajectory", 2),
            w_x=b("weights_penalization_reference_state_trajectory", 1),
            w_u=b("weights_penalization_control_squared", 1),
            w_f=b("weights_penalization_final_state", 1),
            dyn=self.dynamics.flat_params(params, batch, device),
        )

    def c_problem(self, batch: int) -> "_lib.Problem":
        p = _lib.Problem()
        p.model = self.dynamics.model_id
        p.nx = self.num_state_variables
        p.nu = self.num_control_variables
        p.horizon = self.horizon
        p.batch = batch
        p.scheme = int(self.discretization_scheme)
        p.n_params = self.dynamics.num_params
        p.penalize_control_reference = int(bool(self.params["penalize_control_reference"]))
        p.dt = float(torch.as_tensor(self.params["discretization_resolution"]).reshape(-1)[0])
        return p

    def simulate(self, controls, params: Dict[str, Any] = None, scheme: int = None, out=None) -> torch.Tensor:
        """Open-loop rollout of controls (B, N+1, nu) from params["initial_state"]
        on the GPU (dmpc_simulate): x_0 = x0, x_{t+1} = f_d(x_t, u_t).

        scheme: the integrator (0 Euler, 1 midpoint, 2 RK4; default the
        problem's own).  The imitation-learning guess of
        benchmarking/imitation-learning/utils.py:20-31 is scheme 0.
        Ragged batches (params["horizons"]) stop each instance at its own
        horizon and zero the knots past it.  Returns (B, N+1, nx) float64
        (written into `out` when given); no gradient flows through it.
        """
        if params is None:
            params = self.params
        lib = _lib.load_for(self.dynamics)
        us = controls.to(torch.float64).contiguous()
        B, dev = us.shape[0], us.device
        N, nx, nu = self.horizon, self.num_state_variables, self.num_control_variables
        if us.shape[1:] != (N + 1, nu):
            raise ValueError(f"controls must be (B, {N + 1}, {nu}), got {tuple(us.shape)}")
        x0 = batched(params["initial_state"], B, 1, torch.float64, dev).contiguous()
        dyn = self.dynamics.flat_params(params, B, dev).contiguous()
        if out is None:
            out = torch.empty(B, N + 1, nx, dtype=torch.float64, device=dev)
        elif out.shape != (B, N + 1, nx) or out.dtype != torch.float64 or not out.is_contiguous():
            raise ValueError("out must be a contiguous float64 (B, N+1, nx) tensor")
        h = params.get("horizons")
        if h is not None:
            h = torch.as_tensor(h, device=dev).to(torch.int32).reshape(B).contiguous()
            if not torch.cuda.is_current_stream_capturing():  # (the kernel clamps them to [1, N])
                if int(h.min()) < 1 or int(h.max()) > N:
                    raise ValueError(f"horizons must lie in [1, {N}]")
        data = _lib.ProblemData(_lib.ptr(x0), None, None, None, None, None, None, _lib.ptr(dyn),
                                _lib.ptr(h) if h is not None else None)
        prob = self.c_problem(B)
        if scheme is not None:
            prob.scheme = int(scheme)
        rc = lib.dmpc_simulate(prob, data, _lib.ptr(us), _lib.ptr(out), _lib.stream_ptr(dev))
        _lib.check(rc, "dmpc_simulate")
        return out

    def linearize(self, states, controls, params: Dict[str, Any] = None):
        """As_next, As, Bs, Cs at (states, controls) on the GPU (ocp.py:325-398).

        states (B, N+1, nx), controls (B, N+1, nu) device tensors.
        Returns float64 tensors As_next (B,N,nx,nx), As, Bs (B,N,nx,nu), Cs (B,N+1,nx).
        """
        if params is None:
            params = self.params
        lib = _lib.load_for(self.dynamics)
        B = states.shape[0]
        N, nx, nu = self.horizon, self.num_state_variables, self.num_control_variables
        dev = states.device
        d = self.device_problem(params, B, dev)
        d = {k: v.contiguous() for k, v in d.items()}
        data = _lib.ProblemData(_lib.ptr(d["x0"]), _lib.ptr(d["x_final"]), _lib.ptr(d["x_ref"]), _lib.ptr(d["u_ref"]),
                                _lib.ptr(d["w_x"]), _lib.ptr(d["w_u"]), _lib.ptr(d["w_f"]), _lib.ptr(d["dyn"]))
        fd = torch.empty(B, N, nx, dtype=torch.float64, device=dev)
        A = torch.empty(B, N, nx, nx, dtype=torch.float64, device=dev)
        Bm = torch.empty(B, N, nx, nu, dtype=torch.float64, device=dev)
        C = torch.empty(B, N + 1, nx, dtype=torch.float64, device=dev)
        prob = self.c_problem(B)
        xs = states.to(torch.float64).contiguous()
        us = controls.to(torch.float64).contiguous()
        rc = lib.dmpc_linearize(prob, data, _lib.ptr(xs), _lib.ptr(us), _lib.ptr(fd), _lib.ptr(A), _lib.ptr(Bm),
                                _lib.ptr(C), _lib.stream_ptr(dev))
        _lib.check(rc, "dmpc_linearize")
        As_next = -torch.eye(nx, dtype=torch.float64, device=dev).expand(B, N, nx, nx).contiguous()
        return As_next, A, Bm, C
