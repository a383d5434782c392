"""Schur-complement PCG QP layer (diffmpc/solvers/pcg.py:13-445) on the GPU.

Same carriers and method names as the reference's PCGOptimalControl, on
batched float64 device tensors (leading batch axis B).  Each method calls the
corresponding per-launch entry point of libdmpc.so:

  compute_S_Phiinv                          -> dmpc_pcg_assemble
  compute_gamma                             -> dmpc_pcg_gamma
  solve_KKT_Schur                           -> dmpc_pcg_solve (+ dmpc_pcg_recover)
  get_states_controls_from_kkt_multipliers  -> dmpc_pcg_recover

The fused SQP path does not go through this class; it exists so that code
written against the reference's QP layer runs unchanged.
"""
from typing import Any, Dict, NamedTuple, Tuple

import torch

from .. import _lib
from ..utils.load_params import load_solver_params

DEFAULT_PCG_SOLVER_PARAMS = load_solver_params("sqp.yaml")["pcg"]


class QPDynamicsParameters(NamedTuple):
    As_next: torch.Tensor  # (B, N, nx, nx)
    As: torch.Tensor  # (B, N, nx, nx)
    Bs: torch.Tensor  # (B, N, nx, nu)
    Cs: torch.Tensor  # (B, N+1, nx)


class QPCostParameters(NamedTuple):
    Qinv: torch.Tensor  # (B, N+1, nx, nx)
    qvec: torch.Tensor  # (B, N+1, nx)
    Rinv: torch.Tensor  # (B, N, nu, nu)
    rvec: torch.Tensor  # (B, N, nu)


class QPParameters(NamedTuple):
    dynamics: QPDynamicsParameters
    cost: QPCostParameters


class QPSolution(NamedTuple):
    states: torch.Tensor  # (B, N+1, nx)
    controls: torch.Tensor  # (B, N+1, nu)
    kkt_multipliers: torch.Tensor  # (B, N+1, nx)


class QPDynamicsCostPremultipliedMatrices(NamedTuple):
    As_x_Qinv: torch.Tensor
    Asnext_x_Qinvnext: torch.Tensor
    Bs_x_Rinv: torch.Tensor


class SchurComplementMatrices(NamedTuple):
    S: torch.Tensor  # (B, N+1, nx, 3 nx)
    preconditioner_Phiinv: torch.Tensor  # (B, N+1, nx, 3 nx)


class PCGDebugOutput(NamedTuple):
    num_iterations: torch.Tensor  # (B,) int32
    convergence_eta: torch.Tensor  # (B,) float64


def _d(t):
    return t.to(torch.float64).contiguous()


class PCGOptimalControl:
    """Block-Jacobi-preconditioned CG on the Schur complement of the KKT system
    (MPCGPU, arXiv 2309.08079), extended with the A_next_t x_{t+1} term."""

    def __init__(self, problem_horizon: int, problem_num_states: int, problem_num_controls: int,
                 solver_params: Dict[str, Any] = DEFAULT_PCG_SOLVER_PARAMS):
        self._params = solver_params
        self._name = "PCGOptimalControl"
        self._problem_horizon = problem_horizon
        self._problem_num_states = problem_num_states
        self._problem_num_controls = problem_num_controls

    @property
    def name(self) -> str:
        return self._name

    @property
    def params(self) -> Dict:
        return self._params

    @property
    def horizon(self) -> int:
        return self._problem_horizon

    @property
    def num_states(self) -> int:
        return self._problem_num_states

    @property
    def num_controls(self) -> int:
        return self._problem_num_controls

    def zero_qp_parameters(self, batch: int = 1, device=None) -> QPParameters:
        N, nx, nu = self.horizon, self.num_states, self.num_controls
        z = lambda *s: torch.zeros((batch,) + s, dtype=torch.float64, device=device)  # noqa: E731
        return QPParameters(QPDynamicsParameters(z(N, nx, nx), z(N, nx, nx), z(N, nx, nu), z(N + 1, nx)),
                            QPCostParameters(z(N + 1, nx, nx), z(N + 1, nx), z(N, nu, nu), z(N, nu)))

    def zero_dynamics_cost_premultiplied_matrices(self, batch: int = 1, device=None):
        N, nx, nu = self.horizon, self.num_states, self.num_controls
        z = lambda *s: torch.zeros((batch,) + s, dtype=torch.float64, device=device)  # noqa: E731
        return QPDynamicsCostPremultipliedMatrices(z(N, nx, nx), z(N, nx, nx), z(N, nx, nu))

    def zero_schur_complement_matrices(self, batch: int = 1, device=None):
        N, nx = self.horizon, self.num_states
        z = lambda *s: torch.zeros((batch,) + s, dtype=torch.float64, device=device)  # noqa: E731
        return SchurComplementMatrices(z(N + 1, nx, 3 * nx), z(N + 1, nx, 3 * nx))

    def compute_S_Phiinv(self, qp_parameters: QPParameters) -> SchurComplementMatrices:
        """S = -G H^-1 G^T and its block-Jacobi-stair inverse (pcg.py:160-243)."""
        lib = _lib.load()
        d, c = qp_parameters.dynamics, qp_parameters.cost
        B, N, nx = d.As.shape[0], self.horizon, self.num_states
        nu = self.num_controls
        args = [_d(d.As_next), _d(d.As), _d(d.Bs), _d(c.Qinv), _d(c.Rinv)]
        S = torch.empty(B, N + 1, nx, 3 * nx, dtype=torch.float64, device=d.As.device)
        P = torch.empty_like(S)
        rc = lib.dmpc_pcg_assemble(B, N, nx, nu, *[_lib.ptr(a) for a in args], _lib.ptr(S), _lib.ptr(P),
                                   _lib.stream_ptr(S.device))
        _lib.check(rc, "dmpc_pcg_assemble")
        return SchurComplementMatrices(S=S, preconditioner_Phiinv=P)

    def compute_gamma_from_qp(self, qp_parameters: QPParameters) -> torch.Tensor:
        """gamma = C + G H^-1 q (pcg.py:245-288) from the QP blocks."""
        lib = _lib.load()
        d, c = qp_parameters.dynamics, qp_parameters.cost
        B, N, nx, nu = d.As.shape[0], self.horizon, self.num_states, self.num_controls
        args = [_d(d.As_next), _d(d.As), _d(d.Bs), _d(c.Qinv), _d(c.Rinv), _d(d.Cs), _d(c.qvec), _d(c.rvec)]
        g = torch.empty(B, N + 1, nx, dtype=torch.float64, device=d.As.device)
        rc = lib.dmpc_pcg_gamma(B, N, nx, nu, *[_lib.ptr(a) for a in args], _lib.ptr(g), _lib.stream_ptr(g.device))
        _lib.check(rc, "dmpc_pcg_gamma")
        return g

    def compute_gamma(self, qp_parameters: QPParameters, Cs=None, qs=None, rs=None) -> torch.Tensor:
        """gamma with optional replacements of C, q, r (the backward uses C = 0)."""
        d, c = qp_parameters.dynamics, qp_parameters.cost
        qp = QPParameters(d._replace(Cs=d.Cs if Cs is None else Cs),
                          c._replace(qvec=c.qvec if qs is None else qs, rvec=c.rvec if rs is None else rs))
        return self.compute_gamma_from_qp(qp)

    def get_states_controls_from_kkt_multipliers(self, kkt_multipliers: torch.Tensor,
                                                 qp_parameters: QPParameters) -> QPSolution:
        """x, u from lambda (pcg.py:290-348); the last control row is zero."""
        lib = _lib.load()
        d, c = qp_parameters.dynamics, qp_parameters.cost
        B, N, nx, nu = d.As.shape[0], self.horizon, self.num_states, self.num_controls
        lam = _d(kkt_multipliers)
        args = [lam, _d(d.As_next), _d(d.As), _d(d.Bs), _d(c.Qinv), _d(c.qvec), _d(c.Rinv), _d(c.rvec)]
        x = torch.empty(B, N + 1, nx, dtype=torch.float64, device=lam.device)
        u = torch.empty(B, N + 1, nu, dtype=torch.float64, device=lam.device)
        rc = lib.dmpc_pcg_recover(B, N, nx, nu, *[_lib.ptr(a) for a in args], _lib.ptr(x), _lib.ptr(u),
                                  _lib.stream_ptr(x.device))
        _lib.check(rc, "dmpc_pcg_recover")
        return QPSolution(x, u, lam)

    def solve_KKT_Schur(self, qp_parameters: QPParameters, schur_complement_matrices: SchurComplementMatrices,
                        schur_complement_gammas: torch.Tensor,
                        kkt_multipliers_guess: torch.Tensor) -> Tuple[QPSolution, PCGDebugOutput]:
        """PCG on S lam = gamma with the reference's stop rule (pcg.py:350-445)."""
        lib = _lib.load()
        S = _d(schur_complement_matrices.S)
        P = _d(schur_complement_matrices.preconditioner_Phiinv)
        g = _d(schur_complement_gammas)
        lam = _d(kkt_multipliers_guess).clone()
        B, N, nx = S.shape[0], self.horizon, self.num_states
        iters = torch.zeros(B, dtype=torch.int32, device=S.device)
        eta = torch.zeros(B, dtype=torch.float64, device=S.device)
        rc = lib.dmpc_pcg_solve(B, N, nx, _lib.ptr(S), _lib.ptr(P), _lib.ptr(g), _lib.ptr(lam),
                                int(self.params["max_iter"]), float(self.params["tol_epsilon"]), _lib.ptr(iters),
                                _lib.ptr(eta), _lib.stream_ptr(S.device))
        _lib.check(rc, "dmpc_pcg_solve")
        sol = self.get_states_controls_from_kkt_multipliers(lam, qp_parameters)
        return sol, PCGDebugOutput(num_iterations=iters, convergence_eta=eta)
