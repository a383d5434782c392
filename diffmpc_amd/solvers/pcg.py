"""Schur-complement PCG QP layer (diffmpc/solvers/pcg.py:13-445) on the GPU.

Same carriers and method names as the reference's PCGOptimalControl, on
float64 device tensors: one problem as the reference passes it, or a batch
with a leading axis B.  Each method calls the
corresponding per-launch entry point of libdmpc.so:

  compute_S_Phiinv                          -> dmpc_pcg_assemble
  compute_gamma                             -> dmpc_pcg_gamma
  solve_KKT_Schur                           -> dmpc_pcg_solve (+ dmpc_pcg_recover)
  get_states_controls_from_kkt_multipliers  -> dmpc_pcg_recover

The fused SQP path does not go through this class; it exists so that code
written against the reference's QP layer runs unchanged.
"""
from typing import Any, Dict, NamedTuple, Optional, Tuple

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


def _unbatched(qp: "QPParameters") -> bool:
    """A single problem as the reference passes it: As (N, nx, nx)."""
    return qp.dynamics.As.dim() == 3


def _lift(nt):
    """The carrier with a leading batch axis of 1 on every tensor."""
    return type(nt)(*[torch.as_tensor(t).unsqueeze(0) if t is not None else None for t in nt])


def _lift_qp(qp: "QPParameters") -> "QPParameters":
    return QPParameters(_lift(qp.dynamics), _lift(qp.cost))


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

    def zero_qp_parameters(self, batch: Optional[int] = None, device=None) -> QPParameters:
        """All-zero QP blocks: (N, nx, nx) ... as the reference's, or with a
        leading batch axis when `batch` is given."""
        N, nx, nu = self.horizon, self.num_states, self.num_controls
        z = self._zeros(batch, device)
        return QPParameters(QPDynamicsParameters(z(N, nx, nx), z(N, nx, nx), z(N, nx, nu), z(N + 1, nx)),
                            QPCostParameters(z(N + 1, nx, nx), z(N + 1, nx), z(N, nu, nu), z(N, nu)))

    def zero_dynamics_cost_premultiplied_matrices(self, batch: Optional[int] = None, device=None):
        N, nx, nu = self.horizon, self.num_states, self.num_controls
        z = self._zeros(batch, device)
        return QPDynamicsCostPremultipliedMatrices(z(N, nx, nx), z(N, nx, nx), z(N, nx, nu))

    def zero_schur_complement_matrices(self, batch: Optional[int] = None, device=None):
        N, nx = self.horizon, self.num_states
        z = self._zeros(batch, device)
        return SchurComplementMatrices(z(N + 1, nx, 3 * nx), z(N + 1, nx, 3 * nx))

    @staticmethod
    def _zeros(batch, device):
        lead = () if batch is None else (batch,)
        return lambda *s: torch.zeros(lead + s, dtype=torch.float64, device=device)

    def _shapes(self, B: int):
        N, nx, nu = self.horizon, self.num_states, self.num_controls
        return dict(As_next=(B, N, nx, nx), As=(B, N, nx, nx), Bs=(B, N, nx, nu), Cs=(B, N + 1, nx),
                    Qinv=(B, N + 1, nx, nx), qvec=(B, N + 1, nx), Rinv=(B, N, nu, nu), rvec=(B, N, nu),
                    lam=(B, N + 1, nx), S=(B, N + 1, nx, 3 * nx))

    def _args(self, B: int, device, **tensors):
        """Each tensor broadcast to the shape the kernels index, float64,
        contiguous, on `device` (a ValueError names the one that does not
        fit): the entry points read exactly those extents."""
        shapes = self._shapes(B)
        out = []
        for name, t in tensors.items():
            key = "S" if name in ("S", "Phiinv") else ("lam" if name in ("lam", "gamma") else name)
            t = torch.as_tensor(t, dtype=torch.float64, device=device)
            try:
                out.append(t.expand(shapes[key]).contiguous())
            except RuntimeError:
                raise ValueError(f"{name} has shape {tuple(t.shape)}; expected {shapes[key]} "
                                 f"(horizon {self.horizon}, nx {self.num_states}, nu {self.num_controls})") from None
        return out

    def compute_S_Phiinv(self, qp_parameters: QPParameters) -> SchurComplementMatrices:
        """S = -G H^-1 G^T and its block-Jacobi-stair inverse (pcg.py:160-243)."""
        if _unbatched(qp_parameters):
            r = self.compute_S_Phiinv(_lift_qp(qp_parameters))
            return SchurComplementMatrices(r.S[0], r.preconditioner_Phiinv[0])
        lib = _lib.load()
        d, c = qp_parameters.dynamics, qp_parameters.cost
        B, N, nx, nu = d.As.shape[0], self.horizon, self.num_states, self.num_controls
        dev = d.As.device
        args = self._args(B, dev, As_next=d.As_next, As=d.As, Bs=d.Bs, Qinv=c.Qinv, Rinv=c.Rinv)
        S = torch.empty(B, N + 1, nx, 3 * nx, dtype=torch.float64, device=dev)
        P = torch.empty_like(S)
        rc = lib.dmpc_pcg_assemble(B, N, nx, nu, *[_lib.ptr(a) for a in args], _lib.ptr(S), _lib.ptr(P),
                                   _lib.stream_ptr(dev))
        _lib.check(rc, "dmpc_pcg_assemble")
        return SchurComplementMatrices(S=S, preconditioner_Phiinv=P)

    def compute_gamma_from_qp(self, qp_parameters: QPParameters) -> torch.Tensor:
        """gamma = C + G H^-1 q (pcg.py:245-288) from the QP blocks."""
        if _unbatched(qp_parameters):
            return self.compute_gamma_from_qp(_lift_qp(qp_parameters))[0]
        lib = _lib.load()
        d, c = qp_parameters.dynamics, qp_parameters.cost
        B, N, nx, nu = d.As.shape[0], self.horizon, self.num_states, self.num_controls
        dev = d.As.device
        args = self._args(B, dev, As_next=d.As_next, As=d.As, Bs=d.Bs, Qinv=c.Qinv, Rinv=c.Rinv, Cs=d.Cs,
                          qvec=c.qvec, rvec=c.rvec)
        g = torch.empty(B, N + 1, nx, dtype=torch.float64, device=dev)
        rc = lib.dmpc_pcg_gamma(B, N, nx, nu, *[_lib.ptr(a) for a in args], _lib.ptr(g), _lib.stream_ptr(dev))
        _lib.check(rc, "dmpc_pcg_gamma")
        return g

    def compute_gamma(self, first, Cs=None, Q0inv=None, qs=None, rs=None) -> torch.Tensor:
        """gamma = C + G H^-1 q (pcg.py:245-288).

        Called as the reference does, compute_gamma(premultiplied, Cs, Q0inv,
        qs, rs) with the QPDynamicsCostPremultipliedMatrices (A Qinv,
        A_next Qinv_next, B Rinv); or with QPParameters first and optional
        replacements of C, q, r (the backward uses C = 0)."""
        if isinstance(first, QPDynamicsCostPremultipliedMatrices):
            # the same kernel with the products folded into the matrices
            # (Qinv = I, Rinv = I), then row 0, C_0 + Q0inv q_0
            AQ, AnQ, BR = first.As_x_Qinv, first.Asnext_x_Qinvnext, first.Bs_x_Rinv
            lead, N, nx, nu = AQ.shape[:-3], AQ.shape[-3], AQ.shape[-1], BR.shape[-1]
            dev = AQ.device
            eye = lambda n, k: torch.eye(n, dtype=torch.float64, device=dev).expand(lead + (k, n, n))  # noqa: E731
            qp = QPParameters(QPDynamicsParameters(AnQ, AQ, BR, Cs), QPCostParameters(eye(nx, N + 1), qs, eye(nu, N), rs))
            gamma = self.compute_gamma_from_qp(qp)
            C0 = torch.as_tensor(Cs, dtype=torch.float64, device=dev)[..., 0, :]
            q0 = torch.as_tensor(qs, dtype=torch.float64, device=dev)[..., 0, :]
            Q0 = torch.as_tensor(Q0inv, dtype=torch.float64, device=dev)
            gamma[..., 0, :] = C0 + (Q0 @ q0.unsqueeze(-1)).squeeze(-1)
            return gamma
        d, c = first.dynamics, first.cost
        qp = QPParameters(d._replace(Cs=d.Cs if Cs is None else Cs),
                          c._replace(qvec=c.qvec if qs is None else qs, rvec=c.rvec if rs is None else rs))
        return self.compute_gamma_from_qp(qp)

    def get_states_controls_from_kkt_multipliers(self, kkt_multipliers: torch.Tensor,
                                                 qp_parameters: QPParameters) -> QPSolution:
        """x, u from lambda (pcg.py:290-348); the last control row is zero."""
        if _unbatched(qp_parameters):
            r = self.get_states_controls_from_kkt_multipliers(kkt_multipliers.unsqueeze(0), _lift_qp(qp_parameters))
            return QPSolution(r.states[0], r.controls[0], r.kkt_multipliers[0])
        lib = _lib.load()
        d, c = qp_parameters.dynamics, qp_parameters.cost
        B, N, nx, nu = d.As.shape[0], self.horizon, self.num_states, self.num_controls
        args = self._args(B, d.As.device, lam=kkt_multipliers, As_next=d.As_next, As=d.As, Bs=d.Bs, Qinv=c.Qinv,
                          qvec=c.qvec, Rinv=c.Rinv, rvec=c.rvec)
        lam = args[0]
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
        if _unbatched(qp_parameters):
            sol, dbg = self.solve_KKT_Schur(_lift_qp(qp_parameters), _lift(schur_complement_matrices),
                                            schur_complement_gammas.unsqueeze(0), kkt_multipliers_guess.unsqueeze(0))
            return (QPSolution(sol.states[0], sol.controls[0], sol.kkt_multipliers[0]),
                    PCGDebugOutput(dbg.num_iterations[0], dbg.convergence_eta[0]))
        lib = _lib.load()
        B, N, nx = qp_parameters.dynamics.As.shape[0], self.horizon, self.num_states
        S, P, g, lam = self._args(B, qp_parameters.dynamics.As.device, S=schur_complement_matrices.S,
                                  Phiinv=schur_complement_matrices.preconditioner_Phiinv, gamma=schur_complement_gammas,
                                  lam=kkt_multipliers_guess)
        lam = lam.clone()  # in/out
        iters = torch.zeros(B, dtype=torch.int32, device=S.device)
        eta = torch.zeros(B, dtype=torch.float64, device=S.device)
        rc = lib.dmpc_pcg_solve(B, N, nx, _lib.ptr(S), _lib.ptr(P), _lib.ptr(g), _lib.ptr(lam),
                                int(self.params["max_iter"]), float(self.params["tol_epsilon"]), _lib.ptr(iters),
                                _lib.ptr(eta), _lib.stream_ptr(S.device))
        _lib.check(rc, "dmpc_pcg_solve")
        sol = self.get_states_controls_from_kkt_multipliers(lam, qp_parameters)
        return sol, PCGDebugOutput(num_iterations=iters, convergence_eta=eta)
