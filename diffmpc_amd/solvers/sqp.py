"""Batched SQP solver with an implicit-differentiation backward, on the GPU.

Mirrors diffmpc/solvers/sqp.py: ``SQPSolver(program, name, params)``,
``solver.initial_guess(params)``, ``solver.solve(initial_guess, problem_params,
weights) -> SQPSolution`` and the ``SolverReturnStatus`` codes.  Differences
that follow from the batch axis being explicit here instead of a caller's
``jax.vmap``:

* tensors carry a leading batch axis B (an unbatched guess is accepted and
  the result is returned unbatched);
* problem parameters may be shared (no batch axis) or per instance;
* gradients flow, through ``torch.autograd``, to the entries of ``weights``
  only (problem_params are stop-gradient, sqp.py:456-458); shared weights get
  the batch-summed gradient, like ``vmap(in_axes=None)``.

The solve runs in libdmpc.so: mode "fused" is one workgroup per instance per
SQP iteration, "plain" is the per-launch op graph of the reference, and
"auto" (the default) takes fused where the library builds it for the model
and size, else plain (linear models of any size).  Solver parameters use the reference's keys plus the optional
``"mode"``, ``"fused_variant"`` ("auto", "register", "streamed",
"streamed_per_knot"), ``"blocks_per_group"`` (0 = automatic, 1, 2, 4) and
``"instances_per_workgroup"`` (0 or 1 = one, up to 8, register-resident
kernels of models with nx <= 8), which diffmpc_amd.autotune sets, and
``"forward_schedule"`` ("auto": one launch per solve for batches of one
horizon with more instances than CUs, "per_iteration": one linearization and
one QP launch per SQP iteration; the same bits).  ``problem_params["horizons"]`` (B,) makes a ragged batch.
"""
import copy
import enum
from typing import Any, Dict, NamedTuple, Optional

import torch

from .. import _lib
from ..dynamics.base_dynamics import batched
from ..problems.optimal_control_problem import OptimalControlProblem
from ..utils.load_params import load_solver_params
from ..utils.torch_utils import project_matrix_onto_positive_semidefinite_cone
from .pcg import (PCGOptimalControl, QPCostParameters, QPDynamicsCostPremultipliedMatrices, QPDynamicsParameters,
                  QPParameters, QPSolution, SchurComplementMatrices)

PSD_MIN_EIGENVALUE = 1e-9  # sqp.py:191-196

DEFAULT_SOLVER_PARAMS = load_solver_params("sqp.yaml")

# weights keys the backward differentiates (closed-form mixed derivatives)
# problem_params key of the model's flat float32 parameter block, (B, n_params),
# as batch_params precomputes it
FLAT_DYNAMICS_KEY = "_flat_dynamics_parameters"

WEIGHT_KEYS = {
    "weights_penalization_reference_state_trajectory": "w_x",
    "weights_penalization_control_squared": "w_u",
    "weights_penalization_final_state": "w_f",
    "initial_state": "x0",
    "reference_state_trajectory": "x_ref",
    "reference_control_trajectory": "u_ref",
    "final_state": "x_final",
}


class SolverReturnStatus(enum.IntEnum):
    SUCCESS = 0
    ERROR = -100  # NaN detected or not converged


class SavedLinearization(NamedTuple):
    """QP data of the last linearization point (sqp.py:487-492)."""

    AB: torch.Tensor  # (B, N, nx, nx+nu) float32: [A_t | B_t]
    fd: torch.Tensor  # (B, N, nx) float64: f_d(x_t, u_t)
    qinv: torch.Tensor  # (B, N+1, nx) float64: diag Qinv_t
    rinv: torch.Tensor  # (B, nu) float64: diag Rinv


class SQPSolution(NamedTuple):
    states: torch.Tensor
    controls: torch.Tensor
    dual: torch.Tensor
    qp_parameters: Optional[SavedLinearization]
    dynamics_cost_premultiplied_matrices: Any
    schur_complement_matrices: Any
    status: Any
    num_iter: Any
    convergence_error: Any
    convergence_errors: torch.Tensor
    linesearch_alphas: torch.Tensor
    num_pcg_iters: torch.Tensor
    dual_backward_guess: torch.Tensor = None


def _c_solver_params(params: Dict[str, Any]) -> "_lib.SolverParams":
    sp = _lib.SolverParams()
    sp.num_scp_iteration_max = int(params["num_scp_iteration_max"])
    sp.tol_convergence = float(params["tol_convergence"])
    sp.linesearch = int(bool(params["linesearch"]))
    sp.linesearch_eta = float(params["linesearch_eta"])
    alphas = [float(a) for a in params["linesearch_alphas"]]
    if len(alphas) > 8:
        raise ValueError("at most 8 line-search candidates are supported")
    sp.n_alphas = len(alphas)
    for i, a in enumerate(alphas):
        sp.linesearch_alphas[i] = a
    sp.pcg_max_iter = int(params["pcg"]["max_iter"])
    sp.pcg_tol_epsilon = float(params["pcg"]["tol_epsilon"])
    sp.warm_start_backward = int(bool(params["warm_start_backward"]))
    # fused-mode tuning (keys with no reference counterpart; 0 = library default)
    sp.blocks_per_group = int(params.get("blocks_per_group", 0))
    sp.fused_variant = {"auto": 0, "register": 1, "streamed": 2, "streamed_per_knot": 3}[
        params.get("fused_variant", "auto")]
    sp.instances_per_workgroup = int(params.get("instances_per_workgroup", 0))
    sp.forward_schedule = {"auto": 0, "per_iteration": 1}[params.get("forward_schedule", "auto")]
    return sp


def _ragged_horizons(h, B, N, device):
    """Per-instance horizons of a ragged batch (problem_params["horizons"],
    an extension with no reference counterpart): (B,) ints in [1, N], where
    N (the problem's horizon) is the padded length of every per-knot array."""
    if h is None:
        return None
    h = torch.as_tensor(h).reshape(-1)
    if h.numel() != B:
        raise ValueError(f"horizons must have one entry per instance ({B}), got {h.numel()}")
    if not (h.is_cuda and torch.cuda.is_current_stream_capturing()):
        # (while a graph is being captured the values cannot be read back; the
        # kernels clamp them to [1, N])
        hc = h.detach().cpu()
        if int(hc.min()) < 1 or int(hc.max()) > N:
            raise ValueError(f"horizons must lie in [1, {N}]")
    return h.to(device=device, dtype=torch.int32).contiguous()


def _check_shapes(B, N, nx, nu, maxit, n_params, guess, data):
    """Every array the kernels index with fixed strides must have exactly the
    shape of the problem: a guess, dual or history left over from a solver
    with another horizon, size or iteration cap raises here instead of
    reaching the device (where the reference raises a shape error)."""
    want = {
        "states": (B, N + 1, nx), "controls": (B, N + 1, nu), "dual": (B, N + 1, nx),
        "convergence_error": (B,), "convergence_errors": (B, maxit), "linesearch_alphas": (B, maxit),
        "num_pcg_iters": (B, maxit), "dual_backward_guess": (B, N + 1, nx),
        "initial_state": (B, nx), "final_state": (B, nx), "reference_state_trajectory": (B, N + 1, nx),
        "reference_control_trajectory": (B, N + 1, nu),
        "weights_penalization_reference_state_trajectory": (B, nx),
        "weights_penalization_control_squared": (B, nu), "weights_penalization_final_state": (B, nx),
        "dynamics parameters": (B, n_params),
    }
    for src in (guess, data):
        for k, t in src.items():
            if t is not None and tuple(t.shape) != want[k]:
                raise ValueError(f"{k} has shape {tuple(t.shape)}, expected {want[k]} "
                                 f"(B={B}, N={N}, nx={nx}, nu={nu}, num_scp_iteration_max={maxit})")


def _mode(params, prob, lib, ragged: bool) -> int:
    """The execution mode of a solve: params["mode"] "fused" or "plain", or
    "auto" (the default): the fused kernels where the library builds them for
    this model and size, else the per-launch kernels (which take any linear
    model size but no ragged horizons)."""
    m = params.get("mode", "auto")
    if m not in ("auto", "fused", "plain"):
        raise ValueError(f"mode must be 'auto', 'fused' or 'plain', got {m!r}")
    if m != "auto":
        return _lib.MODE_FUSED if m == "fused" else _lib.MODE_PLAIN
    if lib.dmpc_supported(_lib.MODE_FUSED, prob.model, prob.nx, prob.nu, prob.horizon):
        return _lib.MODE_FUSED
    if not ragged and lib.dmpc_supported(_lib.MODE_PLAIN, prob.model, prob.nx, prob.nu, prob.horizon):
        return _lib.MODE_PLAIN
    return _lib.MODE_FUSED  # the solve reports the unsupported size


class _SQPSolveFn(torch.autograd.Function):
    """custom_vjp of sqp.py:562-802 as a torch.autograd.Function."""

    @staticmethod
    def forward(ctx, solver, guess, x0, x_final, x_ref, u_ref, w_x, w_u, w_f, dyn, horizons):
        prog = solver.program
        lib = _lib.load_for(prog.dynamics)
        B = guess["states"].shape[0]
        dev = guess["states"].device
        N, nx, nu = prog.horizon, prog.num_state_variables, prog.num_control_variables
        data_t = [t.contiguous() for t in (x0, x_final, x_ref, u_ref, w_x, w_u, w_f)] + [dyn.contiguous()]
        data_t.append(horizons)
        data = _lib.ProblemData(*[_lib.ptr(t) for t in data_t])
        # in/out arrays: fresh copies of the guess (the guess may be a previous
        # solution that autograd still holds), made by one multi-tensor copy
        dual_bwd_in = guess.get("dual_backward_guess")
        src = [guess["states"], guess["controls"], guess["dual"], guess["convergence_error"],
               guess["convergence_errors"], guess["linesearch_alphas"]]
        if dual_bwd_in is not None:
            src.append(dual_bwd_in)
        src = [t.to(torch.float64) for t in src]
        dst = [torch.empty(t.shape, dtype=torch.float64, device=dev) for t in src]
        torch._foreach_copy_(dst, src)
        x, u, lam, conv_err, conv_errs, ls_alphas = dst[:6]
        dual_bwd_out = (dst[6] if dual_bwd_in is not None
                        else torch.zeros(B, N + 1, nx, dtype=torch.float64, device=dev))
        pcg_iters = guess["num_pcg_iters"].to(torch.int32).clone().contiguous()
        # written by every solve (reset and finish kernels)
        status = torch.empty(B, dtype=torch.int32, device=dev)
        num_iter = torch.empty(B, dtype=torch.int32, device=dev)
        # the saved linearization is written for every knot of a full-horizon
        # batch; ragged batches leave the padding untouched, so it starts at zero
        alloc = torch.empty if horizons is None else torch.zeros
        AB = alloc(B, N, nx, nx + nu, dtype=torch.float32, device=dev)
        fd = alloc(B, N, nx, dtype=torch.float64, device=dev)
        qinv = alloc(B, N + 1, nx, dtype=torch.float64, device=dev)
        rinv = alloc(B, nu, dtype=torch.float64, device=dev)
        sol = _lib.Solution(_lib.ptr(x), _lib.ptr(u), _lib.ptr(lam), _lib.ptr(status), _lib.ptr(num_iter),
                            _lib.ptr(conv_err), _lib.ptr(conv_errs), _lib.ptr(ls_alphas), _lib.ptr(pcg_iters))
        prob = prog.c_problem(B)
        sp = _c_solver_params(solver.params)
        mode = _mode(solver.params, prob, lib, horizons is not None)
        nwork = _lib.workspace_bytes(mode, prob, sp, lib)
        work = torch.empty(max(nwork, 1), dtype=torch.uint8, device=dev)
        saved = _lib.Saved(_lib.ptr(AB), _lib.ptr(fd), _lib.ptr(qinv), _lib.ptr(rinv),
                           _lib.ptr(work) if nwork else None, nwork)
        rc = lib.dmpc_sqp_solve(prob, sp, data, sol, saved, mode, _lib.stream_ptr(dev))
        _lib.check(rc, "dmpc_sqp_solve")
        ctx.solver = solver
        ctx.prob_args = (prob, sp, mode)
        ctx.data_t = data_t
        ctx.saved_lin = (AB, fd, qinv, rinv, work, nwork)
        ctx.dual_bwd_in = dual_bwd_in
        ctx.dual_bwd_out = dual_bwd_out
        ctx.dyn_dtype = dyn.dtype
        ctx.save_for_backward(x, u, lam)
        ctx.mark_non_differentiable(lam, status, num_iter, conv_err, conv_errs, ls_alphas, pcg_iters, AB, fd, qinv,
                                    rinv, dual_bwd_out)
        return x, u, lam, status, num_iter, conv_err, conv_errs, ls_alphas, pcg_iters, AB, fd, qinv, rinv, dual_bwd_out

    @staticmethod
    def backward(ctx, g_x, g_u, *unused):
        lib = _lib.load_for(ctx.solver.program.dynamics)
        x, u, lam = ctx.saved_tensors
        prob, sp, mode = ctx.prob_args
        B = x.shape[0]
        N = prob.horizon
        nx, nu = prob.nx, prob.nu
        dev = x.device
        dl_dx = (g_x if g_x is not None else torch.zeros_like(x)).to(torch.float64).contiguous()
        dl_du = (g_u if g_u is not None else torch.zeros_like(u)).to(torch.float64).contiguous()
        # adjoint warm start: the cotangent that reached this solve's output
        # dual_backward_guess (written by the backward of the solve that used
        # this solution as its guess), sqp.py:748-752
        if ctx.solver.params["warm_start_backward"]:
            lam_b = ctx.dual_bwd_out.clone().contiguous()
        else:
            lam_b = torch.zeros(B, N + 1, nx, dtype=torch.float64, device=dev)
        g = {k: torch.empty(s, dtype=torch.float64, device=dev) for k, s in (
            ("w_x", (B, nx)), ("w_u", (B, nu)), ("w_f", (B, nx)), ("x0", (B, nx)), ("x_ref", (B, N + 1, nx)),
            ("u_ref", (B, N + 1, nu)), ("x_final", (B, nx)))}
        iters = torch.empty(B, dtype=torch.int32, device=dev)  # written by the backward
        want_dyn = ctx.needs_input_grad[9]
        if want_dyn:  # the adjoint primal solution feeds the dynamics-parameter contraction
            adj_x = torch.empty(B, N + 1, nx, dtype=torch.float64, device=dev)
            adj_u = torch.empty(B, N + 1, nu, dtype=torch.float64, device=dev)
        grads = _lib.Grads(_lib.ptr(g["w_x"]), _lib.ptr(g["w_u"]), _lib.ptr(g["w_f"]), _lib.ptr(g["x0"]),
                           _lib.ptr(g["x_ref"]), _lib.ptr(g["u_ref"]), _lib.ptr(g["x_final"]), _lib.ptr(lam_b),
                           _lib.ptr(iters), _lib.ptr(adj_x) if want_dyn else None,
                           _lib.ptr(adj_u) if want_dyn else None)
        data = _lib.ProblemData(*[_lib.ptr(t) for t in ctx.data_t])
        AB, fd, qinv, rinv, work, nwork = ctx.saved_lin
        saved = _lib.Saved(_lib.ptr(AB), _lib.ptr(fd), _lib.ptr(qinv), _lib.ptr(rinv),
                           _lib.ptr(work) if nwork else None, nwork)
        rc = lib.dmpc_sqp_backward(prob, sp, data, _lib.ptr(x), _lib.ptr(u), saved, _lib.ptr(dl_dx), _lib.ptr(dl_du),
                                   grads, mode, _lib.stream_ptr(dev))
        _lib.check(rc, "dmpc_sqp_backward")
        g_dyn = None
        if want_dyn:
            g_dyn64 = torch.zeros(B, prob.n_params, dtype=torch.float64, device=dev)
            rc = lib.dmpc_dynamics_param_grads(prob, data, _lib.ptr(x), _lib.ptr(u), _lib.ptr(lam), _lib.ptr(adj_x),
                                               _lib.ptr(adj_u), _lib.ptr(lam_b), _lib.ptr(g_dyn64),
                                               _lib.stream_ptr(dev))
            _lib.check(rc, "dmpc_dynamics_param_grads")
            g_dyn = g_dyn64.to(ctx.dyn_dtype)
        if ctx.dual_bwd_in is not None:
            with torch.no_grad():
                ctx.dual_bwd_in.copy_(lam_b.to(ctx.dual_bwd_in.dtype))
        ctx.solver.last_backward_pcg_iters = iters
        return (None, None, g["x0"], g["x_final"], g["x_ref"], g["u_ref"], g["w_x"], g["w_u"], g["w_f"], g_dyn, None)


class SQPSolver:
    """SQP for optimal control (equality constraints only) on the GPU."""

    _supported_program_types = [OptimalControlProblem]

    def __init__(self, program: OptimalControlProblem, name: str = "SQPSolver",
                 params: Dict = DEFAULT_SOLVER_PARAMS):
        if params["verbose"]:
            print("Initializing Solver with\n> name    =", name, "\n> program =", program)
        if not any(isinstance(program, t) for t in self._supported_program_types):
            raise NotImplementedError(str(getattr(program, "name", program)) + " is not supported.")
        self._verbose = params["verbose"]
        self._params = copy.deepcopy(params)
        self._program = program
        self._name = name
        self._pcg = PCGOptimalControl(program.horizon, program.num_state_variables, program.num_control_variables,
                                      self._params["pcg"])
        self.last_backward_pcg_iters = None
        self.solve = self._differentiable_solve

    @property
    def name(self) -> str:
        return self._name

    @property
    def params(self) -> Dict:
        return self._params

    @property
    def problem_params(self) -> Dict:
        return self._program.params

    @property
    def maxiter(self) -> int:
        return int(self.params["num_scp_iteration_max"])

    @property
    def program(self) -> OptimalControlProblem:
        return self._program

    @property
    def verbose(self) -> bool:
        return self._verbose

    @property
    def pcg(self) -> PCGOptimalControl:
        return self._pcg

    def initial_guess_primal(self, params: Dict[str, Any] = None, device=None):
        return self.program.initial_guess(params, device=device)

    def initial_guess_dual(self, batch: Optional[int] = None, device=None) -> torch.Tensor:
        shape = (self.program.horizon + 1, self.program.num_state_variables)
        if batch is not None:
            shape = (batch,) + shape
        return torch.zeros(shape, dtype=torch.float64, device=device)

    def initial_guess(self, params: Dict[str, Any] = None, device=None) -> SQPSolution:
        """Straight-line guess (sqp.py:143-158); batched if initial_state is (B, nx)."""
        x, u = self.initial_guess_primal(params, device=device)
        batch = x.shape[0] if x.dim() == 3 else None
        lead = (batch,) if batch is not None else ()
        dev = x.device
        return SQPSolution(
            x, u, self.initial_guess_dual(batch, dev), None, None, None,
            SolverReturnStatus.SUCCESS, 0, 0.0,
            torch.zeros(lead + (self.maxiter,), dtype=torch.float64, device=dev),
            torch.zeros(lead + (self.maxiter,), dtype=torch.float64, device=dev),
            torch.zeros(lead + (self.maxiter,), dtype=torch.int32, device=dev),
            self.initial_guess_dual(batch, dev),
        )

    # base (unbatched) ranks of the per-instance arrays the kernels read
    _BATCH_NDIM = {
        "initial_state": 1, "final_state": 1, "reference_state_trajectory": 2, "reference_control_trajectory": 2,
        "weights_penalization_reference_state_trajectory": 1, "weights_penalization_control_squared": 1,
        "weights_penalization_final_state": 1,
    }

    def batch_params(self, params: Dict[str, Any], batch: int, device=None) -> Dict[str, Any]:
        """params with the per-instance arrays the kernels read expanded to
        contiguous (batch, ...) device tensors (differentiably: gradients of a
        shared value still sum over the batch) and, when params holds the
        model's parameters, their flat float32 block under
        FLAT_DYNAMICS_KEY.  Solving many times with the same parameters
        (closed-loop rollouts, training loops), pass the result instead of the
        shared values: each solve then reads them without first copying them
        out to the batch.  A solve whose weights override a dynamics
        parameter ignores the flat block and flattens the merged values."""
        out = dict(params)
        for k, nd in self._BATCH_NDIM.items():
            if k in params and params[k] is not None:
                out[k] = batched(params[k], batch, nd, torch.float64, device).contiguous()
        dyn = self.program.dynamics
        out.pop(FLAT_DYNAMICS_KEY, None)
        if dyn.param_keys and all(k in params for k in dyn.param_keys):
            out[FLAT_DYNAMICS_KEY] = dyn.flat_params(params, batch, device).contiguous()
        return out

    def make_params_with_weights(self, weights, problem_params=None):
        new_params = copy.copy(self.problem_params if problem_params is None else problem_params)
        for key, value in weights.items():
            new_params[key] = value
        return new_params

    def _differentiable_solve(self, initial_guess: SQPSolution, problem_params: Dict[str, Any] = None,
                              weights: Dict[str, Any] = None, device=None) -> SQPSolution:
        """Solve the batch of problems (sqp.py:436-560, 572-587)."""
        if problem_params is None:
            problem_params = self.problem_params
        weights = {} if weights is None else weights
        prog = self.program
        dyn_keys = set(prog.dynamics.param_keys)
        for k in weights:
            if k not in WEIGHT_KEYS and k not in dyn_keys:
                raise NotImplementedError(f"no gradient rule for weights key {k!r}; differentiable keys: "
                                          f"{sorted(WEIGHT_KEYS) + sorted(dyn_keys)}")
        N, nx, nu = prog.horizon, prog.num_state_variables, prog.num_control_variables
        states = torch.as_tensor(initial_guess.states)
        unbatched = states.dim() == 2
        if device is None:
            device = states.device if states.is_cuda else torch.device("cuda")

        def lead(t, nd, dtype=torch.float64):
            t = torch.as_tensor(t, device=device)
            t = t.unsqueeze(0) if unbatched and t.dim() == nd else t
            return t.to(dtype)

        g = {
            "states": lead(initial_guess.states, 2),
            "controls": lead(initial_guess.controls, 2),
            "dual": lead(initial_guess.dual, 2),
        }
        B = g["states"].shape[0]
        if g["states"].shape[1:] != (N + 1, nx) or g["controls"].shape[1:] != (N + 1, nu):
            raise ValueError("initial guess must have (B, N+1, nx) states and (B, N+1, nu) controls")

        def hist(v, dtype):
            t = torch.as_tensor(v, device=device)
            if t.dim() == 0 or t.numel() == 0:
                return torch.zeros(B, self.maxiter, dtype=dtype, device=device)
            t = t.reshape(-1, t.shape[-1]) if t.dim() >= 1 else t
            # one shared history row is expanded; any other shape is checked below
            return t.expand(B, self.maxiter).to(dtype) if tuple(t.shape) == (1, self.maxiter) else t.to(dtype)

        cev = initial_guess.convergence_error
        if isinstance(cev, torch.Tensor) and cev.is_cuda:
            ce = cev.to(torch.float64).reshape(-1)
        elif isinstance(cev, (int, float)):
            # a device fill, not a host-to-device copy (graph-capture safe)
            ce = torch.full((B,), float(cev), dtype=torch.float64, device=device)
        else:
            ce = torch.as_tensor(cev, dtype=torch.float64, device=device).reshape(-1)
        g["convergence_error"] = ce.expand(B) if ce.numel() == 1 else ce
        g["convergence_errors"] = hist(initial_guess.convergence_errors, torch.float64)
        g["linesearch_alphas"] = hist(initial_guess.linesearch_alphas, torch.float64)
        g["num_pcg_iters"] = hist(initial_guess.num_pcg_iters, torch.int32)
        if initial_guess.dual_backward_guess is not None:
            dbg = torch.as_tensor(initial_guess.dual_backward_guess)
            g["dual_backward_guess"] = dbg if (dbg.dim() == 3 and dbg.is_cuda) else lead(dbg, 2)
        else:
            g["dual_backward_guess"] = None

        # problem_params are stop-gradient (sqp.py:456-457); weights override them
        params = {k: (v.detach() if isinstance(v, torch.Tensor) else v) for k, v in problem_params.items()}
        params = self.make_params_with_weights(weights, params)

        def b(key, nd):
            return batched(params[key], B, nd, torch.float64, device)

        x0 = b("initial_state", 1)
        xf = b("final_state", 1)
        xref = b("reference_state_trajectory", 2)
        uref = b("reference_control_trajectory", 2)
        wx = b("weights_penalization_reference_state_trajectory", 1)
        wu = b("weights_penalization_control_squared", 1)
        wf = b("weights_penalization_final_state", 1)
        # dynamics parameters passed as weights keep their graph: the solve's
        # gradient with respect to the flat parameter block flows back to them
        flat = params.get(FLAT_DYNAMICS_KEY)
        if (isinstance(flat, torch.Tensor) and not (dyn_keys & set(weights))
                and flat.shape == (B, prog.dynamics.num_params) and flat.dtype == torch.float32
                and flat.device == torch.device(device) and flat.is_contiguous()):
            dyn = flat  # from batch_params (its graph, if any, is kept)
        else:
            dyn = prog.dynamics.flat_params({k: (v.detach() if isinstance(v, torch.Tensor) and k not in weights
                                                 else v) for k, v in params.items()}, B, device)
        horizons = _ragged_horizons(params.get("horizons"), B, N, device)
        _check_shapes(B, N, nx, nu, self.maxiter, prog.dynamics.num_params, g,
                      {"initial_state": x0, "final_state": xf, "reference_state_trajectory": xref,
                       "reference_control_trajectory": uref,
                       "weights_penalization_reference_state_trajectory": wx,
                       "weights_penalization_control_squared": wu, "weights_penalization_final_state": wf,
                       "dynamics parameters": dyn})
        if torch.device(device).type != "cuda":
            raise RuntimeError(f"SQPSolver.solve runs on the GPU (libdmpc.so); got device {device}")
        out = _SQPSolveFn.apply(self, g, x0, xf, xref, uref, wx, wu, wf, dyn, horizons)
        (x, u, lam, status, num_iter, conv_err, conv_errs, ls_alphas, pcg_iters, AB, fd, qinv, rinv,
         dual_bwd_out) = out
        saved = SavedLinearization(AB, fd, qinv, rinv)
        if unbatched:
            x, u, lam = x[0], u[0], lam[0]
            status, num_iter, conv_err = status[0], num_iter[0], conv_err[0]
            conv_errs, ls_alphas, pcg_iters, dual_bwd_out = conv_errs[0], ls_alphas[0], pcg_iters[0], dual_bwd_out[0]
        return SQPSolution(x, u, lam, saved, None, None, status, num_iter, conv_err, conv_errs, ls_alphas,
                           pcg_iters, dual_bwd_out)

    # ---- the pieces of one SQP iteration under the reference's names
    # (sqp.py:160-299, 413-424, 562-603).  The solve itself does not call
    # them: it runs the fused kernels.  They take one problem as the
    # reference does ((N+1, nx) trajectories) or a batch.

    def get_cost_terms(self, states, controls, problem_params: Dict[str, Any] = None):
        """(Qs (N+1, nx, nx), qs (N+1, nx), Rs (N, nu, nu), rs (N, nu)) of the
        quadratic cost model, Hessians projected onto eigenvalues >= 1e-9
        (sqp.py:160-197)."""
        QN, qN, Q, q, R, r = self.program.get_cost_linearized_matrices(
            states, controls, self.problem_params if problem_params is None else problem_params)
        Qs = torch.cat([Q, QN.unsqueeze(-3)], dim=-3)
        qs = torch.cat([q, qN.unsqueeze(-2)], dim=-2)
        return (project_matrix_onto_positive_semidefinite_cone(Qs, PSD_MIN_EIGENVALUE), qs,
                project_matrix_onto_positive_semidefinite_cone(R, PSD_MIN_EIGENVALUE), r)

    def get_dynamics_constraints_terms(self, states, controls, problem_params: Dict[str, Any] = None):
        """QPDynamicsParameters (As_next, As, Bs, Cs) of the dynamics linearized
        at (states, controls) (sqp.py:199-228), on the GPU."""
        return QPDynamicsParameters(*self.program.get_dynamics_linearized_matrices_with_states_controls(
            states, controls, self.problem_params if problem_params is None else problem_params))

    def compute_qp_parameters(self, states, controls, problem_params: Dict[str, Any] = None):
        """(QPParameters, QPDynamicsCostPremultipliedMatrices,
        SchurComplementMatrices) at (states, controls) (sqp.py:230-260)."""
        Qs, qs, Rs, rs = self.get_cost_terms(states, controls, problem_params)
        dyn = self.get_dynamics_constraints_terms(states, controls, problem_params)
        Qinv = torch.linalg.inv(Qs)
        Rinv = torch.linalg.inv(Rs)
        qp = QPParameters(dyn, QPCostParameters(Qinv, qs, Rinv, rs))
        pre = QPDynamicsCostPremultipliedMatrices(dyn.As @ Qinv[..., :-1, :, :], dyn.As_next @ Qinv[..., 1:, :, :],
                                                  dyn.Bs @ Rinv)
        return qp, pre, self.pcg.compute_S_Phiinv(qp)

    def solve_one_scp_iteration(self, qp_solution: QPSolution, problem_params: Dict[str, Any] = None):
        """One QP of the SQP iteration at qp_solution (sqp.py:262-299): returns
        (QPSolution, QPParameters, premultiplied matrices, Schur matrices,
        PCG iteration count), through the PCG layer's entry points."""
        qp, pre, schur = self.compute_qp_parameters(qp_solution.states, qp_solution.controls, problem_params)
        gammas = self.pcg.compute_gamma(pre, qp.dynamics.Cs, qp.cost.Qinv[..., 0, :, :], qp.cost.qvec, qp.cost.rvec)
        sol, dbg = self.pcg.solve_KKT_Schur(qp, schur, gammas, qp_solution.kkt_multipliers)
        return sol, qp, pre, schur, dbg.num_iterations

    @staticmethod
    def state_control_difference(states, controls, new_states, new_controls):
        """max(max |new_states - states|, max |new_controls - controls|) over
        each problem's trajectory (sqp.py:413-424)."""
        return torch.maximum((new_states - states).abs().amax(dim=(-2, -1)),
                             (new_controls - controls).abs().amax(dim=(-2, -1)))

    def get_differentiable_solve_function(self):
        """solve(initial_guess, problem_params, weights), differentiable in the
        weights (sqp.py:562-603)."""
        return self.solve

    def qp_parameters(self, solution: SQPSolution) -> QPParameters:
        """Dense QP blocks of the last linearization point of a batched solution
        (sqp.py:230-260); C and q, r are not kept by the solve and come back zero."""
        s = solution.qp_parameters
        B, N, nx, nz = s.AB.shape
        nu = nz - nx
        dev = s.AB.device
        As = s.AB[..., :nx].to(torch.float64)
        Bs = s.AB[..., nx:].to(torch.float64)
        As_next = -torch.eye(nx, dtype=torch.float64, device=dev).expand(B, N, nx, nx).contiguous()
        Qinv = torch.diag_embed(s.qinv)
        Rinv = torch.diag_embed(s.rinv.unsqueeze(1).expand(B, N, nu))
        z = torch.zeros
        return QPParameters(QPDynamicsParameters(As_next, As.contiguous(), Bs.contiguous(),
                                                 z(B, N + 1, nx, dtype=torch.float64, device=dev)),
                            QPCostParameters(Qinv.contiguous(), z(B, N + 1, nx, dtype=torch.float64, device=dev),
                                             Rinv.contiguous(), z(B, N, nu, dtype=torch.float64, device=dev)))

    def schur_complement_matrices(self, solution: SQPSolution) -> SchurComplementMatrices:
        """Explicit S, Phiinv (reference layout) of the last linearization point."""
        return self.pcg.compute_S_Phiinv(self.qp_parameters(solution))
