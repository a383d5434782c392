"""Autotuner for the fused kernels' launch shape.

A fused solve can run as the register-resident kernels with K = 1, 2 or 4
time blocks per lane group (K trades threads per instance against registers
per thread: the on-chip tiling of the block-tridiagonal system), with one
block per group also as 2, 3 or 4 instances per workgroup (teams aligned to
16 lanes, which fill the wavefronts that a short instance of a small model
leaves mostly idle), or as the streamed variant that keeps the blocks in a
device workspace.  A candidate is (variant, blocks per group, instances per
workgroup).  ``tune`` times every candidate the library accepts on the
caller's own problem (forward plus backward: the library's HIP events
around the linearization, QP-iteration and adjoint launches, i.e. the
kernels a launch shape changes -- a batch the register-resident kernels
solve in one launch linearizes inside it -- so host overhead does not blur
small problems; wall clock on the current stream where the library records
no launches), writes the fastest
into the solver's params (``fused_variant``,
``blocks_per_group``, ``instances_per_workgroup``) and caches it per problem
shape (model, sizes, batch, ragged horizons, mode, line search, SQP and PCG
settings).  No reference counterpart: XLA picks its own launch shapes.
"""
import hashlib
from typing import Dict, List, Tuple

import torch

from . import _lib

CANDIDATES: List[Tuple[str, int, int]] = [
    ("register", 1, 1), ("register", 2, 1), ("register", 4, 1), ("streamed", 0, 1),
    ("register", 1, 2), ("register", 1, 3), ("register", 1, 4),
]

_cache: Dict[tuple, Tuple[str, int, int]] = {}


def register_horizon_limit(model: int, nx: int, nu: int, linesearch: bool = False, n_alphas: int = 1,
                           nmax: int = 4096) -> int:
    """Longest horizon the register-resident fused kernels take for a model
    (0 if none): the automatic variant choice needs no workspace there."""
    sp = _lib.SolverParams()
    sp.num_scp_iteration_max, sp.pcg_max_iter, sp.n_alphas = 1, 1, n_alphas
    sp.linesearch = int(linesearch)
    for i in range(n_alphas):
        sp.linesearch_alphas[i] = 1.0 if not linesearch else 0.5 + 0.5 * i / max(n_alphas - 1, 1)

    def fits(n):
        prob = _lib.Problem(model, nx, nu, n, 1, 0, 1 << 20, 0, 0.1)  # n_params: any sufficient count
        return _lib.workspace_bytes(_lib.MODE_FUSED, prob, sp) == 0

    if not fits(1):
        return 0
    lo, hi = 1, nmax
    if fits(hi):
        return hi
    while hi - lo > 1:
        mid = (lo + hi) // 2
        lo, hi = (mid, hi) if fits(mid) else (lo, mid)
    return lo


def _key(solver, params, B):
    """Everything the timings depend on: model and sizes, batch, the ragged
    horizons (the exact mix, by digest), mode, line search, SQP and PCG caps."""
    prog, p = solver.program, solver.params
    h = params.get("horizons") if isinstance(params, dict) else None
    hkey = None if h is None else hashlib.sha1(torch.as_tensor(h).cpu().numpy().tobytes()).hexdigest()
    return (prog.dynamics.model_id, prog.num_state_variables, prog.num_control_variables, prog.horizon, B, hkey,
            p.get("mode", "auto"), bool(p["linesearch"]), tuple(p["linesearch_alphas"]),
            p["num_scp_iteration_max"], float(p["tol_convergence"]), int(p["pcg"]["max_iter"]),
            float(p["pcg"]["tol_epsilon"]))


def _apply(solver, choice):
    variant, k, ipw = choice
    solver.params["fused_variant"], solver.params["blocks_per_group"] = variant, k
    solver.params["instances_per_workgroup"] = 0 if ipw == 1 else ipw


def tune(solver, params, guess, reps: int = 2, backward: bool = True, candidates=CANDIDATES):
    """Pick the fastest fused launch shape for (solver, params); returns
    (choice, {choice: ms per fwd(+bwd)}).  Cached choices are applied without
    timing (the timings dict is then empty)."""
    B = guess.states.shape[0] if guess.states.dim() == 3 else 1
    key = _key(solver, params, B)
    if key in _cache:
        _apply(solver, _cache[key])
        return _cache[key], {}
    wkey = "weights_penalization_reference_state_trajectory"

    def run():
        if guess.dual_backward_guess is not None:
            guess.dual_backward_guess.zero_()
        if not backward:
            with torch.no_grad():
                return solver.solve(guess, params)
        w = params[wkey].detach().clone().requires_grad_(True)
        sol = solver.solve(guess, params, {wkey: w})
        torch.autograd.grad(sol.states.square().sum(), [w])
        return sol

    timings = {}
    saved = (solver.params.get("fused_variant", "auto"), solver.params.get("blocks_per_group", 0),
             solver.params.get("instances_per_workgroup", 0) or 1)
    lib = _lib.load_for(solver.program.dynamics)
    kernels = (_lib.PROF_LINEARIZE, _lib.PROF_QP_ITERATION, _lib.PROF_BACKWARD)
    for choice in candidates:
        _apply(solver, choice)
        try:
            run()
        except _lib.DmpcError as e:
            if "unsupported" in str(e):
                continue
            raise
        torch.cuda.synchronize()
        for k in kernels:
            _lib.profile_read(k, lib)
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        lib.dmpc_profile_enable(1)
        try:
            start.record()
            for _ in range(reps):
                run()
            end.record()
            end.synchronize()
        finally:
            lib.dmpc_profile_enable(0)
        kernel_ms = sum(_lib.profile_read(k, lib)[0] for k in kernels)
        timings[choice] = (kernel_ms if kernel_ms > 0.0 else start.elapsed_time(end)) / reps
    if not timings:
        _apply(solver, saved)
        raise _lib.DmpcError("no fused launch shape supports this problem")
    best = min(timings, key=timings.get)
    _apply(solver, best)
    _cache[key] = best
    return best, timings
