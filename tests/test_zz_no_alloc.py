"""dmpc_sqp_solve_noalloc against dmpc_sqp_solve, through the C ABI.

The allocation-free entry point runs the same kernels with the streamed
variant's workspace in the library's own device array instead of the
caller's buffer, so its outputs must equal the existing entry point's bit
for bit.  Two calls in a row, the second smaller (the array keeps the first
call's contents), on problems that run the streamed variant (which uses the
workspace) and one that runs the register-resident kernels (which do not);
a problem whose workspace exceeds the array is refused before any launch.
"""
import ctypes
import sys

import pytest
import torch

from diffmpc_amd import _lib
from diffmpc_amd import workloads as W
from diffmpc_amd.solvers.sqp import _c_solver_params

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _inputs(name, prob, sp, dev):
    """Solver, C problem / params / data structs and the device tensors behind
    them, for a workload dict."""
    solver, params, guess = W.make_solver(name, prob, sp, device=dev)
    prog = solver.program
    B = guess.states.shape[0]
    keys = ["initial_state", "final_state", "reference_state_trajectory", "reference_control_trajectory",
            "weights_penalization_reference_state_trajectory", "weights_penalization_control_squared",
            "weights_penalization_final_state"]
    data_t = [params[k].to(torch.float64).contiguous() for k in keys]
    data_t.append(prog.dynamics.flat_params(params, B, dev).contiguous())
    data = _lib.ProblemData(*[_lib.ptr(t) for t in data_t], None)
    return solver, guess, prog.c_problem(B), _c_solver_params(solver.params), data, data_t


def _solve(lib, entry, solver, guess, prob, sp, data, work_bytes=None):
    """One call of `entry` ("dmpc_sqp_solve" or "dmpc_sqp_solve_noalloc") on
    fresh copies of the guess; returns the output tensors."""
    dev = guess.states.device
    prog = solver.program
    B, N, nx, nu = guess.states.shape[0], prog.horizon, prog.num_state_variables, prog.num_control_variables
    maxit = solver.maxiter
    out = {
        "x": guess.states.clone(), "u": guess.controls.clone(), "lam": guess.dual.clone(),
        "status": torch.empty(B, dtype=torch.int32, device=dev), "num_iter": torch.empty(B, dtype=torch.int32, device=dev),
        "conv_err": torch.zeros(B, dtype=torch.float64, device=dev),
        "conv_errs": torch.zeros(B, maxit, dtype=torch.float64, device=dev),
        "alphas": torch.zeros(B, maxit, dtype=torch.float64, device=dev),
        "pcg": torch.zeros(B, maxit, dtype=torch.int32, device=dev),
        "AB": torch.empty(B, N, nx, nx + nu, dtype=torch.float32, device=dev),
        "fd": torch.empty(B, N, nx, dtype=torch.float64, device=dev),
        "qinv": torch.empty(B, N + 1, nx, dtype=torch.float64, device=dev),
        "rinv": torch.empty(B, nu, dtype=torch.float64, device=dev),
    }
    sol = _lib.Solution(*[_lib.ptr(out[k]) for k in ("x", "u", "lam", "status", "num_iter", "conv_err", "conv_errs",
                                                     "alphas", "pcg")])
    work = None
    if entry == "dmpc_sqp_solve" and work_bytes:
        work = torch.empty(work_bytes, dtype=torch.uint8, device=dev)
    saved = _lib.Saved(_lib.ptr(out["AB"]), _lib.ptr(out["fd"]), _lib.ptr(out["qinv"]), _lib.ptr(out["rinv"]),
                       _lib.ptr(work) if work is not None else None, work_bytes or 0)
    stream = _lib.stream_ptr(dev)
    if entry == "dmpc_sqp_solve":
        rc = lib.dmpc_sqp_solve(ctypes.byref(prob), ctypes.byref(sp), ctypes.byref(data), ctypes.byref(sol),
                                ctypes.byref(saved), _lib.MODE_FUSED, stream)
    else:
        rc = lib.dmpc_sqp_solve_noalloc(ctypes.byref(prob), ctypes.byref(sp), ctypes.byref(data), ctypes.byref(sol),
                                        ctypes.byref(saved), stream)
    _lib.check(rc, entry)
    torch.cuda.synchronize()
    return out


CASES = [
    # (label, workload): streamed (N > 63 for the quadrotor), larger first
    ("quadrotor B=6 N=100 (streamed)", lambda: W.quadrotor(batch=6, seed=21, horizon=100, sqp_iters=2)),
    ("quadrotor B=3 N=80 (streamed)", lambda: W.quadrotor(batch=3, seed=22, horizon=80, sqp_iters=2)),
]


def test_noalloc_entry_point_equals_the_allocating_one_on_two_calls(cuda):
    lib = _lib.load()
    for label, gen in CASES:
        say(f"[noalloc] {label}")
        name, prob, sp = gen()
        solver, guess, cprob, csp, data, keep = _inputs(name, prob, sp, cuda)
        need = _lib.workspace_bytes(_lib.MODE_FUSED, cprob, csp)
        assert 0 < need <= _lib.NOALLOC_SCRATCH_BYTES, need
        ref = _solve(lib, "dmpc_sqp_solve", solver, guess, cprob, csp, data, work_bytes=need)
        got = _solve(lib, "dmpc_sqp_solve_noalloc", solver, guess, cprob, csp, data)
        for k in ref:
            assert torch.equal(ref[k], got[k]), (label, k)
        assert float(ref["x"].abs().max()) > 0.0


def test_noalloc_register_resident_problem_needs_no_scratch(cuda):
    lib = _lib.load()
    say("[noalloc] quadrotor B=8 N=30 (register-resident)")
    name, prob, sp = W.quadrotor(batch=8, seed=23, horizon=30, sqp_iters=2)
    solver, guess, cprob, csp, data, keep = _inputs(name, prob, sp, cuda)
    assert _lib.workspace_bytes(_lib.MODE_FUSED, cprob, csp) == 0
    ref = _solve(lib, "dmpc_sqp_solve", solver, guess, cprob, csp, data)
    got = _solve(lib, "dmpc_sqp_solve_noalloc", solver, guess, cprob, csp, data)
    for k in ref:
        assert torch.equal(ref[k], got[k]), k


def test_noalloc_refuses_a_workspace_beyond_the_scratch(cuda):
    """Checked on the host before any launch: no kernel runs."""
    lib = _lib.load()
    name, prob, sp = W.quadrotor(batch=2, seed=24, horizon=200, sqp_iters=1)
    solver, guess, cprob, csp, data, keep = _inputs(name, prob, sp, cuda)
    cprob.batch = 4096  # the descriptor only: nothing is launched for it
    assert _lib.workspace_bytes(_lib.MODE_FUSED, cprob, csp) > _lib.NOALLOC_SCRATCH_BYTES
    rc = lib.dmpc_sqp_solve_noalloc(ctypes.byref(cprob), ctypes.byref(csp), ctypes.byref(data),
                                    ctypes.byref(_lib.Solution(*([ctypes.c_void_p(1)] * 9))),
                                    ctypes.byref(_lib.Saved(ctypes.c_void_p(1), ctypes.c_void_p(1),
                                                            ctypes.c_void_p(1), ctypes.c_void_p(1), None, 0)),
                                    _lib.stream_ptr(cuda))
    assert rc == _lib.DMPC_ERR_TOO_LARGE
