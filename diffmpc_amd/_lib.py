"""ctypes binding of libdmpc.so (the C ABI declared in include/dmpc.h).

The library is built in-tree by ``diffmpc_amd/csrc/build.sh`` (or
``__graft_entry__.build()``).  torch is imported first so that the HIP runtime
torch ships is the one the library binds to (both use the SONAME
libamdhip64.so.7).  There is no fallback: if the library is missing, every
solver call raises.
"""
import ctypes
import os

import torch  # noqa: F401  (must be loaded before libdmpc.so)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdmpc.so")

DMPC_OK = 0
DMPC_ERR_ARG = -1
DMPC_ERR_UNSUPPORTED = -2
DMPC_ERR_HIP = -3
DMPC_ERR_TOO_LARGE = -4
NOALLOC_SCRATCH_BYTES = 64 << 20  # DMPC_NOALLOC_SCRATCH_BYTES

MODE_FUSED = 0
MODE_PLAIN = 1

MODEL_LINEAR = 0
MODEL_QUADROTOR = 1
MODEL_SPACECRAFT3 = 2
MODEL_SPACECRAFT7 = 3
MODEL_CARTPOLE = 4
MODEL_CHAIN = 5

_c_double_p = ctypes.POINTER(ctypes.c_double)
_c_float_p = ctypes.POINTER(ctypes.c_float)
_c_int_p = ctypes.POINTER(ctypes.c_int32)


class Problem(ctypes.Structure):
    _fields_ = [
        ("model", ctypes.c_int32), ("nx", ctypes.c_int32), ("nu", ctypes.c_int32),
        ("horizon", ctypes.c_int32), ("batch", ctypes.c_int32), ("scheme", ctypes.c_int32),
        ("n_params", ctypes.c_int32), ("penalize_control_reference", ctypes.c_int32),
        ("dt", ctypes.c_double),
    ]


class SolverParams(ctypes.Structure):
    _fields_ = [
        ("num_scp_iteration_max", ctypes.c_int32), ("tol_convergence", ctypes.c_double),
        ("linesearch", ctypes.c_int32), ("linesearch_eta", ctypes.c_double),
        ("n_alphas", ctypes.c_int32), ("linesearch_alphas", ctypes.c_double * 8),
        ("pcg_max_iter", ctypes.c_int32), ("pcg_tol_epsilon", ctypes.c_double),
        ("warm_start_backward", ctypes.c_int32), ("blocks_per_group", ctypes.c_int32),
        ("fused_variant", ctypes.c_int32), ("instances_per_workgroup", ctypes.c_int32),
        ("forward_schedule", ctypes.c_int32),
    ]


class ProblemData(ctypes.Structure):
    _fields_ = [
        ("x0", ctypes.c_void_p), ("x_final", ctypes.c_void_p), ("x_ref", ctypes.c_void_p),
        ("u_ref", ctypes.c_void_p), ("w_x", ctypes.c_void_p), ("w_u", ctypes.c_void_p),
        ("w_f", ctypes.c_void_p), ("dyn_params", ctypes.c_void_p), ("horizons", ctypes.c_void_p),
    ]


class Solution(ctypes.Structure):
    _fields_ = [
        ("states", ctypes.c_void_p), ("controls", ctypes.c_void_p), ("dual", ctypes.c_void_p),
        ("status", ctypes.c_void_p), ("num_iter", ctypes.c_void_p),
        ("convergence_error", ctypes.c_void_p), ("convergence_errors", ctypes.c_void_p),
        ("linesearch_alphas", ctypes.c_void_p), ("num_pcg_iters", ctypes.c_void_p),
    ]


class Saved(ctypes.Structure):
    _fields_ = [("AB", ctypes.c_void_p), ("fd", ctypes.c_void_p), ("qinv", ctypes.c_void_p), ("rinv", ctypes.c_void_p),
                ("work", ctypes.c_void_p), ("work_bytes", ctypes.c_uint64)]


class Grads(ctypes.Structure):
    _fields_ = [
        ("w_x", ctypes.c_void_p), ("w_u", ctypes.c_void_p), ("w_f", ctypes.c_void_p),
        ("x0", ctypes.c_void_p), ("x_ref", ctypes.c_void_p), ("u_ref", ctypes.c_void_p),
        ("x_final", ctypes.c_void_p), ("dual_backward", ctypes.c_void_p), ("pcg_iters", ctypes.c_void_p),
        ("adj_states", ctypes.c_void_p), ("adj_controls", ctypes.c_void_p),
    ]


# name -> (restype, argtypes); the list is the public C ABI
_SIGNATURES = {
    "dmpc_version": (ctypes.c_char_p, []),
    "dmpc_custom_model_info": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int32)] * 3),
    "dmpc_supported": (ctypes.c_int, [ctypes.c_int32] * 5),
    "dmpc_sqp_solve": (ctypes.c_int, [ctypes.POINTER(Problem), ctypes.POINTER(SolverParams), ctypes.POINTER(ProblemData),
                                      ctypes.POINTER(Solution), ctypes.POINTER(Saved), ctypes.c_int32, ctypes.c_void_p]),
    "dmpc_sqp_solve_noalloc": (ctypes.c_int, [ctypes.POINTER(Problem), ctypes.POINTER(SolverParams),
                                              ctypes.POINTER(ProblemData), ctypes.POINTER(Solution),
                                              ctypes.POINTER(Saved), ctypes.c_void_p]),
    "dmpc_sqp_backward": (ctypes.c_int, [ctypes.POINTER(Problem), ctypes.POINTER(SolverParams), ctypes.POINTER(ProblemData),
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Saved), ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.POINTER(Grads), ctypes.c_int32, ctypes.c_void_p]),
    "dmpc_linearize": (ctypes.c_int, [ctypes.POINTER(Problem), ctypes.POINTER(ProblemData)] + [ctypes.c_void_p] * 7),
    "dmpc_simulate": (ctypes.c_int, [ctypes.POINTER(Problem), ctypes.POINTER(ProblemData)] + [ctypes.c_void_p] * 3),
    "dmpc_dynamics_param_grads": (ctypes.c_int, [ctypes.POINTER(Problem), ctypes.POINTER(ProblemData)]
                                  + [ctypes.c_void_p] * 8),
    "dmpc_pcg_assemble": (ctypes.c_int, [ctypes.c_int32] * 4 + [ctypes.c_void_p] * 8),
    "dmpc_pcg_gamma": (ctypes.c_int, [ctypes.c_int32] * 4 + [ctypes.c_void_p] * 10),
    "dmpc_pcg_solve": (ctypes.c_int, [ctypes.c_int32] * 3 + [ctypes.c_void_p] * 4 + [ctypes.c_int32, ctypes.c_double]
                       + [ctypes.c_void_p] * 3),
    "dmpc_pcg_recover": (ctypes.c_int, [ctypes.c_int32] * 4 + [ctypes.c_void_p] * 11),
    "dmpc_set_blocks_per_group": (ctypes.c_int, [ctypes.c_int32]),
    "dmpc_set_fused_variant": (ctypes.c_int, [ctypes.c_int32]),
    "dmpc_workspace_bytes": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(Problem), ctypes.POINTER(SolverParams),
                                            ctypes.POINTER(ctypes.c_uint64)]),
    "dmpc_profile_enable": (ctypes.c_int, [ctypes.c_int32]),
    "dmpc_profile_read": (ctypes.c_int, [ctypes.c_int32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32)]),
}

PROF_QP_ITERATION = 0
PROF_BACKWARD = 1
PROF_LINEARIZE = 2
PROF_DYN_GRADS = 3


def profile_read(kernel, lib=None):
    """(total ms, launches) of one instrumented kernel since the last read
    (of libdmpc.so, or of the loaded library `lib`)."""
    lib = lib or load()
    ms = ctypes.c_double(0.0)
    n = ctypes.c_int32(0)
    check(lib.dmpc_profile_read(kernel, ctypes.byref(ms), ctypes.byref(n)), "dmpc_profile_read")
    return ms.value, n.value

_lib = None


class DmpcError(RuntimeError):
    pass


_custom_libs = {}


def _open(path):
    if not os.path.exists(path):
        raise DmpcError(
            f"{path} is missing: build it with diffmpc_amd/csrc/build.sh "
            "(there is no CPU fallback for the solver)"
        )
    lib = ctypes.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def load(path=None):
    """Load libdmpc.so (once), or the custom build at `path`; raise if it has
    not been built."""
    global _lib
    if path is not None and os.path.abspath(path) != os.path.abspath(LIB_PATH):
        key = os.path.abspath(path)
        if key not in _custom_libs:
            _custom_libs[key] = _open(key)
        return _custom_libs[key]
    if _lib is None:
        _lib = _open(LIB_PATH)
    return _lib


def load_for(dynamics):
    """The library that runs `dynamics`: the custom build for a user-defined
    model (diffmpc_amd.dynamics.CustomDynamics), else libdmpc.so."""
    path = getattr(dynamics, "library_path", None)
    return load(path() if callable(path) else None)


def workspace_bytes(mode, prob, sp, lib=None):
    """Device workspace (bytes) dmpc_sqp_solve / dmpc_sqp_backward need."""
    n = ctypes.c_uint64(0)
    check((lib or load()).dmpc_workspace_bytes(mode, ctypes.byref(prob), ctypes.byref(sp), ctypes.byref(n)),
          "dmpc_workspace_bytes")
    return n.value


def exported_symbols():
    return sorted(_SIGNATURES)


def check(rc, what):
    if rc == DMPC_OK:
        return
    names = {DMPC_ERR_ARG: "bad argument", DMPC_ERR_UNSUPPORTED: "unsupported configuration", DMPC_ERR_HIP: "HIP error",
             DMPC_ERR_TOO_LARGE: "workspace larger than the allocation-free entry point's scratch"}
    raise DmpcError(f"{what} failed: {names.get(rc, rc)} ({rc})")


def ptr(t):
    """Device pointer of a contiguous CUDA/HIP tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise DmpcError("libdmpc needs device tensors (got a CPU tensor)")
    if not t.is_contiguous():
        raise DmpcError("libdmpc needs contiguous tensors")
    return ctypes.c_void_p(t.data_ptr())


def stream_ptr(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)
