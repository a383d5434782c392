"""User-defined dynamics on the GPU kernels.

The reference lets a user subclass ``Dynamics`` with any JAX ``state_dot``
and takes its Jacobians by autodiff (diffmpc/problems/optimal_control_problem.py:360-381).
Here the device kernels need the model as device code, so a user model is
written once as a Python function over a small math namespace,

    def state_dot(x, u, p, m):          # x[i], u[k], p[j]; m.sin/cos/sqrt/exp
        return [u[0] * m.cos(x[2]), u[0] * m.sin(x[2]), p[0] * u[1]]

and ``CustomDynamics`` uses that function three ways:

* traced with symbolic values into C++ (``cpp_model_source``): the body of a
  ``dmpc::CustomModel`` that the library is compiled with once more
  (``DMPC_MODEL_CUSTOM``); the kernels then run it in float32 dual numbers
  (linearization), float64 (line search) and float64 hyper-dual numbers
  (dynamics-parameter gradients) exactly like the built-in models;
* evaluated on torch tensors for the host API (``state_dot``, simulators);
* evaluated on numpy arrays by the tests' float64 oracle.

Supported operations: + - * / on values and numbers, unary -, integer powers
``v ** n`` (n >= 0), and ``m.sin``, ``m.cos``, ``m.sqrt``, ``m.exp``.
"""
import hashlib
import os
import subprocess
from typing import Any, Callable, Dict, List, Optional, Sequence, Tuple

import torch

from .base_dynamics import Dynamics

MODEL_CUSTOM = 6

_CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "csrc")


# ---------------------------------------------------------------------------
# Tracing into C++
class _Expr:
    __slots__ = ("op", "args", "val")

    def __init__(self, op, args=(), val=None):
        self.op, self.args, self.val = op, tuple(args), val

    @staticmethod
    def wrap(v):
        if isinstance(v, _Expr):
            return v
        if isinstance(v, (int, float)):
            return _Expr("const", val=float(v))
        raise TypeError(f"unsupported operand in a custom model: {type(v).__name__}")

    def __add__(self, o):
        return _Expr("add", (self, _Expr.wrap(o)))

    def __radd__(self, o):
        return _Expr("add", (_Expr.wrap(o), self))

    def __sub__(self, o):
        return _Expr("sub", (self, _Expr.wrap(o)))

    def __rsub__(self, o):
        return _Expr("sub", (_Expr.wrap(o), self))

    def __mul__(self, o):
        return _Expr("mul", (self, _Expr.wrap(o)))

    def __rmul__(self, o):
        return _Expr("mul", (_Expr.wrap(o), self))

    def __truediv__(self, o):
        return _Expr("div", (self, _Expr.wrap(o)))

    def __rtruediv__(self, o):
        return _Expr("div", (_Expr.wrap(o), self))

    def __neg__(self):
        return _Expr("neg", (self,))

    def __pos__(self):
        return self

    def __pow__(self, n):
        if isinstance(n, int) and n >= 0:
            if n == 0:
                return _Expr.wrap(1.0)
            r = self
            for _ in range(n - 1):
                r = r * self
            return r
        if n == 0.5:
            return _Expr("sqrt", (self,))
        raise TypeError("custom models take integer powers >= 0 (or 0.5)")


class _SymMath:
    @staticmethod
    def sin(v):
        return _Expr("sin", (_Expr.wrap(v),))

    @staticmethod
    def cos(v):
        return _Expr("cos", (_Expr.wrap(v),))

    @staticmethod
    def sqrt(v):
        return _Expr("sqrt", (_Expr.wrap(v),))

    @staticmethod
    def exp(v):
        return _Expr("exp", (_Expr.wrap(v),))


class TorchMath:
    """The math namespace for torch evaluation."""
    sin = staticmethod(torch.sin)
    cos = staticmethod(torch.cos)
    sqrt = staticmethod(torch.sqrt)
    exp = staticmethod(torch.exp)


_BINOP = {"add": "+", "sub": "-", "mul": "*", "div": "/"}
_FUNC = {"sin": "dsin", "cos": "dcos", "sqrt": "dsqrt", "exp": "dexp"}


def cpp_model_source(state_dot: Callable, nx: int, nu: int, n_params: int, name: str = "custom") -> str:
    """C++ definition of dmpc::CustomModel for `state_dot` (see the module
    docstring for the function's form)."""
    xs = [_Expr("leaf", val=f"x[{i}]") for i in range(nx)]
    us = [_Expr("leaf", val=f"u[{k}]") for k in range(nu)]
    ps = [_Expr("leaf", val=f"P{j}") for j in range(n_params)]
    out = list(state_dot(xs, us, ps, _SymMath))
    if len(out) != nx:
        raise ValueError(f"state_dot returned {len(out)} components, expected {nx}")
    out = [_Expr.wrap(o) for o in out]
    lines: List[str] = []
    names: Dict[int, str] = {}

    def ref(e: _Expr) -> str:
        if e.op == "leaf":
            return e.val
        if e.op == "const":
            return f"sc<T>(S({e.val!r}))"
        key = id(e)
        if key in names:
            return names[key]
        if e.op in _BINOP:
            a, b = ref(e.args[0]), ref(e.args[1])
            expr = f"{a} {_BINOP[e.op]} {b}"
        elif e.op == "neg":
            expr = f"-{ref(e.args[0])}"
        else:
            expr = f"{_FUNC[e.op]}({ref(e.args[0])})"
        nm = f"v{len(names)}"
        names[key] = nm
        lines.append(f"    const T {nm} = {expr};")
        return nm

    results = [ref(o) for o in out]
    uses_constants = any("sc<T>(S(" in ln for ln in lines + results)
    body = ["    using S = scalar_t<T>;"] if uses_constants else []
    body += [f"    const T P{j} = sc<T>(p[{j}]);" for j in range(n_params)] + lines
    body += [f"    xd[{i}] = {r};" for i, r in enumerate(results)]
    return "\n".join([
        f"// generated by diffmpc_amd.dynamics.custom_dynamics for the model '{name}'",
        "struct CustomModel {",
        f"  static constexpr int NX = {nx}, NU = {nu}, NPARAM = {n_params};",
        "  template <class T, class PA>",
        "  __device__ static void f(const T* x, const T* u, PA p, T* xd) {",
        *body,
        "  }",
        "};",
        "",
    ])


_STAMP = "library_sources.sha256"


def _sources_digest() -> str:
    h = hashlib.sha256()
    for fn in sorted(os.listdir(_CSRC)):
        if fn.endswith((".h", ".hip", ".sh")):
            with open(os.path.join(_CSRC, fn), "rb") as f:
                h.update(fn.encode() + f.read())
    return h.hexdigest()


def build_custom_library(source: str, out_dir: str) -> str:
    """Compile the library with `source` as its CustomModel into
    out_dir/libdmpc_custom.so (hipcc for gfx950, about half a minute)."""
    os.makedirs(out_dir, exist_ok=True)
    header = os.path.join(out_dir, "custom_model.h")
    with open(header, "w") as f:
        f.write(source)
    lib = os.path.join(out_dir, "libdmpc_custom.so")
    env = dict(os.environ, DMPC_CUSTOM_MODEL_FILE=os.path.abspath(header), DMPC_LIB_OUT=os.path.abspath(lib))
    subprocess.run(["bash", os.path.join(_CSRC, "build.sh")], env=env, check=True, stdout=subprocess.DEVNULL)
    with open(os.path.join(out_dir, _STAMP), "w") as f:
        f.write(_sources_digest())
    return lib


def _is_fresh(out_dir: str, source: str) -> bool:
    """The library in out_dir was built from `source` and the current library
    sources."""
    paths = [os.path.join(out_dir, n) for n in ("libdmpc_custom.so", "custom_model.h", _STAMP)]
    if not all(os.path.exists(p) for p in paths):
        return False
    with open(paths[1]) as f, open(paths[2]) as g:
        return f.read() == source and g.read() == _sources_digest()


# in-package location of prebuilt custom libraries (one folder per model name)
PREBUILT_ROOT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "_custom")


class CustomBuild:
    """A model compiled into its own build of the library (model id
    MODEL_CUSTOM).  Needs `source` (the C++ definition of dmpc::CustomModel)
    and `name`; `_build_dir` may pin the folder of the build."""

    source: str = ""
    name: str = "custom"
    _build_dir: Optional[str] = None
    _lib_path: Optional[str] = None

    def prebuilt_dir(self) -> str:
        """The in-package folder a prebuilt library of this model lives in."""
        return os.path.join(PREBUILT_ROOT, self.name)

    def build_dir(self) -> str:
        """The pinned folder; else the in-package prebuilt folder when it holds
        a current build of this source; else a cache folder keyed by the
        source and the library sources."""
        if self._build_dir is not None:
            return self._build_dir
        if _is_fresh(self.prebuilt_dir(), self.source):
            return self.prebuilt_dir()
        key = hashlib.sha256((self.source + _sources_digest()).encode()).hexdigest()[:16]
        root = os.environ.get("DMPC_CUSTOM_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "diffmpc_amd"))
        return os.path.join(root, f"{self.name}-{key}")

    def library_path(self, build: bool = True) -> str:
        """Path of the compiled library for this model (compiled on first use
        when `build`, and again when the library sources have changed since)."""
        if self._lib_path is None:
            d = self.build_dir()
            lib = os.path.join(d, "libdmpc_custom.so")
            if not _is_fresh(d, self.source):
                if not build:
                    raise FileNotFoundError(f"{lib} is not built for this model")
                build_custom_library(self.source, d)
            self._lib_path = lib
        return self._lib_path


# ---------------------------------------------------------------------------
class CustomDynamics(CustomBuild, Dynamics):
    """A user-defined model (see the module docstring).

    state_dot(x, u, p, m): the model; parameters: (key, size) pairs giving
    the flat parameter layout, read from the problem parameters by key (and
    differentiable when passed as weights); build_dir: where the compiled
    library lives (default: a cache directory keyed by the generated source
    and the library sources).
    """

    model_id = MODEL_CUSTOM

    def __init__(self, num_states: int, num_controls: int, state_dot: Callable,
                 parameters: Sequence[Tuple[str, int]] = (), name: str = "custom",
                 names_states: Optional[List[str]] = None, names_controls: Optional[List[str]] = None,
                 build_dir: Optional[str] = None):
        spec = [(str(k), int(n)) for k, n in parameters]
        if not spec:
            spec = [("_no_parameters", 1)]
        self._fn = state_dot
        self._spec = spec
        self.name = name
        self.param_keys = tuple(k for k, _ in spec if k != "_no_parameters")
        super().__init__({
            "num_states": num_states,
            "num_controls": num_controls,
            "names_states": names_states or [f"x{i}" for i in range(num_states)],
            "names_controls": names_controls or [f"u{i}" for i in range(num_controls)],
        })
        self.source = cpp_model_source(state_dot, num_states, num_controls, self.num_params, name)
        self._build_dir = build_dir
        self._lib_path = None

    @property
    def num_params(self) -> int:
        return sum(n for _, n in self._spec)

    def parameter_offsets(self) -> Dict[str, Tuple[int, int]]:
        out, o = {}, 0
        for k, n in self._spec:
            out[k] = (o, n)
            o += n
        return out

    def _flat(self, params: Dict[str, Any], batch: int, dtype, device) -> torch.Tensor:
        cols = []
        for k, n in self._spec:
            if k == "_no_parameters":
                cols.append(torch.zeros(batch, 1, dtype=dtype, device=device))
                continue
            v = torch.as_tensor(params[k], dtype=dtype, device=device)
            if v.numel() == n:
                v = v.reshape(1, n).expand(batch, n)
            else:
                v = v.reshape(batch, n)
            cols.append(v)
        return torch.cat(cols, dim=1)

    def flat_params(self, params: Dict[str, Any], batch: int, device=None) -> torch.Tensor:
        return self._flat(params, batch, torch.float32, device).contiguous()

    def state_dot(self, state, control, params: Dict[str, Any]):
        """x_dot = f(x, u) on torch tensors, broadcast over leading axes.
        Parameters are shared (size n) or per instance ((B, n) with B the
        first leading axis of `state`)."""
        lead = tuple(state.shape[:-1])
        cols = []
        for k, n in self._spec:
            if k == "_no_parameters":
                cols.append(torch.zeros(lead + (1,), dtype=state.dtype, device=state.device))
                continue
            v = torch.as_tensor(params[k], dtype=state.dtype, device=state.device)
            if v.numel() == n:
                v = v.reshape((1,) * len(lead) + (n,))
            else:
                v = v.reshape((v.shape[0],) + (1,) * (len(lead) - 1) + (n,))
            cols.append(v.expand(lead + (n,)))
        p = torch.cat(cols, -1)
        out = self._fn(LastAxis(state), LastAxis(control), LastAxis(p), TorchMath)
        return torch.stack([torch.as_tensor(o, dtype=state.dtype, device=state.device).expand(lead) for o in out], -1)


class LastAxis:
    """Indexes the last axis: v[i] is a[..., i] (how a model function sees
    batched states, controls and parameters)."""

    def __init__(self, a):
        self.a = a

    def __getitem__(self, i):
        return self.a[..., i]
