"""Linear dynamics x' = A x + B u + b (diffmpc/dynamics/linear_dynamics.py:29-70)."""
from typing import Any, Dict, Optional

import torch

from .._lib import MODEL_LINEAR
from .base_dynamics import Dynamics, batched
from .custom_dynamics import MODEL_CUSTOM, CustomBuild

# (nx, nu) the library is built for; other sizes run in a custom build of it
BUILT_SIZES = ((4, 2), (8, 4))

default_parameters: Dict[str, Any] = {
    "num_states": 4,
    "num_controls": 2,
    "names_states": ["pos_x", "pos_y", "vel_x", "vel_y"],
    "names_controls": ["force_x", "force_x"],
}
default_state_dot_parameters: Dict[str, Any] = {
    "A": torch.tensor([[0.0, 0, 1, 0], [0, 0, 0, 1], [0, 0, 0, 0], [0, 0, 0, 0]]),
    "B": torch.tensor([[0.0, 0], [0, 0], [1, 0], [0, 1]]),
    "b": torch.tensor([0.0, 0, 0, 0]),
}


class LinearDynamics(CustomBuild, Dynamics):
    """Linear dynamics class: state_dot = A @ state + B @ control + b.

    Any (nx, nu), as in the reference: the sizes in BUILT_SIZES run in the
    library itself; any other size (nx <= 64) is the same model template
    compiled into a custom build of the library (model id MODEL_CUSTOM,
    compiled on first use and cached, or prebuilt in the package).  Sizes
    with more controls than states (nu > nx) run in the plain (per-launch)
    mode only: the fused kernels recover control rows on the lanes of a
    state block, so mode "auto" picks plain for them and mode "fused"
    raises DmpcError (DMPC_ERR_UNSUPPORTED)."""

    model_id = MODEL_LINEAR
    param_keys = ("A", "B", "b")

    def __init__(self, parameters: Dict[str, Any] = default_parameters):
        super().__init__(parameters)
        nx, nu = self.num_states, self.num_controls
        if (nx, nu) not in BUILT_SIZES:
            self.model_id = MODEL_CUSTOM
            self.name = f"linear{nx}x{nu}"
            self.source = (f"// the linear model with nx = {nx}, nu = {nu}\n"
                           f"using CustomModel = LinearModel<{nx}, {nu}>;\n")

    def library_path(self, build: bool = True) -> Optional[str]:
        """None for the built sizes (the library itself), else the custom build."""
        if self.model_id == MODEL_LINEAR:
            return None
        return super().library_path(build)

    @property
    def num_params(self) -> int:
        nx, nu = self.num_states, self.num_controls
        return nx * nx + nx * nu + nx

    def state_dot(self, state, control, params: Dict[str, Any] = default_state_dot_parameters):
        A = torch.as_tensor(params["A"], dtype=state.dtype, device=state.device)
        B = torch.as_tensor(params["B"], dtype=state.dtype, device=state.device)
        b = torch.as_tensor(params.get("b", torch.zeros(self.num_states)), dtype=state.dtype, device=state.device)
        return (A @ state.unsqueeze(-1)).squeeze(-1) + (B @ control.unsqueeze(-1)).squeeze(-1) + b

    def flat_params(self, params, batch, device=None):
        nx, nu = self.num_states, self.num_controls
        A = batched(params["A"], batch, 2, torch.float32, device).reshape(batch, nx * nx)
        B = batched(params["B"], batch, 2, torch.float32, device).reshape(batch, nx * nu)
        b = batched(params.get("b", torch.zeros(nx)), batch, 1, torch.float32, device)
        return torch.cat([A, B, b], dim=1).contiguous()
