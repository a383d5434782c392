"""Base dynamics class (diffmpc/dynamics/base_dynamics.py:16-131).

Besides the reference's interface (sizes, state/control names, ``state_dot``)
a model here knows the id and the flat float32 parameter layout the device
kernels use (``include/dmpc.h``), and how to build that layout, batched, from
a problem-parameter dict.
"""
from typing import Any, Dict, List

import torch

default_parameters: Dict[str, Any] = {
    "num_states": 2,
    "num_controls": 2,
    "names_states": ["state_1", "state_2"],
    "names_controls": ["control_1", "control_2"],
}


def batched(value, batch: int, base_ndim: int, dtype=torch.float64, device=None) -> torch.Tensor:
    """(batch, *shape) tensor from a shared (base_ndim) or batched (base_ndim + 1) value."""
    t = torch.as_tensor(value, dtype=dtype, device=device)
    if t.dim() == base_ndim:
        t = t.unsqueeze(0).expand((batch,) + tuple(t.shape))
    elif t.dim() != base_ndim + 1 or t.shape[0] != batch:
        raise ValueError(f"expected {base_ndim}-d (shared) or {base_ndim + 1}-d batched value with batch {batch}, got {tuple(t.shape)}")
    return t


class Dynamics:
    """Base dynamics class."""

    model_id: int = -1
    num_params: int = 0
    # problem-parameter keys that flat_params reads: the dynamics parameters
    # a solve can differentiate through when they are passed as weights
    param_keys: tuple = ()

    def __init__(self, parameters: Dict[str, Any] = default_parameters):
        self._params = parameters
        self._state_name_to_state_index_dict = {n: i for i, n in enumerate(self.names_states)}
        self._control_name_to_control_index_dict = {n: i for i, n in enumerate(self.names_controls)}

    @property
    def params(self) -> Dict[str, Any]:
        return self._params

    @property
    def num_states(self) -> int:
        return self.params["num_states"]

    @property
    def num_controls(self) -> int:
        return self.params["num_controls"]

    @property
    def names_states(self) -> List[str]:
        return self.params["names_states"]

    @property
    def names_controls(self) -> List[str]:
        return self.params["names_controls"]

    @property
    def state_name_to_state_index_dict(self) -> Dict[str, Any]:
        return self._state_name_to_state_index_dict

    @property
    def control_name_to_control_index_dict(self) -> Dict[str, Any]:
        return self._control_name_to_control_index_dict

    def get_state_variable_at_state_name(self, state, state_name: str):
        return state[..., self.state_name_to_state_index_dict[state_name]]

    def get_control_variable_at_control_name(self, control, control_name: str):
        return control[..., self.control_name_to_control_index_dict[control_name]]

    def state_dot(self, state, control, params: Dict[str, Any] = None):
        """x_dot = f(x, u); torch, broadcast over leading axes."""
        raise NotImplementedError

    def flat_params(self, params: Dict[str, Any], batch: int, device=None) -> torch.Tensor:
        """(batch, num_params) float32 device parameter block (include/dmpc.h layouts)."""
        raise NotImplementedError
