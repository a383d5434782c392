"""Rate-only spacecraft dynamics (diffmpc/dynamics/spacecraft_dynamics.py:9-66).

State (omega_x, omega_y, omega_z), control torques; w' = J^-1 (tau - w x J w)
with diagonal ``inertia_vector``.
"""
from typing import Any, Dict

import torch

from .._lib import MODEL_SPACECRAFT3
from .base_dynamics import Dynamics, batched

spacecraft_parameters: Dict[str, Any] = {
    "num_states": 3,
    "num_controls": 3,
    "names_states": ["omega_x", "omega_y", "omega_z"],
    "names_controls": ["torque_x", "torque_y", "torque_z"],
}
spacecraft_state_dot_parameters = {"inertia_vector": torch.tensor([5.0, 2.0, 1.0])}


class SpacecraftDynamics(Dynamics):
    """Spacecraft dynamics class."""

    model_id = MODEL_SPACECRAFT3
    num_params = 3
    param_keys = ("inertia_vector",)

    def __init__(self, parameters: Dict[str, Any] = spacecraft_parameters):
        super().__init__(parameters)

    def state_dot(self, state, control, state_dot_params: Dict[str, Any] = spacecraft_state_dot_parameters):
        J = torch.as_tensor(state_dot_params["inertia_vector"], dtype=state.dtype, device=state.device)
        return (control - torch.cross(state, J * state, dim=-1)) / J

    def flat_params(self, params, batch, device=None):
        return batched(params["inertia_vector"], batch, 1, torch.float32, device).contiguous()
