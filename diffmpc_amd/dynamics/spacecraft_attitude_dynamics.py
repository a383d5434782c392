"""Attitude spacecraft: 7 states [q (scalar first), omega], 3 torques.

No reference counterpart (the reference spacecraft is rate-only).  It joins
the normalised quaternion kinematics of diffmpc/dynamics/quadrotor_dynamics.py
(:122, :129) with Euler's equation of spacecraft_dynamics.py:63 for a
diagonal ``inertia_vector``.
"""
from typing import Any, Dict

import torch

from .._lib import MODEL_SPACECRAFT7
from .base_dynamics import Dynamics, batched
from .quadrotor_dynamics import q_left

spacecraft_attitude_parameters: Dict[str, Any] = {
    "num_states": 7,
    "num_controls": 3,
    "names_states": ["q_0", "q_1", "q_2", "q_3", "omega_x", "omega_y", "omega_z"],
    "names_controls": ["torque_x", "torque_y", "torque_z"],
}
spacecraft_attitude_state_dot_parameters = {"inertia_vector": torch.tensor([1.0, 1.0, 1.0])}


class SpacecraftAttitudeDynamics(Dynamics):
    model_id = MODEL_SPACECRAFT7
    num_params = 3
    param_keys = ("inertia_vector",)

    def __init__(self, parameters: Dict[str, Any] = spacecraft_attitude_parameters):
        super().__init__(parameters)

    def state_dot(self, state, control, state_dot_params: Dict[str, Any] = spacecraft_attitude_state_dot_parameters):
        J = torch.as_tensor(state_dot_params["inertia_vector"], dtype=state.dtype, device=state.device)
        q = state[..., :4]
        w = state[..., 4:7]
        q = q / torch.linalg.norm(q, dim=-1, keepdim=True)
        w4 = torch.cat([torch.zeros_like(w[..., :1]), w], -1)
        q_dot = 0.5 * (q_left(q) @ w4.unsqueeze(-1)).squeeze(-1)
        w_dot = (control - torch.cross(w, J * w, dim=-1)) / J
        return torch.cat([q_dot, w_dot], -1)

    def flat_params(self, params, batch, device=None):
        return batched(params["inertia_vector"], batch, 1, torch.float32, device).contiguous()
