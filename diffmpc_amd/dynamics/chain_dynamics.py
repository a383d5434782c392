"""Planar chain of 12 unit masses with springs and dampers (48 states, 12 forces).

No reference counterpart.  Mass 1 hangs from a wall at the origin; link j
(spring ``spring_constants[j]``, damper ``damping[j]``, zero rest length)
joins masses j-1 and j; control j pushes mass j along x.  State
[p_1x, p_1y, ..., p_12y, v_1x, ..., v_12y].
"""
from typing import Any, Dict

import torch

from .._lib import MODEL_CHAIN
from .base_dynamics import Dynamics, batched

M = 12
chain_parameters: Dict[str, Any] = {
    "num_states": 4 * M,
    "num_controls": M,
    "names_states": [f"p{j}_{a}" for j in range(M) for a in "xy"] + [f"v{j}_{a}" for j in range(M) for a in "xy"],
    "names_controls": [f"f{j}" for j in range(M)],
}


class ChainDynamics(Dynamics):
    model_id = MODEL_CHAIN
    num_params = 2 * M
    param_keys = ("spring_constants", "damping")

    def __init__(self, parameters: Dict[str, Any] = chain_parameters):
        super().__init__(parameters)

    def state_dot(self, state, control, params: Dict[str, Any]):
        k = torch.as_tensor(params["spring_constants"], dtype=state.dtype, device=state.device).unsqueeze(-1)
        c = torch.as_tensor(params["damping"], dtype=state.dtype, device=state.device).unsqueeze(-1)
        pos = state[..., : 2 * M].reshape(state.shape[:-1] + (M, 2))
        vel = state[..., 2 * M:].reshape(state.shape[:-1] + (M, 2))
        zero = torch.zeros_like(pos[..., :1, :])
        link = -k * (pos - torch.cat([zero, pos[..., :-1, :]], -2)) - c * (vel - torch.cat([zero, vel[..., :-1, :]], -2))
        acc = link - torch.cat([link[..., 1:, :], torch.zeros_like(link[..., :1, :])], -2)
        acc = acc + torch.stack([control, torch.zeros_like(control)], -1)
        return torch.cat([vel.reshape(state.shape[:-1] + (2 * M,)), acc.reshape(state.shape[:-1] + (2 * M,))], -1)

    def flat_params(self, params, batch, device=None):
        k = batched(params["spring_constants"], batch, 1, torch.float32, device)
        c = batched(params["damping"], batch, 1, torch.float32, device)
        return torch.cat([k, c], dim=1).contiguous()
