"""Cart-pole dynamics (diffmpc/dynamics/cartpole_dynamics.py:25-80)."""
from typing import Any, Dict

import torch

from .._lib import MODEL_CARTPOLE
from .base_dynamics import Dynamics, batched

default_parameters: Dict[str, Any] = {
    "num_states": 4,
    "num_controls": 1,
    "names_states": ["x", "x_dot", "theta", "theta_dot"],
    "names_controls": ["force"],
}
default_state_dot_parameters: Dict[str, Any] = {"masscart": 1.0, "masspole": 0.1, "length": 0.5, "gravity": 9.81}


class CartpoleDynamics(Dynamics):
    model_id = MODEL_CARTPOLE
    num_params = 4
    param_keys = ("masscart", "masspole", "length", "gravity")

    def __init__(self, parameters: Dict[str, Any] = default_parameters):
        super().__init__(parameters)

    def state_dot(self, state, control, params: Dict[str, Any] = default_state_dot_parameters):
        x_dot, th, th_dot = state[..., 1], state[..., 2], state[..., 3]
        force = control[..., 0]
        mc = params.get("masscart", 1.0)
        mp = params.get("masspole", 0.1)
        length = params.get("length", 0.5)
        g = params.get("gravity", 9.81)
        total = mc + mp
        pml = mp * length
        s, c = torch.sin(th), torch.cos(th)
        temp = (force + pml * th_dot ** 2 * s) / total
        th_acc = (g * s - c * temp) / (length * (4.0 / 3.0 - (mp * c ** 2) / total))
        x_acc = temp - pml * th_acc * c / total
        return torch.stack([x_dot, x_acc, th_dot, th_acc], -1)

    def flat_params(self, params, batch, device=None):
        cols = []
        for key, dflt in (("masscart", 1.0), ("masspole", 0.1), ("length", 0.5), ("gravity", 9.81)):
            v = torch.as_tensor(params.get(key, dflt), dtype=torch.float32).reshape(-1)
            cols.append(batched(v[0] if v.numel() == 1 else v, batch, 0, torch.float32, device).reshape(batch, 1))
        return torch.cat(cols, dim=1).contiguous()
