"""Quadrotor dynamics, scalar-first right-handed quaternion.

Follows diffmpc/dynamics/quadrotor_dynamics.py:8-133: state
[p (3), v (3), q (4), w (3)], controls [thrust, tau_x, tau_y, tau_z];
parameters ``mass`` and ``inertia`` (3x3, or its 9 entries row-major).
"""
from typing import Any, Dict

import torch

from .._lib import MODEL_QUADROTOR
from .base_dynamics import Dynamics, batched

GRAVITY = 9.81

quadrotor_parameters: Dict[str, Any] = {
    "num_states": 13,
    "num_controls": 4,
    "names_states": ["pos_x", "pos_y", "pos_z", "vel_x", "vel_y", "vel_z", "q_0", "q_1", "q_2", "q_3",
                     "omega_x", "omega_y", "omega_z"],
    "names_controls": ["thrust", "torque_x", "torque_y", "torque_z"],
}
quadrotor_state_dot_parameters = {
    "mass": 0.1,
    "inertia": torch.tensor([0.1, 0.01, 0.01, 0.01, 0.1, 0.01, 0.01, 0.01, 0.1]),
}


def S(u):
    """Skew-symmetric matrix with S(a) b = a x b."""
    z = torch.zeros_like(u[..., 0])
    return torch.stack([torch.stack([z, -u[..., 2], u[..., 1]], -1),
                        torch.stack([u[..., 2], z, -u[..., 0]], -1),
                        torch.stack([-u[..., 1], u[..., 0], z], -1)], -2)


def q_left(q):
    """Left quaternion-product matrix: q (x) p = q_left(q) p."""
    qw, qx, qy, qz = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    return torch.stack([torch.stack([qw, -qx, -qy, -qz], -1),
                        torch.stack([qx, qw, -qz, qy], -1),
                        torch.stack([qy, qz, qw, -qx], -1),
                        torch.stack([qz, -qy, qx, qw], -1)], -2)


def q_conj(q):
    """Quaternion conjugate."""
    return torch.cat([q[..., :1], -q[..., 1:]], -1)


def get_rotation(q):
    """Rotation matrix of a unit quaternion."""
    qw, qx, qy, qz = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    qw2, qx2, qy2, qz2 = qw * qw, qx * qx, qy * qy, qz * qz
    return torch.stack([
        torch.stack([qw2 + qx2 - qy2 - qz2, 2 * (qx * qy - qw * qz), 2 * (qx * qz + qw * qy)], -1),
        torch.stack([2 * (qx * qy + qw * qz), qw2 - qx2 + qy2 - qz2, 2 * (qy * qz - qw * qx)], -1),
        torch.stack([2 * (qx * qz - qw * qy), 2 * (qy * qz + qw * qx), qw2 - qx2 - qy2 + qz2], -1)], -2)


class QuadrotorDynamics(Dynamics):
    """Quadrotor dynamics class."""

    model_id = MODEL_QUADROTOR
    num_params = 10
    param_keys = ("mass", "inertia")

    def __init__(self, parameters: Dict[str, Any] = quadrotor_parameters):
        super().__init__(parameters)

    def state_dot(self, state, control, state_dot_params: Dict[str, Any] = quadrotor_state_dot_parameters):
        J = torch.as_tensor(state_dot_params["inertia"], dtype=state.dtype, device=state.device)
        J = J.reshape(3, 3) if J.numel() == 9 else J.reshape(J.shape[0], *([1] * (state.dim() - 2)), 3, 3)
        m = torch.as_tensor(state_dot_params["mass"], dtype=state.dtype, device=state.device).reshape(-1)
        m = m[0] if m.numel() == 1 else m.reshape(m.shape[0], *([1] * (state.dim() - 2)))
        Jinv = torch.linalg.inv(J)
        v = state[..., 3:6]
        q = state[..., 6:10]
        w = state[..., 10:13]
        f = control[..., 0]
        tau = control[..., 1:4]
        q = q / torch.linalg.norm(q, dim=-1, keepdim=True)
        R = get_rotation(q)
        e3 = torch.tensor([0.0, 0.0, 1.0], dtype=state.dtype, device=state.device)
        v_dot = -GRAVITY * e3 + (f / m).unsqueeze(-1) * R[..., :, 2]
        w4 = torch.cat([torch.zeros_like(w[..., :1]), w], -1)
        q_dot = 0.5 * (q_left(q) @ w4.unsqueeze(-1)).squeeze(-1)
        Jw = (J @ w.unsqueeze(-1)).squeeze(-1)
        w_dot = (Jinv @ (torch.cross(Jw, w, dim=-1) + tau).unsqueeze(-1)).squeeze(-1)
        return torch.cat([v, v_dot, q_dot, w_dot], -1)

    def flat_params(self, params, batch, device=None):
        m = torch.as_tensor(params["mass"], dtype=torch.float32).reshape(-1)
        if m.numel() == 1:
            m = m.expand(batch)
        elif m.numel() != batch:
            raise ValueError("mass must be a scalar or have one entry per instance")
        J = torch.as_tensor(params["inertia"], dtype=torch.float32)
        if J.numel() % 9:
            raise ValueError("inertia must hold 3x3 entries per instance")
        J = J.reshape(-1, 9)
        if J.shape[0] == 1:
            J = J.expand(batch, 9)
        elif J.shape[0] != batch:
            raise ValueError("inertia must be shared or have one 3x3 block per instance")
        return torch.cat([m.reshape(batch, 1), J], dim=1).to(device).contiguous()
