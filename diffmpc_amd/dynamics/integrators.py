"""Integration schemes (diffmpc/dynamics/integrators.py:12-71), torch host side.

The device kernels integrate with the same three schemes (dmpc_models.h).
"""
import enum
from typing import Dict

import torch

from .base_dynamics import Dynamics


class DiscretizationScheme(enum.IntEnum):
    """Choice of discretization scheme."""

    EULER = 0
    MIDPOINT = 1
    RUNGEKUTTA4 = 2


def predict_next_state(dynamics: Dynamics, dt: float, discretization_scheme: DiscretizationScheme,
                       dynamics_state_dot_params: Dict[str, torch.Tensor], state: torch.Tensor,
                       control: torch.Tensor) -> torch.Tensor:
    """One integration step of x' = f(x, u, params)."""
    f = dynamics.state_dot
    p = dynamics_state_dot_params
    if discretization_scheme == DiscretizationScheme.MIDPOINT:
        xm = state + 0.5 * dt * f(state, control, p)
        return state + dt * f(xm, control, p)
    if discretization_scheme == DiscretizationScheme.RUNGEKUTTA4:
        k1 = f(state, control, p)
        k2 = f(state + 0.5 * dt * k1, control, p)
        k3 = f(state + 0.5 * dt * k2, control, p)
        k4 = f(state + dt * k3, control, p)
        return state + (1.0 / 6.0) * (k1 + 2 * k2 + 2 * k3 + k4) * dt
    return state + dt * f(state, control, p)
