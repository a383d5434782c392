"""Example user-defined models (``CustomDynamics``).

``unicycle``: planar unicycle, state (p_x, p_y, heading), controls (speed,
turn rate), parameter ``turn_gain`` scaling the turn rate.

``rate_spacecraft``: the rate-only spacecraft of
diffmpc/dynamics/spacecraft_dynamics.py:31-66 written as a user model, so a
custom build can be checked against the built-in kernels.
"""
import os

from .custom_dynamics import CustomDynamics

# prebuilt in-tree location of the examples' custom libraries
EXAMPLES_BUILD_ROOT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "_custom")


def unicycle_state_dot(x, u, p, m):
    return [u[0] * m.cos(x[2]), u[0] * m.sin(x[2]), p[0] * u[1]]


def rate_spacecraft_state_dot(w, u, p, m):
    hx, hy, hz = p[0] * w[0], p[1] * w[1], p[2] * w[2]
    return [(u[0] - (w[1] * hz - w[2] * hy)) * (1.0 / p[0]),
            (u[1] - (w[2] * hx - w[0] * hz)) * (1.0 / p[1]),
            (u[2] - (w[0] * hy - w[1] * hx)) * (1.0 / p[2])]


def unicycle(build_dir=None) -> CustomDynamics:
    return CustomDynamics(3, 2, unicycle_state_dot, [("turn_gain", 1)], name="unicycle",
                          names_states=["p_x", "p_y", "heading"], names_controls=["speed", "turn_rate"],
                          build_dir=build_dir or os.path.join(EXAMPLES_BUILD_ROOT, "unicycle"))


def rate_spacecraft(build_dir=None) -> CustomDynamics:
    return CustomDynamics(3, 3, rate_spacecraft_state_dot, [("inertia_vector", 3)], name="rate_spacecraft",
                          names_states=["w_x", "w_y", "w_z"], names_controls=["tau_x", "tau_y", "tau_z"],
                          build_dir=build_dir or os.path.join(EXAMPLES_BUILD_ROOT, "rate_spacecraft"))


EXAMPLES = (unicycle, rate_spacecraft)
