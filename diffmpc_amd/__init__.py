"""diffmpc_amd: batched differentiable MPC (SQP + Schur-complement PCG) on MI355X.

The solver API follows ToyotaResearchInstitute/diffmpc (problems, dynamics,
SQPSolver, SQPSolution); the solve and its adjoint run as HIP kernels in
libdmpc.so behind the C ABI of include/dmpc.h.
"""
from . import _lib  # noqa: F401
from .dynamics import (CartpoleDynamics, ChainDynamics, CustomDynamics, Dynamics, LinearDynamics,  # noqa: F401
                       QuadrotorDynamics, SpacecraftAttitudeDynamics, SpacecraftDynamics)
from .problems import OptimalControlProblem  # noqa: F401
from .solvers import PCGOptimalControl, SolverReturnStatus, SQPSolution, SQPSolver  # noqa: F401
from .utils.load_params import load_problem_params, load_solver_params  # noqa: F401

__version__ = "0.1.0"
