from .optimal_control_problem import OptimalControlProblem  # noqa: F401
