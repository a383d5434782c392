"""Parameter files (diffmpc/utils/load_params.py:26-101): same keys, same rules.

Problem parameters become float64 tensors (except the integer/boolean keys)
and the reference state/control vectors are repeated horizon+1 times.
"""
import os
from typing import Any, Dict

import torch
import yaml

_ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
_NOT_ARRAYS = ["horizon", "discretization_scheme", "penalize_control_reference", "rescale_optimization_variables"]


def load_params(params_yaml_file_path: str) -> Dict[str, Any]:
    with open(params_yaml_file_path, "r") as f:
        return yaml.safe_load(f)


def load_solver_params(params_yaml_filename: str) -> Dict[str, Any]:
    return load_params(os.path.join(_ROOT, "solvers", "params", params_yaml_filename))


def load_problem_params(params_yaml_filename: str) -> Dict[str, Any]:
    params = load_params(os.path.join(_ROOT, "problems", "params", params_yaml_filename))
    for key in params:
        if key not in _NOT_ARRAYS:
            params[key] = torch.as_tensor(params[key], dtype=torch.float64)
    n = int(params["horizon"]) + 1
    params["reference_state_trajectory"] = params["reference_state_trajectory"].unsqueeze(0).repeat(n, 1)
    params["reference_control_trajectory"] = params["reference_control_trajectory"].unsqueeze(0).repeat(n, 1)
    return params


def check_parameters_dictionary_or_raise_errors(params: Dict[str, Any]) -> bool:
    if params is None:
        raise ValueError("[OptimalControlProblem] params should be a dictionary.")
    for key in load_params(os.path.join(_ROOT, "problems", "params", "required_parameters.yaml")):
        if key not in params:
            raise KeyError("[OptimalControlProblem]", key, "is not in params dictionary")
    return True
