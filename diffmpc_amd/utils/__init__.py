from .load_params import check_parameters_dictionary_or_raise_errors, load_params, load_problem_params, load_solver_params
from .torch_utils import (project_matrix_onto_positive_semidefinite_cone, torch_has_gpu, value_and_jacfwd,
                          value_and_jacrev)
