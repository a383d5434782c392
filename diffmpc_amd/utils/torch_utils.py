"""Host helpers of diffmpc/utils/jax_utils.py:12-95 on torch tensors.

The kernels never call these (they use the closed forms of the diagonal cost
and forward-mode dual numbers); they are here for callers of the reference's
utility functions: Jacobians of user functions, the GPU check, and the PSD
projection of a cost Hessian.
"""
from typing import Callable, Tuple

import torch


def value_and_jacfwd(f: Callable[[torch.Tensor], torch.Tensor], x: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
    """(f(x), df/dx (num_outputs, num_variables)) by forward mode
    (jax_utils.py:12-36)."""
    if not callable(f):
        raise TypeError(f"expected a callable, got {type(f).__name__}")
    jac = torch.func.jacfwd(f)(x)
    return f(x), jac.reshape(-1, x.numel())


def value_and_jacrev(f: Callable[[torch.Tensor], torch.Tensor], x: torch.Tensor) -> Tuple[torch.Tensor, torch.Tensor]:
    """(f(x), df/dx (num_outputs, num_variables)) by reverse mode
    (jax_utils.py:39-63)."""
    if not callable(f):
        raise TypeError(f"expected a callable, got {type(f).__name__}")
    jac = torch.func.jacrev(f)(x)
    return f(x), jac.reshape(-1, x.numel())


def torch_has_gpu() -> bool:
    """Whether torch sees a GPU (the counterpart of jax_has_gpu, jax_utils.py:66-72)."""
    return torch.cuda.is_available()


def project_matrix_onto_positive_semidefinite_cone(matrix: torch.Tensor,
                                                   minimum_eigenvalue: float = 0.0) -> torch.Tensor:
    """Clamp the eigenvalues of a symmetric matrix (or a batch of them) at
    minimum_eigenvalue, then symmetrise (jax_utils.py:75-95)."""
    S, V = torch.linalg.eigh(matrix)
    S = torch.clamp(S, min=minimum_eigenvalue)
    P = (V * S.unsqueeze(-2)) @ V.transpose(-1, -2)
    return 0.5 * (P + P.transpose(-1, -2))
