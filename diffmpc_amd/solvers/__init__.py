from .pcg import (PCGDebugOutput, PCGOptimalControl, QPCostParameters, QPDynamicsCostPremultipliedMatrices,  # noqa: F401
                  QPDynamicsParameters, QPParameters, QPSolution, SchurComplementMatrices)
from .sqp import SavedLinearization, SolverReturnStatus, SQPSolution, SQPSolver  # noqa: F401
