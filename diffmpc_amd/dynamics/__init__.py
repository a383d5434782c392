from .base_dynamics import Dynamics  # noqa: F401
from .cartpole_dynamics import CartpoleDynamics  # noqa: F401
from .chain_dynamics import ChainDynamics  # noqa: F401
from .custom_dynamics import CustomDynamics  # noqa: F401
from .integrators import DiscretizationScheme, predict_next_state  # noqa: F401
from .linear_dynamics import LinearDynamics  # noqa: F401
from .quadrotor_dynamics import QuadrotorDynamics  # noqa: F401
from .spacecraft_attitude_dynamics import SpacecraftAttitudeDynamics  # noqa: F401
from .spacecraft_dynamics import SpacecraftDynamics  # noqa: F401
