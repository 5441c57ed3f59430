"""MI355X-native IMEX-DG hot path of Atmospheric_Flows (libatmo_hip.so + host mirror)."""
from ._native import AtmoError, NoConvergence, load  # noqa: F401
from .operator import (BLOCK_DENSITY, BLOCK_PRESSURE, BLOCK_VELOCITY, DeviceContext, DeviceVector,  # noqa: F401
                       EULEROperator, ShellContext, default_params, shell_device_bytes, witch_of_agnesi)
from .ensemble import Ensemble, Member, device_bytes, latency_summary  # noqa: F401,E402
