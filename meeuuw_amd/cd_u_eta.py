"""Drop-in cd_u_eta (cd_u_eta.py:13-99) on the GPU."""
from .uzawa3_solver_L2 import cd_u_eta  # noqa: F401
