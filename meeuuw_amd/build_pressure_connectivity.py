"""Drop-in build_pressure_connectivity (build_pressure_connectivity.py:10-48)
on the GPU (the Q1 pressure space, m_P = 4)."""
import numpy as np

from . import _dev as dev
from . import _lib
from .build_velocity_mesh import GEOMETRY


def build_pressure_connectivity(geometry, nelx, nelz, nel, m_P, device=False):
    if geometry not in GEOMETRY:
        raise ValueError("build_pressure_mesh: unknown geometry")
    if m_P != 4:
        raise NotImplementedError("build_pressure_connectivity: only the Q1 pressure space (m_P=4)")
    nelx, nelz = int(nelx), int(nelz)
    if int(nel) != nelx * nelz:
        raise ValueError("build_pressure_connectivity: nel != nelx*nelz")
    dev.require_gpu()
    icon = dev.empty(4 * nel, np.int32)
    _lib.call("mw_build_pressure_connectivity", GEOMETRY[geometry], nelx, nelz, dev.ptr(icon),
              dev.stream())
    return dev.out(icon, not device, shape=(4, int(nel)))
