"""Drop-in build_velocity_connectivity (build_velocity_connectivity.py:14-136)
on the GPU, for the Q2 velocity space (m_V = 9)."""
import numpy as np

from . import _dev as dev
from . import _lib
from .build_velocity_mesh import GEOMETRY


def build_velocity_connectivity(geometry, m_V, nelx, nelz, nnx, nnz, middleH_nodes,
                                middleV_nodes):
    if m_V == 5:
        raise NotImplementedError("build_velocity_connectivity: the Q1+ space (m_V=5) is not on "
                                  "the GPU path")
    if m_V != 9:
        raise ValueError("build_velocity_mesh: unknown m_V value")
    if geometry not in GEOMETRY:
        raise ValueError("build_velocity_connectivity: unknown geometry")
    nelx, nelz = int(nelx), int(nelz)
    nel = nelx * nelz
    numpy_out = not dev.is_dev(middleH_nodes, middleV_nodes)
    dev.require_gpu()
    annulus = geometry == "annulus"
    mh = None if annulus else dev.to_dev(middleH_nodes, np.uint8)
    mv = None if annulus else dev.to_dev(middleV_nodes, np.uint8)
    icon = dev.empty(9 * nel, np.int32)
    ef = dev.empty(6 * nel, np.uint8)
    _lib.call("mw_build_velocity_connectivity", GEOMETRY[geometry], nelx, nelz, dev.ptr(mh),
              dev.ptr(mv), dev.ptr(icon), dev.ptr(ef), dev.stream())
    f = ef.reshape(6, nel).to(dtype=dev.torch.bool)
    return (dev.out(icon, numpy_out, shape=(9, nel)),
            *(dev.out(f[k], numpy_out) for k in range(6)))  # top, bot, left, right, mH, mV
