"""Drop-in compute_nodal_pressure (compute_nodal_pressure.py:7-21)."""
import numpy as np

from . import _dev as dev
from . import _lib


def compute_nodal_pressure(m_V, nn_V, icon_V, icon_P, p, N_P_n):
    numpy_out = not dev.is_dev(p)
    iv = dev.to_dev(icon_V, np.int32)
    ip = dev.to_dev(icon_P, np.int32)
    nel = iv.shape[1]
    inc = dev.incidence(iv, nn_V, dev.array_key(icon_V))
    pd, nd = dev.to_dev(p, np.float64), dev.to_dev(N_P_n, np.float64)
    q, work = dev.empty(nn_V), dev.empty(9 * nel)
    _lib.call("mw_nodal_pressure", nel, int(nn_V), dev.ptr(ip), dev.ptr(pd), dev.ptr(nd),
              dev.ptr(inc.ptr), dev.ptr(inc.list), dev.ptr(q), dev.ptr(work), dev.stream())
    return dev.out(q, numpy_out)
