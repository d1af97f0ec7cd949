"""Drop-in compute_nodal_heat_flux (compute_nodal_heat_flux.py:10-40)."""
import numpy as np

from . import _dev as dev
from . import _lib


def compute_nodal_heat_flux(icon_T, T, hcond_nodal, nn_T, m_T, nel, dNdr_T_n, dNds_T_n, jcbi00_T,
                            jcbi01_T, jcbi10_T, jcbi11_T):
    """-> (dTdx_n, dTdy_n, qx_n, qy_n)."""
    numpy_out = not dev.is_dev(T)
    it = dev.to_dev(icon_T, np.int32)
    nel = int(nel)
    inc = dev.incidence(it, nn_T, dev.array_key(icon_T))
    # device copies are kept referenced until the kernels have been enqueued
    ins = [dev.to_dev(a, np.float64) for a in (T, hcond_nodal, dNdr_T_n, dNds_T_n, jcbi00_T,
                                               jcbi01_T, jcbi10_T, jcbi11_T)]
    outs = [dev.empty(nn_T) for _ in range(4)]
    work = dev.empty(18 * nel)
    _lib.call("mw_nodal_heat_flux", nel, int(nn_T), dev.ptr(it), *[dev.ptr(a) for a in ins],
              dev.ptr(inc.ptr), dev.ptr(inc.list), *[dev.ptr(o) for o in outs], dev.ptr(work),
              dev.stream())
    return tuple(dev.out(o, numpy_out) for o in outs)
