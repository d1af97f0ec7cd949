"""Drop-in compute_elemental_strain_rate and compute_nodal_strain_rate
(compute_strain_rate.py:17-73) on the GPU."""
import numpy as np

from . import _dev as dev
from . import _lib


def _nodal_gradients(mode, icon_V, u, v, nn_V, nel, dNdr_V_n, dNds_V_n, jcbi00n, jcbi01n, jcbi10n,
                     jcbi11n):
    numpy_out = not dev.is_dev(u)
    ic = dev.to_dev(icon_V, np.int32)
    nel = int(nel)
    inc = dev.incidence(ic, nn_V, dev.array_key(icon_V))
    # device copies are kept referenced until the kernels have been enqueued
    ins = [dev.to_dev(a, np.float64) for a in (u, v if mode == 0 else u, dNdr_V_n, dNds_V_n,
                                               jcbi00n, jcbi01n, jcbi10n, jcbi11n)]
    outs = [dev.empty(nn_V) for _ in range(3 if mode == 0 else 2)]
    work = dev.empty(27 * nel)
    _lib.call("mw_nodal_gradients", mode, nel, int(nn_V), dev.ptr(ic), *[dev.ptr(a) for a in ins],
              dev.ptr(inc.ptr), dev.ptr(inc.list), dev.ptr(outs[0]), dev.ptr(outs[1]),
              dev.ptr(outs[2]) if mode == 0 else None, dev.ptr(work), dev.stream())
    return tuple(dev.out(o, numpy_out) for o in outs)


def compute_nodal_strain_rate(icon_V, u, v, nn_V, m_V, nel, dNdr_V_n, dNds_V_n, jcbi00n, jcbi01n,
                              jcbi10n, jcbi11n):
    """compute_strain_rate.py:50-73 -> (exx_n, eyy_n, exy_n)."""
    return _nodal_gradients(0, icon_V, u, v, nn_V, nel, dNdr_V_n, dNds_V_n, jcbi00n, jcbi01n,
                            jcbi10n, jcbi11n)


def compute_elemental_strain_rate(icon_V, u, v, nn_V, nel, x_V, y_V):
    """compute_strain_rate.py:17-42 -> (exx, eyy, exy) at the element centres."""
    numpy_out = not dev.is_dev(u, v)
    ic = dev.to_dev(icon_V, np.int32)
    ins = [dev.to_dev(a, np.float64) for a in (u, v, x_V, y_V)]
    outs = [dev.empty(int(nel)) for _ in range(3)]
    _lib.call("mw_elemental_strain_rate", int(nel), dev.ptr(ic), *[dev.ptr(a) for a in ins],
              *[dev.ptr(o) for o in outs], dev.stream())
    return tuple(dev.out(o, numpy_out) for o in outs)
