"""Drop-in FP64 diagnostics: compute_global_quantities (postprocessors.py:13-83)
and compute_Nu (postprocessors.py:152-248)."""
import numpy as np

from . import _dev as dev
from . import _lib


def compute_global_quantities(nel, nq_per_element, xq, zq, uq, wq, Tq, rhoq, hcapaq, alphaq, etaq,
                              exxq, ezzq, exzq, dxxq, dzzq, dxzq, volume, JxWq, gxq, gzq):
    """-> (vrms, EK, WAG, TVD, GPE, ITE, TM, Tavrg, eta_avrg)."""
    nel = int(nel)
    f = lambda a: dev.to_dev(np.broadcast_to(a, (nel, 9)) if np.isscalar(a) else a, np.float64)
    # device copies are kept referenced until the kernels have been enqueued
    args = [f(a) for a in (uq, wq, Tq, rhoq, hcapaq, etaq, exxq, ezzq, exzq)]
    jw, gx, gz = f(JxWq), f(gxq), f(gzq)
    out, work = dev.empty(9), dev.empty(1024 * 9 + 9)
    _lib.call("mw_global_quantities", nel, *[dev.ptr(a) for a in args], float(volume),
              dev.ptr(jw), dev.ptr(gx), dev.ptr(gz), dev.ptr(out), dev.ptr(work), dev.stream())
    return tuple(float(v) for v in dev.host(out))


def compute_Nu(Lx, Lz, x_V, z_V, nel, left_element, right_element, top_element, bottom_element,
               icon_V, T, dTdx_nodal, dTdz_nodal, nq_per_dim, qcoords, qweights, hcond_n):
    """-> (avrg_T_left, avrg_T_right, avrg_T_bottom, avrg_T_top, avrg_dTdx_left,
    avrg_dTdx_right, avrg_dTdz_bottom, avrg_dTdz_top, qx_left, qx_right, qz_bottom, qz_top, Nu)."""
    if nq_per_dim != 3:
        raise ValueError("compute_Nu: only the 3-point rule is supported")
    f = lambda a: dev.to_dev(a, np.float64)
    b = lambda a: dev.to_dev(a, np.uint8)
    iv = dev.to_dev(icon_V, np.int32)
    xv, zv = f(x_V), f(z_V)
    flags = [b(a) for a in (left_element, right_element, top_element, bottom_element)]
    fields = [f(a) for a in (T, dTdx_nodal, dTdz_nodal, qcoords, qweights, hcond_n)]
    out, work = dev.empty(13), dev.empty(1024 * 12 + 12)
    _lib.call("mw_compute_Nu", float(Lx), float(Lz), dev.ptr(xv), dev.ptr(zv), int(nel),
              *[dev.ptr(a) for a in flags], dev.ptr(iv), *[dev.ptr(a) for a in fields],
              dev.ptr(out), dev.ptr(work), dev.stream())
    return tuple(float(v) for v in dev.host(out))
