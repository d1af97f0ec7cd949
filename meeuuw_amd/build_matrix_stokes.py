"""Drop-in build_matrix_stokes (build_matrix_stokes.py:18-278) on the GPU."""
import numpy as np

from . import _dev as dev
from . import _lib

CURVED = ("quarter", "half", "eighth", "annulus")


def build_matrix_stokes(bignb_Stokes, bignb_K, bignb_M, bignb_G, nel, nq_per_element, m_V, m_P,
                        ndof_V, Nfem_V, Nfem_P, ndof_V_el, icon_V, icon_P, rhoq, etaq, JxWq,
                        local_to_globalV, gxq, gyq, N_V, N_P, dNdr_V, dNdt_V, dNdr_P, dNdt_P,
                        jcbi00q, jcbi01q, jcbi10q, jcbi11q, eta_e, eta_ref, L_ref, bc_fix_V,
                        bc_val_V, bot_element, top_element, bot_free_slip, top_free_slip,
                        geometry, theta_V, axisymmetric, xq, blocks):
    """Same arguments and return tuple as the reference:
    (VV_Stokes, rhs_f, rhs_h, VV_K, VV_G, VV_GT, VV_M, VV_M_eta, VV_H).
    The arrays of the mode not selected by ``blocks`` are returned empty."""
    rotate = geometry in CURVED and bool(top_free_slip or bot_free_slip)
    if (nq_per_element, m_V, m_P, ndof_V, ndof_V_el) != (9, 9, 4, 2, 18):
        raise ValueError("build_matrix_stokes: only Q2xQ1 with 3x3 quadrature is supported")
    nel = int(nel)
    if blocks:
        if (bignb_K, bignb_G, bignb_M) != (nel * 324, nel * 72, nel * 16):
            raise ValueError("build_matrix_stokes: block sizes do not match nel")
    elif bignb_Stokes != nel * 468:
        raise ValueError("build_matrix_stokes: bignb_Stokes != nel*468")
    numpy_out = not dev.is_dev(rhoq, etaq, icon_V, local_to_globalV)
    f64, i32 = np.float64, np.int32
    ip = dev.to_dev(icon_P, i32)
    l2g = dev.to_dev(local_to_globalV, i32)
    q = [dev.to_dev(a, f64) for a in (rhoq, etaq, JxWq, gxq, gyq)]
    tab = [dev.to_dev(a, f64) for a in (N_V, N_P, dNdr_V, dNdt_V, dNdr_P, dNdt_P)]
    jc = [dev.to_dev(a, f64) for a in (jcbi00q, jcbi01q, jcbi10q, jcbi11q)]
    ee = dev.to_dev(eta_e, f64) if blocks else None
    fix = dev.to_dev(bc_fix_V, np.uint8)
    val = dev.to_dev(bc_val_V, f64)
    inc_v = dev.incidence(l2g, Nfem_V, dev.array_key(local_to_globalV))
    inc_p = dev.incidence(ip, Nfem_P, dev.array_key(icon_P))
    rhs_f, rhs_h = dev.empty(Nfem_V), dev.empty(Nfem_P)
    work = dev.empty(22 * nel)
    if blocks:
        VV = None
        K, G, GT, H = dev.empty(nel * 324), dev.empty(nel * 72), dev.empty(nel * 72), dev.empty(nel * 72)
        M, Me = dev.empty(nel * 16), dev.empty(nel * 16)
    else:
        VV = dev.empty(nel * 468)
        K = G = GT = H = M = Me = None
    p = dev.ptr
    head = (nel, int(Nfem_V), int(Nfem_P), int(bool(blocks)), p(ip), p(l2g), *[p(a) for a in q],
            *[p(a) for a in tab], *[p(a) for a in jc], p(ee), float(eta_ref), float(L_ref),
            p(fix), p(val))
    tail = (p(inc_v.ptr), p(inc_v.list), p(inc_p.ptr), p(inc_p.list), p(VV), p(rhs_f), p(rhs_h),
            p(K), p(G), p(GT), p(M), p(Me), p(H), p(work), dev.stream())
    if rotate or axisymmetric:
        th = be = te = xqd = None
        if rotate:
            if theta_V is None or bot_element is None or top_element is None:
                raise ValueError("build_matrix_stokes: curved free slip needs theta_V, bot_element, "
                                 "top_element")
            if len(theta_V) * 2 != Nfem_V or len(bot_element) != nel or len(top_element) != nel:
                raise ValueError("build_matrix_stokes: theta_V, bot_element or top_element has the "
                                 "wrong size")
            th = dev.to_dev(theta_V, f64)
            be, te = dev.to_dev(bot_element, np.uint8), dev.to_dev(top_element, np.uint8)
        if axisymmetric:
            if xq is None or int(np.prod(xq.shape)) != 9 * nel:
                raise ValueError("build_matrix_stokes: axisymmetric needs xq of nel x 9 entries")
            xqd = dev.to_dev(xq, f64)
        _lib.call("mw_build_matrix_stokes_general", *head, p(be), p(te), int(bool(bot_free_slip)),
                  int(bool(top_free_slip)), int(rotate), p(th), int(bool(axisymmetric)), p(xqd),
                  *tail)
    else:
        _lib.call("mw_build_matrix_stokes", *head, *tail)
    empty = (lambda: np.zeros(0)) if numpy_out else (lambda: dev.empty(0))
    o = lambda t: empty() if t is None else dev.out(t, numpy_out)
    return (o(VV), o(rhs_f), o(rhs_h), o(K), o(G), o(GT), o(M), o(Me), o(H))
