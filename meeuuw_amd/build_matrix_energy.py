"""Drop-in build_matrix_energy (build_matrix_energy.py:11-141) on the GPU."""
import numpy as np

from . import _dev as dev
from . import _lib


def build_matrix_energy(bignb, nel, nq_per_element, m_T, Nfem_T, T, icon_T, rhoq, etaq, Tq, uq, wq,
                        hcondq, hcapaq, alphaq, hprodq, exxq, ezzq, exzq, dpdxq, dpdzq, JxWq, N_T,
                        dNdr_T, dNds_T, jcbi00q, jcbi01q, jcbi10q, jcbi11q, bc_fix_T, bc_val_T, dt,
                        formulation, rho0):
    """Same arguments and return (VV_T, rhs) as the reference.  formulation is
    'BA' or 'EBA' (the latter adds viscous dissipation and adiabatic heating)."""
    if (nq_per_element, m_T) != (9, 9):
        raise ValueError("build_matrix_energy: only Q2 with 3x3 quadrature is supported")
    nel = int(nel)
    if bignb != nel * 81:
        raise ValueError("build_matrix_energy: bignb != nel*81")
    eba = formulation == "EBA"
    numpy_out = not dev.is_dev(T, uq, wq, icon_T)
    f64 = np.float64
    it = dev.to_dev(icon_T, np.int32)
    Td = dev.to_dev(T, f64)
    names = (etaq, Tq, uq, wq, hcondq, hcapaq, alphaq, hprodq, exxq, ezzq, exzq, JxWq)
    q = [dev.to_dev(a, f64) if (a is not None and not np.isscalar(a)) else None for a in names]
    if not eba:  # BA reads neither etaq/Tq/alphaq nor the strain rates
        for k in (0, 1, 6, 8, 9, 10):
            q[k] = None
    tab = [dev.to_dev(a, f64) for a in (N_T, dNdr_T, dNds_T)]
    jc = [dev.to_dev(a, f64) for a in (jcbi00q, jcbi01q, jcbi10q, jcbi11q)]
    fix = dev.to_dev(bc_fix_T, np.uint8)
    val = dev.to_dev(bc_val_T, f64)
    inc = dev.incidence(it, Nfem_T, dev.array_key(icon_T))
    VV, rhs, work = dev.empty(nel * 81), dev.empty(Nfem_T), dev.empty(9 * nel)
    p = dev.ptr
    _lib.call("mw_build_matrix_energy", nel, int(Nfem_T), int(eba), p(Td), p(it), *[p(a) for a in q],
              *[p(a) for a in tab], *[p(a) for a in jc], p(fix), p(val), float(dt), float(rho0),
              p(inc.ptr), p(inc.list), p(VV), p(rhs), p(work), dev.stream())
    return dev.out(VV, numpy_out), dev.out(rhs, numpy_out)
