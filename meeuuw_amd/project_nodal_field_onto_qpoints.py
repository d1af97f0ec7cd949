"""Drop-in nodal -> qpoint projections (project_nodal_field_onto_qpoints.py:11-62)."""
import numpy as np

from . import _dev as dev
from . import _lib


def project_nodal_fields_onto_qpoints(phis, nel, m, N, icon):
    """Several nodal fields at once; m=4 uses the first 4 rows of icon (the
    reference's P variant, also called with icon_V) and N (9, 4)."""
    numpy_out = not dev.is_dev(*phis)
    nel = int(nel)
    ic = dev.to_dev(icon, np.int32)
    Nd = dev.to_dev(np.ascontiguousarray(np.asarray(N)[:, :m]) if not dev.is_dev(N) else N[:, :m],
                    np.float64)
    ph = [dev.to_dev(p, np.float64) for p in phis]
    outs = [dev.empty(nel * 9) for _ in ph]
    pp, _k1 = _lib.ptr_array([dev.ptr(p) for p in ph])
    po, _k2 = _lib.ptr_array([dev.ptr(o) for o in outs])
    _lib.call("mw_nodal_to_qpoints", len(ph), pp, nel, int(m), dev.ptr(Nd), dev.ptr(ic), po,
              dev.stream())
    return [dev.out(o, numpy_out, (nel, 9)) for o in outs]


def project_nodal_Pfield_onto_qpoints(phi_nodal, nq_per_element, nel, m_P, N_P, icon_P):
    """project_nodal_field_onto_qpoints.py:11-23."""
    return project_nodal_fields_onto_qpoints([phi_nodal], nel, m_P, N_P, icon_P)[0]


def project_nodal_Vfield_onto_qpoints(phi_nodal, nq_per_element, nel, m_V, N_V, icon_V):
    """project_nodal_field_onto_qpoints.py:29-41."""
    return project_nodal_fields_onto_qpoints([phi_nodal], nel, m_V, N_V, icon_V)[0]


def project_nodal_Tfield_onto_qpoints(phi_nodal, nq_per_element, nel, m_T, N_T, icon_T):
    """project_nodal_field_onto_qpoints.py:47-59."""
    return project_nodal_fields_onto_qpoints([phi_nodal], nel, m_T, N_T, icon_T)[0]
