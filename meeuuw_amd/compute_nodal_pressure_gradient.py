"""Drop-in compute_nodal_pressure_gradient (compute_nodal_pressure_gradient.py:11-31)."""
from .compute_strain_rate import _nodal_gradients


def compute_nodal_pressure_gradient(icon_V, q, nn_V, m_V, nel, dNdr_V_n, dNds_V_n, jcbi00n,
                                    jcbi01n, jcbi10n, jcbi11n):
    """-> (dpdx_n, dpdy_n)."""
    return _nodal_gradients(1, icon_V, q, None, nn_V, nel, dNdr_V_n, dNds_V_n, jcbi00n, jcbi01n,
                            jcbi10n, jcbi11n)
