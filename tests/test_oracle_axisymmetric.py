"""The oracle's axisymmetric branches (build_matrix_stokes.py:108-125,
compute_pressure_offset.py:33-39, 49-55) pinned by closed forms.  On a box
[0, Lx] x [0, Lz] with x the radius, the 3x3 Gauss rule integrates these
fields exactly, so for constant viscosity:
  u = a x (radial expansion): strain (a, a, 0, 0), x^T K x = 4 a^2 eta pi Lx^2 Lz;
  w = b z: strain (0, 0, b, 0), x^T K x = 2 b^2 eta pi Lx^2 Lz;
  w = c x (shear): strain (0, 0, 0, c), x^T K x = c^2 eta pi Lx^2 Lz;
  the divergence du/dx + u/x + dw/dz of u = a x is 2a, so with p = 1 the
  pressure block gives -2 a pi Lx^2 Lz eta_ref/L_ref;
  the body force sums to rho g pi Lx^2 Lz.
A dropped hoop term, a wrong weight in C, a missing 2 pi x or a hoop entry on
the wrong component fails one of them."""
import numpy as np
import pytest
from scipy import sparse

from conftest import stokes_inputs


def _axi_system(O, nelx, nelz, Lx, Lz, eta=2.5, rho=1.3, gx=0.4, gz=-1.1, bc=None):
    m = O.Mesh(nelx, nelz, Lx=Lx, Lz=Lz)
    one = np.ones((m.nel, 9))
    fix = np.zeros(m.Nfem_V, dtype=bool) if bc is None else bc
    args = list(stokes_inputs(O, m, rho * one, eta * one, gx * one, gz * one,
                              (fix, np.zeros(m.Nfem_V))))
    args[41] = True  # axisymmetric
    out = O.build_matrix_stokes(*args, True)
    coo = O.coo_indices(m)
    K = sparse.coo_matrix((out[3], coo["K"]), shape=(m.Nfem_V, m.Nfem_V)).tocsr()
    G = sparse.coo_matrix((out[4], coo["G"]), shape=(m.Nfem_V, m.Nfem_P)).tocsr()
    return m, K, G, out, args


@pytest.mark.parametrize("nelx,nelz,Lx,Lz", [(6, 4, 1.0, 1.0), (5, 7, 2.0, 0.7)])
def test_axisymmetric_stokes_closed_forms(oracle, nelx, nelz, Lx, Lz):
    eta, rho, gx, gz = 2.5, 1.3, 0.4, -1.1
    m, K, G, out, args = _axi_system(oracle, nelx, nelz, Lx, Lz, eta, rho, gx, gz)
    X, Z = m.x_V, m.z_V
    vol = np.pi * Lx ** 2 * Lz  # integral of 2 pi x over the box
    gs = 1.0 / m.L_ref  # eta_ref / L_ref

    def vec(u, w):
        v = np.empty(m.Nfem_V)
        v[0::2], v[1::2] = u, w
        return v

    a, b, c = 0.7, -1.3, 0.9
    ua, wb, wc = vec(a * X, 0 * X), vec(0 * X, b * Z), vec(0 * X, c * X)
    assert ua @ (K @ ua) == pytest.approx(4 * a * a * eta * vol, rel=1e-12)
    assert wb @ (K @ wb) == pytest.approx(2 * b * b * eta * vol, rel=1e-12)
    assert wc @ (K @ wc) == pytest.approx(c * c * eta * vol, rel=1e-12)
    # strains (a,a,0,0) and (0,0,b,0) are C-orthogonal
    assert abs(ua @ (K @ wb)) < 1e-12 * eta * vol
    # rigid translation along the axis is strain-free
    assert np.abs(K @ vec(0 * X, 1 + 0 * X)).max() < 1e-12 * np.abs(K).max()
    ones_p = np.ones(m.Nfem_P)
    assert ones_p @ (G.T @ ua) == pytest.approx(-2 * a * vol * gs, rel=1e-12)
    assert ones_p @ (G.T @ wb) == pytest.approx(-b * vol * gs, rel=1e-12)
    f = out[1]
    assert f[0::2].sum() == pytest.approx(rho * gx * vol, rel=1e-12)
    assert f[1::2].sum() == pytest.approx(rho * gz * vol, rel=1e-12)
    # the branch forms no pressure mass or H blocks
    assert not np.any(out[6]) and not np.any(out[7]) and not np.any(out[8])


def test_axisymmetric_pressure_offset_volume(oracle):
    m = oracle.Mesh(6, 5, Lx=1.5, Lz=0.8)
    for p, exact in ((np.ones(m.Nfem_P), np.pi * 1.5 ** 2 * 0.8),
                     (m.x_P.copy(), 2 * np.pi * 1.5 ** 3 / 3 * 0.8)):
        off = oracle.compute_pressure_offset("box", "volume", True, m.top_element, m.nel, m.nelx, 9,
                                             m.N_P, m.JxWq, p, m.icon_P, None, m.xq, m.volume)
        assert off == pytest.approx(exact, rel=1e-13)  # the branch does not divide by the volume


def test_axisymmetric_pressure_offset_surface_converges(oracle):
    """p = 1 on the top of a half annulus: the sum is the midpoint rule of
    (1/2) int_0^pi sin = 1, with an error that falls 4x per halving of h."""
    err = []
    for nelx in (16, 32, 64):
        cm = oracle.curved_mesh("half", nelx, 2, 0.55, 1.0)
        m = cm["mesh"]
        theta_P = np.pi / 2 - m.x_P * np.pi
        off = oracle.compute_pressure_offset("half", "surface", True, m.top_element, m.nel, nelx, 9,
                                             m.N_P, m.JxWq, np.ones(m.Nfem_P), m.icon_P, theta_P,
                                             m.xq, 1.0)
        err.append(off - 1.0)
    assert abs(err[-1]) < 1e-3
    assert err[0] / err[1] == pytest.approx(4.0, rel=0.02)
    assert err[1] / err[2] == pytest.approx(4.0, rel=0.02)
