"""GPU curved-geometry PIC functions and least-squares projections against the
oracle.  locate, box interpolation, least squares and the limiter are
compared bit for bit (same expressions, same order; the small solves use the
same pivoted elimination); the polar interpolation and advection use atan2,
whose device and host implementations may differ in the last bit, so they are
compared to 1e-12."""
import numpy as np
import pytest

from oracle import oracle as O
from test_oracle_curved_ls import curved_particles

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("geometry", ["quarter", "half", "eighth"])
def test_locate_annulus_bit_exact(geometry):
    from meeuuw_amd import pic_functions as P
    from meeuuw_amd.locate_particles import locate_particles
    cm = O.curved_mesh(geometry, 24, 10, 0.55, 1.0)
    e, r, t, rad, theta = curved_particles(cm, 9, 5)
    act = np.ones(e.size, dtype=bool)
    act[::17] = False
    want = O.locate_particles___annulus(e.size, act, rad, theta, cm["hrad"], cm["htheta"],
                                        cm["rad_V"], cm["theta_V"], cm["icon_V"], cm["nelx"],
                                        cm["Rinner"])
    got = P.locate_particles___annulus(e.size, act, rad, theta, cm["hrad"], cm["htheta"],
                                       cm["rad_V"], cm["theta_V"], cm["icon_V"], cm["nelx"],
                                       cm["Rinner"])
    for a, b in zip(got, want[:3]):
        np.testing.assert_array_equal(a, b)
    got2 = locate_particles(geometry, e.size, act, None, None, rad, theta, 0, 0, cm["hrad"],
                            cm["htheta"], None, None, cm["rad_V"], cm["theta_V"], cm["icon_V"],
                            cm["nelx"], cm["Rinner"])
    np.testing.assert_array_equal(got2[2], want[2])
    bad = rad.copy()
    bad[5] = 2.0  # outside the mesh
    with pytest.raises(SystemExit):
        P.locate_particles___annulus(e.size, act, bad, theta, cm["hrad"], cm["htheta"],
                                     cm["rad_V"], cm["theta_V"], cm["icon_V"], cm["nelx"],
                                     cm["Rinner"])
    with pytest.raises(SystemExit):
        locate_particles("annulus", e.size, act, None, None, rad, theta, 0, 0, 0, 0, None, None,
                         None, None, cm["icon_V"], cm["nelx"], 0.55)


def test_vel_on_pts_box_and_annulus():
    from meeuuw_amd import pic_functions as P
    rng = np.random.default_rng(3)
    m = O.Mesh(12, 9)
    u, w = rng.standard_normal(m.nn_V), rng.standard_normal(m.nn_V)
    xp, zp = rng.uniform(0, 1, 500), rng.uniform(0, 1, 500)
    want = O.interpolate_vel_on_pts___box(xp, zp, u, w, m.hx, m.hz, m.nelx, m.icon_V, m.x_V, m.z_V)
    got = P.interpolate_vel_on_pt___box(xp, zp, u, w, m.hx, m.hz, m.nelx, m.icon_V, m.x_V, m.z_V)
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)
    one = P.interpolate_vel_on_pt___box(xp[0], zp[0], u, w, m.hx, m.hz, m.nelx, m.icon_V, m.x_V,
                                        m.z_V)
    assert one == (want[0][0], want[1][0], want[2][0])
    cm = O.curved_mesh("half", 16, 6, 0.55, 1.0)
    u, w = rng.standard_normal(cm["rad_V"].size), rng.standard_normal(cm["rad_V"].size)
    e, r, t, rad, theta = curved_particles(cm, 3, 6)
    x, z = rad * np.cos(theta), rad * np.sin(theta)
    got = P.interpolate_vel_on_pt___annulus(x, z, u, w, cm["hrad"], cm["htheta"], cm["nelx"],
                                            cm["icon_V"], cm["rad_V"], cm["theta_V"], 0.55)
    for k in range(0, e.size, 7):
        up, wp, iel = O.interpolate_vel_on_pt___annulus(x[k], z[k], u, w, cm["hrad"],
                                                        cm["htheta"], cm["nelx"], cm["icon_V"],
                                                        cm["rad_V"], cm["theta_V"], 0.55)
        assert got[2][k] == iel
        assert got[0][k] == pytest.approx(up, rel=1e-12, abs=1e-13)
        assert got[1][k] == pytest.approx(wp, rel=1e-12, abs=1e-13)


@pytest.mark.parametrize("geometry,rk", [("quarter", 1), ("quarter", 2), ("half", 1), ("half", 2),
                                         ("eighth", 1), ("annulus", 1), ("quarter", 4)])
def test_curved_advection_vs_oracle(geometry, rk):
    from meeuuw_amd import pic_functions as P
    g = "eighth" if geometry == "annulus" else geometry
    cm = O.curved_mesh(g, 20, 8, 0.55, 1.0)
    rng = np.random.default_rng(rk)
    # a rigid rotation: particles cross elements and leave through the straight
    # sides; with dt = 0.05 no RK2 midpoint leaves the mesh radially (there the
    # reference reads outside its arrays and the GPU path raises instead)
    X, Z = cm["x_V"], cm["z_V"]
    u, w = 0.8 * Z, -0.8 * X
    e, r, t, rad, theta = curved_particles(cm, 6, 11)
    x, z = rad * np.cos(theta), rad * np.sin(theta)
    act = np.ones(e.size, dtype=bool)
    act[::13] = False
    dt = 0.05
    args = lambda xx, zz, ra, th, aa: (rk, dt, e.size, xx, zz, ra, th, aa, u, w, 0.55, 1.0,
                                       cm["hrad"], cm["htheta"], cm["nelx"], cm["icon_V"],
                                       cm["rad_V"], cm["theta_V"])
    want = O.advect_particles_curved(geometry, *args(x.copy(), z.copy(), rad.copy(), theta.copy(),
                                                      act.copy()))
    xs, zs, ra, th, aa = x.copy(), z.copy(), rad.copy(), theta.copy(), act.copy()
    f = getattr(P, f"advect_particles___{geometry}")
    got = f(*args(xs, zs, ra, th, aa))
    assert got[0] is xs and got[6] is aa  # mutated in place and returned
    np.testing.assert_array_equal(got[6], want[6])
    if rk in (1, 2):
        assert (~want[6] & act).any() and want[6].any()
    for a, b in zip(got[:6], want[:6]):
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("nb", [3, 4])
def test_least_squares_bit_exact(nb):
    from meeuuw_amd import pic_functions as P
    rng = np.random.default_rng(10 + nb)
    m = O.Mesh(24, 16)
    nel = m.nel
    x_e = (m.x_V[m.icon_V[0]] + m.x_V[m.icon_V[2]]) / 2
    z_e = (m.z_V[m.icon_V[0]] + m.z_V[m.icon_V[2]]) / 2
    e = rng.permutation(np.repeat(np.arange(nel), 12)).astype(np.int32)
    e[:40] = 0  # particles with iel 0 (e.g. deactivated ones) join element 0, as in the reference
    x = x_e[e] + rng.uniform(-0.5, 0.5, e.size) * m.hx
    z = z_e[e] + rng.uniform(-0.5, 0.5, e.size) * m.hz
    rho = rng.uniform(1, 2, e.size)
    eta = 10 ** rng.uniform(-3, 0, e.size)
    want = (O.compute_ls_coefficients_P1 if nb == 3 else O.compute_ls_coefficients_Q1)(
        nel, x_e, z_e, x, z, e, rho, eta)
    got = (P.compute_ls_coefficients_P1 if nb == 3 else P.compute_ls_coefficients_Q1)(
        nel, x_e, z_e, x, z, e, rho, eta)
    assert len(got) == len(want)
    for a, b in zip(got, want):
        np.testing.assert_array_equal(a, b)
    # the driver's limiter loop (meeuuw.py:1243-1249) as one kernel
    c1, c2 = P.limiter_elements(want[0], want[1], want[2], want[-4], want[-3], m.hx)
    for k in range(0, nel, 5):
        ref = O.limiter(want[0][k], want[1][k], want[2][k], want[-4][k], want[-3][k], m.hx)
        assert (c1[k], c2[k]) == ref
    assert P.limiter(0.9, 1.0, -1.0, 0, 1, 1) == O.limiter(0.9, 1.0, -1.0, 0, 1, 1)


@pytest.mark.parametrize("curved", [False, True])
def test_derived_fields_bit_exact(curved):
    from meeuuw_amd.compute_nodal_pressure_gradient import compute_nodal_pressure_gradient
    from meeuuw_amd.compute_strain_rate import (compute_elemental_strain_rate,
                                                compute_nodal_strain_rate)
    from test_oracle_curved_ls import _node_jacobians
    rng = np.random.default_rng(5)
    if curved:
        cm = O.curved_mesh("half", 40, 12, 0.55, 1.0)
        m, X, Z = cm["mesh"], cm["x_V"], cm["z_V"]
    else:
        m = O.Mesh(64, 48)
        X, Z = m.x_V, m.z_V
    u, v, q = (rng.standard_normal(X.size) for _ in range(3))
    j = _node_jacobians(O, m, X, Z)
    args = (m.icon_V, u, v, X.size, 9, m.nel, m.dNdr_V_n, m.dNdt_V_n, *j)
    for a, b in zip(compute_nodal_strain_rate(*args), O.compute_nodal_strain_rate(*args)):
        np.testing.assert_array_equal(a, b)
    args = (m.icon_V, q, X.size, 9, m.nel, m.dNdr_V_n, m.dNdt_V_n, *j)
    for a, b in zip(compute_nodal_pressure_gradient(*args), O.compute_nodal_pressure_gradient(*args)):
        np.testing.assert_array_equal(a, b)
    args = (m.icon_V, u, v, X.size, m.nel, X, Z)
    for a, b in zip(compute_elemental_strain_rate(*args), O.compute_elemental_strain_rate(*args)):
        np.testing.assert_array_equal(a, b)
