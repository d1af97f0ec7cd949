"""CPU pins of the oracle for the curved-geometry PIC functions and the
least-squares projections (closed forms, brute force, reference quirks)."""
import math

import numpy as np
import pytest

from oracle import oracle as O


def curved_particles(cm, ppe, seed):
    rng = np.random.default_rng(seed)
    nel, nelx = cm["nel"], cm["nelx"]
    e = np.repeat(np.arange(nel), ppe)
    r, t = rng.uniform(-1, 1, e.size), rng.uniform(-1, 1, e.size)
    ielx, ielz = e % nelx, e // nelx
    rad = cm["Rinner"] + (ielz + (t + 1) / 2) * cm["hrad"]
    theta = np.pi / 2 - (ielx + (r + 1) / 2) * cm["htheta"]
    return e, r, t, rad, theta


@pytest.mark.parametrize("geometry", ["quarter", "half", "eighth"])
def test_locate_annulus_round_trip(geometry):
    cm = O.curved_mesh(geometry, 6, 4, 0.55, 1.0)
    e, r, t, rad, theta = curved_particles(cm, 5, 1)
    act = np.ones(e.size, dtype=bool)
    act[3] = False
    r2, t2, iel, bad = O.locate_particles___annulus(e.size, act, rad, theta, cm["hrad"],
                                                    cm["htheta"], cm["rad_V"], cm["theta_V"],
                                                    cm["icon_V"], cm["nelx"], cm["Rinner"])
    assert bad == 0
    assert iel[3] == 0 and r2[3] == 0 and t2[3] == 0
    m = act
    np.testing.assert_array_equal(iel[m], e[m])
    np.testing.assert_allclose(r2[m], r[m], atol=1e-12)
    np.testing.assert_allclose(t2[m], t[m], atol=1e-12)
    # the written formula, evaluated in Python
    ip = 7
    ielx = int((math.pi / 2 - theta[ip]) / cm["htheta"])
    ielz = int((rad[ip] - cm["Rinner"]) / cm["hrad"])
    k = cm["nelx"] * ielz + ielx
    assert iel[ip] == k
    assert r2[ip] == ((-theta[ip] + cm["theta_V"][cm["icon_V"][0, k]]) / cm["htheta"] - 0.5) * 2


def test_interpolation_annulus_exact_for_bilinear_in_polar():
    cm = O.curved_mesh("quarter", 5, 3, 0.55, 1.0)
    f = lambda R, T: 0.3 + 1.7 * R - 0.4 * T + 0.9 * R * T + 0.2 * T * T
    u = f(cm["rad_V"], cm["theta_V"])
    w = -2.0 * u
    e, r, t, rad, theta = curved_particles(cm, 3, 2)
    x, z = rad * np.cos(theta), rad * np.sin(theta)
    for ip in range(0, e.size, 4):
        up, wp, iel = O.interpolate_vel_on_pt___annulus(x[ip], z[ip], u, w, cm["hrad"],
                                                        cm["htheta"], cm["nelx"], cm["icon_V"],
                                                        cm["rad_V"], cm["theta_V"], cm["Rinner"])
        assert iel == e[ip]
        R = math.sqrt(x[ip] ** 2 + z[ip] ** 2)
        T = math.pi / 2 - math.atan2(x[ip], z[ip])
        assert up == pytest.approx(f(R, T), rel=1e-12, abs=1e-12)
        assert wp == pytest.approx(-2 * f(R, T), rel=1e-12, abs=1e-12)


@pytest.mark.parametrize("geometry,rk", [("quarter", 1), ("quarter", 2), ("half", 1), ("half", 2),
                                         ("eighth", 1), ("annulus", 1)])
def test_curved_advection_uniform_flow(geometry, rk):
    g = "eighth" if geometry == "annulus" else geometry
    cm = O.curved_mesh(g, 6, 4, 0.55, 1.0)
    U, W = 0.7, -0.3
    u = np.full(cm["rad_V"].size, U)
    w = np.full(cm["rad_V"].size, W)
    e, r, t, rad, theta = curved_particles(cm, 4, 3)
    x, z = rad * np.cos(theta), rad * np.sin(theta)
    x0, z0 = x.copy(), z.copy()
    act = np.ones(e.size, dtype=bool)
    dt = 0.05
    xs, zs, rs, ts, su, sw, a = O.advect_particles_curved(
        geometry, rk, dt, e.size, x, z, rad.copy(), theta.copy(), act, u, w, cm["Rinner"],
        cm["Router"], cm["hrad"], cm["htheta"], cm["nelx"], cm["icon_V"], cm["rad_V"],
        cm["theta_V"])
    xe, ze = x0 + U * dt, z0 + W * dt
    re = np.sqrt(xe ** 2 + ze ** 2)
    out = (xe < 0) | (re < 0.55) | (re > 1.0)
    if geometry == "quarter":
        out |= ze < 0
    if geometry in ("eighth", "annulus"):
        out |= ze < xe
    if rk == 2:  # the midpoint test measures the radius at A (reference quirk)
        xb, zb = x0 + U * dt / 2, z0 + W * dt / 2
        rA = np.sqrt(x0 ** 2 + z0 ** 2)
        outB = (xb < 0) | (rA < 0.55) | (rA > 1.0)
        if geometry == "quarter":
            outB |= zb < 0
        out |= outB
    ub = np.zeros(e.size, dtype=bool)
    if rk == 2:
        # B passed the (radius-at-A) test but lies outside the mesh: the reference
        # then reads outside its arrays (undefined); the oracle returns NaN there
        rB = np.sqrt(xb ** 2 + zb ** 2)
        ub = ~outB & ((rB < 0.55) | (rB > 1.0))
        assert np.all(np.isnan(xs[ub]))
    ok = ~ub
    assert out[ok].any() and (~out[ok]).any()
    np.testing.assert_array_equal(a[ok], ~out[ok])
    keep = ok & ~out
    np.testing.assert_allclose(xs[keep], xe[keep], rtol=0, atol=1e-13)
    np.testing.assert_allclose(zs[keep], ze[keep], rtol=0, atol=1e-13)
    assert np.all(xs[ok & out] == 0) and np.all(zs[ok & out] == 0)
    np.testing.assert_allclose(rs[keep], re[keep], atol=1e-13)


def test_curved_advection_unimplemented_orders_do_nothing():
    cm = O.curved_mesh("eighth", 4, 3, 0.55, 1.0)
    e, r, t, rad, theta = curved_particles(cm, 2, 4)
    x, z = rad * np.cos(theta), rad * np.sin(theta)
    x0 = x.copy()
    act = np.ones(e.size, dtype=bool)
    u = np.ones(cm["rad_V"].size)
    res = O.advect_particles_curved("eighth", 2, 0.1, e.size, x, z, rad, theta, act, u, u,
                                    0.55, 1.0, cm["hrad"], cm["htheta"], cm["nelx"], cm["icon_V"],
                                    cm["rad_V"], cm["theta_V"])
    np.testing.assert_array_equal(res[0], x0)
    assert np.all(res[4] == 0) and np.all(res[6])


@pytest.mark.parametrize("nb", [3, 4])
def test_least_squares_exact_fits_and_extrema(nb):
    rng = np.random.default_rng(nb)
    m = O.Mesh(4, 3)
    nel = m.nel
    x_e = (m.x_V[m.icon_V[0]] + m.x_V[m.icon_V[2]]) / 2
    z_e = (m.z_V[m.icon_V[0]] + m.z_V[m.icon_V[2]]) / 2
    ppe = 9
    e = rng.permutation(np.repeat(np.arange(nel), ppe)).astype(np.int32)
    sx = rng.uniform(-0.5, 0.5, e.size) * m.hx
    sz = rng.uniform(-0.5, 0.5, e.size) * m.hz
    x, z = x_e[e] + sx, z_e[e] + sz
    ca = rng.standard_normal((nel, 4))
    if nb == 3:
        ca[:, 3] = 0
    rho = ca[e, 0] + ca[e, 1] * sx + ca[e, 2] * sz + ca[e, 3] * sx * sz
    eta = 2 * rho + 1
    cr, ce, rmin, rmax, emin, emax = O.compute_ls_coefficients(nb, nel, x_e, z_e, x, z, e, rho, eta)
    np.testing.assert_allclose(cr, ca[:, :nb], rtol=1e-9, atol=1e-9)
    want_e = 2 * ca[:, :nb].copy()
    want_e[:, 0] += 1
    np.testing.assert_allclose(ce, want_e, rtol=1e-9, atol=1e-9)
    for k in range(nel):
        sel = e == k
        assert rmin[k] == rho[sel].min() and rmax[k] == rho[sel].max()
        assert emin[k] == eta[sel].min() and emax[k] == eta[sel].max()
    # noisy data: agrees with numpy's normal-equation solve (np.linalg.solve, as the reference)
    rho2 = rho + 0.1 * rng.standard_normal(e.size)
    cr2 = O.compute_ls_coefficients(nb, nel, x_e, z_e, x, z, e, rho2, eta)[0]
    for k in range(nel):
        sel = e == k
        B = np.column_stack([np.ones(sel.sum()), sx[sel], sz[sel], sx[sel] * sz[sel]])[:, :nb]
        np.testing.assert_allclose(cr2[k], np.linalg.solve(B.T @ B, B.T @ rho2[sel]), rtol=1e-9)
    # the 10/12-tuples of the reference signatures
    assert len(O.compute_ls_coefficients_P1(nel, x_e, z_e, x, z, e, rho, eta)) == 10
    assert len(O.compute_ls_coefficients_Q1(nel, x_e, z_e, x, z, e, rho, eta)) == 12


def test_limiter_cases():
    assert O.limiter(0.5, 0.1, -0.1, 0, 1, 1) == (0.1, -0.1)        # within bounds: untouched
    c1, c2 = O.limiter(0.9, 1.0, -1.0, 0, 1, 1)                      # clamp, then shrink
    assert c1 == pytest.approx(0.1) and c2 == pytest.approx(-0.1)
    c1, c2 = O.limiter(0.5, 3.0, 0.2, 0, 1, 0.5)                     # clamp c1 only
    assert (abs(c1) + abs(c2)) * 0.5 / 2 <= min(0.5, 0.5) + 1e-15
    assert c1 > 0 and c2 > 0


def _node_jacobians(O, m, x_V, z_V):
    j = [np.zeros((m.nel, 9)) for _ in range(4)]
    O.lib().orc_jacobians(m.nel, 9, O._p(m.dNdr_V_n), O._p(m.dNdt_V_n), None, O._p(O._f64(x_V)),
                          O._p(O._f64(z_V)), O._p(m.icon_V), None, *[O._p(a) for a in j])
    return j


@pytest.mark.parametrize("curved", [False, True])
def test_strain_rate_and_pressure_gradient_exact_for_polynomials(curved):
    if curved:
        cm = O.curved_mesh("quarter", 6, 4, 0.55, 1.0)
        m, X, Z = cm["mesh"], cm["x_V"], cm["z_V"]
        u, v, q = 0.3 + 1.1 * X - 0.7 * Z, -0.2 + 0.4 * X + 0.9 * Z, 2.0 - 3.0 * X + 0.5 * Z
        g = dict(exx=1.1, eyy=0.9, exy=0.5 * (-0.7 + 0.4), px=-3.0 + 0 * X, pz=0.5 + 0 * X)
    else:
        m = O.Mesh(6, 5)
        X, Z = m.x_V, m.z_V
        u = 0.3 + 1.1 * X - 0.7 * Z + 0.5 * X * X + 0.2 * X * Z
        v = -0.2 + 0.4 * X + 0.9 * Z - 0.3 * Z * Z + 0.6 * X * Z
        q = 2.0 - 3.0 * X + 0.5 * Z + X * Z + 0.7 * Z * Z
        g = dict(exx=1.1 + X + 0.2 * Z, eyy=0.9 - 0.6 * Z + 0.6 * X,
                 exy=0.5 * (-0.7 + 0.2 * X) + 0.5 * (0.4 + 0.6 * Z), px=-3.0 + Z, pz=0.5 + X + 1.4 * Z)
    j = _node_jacobians(O, m, X, Z)
    exx, eyy, exy = O.compute_nodal_strain_rate(m.icon_V, u, v, X.size, 9, m.nel, m.dNdr_V_n,
                                                m.dNdt_V_n, *j)
    np.testing.assert_allclose(exx, g["exx"] + 0 * X, atol=1e-9)
    np.testing.assert_allclose(eyy, g["eyy"] + 0 * X, atol=1e-9)
    np.testing.assert_allclose(exy, g["exy"] + 0 * X, atol=1e-9)
    px, pz = O.compute_nodal_pressure_gradient(m.icon_V, q, X.size, 9, m.nel, m.dNdr_V_n,
                                               m.dNdt_V_n, *j)
    np.testing.assert_allclose(px, g["px"], atol=1e-9)
    np.testing.assert_allclose(pz, g["pz"], atol=1e-9)
    ee = O.compute_elemental_strain_rate(m.icon_V, u, v, X.size, m.nel, X, Z)
    xc, zc = X[m.icon_V[8]], Z[m.icon_V[8]]
    if curved:
        np.testing.assert_allclose(ee[0], 1.1, atol=1e-9)
        np.testing.assert_allclose(ee[2], 0.5 * (-0.7 + 0.4), atol=1e-9)
    else:
        np.testing.assert_allclose(ee[0], 1.1 + xc + 0.2 * zc, atol=1e-9)
        np.testing.assert_allclose(ee[1], 0.9 - 0.6 * zc + 0.6 * xc, atol=1e-9)
