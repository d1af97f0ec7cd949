"""Drop-in particle-in-cell routines (pic_functions.py:17-1053) on the GPU.

Same names, positional arguments and return values as the reference.  As in
the reference, advect_particles___box mutates swarm_x, swarm_z and
swarm_active in place (NumPy inputs are updated from the device result) and
also returns them.
"""
import numpy as np

from . import _dev as dev
from . import _lib

AVERAGING = {"arithmetic": 0, "geometric": 1, "harmonic": 2}


def _avg(name):
    if name not in AVERAGING:
        raise ValueError(f"unknown averaging '{name}'")
    return AVERAGING[name]


def interpolate_fields_on_particles(nparticle, swarm_active, swarm_r, swarm_t, swarm_iel, phis, icon):
    """Several nodal fields onto the particles in one pass (fused form of
    interpolate_field_on_particles, pic_functions.py:84-101)."""
    numpy_out = not dev.is_dev(swarm_r, swarm_t, *phis)
    n = int(nparticle)
    ic = dev.to_dev(icon, np.int32)
    act = dev.to_dev(swarm_active, np.uint8)
    r, t = dev.to_dev(swarm_r, np.float64), dev.to_dev(swarm_t, np.float64)
    iel = dev.to_dev(swarm_iel, np.int32)
    ph = [dev.to_dev(p, np.float64) for p in phis]
    outs = [dev.empty(n) for _ in phis]
    pp, _k1 = _lib.ptr_array([dev.ptr(p) for p in ph])
    po, _k2 = _lib.ptr_array([dev.ptr(o) for o in outs])
    _lib.call("mw_interpolate_fields", n, len(ph), dev.ptr(act), dev.ptr(r), dev.ptr(t),
              dev.ptr(iel), pp, dev.ptr(ic), ic.shape[1], po, dev.stream())
    return [dev.out(o, numpy_out) for o in outs]


def interpolate_field_on_particles(nparticle, swarm_active, swarm_r, swarm_t, swarm_iel, phi, icon):
    """pic_functions.py:84-101."""
    return interpolate_fields_on_particles(nparticle, swarm_active, swarm_r, swarm_t, swarm_iel,
                                           [phi], icon)[0]


def locate_particles___box(nparticle, swarm_active, swarm_x, swarm_z, hx, hz, x_V, z_V, icon, nelx):
    """pic_functions.py:107-137.  An element index outside the mesh (the
    reference reads out of bounds there) raises SystemExit."""
    r, t, iel, flags = _locate_box(nparticle, swarm_active, swarm_x, swarm_z, hx, hz, x_V, z_V,
                                   icon, nelx)
    if flags[0]:
        raise SystemExit("r value out of bounds")
    return r, t, iel


def _locate_box(nparticle, swarm_active, swarm_x, swarm_z, hx, hz, x_V, z_V, icon, nelx):
    numpy_out = not dev.is_dev(swarm_x, swarm_z)
    n = int(nparticle)
    ic = dev.to_dev(icon, np.int32)
    act = dev.to_dev(swarm_active, np.uint8)
    x, z = dev.to_dev(swarm_x, np.float64), dev.to_dev(swarm_z, np.float64)
    xv, zv = dev.to_dev(x_V, np.float64), dev.to_dev(z_V, np.float64)
    r, t, iel = dev.empty(n), dev.empty(n), dev.empty(n, np.int32)
    flags = dev.empty(3, np.int32)
    _lib.call("mw_locate_particles_box", n, dev.ptr(act), dev.ptr(x), dev.ptr(z), float(hx),
              float(hz), dev.ptr(xv), dev.ptr(zv), dev.ptr(ic), ic.shape[1], int(nelx), dev.ptr(r),
              dev.ptr(t), dev.ptr(iel), dev.ptr(flags), dev.stream())
    fl = dev.host(flags)
    return dev.out(r, numpy_out), dev.out(t, numpy_out), dev.out(iel, numpy_out), fl


def advect_particles___box(RKorder, dt, nparticle, swarm_x, swarm_z, swarm_active, u, w, Lx, Lz, hx,
                           hz, nelx, icon, x_V, z_V):
    """pic_functions.py:179-292 (RK1, RK2, RK4)."""
    if RKorder not in (1, 2, 4):
        raise ValueError("advect_particles___box: RKorder must be 1, 2 or 4")
    numpy_out = not dev.is_dev(swarm_x, swarm_z)
    n = int(nparticle)
    ic = dev.to_dev(icon, np.int32)
    x, z = dev.to_dev(swarm_x, np.float64), dev.to_dev(swarm_z, np.float64)
    act = dev.to_dev(swarm_active, np.uint8)
    ud, wd = dev.to_dev(u, np.float64), dev.to_dev(w, np.float64)
    xv, zv = dev.to_dev(x_V, np.float64), dev.to_dev(z_V, np.float64)
    su, sw = dev.empty(n), dev.empty(n)
    flags = dev.empty(2, np.int32)
    _lib.call("mw_advect_particles_box", int(RKorder), float(dt), n, dev.ptr(x), dev.ptr(z),
              dev.ptr(act), dev.ptr(ud), dev.ptr(wd), float(Lx), float(Lz), float(hx), float(hz),
              int(nelx), ic.shape[1], dev.ptr(ic), dev.ptr(xv), dev.ptr(zv), dev.ptr(su),
              dev.ptr(sw), dev.ptr(flags), dev.stream())
    if dev.host(flags)[0]:
        raise SystemExit("advect_particles___box: particle position outside the mesh")
    if numpy_out:
        swarm_x[:] = dev.host(x)
        swarm_z[:] = dev.host(z)
        swarm_active[:] = dev.host(act).astype(bool)
        return swarm_x, swarm_z, dev.host(su), dev.host(sw), swarm_active
    if isinstance(swarm_x, type(x)) and swarm_x.dtype == x.dtype:
        swarm_x.copy_(x)
        swarm_z.copy_(z)
        swarm_active.copy_(act.to(swarm_active.dtype))
        return swarm_x, swarm_z, su, sw, swarm_active
    return x, z, su, sw, act


class ElementSort:
    """Particles grouped by element (stable), shared by the projections."""

    def __init__(self, nel, nparticle, swarm_active, swarm_iel):
        self.nel, self.n = int(nel), int(nparticle)
        act = dev.to_dev(swarm_active, np.uint8)
        iel = dev.to_dev(swarm_iel, np.int32)
        self.order = dev.empty(self.n, np.int32)
        self.seg = dev.empty(self.nel + 1, np.int32)
        nb = _lib.lib.mw_sort_workspace_bytes(self.n, self.nel)
        work = dev.empty((nb + 7) // 8)
        _lib.call("mw_sort_particles_by_element", self.n, self.nel, dev.ptr(act), dev.ptr(iel),
                  dev.ptr(self.order), dev.ptr(self.seg), dev.ptr(work), dev.stream())


def project_particles_on_elements(nel, nparticle, swarm_active, swarm_rho, swarm_eta, swarm_iel,
                                  averaging):
    """pic_functions.py:602-650 -> (rho_elemental, eta_elemental, nparticle_per_element)."""
    numpy_out = not dev.is_dev(swarm_rho, swarm_eta)
    srt = ElementSort(nel, nparticle, swarm_active, swarm_iel)
    rho, eta = dev.to_dev(swarm_rho, np.float64), dev.to_dev(swarm_eta, np.float64)
    rho_e, eta_e, cnt = dev.empty(nel), dev.empty(nel), dev.empty(nel, np.int32)
    _lib.call("mw_project_particles_on_elements", int(nel), dev.ptr(srt.seg), dev.ptr(srt.order),
              dev.ptr(rho), dev.ptr(eta), _avg(averaging), dev.ptr(rho_e), dev.ptr(eta_e),
              dev.ptr(cnt), dev.stream())
    return dev.out(rho_e, numpy_out), dev.out(eta_e, numpy_out), dev.out(cnt, numpy_out)


def project_particle_fields_on_nodes(ptype, nel, nn_V, nparticle, swarm_active, swarm_phis, icon,
                                     swarm_iel, swarm_r, swarm_t, averagings, sort=None):
    """Several particle fields onto the V nodes in one pass (fused form of
    project_particle_field_on_nodes_{1..4})."""
    numpy_out = not dev.is_dev(swarm_r, swarm_t, *swarm_phis)
    srt = sort or ElementSort(nel, nparticle, swarm_active, swarm_iel)
    ic = dev.to_dev(icon, np.int32)
    r, t = dev.to_dev(swarm_r, np.float64), dev.to_dev(swarm_t, np.float64)
    ph = [dev.to_dev(p, np.float64) for p in swarm_phis]
    outs = [dev.empty(nn_V) for _ in ph]
    inc = dev.incidence(ic, nn_V, dev.array_key(icon))
    nf = len(ph)
    work = dev.empty(int(nel) * nf * 8)
    pp, _k1 = _lib.ptr_array([dev.ptr(p) for p in ph])
    po, _k2 = _lib.ptr_array([dev.ptr(o) for o in outs])
    import ctypes
    av = (ctypes.c_int * nf)(*[_avg(a) for a in averagings])
    _lib.call("mw_project_particle_fields_on_nodes", int(ptype), int(nel), int(nn_V), nf,
              dev.ptr(srt.seg), dev.ptr(srt.order), pp, ctypes.cast(av, ctypes.c_void_p),
              dev.ptr(ic), dev.ptr(r), dev.ptr(t), dev.ptr(inc.ptr), dev.ptr(inc.list),
              dev.ptr(work), po, dev.stream())
    return [dev.out(o, numpy_out) for o in outs]


def _nodes(ptype):
    def f(nel, nn_V, nparticle, swarm_active, swarm_phi, icon, swarm_iel, swarm_r, swarm_t,
          averaging):
        return project_particle_fields_on_nodes(ptype, nel, nn_V, nparticle, swarm_active,
                                                [swarm_phi], icon, swarm_iel, swarm_r, swarm_t,
                                                [averaging])[0]
    return f


project_particle_field_on_nodes_1 = _nodes(1)
project_particle_field_on_nodes_1.__doc__ = "pic_functions.py:658-724 (all particles, unweighted)."
project_particle_field_on_nodes_2 = _nodes(2)
project_particle_field_on_nodes_2.__doc__ = "pic_functions.py:732-811 (Q1-weighted)."
project_particle_field_on_nodes_3 = _nodes(3)
project_particle_field_on_nodes_3.__doc__ = "pic_functions.py:820-925 (quadrant, unweighted)."
project_particle_field_on_nodes_4 = _nodes(4)
project_particle_field_on_nodes_4.__doc__ = "pic_functions.py:936-1053 (quadrant, weighted)."


# ---------------------------------------------------------------------------
# curved geometries (quarter, half, eighth; the annulus-formula functions)
# ---------------------------------------------------------------------------
def _locate_annulus(nparticle, swarm_active, swarm_rad, swarm_theta, hrad, htheta, rad_V, theta_V,
                    icon, nelx, Rinner):
    numpy_out = not dev.is_dev(swarm_rad, swarm_theta)
    n = int(nparticle)
    ic = dev.to_dev(icon, np.int32)
    act = dev.to_dev(swarm_active, np.uint8)
    ra, th = dev.to_dev(swarm_rad, np.float64), dev.to_dev(swarm_theta, np.float64)
    rv, tv = dev.to_dev(rad_V, np.float64), dev.to_dev(theta_V, np.float64)
    r, t, iel = dev.empty(n), dev.empty(n), dev.empty(n, np.int32)
    flags = dev.empty(3, np.int32)
    _lib.call("mw_locate_particles_annulus", n, dev.ptr(act), dev.ptr(ra), dev.ptr(th),
              float(hrad), float(htheta), dev.ptr(rv), dev.ptr(tv), dev.ptr(ic), ic.shape[1],
              int(nelx), float(Rinner), dev.ptr(r), dev.ptr(t), dev.ptr(iel), dev.ptr(flags),
              dev.stream())
    fl = dev.host(flags)
    return dev.out(r, numpy_out), dev.out(t, numpy_out), dev.out(iel, numpy_out), fl


def locate_particles___annulus(nparticle, swarm_active, swarm_rad, swarm_theta, hrad, htheta,
                               rad_V, theta_V, icon, nelx, Rinner):
    """pic_functions.py:142-173.  An element index outside the mesh (the
    reference reads out of bounds there) raises SystemExit."""
    r, t, iel, flags = _locate_annulus(nparticle, swarm_active, swarm_rad, swarm_theta, hrad,
                                       htheta, rad_V, theta_V, icon, nelx, Rinner)
    if flags[0]:
        raise SystemExit("r value out of bounds")
    return r, t, iel


def _vel_on_pts(polar, xp, zp, u, w, geo, nelx, icon):
    scalar = np.ndim(xp) == 0 and not dev.is_dev(xp)
    x = dev.to_dev(np.atleast_1d(xp) if not dev.is_dev(xp) else xp, np.float64)
    z = dev.to_dev(np.atleast_1d(zp) if not dev.is_dev(zp) else zp, np.float64)
    n = int(x.numel())
    ic = dev.to_dev(icon, np.int32)
    ud, wd = dev.to_dev(u, np.float64), dev.to_dev(w, np.float64)
    g = [dev.to_dev(a, np.float64) if a is not None else None for a in geo[2:4]]
    up, wp, iel = dev.empty(n), dev.empty(n), dev.empty(n, np.int32)
    flags = dev.empty(1, np.int32)
    zero = dev.empty(1)
    h0, h1 = float(geo[0]), float(geo[1])
    args = ([h0, h1, dev.ptr(g[0]), dev.ptr(g[1]), 0.0, 0.0, dev.ptr(zero), dev.ptr(zero), 0.0]
            if not polar else
            [0.0, 0.0, dev.ptr(zero), dev.ptr(zero), h0, h1, dev.ptr(g[0]), dev.ptr(g[1]),
             float(geo[4])])
    _lib.call("mw_interpolate_vel_on_pts", int(polar), n, dev.ptr(x), dev.ptr(z), dev.ptr(ud),
              dev.ptr(wd), *args, int(nelx), ic.shape[1], dev.ptr(ic), dev.ptr(up), dev.ptr(wp),
              dev.ptr(iel), dev.ptr(flags), dev.stream())
    if dev.host(flags)[0]:
        raise SystemExit("interpolate_vel_on_pt: point outside the mesh")
    if scalar:
        return float(dev.host(up)[0]), float(dev.host(wp)[0]), int(dev.host(iel)[0])
    numpy_out = not dev.is_dev(xp)
    return dev.out(up, numpy_out), dev.out(wp, numpy_out), dev.out(iel, numpy_out)


def interpolate_vel_on_pt___box(xp, zp, u, w, hx, hz, nelx, icon, x_V, z_V):
    """pic_functions.py:17-37 -> (up, wp, iel); also accepts arrays of points."""
    return _vel_on_pts(0, xp, zp, u, w, (hx, hz, x_V, z_V), nelx, icon)


def interpolate_vel_on_pt___annulus(xp, yp, u, w, hrad, htheta, nelx, icon, rad_V, theta_V, Rinner):
    """pic_functions.py:40-63 -> (up, wp, iel); also accepts arrays of points."""
    return _vel_on_pts(1, xp, yp, u, w, (hrad, htheta, rad_V, theta_V, Rinner), nelx, icon)


CURVED = {"quarter": 0, "half": 1, "eighth": 2, "annulus": 3}


def _advect_curved(geometry, RKorder, dt, nparticle, swarm_x, swarm_z, swarm_rad, swarm_theta,
                   swarm_active, u, w, Rinner, Router, hrad, htheta, nelx, icon, rad_V, theta_V):
    numpy_out = not dev.is_dev(swarm_x, swarm_z)
    n = int(nparticle)
    ic = dev.to_dev(icon, np.int32)
    x, z = dev.to_dev(swarm_x, np.float64), dev.to_dev(swarm_z, np.float64)
    ra, th = dev.to_dev(swarm_rad, np.float64), dev.to_dev(swarm_theta, np.float64)
    act = dev.to_dev(swarm_active, np.uint8)
    ud, wd = dev.to_dev(u, np.float64), dev.to_dev(w, np.float64)
    rv, tv = dev.to_dev(rad_V, np.float64), dev.to_dev(theta_V, np.float64)
    su, sw = dev.empty(n), dev.empty(n)
    flags = dev.empty(2, np.int32)
    _lib.call("mw_advect_particles_curved", CURVED[geometry], int(RKorder), float(dt), n,
              dev.ptr(x), dev.ptr(z), dev.ptr(ra), dev.ptr(th), dev.ptr(act), dev.ptr(ud),
              dev.ptr(wd), float(Rinner), float(Router), float(hrad), float(htheta), int(nelx),
              ic.shape[1], dev.ptr(ic), dev.ptr(rv), dev.ptr(tv), dev.ptr(su), dev.ptr(sw),
              dev.ptr(flags), dev.stream())
    if dev.host(flags)[0]:
        raise SystemExit(f"advect_particles___{geometry}: particle position outside the mesh")
    if numpy_out:
        for dst, src in ((swarm_x, x), (swarm_z, z), (swarm_rad, ra), (swarm_theta, th)):
            dst[:] = dev.host(src)
        swarm_active[:] = dev.host(act).astype(bool)
        return swarm_x, swarm_z, swarm_rad, swarm_theta, dev.host(su), dev.host(sw), swarm_active
    return x, z, ra, th, su, sw, act


def _curved(geometry):
    def f(RKorder, dt, nparticle, swarm_x, swarm_z, swarm_rad, swarm_theta, swarm_active, u, w,
          Rinner, Router, hrad, htheta, nelx, icon, rad_V, theta_V):
        return _advect_curved(geometry, RKorder, dt, nparticle, swarm_x, swarm_z, swarm_rad,
                              swarm_theta, swarm_active, u, w, Rinner, Router, hrad, htheta, nelx,
                              icon, rad_V, theta_V)
    f.__name__ = f"advect_particles___{geometry}"
    return f


advect_particles___quarter = _curved("quarter")
advect_particles___quarter.__doc__ = "pic_functions.py:400-495 (RK1, RK2)."
advect_particles___half = _curved("half")
advect_particles___half.__doc__ = "pic_functions.py:501-596 (RK1, RK2)."
advect_particles___eighth = _curved("eighth")
advect_particles___eighth.__doc__ = "pic_functions.py:298-343 (RK1)."
advect_particles___annulus = _curved("annulus")
advect_particles___annulus.__doc__ = "pic_functions.py:349-394 (RK1)."


# ---------------------------------------------------------------------------
# least-squares projections (pic_functions.py:1067-1285)
# ---------------------------------------------------------------------------
def _ls(nb, nel, x_e, z_e, swarm_x, swarm_z, swarm_iel, swarm_rho, swarm_eta):
    numpy_out = not dev.is_dev(swarm_x, swarm_z, swarm_rho, swarm_eta)
    n = int(np.shape(swarm_x)[0]) if not dev.is_dev(swarm_x) else int(swarm_x.numel())
    xe, ze = dev.to_dev(x_e, np.float64), dev.to_dev(z_e, np.float64)
    x, z = dev.to_dev(swarm_x, np.float64), dev.to_dev(swarm_z, np.float64)
    iel = dev.to_dev(swarm_iel, np.int32)
    rho, eta = dev.to_dev(swarm_rho, np.float64), dev.to_dev(swarm_eta, np.float64)
    cr, ce, mm = dev.empty(nel * nb), dev.empty(nel * nb), dev.empty(nel * 4)
    nbytes = _lib.lib.mw_ls_workspace_bytes(n, int(nel))
    work = dev.empty((nbytes + 7) // 8)
    _lib.call("mw_ls_coefficients", int(nb), int(nel), dev.ptr(xe), dev.ptr(ze), n, dev.ptr(x),
              dev.ptr(z), dev.ptr(iel), dev.ptr(rho), dev.ptr(eta), dev.ptr(cr), dev.ptr(ce),
              dev.ptr(mm), dev.ptr(work), dev.stream())
    cr = dev.out(cr, numpy_out).reshape(nel, nb)
    ce = dev.out(ce, numpy_out).reshape(nel, nb)
    mm = dev.out(mm, numpy_out).reshape(nel, 4)
    return ([cr[:, k] for k in range(nb)] + [ce[:, k] for k in range(nb)] +
            [mm[:, 0], mm[:, 1], mm[:, 2], mm[:, 3]])


def compute_ls_coefficients_P1(nel, x_e, z_e, swarm_x, swarm_z, swarm_iel, swarm_rho, swarm_eta):
    """pic_functions.py:1097-1178 -> (ls_rho_a, _b, _c, ls_eta_a, _b, _c, rho_min_e,
    rho_max_e, eta_min_e, eta_max_e)"""
    return tuple(_ls(3, nel, x_e, z_e, swarm_x, swarm_z, swarm_iel, swarm_rho, swarm_eta))


def compute_ls_coefficients_Q1(nel, x_e, z_e, swarm_x, swarm_z, swarm_iel, swarm_rho, swarm_eta):
    """pic_functions.py:1188-1285 -> (ls_rho_a, _b, _c, _d, ls_eta_a, _b, _c, _d, rho_min_e,
    rho_max_e, eta_min_e, eta_max_e)"""
    return tuple(_ls(4, nel, x_e, z_e, swarm_x, swarm_z, swarm_iel, swarm_rho, swarm_eta))


def limiter_elements(c_0, c_1, c_2, S, L, h):
    """The driver's per-element limiter loop (meeuuw.py:1243-1249) as one
    kernel: arrays of coefficients and bounds -> limited (c_1, c_2)."""
    numpy_out = not dev.is_dev(c_0, c_1, c_2)
    c0 = dev.to_dev(np.atleast_1d(c_0) if not dev.is_dev(c_0) else c_0, np.float64)
    n = int(c0.numel())
    c1 = dev.to_dev(np.array(np.atleast_1d(c_1), dtype=np.float64) if not dev.is_dev(c_1)
                    else c_1.clone(), np.float64)
    c2 = dev.to_dev(np.array(np.atleast_1d(c_2), dtype=np.float64) if not dev.is_dev(c_2)
                    else c_2.clone(), np.float64)
    S_ = dev.to_dev(np.broadcast_to(np.asarray(S, dtype=np.float64), (n,)).copy()
                    if not dev.is_dev(S) else S, np.float64)
    L_ = dev.to_dev(np.broadcast_to(np.asarray(L, dtype=np.float64), (n,)).copy()
                    if not dev.is_dev(L) else L, np.float64)
    _lib.call("mw_limiter", n, dev.ptr(c0), dev.ptr(c1), dev.ptr(c2), dev.ptr(S_), dev.ptr(L_),
              float(h), dev.stream())
    return dev.out(c1, numpy_out), dev.out(c2, numpy_out)


def limiter(c_0, c_1, c_2, S=0, L=1, h=1):
    """pic_functions.py:1067-1087 (scalar form; limiter_elements is the batched one)."""
    a, b = limiter_elements(c_0, c_1, c_2, S, L, h)
    return float(np.asarray(a).ravel()[0]), float(np.asarray(b).ravel()[0])
