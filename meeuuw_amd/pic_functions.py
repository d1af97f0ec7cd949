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
