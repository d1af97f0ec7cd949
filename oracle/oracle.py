"""CPU FP64 oracle for the meeuuw per-timestep hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker or the timed CPU baseline.  The shipped package never calls it.

Each function restates the reference function of the same name (cited per
function) on top of the C loops in ``mw_oracle.c``; sparse solves use SciPy
SuperLU exactly as the reference does (``coo_matrix(...).tocsr()`` then
``spsolve``, meeuuw.py:1800, 1832, 2233-2235).
"""
import ctypes
import os
import subprocess

import numpy as np
import scipy.sparse as sparse
import scipy.sparse.linalg as sla

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_U8 = ctypes.POINTER(ctypes.c_uint8)

AVERAGING = {"arithmetic": 0, "geometric": 1, "harmonic": 2}
PNORM = {"none": 0, "surface": 1, "volume": 2}


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = ctypes.CDLL(path)
        _LIB.orc_pressure_offset.restype = ctypes.c_double
    return _LIB


def _p(a):
    if a is None:
        return None
    if a.dtype == np.float64:
        return a.ctypes.data_as(_D)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_I)
    if a.dtype in (np.uint8, np.bool_):
        return a.ctypes.data_as(_U8)
    raise TypeError(a.dtype)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _u8(a):
    return np.ascontiguousarray(a, dtype=bool).view(np.uint8)


# ---------------------------------------------------------------------------
# basis functions (basis_functions.py:20-138)
# ---------------------------------------------------------------------------
def basis_functions_V(r, s):
    out = np.zeros(9)
    lib().orc_basis_V(ctypes.c_double(r), ctypes.c_double(s), _p(out))
    return out


def basis_functions_V_dr(r, s):
    out = np.zeros(9)
    lib().orc_basis_V_dr(ctypes.c_double(r), ctypes.c_double(s), _p(out))
    return out


def basis_functions_V_dt(r, s):
    out = np.zeros(9)
    lib().orc_basis_V_dt(ctypes.c_double(r), ctypes.c_double(s), _p(out))
    return out


def basis_functions_P(r, s):
    out = np.zeros(4)
    lib().orc_basis_P(ctypes.c_double(r), ctypes.c_double(s), _p(out))
    return out


R_V = np.array([-1, 1, 1, -1, 0, 1, 0, -1, 0], dtype=np.float64)  # basis_functions.py:16
T_V = np.array([-1, -1, 1, 1, -1, 0, 1, 0, 0], dtype=np.float64)  # basis_functions.py:17


def quadrature_setup(nq_per_dim=3):
    """quadrature_setup.py:12-33 (3-point rule only)."""
    assert nq_per_dim == 3
    qc = np.array([-np.sqrt(3.0 / 5.0), 0.0, np.sqrt(3.0 / 5.0)])
    qw = np.array([5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0])
    return qc, qw


# ---------------------------------------------------------------------------
# mesh and setup
# ---------------------------------------------------------------------------
class Mesh:
    """Box mesh + reference-element tables + per-element Jacobians.

    Restates build_velocity_mesh.py:110-146, build_velocity_connectivity.py:61-94,
    build_pressure_connectivity.py:15-25, basis_functions_setup_q.py:47-79,
    basis_functions_setup_Vnodes.py, meeuuw.py:340-411, 874-881.
    """

    def __init__(self, nelx, nelz, Lx=1.0, Lz=1.0):
        L = lib()
        self.nelx, self.nelz, self.Lx, self.Lz = nelx, nelz, float(Lx), float(Lz)
        self.nel = nel = nelx * nelz
        self.hx, self.hz = self.Lx / nelx, self.Lz / nelz
        self.nnx, self.nnz = 2 * nelx + 1, 2 * nelz + 1
        self.nn_V = self.nnx * self.nnz
        self.nn_P = (nelx + 1) * (nelz + 1)
        self.Nfem_V, self.Nfem_P = 2 * self.nn_V, self.nn_P
        self.Nfem = self.Nfem_V + self.Nfem_P
        self.L_ref = (self.Lx + self.Lz) / 2
        self.volume = self.Lx * self.Lz
        self.x_V = np.zeros(self.nn_V)
        self.z_V = np.zeros(self.nn_V)
        self.icon_V = np.zeros((9, nel), dtype=np.int32)
        self.icon_P = np.zeros((4, nel), dtype=np.int32)
        self.x_P = np.zeros(self.nn_P)
        self.z_P = np.zeros(self.nn_P)
        L.orc_mesh_box(nelx, nelz, ctypes.c_double(self.Lx), ctypes.c_double(self.Lz),
                       _p(self.x_V), _p(self.z_V), _p(self.icon_V), _p(self.icon_P),
                       _p(self.x_P), _p(self.z_P))
        self.qcoords, self.qweights = quadrature_setup(3)
        self.rq, self.tq, self.weightq = np.zeros(9), np.zeros(9), np.zeros(9)
        self.N_V, self.dNdr_V, self.dNdt_V = np.zeros((9, 9)), np.zeros((9, 9)), np.zeros((9, 9))
        self.N_P, self.dNdr_P, self.dNdt_P = np.zeros((9, 4)), np.zeros((9, 4)), np.zeros((9, 4))
        L.orc_qpoint_tables(_p(self.qcoords), _p(self.qweights), 3, _p(self.rq), _p(self.tq),
                            _p(self.weightq), _p(self.N_V), _p(self.dNdr_V), _p(self.dNdt_V),
                            _p(self.N_P), _p(self.dNdr_P), _p(self.dNdt_P))
        self.JxWq = np.zeros((nel, 9))
        self.jcbi00q, self.jcbi01q = np.zeros((nel, 9)), np.zeros((nel, 9))
        self.jcbi10q, self.jcbi11q = np.zeros((nel, 9)), np.zeros((nel, 9))
        L.orc_jacobians(nel, 9, _p(self.dNdr_V), _p(self.dNdt_V), _p(self.weightq),
                        _p(self.x_V), _p(self.z_V), _p(self.icon_V), _p(self.JxWq),
                        _p(self.jcbi00q), _p(self.jcbi01q), _p(self.jcbi10q), _p(self.jcbi11q))
        # tables at the 9 V nodes (basis_functions_setup_Vnodes.py)
        self.N_P_n = np.array([basis_functions_P(R_V[i], T_V[i]) for i in range(9)])
        self.dNdr_V_n = np.array([basis_functions_V_dr(R_V[i], T_V[i]) for i in range(9)])
        self.dNdt_V_n = np.array([basis_functions_V_dt(R_V[i], T_V[i]) for i in range(9)])
        self.jcbi00n, self.jcbi01n = np.zeros((nel, 9)), np.zeros((nel, 9))
        self.jcbi10n, self.jcbi11n = np.zeros((nel, 9)), np.zeros((nel, 9))
        L.orc_jacobians(nel, 9, _p(self.dNdr_V_n), _p(self.dNdt_V_n), None,
                        _p(self.x_V), _p(self.z_V), _p(self.icon_V), None,
                        _p(self.jcbi00n), _p(self.jcbi01n), _p(self.jcbi10n), _p(self.jcbi11n))
        self.xq = project_nodal_Vfield_onto_qpoints(self.x_V, 9, nel, 9, self.N_V, self.icon_V)
        self.zq = project_nodal_Vfield_onto_qpoints(self.z_V, 9, nel, 9, self.N_V, self.icon_V)
        self.local_to_globalV = np.zeros((18, nel), dtype=np.int32)
        L.orc_local_to_global_V(nel, _p(self.icon_V), _p(self.local_to_globalV))
        ie = np.arange(nel)
        self.left_element = (ie % nelx) == 0
        self.right_element = (ie % nelx) == nelx - 1
        self.bot_element = (ie // nelx) == 0
        self.top_element = (ie // nelx) == nelz - 1
        eps = 1e-8  # constants.py eps / experiment0.py:130
        self.top_Vnodes = self.z_V / self.Lz > 1 - eps
        self.bot_Vnodes = self.z_V / self.Lz < eps
        self.bignb_Stokes = nel * (18 * 18 + 2 * 18 * 4)
        self.bignb_K = nel * 18 * 18
        self.bignb_G = nel * 18 * 4
        self.bignb_P = nel * 16
        self.bignb_T = nel * 81

    def bc_free_slip(self):
        """Free slip on all sides (experiment0.py:128-144)."""
        eps = 1e-8
        fix = np.zeros(self.Nfem_V, dtype=bool)
        val = np.zeros(self.Nfem_V)
        fix[0::2] |= self.x_V / self.Lx < eps
        fix[0::2] |= self.x_V / self.Lx > 1 - eps
        fix[1::2] |= self.z_V / self.Lz < eps
        fix[1::2] |= self.z_V / self.Lz > 1 - eps
        return fix, val

    def bc_no_slip(self):
        """No slip on all sides (experiment19.py)."""
        eps = 1e-8
        hull = ((self.x_V / self.Lx < eps) | (self.x_V / self.Lx > 1 - eps) |
                (self.z_V / self.Lz < eps) | (self.z_V / self.Lz > 1 - eps))
        fix = np.repeat(hull, 2)
        return fix, np.zeros(self.Nfem_V)

    def bc_T(self, Tbottom=1.0, Ttop=0.0):
        """experiment0.py:170-176."""
        eps = 1e-8
        fix = np.zeros(self.nn_V, dtype=bool)
        val = np.zeros(self.nn_V)
        b = self.z_V < eps
        t = self.z_V > self.Lz - eps
        fix[b] = True
        val[b] = Tbottom
        fix[t] = True
        val[t] = Ttop
        return fix, val


# ---------------------------------------------------------------------------
# COO index arrays (meeuuw.py:874-1015)
# ---------------------------------------------------------------------------
def coo_indices(mesh):
    L = lib()
    nel = mesh.nel
    l2g = mesh.local_to_globalV
    out = {}
    II = np.zeros(mesh.bignb_Stokes, dtype=np.int32)
    JJ = np.zeros_like(II)
    L.orc_coo_stokes(nel, _p(l2g), _p(mesh.icon_P), mesh.Nfem_V, _p(II), _p(JJ))
    out["Stokes"] = (II, JJ)
    II = np.zeros(mesh.bignb_K, dtype=np.int32)
    JJ = np.zeros_like(II)
    L.orc_coo_K(nel, _p(l2g), _p(II), _p(JJ))
    out["K"] = (II, JJ)
    II = np.zeros(mesh.bignb_G, dtype=np.int32)
    JJ, IIT, JJT = np.zeros_like(II), np.zeros_like(II), np.zeros_like(II)
    L.orc_coo_G(nel, _p(l2g), _p(mesh.icon_P), _p(II), _p(JJ), _p(IIT), _p(JJT))
    out["G"] = (II, JJ)
    out["GT"] = (IIT, JJT)
    II = np.zeros(mesh.bignb_P, dtype=np.int32)
    JJ = np.zeros_like(II)
    L.orc_coo_scalar(nel, 4, _p(mesh.icon_P), _p(II), _p(JJ))
    out["MP"] = (II, JJ)
    II = np.zeros(mesh.bignb_T, dtype=np.int32)
    JJ = np.zeros_like(II)
    L.orc_coo_scalar(nel, 9, _p(mesh.icon_V), _p(II), _p(JJ))
    out["T"] = (II, JJ)
    return out


# ---------------------------------------------------------------------------
# assembly
# ---------------------------------------------------------------------------
def build_matrix_stokes(bignb_Stokes, bignb_K, bignb_M, bignb_G, nel, nq_per_element, m_V, m_P,
                        ndof_V, Nfem_V, Nfem_P, ndof_V_el, icon_V, icon_P, rhoq, etaq, JxWq,
                        local_to_globalV, gxq, gyq, N_V, N_P, dNdr_V, dNdt_V, dNdr_P, dNdt_P,
                        jcbi00q, jcbi01q, jcbi10q, jcbi11q, eta_e, eta_ref, L_ref, bc_fix_V,
                        bc_val_V, bot_element, top_element, bot_free_slip, top_free_slip,
                        geometry, theta_V, axisymmetric, xq, blocks):
    """build_matrix_stokes.py:18-278, plane strain and axisymmetric, including the
    free-slip rotation on curved boundaries (163-203)."""
    axi = bool(axisymmetric)
    curved = geometry in ("quarter", "half", "eighth", "annulus")
    rot = curved and bool(bot_free_slip or top_free_slip)
    rhs_f = np.zeros(Nfem_V)
    rhs_h = np.zeros(Nfem_P)
    if blocks:
        VV_S = np.zeros(0)
        VV_K, VV_G, VV_GT, VV_H = np.zeros(bignb_K), np.zeros(bignb_G), np.zeros(bignb_G), np.zeros(bignb_G)
        VV_M, VV_Me = np.zeros(bignb_M), np.zeros(bignb_M)
    else:
        VV_S = np.zeros(bignb_Stokes)
        VV_K = VV_G = VV_GT = VV_H = VV_M = VV_Me = np.zeros(0)
    a = [_f64(v) for v in (rhoq, etaq, JxWq, gxq, gyq, N_V, N_P, dNdr_V, dNdt_V, dNdr_P, dNdt_P,
                           jcbi00q, jcbi01q, jcbi10q, jcbi11q, eta_e, bc_val_V)]
    lib().orc_build_matrix_stokes(
        nel, _p(_i32(icon_V)), _p(_i32(icon_P)), _p(a[0]), _p(a[1]), _p(a[2]),
        _p(_i32(local_to_globalV)), _p(a[3]), _p(a[4]), _p(a[5]), _p(a[6]), _p(a[7]), _p(a[8]),
        _p(a[9]), _p(a[10]), _p(a[11]), _p(a[12]), _p(a[13]), _p(a[14]), _p(a[15]),
        ctypes.c_double(eta_ref), ctypes.c_double(L_ref), _p(_u8(bc_fix_V)), _p(a[16]),
        Nfem_V, Nfem_P, int(bool(blocks)), _p(VV_S), _p(rhs_f), _p(rhs_h), _p(VV_K), _p(VV_G),
        _p(VV_GT), _p(VV_M), _p(VV_Me), _p(VV_H), int(rot),
        _p(_f64(theta_V)) if rot else None, _p(_u8(bot_element)) if rot else None,
        _p(_u8(top_element)) if rot else None, int(bool(bot_free_slip)),
        int(bool(top_free_slip)), int(axi), _p(_f64(xq)) if axi else None)
    return VV_S, rhs_f, rhs_h, VV_K, VV_G, VV_GT, VV_M, VV_Me, VV_H


def build_matrix_energy(bignb, nel, nq_per_element, m_T, Nfem_T, T, icon_T, rhoq, etaq, Tq, uq, wq,
                        hcondq, hcapaq, alphaq, hprodq, exxq, ezzq, exzq, dpdxq, dpdzq, JxWq, N_T,
                        dNdr_T, dNds_T, jcbi00q, jcbi01q, jcbi10q, jcbi11q, bc_fix_T, bc_val_T, dt,
                        formulation, rho0):
    """build_matrix_energy.py:11-141."""
    VV = np.zeros(bignb)
    rhs = np.zeros(Nfem_T)
    a = [_f64(v) for v in (T, etaq, Tq, uq, wq, hcondq, hcapaq, alphaq, hprodq, exxq, ezzq, exzq,
                           JxWq, N_T, dNdr_T, dNds_T, jcbi00q, jcbi01q, jcbi10q, jcbi11q, bc_val_T)]
    lib().orc_build_matrix_energy(
        nel, Nfem_T, _p(a[0]), _p(_i32(icon_T)), *[_p(v) for v in a[1:20]], _p(_u8(bc_fix_T)),
        _p(a[20]), ctypes.c_double(dt), int(formulation == "EBA"), ctypes.c_double(rho0),
        _p(VV), _p(rhs))
    return VV, rhs


# ---------------------------------------------------------------------------
# particles
# ---------------------------------------------------------------------------
def locate_particles___box(nparticle, swarm_active, swarm_x, swarm_z, hx, hz, x_V, z_V, icon, nelx):
    """pic_functions.py:107-137."""
    icon = _i32(icon)
    nel = icon.shape[1]
    r, t = np.zeros(nparticle), np.zeros(nparticle)
    iel = np.zeros(nparticle, dtype=np.int32)
    bad = lib().orc_locate_box(nparticle, _p(_u8(swarm_active)), _p(_f64(swarm_x)),
                               _p(_f64(swarm_z)), ctypes.c_double(hx), ctypes.c_double(hz),
                               _p(_f64(x_V)), _p(_f64(z_V)), _p(icon), nel, nelx, _p(r), _p(t),
                               _p(iel))
    if bad:
        raise SystemExit("r value out of bounds")
    return r, t, iel


def locate_particles(geometry, nparticle, swarm_active, swarm_x, swarm_z, swarm_rad, swarm_theta,
                     hx, hz, hrad, htheta, x_V, z_V, rad_V, theta_V, icon_V, nelx, Rinner):
    """locate_particles.py:12-66 (box branch)."""
    if geometry != "box":
        raise SystemExit("locate particles not implemented for this geometry")
    r, t, iel = locate_particles___box(nparticle, swarm_active, swarm_x, swarm_z, hx, hz, x_V, z_V,
                                       icon_V, nelx)
    if np.min(r) < -1 or np.max(r) > 1:
        raise SystemExit("r value out of bounds")
    if np.min(t) < -1 or np.max(t) > 1:
        raise SystemExit("t value out of bounds")
    return r, t, iel


def interpolate_field_on_particles(nparticle, swarm_active, swarm_r, swarm_t, swarm_iel, phi, icon):
    """pic_functions.py:84-101."""
    icon = _i32(icon)
    out = np.zeros(nparticle)
    lib().orc_interpolate_field(nparticle, _p(_u8(swarm_active)), _p(_f64(swarm_r)),
                                _p(_f64(swarm_t)), _p(_i32(swarm_iel)), _p(_f64(phi)), _p(icon),
                                icon.shape[1], _p(out))
    return out


def advect_particles___box(RKorder, dt, nparticle, swarm_x, swarm_z, swarm_active, u, w, Lx, Lz,
                           hx, hz, nelx, icon, x_V, z_V):
    """pic_functions.py:179-292.  Mutates swarm_x, swarm_z, swarm_active in place."""
    assert swarm_x.dtype == np.float64 and swarm_z.dtype == np.float64
    assert swarm_active.dtype == np.bool_
    icon = _i32(icon)
    su, sw = np.zeros(nparticle), np.zeros(nparticle)
    bad = lib().orc_advect_box(int(RKorder), ctypes.c_double(dt), nparticle, _p(swarm_x),
                               _p(swarm_z), swarm_active.ctypes.data_as(_U8), _p(_f64(u)),
                               _p(_f64(w)), ctypes.c_double(Lx), ctypes.c_double(Lz),
                               ctypes.c_double(hx), ctypes.c_double(hz), nelx, icon.shape[1],
                               _p(icon), _p(_f64(x_V)), _p(_f64(z_V)), _p(su), _p(sw))
    if bad:
        raise SystemExit("particle position outside the mesh")
    return swarm_x, swarm_z, su, sw, swarm_active


def project_particles_on_elements(nel, nparticle, swarm_active, swarm_rho, swarm_eta, swarm_iel,
                                  averaging):
    """pic_functions.py:602-650."""
    rho_e, eta_e = np.zeros(nel), np.zeros(nel)
    cnt = np.zeros(nel, dtype=np.int32)
    with np.errstate(all="ignore"):
        lib().orc_project_on_elements(nel, nparticle, _p(_u8(swarm_active)), _p(_f64(swarm_rho)),
                                      _p(_f64(swarm_eta)), _p(_i32(swarm_iel)),
                                      AVERAGING[averaging], _p(rho_e), _p(eta_e), _p(cnt))
    return rho_e, eta_e, cnt


def _project_nodes(ptype, nel, nn_V, nparticle, swarm_active, swarm_phi, icon, swarm_iel, swarm_r,
                   swarm_t, averaging):
    out = np.zeros(nn_V)
    lib().orc_project_on_nodes(ptype, nel, nn_V, nparticle, _p(_u8(swarm_active)),
                               _p(_f64(swarm_phi)), _p(_i32(icon)), _p(_i32(swarm_iel)),
                               _p(_f64(swarm_r)), _p(_f64(swarm_t)), AVERAGING[averaging], _p(out))
    return out


def project_particle_field_on_nodes_1(*a):
    """pic_functions.py:658-724."""
    return _project_nodes(1, *a)


def project_particle_field_on_nodes_2(*a):
    """pic_functions.py:732-811."""
    return _project_nodes(2, *a)


def project_particle_field_on_nodes_3(*a):
    """pic_functions.py:820-925."""
    return _project_nodes(3, *a)


def project_particle_field_on_nodes_4(*a):
    """pic_functions.py:936-1053."""
    return _project_nodes(4, *a)


# ---------------------------------------------------------------------------
# nodal -> qpoints (project_nodal_field_onto_qpoints.py:11-62)
# ---------------------------------------------------------------------------
def _nodal_to_q(phi, nel, m, N, icon):
    out = np.zeros((nel, 9))
    N = _f64(N)
    lib().orc_nodal_to_qpoints(_p(_f64(phi)), nel, m, _p(N), _p(_i32(icon)), _p(out))
    return out


def project_nodal_Pfield_onto_qpoints(phi_nodal, nq_per_element, nel, m_P, N_P, icon_P):
    return _nodal_to_q(phi_nodal, nel, m_P, _f64(N_P)[:, :m_P], _i32(icon_P))


def project_nodal_Vfield_onto_qpoints(phi_nodal, nq_per_element, nel, m_V, N_V, icon_V):
    return _nodal_to_q(phi_nodal, nel, m_V, N_V, icon_V)


project_nodal_Tfield_onto_qpoints = project_nodal_Vfield_onto_qpoints


# ---------------------------------------------------------------------------
# reductions
# ---------------------------------------------------------------------------
def compute_pressure_offset(geometry, pressure_normalisation, axisymmetric, top_element, nel, nelx,
                            nq_per_element, N_P, JxWq, p, icon_P, theta_P, xq, volume):
    """compute_pressure_offset.py:11-66."""
    if pressure_normalisation not in PNORM:
        raise ValueError("compute_pressure_offset: unknown pressure_normalisation value")
    axi = bool(axisymmetric)
    return lib().orc_pressure_offset(PNORM[pressure_normalisation], _p(_u8(top_element)), nel, nelx,
                                     _p(_f64(N_P)), _p(_f64(JxWq)), _p(_f64(p)), _p(_i32(icon_P)),
                                     ctypes.c_double(volume), int(axi),
                                     _p(_f64(theta_P)) if axi and theta_P is not None else None,
                                     _p(_f64(xq)) if axi and xq is not None else None)


def compute_nodal_pressure(m_V, nn_V, icon_V, icon_P, p, N_P_n):
    """compute_nodal_pressure.py:7-21."""
    icon_V = _i32(icon_V)
    q = np.zeros(nn_V)
    lib().orc_nodal_pressure(icon_V.shape[1], nn_V, _p(icon_V), _p(_i32(icon_P)), _p(_f64(p)),
                             _p(_f64(N_P_n)), _p(q))
    return q


def compute_nodal_heat_flux(icon_T, T, hcond_nodal, nn_T, m_T, nel, dNdr_T_n, dNds_T_n, jcbi00_T,
                            jcbi01_T, jcbi10_T, jcbi11_T):
    """compute_nodal_heat_flux.py:10-40."""
    o = [np.zeros(nn_T) for _ in range(4)]
    lib().orc_nodal_heat_flux(_p(_i32(icon_T)), _p(_f64(T)), _p(_f64(hcond_nodal)), nn_T, nel,
                              _p(_f64(dNdr_T_n)), _p(_f64(dNds_T_n)), _p(_f64(jcbi00_T)),
                              _p(_f64(jcbi01_T)), _p(_f64(jcbi10_T)), _p(_f64(jcbi11_T)),
                              *[_p(v) for v in o])
    return tuple(o)


def compute_global_quantities(nel, nq_per_element, xq, zq, uq, wq, Tq, rhoq, hcapaq, alphaq, etaq,
                              exxq, ezzq, exzq, dxxq, dzzq, dxzq, volume, JxWq, gxq, gzq):
    """postprocessors.py:13-83 -> (vrms, EK, WAG, TVD, GPE, ITE, TM, Tavrg, eta_avrg)."""
    out = np.zeros(9)
    lib().orc_global_quantities(nel, *[_p(_f64(v)) for v in (uq, wq, Tq, rhoq, hcapaq, etaq, exxq,
                                                             ezzq, exzq)],
                                ctypes.c_double(volume), _p(_f64(JxWq)), _p(_f64(gxq)),
                                _p(_f64(gzq)), _p(out))
    return tuple(out)


def compute_Nu(Lx, Lz, x_V, z_V, nel, left_element, right_element, top_element, bottom_element,
               icon_V, T, dTdx_nodal, dTdz_nodal, nq_per_dim, qcoords, qweights, hcond_n):
    """postprocessors.py:152-248."""
    out = np.zeros(13)
    lib().orc_compute_Nu(ctypes.c_double(Lx), ctypes.c_double(Lz), _p(_f64(x_V)), _p(_f64(z_V)),
                         nel, _p(_u8(left_element)), _p(_u8(right_element)), _p(_u8(top_element)),
                         _p(_u8(bottom_element)), _p(_i32(icon_V)), _p(_f64(T)),
                         _p(_f64(dTdx_nodal)), _p(_f64(dTdz_nodal)), _p(_f64(qcoords)),
                         _p(_f64(qweights)), _p(_f64(hcond_n)), _p(out))
    return tuple(out)


# ---------------------------------------------------------------------------
# solvers
# ---------------------------------------------------------------------------
def stokes_direct(mesh, coo, VV_Stokes, rhs_f, rhs_h, eta_ref=1.0):
    """COO->CSR + spsolve + split (meeuuw.py:1800, 1829-1851)."""
    II, JJ = coo["Stokes"]
    A = sparse.coo_matrix((VV_Stokes, (II, JJ)), shape=(mesh.Nfem, mesh.Nfem)).tocsr()
    b = np.concatenate([rhs_f, rhs_h])
    sol = sla.spsolve(A, b)
    u, w = np.reshape(sol[:mesh.Nfem_V], (mesh.nn_V, 2)).T
    p = sol[mesh.Nfem_V:] * (eta_ref / mesh.L_ref)
    return np.ascontiguousarray(u), np.ascontiguousarray(w), p


def energy_direct(mesh, coo, VV_T, rhs):
    """meeuuw.py:2233-2235."""
    II, JJ = coo["T"]
    A = sparse.coo_matrix((VV_T, (II, JJ)), shape=(mesh.nn_V, mesh.nn_V)).tocsr()
    return sla.spsolve(A, rhs)


def uzawa3_solver_L2(K_mat, G_mat, GT_mat, MP_mat, H_mat, f_rhs, h_rhs, Nfem_P):
    """uzawa3_solver_L2.py:21-99 (innerK = innerMp = 'splu', tol 1e-6, 100 iterations)."""
    tol, niter = 1e-6, 100
    xiV, xiP, alph = np.zeros(niter), np.zeros(niter), np.zeros(niter)
    solP = np.zeros(Nfem_P)
    LU = sla.splu(sparse.csc_matrix(K_mat))
    solV = LU.solve(f_rhs)
    LU_Mp = sla.splu(sparse.csc_matrix(MP_mat))
    r = LU_Mp.solve(GT_mat.dot(solV) - h_rhs)
    p = np.copy(r)
    for k in range(niter):
        d = LU.solve(G_mat.dot(p))
        alpha = r.dot(MP_mat.dot(r)) / p.dot(H_mat.dot(d))
        alph[k] = alpha
        solP += alpha * p
        solV -= alpha * d
        r1 = r + LU_Mp.solve(-alpha * GT_mat.dot(d))
        beta = r1.dot(MP_mat.dot(r1)) / r.dot(MP_mat.dot(r))
        p1 = r1 + beta * p
        xiP[k] = np.linalg.norm(alpha * p)
        xiV[k] = np.linalg.norm(alpha * d)
        if xiP[k] < tol and xiV[k] < tol:
            break
        r, p = np.copy(r1), np.copy(p1)
    return solV, solP, k + 1, xiV, xiP, alph


# ---------------------------------------------------------------------------
# seeded placement (swarm_coordinates_setup.py:71-86 with a seeded generator)
# ---------------------------------------------------------------------------
def random_reduced_coords(nel, ppe, seed):
    """(r, t) ~ U(-1, 1), element-major, r drawn before t, from numpy PCG64(seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    rt = rng.uniform(-1.0, 1.0, size=(nel, ppe, 2))
    return np.ascontiguousarray(rt[:, :, 0].ravel()), np.ascontiguousarray(rt[:, :, 1].ravel())


def place_particles(mesh, r, t, ppe):
    n = mesh.nel * ppe
    x, z = np.zeros(n), np.zeros(n)
    iel = np.zeros(n, dtype=np.int32)
    lib().orc_place_particles(mesh.nel, ppe, _p(_f64(r)), _p(_f64(t)), _p(mesh.x_V), _p(mesh.z_V),
                              _p(mesh.icon_V), _p(x), _p(z), _p(iel))
    return x, z, iel


def cd_u_eta(K, G, GT, M, M_eta, H, f, h, Nfem_P):
    """cd_u_eta.py:13-99 (innerK = innerM = 'splu'; initial w with spsolve(M_eta, ...))."""
    tol, niter = 1e-6, 100
    xiV, xiP, al, be = np.zeros(niter), np.zeros(niter), np.zeros(niter), np.zeros(niter)
    solP = np.zeros(Nfem_P)
    LU_K = sla.splu(sparse.csc_matrix(K))
    solV = LU_K.solve(f)
    LU_M = sla.splu(sparse.csc_matrix(M))
    LU_Me = sla.splu(sparse.csc_matrix(M_eta))
    q = LU_M.solve(-GT @ solV + M @ h)
    w = sla.spsolve(sparse.csc_matrix(M_eta), -GT @ solV + M @ h)
    d = -np.copy(w)
    for k in range(niter):
        z = LU_K.solve(G @ d)
        alpha = q.dot(M.dot(w)) / d.dot(H.dot(z))
        al[k] = alpha
        solP += alpha * d
        solV -= alpha * z
        q1 = LU_M.solve(-GT @ solV + M @ h)
        w1 = LU_Me.solve(M @ q1)
        beta = q1.dot(M.dot(w1)) / q.dot(M.dot(w))
        be[k] = beta
        d1 = -w1 + beta * d
        xiP[k] = np.linalg.norm(alpha * d)
        xiV[k] = np.linalg.norm(alpha * z)
        if xiP[k] < tol and xiV[k] < tol:
            break
        w, d, q = np.copy(w1), np.copy(d1), np.copy(q1)
    return solV, solP, k + 1, xiV, xiP, al, be


# ---------------------------------------------------------------------------
# curved geometries (quarter, half, eighth; annulus-formula functions)
# ---------------------------------------------------------------------------
CURVED = {"quarter": 0, "half": 1, "eighth": 2, "annulus": 3}
OPENING = {"quarter": np.pi / 2, "half": np.pi, "eighth": np.pi / 4}


def curved_mesh(geometry, nelx, nelz, Rinner, Router):
    """Q2 mesh of a quarter/half/eighth annulus: the unit-square box mesh mapped
    to polar coordinates (build_velocity_mesh.py:110-145, 197-203; meeuuw.py:355-378)."""
    m = Mesh(nelx, nelz, 1.0, 1.0)
    opening = OPENING[geometry]
    rad_V = Rinner + m.z_V * (Router - Rinner)
    theta_V = np.pi / 2 - m.x_V * opening
    return dict(mesh=m, rad_V=rad_V, theta_V=theta_V, x_V=rad_V * np.cos(theta_V),
                z_V=rad_V * np.sin(theta_V), hrad=(Router - Rinner) / nelz,
                htheta=opening / nelx, icon_V=m.icon_V, nelx=nelx, nel=m.nel,
                Rinner=float(Rinner), Router=float(Router))


def locate_particles___annulus(nparticle, swarm_active, swarm_rad, swarm_theta, hrad, htheta,
                               rad_V, theta_V, icon, nelx, Rinner):
    """pic_functions.py:142-173 -> (swarm_r, swarm_t, swarm_iel, n_outside)"""
    n = int(nparticle)
    r, t, iel = np.zeros(n), np.zeros(n), np.zeros(n, dtype=np.int32)
    icon = _i32(icon)
    bad = lib().orc_locate_annulus(n, _p(_u8(swarm_active)), _p(_f64(swarm_rad)),
                                   _p(_f64(swarm_theta)), ctypes.c_double(hrad),
                                   ctypes.c_double(htheta), _p(_f64(rad_V)), _p(_f64(theta_V)),
                                   _p(icon), icon.shape[1], int(nelx), ctypes.c_double(Rinner),
                                   _p(r), _p(t), _p(iel))
    return r, t, iel, bad


def interpolate_vel_on_pts___box(xp, zp, u, w, hx, hz, nelx, icon, x_V, z_V):
    """pic_functions.py:17-37 at many points -> (up, wp, iel)"""
    xp, zp = _f64(np.atleast_1d(xp)), _f64(np.atleast_1d(zp))
    n = xp.size
    up, wp, iel = np.zeros(n), np.zeros(n), np.zeros(n, dtype=np.int32)
    icon = _i32(icon)
    lib().orc_vel_on_pts_box(n, _p(xp), _p(zp), _p(_f64(u)), _p(_f64(w)), ctypes.c_double(hx),
                             ctypes.c_double(hz), int(nelx), icon.shape[1], _p(icon),
                             _p(_f64(x_V)), _p(_f64(z_V)), _p(up), _p(wp), _p(iel))
    return up, wp, iel


def interpolate_vel_on_pt___annulus(xp, yp, u, w, hrad, htheta, nelx, icon, rad_V, theta_V,
                                    Rinner):
    """pic_functions.py:40-63 -> (up, wp, iel)"""
    icon = _i32(icon)
    up, wp, iel = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
    lib().orc_vel_on_pt_annulus(ctypes.c_double(xp), ctypes.c_double(yp), _p(_f64(u)),
                                _p(_f64(w)), ctypes.c_double(hrad), ctypes.c_double(htheta),
                                int(nelx), icon.shape[1], _p(icon), _p(_f64(rad_V)),
                                _p(_f64(theta_V)), ctypes.c_double(Rinner), ctypes.byref(up),
                                ctypes.byref(wp), ctypes.byref(iel))
    return up.value, wp.value, iel.value


def advect_particles_curved(geometry, RKorder, dt, nparticle, swarm_x, swarm_z, swarm_rad,
                            swarm_theta, swarm_active, u, w, Rinner, Router, hrad, htheta, nelx,
                            icon, rad_V, theta_V):
    """advect_particles___{quarter,half,eighth,annulus} (pic_functions.py:298-596):
    mutates x, z, rad, theta, active in place -> (x, z, rad, theta, u, w, active)"""
    n = int(nparticle)
    su, sw = np.zeros(n), np.zeros(n)
    act = _u8(swarm_active)
    icon = _i32(icon)
    lib().orc_advect_curved(CURVED[geometry], int(RKorder), ctypes.c_double(dt), n,
                            _p(swarm_x), _p(swarm_z), _p(swarm_rad), _p(swarm_theta), _p(act),
                            _p(_f64(u)), _p(_f64(w)), ctypes.c_double(Rinner),
                            ctypes.c_double(Router), ctypes.c_double(hrad),
                            ctypes.c_double(htheta), int(nelx), icon.shape[1], _p(icon),
                            _p(_f64(rad_V)), _p(_f64(theta_V)), _p(su), _p(sw))
    swarm_active[:] = act.astype(bool)
    return swarm_x, swarm_z, swarm_rad, swarm_theta, su, sw, swarm_active


# ---------------------------------------------------------------------------
# least-squares projections (pic_functions.py:1067-1285)
# ---------------------------------------------------------------------------
def compute_ls_coefficients(nb, nel, x_e, z_e, swarm_x, swarm_z, swarm_iel, swarm_rho, swarm_eta):
    """nb = 3: compute_ls_coefficients_P1; nb = 4: _Q1.  -> (coef_rho (nel, nb),
    coef_eta (nel, nb), rho_min, rho_max, eta_min, eta_max)"""
    cr, ce, mm = np.zeros((nel, nb)), np.zeros((nel, nb)), np.zeros((nel, 4))
    x = _f64(swarm_x)
    lib().orc_ls_coefficients(int(nb), int(nel), _p(_f64(x_e)), _p(_f64(z_e)), x.size, _p(x),
                              _p(_f64(swarm_z)), _p(_i32(swarm_iel)), _p(_f64(swarm_rho)),
                              _p(_f64(swarm_eta)), _p(cr), _p(ce), _p(mm))
    return cr, ce, mm[:, 0], mm[:, 1], mm[:, 2], mm[:, 3]


def compute_ls_coefficients_P1(nel, x_e, z_e, swarm_x, swarm_z, swarm_iel, swarm_rho, swarm_eta):
    """pic_functions.py:1097-1178: the 10-tuple of the reference"""
    cr, ce, rmin, rmax, emin, emax = compute_ls_coefficients(3, nel, x_e, z_e, swarm_x, swarm_z,
                                                             swarm_iel, swarm_rho, swarm_eta)
    return (cr[:, 0], cr[:, 1], cr[:, 2], ce[:, 0], ce[:, 1], ce[:, 2], rmin, rmax, emin, emax)


def compute_ls_coefficients_Q1(nel, x_e, z_e, swarm_x, swarm_z, swarm_iel, swarm_rho, swarm_eta):
    """pic_functions.py:1188-1285: the 12-tuple of the reference"""
    cr, ce, rmin, rmax, emin, emax = compute_ls_coefficients(4, nel, x_e, z_e, swarm_x, swarm_z,
                                                             swarm_iel, swarm_rho, swarm_eta)
    return (cr[:, 0], cr[:, 1], cr[:, 2], cr[:, 3], ce[:, 0], ce[:, 1], ce[:, 2], ce[:, 3],
            rmin, rmax, emin, emax)


def limiter(c_0, c_1, c_2, S=0, L=1, h=1):
    """pic_functions.py:1067-1087 (scalar)"""
    c0 = np.array([c_0], dtype=np.float64)
    c1 = np.array([c_1], dtype=np.float64)
    c2 = np.array([c_2], dtype=np.float64)
    lib().orc_limiter(1, _p(c0), _p(c1), _p(c2), _p(np.array([S], dtype=np.float64)),
                      _p(np.array([L], dtype=np.float64)), ctypes.c_double(h))
    return c1[0], c2[0]


# ---------------------------------------------------------------------------
# derived fields (compute_strain_rate.py, compute_nodal_pressure_gradient.py)
# ---------------------------------------------------------------------------
def _nodal_gradients(mode, icon, u, v, nn, nel, dNdr_n, dNds_n, j00, j01, j10, j11):
    icon = _i32(icon)
    o = [np.zeros(nn) for _ in range(3 if mode == 0 else 2)]
    lib().orc_nodal_gradients(mode, _p(icon), _p(_f64(u)), _p(_f64(v)) if v is not None else None,
                              int(nn), int(nel), _p(_f64(dNdr_n)), _p(_f64(dNds_n)), _p(_f64(j00)),
                              _p(_f64(j01)), _p(_f64(j10)), _p(_f64(j11)), _p(o[0]), _p(o[1]),
                              _p(o[2]) if mode == 0 else None)
    return tuple(o)


def compute_nodal_strain_rate(icon_V, u, v, nn_V, m_V, nel, dNdr_V_n, dNds_V_n, jcbi00n, jcbi01n,
                              jcbi10n, jcbi11n):
    """compute_strain_rate.py:50-73 -> (exx_n, eyy_n, exy_n)"""
    return _nodal_gradients(0, icon_V, u, v, nn_V, nel, dNdr_V_n, dNds_V_n, jcbi00n, jcbi01n,
                            jcbi10n, jcbi11n)


def compute_nodal_pressure_gradient(icon_V, q, nn_V, m_V, nel, dNdr_V_n, dNds_V_n, jcbi00n,
                                    jcbi01n, jcbi10n, jcbi11n):
    """compute_nodal_pressure_gradient.py:11-31 -> (dpdx_n, dpdy_n)"""
    return _nodal_gradients(1, icon_V, q, None, nn_V, nel, dNdr_V_n, dNds_V_n, jcbi00n, jcbi01n,
                            jcbi10n, jcbi11n)


def compute_elemental_strain_rate(icon_V, u, v, nn_V, nel, x_V, y_V):
    """compute_strain_rate.py:17-42 -> (exx, eyy, exy)"""
    icon = _i32(icon_V)
    o = [np.zeros(nel) for _ in range(3)]
    lib().orc_elemental_strain_rate(_p(icon), _p(_f64(u)), _p(_f64(v)), int(nel), _p(_f64(x_V)),
                                    _p(_f64(y_V)), _p(o[0]), _p(o[1]), _p(o[2]))
    return tuple(o)
