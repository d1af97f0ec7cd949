import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def ref_results():
    with open(os.path.join(GOLDEN, "reference_results.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


def donea_huerta_bodyforce(x, y):
    """gravity_model of experiment19.py (Donea & Huerta manufactured solution)."""
    gx = ((12.0 - 24.0 * y) * x ** 4 + (-24.0 + 48.0 * y) * x * x * x
          + (-48.0 * y + 72.0 * y * y - 48.0 * y * y * y + 12.0) * x * x
          + (-2.0 + 24.0 * y - 72.0 * y * y + 48.0 * y * y * y) * x
          + 1.0 - 4.0 * y + 12.0 * y * y - 8.0 * y * y * y)
    gz = ((8.0 - 48.0 * y + 48.0 * y * y) * x * x * x + (-12.0 + 72.0 * y - 72.0 * y * y) * x * x
          + (4.0 - 24.0 * y + 48.0 * y * y - 48.0 * y * y * y + 24.0 * y ** 4) * x
          - 12.0 * y * y + 24.0 * y * y * y - 12.0 * y ** 4)
    return gx, gz


def donea_huerta_exact(x, y):
    """ms_velocity_x/z and ms_pressure for case 19 (postprocessors.py:255-303)."""
    u = x * x * (1.0 - x) ** 2 * (2.0 * y - 6.0 * y * y + 4 * y * y * y)
    w = -y * y * (1.0 - y) ** 2 * (2.0 * x - 6.0 * x * x + 4 * x * x * x)
    p = x * (1.0 - x) - 1.0 / 6.0
    return u, w, p


def stokes_inputs(O, mesh, rhoq, etaq, gxq, gzq, bc):
    """Positional argument tuple of build_matrix_stokes for a box mesh."""
    m = mesh
    fix, val = bc
    eta_e = etaq.mean(axis=1)
    return (m.bignb_Stokes, m.bignb_K, m.bignb_P, m.bignb_G, m.nel, 9, 9, 4, 2, m.Nfem_V, m.Nfem_P,
            18, m.icon_V, m.icon_P, rhoq, etaq, m.JxWq, m.local_to_globalV, gxq, gzq, m.N_V, m.N_P,
            m.dNdr_V, m.dNdt_V, m.dNdr_P, m.dNdt_P, m.jcbi00q, m.jcbi01q, m.jcbi10q, m.jcbi11q,
            eta_e, 1.0, m.L_ref, fix, val, m.bot_element, m.top_element, False, False, "box", None,
            False, m.xq)


def curved_stokes_inputs(O, geometry, nelx, nelz, seed, free_slip=True, bc=True):
    """build_matrix_stokes arguments (without `blocks`) on a quarter/half/eighth
    annulus, r in [0.55, 1]: Jacobians of the isoparametric map at the
    quadrature points (basis_functions_setup_q.py:47-79), random density and
    viscosity, radial gravity, and the velocity conditions of experiment10.py:78-138
    (the sides' normal component and the four corners fixed; the quarter's
    bottom edge z = 0 also fixes w) when `bc`.  Returns (args, curved mesh)."""
    cm = O.curved_mesh(geometry, nelx, nelz, 0.55, 1.0)
    m, X, Z = cm["mesh"], cm["x_V"], cm["z_V"]
    nel = m.nel
    JxW = np.zeros((nel, 9))
    j = [np.zeros((nel, 9)) for _ in range(4)]
    O.lib().orc_jacobians(nel, 9, O._p(m.dNdr_V), O._p(m.dNdt_V), O._p(m.weightq),
                          O._p(O._f64(X)), O._p(O._f64(Z)), O._p(m.icon_V), O._p(JxW),
                          *[O._p(a) for a in j])
    xq = O.project_nodal_Vfield_onto_qpoints(X, 9, nel, 9, m.N_V, m.icon_V)
    zq = O.project_nodal_Vfield_onto_qpoints(Z, 9, nel, 9, m.N_V, m.icon_V)
    rng = np.random.default_rng(seed)
    rho = 1.0 + 0.1 * rng.standard_normal((nel, 9))
    eta = np.exp(rng.uniform(-2.0, 2.0, (nel, 9)))
    rq = np.sqrt(xq ** 2 + zq ** 2)
    gx, gz = -xq / rq, -zq / rq
    fix = np.zeros(m.Nfem_V, dtype=bool)
    if bc:
        eps = 1e-8
        left, right = m.x_V < eps, m.x_V > 1 - eps
        bot, top = m.z_V < eps, m.z_V > 1 - eps
        fix[0::2] |= X / 0.55 < eps
        if geometry == "quarter":
            fix[1::2] |= Z / 0.55 < eps
        corner = (bot | top) & (left | right)
        fix[0::2] |= corner
        fix[1::2] |= corner
    val = np.zeros(m.Nfem_V)
    args = (m.bignb_Stokes, m.bignb_K, m.bignb_P, m.bignb_G, nel, 9, 9, 4, 2, m.Nfem_V, m.Nfem_P,
            18, m.icon_V, m.icon_P, rho, eta, JxW, m.local_to_globalV, gx, gz, m.N_V, m.N_P,
            m.dNdr_V, m.dNdt_V, m.dNdr_P, m.dNdt_P, *j, eta.mean(axis=1), 1.0, 0.775, fix, val,
            m.bot_element, m.top_element, free_slip, free_slip, geometry, cm["theta_V"], False, xq)
    return args, cm


def random_swarm(O, mesh, ppe, seed):
    r, t = O.random_reduced_coords(mesh.nel, ppe, seed)
    x, z, iel = O.place_particles(mesh, r, t, ppe)
    return x, z, r, t, iel


def solcx_material(xq, zq):
    """material_model and gravity_model of experiment21.py:94-107 (SolCx): density
    sin(pi z) cos(pi x), viscosity 1 left of x = 0.5 and 1e6 right of it, g = (0, -1),
    evaluated on the quadrature points (RKorder = -1 puts the particles there)."""
    rho = np.sin(np.pi * zq) * np.cos(np.pi * xq)
    eta = np.where(xq < 0.5, 1.0, 1e6)
    return rho, eta, np.zeros_like(rho), -np.ones_like(rho)


def block_systems(O, n, case):
    """Blocks-mode Stokes system (meeuuw.py:1786-1791) of experiment 19 (Donea-Huerta,
    no slip, eta = rho = 1) or 21 (SolCx, free slip) on an n x n mesh with the
    particles on the quadrature points: (K, G, GT, M, M_eta, H, f, h, Nfem_P).
    eta_e is the harmonic mean of the element's quadrature-point viscosities, the
    default averaging of project_particles_on_elements."""
    from scipy import sparse
    m = O.Mesh(n, n)
    coo = O.coo_indices(m)
    if case == 19:
        gx, gz = donea_huerta_bodyforce(m.xq, m.zq)
        rho = eta = np.ones((m.nel, 9))
        bc = m.bc_no_slip()
    else:
        rho, eta, gx, gz = solcx_material(m.xq, m.zq)
        bc = m.bc_free_slip()
    args = list(stokes_inputs(O, m, rho, eta, gx, gz, bc))
    args[30] = 9.0 / np.sum(1.0 / eta, axis=1)
    out = O.build_matrix_stokes(*args, True)
    nV, nP = m.Nfem_V, m.Nfem_P
    sp_ = lambda v, ij, s: sparse.coo_matrix((v, coo[ij]), shape=s)
    return (sp_(out[3], "K", (nV, nV)).tocsc(), sp_(out[4], "G", (nV, nP)).tocsr(),
            sp_(out[5], "GT", (nP, nV)).tocsr(), sp_(out[6], "MP", (nP, nP)).tocsc(),
            sp_(out[7], "MP", (nP, nP)).tocsc(), sp_(out[8], "GT", (nP, nV)).tocsr(),
            out[1], out[2], nP)
