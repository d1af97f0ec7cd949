"""CPU FP64 oracle of one meeuuw timestep on the box geometry.

TEST INFRASTRUCTURE ONLY (checker and timed CPU baseline).  Restates the body of
the reference time loop, meeuuw.py:1105-2875, in its order, for the
thermal-convection setups used here (experiment0.py physics with either a
constant or a Frank-Kamenetskii viscosity as in experiment3.py:46, 72):
interpolate T on particles -> material model -> elemental and nodal projections
-> qpoint fields -> Stokes assembly + direct solve -> pressure normalisation ->
nodal pressure -> dt -> energy assembly + direct solve -> heat flux, Nu, vrms
-> RK advection -> locate.

Parity unpinned: `population_control` (no reference-held output; the
reference draws from the unseeded random module) and the Picard loop in
`step` with iterations on (no reference-held trace).  `assess_nlconvergence`
is pinned on vectors of the reference's own function.
"""
import math
import time

import numpy as np

from . import oracle as O


class Params:
    def __init__(self, nelx=32, nelz=32, Lx=1.0, Lz=1.0, Ra=1e4, eta_contrast=1.0, eta_min=1e-5,
                 eta_max=1.0, ppe=16, RKorder=2, seed=20261203, alphaT=1e-2, rho0=1.0, CFL=0.25,
                 averaging="harmonic", nodal_projection_type=2, pressure_normalisation="surface",
                 material_mode=0, mat_eta=(1.0, 1.0), mat_rho=(1.0, 1.0), gravity=0.0,
                 rheology=0, tosi_case=1, gamma_T=0.0, gamma_y=0.0, sigma_y=0.0, eta_star=0.0,
                 nonlinear=False, niter_nl=25, tol_nl=1e-3, population_control=False,
                 nparticle_min=0, pc_seed=0, rho_projection="elemental",
                 eta_projection="elemental"):
        self.__dict__.update(locals())
        del self.__dict__["self"]


def initial_temperature(mesh):
    """experiment0.py:91-95 with Tbottom=1, Ttop=0."""
    x, z, Lx, Lz = mesh.x_V, mesh.z_V, mesh.Lx, mesh.Lz
    return (1.0 - 0.0) * (Lz - z) / Lz + 0.0 + 0.01 * np.cos(np.pi * x / Lx) * np.sin(np.pi * z / Lz)


def viscosity(P, T):
    """eta = exp(-ln(contrast) * T) clipped to [eta_min, eta_max] (experiment3.py:46, 72, 105-106);
    eta = 1 for contrast 1 (experiment0.py:256)."""
    if P.eta_contrast == 1.0:
        return np.ones_like(T)
    eta = np.exp(-math.log(P.eta_contrast) * T)
    return np.clip(eta, P.eta_min, P.eta_max)


def tosi_viscosity(P, T, exx, eyy, exy, y):
    """viscosity(), experiment3.py:67-110, elementwise over the particles."""
    c = P.tosi_case
    if c == 1:
        val = np.exp(-P.gamma_T * T)
    elif c == 3:
        val = np.exp(-P.gamma_T * T + P.gamma_y * (1 - y))
    else:
        e = np.sqrt(0.5 * (exx ** 2 + eyy ** 2) + exy ** 2)
        e = np.maximum(e, 1e-12)
        if c == 2:
            eta_lin = np.exp(-P.gamma_T * T)
        else:
            eta_lin = np.exp(-P.gamma_T * T + P.gamma_y * (1 - y))
        eta_plast = P.eta_star + P.sigma_y / (np.sqrt(2.0) * e)
        val = 2.0 / (1.0 / eta_lin + 1.0 / eta_plast)
    val = np.minimum(2.0, val)
    return np.maximum(1.0e-5, val)


def pc_uniform(seed, istep, draw):
    """The population control's draw in [0, 1): splitmix64 of the counter
    seed * 0x9E3779B97F4A7C15 + istep * 0xD1B54A32D192ED03 + draw, top 53 bits.
    It stands in for the reference's unseeded random.uniform stream, so that
    the oracle and the GPU place the same particles."""
    M = (1 << 64) - 1
    z = (seed * 0x9E3779B97F4A7C15 + istep * 0xD1B54A32D192ED03 + draw) & M
    z = (z + 0x9E3779B97F4A7C15) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    z ^= z >> 31
    return (z >> 11) * 2.0 ** -53


def population_control(S, nparticle_e):
    """population_control, pic_functions.py:1623-1825: flagged elements (fewer
    than nparticle_min particles at the start of the step) get the missing
    particles from the inactive ones, in index order, at r, t uniform in
    [-0.95, 0.95], with the material of the nearest particle of the element's
    mask (old element, active).  Returns (needed, available, active)."""
    P, m = S.P, S.mesh
    nmin = P.nparticle_min if P.nparticle_min > 0 else int(P.ppe * 0.85)
    n = S.nparticle
    needed = 0
    for iel in range(m.nel):
        needed += max(0, nmin - int(nparticle_e[iel]))
    if needed == 0:
        return 0, 0, 0
    active = int(np.sum(S.swarm_active))
    available = n - active
    if needed > available:
        raise ValueError("population_control: not available yet")
    avail = [ip for ip in range(n) if not S.swarm_active[ip]]
    counter = 0
    for iel in range(m.nel):
        if nparticle_e[iel] >= nmin:
            continue
        mask = np.logical_and(S.swarm_iel == iel, S.swarm_active)
        if np.sum(mask) == 0:
            raise ValueError("population_control: mask empty")
        ids = np.nonzero(mask)[0]
        for _ in range(nmin - int(nparticle_e[iel])):
            pid = avail[counter]
            r = -0.95 + 1.9 * pc_uniform(P.pc_seed, S.istep, 2 * counter)
            t = -0.95 + 1.9 * pc_uniform(P.pc_seed, S.istep, 2 * counter + 1)
            S.swarm_active[pid] = True
            S.swarm_iel[pid] = iel
            S.swarm_r[pid], S.swarm_t[pid] = r, t
            N = O.basis_functions_V(r, t)
            x = z = 0.0
            for k in range(9):
                x += N[k] * m.x_V[m.icon_V[k, iel]]
                z += N[k] * m.z_V[m.icon_V[k, iel]]
            S.swarm_x[pid], S.swarm_z[pid] = x, z
            dist = (S.swarm_x[ids] - x) ** 2 + (S.swarm_z[ids] - z) ** 2
            S.swarm_mat[pid] = S.swarm_mat[ids[np.argmin(dist)]]
            counter += 1
    return needed, available, active


class State:
    """Everything carried from one step to the next."""

    def __init__(self, P):
        self.P = P
        self.mesh = m = O.Mesh(P.nelx, P.nelz, P.Lx, P.Lz)
        self.coo = O.coo_indices(m)
        self.bc_fix_V, self.bc_val_V = m.bc_free_slip()
        self.bc_fix_T, self.bc_val_T = m.bc_T()
        self.T = initial_temperature(m)
        if P.RKorder == -1:
            # particles collocated with the quadrature points (meeuuw.py:503-509,
            # swarm_coordinates_setup.py:51-65)
            P.ppe = 9
            self.swarm_x = m.xq.ravel().copy()
            self.swarm_z = m.zq.ravel().copy()
            self.swarm_r = np.tile(m.rq, m.nel)
            self.swarm_t = np.tile(m.tq, m.nel)
            self.swarm_iel = np.repeat(np.arange(m.nel, dtype=np.int32), 9)
        else:
            r, t = O.random_reduced_coords(m.nel, P.ppe, P.seed)
            self.swarm_x, self.swarm_z, self.swarm_iel = O.place_particles(m, r, t, P.ppe)
            self.swarm_r, self.swarm_t = r, t
        self.nparticle = m.nel * P.ppe
        self.swarm_active = np.ones(self.nparticle, dtype=bool)
        self.q = np.zeros(m.nn_V)
        self.gxq = np.zeros((m.nel, 9))
        # experiment0.py:274, or a given gravity (the two-material setups)
        self.gzq = np.full((m.nel, 9), -P.gravity if P.gravity > 0 else -P.Ra / P.alphaT)
        # material index per particle (material_mode 1): the reference's particle
        # material weights (swarm_wf, experiment21.py:73-81 pattern) with one
        # material per particle
        self.swarm_mat = np.zeros(self.nparticle, dtype=np.uint8)
        self.dt1_mem = 1e50
        self.dt2_mem = 1e50
        # nodal strain rate (zero before the first solve) and the Picard memories
        # (meeuuw.py:1090-1097, 708)
        self.exx_n = np.zeros(m.nn_V)
        self.ezz_n = np.zeros(m.nn_V)
        self.exz_n = np.zeros(m.nn_V)
        self.u_mem = np.zeros(m.nn_V)
        self.w_mem = np.zeros(m.nn_V)
        self.p_mem = np.zeros(m.nn_P)
        self.T_mem = None  # T.copy() at the first step (after any initial temperature)
        self.nl_history = []
        self.time = 0.0
        self.istep = 0


def assess_nlconvergence(istep, iter_nl, solve_Stokes, solve_T, u, w, p, T, u_mem, w_mem, p_mem,
                         T_mem, tol_nl, inside):
    """assess_steady_state.py:10-74: the relative 2-norm changes xi of u, w, p
    and T since the last iteration, converged when every one is below tol_nl
    (T counts as converged when it is not solved, the flow as not converged
    when it is not solved), new memories of u, w, p, and of T only once the
    iterations end.  Returns (u_mem, w_mem, p_mem, T_mem, inside, xi, the line
    the reference writes to conv.ascii)."""
    rel = lambda a, b: float(np.linalg.norm(a - b, 2) / np.linalg.norm(b, 2))
    if solve_Stokes:
        xu, xw, xp = rel(u_mem, u), rel(w_mem, w), rel(p_mem, p)
        conv = xu < tol_nl and xw < tol_nl and xp < tol_nl
    else:
        xu = xw = xp = 0.0
        conv = False
    xT = rel(T_mem, T) if solve_T else 0.0
    conv = conv and (xT < tol_nl if solve_T else True)
    line = "%.3e %.3e %.3e %.3e %.3e %.3e %d\n" % (istep + iter_nl / 200, xu, xw, xp, xT, tol_nl,
                                                   iter_nl)
    if conv:
        inside = False
    if solve_T and not inside:
        T_mem = T.copy()
    return u.copy(), w.copy(), p.copy(), T_mem, inside, [xu, xw, xp, xT], line


def _iteration(S, iter_nl, inside, tick):
    """Stages of one pass of the time loop up to the energy solve
    (meeuuw.py:1112-2250) and the nodal strain rate (2386-2389); returns the
    fields and whether further Picard iterations follow."""
    P, m = S.P, S.mesh
    nel, nn = m.nel, m.nn_V
    t0 = time.perf_counter()
    n = S.nparticle
    swarm_T = O.interpolate_field_on_particles(n, S.swarm_active, S.swarm_r, S.swarm_t, S.swarm_iel,
                                               S.T, m.icon_V)
    t0 = tick("interp", t0)
    # material model (experiment0.py:230-264), or two materials: eta and
    # rho (1 - alpha T) of each particle's material
    if P.rheology == 1:  # experiment3.py:218-242 with the last nodal strain rate
        sr = [O.interpolate_field_on_particles(n, S.swarm_active, S.swarm_r, S.swarm_t,
                                               S.swarm_iel, f, m.icon_V)
              for f in (S.exx_n, S.ezz_n, S.exz_n)]
        swarm_rho = P.rho0 * (1 - P.alphaT * swarm_T)
        swarm_eta = tosi_viscosity(P, swarm_T, sr[0], sr[1], sr[2], S.swarm_z)
    elif P.material_mode == 1:
        mat = S.swarm_mat.astype(bool)
        swarm_rho = np.where(mat, P.mat_rho[1], P.mat_rho[0]) * (1 - P.alphaT * swarm_T)
        swarm_eta = np.where(mat, P.mat_eta[1], P.mat_eta[0]).astype(np.float64)
    else:
        swarm_rho = P.rho0 * (1 - P.alphaT * swarm_T)
        swarm_eta = viscosity(P, swarm_T)
    swarm_hcond = np.ones(n)
    swarm_hcapa = np.ones(n)
    swarm_hprod = np.zeros(n)
    swarm_alpha = np.full(n, P.alphaT)
    t0 = tick("material", t0)
    rho_e, eta_e, cnt = O.project_particles_on_elements(nel, n, S.swarm_active, swarm_rho,
                                                        swarm_eta, S.swarm_iel, P.averaging)
    if np.min(cnt) == 0:
        raise SystemExit("ABORT: an element contains no particle!")
    t0 = tick("proj_elements", t0)
    proj = {1: O.project_particle_field_on_nodes_1, 2: O.project_particle_field_on_nodes_2,
            3: O.project_particle_field_on_nodes_3, 4: O.project_particle_field_on_nodes_4}
    pf = proj[P.nodal_projection_type]
    args = (nel, nn, n, S.swarm_active)
    rest = (m.icon_V, S.swarm_iel, S.swarm_r, S.swarm_t)
    rho_n = pf(*args, swarm_rho, *rest, "arithmetic")
    eta_n = pf(*args, swarm_eta, *rest, P.averaging)
    p2 = O.project_particle_field_on_nodes_2
    hcond_n = p2(*args, swarm_hcond, *rest, "arithmetic")
    hcapa_n = p2(*args, swarm_hcapa, *rest, "arithmetic")
    hprod_n = p2(*args, swarm_hprod, *rest, "arithmetic")
    alpha_n = p2(*args, swarm_alpha, *rest, "arithmetic")
    t0 = tick("proj_nodes", t0)
    if P.RKorder == -1:  # 'qpts' projection (meeuuw.py:1503-1508)
        rhoq = swarm_rho.reshape(nel, 9).copy()
        etaq = swarm_eta.reshape(nel, 9).copy()
    else:  # 'elemental' projection (meeuuw.py:1509-1511)
        rhoq = np.repeat(rho_e[:, None], 9, axis=1)
        etaq = np.repeat(eta_e[:, None], 9, axis=1)
        # 'nodal' (meeuuw.py:1514-1515, 1539-1540)
        if P.rho_projection == "nodal":
            rhoq = O.project_nodal_Pfield_onto_qpoints(rho_n, 9, nel, 4, m.N_P, m.icon_V)
        if P.eta_projection == "nodal":
            etaq = O.project_nodal_Pfield_onto_qpoints(eta_n, 9, nel, 4, m.N_P, m.icon_V)
        ls = [k for k in (P.rho_projection, P.eta_projection) if k.startswith("least_squares")]
        if ls:  # meeuuw.py:1224-1290 and 1516-1530, 1541-1555
            nb = 3 if "least_squares_P1" in ls else 4
            x_e, z_e = m.x_V[m.icon_V[8]], m.z_V[m.icon_V[8]]
            # the element's active particles (an inactive particle keeps a stale
            # element index and sits at (0, 0); it is left out of the fit)
            iel_act = np.where(S.swarm_active, S.swarm_iel, -1)
            cr, ce, rmin, rmax, emin, emax = O.compute_ls_coefficients(
                nb, nel, x_e, z_e, S.swarm_x, S.swarm_z, iel_act, swarm_rho, swarm_eta)
            if nb == 3:
                for iel in range(nel):
                    cr[iel, 1], cr[iel, 2] = O.limiter(cr[iel, 0], cr[iel, 1], cr[iel, 2],
                                                       rmin[iel], rmax[iel], m.hx)
                    ce[iel, 1], ce[iel, 2] = O.limiter(ce[iel, 0], ce[iel, 1], ce[iel, 2],
                                                       emin[iel], emax[iel], m.hx)
            dx = m.xq - x_e[:, None]
            dz = m.zq - z_e[:, None]
            for kind, c in ((P.rho_projection, cr), (P.eta_projection, ce)):
                if not kind.startswith("least_squares"):
                    continue
                v = c[:, 0:1] + c[:, 1:2] * dx + c[:, 2:3] * dz
                if nb == 4:
                    v = v + c[:, 3:4] * dx * dz
                if c is cr:
                    rhoq = v
                else:
                    etaq = v
    Tq = O.project_nodal_Tfield_onto_qpoints(S.T, 9, nel, 9, m.N_V, m.icon_V)
    hcapaq = O.project_nodal_Pfield_onto_qpoints(hcapa_n, 9, nel, 4, m.N_P, m.icon_V)
    hcondq = O.project_nodal_Pfield_onto_qpoints(hcond_n, 9, nel, 4, m.N_P, m.icon_V)
    hprodq = O.project_nodal_Pfield_onto_qpoints(hprod_n, 9, nel, 4, m.N_P, m.icon_V)
    alphaq = O.project_nodal_Pfield_onto_qpoints(alpha_n, 9, nel, 4, m.N_P, m.icon_V)
    zq = np.zeros((nel, 9))
    t0 = tick("qpoints", t0)
    VV, rhs_f, rhs_h = O.build_matrix_stokes(
        m.bignb_Stokes, m.bignb_K, m.bignb_P, m.bignb_G, nel, 9, 9, 4, 2, m.Nfem_V, m.Nfem_P, 18,
        m.icon_V, m.icon_P, rhoq, etaq, m.JxWq, m.local_to_globalV, S.gxq, S.gzq, m.N_V, m.N_P,
        m.dNdr_V, m.dNdt_V, m.dNdr_P, m.dNdt_P, m.jcbi00q, m.jcbi01q, m.jcbi10q, m.jcbi11q, eta_e,
        1.0, m.L_ref, S.bc_fix_V, S.bc_val_V, m.bot_element, m.top_element, False, False, "box",
        None, False, m.xq, False)[:3]
    t0 = tick("build_stokes", t0)
    u, w, p = O.stokes_direct(m, S.coo, VV, rhs_f, rhs_h)
    t0 = tick("solve_stokes", t0)
    vel = np.sqrt(u ** 2 + w ** 2)
    off = O.compute_pressure_offset("box", P.pressure_normalisation, False, m.top_element, nel,
                                    m.nelx, 9, m.N_P, m.JxWq, p, m.icon_P, None, m.xq, m.volume)
    p = p - off
    t0 = tick("pressure_offset", t0)
    S.q = O.compute_nodal_pressure(9, nn, m.icon_V, m.icon_P, p, m.N_P_n)
    t0 = tick("nodal_pressure", t0)
    if P.nonlinear:
        S.u_mem, S.w_mem, S.p_mem, S.T_mem, inside, xi, _ = assess_nlconvergence(
            S.istep, iter_nl, True, True, u, w, p, S.T, S.u_mem, S.w_mem, S.p_mem, S.T_mem,
            P.tol_nl, inside)
        S.nl_history.append(xi)
    else:
        S.T_mem = S.T.copy()
    h = min(m.hx, m.hz)
    dt1 = P.CFL * h / np.max(vel)
    dt2 = P.CFL * h ** 2 / (np.average(swarm_hcond) / np.average(swarm_hcapa) / np.average(swarm_rho))
    dt1 = min(dt1, 1.25 * S.dt1_mem)
    dt2 = min(dt2, 1.25 * S.dt2_mem)
    dt = min(dt1, dt2, 1e30)
    if not inside:
        S.dt1_mem, S.dt2_mem = dt1, dt2
    t0 = tick("dt", t0)
    uq = O.project_nodal_Vfield_onto_qpoints(u, 9, nel, 9, m.N_V, m.icon_V)
    wq = O.project_nodal_Vfield_onto_qpoints(w, 9, nel, 9, m.N_V, m.icon_V)
    t0 = tick("qpoints", t0)
    VV_T, rhs_T = O.build_matrix_energy(m.bignb_T, nel, 9, 9, nn, S.T_mem, m.icon_V, rhoq, etaq, Tq,
                                        uq, wq, hcondq, hcapaq, alphaq, hprodq, zq, zq, zq, zq, zq,
                                        m.JxWq, m.N_V, m.dNdr_V, m.dNdt_V, m.jcbi00q, m.jcbi01q,
                                        m.jcbi10q, m.jcbi11q, S.bc_fix_T, S.bc_val_T, dt, "BA",
                                        P.rho0)
    t0 = tick("build_energy", t0)
    S.T = O.energy_direct(m, S.coo, VV_T, rhs_T)
    t0 = tick("solve_energy", t0)
    # nodal strain rate (meeuuw.py:2386-2389): this step's global quantities and
    # the next material evaluation
    S.exx_n, S.ezz_n, S.exz_n = O.compute_nodal_strain_rate(
        m.icon_V, u, w, nn, 9, nel, m.dNdr_V_n, m.dNdt_V_n, m.jcbi00n, m.jcbi01n, m.jcbi10n,
        m.jcbi11n)
    t0 = tick("strain_rate", t0)
    return dict(u=u, w=w, p=p, dt=dt, rho_e=rho_e, eta_e=eta_e, rho_n=rho_n, eta_n=eta_n,
                hcond_n=hcond_n, uq=uq, wq=wq, Tq=Tq, rhoq=rhoq, hcapaq=hcapaq, alphaq=alphaq,
                etaq=etaq, zq=zq, n=n, cnt=cnt), inside


def step(S, timings=None):
    """One timestep (with its Picard iterations when P.nonlinear); returns a
    dict of the diagnostics and fields."""
    P, m = S.P, S.mesh
    nel, nn = m.nel, m.nn_V
    tm = timings if timings is not None else {}

    def tick(name, t0):
        tm[name] = tm.get(name, 0.0) + time.perf_counter() - t0
        return time.perf_counter()

    if S.T_mem is None:
        S.T_mem = S.T.copy()
    S.nl_history = []
    niter = P.niter_nl if P.nonlinear else 1
    for iter_nl in range(niter):
        # meeuuw.py:2084-2089
        inside = bool(P.nonlinear) and iter_nl < niter - 1
        it, inside = _iteration(S, iter_nl, inside, tick)
        if not inside:
            break
    u, w, p, dt, n = it["u"], it["w"], it["p"], it["dt"], it["n"]
    hcond_n, uq, wq, Tq = it["hcond_n"], it["uq"], it["wq"], it["Tq"]
    rhoq, hcapaq, alphaq, etaq, zq = it["rhoq"], it["hcapaq"], it["alphaq"], it["etaq"], it["zq"]
    t0 = time.perf_counter()
    dTdx, dTdz, qx, qz = O.compute_nodal_heat_flux(m.icon_V, S.T, hcond_n, nn, 9, nel, m.dNdr_V_n,
                                                   m.dNdt_V_n, m.jcbi00n, m.jcbi01n, m.jcbi10n,
                                                   m.jcbi11n)
    t0 = tick("heat_flux", t0)
    nu = O.compute_Nu(m.Lx, m.Lz, m.x_V, m.z_V, nel, m.left_element, m.right_element,
                      m.top_element, m.bot_element, m.icon_V, S.T, dTdx, dTdz, 3, m.qcoords,
                      m.qweights, hcond_n)
    t0 = tick("Nu", t0)
    # strain rate on the quadrature points (meeuuw.py:2496-2498)
    exxq, ezzq, exzq = (O.project_nodal_Vfield_onto_qpoints(f, 9, nel, 9, m.N_V, m.icon_V)
                        for f in (S.exx_n, S.ezz_n, S.exz_n))
    gq = O.compute_global_quantities(nel, 9, m.xq, m.zq, uq, wq, Tq, rhoq, hcapaq, alphaq, etaq,
                                     exxq, ezzq, exzq, zq, zq, zq, m.volume, m.JxWq, S.gxq, S.gzq)
    t0 = tick("vrms", t0)
    if P.RKorder > 0:
        S.swarm_x, S.swarm_z, su, sw, S.swarm_active = O.advect_particles___box(
            P.RKorder, dt, n, S.swarm_x, S.swarm_z, S.swarm_active, u, w, m.Lx, m.Lz, m.hx, m.hz,
            m.nelx, m.icon_V, m.x_V, m.z_V)
        t0 = tick("advect", t0)
        if P.population_control:  # meeuuw.py:2847-2851
            S.pc_stats = population_control(S, it["cnt"])
            t0 = tick("population_control", t0)
        S.swarm_r, S.swarm_t, S.swarm_iel = O.locate_particles(
            "box", n, S.swarm_active, S.swarm_x, S.swarm_z, None, None, m.hx, m.hz, 0, 0, m.x_V,
            m.z_V, None, None, m.icon_V, m.nelx, 0)
        t0 = tick("locate", t0)
    out = dict(u=u, w=w, p=p, q=S.q, T=S.T, dt=dt, vrms=gq[0], Nu=nu[12], rho_e=it["rho_e"],
               eta_e=it["eta_e"], rho_n=it["rho_n"], eta_n=it["eta_n"], time=S.time,
               istep=S.istep, Tavrg=gq[7], EK=gq[1], WAG=gq[2], TVD=gq[3], TM=gq[6],
               eta_avrg=gq[8], nl_its=iter_nl + 1,
               nl_history=np.array(S.nl_history).reshape(-1, 4))
    S.time += dt
    S.istep += 1
    return out
