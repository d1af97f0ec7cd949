"""Device-resident timestep of the box-geometry convection model.

``Simulation`` owns the whole state on the GPU (mesh, particles, fields,
solver hierarchy) and advances it with one C call per step
(mw_sim_step).  The stages are those of the reference time loop,
meeuuw.py:1105-2875; see include/meeuuw_hip.h for the contract.
"""
import ctypes

import numpy as np

from . import _lib

SOLVERS = {"plain": 0, "mixed": 1, "learned": 2}
AVERAGING = {"arithmetic": 0, "geometric": 1, "harmonic": 2}
PNORM = {"none": 0, "surface": 1, "volume": 2}
QPROJ = {"elemental": 0, "nodal": 1, "least_squares_P1": 2, "least_squares_Q1": 3}

STATS = ["dt", "vrms", "Nu", "stokes_outer", "stokes_vcycles", "stokes_relres", "energy_its",
         "energy_relres", "Tavrg", "time", "pressure_offset", "mincount", "nactive", "learned",
         "learn_ms", "learn_ratio", "learn_its_plain", "learn_its", "EK", "WAG", "TVD", "TM",
         "eta_avrg", "dt1", "dt2", "maxcount", "eta_e_min", "eta_e_max", "nl_its", "nl_xi_u",
         "nl_xi_w", "nl_xi_p", "nl_xi_T", "stokes_fgmres64", "pc_needed", "pc_available",
         "pc_active", "learn_probe_ms_plain", "learn_probe_ms"]
NSTATS = 48

FIELDS = {"T": 0, "u": 1, "w": 2, "p": 3, "q": 4, "swarm_x": 5, "swarm_z": 6, "swarm_r": 7,
          "swarm_t": 8, "swarm_iel": 9, "swarm_active": 10, "rho_e": 11, "eta_e": 12, "rho_n": 13,
          "eta_n": 14, "dTdx": 15, "dTdz": 16, "VV": 17, "rhs_f": 18, "rhs_h": 19,
          "stokes_history": 20, "exx_n": 21, "ezz_n": 22, "exz_n": 23, "nl_history": 24,
          "swarm_mat": 25}
_FIELD_DTYPE = {"swarm_iel": np.int32, "swarm_active": np.uint8, "swarm_mat": np.uint8}


class SimConfig(ctypes.Structure):
    """Mirror of mw_sim_config."""
    _fields_ = [("nelx", ctypes.c_int), ("nelz", ctypes.c_int), ("Lx", ctypes.c_double),
                ("Lz", ctypes.c_double), ("ppe", ctypes.c_int), ("rk", ctypes.c_int),
                ("Ra", ctypes.c_double), ("alphaT", ctypes.c_double), ("rho0", ctypes.c_double),
                ("CFL", ctypes.c_double), ("Tbottom", ctypes.c_double), ("Ttop", ctypes.c_double),
                ("eta_contrast", ctypes.c_double), ("eta_min", ctypes.c_double),
                ("eta_max", ctypes.c_double), ("averaging", ctypes.c_int),
                ("nodal_projection_type", ctypes.c_int), ("pressure_normalisation", ctypes.c_int),
                ("solver", ctypes.c_int), ("stokes_tol", ctypes.c_double),
                ("energy_tol", ctypes.c_double), ("fgmres_restart", ctypes.c_int),
                ("cheb_degree", ctypes.c_int), ("warm_start", ctypes.c_int),
                ("pc_vcycles", ctypes.c_int), ("learn_pairs", ctypes.c_int),
                ("learn_steps", ctypes.c_int), ("learn_batch", ctypes.c_int),
                ("learn_vcycles", ctypes.c_int), ("learn_lr", ctypes.c_double),
                ("learn_seed", ctypes.c_int), ("fgmres_graphs", ctypes.c_int),
                ("gmg_tail", ctypes.c_int), ("cheb_degree_coarse", ctypes.c_int),
                ("gmg_bottom", ctypes.c_int), ("extrapolate", ctypes.c_int),
                ("schur_cheb", ctypes.c_int), ("schur_scale", ctypes.c_double),
                ("energy_mf", ctypes.c_int), ("fgmres_inner", ctypes.c_double),
                ("gmg_power_its", ctypes.c_int), ("gs_passes", ctypes.c_int),
                ("material_mode", ctypes.c_int), ("mat_eta", ctypes.c_double * 2),
                ("mat_rho", ctypes.c_double * 2), ("gravity", ctypes.c_double),
                ("gmg_eta_coarsen", ctypes.c_int), ("fgmres_maxit", ctypes.c_int),
                ("pcg_maxit", ctypes.c_int), ("rheology", ctypes.c_int),
                ("tosi_case", ctypes.c_int), ("tosi_gamma_T", ctypes.c_double),
                ("tosi_gamma_y", ctypes.c_double), ("tosi_sigma_y", ctypes.c_double),
                ("tosi_eta_star", ctypes.c_double), ("nonlinear", ctypes.c_int),
                ("niter_nl", ctypes.c_int), ("tol_nl", ctypes.c_double),
                ("population_control", ctypes.c_int), ("nparticle_min", ctypes.c_int),
                ("pc_seed", ctypes.c_int), ("rho_projection", ctypes.c_int),
                ("eta_projection", ctypes.c_int), ("gmg_coarse_dofs", ctypes.c_int)]


def make_config(nelx=32, nelz=32, Lx=1.0, Lz=1.0, ppe=16, rk=2, Ra=1e4, alphaT=1e-2, rho0=1.0,
                CFL=0.25, Tbottom=1.0, Ttop=0.0, eta_contrast=1.0, eta_min=1e-5, eta_max=1.0,
                averaging="harmonic", nodal_projection_type=2, pressure_normalisation="surface",
                solver="mixed", stokes_tol=1e-8, energy_tol=1e-10, fgmres_restart=20,
                cheb_degree=2, warm_start=True, pc_vcycles=1, learn_pairs=512, learn_steps=400,
                learn_batch=8, learn_vcycles=2, learn_lr=2e-3, learn_seed=20261203, fgmres_graphs=False,
                gmg_tail=False, cheb_degree_coarse=0, gmg_bottom=3, extrapolate=1, schur_cheb=3,
                schur_scale=2.5, energy_mf=None, fgmres_inner=0.8, gmg_power_its=2,
                gs_passes=0, material_mode=0, mat_eta=(1.0, 1.0), mat_rho=(1.0, 1.0),
                gravity=0.0, gmg_eta_coarsen=0, fgmres_maxit=0, pcg_maxit=0, rheology=0,
                tosi_case=1, tosi_gamma_T=0.0, tosi_gamma_y=0.0, tosi_sigma_y=0.0,
                tosi_eta_star=0.0, nonlinear=False, niter_nl=0, tol_nl=0.0,
                population_control=False, nparticle_min=0, pc_seed=0,
                rho_projection="elemental", eta_projection="elemental", gmg_coarse_dofs=600):
    """Defaults follow set_default_parameters.py (CFL 0.25, harmonic averaging,
    nodal projection 2, surface normalisation) and experiment0.py (alphaT 1e-2,
    rho0 1, T 1 at the bottom and 0 at the top).  energy_mf None: matrix-free
    energy operator for the mixed and learned solvers, the assembled one
    (build_matrix_energy) for the plain solver.  gmg_coarse_dofs 600: the
    multigrid stops at the first level with at most 600 unknowns (8x8 elements
    of a square box) and solves it with its dense inverse, formed one solve
    ahead; 60 keeps the levels down to 2x2 elements, the last three fused with
    the coarsest solve in one workgroup (measured at 256x256: 3.51 against 3.68
    ms per step, 13.4 against 12.7 FGMRES iterations)."""
    if energy_mf is None:
        energy_mf = solver != "plain"
    return SimConfig(int(nelx), int(nelz), float(Lx), float(Lz), int(ppe), int(rk), float(Ra),
                     float(alphaT), float(rho0), float(CFL), float(Tbottom), float(Ttop),
                     float(eta_contrast), float(eta_min), float(eta_max), AVERAGING[averaging],
                     int(nodal_projection_type), PNORM[pressure_normalisation], SOLVERS[solver],
                     float(stokes_tol), float(energy_tol), int(fgmres_restart), int(cheb_degree),
                     int(bool(warm_start)), int(pc_vcycles), int(learn_pairs), int(learn_steps),
                     int(learn_batch), int(learn_vcycles), float(learn_lr), int(learn_seed),
                     int(bool(fgmres_graphs)), int(bool(gmg_tail)), int(cheb_degree_coarse),
                     int(gmg_bottom), int(extrapolate), int(schur_cheb),
                     float(schur_scale), int(bool(energy_mf)), float(fgmres_inner),
                     int(gmg_power_its), int(gs_passes), int(material_mode),
                     (ctypes.c_double * 2)(*mat_eta), (ctypes.c_double * 2)(*mat_rho),
                     float(gravity), int(gmg_eta_coarsen), int(fgmres_maxit), int(pcg_maxit),
                     int(rheology), int(tosi_case), float(tosi_gamma_T), float(tosi_gamma_y),
                     float(tosi_sigma_y), float(tosi_eta_star), int(bool(nonlinear)),
                     int(niter_nl), float(tol_nl), int(bool(population_control)),
                     int(nparticle_min), int(pc_seed), QPROJ[rho_projection],
                     QPROJ[eta_projection], int(gmg_coarse_dofs))


OPERATORS = {"stokes32": 0, "prec32": 1, "restrict32": 2, "prolong32": 3, "mask32": 4,
             "prec32_split": 5}


class Simulation:
    def __init__(self, config, stream=None, _handle=None):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("meeuuw_amd.timestep needs a ROCm GPU (no CPU fallback)")
        self.cfg = config
        self._owned = _handle is None
        if _handle is not None:
            self._h = _handle  # an instance of a Batch, which owns it
        else:
            if stream is None:
                stream = torch.cuda.current_stream().cuda_stream
            self._h = ctypes.c_void_p()
            _lib.call("mw_sim_create", ctypes.byref(config), ctypes.c_void_p(stream),
                      ctypes.byref(self._h))
        self.nel = config.nelx * config.nelz
        self.nn = (2 * config.nelx + 1) * (2 * config.nelz + 1)
        self.nparticle = self.nel * (9 if config.rk == -1 else config.ppe)
        self.stats = dict.fromkeys(STATS, 0.0)

    def close(self):
        if self._h and self._owned:
            _lib.call("mw_sim_destroy", self._h)
        self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_temperature(self, T):
        T = np.ascontiguousarray(T, dtype=np.float64)
        if T.size != self.nn:
            raise ValueError("set_temperature: expected nn_V values")
        _lib.call("mw_sim_set_temperature", self._h, T.ctypes.data)

    def set_swarm_reduced(self, r, t):
        r = np.ascontiguousarray(r, dtype=np.float64)
        t = np.ascontiguousarray(t, dtype=np.float64)
        if r.size != self.nparticle or t.size != self.nparticle:
            raise ValueError("set_swarm_reduced: expected nel*ppe values")
        _lib.call("mw_sim_set_swarm_reduced", self._h, r.ctypes.data, t.ctypes.data)

    def set_swarm(self, x, z, r, t, iel, active):
        """Explicit particle state (mw_sim_set_swarm): positions, reduced
        coordinates, elements and active flags of all nel*ppe particles."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, z, r, t)]
        iel = np.ascontiguousarray(iel, dtype=np.int32)
        act = np.ascontiguousarray(np.asarray(active).astype(np.uint8))
        if any(a.size != self.nparticle for a in (*arrs, iel, act)):
            raise ValueError("set_swarm: expected nel*ppe values per array")
        _lib.call("mw_sim_set_swarm", self._h, self.nparticle, *(a.ctypes.data for a in arrs),
                  iel.ctypes.data, act.ctypes.data)

    def set_swarm_material(self, mat):
        """Material index (0/1) of every particle, particle order (material_mode 1)."""
        m = np.ascontiguousarray(np.asarray(mat).astype(np.uint8))
        if m.size != self.nparticle:
            raise ValueError("set_swarm_material: expected one value per particle")
        _lib.call("mw_sim_set_swarm_material", self._h, m.ctypes.data)

    def seed_swarm(self, seed):
        """(r, t) ~ U(-1, 1), element-major, r before t, from numpy PCG64(seed)."""
        rng = np.random.Generator(np.random.PCG64(seed))
        rt = rng.uniform(-1.0, 1.0, size=(self.nel, self.cfg.ppe, 2))
        self.set_swarm_reduced(rt[:, :, 0].ravel(), rt[:, :, 1].ravel())

    def step(self):
        buf = (ctypes.c_double * NSTATS)()
        rc = _lib.lib.mw_sim_step(self._h, buf)
        for i, k in enumerate(STATS):
            self.stats[k] = buf[i]
        _lib.check(rc)
        self._last_stats = buf
        return dict(self.stats)

    def run(self, nstep, output=None):
        """nstep timesteps (the reference's time loop, meeuuw.py:1105); with an
        Output, its files are written after every step, and the solution and
        swarm of the last step always (meeuuw.py:2913, 2993)."""
        for istep in range(nstep):
            self.step()
            if output is not None:
                output.write(last=istep == nstep - 1)
        if output is not None:
            output.flush()
        return dict(self.stats)

    def apply_operator(self, op, x, y):
        """y = op(x) for the mixed solver's operators (mw_sim_operator):
        x and y are float32 device tensors of the operator's sizes."""
        _lib.call("mw_sim_operator", self._h, int(OPERATORS[op]), ctypes.c_void_p(x.data_ptr()),
                  ctypes.c_void_p(y.data_ptr()))

    def field(self, name):
        f = FIELDS[name]
        n = _lib.lib.mw_sim_field_size(self._h, f)
        out = np.empty(n, dtype=_FIELD_DTYPE.get(name, np.float64))
        _lib.call("mw_sim_get_field", self._h, f, out.ctypes.data, n)
        return out


class Output:
    """The reference's output files of a Simulation (mw_output_*): time series
    (vrms.ascii, Nu.ascii, dt.ascii, ...) every step, solution_NNNN.vtu every
    ``every_solution`` steps and SWARM/swarm_NNNN.vtu every ``every_swarm``
    steps (0: never), solution.pvd and swarm.pvd collections; written by a
    native thread from pinned snapshots, off the stepping path.  Call
    ``write()`` after each ``sim.step()``; ``close()`` before closing the
    simulation."""

    def __init__(self, sim, folder, every_solution=1, every_swarm=1, material_names=None):
        self.sim = sim
        self._h = ctypes.c_void_p()
        names = None if material_names is None else ",".join(material_names).encode()
        _lib.call("mw_output_create", sim._h, str(folder).encode(), int(every_solution),
                  int(every_swarm), names, ctypes.byref(self._h))

    def write(self, last=False):
        if getattr(self.sim, "_last_stats", None) is None:
            raise RuntimeError("Output.write: no step taken yet")
        _lib.call("mw_output_step", self._h, self.sim._last_stats, int(bool(last)))

    def flush(self):
        _lib.call("mw_output_flush", self._h)

    def close(self):
        if self._h:
            h, self._h = self._h, ctypes.c_void_p()
            _lib.call("mw_output_destroy", h)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Independent timesteps stepped together (mw_batch_*): instance i is
    ``batch.sims[i]``, a Simulation whose setup (temperature, particles) and
    field reads work as for a lone one; ``step(k)`` advances every instance
    by k timesteps with ``threads`` host threads sharing the instances, so
    their kernels overlap on the device.  An instance's results equal those
    of the same configuration run alone."""

    def __init__(self, configs, threads=8):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("meeuuw_amd.timestep needs a ROCm GPU (no CPU fallback)")
        self.n = len(configs)
        arr = (SimConfig * self.n)(*configs)
        self._h = ctypes.c_void_p()
        _lib.call("mw_batch_create", self.n, arr, int(threads), ctypes.byref(self._h))
        self.sims = []
        for i, c in enumerate(configs):
            h = ctypes.c_void_p()
            _lib.call("mw_batch_sim", self._h, i, ctypes.byref(h))
            self.sims.append(Simulation(c, _handle=h))

    def step(self, steps=1):
        """Advance every instance; returns each instance's last-step statistics."""
        buf = (ctypes.c_double * (NSTATS * self.n))()
        status = (ctypes.c_int * self.n)()
        rc = _lib.lib.mw_batch_step(self._h, int(steps), buf, status)
        out = []
        for i, sim in enumerate(self.sims):
            for j, k in enumerate(STATS):
                sim.stats[k] = buf[i * NSTATS + j]
            out.append(dict(sim.stats))
        _lib.check(rc)
        return out

    def close(self):
        for sim in self.sims:
            sim.close()
        if self._h:
            _lib.call("mw_batch_destroy", self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def coarsest_dofs(nelx, nelz):
    """Unknowns of the coarsest multigrid level of an nelx x nelz box (the
    levels halve while both counts are even and at least 4)."""
    while nelx % 2 == 0 and nelz % 2 == 0 and nelx >= 4 and nelz >= 4:
        nelx //= 2
        nelz //= 2
    return 2 * (2 * nelx + 1) * (2 * nelz + 1)


def batch_configs(n=64, seed=20261203, nelx_choices=(32, 48, 64, 96, 128), aspects=(1, 2, 3, 4),
                  Ra=(1e4, 1e7), contrast=(1.0, 1e4), ppe=16, **kw):
    """Configurations of a batch of independent instances drawn from one seed:
    nelx from nelx_choices, aspect ratio Lx/Lz from aspects (Lz = 1), Ra and
    the viscosity contrast log-uniform in their ranges.  nelz is nelx/aspect
    rounded to a multiple of 4, or of 8, 16, 32 where that is needed for the
    coarsest multigrid level to stay within 256 unknowns (its dense solve)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(n):
        nelx = int(rng.choice(nelx_choices))
        aspect = int(rng.choice(aspects))
        ra = float(10.0 ** rng.uniform(np.log10(Ra[0]), np.log10(Ra[1])))
        ct = float(10.0 ** rng.uniform(np.log10(contrast[0]), np.log10(contrast[1])))
        for m in (4, 8, 16, 32):
            nelz = max(m, m * int(round(nelx / aspect / m)))
            if coarsest_dofs(nelx, nelz) <= 256:
                break
        cfg = make_config(nelx=nelx, nelz=nelz, Lx=float(aspect), Lz=1.0, Ra=ra, rk=4, ppe=ppe,
                          eta_contrast=ct, eta_min=1.0 / ct, eta_max=1.0, **kw)
        out.append(cfg)
    return out
