"""Schedule autotuner for batches of timesteps.

The solver's schedule (Chebyshev degree of the finest multigrid level, the
size of the coarsest level, how many of the last levels run fused with the
coarsest solve, FGMRES restart length)
changes the cost of a step differently at different mesh sizes: a small mesh
is launch-bound and takes more iterations per step, a large one is
bandwidth-bound.  For each (nelx, nelz) class of a batch, ``tune`` runs a
probe instance of that class under every candidate schedule (one warm-up
step, then timed steps, on the device) and keeps the fastest one that met the
Stokes tolerance; ``tuned_configs`` applies the choices to the whole batch.
The schedules change the solver's path only: every candidate solves to the
configured tolerance.
"""
import ctypes
import itertools
import time

from . import timestep as TS

# degree 3 on the finest level only (the coarse levels keep 2, which the fused
# multigrid bottom needs); coarsest level at <= 600 unknowns (dense inverse over
# the whole GPU) or <= 60 (levels down to 2x2 elements, the last ones fused with
# the coarsest solve in one workgroup)
DEFAULT_SPACE = {
    "cheb_degree": (2, 3),
    "cheb_degree_coarse": (2,),
    "gmg_bottom": (3, 1),
    "gmg_coarse_dofs": (600, 60),
    "fgmres_restart": (20, 32),
}


def _redundant(kw):
    """gmg_bottom acts only on coarsest levels of at most 64 unknowns, so with a
    larger gmg_coarse_dofs its values are one schedule."""
    return (kw.get("gmg_coarse_dofs", 60) > 64 and "gmg_bottom" in kw and
            kw["gmg_bottom"] != DEFAULT_SPACE["gmg_bottom"][0])


def _with(cfg, **kw):
    c = TS.SimConfig()
    ctypes.memmove(ctypes.addressof(c), ctypes.addressof(cfg), ctypes.sizeof(cfg))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def _time(cfg, steps, seed):
    import torch
    sim = TS.Simulation(cfg)
    try:
        sim.seed_swarm(seed)
        sim.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            st = sim.step()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / steps * 1e3
        ok = st["stokes_relres"] <= cfg.stokes_tol
    finally:
        sim.close()
    return ms, ok


def tune(cfg, space=None, steps=2, seed=7):
    """The fastest schedule for cfg's mesh class: (settings dict, ms per
    step, {settings tuple: ms} of every candidate)."""
    space = space or DEFAULT_SPACE
    keys = sorted(space)
    trials = {}
    best = None
    for values in itertools.product(*(space[k] for k in keys)):
        kw = dict(zip(keys, values))
        if _redundant(kw):
            continue
        ms, ok = _time(_with(cfg, **kw), steps, seed)
        trials[values] = ms
        if ok and (best is None or ms < best[1]):
            best = (kw, ms)
    if best is None:
        raise RuntimeError("autotune: no candidate schedule met the Stokes tolerance")
    return best[0], best[1], trials


def tuned_configs(configs, space=None, steps=2, seed=7):
    """The batch's configurations with each mesh class's fastest schedule
    (one probe per class: the first configuration of that class), and the
    per-class choices {(nelx, nelz): settings}."""
    choice = {}
    out = []
    for c in configs:
        key = (c.nelx, c.nelz)
        if key not in choice:
            choice[key] = tune(c, space, steps, seed)[0]
        out.append(_with(c, **choice[key]))
    return out, choice
