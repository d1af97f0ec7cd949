"""Learned Schur-complement correction (solver "learned") on the GPU.

The network is trained on the first Stokes solve and judged by the classical
preconditioner: FGMRES on two probe systems with and without it.  The test
checks that training ran, that the judge's decision follows its iteration
counts, and that the timestep's fields match the mixed solver's (both solve
to the same relative residual).  It prints the measured gain (the iteration
ratio), which is what the network brings over the method without it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run(solver, n, steps):
    from meeuuw_amd import timestep as TS
    cfg = TS.make_config(nelx=n, nelz=n, Ra=1e6, eta_contrast=1e5, rk=4, ppe=9, solver=solver,
                         stokes_tol=1e-8, learn_steps=150, learn_pairs=256)
    sim = TS.Simulation(cfg)
    sim.seed_swarm(20261203)
    stats = [sim.step() for _ in range(steps)]
    out = {k: sim.field(k) for k in ("u", "w", "p", "T")}
    sim.close()
    return stats, out


def test_learned_solver_trains_is_judged_and_matches_mixed():
    n, steps = 48, 4
    st_l, f_l = run("learned", n, steps)
    st_m, f_m = run("mixed", n, steps)
    s0 = st_l[0]
    print("training ms", s0["learn_ms"], "validation loss ratio", s0["learn_ratio"],
          "probe iterations plain", s0["learn_its_plain"], "learned", s0["learn_its"])
    print("stokes iterations learned", [s["stokes_vcycles"] for s in st_l],
          "mixed", [s["stokes_vcycles"] for s in st_m])
    assert s0["learn_ms"] > 0
    assert s0["learn_its_plain"] > 0 and s0["learn_its"] > 0
    used = s0["learn_its"] < 0.9 * s0["learn_its_plain"]
    assert all(s["learned"] == (1.0 if used else 0.0) for s in st_l[1:])
    for s in st_l + st_m:
        assert s["stokes_relres"] <= 1e-8
    for k in ("u", "w", "p", "T"):
        a, b = f_l[k], f_m[k]
        assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(b), k
