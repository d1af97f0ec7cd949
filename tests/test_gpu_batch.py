"""Batch of independent timesteps (mw_batch_*): every instance of a batch of
uneven sizes, aspect ratios, Rayleigh numbers and viscosity contrasts, stepped
with several host threads sharing the instances, gives the same bits as the
same configuration stepped alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_batch_instances_match_lone_runs():
    from meeuuw_amd import timestep as TS
    cfgs = TS.batch_configs(n=5, seed=3, nelx_choices=(16, 32, 48), solver="mixed",
                            stokes_tol=1e-8)
    assert len({(c.nelx, c.nelz) for c in cfgs}) > 1
    batch = TS.Batch(cfgs, threads=3)
    for i, sim in enumerate(batch.sims):
        sim.seed_swarm(100 + i)
    st_b = [batch.step(1) for _ in range(2)]
    fields_b = [(s.field("u"), s.field("p"), s.field("T")) for s in batch.sims]
    batch.close()
    for i, c in enumerate(cfgs):
        sim = TS.Simulation(c)
        sim.seed_swarm(100 + i)
        st = [sim.step() for _ in range(2)]
        u, p, T = sim.field("u"), sim.field("p"), sim.field("T")
        sim.close()
        print(i, c.nelx, c.nelz, st[-1]["stokes_vcycles"], st_b[-1][i]["stokes_vcycles"])
        for k in range(2):
            assert st[k]["vrms"] == st_b[k][i]["vrms"]
            assert st[k]["stokes_vcycles"] == st_b[k][i]["stokes_vcycles"]
        np.testing.assert_array_equal(u, fields_b[i][0])
        np.testing.assert_array_equal(p, fields_b[i][1])
        np.testing.assert_array_equal(T, fields_b[i][2])


def test_batch_reports_the_failing_instance():
    """An instance whose particles were never set fails; the batch says which."""
    from meeuuw_amd import timestep as TS
    cfgs = TS.batch_configs(n=2, seed=5, nelx_choices=(16,))
    batch = TS.Batch(cfgs, threads=2)
    batch.sims[0].seed_swarm(1)
    with pytest.raises(Exception, match="instance 1"):
        batch.step(1)
    batch.close()


def test_autotuner_picks_a_schedule_that_meets_the_tolerance():
    """tune() times every candidate on a probe of the class and returns the
    fastest that solved to the tolerance; tuned_configs applies it per class."""
    from meeuuw_amd import autotune
    from meeuuw_amd import timestep as TS
    cfgs = TS.batch_configs(n=3, seed=11, nelx_choices=(16, 32))
    space = {"gmg_bottom": (3, 1), "fgmres_restart": (20, 32)}
    kw, ms, trials = autotune.tune(cfgs[0], space, steps=1)
    assert set(kw) == set(space) and len(trials) == 4 and ms == min(trials.values())
    tuned, choice = autotune.tuned_configs(cfgs, space, steps=1)
    assert set(choice) == {(c.nelx, c.nelz) for c in cfgs}
    for c, t in zip(cfgs, tuned):
        assert (t.nelx, t.nelz, t.Ra, t.eta_contrast) == (c.nelx, c.nelz, c.Ra, c.eta_contrast)
        for k, v in choice[(c.nelx, c.nelz)].items():
            assert getattr(t, k) == v
    batch = TS.Batch(tuned, threads=2)
    for i, s in enumerate(batch.sims):
        s.seed_swarm(i)
    st = batch.step(1)
    assert all(x["stokes_relres"] <= 1e-8 for x in st)
    batch.close()


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


# Instances of the batch workload's shape, one per aspect ratio Lx/Lz (Lz = 1),
# at Rayleigh numbers and viscosity contrasts of its range (the contrast is
# exp(-ln(c) T) clipped to [1/c, 1]), RK4, 16 particles per element: each
# against the FP64 oracle (direct solves) over two steps.  The flow is driven
# by the 1e-2 * T part of a right-hand side dominated by the hydrostatic load,
# so a relative residual r leaves a velocity error of up to ~5e4 r: the mild
# contrasts are solved to 1e-13; at contrast 1e4 FP64 round-off in the residual
# itself sits near 1e-13, and 1e-12 already gives fields within 1e-8.
ASPECT_CASES = [(2, 32, 16, 1e7, 1e4, 1e-12), (3, 48, 16, 3e5, 1e2, 1e-13),
                (4, 64, 16, 1e6, 1e3, 1e-13)]


@pytest.mark.parametrize("aspect,nelx,nelz,Ra,ct,tol", ASPECT_CASES)
def test_aspect_ratio_instances_match_oracle(oracle, aspect, nelx, nelz, Ra, ct, tol):
    from oracle import step as S
    from meeuuw_amd import timestep as TS
    P = S.Params(nelx=nelx, nelz=nelz, Lx=float(aspect), Lz=1.0, Ra=Ra, RKorder=4, ppe=16,
                 eta_contrast=ct, eta_min=1.0 / ct, eta_max=1.0, seed=31)
    ost = S.State(P)
    cfg = TS.make_config(nelx=nelx, nelz=nelz, Lx=float(aspect), Lz=1.0, Ra=Ra, rk=4, ppe=16,
                         eta_contrast=ct, eta_min=1.0 / ct, eta_max=1.0, solver="mixed",
                         stokes_tol=tol, energy_tol=1e-13)
    sim = TS.Simulation(cfg)
    sim.set_temperature(ost.T)
    sim.set_swarm_reduced(ost.swarm_r, ost.swarm_t)
    print(f"aspect {aspect}: {nelx}x{nelz} Ra {Ra} contrast {ct}", flush=True)
    for k in range(2):
        st = sim.step()
        o = S.step(ost)
        print(k, st["stokes_relres"], st["vrms"], o["vrms"], flush=True)
        for f in ("u", "w", "p", "T"):
            assert rel(sim.field(f), o[f]) < 1e-8, f
        assert st["dt"] == pytest.approx(o["dt"], rel=1e-9)
        assert st["vrms"] == pytest.approx(o["vrms"], rel=1e-8)
        np.testing.assert_allclose(sim.field("swarm_x"), ost.swarm_x, rtol=0, atol=1e-9)
    sim.close()


def test_full_batch_of_64_instances():
    """The batch workload at its real shape (64 instances, nelx 32-128, aspect
    1-4, Ra 1e4-1e7, contrast 1-1e4, 16 particles per element, RK4): three
    steps with every Stokes solve at relative residual <= 1e-8 and every field
    finite, and three sampled instances bit-identical to their lone runs."""
    from meeuuw_amd import timestep as TS
    cfgs = TS.batch_configs(n=64)
    assert {round(c.Lx) for c in cfgs} == {1, 2, 3, 4}
    batch = TS.Batch(cfgs, threads=8)
    for i, s in enumerate(batch.sims):
        s.seed_swarm(1000 + i)
    print("64-instance batch: 3 steps", flush=True)
    sts = [batch.step(1) for _ in range(3)]
    for st in sts:
        assert all(x["stokes_relres"] <= 1e-8 for x in st)
        assert all(x["energy_relres"] <= 1e-8 for x in st)
        assert all(x["nactive"] > 0 and x["mincount"] >= 1 for x in st)
    picks = [0, 31, 63]
    got = {i: [batch.sims[i].field(f) for f in ("u", "p", "T", "swarm_x")] for i in range(64)}
    batch.close()
    for i in range(64):
        assert all(np.isfinite(a).all() for a in got[i]), i
    for i in picks:
        sim = TS.Simulation(cfgs[i])
        sim.seed_swarm(1000 + i)
        for _ in range(3):
            sim.step()
        for f, a in zip(("u", "p", "T", "swarm_x"), got[i]):
            np.testing.assert_array_equal(sim.field(f), a)
        sim.close()
