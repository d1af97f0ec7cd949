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
