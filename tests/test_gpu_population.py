"""Population control (pic_functions.py:1623-1825) in the device timestep,
against the oracle's restatement on the same inputs and the same draws.

The swarm starts with a seeded 30 % of its particles inactive, so that most
elements fall below nparticle_min and take particles from the inactive pool:
which particles are reused (increasing index, flagged elements in increasing
order), their element, reduced coordinates and position, and the material
copied from the nearest active particle of the element must all agree.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(nelx, nelz, ppe, nmin, seed):
    from oracle import step as S
    P = S.Params(nelx=nelx, nelz=nelz, Ra=1e4, ppe=ppe, RKorder=2, seed=seed,
                 material_mode=1, mat_eta=(1.0, 2.0), mat_rho=(1.0, 1.0),
                 population_control=True, nparticle_min=nmin, pc_seed=11)
    ost = S.State(P)
    rng = np.random.Generator(np.random.PCG64(seed + 1))
    off = rng.random(ost.nparticle) < 0.3
    ost.swarm_active[off] = False
    ost.swarm_x[off] = 0.0
    ost.swarm_z[off] = 0.0
    ost.swarm_mat = (ost.swarm_x > 0.5).astype(np.uint8)
    return S, ost


@pytest.mark.parametrize("solver", ["mixed", "plain"])
def test_population_control_matches_oracle(solver, tmp_path):
    from meeuuw_amd import timestep as TS
    nelx, nelz, ppe, nmin = 8, 8, 16, 14
    S, ost = _setup(nelx, nelz, ppe, nmin, 23)
    cfg = TS.make_config(nelx=nelx, nelz=nelz, ppe=ppe, rk=2, Ra=1e4, solver=solver,
                         stokes_tol=1e-11 if solver == "plain" else 1e-12, energy_tol=1e-13,
                         material_mode=1, mat_eta=(1.0, 2.0), mat_rho=(1.0, 1.0),
                         population_control=True, nparticle_min=nmin, pc_seed=11)
    sim = TS.Simulation(cfg)
    sim.set_temperature(ost.T)
    sim.set_swarm(ost.swarm_x, ost.swarm_z, ost.swarm_r, ost.swarm_t, ost.swarm_iel,
                  ost.swarm_active)
    sim.set_swarm_material(ost.swarm_mat)
    out = TS.Output(sim, tmp_path, every_solution=0, every_swarm=0)
    print(f"population control {solver}: {int((~ost.swarm_active).sum())} inactive of "
          f"{ost.nparticle}", flush=True)
    rows = []
    for k in range(3):
        before = ost.swarm_active.copy()
        S.step(ost)
        s = sim.step()
        out.write()
        rows.append("%d %d %d %d\n" % ((k,) + tuple(int(v) for v in (ost.pc_stats[2], ost.pc_stats[0],
                                                                    ost.pc_stats[1]))))
        need, avail, act = ost.pc_stats
        print(k, (s["pc_needed"], s["pc_available"], s["pc_active"]), ost.pc_stats, flush=True)
        assert (s["pc_needed"], s["pc_available"], s["pc_active"]) == (need, avail, act)
        a = sim.field("swarm_active").astype(bool)
        assert np.array_equal(a, ost.swarm_active)
        assert np.array_equal(sim.field("swarm_iel")[a], ost.swarm_iel[a])
        assert np.array_equal(sim.field("swarm_mat")[a], ost.swarm_mat[a])
        new = a & ~before  # placed by this step's population control: drawn, not advected
        for f, ref in (("swarm_x", ost.swarm_x), ("swarm_z", ost.swarm_z),
                       ("swarm_r", ost.swarm_r), ("swarm_t", ost.swarm_t)):
            v = sim.field(f)
            assert np.array_equal(v[new], ref[new]), f
            # the others moved with velocities known to the solve's accuracy
            np.testing.assert_allclose(v[a], ref[a], rtol=0, atol=1e-7, err_msg=f)
        assert new.sum() == need
    out.close()
    # active_particles_number.ascii: "%d %d %d %d" % (istep, active, needed, available)
    with open(tmp_path / "active_particles_number.ascii") as f:
        assert f.read() == "#time active nparticle_needed nparticle_available )\n" + "".join(rows)
    sim.close()


def test_population_control_reports_shortage():
    """More particles needed than inactive ones left: the step fails as the
    reference does (pic_functions.py:1742-1750), on the GPU and in the oracle."""
    from meeuuw_amd import timestep as TS
    nelx, nelz, ppe = 4, 4, 8
    S, ost = _setup(nelx, nelz, ppe, ppe, 5)
    ost.swarm_active[:] = True
    ost.swarm_active[0] = False  # element 0 needs one particle, one is available: fine
    cfg = TS.make_config(nelx=nelx, nelz=nelz, ppe=ppe, rk=2, Ra=1e4, solver="mixed",
                         population_control=True, nparticle_min=ppe, pc_seed=1)
    sim = TS.Simulation(cfg)
    sim.set_temperature(ost.T)
    sim.set_swarm(ost.swarm_x, ost.swarm_z, ost.swarm_r, ost.swarm_t, ost.swarm_iel,
                  ost.swarm_active)
    s = sim.step()
    assert (s["pc_needed"], s["pc_available"]) == (1, 1)
    assert sim.field("swarm_active").all()
    sim.close()
    # every element one short, one particle available
    cfg = TS.make_config(nelx=nelx, nelz=nelz, ppe=ppe, rk=2, Ra=1e4, solver="mixed",
                         population_control=True, nparticle_min=ppe + 1, pc_seed=1)
    sim = TS.Simulation(cfg)
    sim.set_temperature(ost.T)
    ost.swarm_active[0] = False
    sim.set_swarm(ost.swarm_x, ost.swarm_z, ost.swarm_r, ost.swarm_t, ost.swarm_iel,
                  ost.swarm_active)
    with pytest.raises(SystemExit, match="not available"):
        sim.step()
    sim.close()
    with pytest.raises(ValueError, match="not available"):
        S.population_control(ost, np.full(nelx * nelz, ppe - 2))
