"""Failure reporting and step statistics of the device-resident timestep.

A NaN temperature makes every solve fail: the step must return a non-zero
status (no NaN residual may pass a convergence test).  The particle-count
statistics of a fresh, uniformly seeded swarm are known exactly: every element
holds ppe particles and all of them are active (meeuuw.py:1194-1202).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _sim(solver, nel=16, ppe=9):
    from meeuuw_amd import timestep as TS
    cfg = TS.make_config(nelx=nel, nelz=nel, ppe=ppe, rk=2, solver=solver, stokes_tol=1e-8)
    sim = TS.Simulation(cfg)
    sim.seed_swarm(5)
    x = np.linspace(0, 1, 2 * nel + 1)
    X, Z = np.meshgrid(x, x)
    return sim, (1 - Z) + 0.01 * np.cos(np.pi * X) * np.sin(np.pi * Z)


@pytest.mark.parametrize("solver", ["plain", "mixed"])
def test_nan_temperature_fails_the_step(solver):
    from meeuuw_amd._lib import MeeuuwHipError
    sim, T = _sim(solver)
    T = T.ravel().copy()
    T[T.size // 2] = np.nan
    sim.set_temperature(T)
    print(f"nan temperature, solver {solver}", flush=True)
    with pytest.raises(MeeuuwHipError) as err:
        sim.step()
    print(err.value, flush=True)
    assert err.value.code == 5  # MW_ENOCONV
    sim.close()


@pytest.mark.parametrize("solver", ["plain", "mixed"])
def test_particle_count_statistics(solver):
    nel, ppe = 16, 9
    sim, T = _sim(solver, nel, ppe)
    sim.set_temperature(T.ravel())
    st = sim.step()
    assert st["mincount"] == ppe
    assert st["nactive"] == nel * nel * ppe
    active = int(sim.field("swarm_active").sum())  # after one advection
    st = sim.step()  # counts the swarm as the first step left it
    assert 1 <= st["mincount"] <= ppe
    assert st["nactive"] == active
    sim.close()
