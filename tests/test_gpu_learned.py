"""Learned Schur-complement correction (solver "learned") on the GPU.

The network is trained on the first Stokes solve and judged by the classical
preconditioner: FGMRES on two probe systems with and without it (used only if
it saves iterations and time).  test_learned_solver_judge_and_fields checks
that training ran, that the judge's decision follows its iteration counts and
probe times, and that the fields match the
mixed solver's (on the Frank-Kamenetskii field the judge rejects the network).
test_learned_gain_on_two_material_field measures what the trained network
gains over the method without it on the two-material blob field, where the
classical Schur-complement block is weakest."""
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


def test_learned_solver_judge_and_fields():
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
    assert s0["learn_probe_ms_plain"] > 0 and s0["learn_probe_ms"] > 0
    used = (s0["learn_its"] < 0.9 * s0["learn_its_plain"] and
            s0["learn_probe_ms"] < 0.95 * s0["learn_probe_ms_plain"])
    assert all(s["learned"] == (1.0 if used else 0.0) for s in st_l[1:])
    for s in st_l + st_m:
        assert s["stokes_relres"] <= 1e-8
    for k in ("u", "w", "p", "T"):
        a, b = f_l[k], f_m[k]
        assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(b), k


def test_learned_gain_on_two_material_field():
    """256x64 elements of the two-material workload (aspect 4, 36 particles per
    element, blobs of viscosity ratio 1e2 and density contrast 1 %): the judge
    accepts the network, and the learned timestep needs fewer than 0.75 times
    the FGMRES iterations of the mixed timestep on the steps after training,
    both to the same relative residual 1e-8 (measured: 393 -> 180-202)."""
    import torch
    from meeuuw_amd import timestep as TS
    from meeuuw_amd import workloads as W
    nelx, nelz = 256, 64
    r, t, x, z = W.box_particles(nelx, nelz, 4.0, 1.0, 36, seed=20261203)
    mat = W.blob_materials(x, z, 4.0, 1.0)
    its, fields = {}, {}
    for solver in ("mixed", "learned"):
        cfg = W.two_material_config(nelx=nelx, nelz=nelz, ppe=36, solver=solver,
                                    stokes_tol=1e-8, eta_ratio=1e2, fgmres_restart=64,
                                    fgmres_maxit=2000)
        sim = TS.Simulation(cfg)
        sim.set_swarm_reduced(r, t)
        sim.set_swarm_material(mat)
        st = [sim.step() for _ in range(3)]
        torch.cuda.synchronize()
        print(solver, [(s["stokes_vcycles"], s["learned"]) for s in st], st[0]["learn_ratio"],
              st[0]["learn_its_plain"], st[0]["learn_its"], flush=True)
        assert all(s["stokes_relres"] <= 1e-8 for s in st)
        its[solver] = [s["stokes_vcycles"] for s in st[1:]]
        fields[solver] = sim.field("p")
        if solver == "learned":
            assert st[0]["learn_ratio"] < 1.0  # held-out Schur-block loss below the classical one
            assert all(s["learned"] == 1.0 for s in st[1:])
        sim.close()
    assert sum(its["learned"]) < 0.75 * sum(its["mixed"])
    a, b = fields["learned"], fields["mixed"]
    assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b)
