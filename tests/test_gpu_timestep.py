"""GPU timestep (mw_sim_step) against the CPU oracle's timestep.

The oracle solves the Stokes and energy systems directly (SciPy SuperLU, as
the reference does); the GPU solves them iteratively to a relative residual
of 1e-12 here, so fields agree to the tolerance written in each assertion
(relative L2 1e-8).  Pressure is compared after the
surface normalisation, which removes the solver-dependent constant.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _pair(oracle, solver, nel, rk, ppe=16, contrast=1.0, Ra=1e4, seed=7, tol=1e-12):
    from oracle import step as S
    from meeuuw_amd import timestep as TS
    P = S.Params(nelx=nel, nelz=nel, Ra=Ra, RKorder=rk, ppe=ppe, eta_contrast=contrast, seed=seed)
    ost = S.State(P)
    cfg = TS.make_config(nelx=nel, nelz=nel, Ra=Ra, rk=rk, ppe=ppe, eta_contrast=contrast,
                         solver=solver, stokes_tol=tol, energy_tol=1e-13)
    sim = TS.Simulation(cfg)
    sim.set_temperature(ost.T)
    if rk != -1:
        sim.set_swarm_reduced(ost.swarm_r, ost.swarm_t)
    return ost, sim


@pytest.mark.parametrize("solver", ["plain", "mixed"])
@pytest.mark.parametrize("n", [16, 32])
def test_blankenbach_step0(oracle, ref_results, solver, n):
    """Blankenbach 1a, particles on the qpoints: the GPU step reproduces the
    reference's published step-0 vrms, Nu and dt (RESULTS/exp00/case1a).
    Solved to 1e-13: at a relative residual of 1e-12 the velocity error of an
    iterative solve can reach the 1e-8 bar (cond(A) ~ 1e4 here), depending on
    the path the solver took."""
    ost, sim = _pair(oracle, solver, n, -1, tol=1e-13)
    st = sim.step()
    ref = ref_results["blankenbach_1a"][str(n)]
    print(solver, n, st)
    assert abs(st["vrms"] - ref["vrms_step0"]) <= 5.01e-7
    assert abs(st["Nu"] - ref["Nu_step0"]) <= 5.01e-7
    assert st["dt"] == pytest.approx(ref["time_after_step0"], rel=1e-6)
    assert st["stokes_relres"] <= 1e-13
    from oracle import step as S
    o = S.step(ost)
    assert rel(sim.field("u"), o["u"]) < 1e-8
    assert rel(sim.field("w"), o["w"]) < 1e-8
    assert rel(sim.field("p"), o["p"]) < 1e-8
    assert rel(sim.field("T"), o["T"]) < 1e-8
    assert st["vrms"] == pytest.approx(o["vrms"], rel=1e-9)
    assert st["Nu"] == pytest.approx(o["Nu"], rel=1e-9)


@pytest.mark.parametrize("solver", ["plain", "mixed"])
def test_random_particles_rk2(oracle, solver):
    """Three steps, 16 random particles per element, RK2, isoviscous.  Solved
    to 1e-13: the flow is ~1 % of the hydrostatic right-hand side, so a relative
    residual r leaves a velocity error of up to ~2e4 r (measured: 1.7e-8 at
    r = 7.8e-13), and the 1e-8 bar must not depend on where the solver stops."""
    from oracle import step as S
    ost, sim = _pair(oracle, solver, 16, 2, tol=1e-13)
    for k in range(3):
        st = sim.step()
        o = S.step(ost)
        print(solver, k, st, o["vrms"], o["Nu"])
        assert np.array_equal(sim.field("rho_e"), o["rho_e"]) or rel(sim.field("rho_e"), o["rho_e"]) < 1e-13
        for f in ("u", "w", "p", "T"):
            assert rel(sim.field(f), o[f]) < 1e-8, f
        assert st["dt"] == pytest.approx(o["dt"], rel=1e-9)
        # positions move by dt*u with u from two different solvers (1e-12
        # residual): they agree to dt * |du| ~ 1e-3 * 1e-8
        np.testing.assert_allclose(sim.field("swarm_x"), ost.swarm_x, rtol=0, atol=1e-9)
        np.testing.assert_allclose(sim.field("swarm_z"), ost.swarm_z, rtol=0, atol=1e-9)
        assert np.mean(sim.field("swarm_iel") != ost.swarm_iel) < 1e-3


@pytest.mark.parametrize("solver", ["plain", "mixed"])
def test_plain_cpu_case_ten_steps(oracle, solver):
    """The plain CPU case over its full length: 32x32, isoviscous, Ra 1e4, 16
    random particles per element, RK2, Crank-Nicolson energy, 10 steps, every
    field within the 1e-8 bar of the oracle (direct solves) after every step
    and the time series (vrms, Nu, dt) to 1e-8."""
    from oracle import step as S
    ost, sim = _pair(oracle, solver, 32, 2, tol=1e-13)
    for k in range(10):
        st = sim.step()
        o = S.step(ost)
        for f in ("u", "w", "p", "T"):
            assert rel(sim.field(f), o[f]) < 1e-8, (k, f, rel(sim.field(f), o[f]))
        assert st["dt"] == pytest.approx(o["dt"], rel=1e-8)
        assert st["vrms"] == pytest.approx(o["vrms"], rel=1e-8)
        assert st["Nu"] == pytest.approx(o["Nu"], rel=1e-8)
    print(solver, "step 10", st)


@pytest.mark.parametrize("solver", ["plain", "mixed"])
def test_frank_kamenetskii_contrast(oracle, solver):
    """Viscosity contrast 1e5 (eta = exp(-ln(1e5) T) clipped), Ra 1e6, RK4, 32x32, two steps."""
    from oracle import step as S
    # at contrast 1e5 a 1e-12 residual is below what FP64 round-off in K u
    # allows (||A|| ||x|| / ||b|| ~ 1e6), and a relative residual r leaves a
    # velocity error of up to ~500 r: solve to 1e-11, compare fields at 1e-8
    ost, sim = _pair(oracle, solver, 32, 4, ppe=25, contrast=1e5, Ra=1e6, tol=1e-11)
    for k in range(2):
        try:
            st = sim.step()
        finally:
            print(solver, k, sim.stats)
        o = S.step(ost)
        print(solver, k, st, o["vrms"], o["Nu"])
        # eta = exp(-ln(1e5) T): a 1e-10 difference in T (solver tolerance) is
        # amplified 11.5 times; on the first step T is identical
        np.testing.assert_allclose(sim.field("eta_e"), o["eta_e"], rtol=1e-13 if k == 0 else 1e-7)
        for f in ("u", "w", "p", "T"):
            assert rel(sim.field(f), o[f]) < 1e-8, f


def test_fgmres_iteration_graphs_are_bit_identical():
    """Replaying each Arnoldi step of the mixed solver as a captured graph gives
    the same bits as launching its kernels on the stream."""
    from meeuuw_amd import timestep as TS
    out = []
    for graphs in (True, False):
        cfg = TS.make_config(nelx=64, nelz=64, Ra=1e6, eta_contrast=1e5, rk=2, ppe=9,
                             solver="mixed", stokes_tol=1e-8, fgmres_graphs=graphs)
        sim = TS.Simulation(cfg)
        sim.seed_swarm(7)
        st = [sim.step() for _ in range(3)]
        out.append((st, sim.field("u"), sim.field("p"), sim.field("T")))
        sim.close()
    (sa, ua, pa, Ta), (sb, ub, pb, Tb) = out
    assert [s["stokes_vcycles"] for s in sa] == [s["stokes_vcycles"] for s in sb]
    np.testing.assert_array_equal(ua, ub)
    np.testing.assert_array_equal(pa, pb)
    np.testing.assert_array_equal(Ta, Tb)


@pytest.mark.parametrize("solver,nel", [("plain", 64), ("mixed", 64), ("mixed", 32),
                                        ("mixed", 16)])
def test_multigrid_bottom_workgroup_is_bit_identical(solver, nel):
    """The last one, two or three smoothed levels and the dense coarsest solve
    as one workgroup repeat the arithmetic of the per-level kernels:
    bit-identical timestep (64x64: levels 16, 8, 4 fused; 32x32: from the first
    coarse level; 16x16: two levels)."""
    from meeuuw_amd import timestep as TS
    out = []
    for bottom in (3, 2, 1, 0):
        cfg = TS.make_config(nelx=nel, nelz=nel, Ra=1e6, eta_contrast=1e5, rk=2, ppe=9,
                             solver=solver, stokes_tol=1e-8, gmg_bottom=bottom,
                             gmg_coarse_dofs=60)
        sim = TS.Simulation(cfg)
        sim.seed_swarm(7)
        st = [sim.step() for _ in range(2)]
        out.append((st, sim.field("u"), sim.field("p"), sim.field("T")))
        sim.close()
    sb, ub, pb, Tb = out[-1]
    for sa, ua, pa, Ta in out[:-1]:
        assert [s["stokes_vcycles"] for s in sa] == [s["stokes_vcycles"] for s in sb]
        np.testing.assert_array_equal(ua, ub)
        np.testing.assert_array_equal(pa, pb)
        np.testing.assert_array_equal(Ta, Tb)


@pytest.mark.parametrize("solver", ["plain", "mixed"])
def test_fused_multigrid_tail_is_bit_identical(solver):
    """The single-workgroup tail of the V-cycle repeats the per-level kernels'
    arithmetic in the same order: the timestep is bit-identical either way."""
    from meeuuw_amd import timestep as TS
    out = []
    for tail in (True, False):
        cfg = TS.make_config(nelx=64, nelz=64, Ra=1e6, eta_contrast=1e5, rk=2, ppe=9,
                             solver=solver, stokes_tol=1e-8, gmg_tail=tail, gmg_coarse_dofs=60)
        sim = TS.Simulation(cfg)
        sim.seed_swarm(7)
        st = [sim.step() for _ in range(2)]
        out.append((st, sim.field("u"), sim.field("p"), sim.field("T")))
        sim.close()
    (sa, ua, pa, Ta), (sb, ub, pb, Tb) = out
    assert [s["stokes_vcycles"] for s in sa] == [s["stokes_vcycles"] for s in sb]
    np.testing.assert_array_equal(ua, ub)
    np.testing.assert_array_equal(pa, pb)
    np.testing.assert_array_equal(Ta, Tb)


def test_matrix_free_energy_matches_assembled():
    """The matrix-free energy operator (quadrature-point coefficients, fused
    BiCGStab kernels) solves the same Crank-Nicolson system as the assembled
    build_matrix_energy values: with both solved to a relative residual of
    1e-13 the temperatures agree to 1e-11 (relative L2) over three steps, and
    the velocities (Stokes solved to 1e-12, whose velocity error reaches ~1e2
    times the residual at this contrast) to 1e-8."""
    from meeuuw_amd import timestep as TS
    out = []
    for mf in (True, False):
        cfg = TS.make_config(nelx=48, nelz=32, Lx=1.0, Ra=1e6, eta_contrast=1e5, rk=2, ppe=9,
                             solver="mixed", stokes_tol=1e-12, energy_tol=1e-13, energy_mf=mf)
        sim = TS.Simulation(cfg)
        sim.seed_swarm(11)
        st = [sim.step() for _ in range(3)]
        out.append((st, sim.field("T"), sim.field("u")))
        sim.close()
    (sa, Ta, ua), (sb, Tb, ub) = out
    for a, b in zip(sa, sb):
        print(a["energy_its"], a["energy_relres"], b["energy_its"], b["energy_relres"])
        assert a["energy_relres"] <= 1e-12 and b["energy_relres"] <= 1e-12
        assert abs(a["energy_its"] - b["energy_its"]) <= 1
    assert rel(Ta, Tb) < 1e-11
    assert rel(ua, ub) < 1e-8


@pytest.mark.parametrize("nelx,nelz,steps", [(48, 32, 3), (40, 24, 2), (64, 64, 4), (36, 20, 3)])
def test_fused_schur_block_matches_split_launches(nelx, nelz, steps):
    """The preconditioner's Schur block in one launch (Chebyshev steps on the
    pressure mass matrix recomputed on overlapping patches, then the block
    front) gives the bits of the launch-per-step path, on meshes whose node
    counts are not multiples of the 16-node tile."""
    import torch
    from meeuuw_amd import timestep as TS
    cfg = TS.make_config(nelx=nelx, nelz=nelz, Lx=1.5, Ra=1e5, eta_contrast=1e4, rk=2, ppe=9,
                         solver="mixed", schur_cheb=steps)
    sim = TS.Simulation(cfg)
    try:
        sim.seed_swarm(5)
        sim.step()
        nnp = (nelx + 1) * (nelz + 1)
        g = torch.Generator(device="cuda").manual_seed(3)
        x = torch.randn(2 * sim.nn + nnp, device="cuda", generator=g, dtype=torch.float32)
        y1 = torch.full_like(x, float("nan"))
        y2 = torch.full_like(x, float("nan"))
        sim.apply_operator("prec32", x, y1)
        sim.apply_operator("prec32_split", x, y2)
        torch.cuda.synchronize()
        assert torch.isfinite(y1).all()
        assert torch.equal(y1, y2)
    finally:
        sim.close()


@pytest.mark.parametrize("nel,cd", [(8, 600), (4, 200)])
def test_dense_coarse_inverse_solves_the_velocity_block(nel, cd):
    """A mesh whose finest level is the coarsest (at most gmg_coarse_dofs
    unknowns) is solved by its dense inverse alone (blocked Gauss-Jordan: 578
    unknowns in 37 pivot blocks, 162 in 11): the preconditioner's velocity
    block is then exact, so the operator applied to the preconditioned vector
    gives back the input's velocity on every free dof (x_u = K z_u + G z_p with
    K z_u = x_u - G z_p), to FP32 accuracy."""
    import torch
    from meeuuw_amd import timestep as TS
    cfg = TS.make_config(nelx=nel, nelz=nel, Ra=1e5, eta_contrast=1e2, rk=2, ppe=9,
                         solver="mixed", gmg_coarse_dofs=cd)
    sim = TS.Simulation(cfg)
    try:
        sim.seed_swarm(3)
        sim.step()
        nnx = 2 * nel + 1
        nnp = (nel + 1) ** 2
        g = torch.Generator(device="cuda").manual_seed(1)
        x = torch.randn(2 * sim.nn + nnp, device="cuda", generator=g, dtype=torch.float32)
        z = torch.empty_like(x)
        y = torch.empty_like(x)
        sim.apply_operator("prec32", x, z)
        sim.apply_operator("stokes32", z, y)
        torch.cuda.synchronize()
        I, J = np.meshgrid(np.arange(nnx), np.arange(nnx))
        fixu = ((I == 0) | (I == nnx - 1)).ravel()
        fixw = ((J == 0) | (J == nnx - 1)).ravel()
        free = np.stack([~fixu, ~fixw], 1).ravel()
        xu = x[:2 * sim.nn].cpu().numpy()[free]
        yu = y[:2 * sim.nn].cpu().numpy()[free]
        assert np.linalg.norm(yu - xu) / np.linalg.norm(xu) < 2e-4
    finally:
        sim.close()
