"""Visco-plastic convection of Tosi et al. (2015), experiment3.py, on the
device timestep (rheology 1), with and without Picard iterations.

Parity: against the oracle (oracle/step.py: tosi_viscosity, the nodal strain
rate, the Picard state machine of meeuuw.py:2084-2101, 3409-3415) step by step
at 16x16, and against the reference's own case-2 run at 32x32 to t = 0.25
(RESULTS/exp03/case2/vrms.ascii, Nu.ascii, extracted into
tests/golden/reference_results.json).
"""
import json
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SIGMA_Y = {1: 1.0, 2: 1.0, 3: 1.0, 4: 1.0, 5: 4.0}  # experiment3.py:8-28
GAMMA_Y = {1: 1.0, 2: 1.0, 3: 10.0, 4: 10.0, 5: 10.0}


def _params(case):
    return dict(gamma_T=math.log(1e5), gamma_y=math.log(GAMMA_Y[case]), sigma_y=SIGMA_Y[case],
                eta_star=1e-3)


def _oracle(case, n, **kw):
    from oracle import step as S
    P = S.Params(nelx=n, nelz=n, Ra=1e2, alphaT=1e-4, CFL=0.75, RKorder=-1, rheology=1,
                 tosi_case=case, **_params(case), **kw)
    return S, S.State(P)


def _sim(case, n, solver, tol, ost=None, **kw):
    from meeuuw_amd import timestep as TS
    p = _params(case)
    cfg = TS.make_config(nelx=n, nelz=n, rk=-1, Ra=1e2, alphaT=1e-4, CFL=0.75, rheology=1,
                         tosi_case=case, tosi_gamma_T=p["gamma_T"], tosi_gamma_y=p["gamma_y"],
                         tosi_sigma_y=p["sigma_y"], tosi_eta_star=p["eta_star"], solver=solver,
                         stokes_tol=tol, energy_tol=1e-13, pcg_maxit=5000, fgmres_maxit=2000, **kw)
    sim = TS.Simulation(cfg)
    if ost is not None:
        sim.set_temperature(ost.T)
    return sim


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


# The flow is driven by the alphaT = 1e-4 thermal part of a hydrostatic
# right-hand side (g_z = -Ra / alphaT = -1e6) through a viscosity contrast of
# 1e5: the relative residual cannot go much below 1e-11 in FP64 here (1e-12
# and 1e-13 stall for the plastic cases), and 1e-11 leaves velocity errors of
# ~1e-6 relative.  Velocities are compared to `vel`, pressure and temperature
# 100x tighter; a wrong law, strain rate or Picard step moves them by 1e-2 or more.
@pytest.mark.parametrize("case,solver,tol,vel", [(2, "plain", 1e-11, 1e-5),
                                                 (2, "mixed", 1e-11, 1e-5),
                                                 (4, "mixed", 1e-11, 1e-5),
                                                 (5, "mixed", 1e-11, 1e-5),
                                                 (1, "mixed", 1e-11, 1e-5),
                                                 (3, "mixed", 1e-11, 1e-5)])
def test_tosi_matches_oracle(case, solver, tol, vel):
    S, ost = _oracle(case, 16)
    sim = _sim(case, 16, solver, tol, ost)
    print(f"tosi case {case} 16x16 {solver}", flush=True)
    for k in range(4):
        o = S.step(ost)
        s = sim.step()
        for key in ("vrms", "Nu", "dt"):
            assert abs(s[key] - o[key]) <= vel * abs(o[key]), (k, key, s[key], o[key])
        assert s["nl_its"] == 1
    errs = {f: _rel(sim.field(f), ref) for f, ref in (("u", o["u"]), ("w", o["w"]),
                                                       ("p", o["p"]), ("T", ost.T))}
    print("field errors", errs, flush=True)
    assert _rel(sim.field("T"), ost.T) < 1e-2 * vel
    for f, ref in (("u", o["u"]), ("w", o["w"])):
        assert _rel(sim.field(f), ref) < vel, f
    assert _rel(sim.field("p"), o["p"]) < 1e-2 * vel
    for f, ref in (("eta_e", o["eta_e"]), ("eta_n", o["eta_n"])):
        assert _rel(sim.field(f), ref) < vel, f
    if case in (2, 4, 5):
        for f, ref in (("exx_n", ost.exx_n), ("ezz_n", ost.ezz_n), ("exz_n", ost.exz_n)):
            assert _rel(sim.field(f), ref) < vel, f
    sim.close()


@pytest.mark.parametrize("case,tol_nl", [(2, 1e-3), (4, 5e-2)])
def test_picard_iterations_match_oracle(case, tol_nl, tmp_path):
    """Picard iterations: the relative changes of each iteration and the
    iteration count (assess_steady_state.py:34-66), the conv.ascii lines
    (assess_steady_state.py:54-56), then the fields."""
    from meeuuw_amd import timestep as TS
    S, ost = _oracle(case, 16, nonlinear=True, niter_nl=5, tol_nl=tol_nl)
    sim = _sim(case, 16, "mixed", 1e-11, ost, nonlinear=True, niter_nl=5, tol_nl=tol_nl)
    out = TS.Output(sim, tmp_path, every_solution=0, every_swarm=0)
    print(f"picard case {case} tol_nl {tol_nl}", flush=True)
    rows = []
    for k in range(3):
        o = S.step(ost)
        s = sim.step()
        out.write()
        rows += [(k + i / 200, *o["nl_history"][i], tol_nl, i) for i in range(o["nl_its"])]
        h = sim.field("nl_history").reshape(-1, 4)
        print(k, s["nl_its"], o["nl_its"], h[-1], o["nl_history"][-1], flush=True)
        assert s["nl_its"] == o["nl_its"] and h.shape == o["nl_history"].shape
        # xi_u and xi_w carry the solves' velocity error (~1e-6 relative);
        # xi_T (temperature) is tight; xi_p is the change of a nearly
        # hydrostatic pressure (~1e-8 relative), at the solves' accuracy
        np.testing.assert_allclose(h[:, :2], o["nl_history"][:, :2], rtol=1e-3, atol=2e-5)
        np.testing.assert_allclose(h[:, 3], o["nl_history"][:, 3], rtol=1e-5)
        np.testing.assert_allclose(h[:, 2], o["nl_history"][:, 2], atol=1e-6)
        assert abs(s["vrms"] - o["vrms"]) <= 1e-5 * o["vrms"]
        assert abs(s["dt"] - o["dt"]) <= 1e-5 * o["dt"]
        assert abs(s["time"] - o["time"]) <= 1e-12 + 1e-5 * o["time"]
    # three steps of five iterations carry the velocity error into T
    assert _rel(sim.field("T"), ost.T) < 1e-6
    assert _rel(sim.field("u"), o["u"]) < 1e-4
    out.close()
    with open(tmp_path / "conv.ascii") as f:
        lines = f.read().splitlines()
    assert len(lines) == len(rows)
    for line, r in zip(lines, rows):
        c = line.split()
        assert len(c) == 7 and c[0] == "%.3e" % r[0] and c[5] == "%.3e" % tol_nl
        assert c[6] == "%d" % r[6]
        np.testing.assert_allclose([float(v) for v in c[1:5]], r[1:5], rtol=2e-3, atol=2e-5)
    sim.close()


def test_tosi_case2_matches_reference_run():
    """Case 2 at 32x32 from t = 0 to 0.25 against the reference's own run
    (RESULTS/exp03/case2): its first rows (the step-0 Stokes solve, then the
    first steps with the lagged strain rate) and the mobile-lid steady state."""
    with open(os.path.join(HERE, "golden", "reference_results.json")) as f:
        g = json.load(f)["tosi"]["case2_32"]
    vr = np.array(g["vrms"]["head"])
    sim = _sim(2, 32, "mixed", 1e-10)
    from oracle import step as S
    from oracle import oracle as O
    sim.set_temperature(S.initial_temperature(O.Mesh(32, 32, 1.0, 1.0)))
    hist = []
    while True:
        s = sim.step()
        hist.append((s["time"] + s["dt"], s["vrms"], s["Nu"]))  # rows hold the time after the step
        if hist[-1][0] >= 0.25 or len(hist) > 20000:
            break
    h = np.array(hist)
    print("first rows", h[:4], vr[:4], flush=True)
    print("steps", len(h), g["vrms"]["rows"], "last", h[-1], g["vrms"]["tail"][-1],
          g["Nu"]["tail"][-1], flush=True)
    # step 0: the same Stokes problem (strain rate zero, eta = min(2, 2 eta_lin)) and dt
    assert abs(h[0, 0] - vr[0, 0]) <= 1e-6 * vr[0, 0]
    assert abs(h[0, 1] - vr[0, 1]) <= 1e-5 * vr[0, 1]
    # the next rows use the nodal strain rate of the previous step; the solve's
    # relative residual 1e-10 leaves ~3e-5 in the velocity
    np.testing.assert_allclose(h[1:10, 0], vr[1:10, 0], rtol=1e-4)
    np.testing.assert_allclose(h[1:10, 1], vr[1:10, 1], rtol=1e-4)
    # the same number of steps to t = 0.25, and the mobile-lid state there
    assert abs(len(h) - g["vrms"]["rows"]) <= 2
    tail = np.array(g["vrms"]["tail"])
    np.testing.assert_allclose(h[-1, 0], tail[-1, 0], rtol=1e-4)
    np.testing.assert_allclose(h[-1, 1], tail[-1, 1], rtol=1e-4)
    np.testing.assert_allclose(h[-1, 2], g["Nu"]["tail"][-1][1], rtol=1e-4)
    sim.close()
