"""Benchmark: timesteps/s of the device-resident convection timestep.

Workload (default): 256x256 Q2xQ1 elements on the unit square, Ra = 1e6,
Frank-Kamenetskii viscosity eta = exp(-ln(1e5) T) clipped to [1e-5, 1],
25 particles per element placed uniformly at random from a seed, RK4,
Crank-Nicolson energy, Stokes solved to a relative residual of 1e-8.
Prints ONE JSON line.  `value` is the timesteps/s of the learned-preconditioner
timestep (solver "learned": FP32 multigrid-FGMRES in FP64 defect correction,
with the learned Schur-complement correction wherever the classical
preconditioner judges that it helps; the network is trained on the first
solve, inside the warm-up step).  The mixed-precision timestep without the
network and the plain FP64 Uzawa timestep are timed in the same run and
reported beside it.

CPU baseline: the FP64 CPU oracle (C loops + SciPy SuperLU direct solves, one
thread) runs in a subprocess started first, concurrently with the GPU legs, on
one host core: one timestep of the same 256x256 workload (the metric's unit),
two steps at 64x64 (whose fields also give the GPU path's field errors), and
the 32x32 plain-method case (10 RK2 steps, isoviscous, Ra 1e4).  Whatever has
not finished when the GPU legs end plus --cpu-budget seconds is reported as
missing.  Under "batch": the throughput of 64 independent instances of uneven
size stepped together (mw_batch_*), sequentially, concurrently, and with the
schedules the autotuner picks per mesh class.  Under "two_material": the
mixed and learned timesteps on the two-material blob workload's shape
(512x128, aspect 4, 36 particles per element) at viscosity ratio --tm-ratio.

`python bench.py --kernel-table TRACE.csv` prints, from a rocprofv3
--kernel-trace CSV of a bench run, the achieved GB/s of the top kernels against
their algorithmic bytes (meeuuw_amd/roofline.py) and the bytes per timestep.
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PROF_NAMES = ["stokes_fp64_operator", "cheb_fine", "stokes_elements", "energy_vv", "fgmres_gs",
              "advect_rk4"]
TIMING_REPS = 200
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "cheb_fine_256x256_hbm_traffic.txt")
STATS_FILE = os.path.join(ROOT, "profiles", "timestep_256x256_fk1e5_kernel_stats.txt")
TABLE_FILE = os.path.join(ROOT, "profiles", "timestep_256x256_fk1e5_kernel_table.txt")


def measured_traffic(nel):
    """HBM bytes per launch of the finest-level Chebyshev step from the committed
    counter summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated), for the 256x256
    workload it was measured on."""
    if nel != 256 or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith("traffic_bytes_per_launch"):
            return float(line.split()[1])
    return None


def committed_kernel_table(nel):
    """Per-step algorithmic bytes and the first row (the kernel group with the
    largest share of kernel time) of the committed kernel table of the 256x256
    workload."""
    if nel != 256 or not os.path.exists(TABLE_FILE):
        return None
    out = {"top": None}
    for line in open(TABLE_FILE):
        if line.startswith("# per step:"):
            w = line.split()
            out["bytes_per_step"] = float(w[w.index("bytes") + 1]) * 1e6
            out["covered"] = float(line.split("(")[1].split()[0]) / 100.0
        elif (not line.startswith("#") and line.strip() and out["top"] is None
              and "k_gj_sweep" not in line):  # the sweeps run on a side stream
            w = line.split()
            out["top"] = {"kernel": " ".join(w[:-6]), "avg_us": float(w[-5]),
                          "bytes_per_launch": None if w[-4] == "-" else float(w[-4]),
                          "GBps": None if w[-3] == "-" else float(w[-3]),
                          "frac": None if w[-2] == "-" else float(w[-2]),
                          "share_of_kernel_time": float(w[-1]) / 100.0,
                          "bound": "launch and memory latency (tiles of the coarse levels)"}
    return out if "bytes_per_step" in out else None


def profiled_avg_us(nel):
    """rocprofv3 average duration of the finest-level Chebyshev step from the
    committed kernel-stats summary of the 256x256 workload."""
    if nel != 256 or not os.path.exists(STATS_FILE):
        return None
    for line in open(STATS_FILE):
        if line.startswith("#   rocprofv3 average duration"):
            return float(line.split()[-2])
    return None


def algorithmic_bytes(name, nelx, nelz, ppe):
    """HBM bytes one launch must move at minimum (every input read once, every
    output written once)."""
    nel = nelx * nelz
    nn = (2 * nelx + 1) * (2 * nelz + 1)
    nnP = (nelx + 1) * (nelz + 1)
    NV = 2 * nn
    npart = nel * ppe
    if name == "stokes_fp64_operator":  # matrix-free residual: u, p, f, h, eta_q, mask in; ru, rp out
        return NV * (3 * 8 + 1) + nnP * 3 * 8 + nel * 9 * 8
    if name == "cheb_fine":  # FP32: x, b, 1/diag, d in; d, x out; fix mask; element eta
        return NV * (4 * 6 + 1) + nel * 4
    if name == "energy_vv":
        return nel * 81 * 8 + 2 * nn * 8
    if name == "advect_rk4":  # x, z, active in / x, z, active, u, w out; nodal u, w; icon
        return npart * (17 + 33) + nn * 16 + nel * 36
    return None


def cpu_info():
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


# ------------------------------------------------------------------ CPU leg ---
def cpu_leg(out_path, nel_main, ppe, seed, fields_path):
    """Runs in the subprocess: the oracle legs in order, one JSON line each
    (flushed as it ends)."""
    import numpy as np
    from oracle import step as S

    def emit(d):
        with open(out_path, "a") as f:
            f.write(json.dumps(d) + "\n")

    # 64x64, two steps: timing and the fields the GPU path is compared with
    P = S.Params(nelx=64, nelz=64, Ra=1e6, eta_contrast=1e5, RKorder=4, ppe=ppe, seed=seed)
    st = S.State(P)
    tm, flds = {}, {}
    t0 = time.perf_counter()
    for k in range(2):
        o = S.step(st, tm)
        for f in ("u", "w", "p", "T"):
            flds[f"{f}{k}"] = o[f]
    t = time.perf_counter() - t0
    np.savez(fields_path, **flds)
    emit({"leg": "64x64", "steps": 2, "seconds": t, "stages": tm})
    # the plain method's own case: 32x32, isoviscous, Ra 1e4, RK2, 16 ppe, 10 steps
    P = S.Params(nelx=32, nelz=32, Ra=1e4, eta_contrast=1.0, RKorder=2, ppe=16, seed=seed)
    st = S.State(P)
    tm = {}
    t0 = time.perf_counter()
    for _ in range(10):
        S.step(st, tm)
    emit({"leg": "32x32", "steps": 10, "seconds": time.perf_counter() - t0, "stages": tm})
    # the metric's workload: one step at 256x256
    P = S.Params(nelx=nel_main, nelz=nel_main, Ra=1e6, eta_contrast=1e5, RKorder=4, ppe=ppe,
                 seed=seed)
    st = S.State(P)
    tm = {}
    t0 = time.perf_counter()
    S.step(st, tm)
    emit({"leg": f"{nel_main}x{nel_main}", "steps": 1, "seconds": time.perf_counter() - t0,
          "stages": tm})


def start_cpu_leg(args, tmpdir):
    out = os.path.join(tmpdir, "cpu.jsonl")
    fields = os.path.join(tmpdir, "oracle64.npz")
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-leg", out, "--cpu-fields", fields,
           "--nel", str(args.nel), "--ppe", str(args.ppe), "--seed", str(args.seed)]
    proc = subprocess.Popen(cmd, env=env, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE)
    return proc, out, fields


def read_cpu_legs(path):
    legs = {}
    if os.path.exists(path):
        for line in open(path):
            d = json.loads(line)
            legs[d["leg"]] = d
    return legs


# ------------------------------------------------------------------ GPU legs --
def run_solver(solver, args, steps, warmup, prof_steps=0, nel=None, tol=None, time_kernel=True,
               keep=False):
    import ctypes
    import torch
    from meeuuw_amd import _lib
    from meeuuw_amd import timestep as TS
    n = nel or args.nel
    cfg = TS.make_config(nelx=n, nelz=n, Ra=1e6, eta_contrast=1e5, eta_min=1e-5, eta_max=1.0,
                         rk=4, ppe=args.ppe, solver=solver, stokes_tol=tol or args.tol,
                         energy_tol=1e-10, fgmres_restart=20, cheb_degree=2)
    sim = TS.Simulation(cfg)
    sim.seed_swarm(args.seed)
    first = None
    for _ in range(warmup):
        st = sim.step()
        first = first or st
        print(f"[bench] {solver} {n}x{n} warmup step {st}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = []
    for k in range(steps):
        hist.append(sim.step())
        print(f"[bench] {solver} {n}x{n} step {k} {hist[-1]}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    prof = None
    kernel_us = None
    if solver != "plain" and time_kernel:
        # the finest-level Chebyshev step, timed over back-to-back launches at
        # its real shape with HIP events on the simulation's stream
        us = ctypes.c_double()
        _lib.call("mw_sim_time_kernel", sim._h, 0, TIMING_REPS, ctypes.byref(us))
        kernel_us = us.value
    if prof_steps > 0:
        _lib.call("mw_profile_enable", 1)
        buf = (ctypes.c_double * 12)()
        _lib.call("mw_profile_read", buf, 1)  # reset counters
        for k in range(prof_steps):
            sim.step()
        torch.cuda.synchronize()
        _lib.call("mw_profile_read", buf, 1)
        _lib.call("mw_profile_enable", 0)
        prof = {PROF_NAMES[c]: {"launches": int(buf[2 * c]), "ms": buf[2 * c + 1]}
                for c in range(len(PROF_NAMES))}
    fields = None
    if keep:
        fields = {f: sim.field(f) for f in ("u", "w", "p", "T")}
    sim.close()
    return t, hist, prof, first, kernel_us, fields


TIMING_NAMES = {1: "stokes_assembly", 2: "stokes_solve", 3: "nodal_pressure", 4: "energy_assembly",
                5: "energy_solve", 6: "global_quantities", 7: "heat_flux_Nu",
                10: "write_solution_vtu", 11: "nodal_strain_rate", 12: "pressure_normalisation",
                13: "advection",
                15: "sort_T_interp_material", 16: "locate", 17: "elemental_rho_eta",
                18: "nodal_projection", 19: "timestep_dt", 20: "write_swarm_vtu",
                21: "fields_on_qpoints"}


def run_output(args, steps, every, every_swarm):
    """The mixed solver at the main workload with the reference-format output
    on: time series every step, solution VTU every `every` steps and swarm VTU
    every `every_swarm` steps (and both on the last), written by the library's
    writer thread.  Returns the stepping
    rate, the time the writer needed after the loop, and the per-stage times
    from timings.ascii (HIP events at the stage boundaries)."""
    import numpy as np
    import torch
    from meeuuw_amd import timestep as TS
    n = args.nel
    cfg = TS.make_config(nelx=n, nelz=n, Ra=1e6, eta_contrast=1e5, eta_min=1e-5, eta_max=1.0,
                         rk=4, ppe=args.ppe, solver="mixed", stokes_tol=args.tol,
                         energy_tol=1e-10, fgmres_restart=20, cheb_degree=2)
    sim = TS.Simulation(cfg)
    sim.seed_swarm(args.seed)
    folder = tempfile.mkdtemp(prefix="mw_out_")
    out = TS.Output(sim, folder, every_solution=every, every_swarm=every_swarm)
    sim.step()
    out.write()
    out.flush()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        sim.step()
        out.write(last=k == steps - 1)
        print(f"[bench] output leg step {k}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t_loop = time.perf_counter() - t0
    out.flush()
    t_all = time.perf_counter() - t0
    out.close()
    sim.close()
    rows = np.loadtxt(os.path.join(folder, "timings.ascii"), ndmin=2)[1:]
    files, nbytes = 0, 0
    for root, _, names in os.walk(folder):
        for f in names:
            files += 1
            nbytes += os.path.getsize(os.path.join(root, f))
    shutil.rmtree(folder, ignore_errors=True)
    stage_ms = {name: 1e3 * float(rows[:, i].mean()) for i, name in TIMING_NAMES.items()}
    sol_w, swm_w = rows[:, 10] > 0, rows[:, 20] > 0
    return {"workload": f"mixed solver, {n}x{n}, time series every step, solution VTU every "
                        f"{every} and swarm VTU every {every_swarm} steps (experiment0.py's "
                        "output cadence; the reference's ASCII formats)",
            "steps": steps, "ms_per_step": 1e3 * t_loop / steps, "value": steps / t_loop,
            "writer_drain_s_after_loop": t_all - t_loop, "files": files, "bytes": nbytes,
            "stage_ms_mean": stage_ms,
            "write_solution_vtu_s": float(rows[sol_w, 10].mean()) if sol_w.any() else None,
            "write_swarm_vtu_s": float(rows[swm_w, 20].mean()) if swm_w.any() else None}


def field_errors(args, oracle_npz):
    """Relative L2 errors of u, w, p, T of the mixed timestep at 64x64 (same
    physics, seed and particles as the timed workload) against the FP64 oracle
    after two steps, solved to the timed tolerance and to 1e-12."""
    import numpy as np
    ref = np.load(oracle_npz)
    out = {}
    for tol in (args.tol, 1e-12):
        import torch
        from meeuuw_amd import timestep as TS
        cfg = TS.make_config(nelx=64, nelz=64, Ra=1e6, eta_contrast=1e5, eta_min=1e-5,
                             eta_max=1.0, rk=4, ppe=args.ppe, solver="mixed", stokes_tol=tol,
                             energy_tol=1e-13 if tol < 1e-10 else 1e-10)
        sim = TS.Simulation(cfg)
        sim.seed_swarm(args.seed)
        errs = []
        for k in range(2):
            st = sim.step()
            e = {}
            for f in ("u", "w", "p", "T"):
                a, b = sim.field(f), ref[f"{f}{k}"]
                e[f] = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
            e["stokes_relres"] = st["stokes_relres"]
            errs.append(e)
        torch.cuda.synchronize()
        sim.close()
        out[f"stokes_tol_{tol:g}"] = errs
    return out


def run_batch(n, steps, threads, tune):
    """Throughput of a batch of independent instances (uneven sizes, aspect
    ratios, Ra and viscosity contrasts from one seed) in instance-steps/s:
    one warm-up step, then `steps` timed steps of every instance."""
    import torch
    from meeuuw_amd import autotune
    from meeuuw_amd import timestep as TS
    cfgs = TS.batch_configs(n=n)
    choice, tune_s = None, 0.0
    if tune:
        t0 = time.perf_counter()
        cfgs, choice = autotune.tuned_configs(cfgs)
        tune_s = time.perf_counter() - t0
    b = TS.Batch(cfgs, threads=threads)
    for i, s in enumerate(b.sims):
        s.seed_swarm(1000 + i)
    b.step(1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    st = b.step(steps)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    its = sum(x["stokes_vcycles"] for x in st) / n
    relres = max(x["stokes_relres"] for x in st)
    b.close()
    print(f"[bench] batch threads={threads} tune={tune}: {n * steps / t:.1f} instance-steps/s",
          file=sys.stderr, flush=True)
    return n * steps / t, its, relres, choice, tune_s


def run_two_material(args):
    """The two-material workload's shape (512x128 elements, aspect 4, 36 particles
    per element, thresholded Gaussian random blobs, density contrast 1 %) at the
    viscosity ratio args.tm_ratio: mixed and learned timesteps, one warm-up step
    (the network trains inside it) and args.tm_steps timed steps each."""
    import torch
    from meeuuw_amd import timestep as TS
    from meeuuw_amd import workloads as W
    nelx, nelz, ppe = args.tm_nelx, args.tm_nelx // 4, 36
    r, t, x, z = W.box_particles(nelx, nelz, 4.0, 1.0, ppe, seed=args.seed)
    mat = W.blob_materials(x, z, 4.0, 1.0)
    out = {"workload": f"{nelx}x{nelz} Q2xQ1, aspect 4, {ppe} particles per element, two "
                       f"materials (Gaussian random field, correlation length 0.1, thresholded "
                       f"at 0; blob fraction {mat.mean():.3f}), viscosity ratio "
                       f"{args.tm_ratio:g}, density contrast 1 %, RK4, Stokes rel. residual "
                       f"{args.tol}", "steps": args.tm_steps}
    for solver in ("mixed", "learned"):
        cfg = W.two_material_config(nelx=nelx, nelz=nelz, ppe=ppe, solver=solver,
                                    stokes_tol=args.tol, eta_ratio=args.tm_ratio,
                                    fgmres_restart=64, fgmres_maxit=2000)
        sim = TS.Simulation(cfg)
        sim.set_swarm_reduced(r, t)
        sim.set_swarm_material(mat)
        try:
            first = sim.step()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st = [sim.step() for _ in range(args.tm_steps)]
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            res = {"ms_per_step": 1e3 * dt / args.tm_steps,
                   "fgmres_its_per_step": sum(x["stokes_vcycles"] for x in st) / len(st),
                   "stokes_relres_max": max(x["stokes_relres"] for x in st)}
            if solver == "learned":
                res.update({"used": bool(st[-1]["learned"]), "training_s": first["learn_ms"] / 1e3,
                            "validation_loss_ratio": first["learn_ratio"],
                            "probe_fgmres_its_without": first["learn_its_plain"],
                            "probe_fgmres_its_with": first["learn_its"],
                            "probe_ms_without": first["learn_probe_ms_plain"],
                            "probe_ms_with": first["learn_probe_ms"]})
        except Exception as e:  # reported, not fatal: the workload is beside the metric
            res = {"failed": str(e)}
        sim.close()
        out[solver] = res
        print(f"[bench] two-material {solver}: {res}", file=sys.stderr, flush=True)
    if "ms_per_step" in out["mixed"] and "ms_per_step" in out["learned"]:
        out["learned_speedup"] = out["mixed"]["ms_per_step"] / out["learned"]["ms_per_step"]
        out["learned_iteration_ratio"] = (out["learned"]["fgmres_its_per_step"]
                                          / out["mixed"]["fgmres_its_per_step"])
    return out


def run_assembly(args, reps=20):
    """The build_matrix_stokes (VV_Stokes and blocks) and build_matrix_energy
    drop-ins on the args.nel^2 box with every input already on the device:
    milliseconds per call from HIP events on the stream the kernels run on
    (the element kernel, the right-hand-side gathers and the output
    allocations).  The element kernels' own durations come from rocprofv3
    (--kernel-trace --stats over this leg).  Inputs are synthetic values of the
    reference's shapes on the real box connectivity."""
    import torch
    from meeuuw_amd.assembly_indices import local_to_globalV
    from meeuuw_amd.build_matrix_energy import build_matrix_energy
    from meeuuw_amd.build_matrix_stokes import build_matrix_stokes
    from meeuuw_amd.build_pressure_connectivity import build_pressure_connectivity
    from meeuuw_amd.build_velocity_connectivity import build_velocity_connectivity
    n = args.nel
    nel, nnx = n * n, 2 * n + 1
    nn_V, nn_P = nnx * nnx, (n + 1) * (n + 1)
    cuda = dict(device="cuda", dtype=torch.float64)
    flags = torch.zeros(nn_V, dtype=torch.bool, device="cuda")
    icon_V, top, bot = build_velocity_connectivity("box", 9, n, n, nnx, nnx, flags, flags)[:3]
    icon_P = build_pressure_connectivity("box", n, n, nel, 4, device=True)
    l2g = local_to_globalV(icon_V)
    gen = torch.Generator(device="cuda").manual_seed(args.seed)
    q = lambda lo, hi: lo + (hi - lo) * torch.rand((nel, 9), generator=gen, **cuda)
    tab = lambda c: torch.rand((9, c), generator=gen, **cuda)
    h = 1.0 / n
    w = torch.tensor([5 / 9, 8 / 9, 5 / 9], **cuda)
    jxw = (torch.outer(w, w).reshape(9) * h * h / 4).repeat(nel, 1)
    j00 = torch.full((nel, 9), 2 / h, **cuda)
    j01 = torch.zeros((nel, 9), **cuda)
    ix = torch.arange(nn_V, device="cuda") % nnx
    iz = torch.arange(nn_V, device="cuda") // nnx
    fix = torch.zeros(2 * nn_V, dtype=torch.bool, device="cuda")
    fix[0::2] = (ix == 0) | (ix == nnx - 1)
    fix[1::2] = (iz == 0) | (iz == nnx - 1)
    val = torch.zeros(2 * nn_V, **cuda)
    eta = q(1e-5, 1.0)
    sargs = (nel * 468, nel * 324, nel * 16, nel * 72, nel, 9, 9, 4, 2, 2 * nn_V, nn_P, 18, icon_V,
             icon_P, q(0.9, 1.1), eta, jxw, l2g, q(-0.1, 0.1), q(-1.1, -0.9), tab(9), tab(4),
             tab(9), tab(9), tab(4), tab(4), j00, j01, j01, j00, eta.mean(dim=1), 1.0, 1.0, fix,
             val, bot, top, False, False, "box", None, False, None)
    tfix = torch.zeros(nn_V, dtype=torch.bool, device="cuda")
    tfix[(iz == 0) | (iz == nnx - 1)] = True
    eargs = (nel * 81, nel, 9, 9, nn_V, torch.rand(nn_V, generator=gen, **cuda), icon_V,
             q(0.9, 1.1), eta, q(0, 1), q(-50, 50), q(-50, 50), q(0.5, 2), q(0.5, 2), q(1e-3, 3e-3),
             q(0, 1), q(-1, 1), q(-1, 1), q(-1, 1), q(-1, 1), q(-1, 1), jxw, tab(9), tab(9), tab(9),
             j00, j01, j01, j00, tfix, torch.zeros(nn_V, **cuda), 1e-3, "BA", 1.0)
    calls = {"build_matrix_stokes (VV_Stokes)": lambda: build_matrix_stokes(*sargs, False),
             "build_matrix_stokes (blocks)": lambda: build_matrix_stokes(*sargs, True),
             "build_matrix_energy": lambda: build_matrix_energy(*eargs)}
    out = {"workload": f"{n}x{n} box, {nel} elements, inputs on the device"}
    st = torch.cuda.current_stream()
    for name, f in calls.items():
        for _ in range(2):
            f()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            f()
        e1.record(st)
        e1.synchronize()
        out[name] = {"ms_per_call": e0.elapsed_time(e1) / reps, "calls": reps}
        print(f"[bench] assembly {name}: {out[name]}", file=sys.stderr, flush=True)
    return out


def print_kernel_table(path, args):
    from meeuuw_amd import roofline as R
    rows, summ = R.kernel_table(path, R.Box(args.nel, args.nel, args.ppe), steps=args.table_steps,
                                top=args.table_top)
    print(f"# {'kernel':52s} {'launch/step':>11s} {'avg_us':>8s} {'bytes':>12s} {'GB/s':>7s} "
          f"{'frac':>6s} {'time%':>6s}")
    for r in rows:
        by, gb, fr = r["bytes_per_launch"], r["GBps"], r["frac"]
        print(f"  {r['kernel'][:52]:52s} {r['launches_per_step']:11.1f} {r['avg_us']:8.2f} "
              f"{by if by else '-':>12} {f'{gb:.0f}' if gb else '-':>7} "
              f"{f'{fr:.3f}' if fr else '-':>6} {100 * r['share_of_kernel_time']:6.1f}")
    print(f"# per step: kernel time {summ['kernel_ms_per_step']:.3f} ms, algorithmic bytes "
          f"{summ['algorithmic_bytes_per_step'] / 1e6:.1f} MB over the kernels with a formula "
          f"({100 * summ['kernel_time_with_formula']:.1f} % of kernel time)")


def print_kernel_stats(path, args):
    from meeuuw_amd import roofline as R
    lines, rows = R.stats_summary(path, top=args.table_top)
    print("\n".join(lines))
    cheb = [x for x in rows if "k_level_op<float, 2, true>" in x[5]]
    if cheb:
        print("#\n# dominant streaming kernel: finest-level Chebyshev step k_level_op<float, 2, true>")
        print(f"#   rocprofv3 average duration, all {cheb[0][1]} launches {cheb[0][2]:15.2f} us")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--nel", type=int, default=256)
    ap.add_argument("--ppe", type=int, default=25)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=200.0,
                    help="seconds to wait for the CPU oracle legs after the GPU legs end")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--prof-steps", type=int, default=2)
    ap.add_argument("--no-kernel-timing", action="store_true",
                    help="skip the back-to-back timing launches (for profiled runs)")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="first copy a buffer of this many MiB device to device (a kernel of "
                         "known traffic, to calibrate the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--batch", type=int, default=64,
                    help="instances of the batched-timestep measurement (0: none)")
    ap.add_argument("--batch-steps", type=int, default=50)
    ap.add_argument("--only", choices=["learned", "mixed", "two_material", "output", "assembly"],
                    default=None,
                    help="time one solver only (for profiling runs); the JSON line then "
                         "reports that solver as value (two_material: the two-material leg "
                         "alone, its learned timesteps/s as value; output: the output leg)")
    ap.add_argument("--tm-ratio", type=float, default=1e2,
                    help="viscosity ratio of the two-material measurement (0: none)")
    ap.add_argument("--tm-nelx", type=int, default=512)
    ap.add_argument("--tm-steps", type=int, default=3)
    ap.add_argument("--output-every", type=int, default=10,
                    help="output leg: solution VTU every N steps (0: no output leg)")
    ap.add_argument("--output-swarm-every", type=int, default=100)
    ap.add_argument("--output-steps", type=int, default=100)
    ap.add_argument("--kernel-table", default=None, metavar="TRACE_CSV")
    ap.add_argument("--kernel-stats", default=None, metavar="STATS_CSV")
    ap.add_argument("--table-steps", type=float, default=1.0)
    ap.add_argument("--table-top", type=int, default=14)
    ap.add_argument("--cpu-leg", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-fields", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_leg:
        cpu_leg(args.cpu_leg, args.nel, args.ppe, args.seed, args.cpu_fields)
        return
    if args.kernel_table:
        print_kernel_table(args.kernel_table, args)
        return
    if args.kernel_stats:
        print_kernel_stats(args.kernel_stats, args)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    tmpdir = tempfile.mkdtemp(prefix="mw_bench_")
    cpu_proc = None
    if not args.no_cpu and args.only is None:
        cpu_proc, cpu_out, cpu_fields = start_cpu_leg(args, tmpdir)
    if args.copy_probe_mb > 0:
        import torch
        a = torch.ones(args.copy_probe_mb << 18, device="cuda", dtype=torch.float32)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    if args.only == "two_material":
        tmo = run_two_material(args)
        lv = tmo["learned"].get("ms_per_step")
        print(json.dumps({"metric": "two-material timesteps/s (learned)",
                          "value": 1e3 / lv if lv else None, "unit": "timesteps/s", "n_gpus": 1,
                          "two_material": tmo}))
        return
    if args.only == "assembly":
        print(json.dumps({"metric": "assembly drop-ins at the benchmark mesh", "n_gpus": 1,
                          "assembly": run_assembly(args)}))
        return
    if args.only == "output":
        print(json.dumps({"metric": "timesteps/s with output", "n_gpus": 1,
                          "output": run_output(args, args.output_steps, args.output_every,
                                               args.output_swarm_every)}))
        return
    main_solver = args.only or "learned"
    t_l, hist, prof, first, kernel_us, _ = run_solver(
        main_solver, args, args.steps, max(args.warmup, 1), prof_steps=args.prof_steps,
        time_kernel=not args.no_kernel_timing)
    value = args.steps / t_l
    mixed = None
    if args.only is None:
        t_m, hist_m, _, _, _, _ = run_solver("mixed", args, args.steps, max(args.warmup, 1))
        mixed = {"value": args.steps / t_m, "ms_per_step": 1e3 * t_m / args.steps,
                 "stokes_vcycles_per_step": sum(h["stokes_vcycles"] for h in hist_m) / len(hist_m)}
    plain = None
    if not args.no_plain and args.plain_steps > 0 and args.only is None:
        t_pl, hist_pl, _, _, _, _ = run_solver("plain", args, args.plain_steps, args.warmup)
        plain = {"value": args.plain_steps / t_pl, "steps": args.plain_steps,
                 "ms_per_step": 1e3 * t_pl / args.plain_steps,
                 "stokes_vcycles_per_step": sum(h["stokes_vcycles"] for h in hist_pl) / len(hist_pl)}
    learned = {"used": bool(hist[-1]["learned"]), "training_s": first["learn_ms"] / 1e3,
               "validation_loss_ratio": first["learn_ratio"],
               "probe_fgmres_its_without": first["learn_its_plain"],
               "probe_fgmres_its_with": first["learn_its"],
               "probe_ms_without": first["learn_probe_ms_plain"],
               "probe_ms_with": first["learn_probe_ms"],
               "judge": "used if its_with < 0.9 its_without and ms_with < 0.95 ms_without"}
    # roofline of the finest-level Chebyshev step (the dominant streaming kernel
    # of the solve; the coarse-level kernels that take more time in total are
    # latency-bound, see "dominant_by_time" and "timestep_total")
    prof = prof or {k: {"launches": 0, "ms": 0.0} for k in PROF_NAMES}
    bytes_per = algorithmic_bytes("cheb_fine", args.nel, args.nel, args.ppe)
    achieved = bytes_per / (kernel_us * 1e-6) / 1e9 if kernel_us else None
    prof_us = profiled_avg_us(args.nel)
    roofline = {"kernel": "cheb_fine (finest-level FP32 Chebyshev step)", "bound": "hbm",
                "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                "frac": achieved / 8000.0 if achieved else None,
                "traffic": measured_traffic(args.nel),
                "avg_launch_us_hip_events": kernel_us, "timed_launches": TIMING_REPS,
                "avg_launch_us_rocprof_committed": prof_us,
                "frac_from_rocprof_committed": (bytes_per / (prof_us * 1e-6) / 1e9 / 8000.0
                                                if prof_us else None),
                "algorithmic_bytes": bytes_per}
    table = committed_kernel_table(args.nel)
    if table:
        # the whole step: algorithmic bytes per step over the kernels with a byte
        # formula (committed kernel table of the same workload) / this run's step time
        ms = 1e3 * t_l / args.steps
        roofline["timestep_total"] = {
            "algorithmic_bytes_per_step": table["bytes_per_step"],
            "share_of_kernel_time_with_formula": table["covered"],
            "ms_per_step": ms,
            "achieved_GBps": table["bytes_per_step"] / (ms * 1e-3) / 1e9,
            "frac": table["bytes_per_step"] / (ms * 1e-3) / 1e9 / 8000.0}
        roofline["dominant_by_time"] = table["top"]
    batch = None
    if args.batch > 0 and args.only is None:
        par, its_b, rr_b, _, _ = run_batch(args.batch, args.batch_steps, 8, False)
        seq, _, _, _, _ = run_batch(args.batch, args.batch_steps, 1, False)
        tun, its_t, rr_t, choice, tune_s = run_batch(args.batch, args.batch_steps, 8, True)
        batch = {"workload": f"{args.batch} independent instances in one batch: nelx from "
                             "{32,48,64,96,128}, aspect Lx/Lz from {1,2,3,4} (nelz = nelx/aspect "
                             "rounded to a multiple of 4..32), Ra log-uniform in [1e4,1e7], "
                             "viscosity contrast log-uniform in [1,1e4], 16 particles per element, "
                             "RK4, mixed FP32/FP64 solver to 1e-8, one seed",
                 "unit": "instance-timesteps/s", "steps": args.batch_steps, "threads": 8,
                 "value": par, "sequential_value": seq, "speedup_vs_sequential": par / seq,
                 "stokes_vcycles_per_step": its_b, "stokes_relres_max": rr_b,
                 "autotuned_value": tun, "autotuned_stokes_vcycles_per_step": its_t,
                 "autotuned_relres_max": rr_t, "autotune_s": tune_s,
                 "autotuned_schedules": {f"{k[0]}x{k[1]}": v for k, v in choice.items()}}
    output = None
    if args.output_every > 0 and args.only is None:
        output = run_output(args, args.output_steps, args.output_every, args.output_swarm_every)
        if mixed:
            output["rate_vs_mixed_without_output"] = output["value"] / mixed["value"]
    two_material = None
    if args.tm_ratio > 0 and args.only is None:
        two_material = run_two_material(args)
    assembly = None
    if args.only is None:
        try:  # the drop-in assembly calls at the benchmark mesh, beside the metric
            assembly = run_assembly(args)
        except Exception as e:
            assembly = {"failed": str(e)}
    cpu, ferr = None, None
    if cpu_proc is not None:
        deadline = time.time() + args.cpu_budget
        legs = read_cpu_legs(cpu_out)
        if "64x64" not in legs:  # the field errors need the 64x64 oracle fields
            while "64x64" not in legs and cpu_proc.poll() is None and time.time() < deadline:
                time.sleep(0.5)
                legs = read_cpu_legs(cpu_out)
        if "64x64" in legs:
            ferr = field_errors(args, cpu_fields)
        main_leg = f"{args.nel}x{args.nel}"
        while cpu_proc.poll() is None and time.time() < deadline:
            time.sleep(1.0)
        if cpu_proc.poll() is None:
            cpu_proc.kill()
            cpu_proc.wait()
        legs = read_cpu_legs(cpu_out)
        model, ncpu = cpu_info()
        extra = {k: {"timesteps_per_s": v["steps"] / v["seconds"], "steps": v["steps"],
                     "stages_s": {s: round(x, 4) for s, x in v["stages"].items()}}
                 for k, v in legs.items() if k != main_leg}
        m = legs.get(main_leg)
        cpu = {"value": (m["steps"] / m["seconds"]) if m else None, "unit": "timesteps/s",
               "cores": 1, "kind": "port",
               "sample": (f"1 oracle timestep of the {main_leg} workload (eta contrast 1e5, Ra 1e6, "
                          f"RK4, {args.ppe} particles per element; C loops + SciPy SuperLU direct "
                          "solves, 1 thread), run concurrently with the GPU legs"
                          + ("" if m else f"; not finished within the {args.cpu_budget:.0f} s "
                                          "budget after the GPU legs")),
               "cpu_model": model, "host_cores_available": ncpu,
               "stages_s": ({s: round(x, 4) for s, x in m["stages"].items()} if m else None),
               "other_samples": extra}
    out = {
        "metric": "timesteps/s at 256x256 Q2Q1 eta-contrast 1e5, learned-PC vs plain, res 1e-8",
        "value": value, "unit": "timesteps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": max(args.warmup, 1), "ms_per_step": 1e3 * t_l / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{args.nel}x{args.nel} Q2xQ1 box, Ra 1e6, FK eta contrast 1e5, "
                               f"{args.ppe} particles/element, RK4, Stokes rel. residual {args.tol}",
                   "solver": "learned: FP32 GMG-FGMRES in FP64 defect correction with the "
                             "learned Schur-complement correction where it is judged useful",
                   "nelx": args.nel, "nelz": args.nel, "particles": args.nel ** 2 * args.ppe},
        "stokes_vcycles_per_step": sum(h["stokes_vcycles"] for h in hist) / len(hist),
        "stokes_relres_max": max(h["stokes_relres"] for h in hist),
        "vrms_last": hist[-1]["vrms"], "Nu_last": hist[-1]["Nu"],
        "learned_correction": learned,
        "mixed_no_network": mixed,
        "plain_fp64_uzawa": plain,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
        "field_errors_vs_fp64_oracle_64x64": ferr,
        "two_material": two_material,
        "assembly_dropins": assembly,
        "output": output,
        "kernels": prof, "roofline": roofline, "cpu_baseline": cpu, "batch": batch,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
