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
reported beside it, with the CPU oracle as the CPU baseline.  Kernel timings
for the roofline come from a separate, profiled pass after the timed steps.
Under "batch": the throughput of 64 independent instances of uneven size
stepped together (mw_batch_*), sequentially, concurrently, and with the
schedules the autotuner picks per mesh class.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PROF_NAMES = ["stokes_fp64_operator", "cheb_fine", "stokes_elements", "energy_vv", "fgmres_gs",
              "advect_rk4"]
TIMING_REPS = 200
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "cheb_fine_256x256_hbm_traffic.txt")


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


def cpu_baseline(nel_sample, steps):
    """Time the CPU oracle's full timestep (C loops + SciPy SuperLU) on one core
    in a subprocess; returns (timesteps/s, per-stage seconds)."""
    code = f"""
import json, time, sys
sys.path.insert(0, {ROOT!r})
from oracle import step as S
P = S.Params(nelx={nel_sample}, nelz={nel_sample}, Ra=1e6, eta_contrast=1e5, RKorder=4, ppe=25)
st = S.State(P)
tm = {{}}
t0 = time.perf_counter()
for k in range({steps}):
    S.step(st, tm)
t = time.perf_counter() - t0
print(json.dumps({{"seconds": t, "stages": tm}}))
"""
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env,
                       timeout=900)
    if r.returncode != 0:
        raise RuntimeError(r.stderr[-2000:])
    out = json.loads(r.stdout.strip().splitlines()[-1])
    return steps / out["seconds"], out["stages"]


def run_solver(solver, args, steps, warmup, prof_steps=0):
    import ctypes
    import torch
    from meeuuw_amd import _lib
    from meeuuw_amd import timestep as TS
    n = args.nel
    cfg = TS.make_config(nelx=n, nelz=n, Ra=1e6, eta_contrast=1e5, eta_min=1e-5, eta_max=1.0,
                         rk=4, ppe=args.ppe, solver=solver, stokes_tol=args.tol,
                         energy_tol=1e-10, fgmres_restart=20, cheb_degree=2)
    sim = TS.Simulation(cfg)
    sim.seed_swarm(args.seed)
    first = None
    for _ in range(warmup):
        st = sim.step()
        first = first or st
        print(f"[bench] {solver} warmup step {st}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = []
    for k in range(steps):
        hist.append(sim.step())
        print(f"[bench] {solver} step {k} {hist[-1]}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    prof = None
    kernel_us = None
    if solver != "plain":
        # the dominant streaming kernel (finest-level Chebyshev step), timed over
        # back-to-back launches at its real shape with HIP events on the sim stream
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
    sim.close()
    return t, hist, prof, first, kernel_us


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
    return n * steps / t, its, relres, choice, tune_s


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
    ap.add_argument("--cpu-nel", type=int, default=64)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--prof-steps", type=int, default=2)
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="first copy a buffer of this many MiB device to device (a kernel of "
                         "known traffic, to calibrate the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--batch", type=int, default=64,
                    help="instances of the batched-timestep measurement (0: none)")
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--only", choices=["learned", "mixed"], default=None,
                    help="time one solver only (for profiling runs); the JSON line then "
                         "reports that solver as value")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if args.copy_probe_mb > 0:
        import torch
        a = torch.ones(args.copy_probe_mb << 18, device="cuda", dtype=torch.float32)
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    main_solver = args.only or "learned"
    t_l, hist, prof, first, kernel_us = run_solver(main_solver, args, args.steps,
                                                   max(args.warmup, 1), prof_steps=args.prof_steps)
    value = args.steps / t_l
    mixed = None
    if args.only is None:
        t_m, hist_m, _, _, _ = run_solver("mixed", args, args.steps, max(args.warmup, 1))
        mixed = {"value": args.steps / t_m, "ms_per_step": 1e3 * t_m / args.steps,
                 "stokes_vcycles_per_step": sum(h["stokes_vcycles"] for h in hist_m) / len(hist_m)}
    plain = None
    if not args.no_plain and args.plain_steps > 0 and args.only is None:
        t_pl, hist_pl, _, _, _ = run_solver("plain", args, args.plain_steps, args.warmup)
        plain = {"value": args.plain_steps / t_pl, "steps": args.plain_steps,
                 "ms_per_step": 1e3 * t_pl / args.plain_steps,
                 "stokes_vcycles_per_step": sum(h["stokes_vcycles"] for h in hist_pl) / len(hist_pl)}
    learned = {"used": bool(hist[-1]["learned"]), "training_s": first["learn_ms"] / 1e3,
               "validation_loss_ratio": first["learn_ratio"],
               "probe_fgmres_its_without": first["learn_its_plain"],
               "probe_fgmres_its_with": first["learn_its"],
               "judge": "used if its_with < 0.9 its_without"}
    # roofline of the finest-level Chebyshev step (the dominant streaming kernel
    # of the solve, see the kernel-class breakdown in "kernels")
    prof = prof or {k: {"launches": 0, "ms": 0.0} for k in PROF_NAMES}
    bytes_per = algorithmic_bytes("cheb_fine", args.nel, args.nel, args.ppe)
    achieved = bytes_per / (kernel_us * 1e-6) / 1e9 if kernel_us else None
    roofline = {"kernel": "cheb_fine (finest-level FP32 Chebyshev step)", "bound": "hbm",
                "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                "frac": achieved / 8000.0 if achieved else None,
                "traffic": measured_traffic(args.nel), "avg_launch_us": kernel_us,
                "timed_launches": TIMING_REPS, "algorithmic_bytes": bytes_per}
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
    cpu = None
    if not args.no_cpu:
        v, stages = cpu_baseline(args.cpu_nel, args.cpu_steps)
        cpu = {"value": v, "unit": "timesteps/s", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_steps} oracle timesteps at {args.cpu_nel}x{args.cpu_nel} "
                         f"(same physics: eta contrast 1e5, Ra 1e6, RK4, {args.ppe} particles per "
                         f"element; C loops + SciPy SuperLU direct solves, 1 thread)",
               "stages_s": {k: round(x, 4) for k, x in stages.items()}}
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
        "kernels": prof, "roofline": roofline, "cpu_baseline": cpu, "batch": batch,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
