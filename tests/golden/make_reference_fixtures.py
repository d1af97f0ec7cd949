"""Extract the reference's published numeric results into a small JSON fixture.

Reads data files that the reference repository holds under RESULTS/ (text
tables written by its own runs) and writes tests/golden/reference_results.json.
Run in the build container, where the reference is mounted:
    python tests/golden/make_reference_fixtures.py /root/reference
"""
import json
import os
import re
import sys


def second_line_values(path):
    with open(path) as f:
        lines = [l for l in f if l.strip() and not l.startswith("#")]
    return [float(v) for v in lines[0].split()]


def errors_table(path):
    out = {}
    with open(path) as f:
        for line in f:
            m = re.search(r"errv,errp=\s*([-\d.eE+]+)\s+([-\d.eE+]+)\s+(\d+)", line)
            if m:
                out[int(m.group(3))] = [float(m.group(1)), float(m.group(2))]
    return out


def trace(path):
    """Rows of a solver_convergence_N.ascii file: xiV, xiP, alpha[, beta] per iteration."""
    with open(path) as f:
        return [[float(v) for v in l.split()] for l in f if l.strip() and not l.startswith("#")]


def series(path, head=10, tail=10):
    """First and last rows of a two-column time series, and its row count."""
    with open(path) as f:
        rows = [[float(v) for v in l.split()] for l in f if l.strip() and not l.startswith("#")]
    return {"head": rows[:head], "tail": rows[-tail:], "rows": len(rows)}


def main(ref):
    R = os.path.join(ref, "RESULTS")
    fx = {"source": "reference RESULTS/ tables (see 'files')", "files": {}}
    bb = {}
    for n in (16, 24, 32, 48, 64):
        d = os.path.join(R, "exp00", "case1a", f"OUTPUT_{n}")
        t_v, vrms = second_line_values(os.path.join(d, "vrms.ascii"))
        t_n, nu = second_line_values(os.path.join(d, "Nu.ascii"))
        bb[n] = {"time_after_step0": t_v, "vrms_step0": vrms, "Nu_step0": nu}
        fx["files"][f"blankenbach_{n}"] = f"RESULTS/exp00/case1a/OUTPUT_{n}/vrms.ascii:2, Nu.ascii:2"
    fx["blankenbach_1a"] = bb
    fx["donea_huerta_errors"] = errors_table(os.path.join(R, "exp19", "errors.ascii"))
    fx["files"]["donea_huerta_errors"] = "RESULTS/exp19/errors.ascii"
    # per-iteration block-solver traces (particles on the quadrature points):
    # exp19 Donea-Huerta with uzawa3_solver_L2, exp21 SolCx with cd_u_eta and
    # with uzawa3_solver_L2
    traces = {}
    for n in (16, 32, 48):
        traces[f"exp19_uzawa3_{n}"] = trace(os.path.join(R, "exp19", f"solver_convergence_{n}.ascii"))
        traces[f"exp21_cd_u_eta_{n}"] = trace(os.path.join(R, "exp21", f"solver_convergence_{n}.ascii"))
        traces[f"exp21_uzawa3_{n}"] = trace(
            os.path.join(R, "exp21", "uzawa3", f"solver_convergence_{n}.ascii"))
    fx["solver_traces"] = traces
    fx["files"]["solver_traces"] = ("RESULTS/exp19/solver_convergence_N.ascii, "
                                    "RESULTS/exp21/solver_convergence_N.ascii, "
                                    "RESULTS/exp21/uzawa3/solver_convergence_N.ascii")
    # Tosi et al. (2015) visco-plastic convection (experiment3.py): case 2 at
    # 32x32 to t = 0.25, case 1 at 64x64 (start of the run)
    tosi = {"case2_32": {"vrms": series(os.path.join(R, "exp03", "case2", "vrms.ascii")),
                         "Nu": series(os.path.join(R, "exp03", "case2", "Nu.ascii"))},
            "case1_64": {"vrms": series(os.path.join(R, "exp03", "case1", "64", "vrms.ascii")),
                         "Nu": series(os.path.join(R, "exp03", "case1", "64", "Nu.ascii"))}}
    fx["tosi"] = tosi
    fx["files"]["tosi"] = ("RESULTS/exp03/case2/{vrms,Nu}.ascii, "
                           "RESULTS/exp03/case1/64/{vrms,Nu}.ascii")
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_results.json"),
              "w") as f:
        json.dump(fx, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
