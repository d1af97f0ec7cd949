"""Algorithmic HBM bytes of the timestep's kernels, for roofline figures.

Compulsory traffic only: every input array read once, every output written
once, nodal and elemental arrays gathered once (a gather of a nodal field costs
the field once, not once per element).  Sizes: f64 8 B, f32 4 B, int32 4 B,
uint8 1 B.  Formulas follow the kernels' own arguments; the particle and
assembly kernels use the per-unit figures of the reference-signature kernels
(RK4 advection N_p*50 + 16 nn + 52 nel, locate N_p*37 + 20 nel, ...).

``kernel_table`` reads a rocprofv3 kernel trace (``--kernel-trace`` CSV), groups
the dispatches by kernel and launch shape, attaches the bytes of each group
(the multigrid level follows from the launch shape) and reports achieved GB/s
against the MI355X HBM peak.
"""
import collections
import csv

HBM_PEAK_GBPS = 8000.0


class Box:
    """Sizes of an nelx x nelz Q2xQ1 box with ppe particles per element."""

    def __init__(self, nelx, nelz, ppe):
        self.nelx, self.nelz, self.ppe = nelx, nelz, ppe
        self.nel = nelx * nelz
        self.nn = (2 * nelx + 1) * (2 * nelz + 1)
        self.NV = 2 * self.nn
        self.nnP = (nelx + 1) * (nelz + 1)
        self.N = self.NV + self.nnP  # Stokes unknowns
        self.np = self.nel * ppe

    def coarsest(self, max_dof=600):
        """Unknowns of the coarsest multigrid level (the first with at most
        max_dof unknowns, the levels halving while both counts are even and >= 4)."""
        ex, ez = self.nelx, self.nelz
        while 2 * (2 * ex + 1) * (2 * ez + 1) > max_dof and ex % 2 == 0 and ez % 2 == 0 \
                and ex >= 4 and ez >= 4:
            ex, ez = ex // 2, ez // 2
        return 2 * (2 * ex + 1) * (2 * ez + 1)

    def level(self, l):
        """(ndof, nel) of multigrid level l (level 0: the mesh itself)."""
        ex, ez = self.nelx >> l, self.nelz >> l
        return 2 * (2 * ex + 1) * (2 * ez + 1), ex * ez


def level_of_tiles(box, tiles_x, edge=16):
    """Multigrid level whose node grid is covered by tiles_x tiles of `edge` nodes."""
    for l in range(12):
        ex = box.nelx >> l
        if ex == 0:
            break
        if -(-(2 * ex + 1) // edge) == tiles_x:
            return l
    return None


def kernel_bytes(name, tx, box):
    """Algorithmic bytes of one launch of kernel `name` whose grid has tx
    workgroups along x, or None for kernels without a formula here."""
    b = box
    if "k_level_op<float, 2" in name or "k_cheb01<float" in name:
        l = level_of_tiles(b, tx)
        if l is None:
            return None
        nd, ne = b.level(l)
        return (25 if "k_level_op" in name else 17) * nd + 4 * ne
    if "k_resid_restrict8<float" in name or "k_prolong_cheb0<float" in name:
        # 8x8 coarse tiles (restriction) or 16x16 fine tiles (prolongation)
        l = level_of_tiles(b, tx, 8) if "restrict" in name else level_of_tiles(b, tx)
        if "restrict" in name and l is not None:
            l -= 1  # the tiles cover the coarse level
        if l is None or l < 0:
            return None
        nd, ne = b.level(l)
        ndc, _ = b.level(l + 1)
        if "restrict" in name:
            return 9 * nd + 4 * ne + 4 * ndc
        return 21 * nd + 4 * ne + 4 * ndc
    if "k_dense_mv<float>" in name:  # the coarsest inverse (f32) and b in, x out
        n = b.coarsest()
        return 4 * n * n + 8 * n
    if "k_gj_sweep" in name:  # one blocked Gauss-Jordan sweep reads and writes the padded matrix
        npad = -(-b.coarsest() // 64) * 64
        return 2 * 8 * npad * npad
    if "k_stokes_fine<float" in name:  # y = A z: z, y (f32), eta, fixed-dof mask
        return 8 * b.N + 4 * b.nel + b.NV
    if "k_schur_front_cheb" in name:  # v (f32), element weights (f64), diag (f64), t, z_p out
        return 8 * b.N + 8 * b.nel + 16 * b.nnP + b.NV
    if "k_stokes_residual64" in name:  # u, p, f, h, eta_q in; ru, rp out; mask
        return b.NV * (3 * 8 + 1) + b.nnP * 3 * 8 + b.nel * 9 * 8
    for tag, extra in (("k_mdot_partial<float, ", 1), ("k_gs_project<float, ", 2),
                       ("k_gs_update_scale<float, ", 2)):
        if tag in name:
            m = int(name.split(tag)[1].split(">")[0])
            return (m + extra) * b.N * 4
    if "k_energy_op" in name:  # 36 coefficients per element; r, p, v, 1/diag in; p, p^, v out
        return b.nel * 288 + 7 * 8 * b.nn
    if "k_energy_update" in name:
        return 8 * 8 * b.nn
    if "k_energy_rhs" in name:  # T_old, coefficients, fixed values; rhs and diagonal out
        return b.nel * 288 + 5 * 8 * b.nn
    if "k_advect_box" in name:
        return b.np * 50 + 16 * b.nn + 52 * b.nel
    if "k_locate_box" in name:
        return b.np * 37 + 20 * b.nel
    if "k_material_sorted" in name:  # order, iel, r, t in; T gathered; 8 f64 out
        return b.np * (4 + 4 + 16 + 64) + 8 * b.nn
    if "k_project_elements" in name:
        return b.np * 16 + b.nel * 24
    if "k_node_partials" in name:  # r, t and 6 fields in; 48 partial sums per element out
        return b.np * 64 + b.nel * (8 + 48 * 8)
    if "k_node_finish" in name:
        return b.nel * 48 * 8 + b.nn * 4 + 9 * b.nel * 4 + 6 * 8 * b.nn
    if "k_node_fill" in name:
        return 9 * b.nel * 4 + 6 * 8 * b.nn
    if "k_sort_count" in name or "k_sort_scatter" in name:
        return b.np * 5 + 4 * b.nel + (4 * b.np if "scatter" in name else 0)
    if "k_sort_rank" in name:
        return b.np * 13 + 4 * b.nel
    if "k_ngrad_elem" in name:  # u, v gathered; 4 node Jacobians in; 3 per-(element, node) values out
        return 16 * b.nn + 9 * b.nel * 8 * (4 + 3)
    if "k_ngrad_node" in name:  # the per-(element, node) values, incidence in; 3 nodal fields out
        return 27 * 8 * b.nel + 36 * b.nel + 4 * b.nn + 24 * b.nn
    if "k_gq" in name:  # 12 quadrature-point fields in
        return 12 * 72 * b.nel
    if "k_stokes_elements" in name:  # VV-writing build_matrix_stokes (blocks=False)
        return (b.nel * (9 * 9 * 8 + 36 + 16 + 72 + 8 + 2) + b.NV * 9 + b.nel * 468 * 8
                + b.N * 8)
    return None


def kernel_table(trace_csv, box, steps, top=12):
    """Rows (name, launches/step, avg us, bytes per launch, GB/s, fraction of
    peak) for the `top` kernel groups by time, and the algorithmic bytes and
    kernel time per step over the groups with a formula."""
    groups = collections.OrderedDict()
    with open(trace_csv) as f:
        for r in csv.DictReader(f):
            key = (r["Kernel_Name"], int(r["Grid_Size_X"]), int(r["Grid_Size_Y"]),
                   int(r["Workgroup_Size_X"]))
            g = groups.setdefault(key, [0, 0.0])
            g[0] += 1
            g[1] += (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9
    rows, tot_b, tot_t, cov_t = [], 0.0, 0.0, 0.0
    for (name, gx, gy, wg), (n, t) in groups.items():
        by = kernel_bytes(name, gx // max(wg, 1), box)
        tot_t += t
        if by is not None:
            tot_b += by * n
            cov_t += t
        rows.append((name, gx, gy, n, t, by))
    rows.sort(key=lambda x: -x[4])
    out = []
    for name, gx, gy, n, t, by in rows[:top]:
        avg = t / n
        gbps = by / avg / 1e9 if by else None
        out.append({"kernel": name.split("(")[0].replace("void ", ""), "grid": [gx, gy],
                    "launches_per_step": n / steps, "avg_us": avg * 1e6,
                    "bytes_per_launch": by, "GBps": gbps,
                    "frac": gbps / HBM_PEAK_GBPS if gbps else None,
                    "share_of_kernel_time": t / tot_t})
    summary = {"kernel_ms_per_step": tot_t * 1e3 / steps,
               "algorithmic_bytes_per_step": tot_b / steps,
               "kernel_time_with_formula": cov_t / tot_t}
    return out, summary


def stats_summary(stats_csv, top=40):
    """Text rows of a rocprofv3 --stats kernel table (total ms, share, calls,
    average/min/max us, name) for the `top` kernels by total time, and the
    total kernel time and dispatch count."""
    rows = []
    with open(stats_csv) as f:
        for r in csv.DictReader(f):
            rows.append((float(r["TotalDurationNs"]) * 1e-6, int(r["Calls"]),
                         float(r["AverageNs"]) * 1e-3, float(r["MinNs"]) * 1e-3,
                         float(r["MaxNs"]) * 1e-3, r["Name"]))
    rows.sort(key=lambda x: -x[0])
    total = sum(x[0] for x in rows)
    lines = [f"# total kernel time {total:.2f} ms over {sum(x[1] for x in rows)} dispatches",
             "#  total_ms   pct   calls    avg_us   min_us    max_us  kernel"]
    for ms, n, avg, mn, mx, name in rows[:top]:
        lines.append(f"  {ms:8.3f} {100 * ms / total:5.1f} {n:7d} {avg:9.2f} {mn:8.2f} {mx:9.2f}  "
                     f"{name[:110]}")
    return lines, rows
