"""Reference-format output files (mw_output_*), written by the native writer
thread from pinned snapshots.

Expected text is formatted here from the statistics mw_sim_step returned and
the fields read back after the same step, with the reference's own format
strings: time series meeuuw.py:1201-1203, 2148, 2308, 2518-2525; solution VTU
output_solution_to_vtu.py:84-433 (numpy tofile(sep=" ", format=...) joins
values with one space); swarm VTU output_swarm_to_vtu.py:60-209 (coordinates
as tofile(sep=" ") without a format, i.e. repr); pvd write_in_pvd_files.py.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _arr(text, name):
    """Text between the DataArray tag named `name` and its closing tag."""
    i = text.index("Name='%s'" % name)
    i = text.index("\n", i) + 1
    return text[i:text.index("</DataArray>", i)]


def _join(v, f):
    return " ".join(f % x for x in v)


def test_output_files_match_reference_formats(tmp_path):
    from meeuuw_amd import timestep as ts
    from meeuuw_amd.workloads import blob_materials
    nelx, nelz, ppe, nstep = 12, 8, 9, 3
    cfg = ts.make_config(nelx=nelx, nelz=nelz, Lx=1.5, ppe=ppe, rk=2, material_mode=1,
                         mat_eta=(1.0, 10.0), mat_rho=(1.0, 1.01), gravity=1.0, alphaT=0.0)
    sim = ts.Simulation(cfg)
    mesh = O.Mesh(nelx, nelz, 1.5, 1.0)
    sim.set_temperature(np.clip(1.0 - mesh.z_V, 0.0, 1.0))
    sim.seed_swarm(7)
    mat = blob_materials(sim.field("swarm_x"), sim.field("swarm_z"), 1.5, 1.0)
    sim.set_swarm_material(mat)
    out = ts.Output(sim, tmp_path, every_solution=2, every_swarm=2,
                    material_names=["matrix", "blob"])
    stats, fields = [], {}
    for istep in range(nstep):
        stats.append(sim.step())
        out.write(last=istep == nstep - 1)
        if istep in (0, 2):  # read after the snapshot was queued: it must not see later steps
            fields[istep] = {k: sim.field(k) for k in ("u", "w", "q", "T", "eta_n", "rho_n",
                                                       "eta_e", "rho_e", "swarm_x", "swarm_z",
                                                       "swarm_active")}
    out.close()
    sim.close()

    def lines(name):
        with open(os.path.join(tmp_path, name)) as f:
            return f.read()

    t = [s["time"] for s in stats]
    assert lines("vrms.ascii") == "#time,vrms\n" + "".join(
        "%.4e %.6e \n" % (s["time"], s["vrms"]) for s in stats)
    assert lines("Nu.ascii") == "#time Nu\n" + "".join(
        "%e %.6e \n" % (s["time"], s["Nu"]) for s in stats)
    assert lines("dt.ascii") == "#time dt dt1 dt2\n# \n" + "".join(
        "%e %e %e %e\n" % (s["time"], s["dt"], s["dt1"], s["dt2"]) for s in stats)
    for name, head, key, f in (("T_avrg.ascii", "#time <T>\n", "Tavrg", "%.6e"),
                               ("kinetic_energy.ascii", "# time kinetic_energy\n", "EK", "%.4e"),
                               ("total_mass.ascii", "# time total mass\n", "TM", "%.4e"),
                               ("tvd.ascii", "# time total viscous dissipation\n", "TVD", "%.4e"),
                               ("wag.ascii", "#time WAG\n", "WAG", "%.4e"),
                               ("eta_avrg.ascii", "#time eta_avrg\n", "eta_avrg", "%.4e")):
        assert lines(name) == head + "".join(("%.4e " + f + " \n") % (s["time"], s[key])
                                             for s in stats), name
    assert lines("delta_wag_tvd.ascii") == "#time delta max(abs(WAG),TVD)\n" + "".join(
        "%.4e %.4e %.4e\n" % (s["time"], s["WAG"] + s["TVD"], max(abs(s["WAG"]), s["TVD"]))
        for s in stats)
    assert lines("stats_particle.ascii") == "#time min(nparticle_e) max(nparticle_e)\n" + "".join(
        "%.4e %d %d\n" % (s["time"], s["mincount"], s["maxcount"]) for s in stats)
    assert lines("stats_eta_e.ascii") == "#time min(eta_e) max(eta_e)\n" + "".join(
        "%.4e %.4e %.4e\n" % (s["time"], s["eta_e_min"], s["eta_e_max"]) for s in stats)
    # timings.ascii (meeuuw.py:3371-3375): istep, then the stage times in seconds
    # at the reference's timings[] indices; output times (10, 20) on written steps
    rows = [ln for ln in lines("timings.ascii").split("\n") if ln]
    assert len(rows) == nstep
    staged = {1, 2, 3, 4, 5, 6, 7, 11, 12, 13, 15, 16, 17, 18, 19, 21}
    for i, ln in enumerate(rows):
        assert ln.endswith(" ") and " ".join("%e" % float(x) for x in ln.split()) == ln[:-1]
        v = np.array(ln.split(), dtype=float)
        assert v.size == 40 and v[0] == i
        assert all(v[k] > 0 for k in staged), ln
        written = {10, 20} if i in (0, 2) else set()
        assert all(v[k] > 0 for k in written) and not any(
            v[k] for k in range(1, 40) if k not in staged | written), ln
    # the statistics themselves: extrema and counts of the step's fields
    f2 = fields[2]
    assert stats[2]["eta_e_min"] == f2["eta_e"].min() and stats[2]["eta_e_max"] == f2["eta_e"].max()
    assert stats[0]["mincount"] <= ppe <= stats[0]["maxcount"]
    # averages of material viscosities 1 and 10 (up to rounding)
    assert 1.0 < stats[2]["eta_e_max"] <= 10.0 * (1 + 1e-12)
    assert 1.0 - 1e-12 <= stats[2]["eta_e_min"] < 10.0

    # solution and swarm files: steps 0 and 2 (every 2nd, and the last), not 1
    names = sorted(os.listdir(tmp_path))
    assert "solution_0000.vtu" in names and "solution_0002.vtu" in names
    assert "solution_0001.vtu" not in names
    assert sorted(os.listdir(tmp_path / "SWARM")) == ["swarm_0000.vtu", "swarm_0002.vtu"]
    head = ('<?xml version="1.0"?> \n<VTKFile type="Collection" version="0.1" '
            'ByteOrder="LittleEndian"> \n  <Collection> \n')
    tail = "  </Collection>\n</VTKFile>\n"
    for pvd, pat in (("solution.pvd", "solution_%04d.vtu"), ("swarm.pvd", "SWARM/swarm_%04d.vtu")):
        assert lines(pvd) == head + "".join(
            '    <DataSet timestep="%s" group="" part="0" file="%s"/>  \n' % (t[i], pat % i)
            for i in (0, 2)) + tail

    nn, nel = mesh.x_V.size, nelx * nelz
    for istep in (0, 2):
        f = fields[istep]
        v = lines("solution_%04d.vtu" % istep)
        assert v.startswith("<VTKFile type='UnstructuredGrid' version='0.1' byte_order='BigEndian'>"
                            " \n<UnstructuredGrid> \n<Piece NumberOfPoints=' %5d ' "
                            "NumberOfCells=' %5d '> \n" % (nn, nel))
        pts = "".join("%.5e %.1e %.5e \n" % (a, 0.0, b) for a, b in zip(mesh.x_V, mesh.z_V))
        assert ("<DataArray type='Float32' NumberOfComponents='3' Format='ascii'> \n" + pts
                + "</DataArray>\n") in v
        assert _arr(v, "Velocity ( )") == "".join(
            "%.4e %.1e %.4e \n" % (a, 0.0, b) for a, b in zip(f["u"], f["w"]))
        assert _arr(v, "Pressure") == _join(f["q"], "%.4e")
        assert _arr(v, "Temperature") == "".join("%.4e \n" % x for x in f["T"])
        pd = v[v.index("<PointData"):v.index("</PointData>")]
        cd = v[v.index("<CellData"):v.index("</CellData>")]
        assert _arr(pd, "Viscosity") == _join(f["eta_n"], "%.4e")
        assert _arr(pd, "Density") == _join(f["rho_n"], "%.5e")
        assert _arr(pd, "Heat flux").count("\n") == nn
        assert _arr(cd, "Viscosity") == _join(f["eta_e"], "%.4e")
        assert _arr(cd, "Density") == _join(f["rho_e"], "%.5e")
        cnt = np.array(_arr(cd, "nb particles").split(), dtype=np.int64)
        assert cnt.size == nel and cnt.sum() == nel * ppe
        assert _arr(v, "connectivity") == "".join(
            "%d %d %d %d %d %d %d %d %d\n" % tuple(mesh.icon_V[:, e]) for e in range(nel))
        assert _arr(v, "offsets") == "".join("%d \n" % ((e + 1) * 9) for e in range(nel))
        assert _arr(v, "types") == "28 \n" * nel
        assert v.endswith("</Cells>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n")

        s = lines("SWARM/swarm_%04d.vtu" % istep)
        act = f["swarm_active"].astype(bool)
        na = int(act.sum())
        assert "<Piece NumberOfPoints=' %5d ' NumberOfCells=' %5d '> \n" % (na, na) in s
        coords = np.array([f["swarm_x"][act], np.zeros(na), f["swarm_z"][act]]).T
        expect = " ".join(repr(float(x)) for x in coords.ravel())
        i = s.index("<Points> \n<DataArray type='Float32' NumberOfComponents='3' Format='ascii'> \n")
        i = s.index("\n", s.index("<DataArray", i)) + 1
        assert s[i:s.index("</DataArray>", i)] == expect
        m = mat[act]
        assert _arr(s, "matrix") == " ".join("%.3e" % x for x in (m == 0).astype(float))
        assert _arr(s, "blob") == " ".join("%.3e" % x for x in (m == 1).astype(float))
        assert _arr(s, "id") == " ".join("%d" % i for i in np.nonzero(act)[0])
        assert _arr(s, "connectivity") == _join(np.arange(0, na + 1), "%d")
        assert _arr(s, "offsets") == _join(np.arange(1, na + 2), "%d")
        assert _arr(s, "types") == _join(np.full(na, 1), "%d")


def test_output_rejects_bad_arguments(tmp_path):
    from meeuuw_amd import timestep as ts
    cfg = ts.make_config(nelx=4, nelz=4, ppe=4)
    sim = ts.Simulation(cfg)
    with pytest.raises(ValueError, match="material names"):
        ts.Output(sim, tmp_path, material_names=["a", "b"])  # one material in material_mode 0
    out = ts.Output(sim, tmp_path)
    with pytest.raises(RuntimeError, match="no step"):
        out.write()
    out.close()
    sim.close()
