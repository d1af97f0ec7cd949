// On-disk output of the device-resident timestep, written off the stepping
// path, in the reference's formats.
//
//   time series        open_files.py:10-66 (names, headers); lines as
//                      meeuuw.py:1201 (stats_particle), 1203 (stats_eta_e),
//                      2148 (dt), 2308 (Nu), 2518-2525 (vrms, total_mass,
//                      kinetic_energy, tvd, wag, delta_wag_tvd, T_avrg, eta_avrg)
//   solution_NNNN.vtu  output_solution_to_vtu.py:84-433, ASCII: points, the
//                      nodal velocity, pressure q, temperature, viscosity,
//                      density and heat flux; elemental viscosity, density and
//                      particle count; 9-node cells in icon_V order, VTK type 28
//   SWARM/swarm_NNNN.vtu  output_swarm_to_vtu.py:60-209, ASCII: the active
//                      particles' positions, material weights and index; the
//                      vertex cells exactly as the reference writes them
//   solution.pvd,      write_in_pvd_files.py:15-36 and meeuuw.py:3421-3424
//   swarm.pvd          (one DataSet per written step, timestep = the step's
//                      start time in Python's float repr)
// Arrays the timestep does not hold (strain rate, stress, divergence, particle
// strain, paint, ...) are not written.  The timestep is nondimensional: the
// velocity, time and temperature scales are 1 and 0 K offset, and the unit
// strings are " " (set_default_parameters.py:8-12).
//
// Off the stepping path: after mw_sim_step, mw_output_step queues the step's
// statistics for a writer thread.  On a step whose solution or swarm is due it
// copies the device fields into a device staging slot on the simulation's
// stream (microseconds at HBM speed), a second stream copies the slot to
// pinned host memory while the simulation goes on, and the writer thread
// formats the files from there once that copy has landed.  The host waits only
// when the slot it wants is still being written out by an earlier snapshot.
#include <hip/hip_runtime.h>

#include <charconv>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <sys/stat.h>
#include <thread>
#include <vector>

#include "mw_common.h"
#include "mw_sim.h"
#include "mw_text.h"
#include "meeuuw_hip.h"

namespace mw {

std::string py_float_repr(double v) {
  char b[40];
  return std::string(b, py_repr_chars(v, b));
}

namespace {

// text accumulated in memory, written with one fwrite
struct Text {
  std::string s;
  void put(const char* t) { s += t; }
  void put(const std::string& t) { s += t; }
  template <typename... A>
  void fmt(const char* f, A... a) {
    char b[160];
    int n = snprintf(b, sizeof(b), f, a...);
    s.append(b, n < (int)sizeof(b) ? n : (int)sizeof(b) - 1);
  }
  void repr(double v);
  void e(double v, int prec) {  // printf %.<prec>e
    char b[40];
    s.append(b, sci_chars(v, prec, b));
  }
  void i(long v) {
    char b[24];
    s.append(b, int_chars(v, b));
  }
  int write(const std::string& path) const {
    FILE* f = fopen(path.c_str(), "w");
    if (!f) return MW_EINVAL;
    size_t w = fwrite(s.data(), 1, s.size(), f);
    int rc = (w == s.size()) ? MW_OK : MW_EINVAL;
    if (fclose(f) != 0) rc = MW_EINVAL;
    return rc;
  }
};

// Formats items [0, n) with f(text, i) on up to 8 host threads, in contiguous
// chunks, and appends the chunks in order: the same text as one sequential loop.
template <class F>
void par_append(Text& out, long n, F f) {
  const unsigned hw = std::thread::hardware_concurrency();
  const int T = n < 16384 ? 1 : (int)std::min<unsigned>(8, hw ? hw : 1);
  if (T <= 1) {
    for (long i = 0; i < n; i++) f(out, i);
    return;
  }
  std::vector<Text> part(T);
  std::vector<std::thread> th;
  for (int k = 0; k < T; k++)
    th.emplace_back([&, k] {
      const long a = n * k / T, b = n * (k + 1) / T;
      part[k].s.reserve((size_t)(b - a) * 24);
      for (long i = a; i < b; i++) f(part[k], i);
    });
  for (auto& t : th) t.join();
  for (auto& p : part) out.s += p.s;
}

// numpy's arr.tofile(f, sep=" ", format="%.<prec>e"): values joined by one space
void par_join(Text& out, const double* v, long n, int prec) {
  par_append(out, n, [&](Text& t, long i) {
    if (i) t.s.push_back(' ');
    t.e(v[i], prec);
  });
}

void Text::repr(double v) {
  char b[40];
  s.append(b, py_repr_chars(v, b));
}

struct Series {
  FILE* f = nullptr;
  int open(const std::string& path, const char* header) {
    f = fopen(path.c_str(), "w");
    if (!f) return fail(MW_EINVAL, "output: cannot open " + path);
    fputs(header, f);
    return MW_OK;
  }
  ~Series() {
    if (f) fclose(f);
  }
};

enum { SNAP_SOLUTION = 1, SNAP_SWARM = 2 };

struct Job {
  int istep = 0;
  double stats[MW_NSTATS];
  double tim[MW_NTIMINGS];  // the step's stage times, seconds (timings[] of the reference)
  std::vector<double> nl;   // (xi_u, xi_w, xi_p, xi_T) per Picard iteration
  int snap = 0;  // SNAP_* bits
  int slot = 0;  // staging slot of the snapshot
};

}  // namespace

struct Output {
  Sim* sim = nullptr;
  std::string dir;
  int every_solution = 0, every_swarm = 0, nmat = 1;
  std::vector<std::string> mat_names;
  long nn = 0, nel = 0, np = 0;
  std::vector<double> x, z;       // node coordinates
  std::vector<int32_t> icon;      // 9 x nel, icon_V[k * nel + e]
  static constexpr int NSLOT = 2;
  size_t slot_bytes = 0;
  DBuf<double> dstage[NSLOT];
  unsigned char* hstage[NSLOT] = {nullptr, nullptr};
  hipEvent_t ev_snap[NSLOT] = {}, ev_copy[NSLOT] = {};
  bool slot_busy[NSLOT] = {false, false};
  int next_slot = 0;
  hipStream_t side = nullptr;
  Series vrms, nu, dtf, tavg, ek, tm, tvd, wag, delta, etaavg, pstat, etae, timings, conv, pact;
  FILE* pvd_sol = nullptr;
  FILE* pvd_swarm = nullptr;
  std::mutex mu;
  std::condition_variable cv, cv_done;
  std::deque<Job> q;
  bool stop = false, busy = false;
  int err = MW_OK;
  std::string errmsg;
  std::thread th;
  // byte offsets of the fields in a staging slot
  size_t o_u = 0, o_w = 0, o_q = 0, o_T = 0, o_eta = 0, o_rho = 0, o_qx = 0, o_qz = 0,
         o_etae = 0, o_rhoe = 0, o_cnt = 0, o_sx = 0, o_sz = 0, o_sact = 0, o_smat = 0;

  int create(Sim* s, const char* d, int es, int ew, const char* names);
  int step(const double* stats, int last);
  int flush();
  void run();
  int write_lines(const Job& j);
  int write_solution(const Job& j, const unsigned char* h);
  int write_swarm(const Job& j, const unsigned char* h);
  ~Output();
};

static size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }

int Output::create(Sim* s, const char* d, int es, int ew, const char* names) {
  sim = s;
  dir = d;
  every_solution = es;
  every_swarm = ew;
  nn = s->nn;
  nel = s->nel;
  np = s->np;
  nmat = s->cfg.material_mode == 1 ? 2 : 1;
  {
    std::string all = names ? names : "";
    size_t a = 0;
    while (!all.empty() && a <= all.size()) {
      size_t b = all.find(',', a);
      if (b == std::string::npos) b = all.size();
      mat_names.push_back(all.substr(a, b - a));
      a = b + 1;
    }
    if (!mat_names.empty() && (int)mat_names.size() != nmat)
      return fail(MW_EINVAL, "output_create: " + std::to_string(nmat) +
                                 " material names expected");
    for (const std::string& n : mat_names)
      if (n.find_first_of("'<>&\"") != std::string::npos)
        return fail(MW_EINVAL, "output_create: material names may not hold ' < > & \"");
    if (mat_names.empty())
      for (int m = 0; m < nmat; m++) mat_names.push_back("material " + std::to_string(m));
  }
  mkdir(dir.c_str(), 0755);
  struct stat sb;
  if (stat(dir.c_str(), &sb) != 0 || !S_ISDIR(sb.st_mode))
    return fail(MW_EINVAL, "output_create: cannot create " + dir);
  if (ew > 0) mkdir((dir + "/SWARM").c_str(), 0755);
  RC(vrms.open(dir + "/vrms.ascii", "#time,vrms\n"));
  RC(dtf.open(dir + "/dt.ascii", "#time dt dt1 dt2\n# \n"));
  RC(pstat.open(dir + "/stats_particle.ascii", "#time min(nparticle_e) max(nparticle_e)\n"));
  RC(tm.open(dir + "/total_mass.ascii", "# time total mass\n"));
  RC(ek.open(dir + "/kinetic_energy.ascii", "# time kinetic_energy\n"));
  RC(tvd.open(dir + "/tvd.ascii", "# time total viscous dissipation\n"));
  RC(wag.open(dir + "/wag.ascii", "#time WAG\n"));
  RC(tavg.open(dir + "/T_avrg.ascii", "#time <T>\n"));
  RC(etaavg.open(dir + "/eta_avrg.ascii", "#time eta_avrg\n"));
  RC(delta.open(dir + "/delta_wag_tvd.ascii", "#time delta max(abs(WAG),TVD)\n"));
  RC(etae.open(dir + "/stats_eta_e.ascii", "#time min(eta_e) max(eta_e)\n"));
  RC(nu.open(dir + "/Nu.ascii", "#time Nu\n"));
  RC(timings.open(dir + "/timings.ascii", ""));
  RC(conv.open(dir + "/conv.ascii", ""));
  RC(pact.open(dir + "/active_particles_number.ascii",
               "#time active nparticle_needed nparticle_available )\n"));
  RC(s->set_stage_timing(true));
  const char* head =
      "<?xml version=\"1.0\"?> \n<VTKFile type=\"Collection\" version=\"0.1\" "
      "ByteOrder=\"LittleEndian\"> \n  <Collection> \n";
  if (es > 0) {
    pvd_sol = fopen((dir + "/solution.pvd").c_str(), "w");
    if (!pvd_sol) return fail(MW_EINVAL, "output_create: cannot open solution.pvd");
    fputs(head, pvd_sol);
  }
  if (ew > 0) {
    pvd_swarm = fopen((dir + "/swarm.pvd").c_str(), "w");
    if (!pvd_swarm) return fail(MW_EINVAL, "output_create: cannot open swarm.pvd");
    fputs(head, pvd_swarm);
  }
  x.resize(nn);
  z.resize(nn);
  icon.resize(9 * nel);
  MW_HIP(hipStreamSynchronize(s->st));
  MW_HIP(hipMemcpy(x.data(), s->x_V.p, nn * 8, hipMemcpyDeviceToHost));
  MW_HIP(hipMemcpy(z.data(), s->z_V.p, nn * 8, hipMemcpyDeviceToHost));
  MW_HIP(hipMemcpy(icon.data(), s->icon_V.p, 9 * nel * 4, hipMemcpyDeviceToHost));
  size_t o = 0;
  auto take = [&](size_t bytes) {
    size_t r = o;
    o += al256(bytes);
    return r;
  };
  if (es > 0) {
    o_u = take(nn * 8); o_w = take(nn * 8); o_q = take(nn * 8); o_T = take(nn * 8);
    o_eta = take(nn * 8); o_rho = take(nn * 8); o_qx = take(nn * 8); o_qz = take(nn * 8);
    o_etae = take(nel * 8); o_rhoe = take(nel * 8); o_cnt = take(nel * 4);
  }
  if (ew > 0) {
    o_sx = take(np * 8); o_sz = take(np * 8); o_sact = take(np); o_smat = take(np);
  }
  slot_bytes = o;
  if (slot_bytes > 0) {
    for (int k = 0; k < NSLOT; k++) {
      RC(dstage[k].alloc(slot_bytes / 8));
      MW_HIP(hipHostMalloc((void**)&hstage[k], slot_bytes, hipHostMallocDefault));
      MW_HIP(hipEventCreateWithFlags(&ev_snap[k], hipEventDisableTiming));
      MW_HIP(hipEventCreateWithFlags(&ev_copy[k], hipEventDisableTiming));
    }
    MW_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));  // default priority
  }
  th = std::thread([this] { run(); });
  return MW_OK;
}

int Output::step(const double* stats, int last) {
  Job j;
  memcpy(j.stats, stats, sizeof(j.stats));
  j.istep = sim->istep - 1;  // the step just taken
  if (j.istep < 0) return fail(MW_EINVAL, "output_step: no step taken yet");
  static_assert(MW_NTIMINGS == Sim::NTIMINGS, "timings layout");
  memcpy(j.tim, sim->stage_s, sizeof(j.tim));
  j.tim[0] = j.istep;  // istep + iter_nl / 100, written once per step (iter_nl 0)
  if (sim->cfg.nonlinear) j.nl = sim->nl_hist;
  // meeuuw.py:2913, 2993: every n-th step and the last one
  const bool sol = every_solution > 0 && (j.istep % every_solution == 0 || last);
  const bool swm = every_swarm > 0 && (j.istep % every_swarm == 0 || last);
  if (sol || swm) {
    const int k = next_slot;
    next_slot = (next_slot + 1) % NSLOT;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv_done.wait(lk, [&] { return !slot_busy[k] || err != MW_OK; });
      if (err != MW_OK) return fail(err, errmsg);
      slot_busy[k] = true;
    }
    unsigned char* dst = reinterpret_cast<unsigned char*>(dstage[k].p);
    hipStream_t st = sim->st;
    auto d2d = [&](size_t off, const void* src, size_t bytes) {
      return hipMemcpyAsync(dst + off, src, bytes, hipMemcpyDeviceToDevice, st);
    };
    int rc = MW_OK;
    auto chk = [&](hipError_t e) {
      if (e != hipSuccess && rc == MW_OK)
        rc = fail(MW_EHIP, std::string("output_step: ") + hipGetErrorString(e));
    };
    if (sol) {
      chk(d2d(o_u, sim->un.p, nn * 8));
      chk(d2d(o_w, sim->wn.p, nn * 8));
      chk(d2d(o_q, sim->q.p, nn * 8));
      chk(d2d(o_T, sim->T.p, nn * 8));
      chk(d2d(o_eta, sim->eta_n.p, nn * 8));
      chk(d2d(o_rho, sim->rho_n.p, nn * 8));
      chk(d2d(o_qx, sim->qxn.p, nn * 8));
      chk(d2d(o_qz, sim->qzn.p, nn * 8));
      chk(d2d(o_etae, sim->eta_e.p, nel * 8));
      chk(d2d(o_rhoe, sim->rho_e.p, nel * 8));
      chk(d2d(o_cnt, sim->cnt.p, nel * 4));
    }
    if (swm) {
      chk(d2d(o_sx, sim->sx.p, np * 8));
      chk(d2d(o_sz, sim->sz.p, np * 8));
      chk(d2d(o_sact, sim->sact.p, np));
      chk(d2d(o_smat, sim->smat.p, np));
    }
    chk(hipEventRecord(ev_snap[k], st));
    chk(hipStreamWaitEvent(side, ev_snap[k], 0));
    chk(hipMemcpyAsync(hstage[k], dst, slot_bytes, hipMemcpyDeviceToHost, side));
    chk(hipEventRecord(ev_copy[k], side));
    if (rc != MW_OK) {
      // nothing was queued for this slot: make sure no copy is still landing in it
      (void)hipStreamSynchronize(side);
      std::lock_guard<std::mutex> lk(mu);
      slot_busy[k] = false;
      return rc;
    }
    j.snap = (sol ? SNAP_SOLUTION : 0) | (swm ? SNAP_SWARM : 0);
    j.slot = k;
  }
  {
    std::lock_guard<std::mutex> lk(mu);
    if (err != MW_OK) return fail(err, errmsg);
    q.push_back(j);
  }
  cv.notify_one();
  return MW_OK;
}

void Output::run() {
  for (;;) {
    Job j;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv.wait(lk, [&] { return stop || !q.empty(); });
      if (q.empty()) return;  // stop, nothing left
      j = q.front();
      q.pop_front();
      busy = true;
    }
    int rc = write_lines(j);
    if (j.snap) {
      using clk = std::chrono::steady_clock;
      if (hipEventSynchronize(ev_copy[j.slot]) != hipSuccess && rc == MW_OK) rc = MW_EHIP;
      auto t0 = clk::now();
      if (rc == MW_OK && (j.snap & SNAP_SOLUTION)) rc = write_solution(j, hstage[j.slot]);
      auto t1 = clk::now();
      if (rc == MW_OK && (j.snap & SNAP_SWARM)) rc = write_swarm(j, hstage[j.slot]);
      auto t2 = clk::now();
      if (j.snap & SNAP_SOLUTION)  // output solution to vtu
        j.tim[10] = std::chrono::duration<double>(t1 - t0).count();
      if (j.snap & SNAP_SWARM)  // output swarm to vtu
        j.tim[20] = std::chrono::duration<double>(t2 - t1).count();
    }
    if (rc == MW_OK) {  // meeuuw.py:3371-3375: the step's timings, "%e" joined by " "
      for (int i = 0; i < MW_NTIMINGS; i++) fprintf(timings.f, i ? " %e" : "%e", j.tim[i]);
      fputs(" \n", timings.f);
      if (fflush(timings.f) != 0) rc = MW_EINVAL;
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      if (j.snap) slot_busy[j.slot] = false;
      if (rc != MW_OK && err == MW_OK) {
        err = rc;
        errmsg = "output: writing the files of step " + std::to_string(j.istep) + " in " + dir +
                 " failed";
      }
      busy = false;
    }
    cv_done.notify_all();
  }
}

int Output::write_lines(const Job& j) {
  const double* s = j.stats;
  const double t = s[MW_STAT_TIME];
  const double W = s[MW_STAT_WAG], V = s[MW_STAT_TVD];
  const double aw = W < 0 ? -W : W;
  fprintf(pstat.f, "%.4e %d %d\n", t, (int)s[MW_STAT_MINCOUNT], (int)s[MW_STAT_MAXCOUNT]);
  fprintf(etae.f, "%.4e %.4e %.4e\n", t, s[MW_STAT_ETAE_MIN], s[MW_STAT_ETAE_MAX]);
  fprintf(dtf.f, "%e %e %e %e\n", t, s[MW_STAT_DT], s[MW_STAT_DT1], s[MW_STAT_DT2]);
  fprintf(nu.f, "%e %.6e \n", t, s[MW_STAT_NU]);
  fprintf(vrms.f, "%.4e %.6e \n", t, s[MW_STAT_VRMS]);
  fprintf(tm.f, "%.4e %.4e \n", t, s[MW_STAT_TM]);
  fprintf(ek.f, "%.4e %.4e \n", t, s[MW_STAT_EK]);
  fprintf(tvd.f, "%.4e %.4e \n", t, V);
  fprintf(wag.f, "%.4e %.4e \n", t, W);
  fprintf(delta.f, "%.4e %.4e %.4e\n", t, W + V, aw > V ? aw : V);
  fprintf(tavg.f, "%.4e %.6e \n", t, s[MW_STAT_TAVRG]);
  fprintf(etaavg.f, "%.4e %.4e \n", t, s[MW_STAT_ETA_AVRG]);
  // assess_steady_state.py:54-56, one line per Picard iteration
  const double tol = sim->cfg.tol_nl > 0 ? sim->cfg.tol_nl : 1e-3;
  for (size_t k = 0; k + 4 <= j.nl.size(); k += 4)
    fprintf(conv.f, "%.3e %.3e %.3e %.3e %.3e %.3e %d\n", j.istep + (k / 4) / 200.0, j.nl[k],
            j.nl[k + 1], j.nl[k + 2], j.nl[k + 3], tol, (int)(k / 4));
  // pic_functions.py:1674-1675, on steps where population control added particles
  if (s[MW_STAT_PC_NEEDED] > 0)
    fprintf(pact.f, "%d %d %d %d\n", j.istep, (int)s[MW_STAT_PC_ACTIVE], (int)s[MW_STAT_PC_NEEDED],
            (int)s[MW_STAT_PC_AVAILABLE]);
  for (Series* f : {&pstat, &etae, &dtf, &nu, &vrms, &tm, &ek, &tvd, &wag, &delta, &tavg, &etaavg,
                    &conv, &pact})
    if (fflush(f->f) != 0) return MW_EINVAL;
  return MW_OK;
}

int Output::write_solution(const Job& j, const unsigned char* h) {
  auto D = [&](size_t off) { return reinterpret_cast<const double*>(h + off); };
  const double *u = D(o_u), *w = D(o_w), *qx = D(o_qx), *qz = D(o_qz), *T = D(o_T);
  const int32_t* cnt = reinterpret_cast<const int32_t*>(h + o_cnt);
  const long E = nel;
  Text o;
  o.s.reserve((size_t)nn * 150 + (size_t)nel * 90);
  o.put("<VTKFile type='UnstructuredGrid' version='0.1' byte_order='BigEndian'> \n");
  o.put("<UnstructuredGrid> \n");
  o.fmt("<Piece NumberOfPoints=' %5d ' NumberOfCells=' %5d '> \n", (int)nn, (int)nel);
  o.put("<Points> \n<DataArray type='Float32' NumberOfComponents='3' Format='ascii'> \n");
  // "%.5e %.1e %.5e \n" % (x, 0, z)
  par_append(o, nn, [&](Text& t, long i) {
    t.e(x[i], 5);
    t.put(" 0.0e+00 ");
    t.e(z[i], 5);
    t.put(" \n");
  });
  o.put("</DataArray>\n</Points> \n<PointData Scalars='scalars'>\n");
  o.put("<DataArray type='Float32' NumberOfComponents='3' Name='Velocity ( )' Format='ascii'> \n");
  par_append(o, nn, [&](Text& t, long i) {
    t.e(u[i], 4);
    t.put(" 0.0e+00 ");
    t.e(w[i], 4);
    t.put(" \n");
  });
  o.put("</DataArray>\n<DataArray type='Float32' Name='Pressure' Format='ascii'> \n");
  par_join(o, D(o_q), nn, 4);
  o.put("</DataArray>\n<DataArray type='Float32' Name='Temperature' Format='ascii'> \n");
  par_append(o, nn, [&](Text& t, long i) {
    t.e(T[i], 4);
    t.put(" \n");
  });
  o.put("</DataArray>\n<DataArray type='Float32' Name='Viscosity' Format='ascii'> \n");
  par_join(o, D(o_eta), nn, 4);
  o.put("</DataArray>\n<DataArray type='Float32' Name='Density' Format='ascii'> \n");
  par_join(o, D(o_rho), nn, 5);
  o.put("</DataArray>\n"
        "<DataArray type='Float32' NumberOfComponents='3' Name='Heat flux' Format='ascii'> \n");
  par_append(o, nn, [&](Text& t, long i) {
    t.e(qx[i], 3);
    t.put(" 0.0e+00 ");
    t.e(qz[i], 3);
    t.put(" \n");
  });
  o.put("</DataArray>\n</PointData>\n<CellData Scalars='scalars'>\n");
  o.put("<DataArray type='Float32' Name='Viscosity' Format='ascii'> \n");
  par_join(o, D(o_etae), nel, 4);
  o.put("</DataArray>\n<DataArray type='Float32' Name='Density' Format='ascii'> \n");
  par_join(o, D(o_rhoe), nel, 5);
  o.put("</DataArray>\n<DataArray type='Int32' Name='nb particles' Format='ascii'> \n");
  par_append(o, nel, [&](Text& t, long e) {
    if (e) t.s.push_back(' ');
    t.i(cnt[e]);
  });
  o.put("</DataArray>\n</CellData>\n<Cells>\n");
  o.put("<DataArray type='Int32' Name='connectivity' Format='ascii'> \n");
  par_append(o, nel, [&](Text& t, long e) {
    const int32_t* c = icon.data() + e;
    for (int k = 0; k < 9; k++) {
      t.i(c[k * E]);
      t.s.push_back(k < 8 ? ' ' : '\n');
    }
  });
  o.put("</DataArray>\n<DataArray type='Int32' Name='offsets' Format='ascii'> \n");
  par_append(o, nel, [&](Text& t, long e) {
    t.i((e + 1) * 9);
    t.put(" \n");
  });
  o.put("</DataArray>\n<DataArray type='Int32' Name='types' Format='ascii'>\n");
  for (long e = 0; e < nel; e++) o.put("28 \n");
  o.put("</DataArray>\n</Cells>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n");
  char name[64];
  snprintf(name, sizeof(name), "solution_%04d.vtu", j.istep);
  RC(o.write(dir + "/" + name));
  fprintf(pvd_sol, "    <DataSet timestep=\"%s\" group=\"\" part=\"0\" file=\"%s\"/>  \n",
          py_float_repr(j.stats[MW_STAT_TIME]).c_str(), name);
  return fflush(pvd_sol) == 0 ? MW_OK : MW_EINVAL;
}

int Output::write_swarm(const Job& j, const unsigned char* h) {
  const double* sx = reinterpret_cast<const double*>(h + o_sx);
  const double* sz = reinterpret_cast<const double*>(h + o_sz);
  const uint8_t* act = h + o_sact;
  const uint8_t* mat = h + o_smat;
  std::vector<long> idx;  // the active particles, in particle order
  idx.reserve(np);
  for (long i = 0; i < np; i++)
    if (act[i]) idx.push_back(i);
  const long na = (long)idx.size();
  Text o;
  o.s.reserve((size_t)na * 100);
  o.put("<VTKFile type='UnstructuredGrid' version='0.1' byte_order='BigEndian'> \n");
  o.put("<UnstructuredGrid> \n");
  o.fmt("<Piece NumberOfPoints=' %5d ' NumberOfCells=' %5d '> \n", (int)na, (int)na);
  o.put("<Points> \n<DataArray type='Float32' NumberOfComponents='3' Format='ascii'> \n");
  // numpy tofile(sep=" ") of the (na, 3) array (x, 0, z): repr of every value
  par_append(o, na, [&](Text& t, long k) {
    if (k) t.s.push_back(' ');
    t.repr(sx[idx[k]]);
    t.put(" 0.0 ");
    t.repr(sz[idx[k]]);
  });
  o.put("</DataArray>\n</Points> \n<PointData Scalars='scalars'>\n");
  for (int m = 0; m < nmat; m++) {  // material weights (particle_layout's swarm_wf)
    o.put("<DataArray type='Float32' Name='" + mat_names[m] + "' Format='binary'> \n");
    par_append(o, na, [&](Text& t, long k) {
      if (k) t.s.push_back(' ');
      t.put((nmat == 1 || mat[idx[k]] == m) ? "1.000e+00" : "0.000e+00");
    });
    o.put("</DataArray>\n");
  }
  o.put("<DataArray type='Float32' Name='id' Format='binary'> \n");
  par_append(o, na, [&](Text& t, long k) {
    if (k) t.s.push_back(' ');
    t.i(idx[k]);
  });
  o.put("</DataArray>\n</PointData>\n<Cells>\n");
  // output_swarm_to_vtu.py:186-199: na + 1 connectivity entries 0..na,
  // offsets 1..na+1, na vertex types
  o.put("<DataArray type='Int32' Name='connectivity' Format='ascii'> \n");
  par_append(o, na + 1, [&](Text& t, long i) {
    if (i) t.s.push_back(' ');
    t.i(i);
  });
  o.put("</DataArray>\n<DataArray type='Int32' Name='offsets' Format='ascii'> \n");
  par_append(o, na + 1, [&](Text& t, long i) {
    if (i) t.s.push_back(' ');
    t.i(i + 1);
  });
  o.put("</DataArray>\n<DataArray type='Int32' Name='types' Format='ascii'>\n");
  for (long i = 0; i < na; i++) o.put(i ? " 1" : "1");
  o.put("</DataArray>\n</Cells>\n</Piece>\n</UnstructuredGrid>\n</VTKFile>\n");
  char name[64];
  snprintf(name, sizeof(name), "SWARM/swarm_%04d.vtu", j.istep);
  RC(o.write(dir + "/" + name));
  fprintf(pvd_swarm, "    <DataSet timestep=\"%s\" group=\"\" part=\"0\" file=\"%s\"/>  \n",
          py_float_repr(j.stats[MW_STAT_TIME]).c_str(), name);
  return fflush(pvd_swarm) == 0 ? MW_OK : MW_EINVAL;
}

int Output::flush() {
  std::unique_lock<std::mutex> lk(mu);
  cv_done.wait(lk, [&] { return (q.empty() && !busy) || err != MW_OK; });
  if (err != MW_OK) return fail(err, errmsg);
  return MW_OK;
}

Output::~Output() {
  {
    std::lock_guard<std::mutex> lk(mu);
    stop = true;
  }
  cv.notify_all();
  if (th.joinable()) th.join();
  if (sim) (void)sim->set_stage_timing(false);
  const char* tail = "  </Collection>\n</VTKFile>\n";
  if (pvd_sol) {
    fputs(tail, pvd_sol);
    fclose(pvd_sol);
  }
  if (pvd_swarm) {
    fputs(tail, pvd_swarm);
    fclose(pvd_swarm);
  }
  if (side) (void)hipStreamSynchronize(side);
  for (int k = 0; k < NSLOT; k++) {
    if (hstage[k]) (void)hipHostFree(hstage[k]);
    if (ev_snap[k]) (void)hipEventDestroy(ev_snap[k]);
    if (ev_copy[k]) (void)hipEventDestroy(ev_copy[k]);
  }
  if (side) (void)hipStreamDestroy(side);
}

}  // namespace mw

extern "C" {

int mw_output_create(void* sim, const char* dir, int every_solution, int every_swarm,
                     const char* material_names, void** out) {
  using namespace mw;
  if (!sim || !dir || !out || every_solution < 0 || every_swarm < 0)
    return fail(MW_EINVAL, "output_create: bad arguments");
  *out = nullptr;
  auto* o = new Output();
  int rc = o->create((Sim*)sim, dir, every_solution, every_swarm, material_names);
  if (rc != MW_OK) {
    delete o;
    return rc;
  }
  *out = o;
  return MW_OK;
}

int mw_output_step(void* out, const double* stats_host, int last) {
  if (!out || !stats_host) return mw::fail(MW_EINVAL, "output_step: null argument");
  return ((mw::Output*)out)->step(stats_host, last);
}

int mw_output_flush(void* out) {
  if (!out) return mw::fail(MW_EINVAL, "output_flush: null argument");
  return ((mw::Output*)out)->flush();
}

int mw_output_destroy(void* out) {
  if (!out) return MW_OK;
  auto* o = (mw::Output*)out;
  int rc = o->flush();
  delete o;
  return rc;
}

}  // extern "C"
