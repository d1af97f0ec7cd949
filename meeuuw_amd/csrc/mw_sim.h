// Device-resident timestep of the box-geometry convection model.
#pragma once

#include <hip/hip_runtime.h>

#include "meeuuw_hip.h"
#include "mw_cnn.h"
#include "mw_solver.h"

#include <vector>

namespace mw {

constexpr int FGMRES_MAXM = 64;  // largest FGMRES restart length

struct Sim {
  mw_sim_config cfg;
  hipStream_t st = nullptr;
  BoxGeom g;
  int nel = 0, np = 0;
  long nn = 0, nnP = 0, NV = 0;  // NV = 2*nn velocity dofs
  double hx = 0, hz = 0, volume = 0, L_ref = 1, time = 0;
  double dt1_mem = 1e50, dt2_mem = 1e50;
  int istep = 0;
  bool warm = false;
  int have_prev = 0;              // solutions held in u_prev, p_prev (and u_prev2, p_prev2)
  double dt_hist[3] = {0.0, 0.0, 0.0};  // the last three time steps (older first)
  bool swarm_set = false;

  // mesh and reference tables
  DBuf<double> x_V, z_V;
  DBuf<int> icon_V, icon_P, l2g;
  DBuf<int> incv_ptr, incv_list, incp_ptr, incp_list, incn_ptr, incn_list;
  DBuf<double> N_V, dNdr_V, dNdt_V, N_P, dNdr_P, dNdt_P, N_P_n, dNdr_n, dNdt_n, qc, qw;
  DBuf<double> JxWq, j00q, j01q, j10q, j11q, j00n, j01n, j10n, j11n;
  DBuf<uint8_t> fixV, fixT, left, right, top, bottom;
  DBuf<double> valV, valT, gxq, gzq;
  // swarm
  // particle state in particle order (sx .. st_, su, sw) and, in the order of the
  // element sort, the per-step particle properties and reduced coordinates
  DBuf<double> sx, sz, sr, st_, sr_s, st_s, srho, seta, shcond, shcapa, shprod, salpha, su, sw;
  DBuf<int> siel, order, seg;
  DBuf<uint8_t> sact, smat;  // active flag, material (particle order)
  DBuf<double> sortwork;
  bool sort_clean = false;  // sortwork's counters are zero (after the first sort)
  // fields
  DBuf<double> T, u, p, q, rho_e, eta_e, eta_mg, rho_n, eta_n, hcond_n, hcapa_n, hprod_n, alpha_n;
  DBuf<double> dTdx, dTdz, qxn, qzn, un, wn;
  DBuf<double> exxn, ezzn, exzn;  // nodal strain rate of the last Stokes solve
  DBuf<double> exxq, ezzq, exzq;  // ... on the quadrature points
  // Picard iterations: T_mem, and the iterate before the last (meeuuw.py:1095-1097, 708)
  DBuf<double> Tm, u_mem, p_mem;
  DBuf<double> Tprev;       // the previous step's temperature (energy initial guess)
  bool have_tprev = false;
  bool tm_set = false;            // T_mem holds the initial temperature
  bool nl_keep_guess = false;     // inner iteration: the Stokes solve starts from the last iterate
  std::vector<double> nl_hist;    // (xi_u, xi_w, xi_p, xi_T) per iteration of the last step
  int assess_nl(int iter, bool* inside);
  // population control scratch (only with cfg.population_control)
  DBuf<int> pc_need, pc_offs, pc_flag, pc_pos, pc_avail, pc_tsum, pc_e, pc_tmpl, pc_bad;
  DBuf<double> pc_x, pc_z, pc_r, pc_t;
  int population_control();
  // least-squares qpoint fields: element centres, qpoint coordinates, fits
  DBuf<double> xe, ze, xqc, zqc, lscr, lsce, lsmm;
  DBuf<double> nlsum;             // 8 sums of the assessment, then its block partials
  int iteration(int iter_nl, bool nonlinear, bool* inside, double* dt_out, double* poff_out);
  DBuf<int> cnt;
  DBuf<double> rhoq, etaq, Tq, uq, wq, hcondq, hcapaq, hprodq, alphaq, zq;
  // assembly
  DBuf<double> VV, rhs_f, rhs_h, VVT, rhsT, awork, nodework;
  // solvers
  Reducer R;
  Gmg<double> gmg64;
  Gmg<float> gmg32;
  DBuf<double> Gh, mdiag, minv, invdT;
  DBuf<double> r64tab;  // quadrature tables of the matrix-free FP64 Stokes residual
  DBuf<double> u_prev, p_prev;    // previous Stokes solution (initial-guess extrapolation)
  DBuf<double> u_prev2, p_prev2;  // the one before it (quadratic extrapolation)
  DBuf<float> pm_xa, pm_xb, pm_d;  // Chebyshev solve on the pressure mass matrix
  bool split_schur = false;  // Schur block in a launch per Chebyshev step (MW_OP_PREC32_SPLIT)
  DBuf<double> pm_w, pm_dg, pm_dinv;  // its element weights and diagonal
  DBuf<double> wu[8], wp[6];      // FP64 work vectors (velocity / pressure sized)
  DBuf<double> wT[8];             // energy work vectors
  DBuf<float> V32, Z32, f32a, f32b, f32c, f32d, r32, x32;  // FGMRES basis and work
  DBuf<double> hcol;              // Gram-Schmidt coefficients (two passes, 64 each)
  DBuf<double> hzero;             // zeros (the first-pass coefficients of one-pass Gram-Schmidt)
  DBuf<double> gm;                // device Hessenberg matrix, rotations, rhs, solution
  double* hest = nullptr;         // pinned host copies of the residual estimates
  double* hest_dev = nullptr;     // the same memory as the kernels address it
  hipEvent_t ev[2] = {nullptr, nullptr};
  // matrix-free energy solve (cfg.energy_mf)
  EnergyMF emf;
  DBuf<double> etab, ecoef, edT, epart, escal;
  double* ehost = nullptr;        // pinned host mirror of its residual norms
  bool energy_pending = false;    // true-residual check to finish after the step's sync
  double energy_tol_abs = 0.0;
  DBuf<int> flags;                // 0 outside(locate) 1 r 2 t 3 outside(advect) 4 empty
                                  // 5 min count 6 active particles 7 max count
  DBuf<unsigned long long> eebits;  // eta_e min, max (bit patterns of positive doubles)
  DBuf<double> dscal;             // device scalars
  double stats[MW_NSTATS] = {0};
  std::vector<double> history;    // Stokes residual history of the last solve (relative)

  // learned Schur-complement correction (solver 2)
  Cnn net;
  bool learn_done = false;        // trained (after the first solve)
  bool learned = false;           // the trained network passed validation and is in use
  bool collecting = false;        // store the pressure parts of the Arnoldi vectors
  long ncollect = 0;
  DBuf<float> Qs;                 // learn_pairs x nnP collected vectors
  DBuf<float> lQ, lQm, lYc, lY0, lY, lC, lSY, lGd, lDY, lDC;  // batch x nnP
  DBuf<float> lt, lu, ls, lq, lb, leta;  // velocity-sized work, a right-hand side, log10 eta
  DBuf<double> lsc, lloss;        // per-sample scalars and losses
  DBuf<double> lpart;             // block partials and statistics of the inference path
  int learn_setup();
  int learn_train();
  int collect_pairs(long count);
  void schur_learned(const float* vp, float* zp);
  void schur_apply32(const float* y, float* out);
  void learn_features(int B, const float* Q);
  void update_leta();

  int create(const mw_sim_config& c, hipStream_t s);
  int set_temperature(const double* Th);
  int set_swarm_reduced(const double* rh, const double* th);
  int set_swarm_material(const uint8_t* mh);
  int set_swarm(long n, const double* xh, const double* zh, const double* rh, const double* th,
                const int* ielh, const uint8_t* acth);
  int step();
  int check_empty_elements();
  int get(int field, void* out, long n);
  long size(int field) const;
  int op(int which, const float* x, float* y);

  // stages
  int stokes_solve_plain();
  int stokes_solve_mixed();
  int energy_solve(double dt);
  int pcg_K(const double* b, double* x, double rtol, int maxit, int* its);
  // FP64 residual, its host norm^2 and (bnrm) the right-hand side norm
  int stokes_residual(double* ru, double* rp, double* nrm2, double* bnrm = nullptr);
  // hist_scale < 0: no residual history is recorded
  int fgmres32(const float* r0, float* x, int m, double tol, int maxit, int* its,
               double hist_scale);
  // FP64 FGMRES with the FP32 preconditioner, from the current (u, p): the
  // fallback when FP32 defect correction stops contracting
  DBuf<double> V64, zb64, w64, zero64, h64;
  int fgmres64(double abs_tol, int maxit, double bnrm, double* res2, int* its);
  void prec32(const float* v, float* z);
  int arnoldi_step(int j, int m);
  int initial_guess();
  int build_iteration_graphs(int m);
  std::vector<hipGraphExec_t> it_exec;  // one captured Arnoldi step per column
  bool use_graphs = false;
  bool gs_once = false;  // one classical Gram-Schmidt pass per Arnoldi step
  void apply32(const float* z, float* y);
  double host_dot(const double* a, const double* b, long n);
  int host_norms2(int k, const double* const* v, const long* n, double* out);
  // per-stage times of the last step from HIP events at the stage boundaries,
  // in seconds, at the reference's timings[] indices (print_timings.py)
  static constexpr int NSEV = 160, NTIMINGS = 40;
  bool stage_timing = false;
  hipEvent_t sev[NSEV] = {};
  int sev_slot[NSEV] = {};
  int nsev = 0;
  double stage_s[NTIMINGS] = {};
  int set_stage_timing(bool on);
  void mark(int slot) {
    if (stage_timing && nsev < NSEV) {
      (void)hipEventRecord(sev[nsev], st);
      sev_slot[nsev++] = slot;
    }
  }
  void collect_stage_times();
  ~Sim();
};

}  // namespace mw
