// Device linear algebra for the timestep solvers: vector kernels with
// deterministic reductions, the structured geometric-multigrid hierarchy for
// the Q2 velocity block, element-by-element operators read from the assembled
// COO values, and the Krylov drivers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace mw {

// ---------------------------------------------------------------- memory ---
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  int alloc(size_t count);  // returns status
  operator T*() const { return p; }
};

// ------------------------------------------------------------- reductions --
// Deterministic dot products: fixed grid, block partials, one final block.
constexpr int DOT_BLOCKS = 512;
constexpr int DOT_BS = 256;

// blocks of the partial-sum kernels over n elements
inline int dot_blocks(long n) {
  const long b = (n + DOT_BS - 1) / DOT_BS;
  return (int)(b < DOT_BLOCKS ? (b < 1 ? 1 : b) : DOT_BLOCKS);
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// sum_b partial[b*k + j] over nblk <= DOT_BLOCKS blocks by one wavefront
// (the sum is complete in lane 0): lane-strided terms in increasing block
// order, then a tree sum.  All loads are issued before the adds (indices past
// nblk re-read a valid partial and are not added).
__device__ __forceinline__ double reduce_column(const double* __restrict__ partial, int nblk,
                                                int k, int j, int lane) {
  constexpr int NT = DOT_BLOCKS / 64;
  double v[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int b = lane + 64 * t;
    v[t] = partial[(long)(b < nblk ? b : nblk - 1) * k + j];
  }
  double s = 0.0;
#pragma unroll
  for (int t = 0; t < NT; t++) s = (lane + 64 * t < nblk) ? s + v[t] : s;
  return wave_sum(s);
}

struct Reducer {
  DBuf<double> partial;  // DOT_BLOCKS * 33
  DBuf<double> result;   // 64 slots of device scalars
  DBuf<double> host_pinned_dummy;
  double* hbuf = nullptr;  // pinned host mirror (64)
  double* hbuf_dev = nullptr;  // the same memory as kernels address it
  int init();
  ~Reducer();
};

template <class T>
int dot(const T* a, const T* b, long n, double* out, Reducer& R, hipStream_t st);
// out[j] = <V_j, w> for j < k (columns of V with leading dimension ld)
template <class T>
int multi_dot(const T* V, long ld, int k, const T* w, long n, double* out, Reducer& R,
              hipStream_t st);

// fused classical Gram-Schmidt passes (k <= 32): w_out = w_in - V h (h null:
// no update), out[j] = <V_j, w_out> (and out[k] = <w_out, w_out> with_norm);
// w -= V h with out[0] = <w, w>; out = (w - V h) * scale[0]
template <class T>
int gs_project(const T* V, long ld, int k, const double* h, const T* w_in, T* w_out, long n,
               double* out, Reducer& R, hipStream_t st, bool with_norm = false,
               bool finish = true);
template <class T>
int gs_update_norm(const T* V, long ld, int k, const double* h, T* w, long n, double* out,
                   Reducer& R, hipStream_t st);
template <class T>
int gs_update_scale(const T* V, long ld, int k, const double* h, const T* w, long n,
                    const double* scale, T* out, hipStream_t st);

// y = a*x + b*y with device scalar multipliers (nullptr -> 1)
template <class T>
void axpby(long n, double a, const double* da, const T* x, double b, const double* db, T* y,
           hipStream_t st);
template <class T>
void fill(long n, T v, T* y, hipStream_t st);
template <class T, class U>
void convert(long n, const T* x, U* y, hipStream_t st);

// ------------------------------------------------------ structured levels --
// One level of the structured Q2 box hierarchy.  The velocity operator of an
// element is eta_e * Khat (Khat: 18x18 stiffness of a unit-viscosity
// hx*hz rectangle), with rows and columns of fixed dofs replaced by the
// diagonal, exactly as the assembled matrix has them.
template <class T>
struct Level {
  int nelx = 0, nelz = 0, nnx = 0, nnz = 0;
  long nn = 0, ndof = 0;
  double hx = 0, hz = 0;
  DBuf<T> eta;      // nel
  DBuf<uint8_t> fix;  // ndof
  DBuf<T> diag;     // ndof
  DBuf<T> Khat;     // 18*18
  DBuf<T> invdiag;  // ndof
  DBuf<T> r, x, x2, d, b;  // work vectors
  DBuf<T> coef;           // Chebyshev (c1, c2) per smoothing step
  DBuf<T> pvec;           // power-iteration vector, kept between updates
  double lmax = 0;  // largest eigenvalue of D^-1 A
  bool fine = false;
};

constexpr int MAX_LEVELS = 16;

template <class T>
struct Gmg {
  std::vector<Level<T>*> lev;
  int nsmooth = 2;           // Chebyshev degree (pre and post)
  int nsmooth_coarse = 0;    // degree on the levels below the finest (0: nsmooth)
  int steps(int l) const { return (l == 0 || nsmooth_coarse <= 0) ? nsmooth : nsmooth_coarse; }
  bool fuse_tail = true;     // run the small levels as one single-workgroup kernel
  bool inline_coarse = false;  // launch the coarse levels' kernels directly (while capturing)
  int tail_from = 0;         // first fused level (0: none)
  bool bottom = false;       // last smoothed level + coarsest solve as one workgroup
  int bottom_ml = 0;         // or the last 2..3 smoothed levels + coarsest as one workgroup
  int allow_bottom = 3;      // smoothed levels fused with the coarsest solve at most (0: none)
  int coarse_max_dof = 60;   // coarsening stops at the first level with at most this many
                             // unknowns (dense inverse; above 64 by blocked Gauss-Jordan)
  size_t tail_lds = 0;       // LDS bytes of the fused tail
  DBuf<T> coarse_inv;        // dense inverse of the coarsest operator
  DBuf<T> coarse_next;       // the next one (blocked Gauss-Jordan, formed one solve ahead)
  bool have_next = false;
  int side_k = -1;           // next pivot block of that inverse to queue (-1: none pending)
  bool solver_pumps = false; // the solver queues the sweeps from its iteration loop
  int pump_side(int nmax);   // queue up to nmax pending sweeps (then the extraction)
  DBuf<double> coarse_a;     // FP64 dense work (n*n)
  DBuf<double> pnorm;        // power-iteration norms, 2 per level
  int power_its = 12;        // power iterations of the first update (from a fixed hash vector)
  int power_its_warm = 0;    // later updates, from the last update's vectors (0: none kept)
  int eta_coarsen = 0;       // coarse viscosity from the 4 children: 0 arithmetic, 1 geometric,
                             // 2 harmonic mean
  bool have_pvec = false;
  int ncoarse = 0;
  hipGraph_t graph = nullptr;          // levels 1.. of the V-cycle
  hipGraphExec_t graph_exec = nullptr;
  // the coarsest inverse is formed on a side stream (one workgroup) while the
  // caller's stream goes on; the first V-cycle after an update waits for it
  hipStream_t side = nullptr;
  hipEvent_t ev_dense = nullptr, ev_inv = nullptr;
  bool inv_pending = false;
  ~Gmg();
  // build the hierarchy for an nelx x nelz box with fine fixed-dof mask
  int setup(int nelx, int nelz, double Lx, double Lz, const uint8_t* fine_fix, hipStream_t st);
  // per-solve update: fine element viscosities (double) -> all levels,
  // diagonals, eigenvalue estimates, coarse inverse
  int update(const double* eta_fine, Reducer& R, hipStream_t st);
  // z = V-cycle(r) on the fine level (z overwritten)
  int vcycle(const T* r, T* z, hipStream_t st);
  int vcycle_level(int l, hipStream_t st);  // solves lev[l]: b -> x
  // transfers between level l and l+1: bc = P^T rf (masked), xf += P xc (free fine dofs)
  void restrict_to(int l, const T* rf, T* bc, hipStream_t st);
  void prolong_add(int l, const T* xc, T* xf, hipStream_t st);
  // `reps` back-to-back Chebyshev steps (step index 1: x and d both read) on
  // level l, ping-ponging x and x2; for kernel timing at the real shape
  void smooth_steps(int l, int reps, hipStream_t st);
};

// y = A_l x on a structured level (masked)
template <class T>
void level_apply(const Level<T>& L, const T* x, T* y, hipStream_t st);
// Stokes operator on the finest level L (velocity block eta_e * Khat) with the
// divergence template Gh (18x4): y_u = K z_u + G z_p, y_p = G^T z_u, vectors
// laid out [u (nv), p], in one launch
template <class T>
void stokes_apply_fine(const Level<T>& L, const double* Gh, const T* z, T* y, long nv,
                       hipStream_t st);
// z_p = -minv v_p (or -xs when given) and t = v_u - G z_p (fixed rows: v_u), in one launch
template <class T>
void schur_front(const Level<T>& L, const double* Gh, const double* minv, const T* v, T* zp,
                 T* t, long nv, hipStream_t st, const T* xs = nullptr, double scale = 1.0);
// element weights w = area / (36 eta_e) (nel), the diagonal dg of M_eta (Q1
// pressure mass matrix weighted by 1/eta_e, eta per element) and 1/dg (nnP)
int pmass_prepare(int nelx, int nelz, double area, const double* eta, double* w, double* dg,
                  double* dinv, hipStream_t st);
// steps >= 2 Chebyshev-Jacobi steps from 0 on M_eta x = b with the arrays of
// pmass_prepare; *x: xa or xb, where the last step put it
template <class T>
int pmass_cheb_solve(int nelx, int nelz, const double* w, const double* dg, const double* dinv,
                     const T* b, T* xa, T* xb, T* d, int steps, T** x, hipStream_t st);
// pmass_cheb_solve followed by schur_front (z_p = -scale x), in one launch
// and with the same results; steps in [2, SCHUR_FUSED_MAX]
constexpr int SCHUR_FUSED_MAX = 4;
template <class T>
int schur_front_cheb(const Level<T>& L, const double* Gh, const double* w, const double* dg,
                     const double* dinv, const T* v, T* zp, T* t, long nv, int steps,
                     double scale, hipStream_t st);

// ----------------------------------------------- assembled (COO) operators --
// Element-by-element application of the assembled Stokes matrix from its COO
// values in the reference order (468 per element), gathered per node in
// increasing element order (deterministic).  Structured box only.
struct BoxGeom {
  int nelx, nelz, nnx, nnz;
  long nn, nnP;
};
// yu = K u + G p ; yp = G^T u   (any output may be null)
// ru = f - (K u + G p), rp = h - G^T u on the box, matrix-free in FP64 from
// per-quadrature-point viscosities (the assembled operator, Dirichlet rows
// masked as assembled); tab (207 doubles) from residual64_tables
int stokes_residual64(const BoxGeom& g, const uint8_t* fix, const double* tab, const double* etaq,
                      const double* u, const double* p, const double* f, const double* h,
                      double gscale, double* ru, double* rp, hipStream_t st);
// tab = [dN/dx (81), dN/dz (81), N_P (36), JxW (9)] at the quadrature points of
// element 0 (every element of the box is the same rectangle)
int residual64_tables(const double* dNdr_V, const double* dNdt_V, const double* N_P,
                      const double* j00q, const double* j01q, const double* j10q,
                      const double* j11q, const double* JxWq, double* tab, hipStream_t st);
void stokes_apply_vv(const BoxGeom& g, const double* VV, const double* u, const double* p,
                     double* yu, double* yp, hipStream_t st);
// y = A T with the energy COO values (81 per element)
void energy_apply_vv(const BoxGeom& g, const double* VV, const double* x, double* y,
                     hipStream_t st);
// inverse diagonals of the assembled K (from VV) and of the energy matrix
void stokes_invdiag_vv(const BoxGeom& g, const double* VV, double* invdiagK, hipStream_t st);
void energy_invdiag_vv(const BoxGeom& g, const double* VV, double* invdiag, hipStream_t st);

// Crank-Nicolson energy operator on the box, matrix-free in FP64 from its
// quadrature-point coefficients, and its Jacobi-preconditioned BiCGStab solve
struct EnergyMF {
  BoxGeom g;
  const uint8_t* fix = nullptr;  // Dirichlet mask (nn)
  const double* tab = nullptr;   // N[81], dN/dx[81], dN/dz[81] (q*9 + node), JxW[9]
  double* coef = nullptr;        // nel*36: per element [cm, cu, cw, ck][9 points]
  double* d = nullptr;           // diagonal of the assembled operator (nn)
  double* invd = nullptr;        // its inverse
  double* part = nullptr;        // block partial sums (energy_mf_partials)
  double* S = nullptr;           // device scalars of the solve (8)
  double* host = nullptr;        // pinned host mirror (4) and its device address
  double* host_dev = nullptr;
};
long energy_mf_partials(const BoxGeom& g);
int energy_mf_tables(const double* N_V, const double* dNdr_V, const double* dNdt_V,
                     const double* j00q, const double* j01q, const double* j10q,
                     const double* j11q, const double* JxWq, double* tab, hipStream_t st);
// b (right-hand side with the Dirichlet lifting), E.d, E.invd and E.coef for
// one solve from the old temperature and the quadrature-point fields
int energy_mf_rhs(const EnergyMF& E, const double* Told, const double* val, const double* uq,
                  const double* wq, const double* hcapaq, const double* hcondq,
                  const double* hprodq, double rho0, double dt, double* b, hipStream_t st);
// solve A x = b from x to |b - A x| <= tol_rel |b| (w: 8 work vectors of nn);
// the true residual is queued after the loop and read with energy_mf_true_residual
// once the stream has been synchronised
int energy_mf_solve(const EnergyMF& E, double* x, const double* b, double* const w[8],
                    double tol_rel, int maxit, hipEvent_t ev[2], int* its, hipStream_t st);
double energy_mf_true_residual(const EnergyMF& E);
double energy_mf_bnorm(const EnergyMF& E);

// lumped pressure mass matrix scaled by 1/eta_e (Schur complement approximation)
void pressure_mass_eta_lumped(const BoxGeom& g, double area, const double* eta_e, double* mdiag,
                              hipStream_t st);

}  // namespace mw

namespace mw {

// Live kernel timing with HIP events on the launching stream.  A class id is
// attached to a launch site; when profiling is on, every launch of that class
// is bracketed by two events and the elapsed times are summed after the step.
enum ProfClass {
  PROF_STOKES_VV = 0,   // FP64 Stokes operator: assembled (velocity rows) or matrix-free residual
  PROF_CHEB_FINE,       // Chebyshev step on the finest multigrid level
  PROF_ASSEMBLY,        // Stokes element kernel (matrices and right-hand sides, or the latter)
  PROF_ENERGY_VV,       // assembled energy operator
  PROF_GS,              // FGMRES Gram-Schmidt update
  PROF_ADVECT,          // RK advection
  PROF_NCLASSES
};

struct Prof {
  bool on = false;
  std::vector<hipEvent_t> ev[PROF_NCLASSES];
  int used[PROF_NCLASSES] = {0};
  double total_ms[PROF_NCLASSES] = {0};
  long count[PROF_NCLASSES] = {0};
  // capture on this thread (graph building) records no events; per thread, so
  // batch worker threads never write shared state for it
  static thread_local int muted;
  void begin(int c, hipStream_t st);
  void end(int c, hipStream_t st);
  void collect();  // after a stream synchronisation
  void reset();
  ~Prof();
};
Prof& prof();

}  // namespace mw
