// Particle-in-cell kernels of the curved geometries (quarter, half and eighth
// annulus sectors, and the annulus-formula functions) and the least-squares
// projections of particle fields onto elements.
//
// Same conventions as the box kernels: one thread per particle (or element),
// FP64 throughout, the reference's expressions in its order (built with
// -ffp-contract=off), element indices that leave the mesh reported through
// device flags instead of the reference's out-of-bounds reads.
#include <hip/hip_runtime.h>

#include "mw_capi_impl.h"
#include "mw_common.h"

namespace mw {

constexpr double HALF_PI = 1.5707963267948966;  // np.pi / 2

// locate_particles___annulus (pic_functions.py:142-173)
__global__ void k_locate_annulus(int np, const uint8_t* __restrict__ active,
                                 const double* __restrict__ rad, const double* __restrict__ theta,
                                 double hrad, double htheta, const double* __restrict__ rad_V,
                                 const double* __restrict__ theta_V, const int* __restrict__ icon,
                                 int nel, int nelx, double Rinner, double* __restrict__ r,
                                 double* __restrict__ t, int* __restrict__ iel_out,
                                 int* __restrict__ flags) {
  int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np) return;
  double rr = 0.0, tt = 0.0;
  int iel = 0;
  if (active[ip]) {
    double th = theta[ip], ra = rad[ip];
    int ielx = (int)((HALF_PI - th) / htheta);
    int ielz = (int)((ra - Rinner) / hrad);
    iel = nelx * ielz + ielx;
    if (iel < 0 || iel >= nel) {
      atomicAdd(&flags[0], 1);
      rr = tt = __builtin_nan("");
    } else {
      int n0 = icon[iel];
      rr = ((-th + theta_V[n0]) / htheta - 0.5) * 2;
      tt = ((ra - rad_V[n0]) / hrad - 0.5) * 2;
      if (rr < -1 || rr > 1) atomicAdd(&flags[1], 1);
      if (tt < -1 || tt > 1) atomicAdd(&flags[2], 1);
    }
  }
  r[ip] = rr;
  t[ip] = tt;
  iel_out[ip] = iel;
}

struct PolarCtx {
  const double *u, *w, *rad_V, *theta_V;
  const int* icon;
  double hrad, htheta, Rinner;
  int nelx, nel;
};

__device__ __forceinline__ void interp9(const double* __restrict__ u, const double* __restrict__ w,
                                        const int* __restrict__ icon, int nel, int iel,
                                        double rp, double tp, double& up, double& wp) {
  double N[9], uv[9], wv[9];
  basis_V(rp, tp, N);
#pragma unroll
  for (int k = 0; k < 9; k++) {
    int n = icon[k * (long)nel + iel];
    uv[k] = u[n];
    wv[k] = w[n];
  }
  up = dot9(N, uv);
  wp = dot9(N, wv);
}

// interpolate_vel_on_pt___annulus (pic_functions.py:40-63); false when the
// element index leaves the mesh
__device__ __forceinline__ bool vel_on_pt_polar(const PolarCtx& c, double xp, double yp,
                                                double& up, double& wp, int& iel) {
  double radp = sqrt(xp * xp + yp * yp);
  double thetap = HALF_PI - atan2(xp, yp);
  int ielx = (int)((HALF_PI - thetap) / c.htheta);
  int ielz = (int)((radp - c.Rinner) / c.hrad);
  iel = c.nelx * ielz + ielx;
  if (iel < 0 || iel >= c.nel) {
    up = wp = __builtin_nan("");
    return false;
  }
  int n0 = c.icon[iel];
  double rp = ((-thetap + c.theta_V[n0]) / c.htheta - 0.5) * 2;
  double tp = ((radp - c.rad_V[n0]) / c.hrad - 0.5) * 2;
  interp9(c.u, c.w, c.icon, c.nel, iel, rp, tp, up, wp);
  return true;
}

// deactivation after a move: 0 quarter, 1 half, 2 eighth and annulus
__device__ __forceinline__ bool curved_outside(int variant, double x, double z, double rad,
                                               double Rin, double Rout) {
  if (variant == 0) return x < 0 || z < 0 || rad < Rin || rad > Rout;
  if (variant == 1) return x < 0 || rad < Rin || rad > Rout;
  return x < 0 || rad < Rin || rad > Rout || z < x;
}

// advect_particles___{quarter,half,eighth,annulus} (pic_functions.py:298-596):
// RK1 for every variant, RK2 for quarter and half (whose midpoint test measures
// the radius at A, as the reference); other orders leave particles in place.
__global__ void k_advect_curved(int variant, int rk, double dt, int np, double* __restrict__ x,
                                double* __restrict__ z, double* __restrict__ rad,
                                double* __restrict__ theta, uint8_t* __restrict__ active,
                                PolarCtx c, double Rout, double* __restrict__ su,
                                double* __restrict__ sw, int* __restrict__ flags) {
  int ip = blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= np) return;
  double ou = 0.0, ow = 0.0;
  bool act = active[ip];
  double xp = x[ip], zp = z[ip];
  int iel;
  if (act && rk == 1) {
    if (!vel_on_pt_polar(c, xp, zp, ou, ow, iel)) atomicAdd(&flags[0], 1);
    xp += ou * dt;
    zp += ow * dt;
    double ra = sqrt(xp * xp + zp * zp);
    rad[ip] = ra;
    theta[ip] = HALF_PI - atan2(xp, zp);
    if (curved_outside(variant, xp, zp, ra, c.Rinner, Rout)) act = false;
  } else if (act && rk == 2 && variant <= 1) {
    double xA = xp, zA = zp, uA, wA, uB, wB;
    if (!vel_on_pt_polar(c, xA, zA, uA, wA, iel)) atomicAdd(&flags[0], 1);
    double xB = xA + uA * dt / 2.0, yB = zA + wA * dt / 2.0;
    double rB = sqrt(xA * xA + zA * zA);
    bool outB = variant == 0 ? (xB < 0 || yB < 0 || rB < c.Rinner || rB > Rout)
                             : (xB < 0 || rB < c.Rinner || rB > Rout);
    if (outB) {
      act = false;
    } else {
      if (!vel_on_pt_polar(c, xB, yB, uB, wB, iel)) atomicAdd(&flags[0], 1);
      xp = xA + uB * dt;
      zp = zA + wB * dt;
      ou = uB;
      ow = wB;
      double ra = sqrt(xp * xp + zp * zp);
      rad[ip] = ra;
      theta[ip] = HALF_PI - atan2(xp, zp);
      if (curved_outside(variant, xp, zp, ra, c.Rinner, Rout)) act = false;
    }
  }
  if (!act) xp = zp = 0.0;
  x[ip] = xp;
  z[ip] = zp;
  active[ip] = act;
  su[ip] = ou;
  sw[ip] = ow;
}

// interpolate_vel_on_pt___box (pic_functions.py:17-37) and ___annulus at many points
__global__ void k_vel_on_pts(int polar, int n, const double* __restrict__ xp,
                             const double* __restrict__ zp, PolarCtx c, double hx, double hz,
                             const double* __restrict__ x_V, const double* __restrict__ z_V,
                             double* __restrict__ up, double* __restrict__ wp,
                             int* __restrict__ iel_out, int* __restrict__ flags) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double uu, ww;
  int iel;
  bool ok;
  if (polar) {
    ok = vel_on_pt_polar(c, xp[i], zp[i], uu, ww, iel);
  } else {
    int ielx = (int)(xp[i] / hx);
    int ielz = (int)(zp[i] / hz);
    iel = c.nelx * ielz + ielx;
    ok = iel >= 0 && iel < c.nel;
    if (ok) {
      int n0 = c.icon[iel];
      double rp = ((xp[i] - x_V[n0]) / hx - 0.5) * 2;
      double tp = ((zp[i] - z_V[n0]) / hz - 0.5) * 2;
      interp9(c.u, c.w, c.icon, c.nel, iel, rp, tp, uu, ww);
    } else {
      uu = ww = __builtin_nan("");
    }
  }
  if (!ok) atomicAdd(&flags[0], 1);
  up[i] = uu;
  wp[i] = ww;
  iel_out[i] = iel;
}

// Gaussian elimination with partial pivoting, n <= 4 (np.linalg.solve is LAPACK
// gesv: the same LU, equal up to rounding)
__device__ __forceinline__ void small_solve(int n, double (&A)[4][4], double (&b)[4]) {
  for (int k = 0; k < n; k++) {
    int p = k;
    for (int i = k + 1; i < n; i++)
      if (fabs(A[i][k]) > fabs(A[p][k])) p = i;
    if (p != k) {
      for (int j = 0; j < 4; j++) {
        double s = A[k][j];
        A[k][j] = A[p][j];
        A[p][j] = s;
      }
      double s = b[k];
      b[k] = b[p];
      b[p] = s;
    }
    for (int i = k + 1; i < n; i++) {
      double f = A[i][k] / A[k][k];
      for (int j = k; j < n; j++) A[i][j] -= f * A[k][j];
      b[i] -= f * b[k];
    }
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = b[i];
    for (int j = i + 1; j < n; j++) s -= A[i][j] * b[j];
    b[i] = s / A[i][i];
  }
}

// compute_ls_coefficients_P1 (nb 3) / _Q1 (nb 4) (pic_functions.py:1097-1285):
// one thread per element over its particles in increasing particle order
// (stable sort by element; every particle counts, as the reference's mask).
// rho, eta are indexed by particle (sorted_props 0) or by sorted position j
// (sorted_props 1: the timestep keeps particle properties in sort order).
__global__ void k_ls_coefficients(int nb, int nel, const double* __restrict__ x_e,
                                  const double* __restrict__ z_e, const int* __restrict__ seg,
                                  const int* __restrict__ order, const double* __restrict__ x,
                                  const double* __restrict__ z, const double* __restrict__ rho,
                                  const double* __restrict__ eta, double* __restrict__ cr,
                                  double* __restrict__ ce, double* __restrict__ mm,
                                  int sorted_props) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nel) return;
  double A[4][4] = {}, br[4] = {}, be[4] = {};
  double rmin = __builtin_inf(), rmax = -__builtin_inf(), emin = __builtin_inf(),
         emax = -__builtin_inf();
  const int b0 = seg[e], b1 = seg[e + 1];
  for (int j = b0; j < b1; j++) {
    int ip = order[j];
    double sx = x[ip] - x_e[e], sz = z[ip] - z_e[e];
    const int q = sorted_props ? j : ip;
    double ro = rho[q], et = eta[q];
    A[0][1] += sx;
    A[0][2] += sz;
    A[1][1] += sx * sx;
    A[1][2] += sx * sz;
    A[2][2] += sz * sz;
    if (nb == 4) {
      A[0][3] += sx * sz;
      A[1][3] += sx * sx * sz;
      A[2][3] += sx * sz * sz;
      A[3][3] += sx * sx * sz * sz;
    }
    be[0] += et;
    be[1] += et * sx;
    be[2] += et * sz;
    br[0] += ro;
    br[1] += ro * sx;
    br[2] += ro * sz;
    if (nb == 4) {
      be[3] += et * sx * sz;
      br[3] += ro * sx * sz;
    }
    rmin = fmin(rmin, ro);
    rmax = fmax(rmax, ro);
    emin = fmin(emin, et);
    emax = fmax(emax, et);
  }
  A[0][0] = (double)(b1 - b0);
  for (int i = 0; i < nb; i++)
    for (int j = 0; j < i; j++) A[i][j] = A[j][i];
  double A2[4][4];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) A2[i][j] = A[i][j];
  small_solve(nb, A, be);
  small_solve(nb, A2, br);
  for (int k = 0; k < nb; k++) {
    ce[(long)e * nb + k] = be[k];
    cr[(long)e * nb + k] = br[k];
  }
  mm[4L * e + 0] = rmin;
  mm[4L * e + 1] = rmax;
  mm[4L * e + 2] = emin;
  mm[4L * e + 3] = emax;
}

// limiter (pic_functions.py:1067-1087) element-wise; c1, c2 in place.  The
// coefficients of element i are c*[i * cs], its bounds S[i * ms], L[i * ms].
__global__ void k_limiter(int n, const double* __restrict__ c0, double* __restrict__ c1,
                          double* __restrict__ c2, const double* __restrict__ S,
                          const double* __restrict__ L, double h, int cs, int ms) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const long i = (long)e * cs, im = (long)e * ms;
  double a = c1[i], b = c2[i];
  double ku = fmin(c0[i] - S[im], L[im] - c0[i]);
  double kl = (fabs(a) + fabs(b)) * h / 2;
  if (kl > ku) {
    double cmax = ku * 2 / h;
    a = (a < 0 ? -1.0 : 1.0) * fmin(fabs(a), cmax);
    b = (b < 0 ? -1.0 : 1.0) * fmin(fabs(b), cmax);
  }
  kl = (fabs(a) + fabs(b)) * h / 2;
  if (kl > ku) {
    double cc = (kl - ku) / h;
    a = (a < 0 ? -1.0 : 1.0) * (fabs(a) - cc);
    b = (b < 0 ? -1.0 : 1.0) * (fabs(b) - cc);
  }
  c1[i] = a;
  c2[i] = b;
}

}  // namespace mw

using namespace mw;

#define ST(s) ((hipStream_t)(s))

namespace mw {

int ls_coefficients_sorted(int nb, int nel, const double* x_e, const double* z_e, const int* seg,
                           const int* order, const double* x, const double* z,
                           const double* rho_sorted, const double* eta_sorted, double* cr,
                           double* ce, double* mm, hipStream_t st) {
  k_ls_coefficients<<<nblocks(nel, 128), 128, 0, st>>>(nb, nel, x_e, z_e, seg, order, x, z,
                                                       rho_sorted, eta_sorted, cr, ce, mm, 1);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

int limiter_strided(int n, const double* c0, double* c1, double* c2, const double* S,
                    const double* L, double h, int cs, int ms, hipStream_t st) {
  k_limiter<<<nblocks(n, 256), 256, 0, st>>>(n, c0, c1, c2, S, L, h, cs, ms);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

}  // namespace mw

extern "C" {

int mw_locate_particles_annulus(int np, const uint8_t* active, const double* rad,
                                const double* theta, double hrad, double htheta,
                                const double* rad_V, const double* theta_V,
                                const int32_t* icon_V, int nel, int nelx, double Rinner,
                                double* r, double* t, int32_t* iel, int32_t* flags,
                                void* stream) {
  if (np < 0 || nel <= 0 || nelx <= 0 || !(hrad > 0) || !(htheta > 0) || !flags)
    return fail(MW_EINVAL, "locate_particles_annulus: bad arguments");
  MW_HIP(hipMemsetAsync(flags, 0, 3 * sizeof(int32_t), ST(stream)));
  if (np == 0) return MW_OK;
  if (!active || !rad || !theta || !rad_V || !theta_V || !icon_V || !r || !t || !iel)
    return fail(MW_EINVAL, "locate_particles_annulus: null pointer argument");
  k_locate_annulus<<<nblocks(np, 256), 256, 0, ST(stream)>>>(np, active, rad, theta, hrad, htheta,
                                                            rad_V, theta_V, icon_V, nel, nelx,
                                                            Rinner, r, t, iel, flags);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

int mw_advect_particles_curved(int variant, int rk, double dt, int np, double* x, double* z,
                               double* rad, double* theta, uint8_t* active, const double* u,
                               const double* w, double Rinner, double Router, double hrad,
                               double htheta, int nelx, int nel, const int32_t* icon_V,
                               const double* rad_V, const double* theta_V, double* su,
                               double* sw, int32_t* flags, void* stream) {
  if (variant < 0 || variant > 3 || np < 0 || nel <= 0 || nelx <= 0 || !(hrad > 0) ||
      !(htheta > 0) || !flags)
    return fail(MW_EINVAL, "advect_particles_curved: bad arguments");
  MW_HIP(hipMemsetAsync(flags, 0, 2 * sizeof(int32_t), ST(stream)));
  if (np == 0) return MW_OK;
  if (!x || !z || !rad || !theta || !active || !u || !w || !icon_V || !rad_V || !theta_V ||
      !su || !sw)
    return fail(MW_EINVAL, "advect_particles_curved: null pointer argument");
  PolarCtx c{u, w, rad_V, theta_V, icon_V, hrad, htheta, Rinner, nelx, nel};
  k_advect_curved<<<nblocks(np, 256), 256, 0, ST(stream)>>>(variant, rk, dt, np, x, z, rad, theta,
                                                           active, c, Router, su, sw, flags);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

int mw_interpolate_vel_on_pts(int polar, int n, const double* xp, const double* zp,
                              const double* u, const double* w, double hx, double hz,
                              const double* x_V, const double* z_V, double hrad, double htheta,
                              const double* rad_V, const double* theta_V, double Rinner,
                              int nelx, int nel, const int32_t* icon_V, double* up, double* wp,
                              int32_t* iel, int32_t* flags, void* stream) {
  if (n < 0 || nel <= 0 || nelx <= 0 || !flags)
    return fail(MW_EINVAL, "interpolate_vel_on_pts: bad arguments");
  if (polar ? !(hrad > 0 && htheta > 0 && rad_V && theta_V) : !(hx > 0 && hz > 0 && x_V && z_V))
    return fail(MW_EINVAL, "interpolate_vel_on_pts: bad geometry arguments");
  MW_HIP(hipMemsetAsync(flags, 0, sizeof(int32_t), ST(stream)));
  if (n == 0) return MW_OK;
  if (!xp || !zp || !u || !w || !icon_V || !up || !wp || !iel)
    return fail(MW_EINVAL, "interpolate_vel_on_pts: null pointer argument");
  PolarCtx c{u, w, rad_V, theta_V, icon_V, hrad, htheta, Rinner, nelx, nel};
  k_vel_on_pts<<<nblocks(n, 256), 256, 0, ST(stream)>>>(polar, n, xp, zp, c, hx, hz, x_V, z_V, up,
                                                        wp, iel, flags);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

int mw_ls_coefficients(int nb, int nel, const double* x_e, const double* z_e, int np,
                       const double* x, const double* z, const int32_t* iel, const double* rho,
                       const double* eta, double* coef_rho, double* coef_eta, double* minmax,
                       void* work, void* stream) {
  if ((nb != 3 && nb != 4) || nel <= 0 || np <= 0)
    return fail(MW_EINVAL, "ls_coefficients: bad sizes");
  if (!x_e || !z_e || !x || !z || !iel || !rho || !eta || !coef_rho || !coef_eta || !minmax ||
      !work)
    return fail(MW_EINVAL, "ls_coefficients: null pointer argument");
  uintptr_t b = ((uintptr_t)work + 255) & ~(uintptr_t)255;
  auto take = [&](size_t bytes) {
    void* p = (void*)b;
    b += (bytes + 255) & ~(size_t)255;
    return p;
  };
  int* order = (int*)take(sizeof(int) * (size_t)np);
  int* seg = (int*)take(sizeof(int) * (size_t)(nel + 1));
  void* temp = take(sort_work_bytes(np, nel));
  RC(sort_by_element(np, nel, nullptr, iel, order, seg, temp, ST(stream)));
  k_ls_coefficients<<<nblocks(nel, 128), 128, 0, ST(stream)>>>(
      nb, nel, x_e, z_e, seg, order, x, z, rho, eta, coef_rho, coef_eta, minmax, 0);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

size_t mw_ls_workspace_bytes(int np, int nel) {
  if (np <= 0 || nel <= 0) return 0;
  return (((size_t)np * sizeof(int) + 255) & ~(size_t)255) +
         (((size_t)(nel + 1) * sizeof(int) + 255) & ~(size_t)255) + sort_work_bytes(np, nel) +
         768;
}

int mw_limiter(int n, const double* c0, double* c1, double* c2, const double* S,
               const double* L, double h, void* stream) {
  if (n < 0 || !(h > 0)) return fail(MW_EINVAL, "limiter: bad arguments");
  if (n == 0) return MW_OK;
  if (!c0 || !c1 || !c2 || !S || !L) return fail(MW_EINVAL, "limiter: null pointer argument");
  k_limiter<<<nblocks(n, 256), 256, 0, ST(stream)>>>(n, c0, c1, c2, S, L, h, 1, 1);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

}  // extern "C"
