// Mesh and connectivity of the Q2 x Q1 discretisation on the device: box,
// the curved sectors (eighth, quarter, half) and the full annulus.
//
// Reference behaviour restated here:
//   build_velocity_mesh (m_V = 9)       build_velocity_mesh.py:17-46, 110-200
//   build_velocity_connectivity (Q2)    build_velocity_connectivity.py:14-136
//   build_pressure_mesh                 build_pressure_mesh.py:10-88
//   build_pressure_connectivity         build_pressure_connectivity.py:10-48
//
// One thread per node or element.  Node coordinates are formed first on the
// reference (i, j) lattice with the reference's expressions (i * hx / 2 for
// the Q2 nodes, i * hx for the Q1 nodes), the boundary and mid-line flags are
// taken there, and the curved geometries then map (x, z) to (theta, r) as the
// reference does.  Integer outputs (connectivity, flags) are exact; the
// curved coordinates go through the device's cos, sin and atan2.
#include "mw_common.h"
#include "mw_capi_impl.h"

namespace mw {

constexpr double MESH_EPS = 1e-9;  // constants.py:7
constexpr double MESH_PI = 3.141592653589793;  // the double nearest pi (numpy's pi)
#define ST(s) ((hipStream_t)(s))


// the sector and annulus maps of build_velocity_mesh.py:185-200 (and of
// build_pressure_mesh.py:62-79, the same expressions)
__device__ __forceinline__ void curved_map(int geom, double Rinner, double Router,
                                           double opening_angle, double& x, double& z,
                                           double& rad, double& theta) {
  if (geom == MW_GEOM_ANNULUS) {
    const double xi = x, zi = z;
    const double t = xi * 2 * MESH_PI;
    x = cos(t) * (Rinner + zi * (Router - Rinner));
    z = sin(t) * (Rinner + zi * (Router - Rinner));
    rad = Rinner + zi * (Router - Rinner);
    theta = atan2(z, x);
    if (theta < 0.0) theta += 2.0 * MESH_PI;
  } else if (geom != MW_GEOM_BOX) {
    rad = Rinner + z * (Router - Rinner);
    theta = MESH_PI / 2 - x * opening_angle;
    x = rad * cos(theta);
    z = rad * sin(theta);
  }
}

// flags: 7 rows of nn (top, bot, left, right, middleH, middleV, hull)
__global__ void k_velocity_mesh(int geom, int nelx, int nelz, int nnx, int nnz, double Lx,
                                double Lz, double hx, double hz, double Rinner, double Router,
                                double opening_angle, double* __restrict__ x_V,
                                double* __restrict__ z_V, double* __restrict__ rad_V,
                                double* __restrict__ theta_V, uint8_t* __restrict__ flags) {
  const long c = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long nn = (long)nnx * nnz;
  if (c >= nn) return;
  const int i = (int)(c % nnx), j = (int)(c / nnx);
  double x = i * hx / 2, z = j * hz / 2;
  bool top = j == 2 * nelz, bot = j == 0, left = false, right = false, mh = false, mv = false;
  if (geom != MW_GEOM_ANNULUS) {
    left = i == 0;
    right = i == 2 * nelx;
    mv = fabs(x / Lx - 0.5) < MESH_EPS;
    mh = fabs(z / Lz - 0.5) < MESH_EPS;
  }
  const bool hull = top || bot || left || right;
  double rad = 0.0, theta = 0.0;
  curved_map(geom, Rinner, Router, opening_angle, x, z, rad, theta);
  x_V[c] = x;
  z_V[c] = z;
  rad_V[c] = rad;
  theta_V[c] = theta;
  flags[0 * nn + c] = top;
  flags[1 * nn + c] = bot;
  flags[2 * nn + c] = left;
  flags[3 * nn + c] = right;
  flags[4 * nn + c] = mh;
  flags[5 * nn + c] = mv;
  flags[6 * nn + c] = hull;
}

// icon (9 x nel, element fastest) and the element flags: 6 rows of nel (top,
// bot, left, right, middleH, middleV)
__global__ void k_velocity_icon(int geom, int nelx, int nelz, int nnx,
                                const uint8_t* __restrict__ middleH,
                                const uint8_t* __restrict__ middleV, int* __restrict__ icon,
                                uint8_t* __restrict__ eflags) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int nel = nelx * nelz;
  if (e >= nel) return;
  const int i = e % nelx, j = e / nelx;
  int n[9];
  bool top = j == nelz - 1, bot = j == 0, left = false, right = false, mh = false, mv = false;
  if (geom == MW_GEOM_ANNULUS) {
    const int nelt = nelx;
    n[0] = 2 * e + 2 + 2 * j * nelt;
    n[1] = 2 * e + 2 * j * nelt;
    n[2] = n[1] + 4 * nelt;
    n[3] = n[1] + 4 * nelt + 2;
    n[4] = n[0] - 1;
    n[5] = n[1] + 2 * nelt;
    n[6] = n[2] + 1;
    n[7] = n[5] + 2;
    n[8] = n[5] + 1;
    if (i == nelt - 1) {  // the last element of a ring closes it
      n[0] -= 2 * nelt;
      n[7] -= 2 * nelt;
      n[3] -= 2 * nelt;
    }
  } else {
    const int b = i * 2 + j * 2 * nnx;
    n[0] = b;
    n[1] = b + 2;
    n[2] = b + 2 + nnx * 2;
    n[3] = b + nnx * 2;
    n[4] = b + 1;
    n[5] = b + 2 + nnx;
    n[6] = b + 1 + nnx * 2;
    n[7] = b + nnx;
    n[8] = b + 1 + nnx;
    left = i == 0;
    right = i == nelx - 1;
    mh = middleH[n[0]] || middleH[n[2]];
    mv = middleV[n[0]] || middleV[n[1]];
  }
#pragma unroll
  for (int k = 0; k < 9; k++) icon[k * (long)nel + e] = n[k];
  eflags[0 * nel + e] = top;
  eflags[1 * nel + e] = bot;
  eflags[2 * nel + e] = left;
  eflags[3 * nel + e] = right;
  eflags[4 * nel + e] = mh;
  eflags[5 * nel + e] = mv;
}

// flags: 5 rows of nn (left, right, top, bot, hull)
__global__ void k_pressure_mesh(int geom, int nelx, int nelz, int nnx, int nnz, double hx,
                                double hz, double Rinner, double Router, double opening_angle,
                                double* __restrict__ x_P, double* __restrict__ z_P,
                                double* __restrict__ rad_P, double* __restrict__ theta_P,
                                uint8_t* __restrict__ flags) {
  const long c = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long nn = (long)nnx * nnz;
  if (c >= nn) return;
  const int i = (int)(c % nnx), j = (int)(c / nnx);
  double x = i * hx, z = j * hz;
  const bool top = j == nelz, bot = j == 0;
  bool left = false, right = false, hull = false;
  if (geom != MW_GEOM_ANNULUS) {
    left = i == 0;
    right = i == nelx;
    hull = top || bot || left || right;  // the annulus leaves its hull unset
  }
  double rad = 0.0, theta = 0.0;
  curved_map(geom, Rinner, Router, opening_angle, x, z, rad, theta);
  x_P[c] = x;
  z_P[c] = z;
  rad_P[c] = rad;
  theta_P[c] = theta;
  flags[0 * nn + c] = left;
  flags[1 * nn + c] = right;
  flags[2 * nn + c] = top;
  flags[3 * nn + c] = bot;
  flags[4 * nn + c] = hull;
}

__global__ void k_pressure_icon(int geom, int nelx, int nelz, int* __restrict__ icon) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int nel = nelx * nelz;
  if (e >= nel) return;
  const int i = e % nelx, j = e / nelx;
  int n0, n1, n2, n3;
  if (geom == MW_GEOM_ANNULUS) {
    const int nelt = nelx;
    int a1 = e, a2 = e + 1, a3 = i + (j + 1) * nelt + 1, a4 = i + (j + 1) * nelt;
    if (i == nelt - 1) {
      a2 -= nelt;
      a3 -= nelt;
    }
    n0 = a2;
    n1 = a1;
    n2 = a4;
    n3 = a3;
  } else {
    n0 = i + j * (nelx + 1);
    n1 = i + 1 + j * (nelx + 1);
    n2 = i + 1 + (j + 1) * (nelx + 1);
    n3 = i + (j + 1) * (nelx + 1);
  }
  icon[0 * nel + e] = n0;
  icon[1 * nel + e] = n1;
  icon[2 * nel + e] = n2;
  icon[3 * nel + e] = n3;
}

static bool geom_ok(int g) { return g >= MW_GEOM_BOX && g <= MW_GEOM_ANNULUS; }

}  // namespace mw

using namespace mw;

extern "C" {

int mw_build_velocity_mesh(int geometry, int nelx, int nelz, double Lx, double Lz, double Rinner,
                           double Router, double opening_angle, double* x_V, double* z_V,
                           double* rad_V, double* theta_V, uint8_t* flags, void* stream) {
  if (!geom_ok(geometry)) return fail(MW_EINVAL, "build_velocity_mesh: unknown geometry");
  if (nelx < 1 || nelz < 1 || !x_V || !z_V || !rad_V || !theta_V || !flags)
    return fail(MW_EINVAL, "build_velocity_mesh: bad arguments");
  const int nnx = geometry == MW_GEOM_ANNULUS ? 2 * nelx : 2 * nelx + 1, nnz = 2 * nelz + 1;
  const long nn = (long)nnx * nnz;
  const double hx = Lx / nelx, hz = Lz / nelz;
  k_velocity_mesh<<<nblocks(nn, 256), 256, 0, ST(stream)>>>(geometry, nelx, nelz, nnx, nnz, Lx,
                                                            Lz, hx, hz, Rinner, Router,
                                                            opening_angle, x_V, z_V, rad_V,
                                                            theta_V, flags);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

int mw_build_velocity_connectivity(int geometry, int nelx, int nelz, const uint8_t* middleH_nodes,
                                   const uint8_t* middleV_nodes, int32_t* icon_V,
                                   uint8_t* elem_flags, void* stream) {
  if (!geom_ok(geometry)) return fail(MW_EINVAL, "build_velocity_connectivity: unknown geometry");
  if (nelx < 1 || nelz < 1 || !icon_V || !elem_flags ||
      (geometry != MW_GEOM_ANNULUS && (!middleH_nodes || !middleV_nodes)))
    return fail(MW_EINVAL, "build_velocity_connectivity: bad arguments");
  const int nnx = geometry == MW_GEOM_ANNULUS ? 2 * nelx : 2 * nelx + 1;
  const long nel = (long)nelx * nelz;
  k_velocity_icon<<<nblocks(nel, 256), 256, 0, ST(stream)>>>(geometry, nelx, nelz, nnx,
                                                             middleH_nodes, middleV_nodes, icon_V,
                                                             elem_flags);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

int mw_build_pressure_mesh(int geometry, int nelx, int nelz, double hx, double hz, double Rinner,
                           double Router, double opening_angle, double* x_P, double* z_P,
                           double* rad_P, double* theta_P, uint8_t* flags, void* stream) {
  if (!geom_ok(geometry)) return fail(MW_EINVAL, "build_pressure_mesh: unknown geometry");
  if (nelx < 1 || nelz < 1 || !x_P || !z_P || !rad_P || !theta_P || !flags)
    return fail(MW_EINVAL, "build_pressure_mesh: bad arguments");
  const int nnx = geometry == MW_GEOM_ANNULUS ? nelx : nelx + 1, nnz = nelz + 1;
  const long nn = (long)nnx * nnz;
  k_pressure_mesh<<<nblocks(nn, 256), 256, 0, ST(stream)>>>(geometry, nelx, nelz, nnx, nnz, hx,
                                                            hz, Rinner, Router, opening_angle,
                                                            x_P, z_P, rad_P, theta_P, flags);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

int mw_build_pressure_connectivity(int geometry, int nelx, int nelz, int32_t* icon_P,
                                   void* stream) {
  if (!geom_ok(geometry)) return fail(MW_EINVAL, "build_pressure_connectivity: unknown geometry");
  if (nelx < 1 || nelz < 1 || !icon_P)
    return fail(MW_EINVAL, "build_pressure_connectivity: bad arguments");
  const long nel = (long)nelx * nelz;
  k_pressure_icon<<<nblocks(nel, 256), 256, 0, ST(stream)>>>(geometry, nelx, nelz, icon_P);
  MW_CHECK_LAUNCH();
  return MW_OK;
}

}  // extern "C"
