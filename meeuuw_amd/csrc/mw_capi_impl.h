// Internal (C++) declarations shared by the kernel translation units and the
// C entry points in mw_capi.cpp.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mw {

struct StokesArgs {
  int nel, Nfem_V, Nfem_P, blocks;
  const int *icon_P, *l2g;
  const double *rhoq, *etaq, *JxWq, *gxq, *gyq;
  const double *N_V, *N_P, *dNdr_V, *dNdt_V, *dNdr_P, *dNdt_P;
  const double *j00, *j01, *j10, *j11, *eta_e;
  double eta_ref, L_ref;
  const uint8_t* bc_fix;
  const double* bc_val;
  const int *inc_v_ptr, *inc_v_list, *inc_p_ptr, *inc_p_list;
  double *VV, *rhs_f, *rhs_h, *VV_K, *VV_G, *VV_GT, *VV_M, *VV_Me, *VV_H;
  double* work;  // nel*22 doubles
  int rhs_only;  // 1: right-hand sides only (VV not formed, may be null; blocks must be 0)
  // free slip on curved boundaries (quarter, half, eighth, annulus); null theta: none
  const double* theta;  // theta_V (nn_V)
  const uint8_t *bot_el, *top_el;  // bot_element, top_element (nel)
  int bot_fs, top_fs;
  int axi;           // axisymmetric branch (xq: radius at the quadrature points, nel*9)
  const double* xq;
};

struct EnergyArgs {
  int nel, Nfem_T, eba;
  const double* T;
  const int* icon_T;
  const double *etaq, *Tq, *uq, *wq, *hcondq, *hcapaq, *alphaq, *hprodq, *exxq, *ezzq, *exzq,
      *JxWq;
  const double *N_T, *dNdr_T, *dNds_T, *j00, *j01, *j10, *j11;
  const uint8_t* bc_fix;
  const double* bc_val;
  double dt, rho0;
  const int *inc_ptr, *inc_list;
  double *VV, *rhs;
  double* work;  // nel*9 doubles
};

// assembly
int local_to_global(int nel, const int* icon_V, int* l2g, hipStream_t st);
int coo_stokes(int nel, const int* l2g, const int* icon_P, int Nfem_V, int* II, int* JJ,
               hipStream_t st);
int coo_outer(int nel, int nr, int nc, const int* rmap, const int* cmap, int* II, int* JJ,
              hipStream_t st);
size_t incidence_temp_bytes(int nn);
int incidence_build(int nel, int m, const int* icon, int nn, int* ptr, int* list, void* work,
                    hipStream_t st);
int gather_sum(int nn, int m, const int* ptr, const int* list, const double* vel, double* out,
               hipStream_t st);
int build_matrix_stokes(const StokesArgs& a, hipStream_t st);
// rhs_f and rhs_h of build_matrix_stokes when every bc_val is zero (work: 18*nel)
int stokes_rhs_zero_lift(const StokesArgs& a, hipStream_t st);
int build_matrix_energy(const EnergyArgs& a, hipStream_t st);

// particles
int locate_box(int np, const uint8_t* active, const double* x, const double* z, double hx,
               double hz, const double* x_V, const double* z_V, const int* icon_V, int nel,
               int nelx, double* r, double* t, int* iel, int* flags, hipStream_t st);
int interpolate_fields(int np, int nf, const uint8_t* active, const double* r, const double* t,
                       const int* iel, const double* const* phi, const int* icon, int nel,
                       double* const* out, hipStream_t st);
int advect_box(int rk, double dt, int np, double* x, double* z, uint8_t* active, const double* u,
               const double* w, double Lx, double Lz, double hx, double hz, int nelx, int nel,
               const int* icon, const double* x_V, const double* z_V, double* su, double* sw,
               int* flags, hipStream_t st, const double* uw = nullptr);
// (uw given: the mesh is the box mesh's own row-major node grid, icon, u and w
// are not read and the velocity comes from uw = (u, w) per node)
// Stable sort of particles by element (counting sort; inactive particles, and
// particles whose element index is outside [0, nel), last in unspecified order).
// order: np, seg: nel + 1.  work: sort_work_bytes(np, nel) bytes.  order == NULL
// in the projections below means the identity (fields already in sorted order).
size_t sort_work_bytes(int np, int nel);
// cnt_zero: the workspace's counters are known to be zero (the previous sort
// with this workspace left them so)
int sort_by_element(int np, int nel, const uint8_t* active, const int* iel, int* order, int* seg,
                    void* work, hipStream_t st, bool cnt_zero = false);
// the same with explicit scratch: cnt (nel + 1 ints, all zero on entry when
// cnt_zero, and all zero again on return), tsum (scan_tile_sums_ints(nel) ints)
// and tmp (np ints)
int sort_by_element_scratch(int np, int nel, const uint8_t* active, const int* iel, int* order,
                            int* seg, int* cnt, int* tsum, int* tmp, bool cnt_zero,
                            hipStream_t st);
// out[i] = in[0] + ... + in[i-1] for i in [0, n] (n + 1 outputs); tsum:
// scan_tile_sums_ints(n) ints of scratch
int exclusive_scan_i32(int n, const int* in, int* out, int* tsum, hipStream_t st);
// compute_ls_coefficients_P1/_Q1 over the segments of an element sort, with
// particle properties in sort order; limiter over strided coefficient rows
int ls_coefficients_sorted(int nb, int nel, const double* x_e, const double* z_e, const int* seg,
                           const int* order, const double* x, const double* z,
                           const double* rho_sorted, const double* eta_sorted, double* cr,
                           double* ce, double* mm, hipStream_t st);
int limiter_strided(int n, const double* c0, double* c1, double* c2, const double* S,
                    const double* L, double h, int cs, int ms, hipStream_t st);
size_t scan_tile_sums_ints(int n);
int project_on_elements(int nel, const int* seg, const int* order, const double* rho,
                        const double* eta, int averaging, double* rho_e, double* eta_e, int* cnt,
                        hipStream_t st);
int project_on_nodes(int type, int nel, int nn, int nf, const int* seg, const int* order,
                     const double* const* phi, const int* averaging, const int* icon,
                     const double* r, const double* t, const int* inc_ptr, const int* inc_list,
                     double* work, double* const* out, hipStream_t st);

// nodal fields, reductions
int nodal_to_qpoints(int nf, const double* const* phi, int nel, int m, const double* N,
                     const int* icon, double* const* phiq, hipStream_t st,
                     int grid_nelx = 0);
int pressure_offset(int mode, const uint8_t* top_element, int nel, int nelx, const double* N_P,
                    const double* JxWq, const double* p, const int* icon_P, double volume,
                    double* out, double* work, hipStream_t st);
int pressure_offset_axi(int mode, const uint8_t* top_element, int nel, const double* N_P,
                        const double* JxWq, const double* p, const int* icon_P,
                        const double* theta_P, const double* xq, double* out, double* work,
                        hipStream_t st);
int nodal_pressure(int nel, int nn, const int* icon_V, const int* icon_P, const double* p,
                   const double* N_P_n, const int* inc_ptr, const int* inc_list, double* work,
                   double* q, hipStream_t st);
int nodal_heat_flux(const int* icon_T, const double* T, const double* hcond_n, int nn, int nel,
                    const double* dNdr_n, const double* dNds_n, const double* j00,
                    const double* j01, const double* j10, const double* j11, const int* inc_ptr,
                    const int* inc_list, double* work, double* dTdx, double* dTdy, double* qx,
                    double* qy, hipStream_t st);
int global_quantities(int nel, const double* uq, const double* wq, const double* Tq,
                      const double* rhoq, const double* hcapaq, const double* etaq,
                      const double* exxq, const double* ezzq, const double* exzq, double volume,
                      const double* JxWq, const double* gxq, const double* gzq, double* out,
                      double* work, hipStream_t st);
int compute_Nu(double Lx, double Lz, const double* x_V, const double* z_V, int nel,
               const uint8_t* left, const uint8_t* right, const uint8_t* top,
               const uint8_t* bottom, const int* icon, const double* T, const double* dTdx,
               const double* dTdz, const double* qc, const double* qw, const double* hcond_n,
               double* out, double* work, hipStream_t st);

}  // namespace mw
