/*
 * meeuuw_hip -- C ABI of the MI355X (gfx950) implementation of the meeuuw
 * per-timestep hot path.
 *
 * Conventions
 *   - Every array argument is a DEVICE pointer (HIP global memory) unless the
 *     parameter name ends in _host.  Pointer-array parameters (const double*
 *     const* phi, ...) are HOST arrays whose elements are device pointers.
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).  All
 *     work is enqueued on that stream; functions return without synchronising
 *     unless stated.  Functions that report a host-visible count (flags) say so.
 *   - Layouts are the reference's: icon (m, nel) row-major, qpoint fields
 *     (nel, 9), reference tables (9, m), bool arrays as uint8.
 *   - Every function returns 0 on success or a status code (below);
 *     mw_last_error() gives the message of the last failure on this thread.
 *
 * Each entry point names the reference function (file:line) whose behaviour
 * it provides; the Python package meeuuw_amd wraps them behind the
 * reference's own signatures.
 */
#ifndef MEEUUW_HIP_H
#define MEEUUW_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MW_OK = 0,
  MW_EINVAL = 1,
  MW_EHIP = 2,
  MW_EOUTSIDE = 3,
  MW_EEMPTY = 4,
  MW_ENOCONV = 5,
  MW_ERANGE = 6,
  MW_EUNSUPPORTED = 7
};

int mw_version(void);
const char* mw_last_error(void);

/* ---------------------------------------------------------------- setup --- */

/* local_to_globalV, meeuuw.py:874-881: l2g (18, nel) */
int mw_local_to_global_V(int nel, const int32_t* icon_V, int32_t* l2g, void* stream);

/* II_Stokes/JJ_Stokes, meeuuw.py:890-912: nel*468 entries */
int mw_coo_stokes(int nel, const int32_t* l2g, const int32_t* icon_P, int Nfem_V, int32_t* II,
                  int32_t* JJ, void* stream);

/* II_K/JJ_K, meeuuw.py:920-935: nel*324 entries */
int mw_coo_K(int nel, const int32_t* l2g, int32_t* II, int32_t* JJ, void* stream);

/* II_G/JJ_G and II_GT/JJ_GT, meeuuw.py:943-970: nel*72 entries each */
int mw_coo_G(int nel, const int32_t* l2g, const int32_t* icon_P, int32_t* II_G, int32_t* JJ_G,
             int32_t* II_GT, int32_t* JJ_GT, void* stream);

/* II_MP/JJ_MP (m=4, icon_P) and II_T/JJ_T (m=9, icon_T), meeuuw.py:978-1015 */
int mw_coo_scalar(int nel, int m, const int32_t* icon, int32_t* II, int32_t* JJ, void* stream);

/* Node -> (element, local index) incidence of a connectivity (m, nel), used by
 * every dof-parallel gather.  ptr: nn+1, list: m*nel entries (iel<<5 | k),
 * sorted by element.  work: mw_incidence_workspace_bytes(nn) bytes.
 * Synchronises `stream` (validates the connectivity).  No reference
 * counterpart: it replaces the scatter order of the element loops. */
size_t mw_incidence_workspace_bytes(int nn);
int mw_incidence_build(int nel, int m, const int32_t* icon, int nn, int32_t* ptr, int32_t* list,
                       void* work, void* stream);

/* ------------------------------------------------------------- assembly --- */

/* build_matrix_stokes, build_matrix_stokes.py:18-278, plane strain without the
 * curved free-slip rotation (the general form is mw_build_matrix_stokes_general).
 * blocks=0 fills VV_Stokes (nel*468); blocks=1 fills VV_K
 * (nel*324), VV_G, VV_GT, VV_H (nel*72 each), VV_M, VV_M_eta (nel*16).
 * inc_v_*: incidence of l2g (m=18, nn=Nfem_V); inc_p_*: of icon_P (m=4,
 * nn=Nfem_P).  work: nel*22 doubles. */
int mw_build_matrix_stokes(int nel, int Nfem_V, int Nfem_P, int blocks, const int32_t* icon_P,
                           const int32_t* l2g, const double* rhoq, const double* etaq,
                           const double* JxWq, const double* gxq, const double* gyq,
                           const double* N_V, const double* N_P, const double* dNdr_V,
                           const double* dNdt_V, const double* dNdr_P, const double* dNdt_P,
                           const double* jcbi00q, const double* jcbi01q, const double* jcbi10q,
                           const double* jcbi11q, const double* eta_e, double eta_ref,
                           double L_ref, const uint8_t* bc_fix_V, const double* bc_val_V,
                           const int32_t* inc_v_ptr, const int32_t* inc_v_list,
                           const int32_t* inc_p_ptr, const int32_t* inc_p_list,
                           double* VV_Stokes, double* rhs_f, double* rhs_h, double* VV_K,
                           double* VV_G, double* VV_GT, double* VV_M, double* VV_M_eta,
                           double* VV_H, double* work, void* stream);

/* build_matrix_stokes with every branch of build_matrix_stokes.py:18-278: the
 * arguments of mw_build_matrix_stokes plus, in the reference's order,
 * bot_element, top_element, bot_free_slip, top_free_slip, curved (geometry is
 * quarter, half, eighth or annulus), theta_V, axisymmetric and xq (nel*9).
 * Curved free slip (163-203): at each node of a bot_element (local nodes 0, 1,
 * 4) or top_element (2, 3, 6) whose two dofs are free, the element system is
 * rotated to (normal, tangential) components with theta_V of the node and the
 * normal dof is eliminated, before the Dirichlet rows; the node of local dof
 * 2i is l2g[2i]/2 (local_to_globalV, meeuuw.py:875-880).  The rotated entries
 * agree with the reference's dense products to rounding (sums of two
 * products).  Axisymmetric (108-125): B rows (du/dx, u/x, dw/dz, shear),
 * C = diag(2,2,2,1), weight 2 pi x; VV_M, VV_M_eta and VV_H are zero, as the
 * reference forms no such blocks there.  bot/top_element and theta_V may be
 * null without free slip, xq without axisymmetric. */
int mw_build_matrix_stokes_general(
    int nel, int Nfem_V, int Nfem_P, int blocks, const int32_t* icon_P, const int32_t* l2g,
    const double* rhoq, const double* etaq, const double* JxWq, const double* gxq,
    const double* gyq, const double* N_V, const double* N_P, const double* dNdr_V,
    const double* dNdt_V, const double* dNdr_P, const double* dNdt_P, const double* jcbi00q,
    const double* jcbi01q, const double* jcbi10q, const double* jcbi11q, const double* eta_e,
    double eta_ref, double L_ref, const uint8_t* bc_fix_V, const double* bc_val_V,
    const uint8_t* bot_element, const uint8_t* top_element, int bot_free_slip,
    int top_free_slip, int curved, const double* theta_V, int axisymmetric, const double* xq,
    const int32_t* inc_v_ptr,
    const int32_t* inc_v_list, const int32_t* inc_p_ptr, const int32_t* inc_p_list,
    double* VV_Stokes, double* rhs_f, double* rhs_h, double* VV_K, double* VV_G, double* VV_GT,
    double* VV_M, double* VV_M_eta, double* VV_H, double* work, void* stream);

/* build_matrix_energy, build_matrix_energy.py:11-141.  eba=0 is 'BA'.
 * inc_*: incidence of icon_T (m=9, nn=Nfem_T).  work: nel*9 doubles. */
int mw_build_matrix_energy(int nel, int Nfem_T, int eba, const double* T, const int32_t* icon_T,
                           const double* etaq, const double* Tq, const double* uq,
                           const double* wq, const double* hcondq, const double* hcapaq,
                           const double* alphaq, const double* hprodq, const double* exxq,
                           const double* ezzq, const double* exzq, const double* JxWq,
                           const double* N_T, const double* dNdr_T, const double* dNds_T,
                           const double* jcbi00q, const double* jcbi01q, const double* jcbi10q,
                           const double* jcbi11q, const uint8_t* bc_fix_T,
                           const double* bc_val_T, double dt, double rho0,
                           const int32_t* inc_ptr, const int32_t* inc_list, double* VV_T,
                           double* rhs, double* work, void* stream);

/* ------------------------------------------------------------ particles --- */

/* locate_particles___box, pic_functions.py:107-137.  flags (3 int32, device):
 * [0] active particles whose element index leaves [0, nel) (the reference
 * reads out of bounds there), [1] particles with r outside [-1, 1], [2] with t
 * outside [-1, 1] (locate_particles.py:61-64).  flags are zeroed by the call. */
int mw_locate_particles_box(int np, const uint8_t* active, const double* x, const double* z,
                            double hx, double hz, const double* x_V, const double* z_V,
                            const int32_t* icon_V, int nel, int nelx, double* r, double* t,
                            int32_t* iel, int32_t* flags, void* stream);

/* interpolate_field_on_particles, pic_functions.py:84-101, for nf <= 8 fields
 * in one pass.  phi_host/out_host: host arrays of nf device pointers. */
int mw_interpolate_fields(int np, int nf, const uint8_t* active, const double* r,
                          const double* t, const int32_t* iel, const double* const* phi_host,
                          const int32_t* icon, int nel, double* const* out_host, void* stream);

/* locate_particles___annulus, pic_functions.py:142-173 (the locate of the
 * quarter, half and eighth geometries, locate_particles.py:38-51): element from
 * the polar coordinates (rad, theta); flags as mw_locate_particles_box. */
int mw_locate_particles_annulus(int np, const uint8_t* active, const double* rad,
                                const double* theta, double hrad, double htheta,
                                const double* rad_V, const double* theta_V,
                                const int32_t* icon_V, int nel, int nelx, double Rinner,
                                double* r, double* t, int32_t* iel, int32_t* flags,
                                void* stream);

/* advect_particles___quarter / ___half / ___eighth / ___annulus,
 * pic_functions.py:298-596 (variant 0, 1, 2, 3).  RK1 for every variant and
 * RK2 for quarter and half, as the reference (other orders leave the particles
 * in place); x, z, rad, theta, active updated in place; flags[0] counts
 * velocity evaluations whose element index left the mesh. */
int mw_advect_particles_curved(int variant, int rk, double dt, int np, double* x, double* z,
                               double* rad, double* theta, uint8_t* active, const double* u,
                               const double* w, double Rinner, double Router, double hrad,
                               double htheta, int nelx, int nel, const int32_t* icon_V,
                               const double* rad_V, const double* theta_V, double* su,
                               double* sw, int32_t* flags, void* stream);

/* interpolate_vel_on_pt___box (polar = 0, pic_functions.py:17-37) and
 * interpolate_vel_on_pt___annulus (polar = 1, pic_functions.py:40-63) at n
 * points; the arguments of the other geometry are ignored. */
int mw_interpolate_vel_on_pts(int polar, int n, const double* xp, const double* zp,
                              const double* u, const double* w, double hx, double hz,
                              const double* x_V, const double* z_V, double hrad, double htheta,
                              const double* rad_V, const double* theta_V, double Rinner,
                              int nelx, int nel, const int32_t* icon_V, double* up, double* wp,
                              int32_t* iel, int32_t* flags, void* stream);

/* compute_ls_coefficients_P1 (nb = 3, pic_functions.py:1097-1178) and _Q1
 * (nb = 4, pic_functions.py:1188-1285): per-element least-squares fits of rho
 * and eta (coef_* nel x nb, row-major) and their extrema (minmax nel x 4: rho
 * min, rho max, eta min, eta max).  Every particle with iel == e counts, as
 * the reference's mask.  work: mw_ls_workspace_bytes device bytes. */
int mw_ls_coefficients(int nb, int nel, const double* x_e, const double* z_e, int np,
                       const double* x, const double* z, const int32_t* iel, const double* rho,
                       const double* eta, double* coef_rho, double* coef_eta, double* minmax,
                       void* work, void* stream);
size_t mw_ls_workspace_bytes(int np, int nel);

/* limiter, pic_functions.py:1067-1087, applied element-wise as the driver's
 * loop does (meeuuw.py:1243-1249); c1, c2 updated in place. */
int mw_limiter(int n, const double* c0, double* c1, double* c2, const double* S,
               const double* L, double h, void* stream);

/* advect_particles___box, pic_functions.py:179-292 (RK 1, 2, 4).  x, z,
 * active are updated in place; su, sw receive the particle velocities.
 * flags[0] (zeroed by the call) counts particles whose stage position left
 * the mesh's element range. */
int mw_advect_particles_box(int rk, double dt, int np, double* x, double* z, uint8_t* active,
                            const double* u, const double* w, double Lx, double Lz, double hx,
                            double hz, int nelx, int nel, const int32_t* icon_V,
                            const double* x_V, const double* z_V, double* su, double* sw,
                            int32_t* flags, void* stream);

/* Stable sort of particle indices by element.  order: np, seg: nel+1
 * (seg[e]..seg[e+1] are element e's particles in increasing index); inactive
 * particles, and active ones whose element index is outside [0, nel), follow
 * from seg[nel] on, in unspecified order.
 * work: mw_sort_workspace_bytes(np, nel) bytes.  Feeds the projections. */
size_t mw_sort_workspace_bytes(int np, int nel);
int mw_sort_particles_by_element(int np, int nel, const uint8_t* active, const int32_t* iel,
                                 int32_t* order, int32_t* seg, void* work, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate (code
 * under HIP graph capture, tight stepping loops, harnesses that forbid
 * hipMalloc): mw_sort_particles_by_element with its workspace in a fixed
 * device array of the library (about 20 MB per GPU, held until the process
 * ends) instead of `work`.  np <= 4194304 and nel <= 1048576, else MW_ERANGE.
 * Results are those of mw_sort_particles_by_element.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU race. */
int mw_sort_particles_by_element_noalloc(int np, int nel, const uint8_t* active,
                                         const int32_t* iel, int32_t* order, int32_t* seg,
                                         void* stream);

/* project_particles_on_elements, pic_functions.py:602-650.
 * averaging: 0 arithmetic, 1 geometric, 2 harmonic. */
int mw_project_particles_on_elements(int nel, const int32_t* seg, const int32_t* order,
                                     const double* rho, const double* eta, int averaging,
                                     double* rho_e, double* eta_e, int32_t* count, void* stream);

/* project_particle_field_on_nodes_{1,2,3,4}, pic_functions.py:658-1053, for
 * nf <= 8 fields at once (averaging_host: nf codes).  inc_*: incidence of
 * icon_V (m=9, nn).  work: nel*nf*8 doubles. */
int mw_project_particle_fields_on_nodes(int type, int nel, int nn, int nf, const int32_t* seg,
                                        const int32_t* order, const double* const* phi_host,
                                        const int32_t* averaging_host, const int32_t* icon_V,
                                        const double* r, const double* t,
                                        const int32_t* inc_ptr, const int32_t* inc_list,
                                        double* work, double* const* out_host, void* stream);

/* --------------------------------------------------- fields, reductions --- */

/* project_nodal_{P,V,T}field_onto_qpoints, project_nodal_field_onto_qpoints.py:11-62:
 * m=4 (first 4 rows of icon) or m=9; N: (9, m). */
int mw_nodal_to_qpoints(int nf, const double* const* phi_host, int nel, int m, const double* N,
                        const int32_t* icon, double* const* phiq_host, void* stream);

/* compute_pressure_offset, compute_pressure_offset.py:11-66 (plane):
 * mode 0 none, 1 surface, 2 volume.  out: 1 double (device). work: 1024 doubles. */
int mw_pressure_offset(int mode, const uint8_t* top_element, int nel, int nelx,
                       const double* N_P, const double* JxWq, const double* p,
                       const int32_t* icon_P, double volume, double* out, double* work,
                       void* stream);

/* compute_pressure_offset with axisymmetric=True, compute_pressure_offset.py:33-39,
 * 49-55: mode 1 sums 0.5 * pmean * sin(pi/2 - mean theta_P) * dtheta over the
 * top edges (pressure nodes 2, 3 of each top_element), mode 2 the integral of p
 * weighted by 2 pi xq (nel*9); as in the reference neither is divided by a
 * length or volume.  out, work as mw_pressure_offset. */
int mw_pressure_offset_axisymmetric(int mode, const uint8_t* top_element, int nel,
                                    const double* N_P, const double* JxWq, const double* p,
                                    const int32_t* icon_P, const double* theta_P,
                                    const double* xq, double* out, double* work, void* stream);

/* compute_nodal_pressure, compute_nodal_pressure.py:7-21.  inc_*: incidence of
 * icon_V.  work: nel*9 doubles. */
int mw_nodal_pressure(int nel, int nn, const int32_t* icon_P, const double* p,
                      const double* N_P_n, const int32_t* inc_ptr, const int32_t* inc_list,
                      double* q, double* work, void* stream);

/* compute_nodal_heat_flux, compute_nodal_heat_flux.py:10-40.  jcbi**: (nel, 9)
 * at the T nodes.  work: nel*18 doubles. */
int mw_nodal_heat_flux(int nel, int nn, const int32_t* icon_T, const double* T,
                       const double* hcond_n, const double* dNdr_n, const double* dNds_n,
                       const double* jcbi00, const double* jcbi01, const double* jcbi10,
                       const double* jcbi11, const int32_t* inc_ptr, const int32_t* inc_list,
                       double* dTdx, double* dTdy, double* qx, double* qy, double* work,
                       void* stream);

/* compute_nodal_strain_rate, compute_strain_rate.py:50-73 (mode 0: u, v ->
 * o0 exx, o1 eyy, o2 exy) and compute_nodal_pressure_gradient,
 * compute_nodal_pressure_gradient.py:11-31 (mode 1: u = q -> o0 dpdx, o1 dpdy;
 * v and o2 unused).  jcbi**: (nel, 9) at the V nodes; inc_*: node incidence
 * (mw_incidence_build); work: nel*27 doubles. */
int mw_nodal_gradients(int mode, int nel, int nn, const int32_t* icon_V, const double* u,
                       const double* v, const double* dNdr_n, const double* dNds_n,
                       const double* jcbi00, const double* jcbi01, const double* jcbi10,
                       const double* jcbi11, const int32_t* inc_ptr, const int32_t* inc_list,
                       double* o0, double* o1, double* o2, double* work, void* stream);

/* compute_elemental_strain_rate, compute_strain_rate.py:17-42 (element centre;
 * the 2x2 Jacobian inverted in closed form). */
int mw_elemental_strain_rate(int nel, const int32_t* icon_V, const double* u, const double* v,
                             const double* x_V, const double* y_V, double* exx, double* eyy,
                             double* exy, void* stream);

/* compute_global_quantities, postprocessors.py:13-83.  out (device, 9):
 * vrms, EK, WAG, TVD, GPE, ITE, TM, Tavrg, eta_avrg.  work: 1024*9+9 doubles. */
int mw_global_quantities(int nel, const double* uq, const double* wq, const double* Tq,
                         const double* rhoq, const double* hcapaq, const double* etaq,
                         const double* exxq, const double* ezzq, const double* exzq,
                         double volume, const double* JxWq, const double* gxq,
                         const double* gzq, double* out, double* work, void* stream);

/* compute_Nu, postprocessors.py:152-248 (box).  out (device, 13): T_left,
 * T_right, T_bottom, T_top, dTdx_left, dTdx_right, dTdz_bottom, dTdz_top,
 * qx_left, qx_right, qz_bottom, qz_top, Nu.  qc, qw: 3 Gauss points/weights.
 * work: 1024*12+12 doubles. */
int mw_compute_Nu(double Lx, double Lz, const double* x_V, const double* z_V, int nel,
                  const uint8_t* left, const uint8_t* right, const uint8_t* top,
                  const uint8_t* bottom, const int32_t* icon_V, const double* T,
                  const double* dTdx, const double* dTdz, const double* qc, const double* qw,
                  const double* hcond_n, double* out, double* work, void* stream);

/* ------------------------------------------------------- block solvers --- */

/* uzawa3_solver_L2, uzawa3_solver_L2.py:21-99, on CSR matrices in device
 * memory: K (nV x nV), G (nV x nP), GT, MP, H (nP x ...).  SuperLU solves with
 * K and MP are replaced by Jacobi-PCG to relative residual inner_rtol.  The
 * outer loop (at most 100 iterations, stop when both increments < 1e-6) and
 * the per-iteration arrays (100 doubles each, host) are the reference's.
 * Synchronises the stream. */
int mw_uzawa3_solver_L2(int nV, int nP, const int32_t* K_ptr, const int32_t* K_col,
                        const double* K_val, const int32_t* G_ptr, const int32_t* G_col,
                        const double* G_val, const int32_t* GT_ptr, const int32_t* GT_col,
                        const double* GT_val, const int32_t* MP_ptr, const int32_t* MP_col,
                        const double* MP_val, const int32_t* H_ptr, const int32_t* H_col,
                        const double* H_val, const double* f, const double* h, double* solV,
                        double* solP, double inner_rtol, int32_t* niter_host, double* xiV_host,
                        double* xiP_host, double* alpha_host, void* stream);

/* cd_u_eta, cd_u_eta.py:13-99 (M: pressure mass matrix, M_eta: M scaled by
 * 1/eta_e), same conventions; beta_host receives the 100 beta values. */
int mw_cd_u_eta(int nV, int nP, const int32_t* K_ptr, const int32_t* K_col, const double* K_val,
                const int32_t* G_ptr, const int32_t* G_col, const double* G_val,
                const int32_t* GT_ptr, const int32_t* GT_col, const double* GT_val,
                const int32_t* M_ptr, const int32_t* M_col, const double* M_val,
                const int32_t* Me_ptr, const int32_t* Me_col, const double* Me_val,
                const int32_t* H_ptr, const int32_t* H_col, const double* H_val, const double* f,
                const double* h, double* solV, double* solP, double inner_rtol,
                int32_t* niter_host, double* xiV_host, double* xiP_host, double* alpha_host,
                double* beta_host, void* stream);

/* ------------------------------------------------------- mesh and icon --- */
/* Geometry codes of the mesh builders (the reference's geometry strings). */
enum { MW_GEOM_BOX = 0, MW_GEOM_EIGHTH, MW_GEOM_QUARTER, MW_GEOM_HALF, MW_GEOM_ANNULUS };

/* build_velocity_mesh (build_velocity_mesh.py:17-227) for the Q2 space
 * (m_V = 9): nn_V = nnx * nnz nodes, nnx = 2 nelx + 1 (2 nelx for the
 * annulus), nnz = 2 nelz + 1, hx = Lx / nelx, hz = Lz / nelz; device arrays
 * of nn_V: x_V, z_V, rad_V, theta_V (0 for the box), and flags (uint8, 7 rows
 * of nn_V: top, bot, left, right, middleH, middleV, hull).  Box coordinates
 * are the reference's bit for bit; the curved ones use the device's cos, sin
 * and atan2. */
int mw_build_velocity_mesh(int geometry, int nelx, int nelz, double Lx, double Lz, double Rinner,
                           double Router, double opening_angle, double* x_V, double* z_V,
                           double* rad_V, double* theta_V, uint8_t* flags, void* stream);
/* build_velocity_connectivity (build_velocity_connectivity.py:14-136), Q2:
 * icon_V (int32, 9 x nel, element fastest) and elem_flags (uint8, 6 rows of
 * nel: top, bot, left, right, middleH, middleV) from the node mid-line flags
 * of mw_build_velocity_mesh (not read for the annulus, may be NULL there). */
int mw_build_velocity_connectivity(int geometry, int nelx, int nelz, const uint8_t* middleH_nodes,
                                   const uint8_t* middleV_nodes, int32_t* icon_V,
                                   uint8_t* elem_flags, void* stream);
/* build_pressure_mesh (build_pressure_mesh.py:10-88): nn_P = nnx * (nelz + 1),
 * nnx = nelx + 1 (nelx for the annulus); x_P, z_P, rad_P, theta_P and flags
 * (uint8, 5 rows of nn_P: left, right, top, bot, hull). */
int mw_build_pressure_mesh(int geometry, int nelx, int nelz, double hx, double hz, double Rinner,
                           double Router, double opening_angle, double* x_P, double* z_P,
                           double* rad_P, double* theta_P, uint8_t* flags, void* stream);
/* build_pressure_connectivity (build_pressure_connectivity.py:10-48), m_P = 4:
 * icon_P (int32, 4 x nel). */
int mw_build_pressure_connectivity(int geometry, int nelx, int nelz, int32_t* icon_P,
                                   void* stream);

/* ------------------------------------------------- device-resident step --- */

/* One timestep of the box-geometry thermal-convection model, entirely on the
 * device, in the order of the reference time loop (meeuuw.py:1105-2875):
 * interpolate T on particles -> material model -> elemental and nodal
 * projections -> qpoint fields -> Stokes assembly (build_matrix_stokes) ->
 * Stokes solve -> pressure normalisation -> nodal pressure -> dt -> energy
 * assembly (build_matrix_energy) and solve -> heat flux, Nu, vrms -> RK
 * advection -> locate.  The material model is the Boussinesq law of
 * experiment0.py:230-264 with eta = exp(-ln(eta_contrast) T) clipped to
 * [eta_min, eta_max] (experiment3.py:46, 72) or eta = 1 for contrast 1.
 * No reference counterpart as a single function: it replaces the driver's
 * loop body.  Host arrays end in _host. */
typedef struct mw_sim_config {
  int nelx, nelz;
  double Lx, Lz;
  int ppe;          /* particles per element (ignored when rk == -1) */
  int rk;           /* 1, 2, 4; -1: particles on the quadrature points */
  double Ra, alphaT, rho0, CFL, Tbottom, Ttop;
  double eta_contrast, eta_min, eta_max;
  int averaging;               /* 0 arithmetic, 1 geometric, 2 harmonic */
  int nodal_projection_type;   /* 1..4 */
  int pressure_normalisation;  /* 0 none, 1 surface, 2 volume */
  int solver;       /* 0: FP64 Uzawa-CG with GMG-PCG inner solves;
                       1: FP32 GMG-preconditioned FGMRES in FP64 defect correction;
                       2: as 1, with the learned Schur-complement correction (trained on
                          the first solve, kept only if it beats the plain preconditioner) */
  double stokes_tol;   /* relative residual of the full Stokes system */
  double energy_tol;   /* relative residual of the energy system */
  int fgmres_restart;  /* 2..64 */
  int cheb_degree;
  int warm_start;      /* start each solve from the previous solution */
  int pc_vcycles;      /* V-cycles per velocity-block preconditioner application (1..8) */
  /* learned correction (solver 2) */
  int learn_pairs;     /* training vectors collected from the solves (>= 16) */
  int learn_steps;     /* Adam steps */
  int learn_batch;     /* samples per step (1..16) */
  int learn_vcycles;   /* V-cycles per velocity solve inside the training Schur complement */
  double learn_lr;     /* initial Adam learning rate (cosine schedule) */
  int learn_seed;
  int fgmres_graphs;   /* 1: each FGMRES Arnoldi step (preconditioner, operator,
                          Gram-Schmidt, Givens) replays as one captured graph
                          (bit-identical; measured 1-2% slower than stream launches
                          on MI355X, so 0 is the default) */
  int gmg_tail;        /* 1: the multigrid levels below 600 unknowns run as one
                          single-workgroup LDS-resident kernel (bit-identical, but not
                          faster on MI355X: one CU cannot hide the LDS latency);
                          0: one kernel per level operation (default) */
  int cheb_degree_coarse; /* Chebyshev degree on the levels below the finest (0: cheb_degree) */
  int gmg_bottom;      /* with a coarsest level of at most 64 unknowns (gmg_coarse_dofs
                          <= 64): smoothed multigrid levels run together with the coarsest
                          solve as one workgroup where they fit, bit-identical to the
                          per-level kernels: 0 none; 1 the last level (up to 16x16 nodes);
                          2 or 3 (the Python default) the last two or three levels (up to
                          9x9, 17x17 and 33x33 nodes) in one 1024-thread kernel */
  int extrapolate;     /* with warm_start: start each Stokes solve from the last
                          solution (0), the linear extrapolation in time of the last two
                          (1, default) or the quadratic one of the last three (2); with 1
                          or 2 the matrix-free energy solve (outside Picard iterations)
                          starts from the temperature extrapolated linearly in time */
  int schur_cheb;      /* Schur block of the mixed preconditioner: 0 the lumped
                          1/eta-weighted pressure mass matrix; k >= 2: k Chebyshev-Jacobi
                          steps on the consistent one (3 in the Python defaults) */
  double schur_scale;  /* the Schur block's inverse is scaled by this (2.5 in the Python defaults) */
  int energy_mf;       /* 1: the energy system is applied matrix-free from its quadrature-point
                          coefficients (FP64, fused Jacobi-BiCGStab kernels; the Python default
                          for the mixed and learned solvers); 0: from the assembled values of
                          build_matrix_energy (the Python default for the plain solver) */
  double fgmres_inner; /* mixed/learned: each FP32 FGMRES solve is asked for this fraction of the
                          reduction still needed to reach stokes_tol (<= 0: 0.8, the Python
                          default; the FP64 defect correction checks stokes_tol itself) */
  int gmg_power_its;   /* power iterations per step for the multigrid smoothers' largest
                          eigenvalue, started from the previous step's vectors (the first
                          step: 12 from a fixed start); 0: 12 from the fixed start every step */
  int gs_passes;       /* FGMRES Gram-Schmidt: 2 classical twice, 1 classical once (one sweep
                          and one reduction fewer per iteration); 0: once from 192x192
                          elements up, twice below (where the FP32 basis loses
                          orthogonality and costs an extra defect-correction step) */
  int material_mode;   /* 0: the thermal law above; 1: two materials carried by the particles
                          (mw_sim_set_swarm_material): eta = mat_eta[m],
                          rho = mat_rho[m] (1 - alphaT T), as the reference's material plugins
                          with particle material weights (experiment21.py:94-107 pattern) */
  double mat_eta[2], mat_rho[2];
  double gravity;      /* > 0: g_z = -gravity; else g_z = -Ra / alphaT (experiment0.py:274) */
  int gmg_eta_coarsen; /* multigrid coarse-level viscosity from the 4 children: 0 arithmetic
                          (default), 1 geometric, 2 harmonic mean */
  int fgmres_maxit;    /* FGMRES iterations per defect-correction step at most (<= 0: 300) */
  int pcg_maxit;       /* plain solver: V-cycle-preconditioned CG iterations per velocity solve
                          at most (<= 0: 500) */
  /* viscosity law (material_mode 0): 0 iso / Frank-Kamenetskii (eta_contrast); 1 the
     visco-plastic law of Tosi et al. (2015), experiment3.py:67-110, case tosi_case 1..5:
     eta_lin = exp(-gamma_T T + gamma_y (1 - z)) (gamma_y only in cases 3-5), cases 2, 4, 5
     harmonic with eta_plast = eta_star + sigma_y / (sqrt(2) e), e the effective strain rate
     interpolated on the particles from the nodal strain rate of the previous Stokes solve
     (compute_strain_rate.py:50-73, zero before the first), clipped to [1e-5, 2] */
  int rheology;
  int tosi_case;
  double tosi_gamma_T, tosi_gamma_y, tosi_sigma_y, tosi_eta_star;
  /* Picard iterations within a timestep (meeuuw.py:1105, 2084-2101, 3409-3415;
     assess_steady_state.py:10-74): repeat material, Stokes and energy until the relative
     changes of u, w, p (and T against T_mem) fall below tol_nl, at most niter_nl times;
     advection, the dt memory and the time advance only after the last iteration */
  int nonlinear;
  int niter_nl;        /* <= 0: 25 */
  double tol_nl;       /* <= 0: 1e-3 */
  /* population control (pic_functions.py:1623-1825, called at meeuuw.py:2847-2851 between
     advection and locate): elements whose particle count at the start of the step is below
     nparticle_min get nparticle_min - count inactive particles, taken in increasing index
     order and given to the flagged elements in increasing element order, at reduced
     coordinates r, t uniform in [-0.95, 0.95] and the material of the nearest particle that
     was active in the element; fails (MW_ERANGE) when fewer inactive particles are left
     than needed, as the reference does.  r, t come from a counter-based generator:
     u = splitmix64(pc_seed * 0x9E3779B97F4A7C15 + istep * 0xD1B54A32D192ED03 + 2k + c)
     >> 11 times 2^-53 (c = 0 for r, 1 for t, k the new particle's rank), -0.95 + 1.9 u */
  int population_control;
  int nparticle_min;   /* <= 0: int(0.85 ppe) (set_default_parameters.py:96) */
  int pc_seed;
  /* density and viscosity at the quadrature points (meeuuw.py:1503-1560; rk > 0; rk = -1
     always uses the particles on the points): 0 'elemental' (default), 1 'nodal' (the
     nodal projection through the Q1 functions), 2 'least_squares_P1' (per-element
     linear fit with the slope limiter, meeuuw.py:1224-1249), 3 'least_squares_Q1'
     (bilinear fit, no limiter, 1276-1290).  The two fits, when both are least squares,
     must be of the same kind (the reference fails for a P1/Q1 mix).  The fits take the
     element's active particles. */
  int rho_projection;
  int eta_projection;
  int gmg_coarse_dofs; /* the multigrid coarsens until a level has at most this many velocity
                          unknowns (<= 0: 60, the 2x2-element level of a square box; at most
                          2304); that level is solved with its dense inverse, formed on a
                          side stream each solve: in one workgroup up to 64 unknowns, by a
                          blocked Gauss-Jordan over the whole GPU above (the fused bottom
                          kernels need the small one) */
} mw_sim_config;

enum {
  MW_STAT_DT = 0, MW_STAT_VRMS, MW_STAT_NU, MW_STAT_STOKES_OUTER, MW_STAT_STOKES_VCYCLES,
  MW_STAT_STOKES_RELRES, MW_STAT_ENERGY_ITS, MW_STAT_ENERGY_RELRES, MW_STAT_TAVRG,
  MW_STAT_TIME, MW_STAT_POFFSET, MW_STAT_MINCOUNT, MW_STAT_NACTIVE,
  MW_STAT_LEARNED,     /* 1 if this step's Stokes solve used the learned correction */
  MW_STAT_LEARN_MS,    /* wall time of the training (on the step that trained), ms */
  MW_STAT_LEARN_RATIO, /* validation loss with the network / without it (< 1: better) */
  MW_STAT_LEARN_ITS_PLAIN, /* FGMRES iterations on the probe systems without the network */
  MW_STAT_LEARN_ITS,   /* ... and with it (the network is used if < 0.9 x plain and its
                          probe time < 0.95 x plain) */
  /* compute_global_quantities (postprocessors.py:13-83) of the step */
  MW_STAT_EK, MW_STAT_WAG, MW_STAT_TVD, MW_STAT_TM, MW_STAT_ETA_AVRG,
  MW_STAT_DT1, MW_STAT_DT2,    /* the advective and diffusive limits (meeuuw.py:2112-2139) */
  MW_STAT_MAXCOUNT,            /* most particles in an element (meeuuw.py:1201) */
  MW_STAT_ETAE_MIN, MW_STAT_ETAE_MAX,  /* elemental viscosity extrema (meeuuw.py:1203) */
  MW_STAT_NL_ITS,              /* Picard iterations of the step (1 without nonlinear) */
  MW_STAT_NL_XI_U, MW_STAT_NL_XI_W, MW_STAT_NL_XI_P, MW_STAT_NL_XI_T,  /* last relative changes */
  MW_STAT_STOKES_FGMRES64,     /* mixed solver: FP64 FGMRES iterations after FP32 defect
                                  correction stopped contracting (0: not needed) */
  MW_STAT_PC_NEEDED, MW_STAT_PC_AVAILABLE, MW_STAT_PC_ACTIVE,  /* population control: particles
                                  needed, inactive ones available, active before it */
  MW_STAT_LEARN_PROBE_MS_PLAIN, MW_STAT_LEARN_PROBE_MS,  /* device time of the probe solves
                                  without and with the network, ms */
  MW_NSTATS = 48
};

enum {
  MW_F_T = 0, MW_F_U, MW_F_W, MW_F_P, MW_F_Q, MW_F_SX, MW_F_SZ, MW_F_SR, MW_F_ST, MW_F_SIEL,
  MW_F_SACTIVE, MW_F_RHO_E, MW_F_ETA_E, MW_F_RHO_N, MW_F_ETA_N, MW_F_DTDX, MW_F_DTDZ,
  MW_F_VV, MW_F_RHS_F, MW_F_RHS_H,
  MW_F_STOKES_HISTORY,  /* mixed solver: FGMRES residual estimates, then the true residual after
                           each defect-correction step, relative to the right-hand side */
  MW_F_EXX_N, MW_F_EZZ_N, MW_F_EXZ_N,  /* nodal strain rate (rheology 1) */
  MW_F_NL_HISTORY,      /* Picard iterations of the last step: (xi_u, xi_w, xi_p, xi_T) each */
  MW_F_SMAT,            /* particle material (uint8) */
  MW_NFIELDS
};

int mw_sim_create(const mw_sim_config* cfg, void* stream, void** sim);
int mw_sim_destroy(void* sim);
/* initial temperature on the V nodes (nn_V values) */
int mw_sim_set_temperature(void* sim, const double* T_host);
/* particles from reduced coordinates, element-major (swarm_coordinates_setup.py:71-86) */
int mw_sim_set_swarm_reduced(void* sim, const double* r_host, const double* t_host);
/* explicit particle state (n must equal nel*ppe) */
int mw_sim_set_swarm(void* sim, long n, const double* x_host, const double* z_host,
                     const double* r_host, const double* t_host, const int32_t* iel_host,
                     const uint8_t* active_host);
/* material index (0 or 1) of every particle, in particle order (material_mode 1);
 * the particles keep it as they move.  Default: all 0. */
int mw_sim_set_swarm_material(void* sim, const uint8_t* mat_host);
/* advance one timestep; stats_host receives MW_NSTATS doubles (may be NULL) */
int mw_sim_step(void* sim, double* stats_host);
/* copy a field (MW_F_*) to the host; n must equal mw_sim_field_size */
long mw_sim_field_size(void* sim, int field);
int mw_sim_get_field(void* sim, int field, void* host_out, long n);

/* Per-stage times of the last step, the device counterpart of the reference's
 * timings[] (meeuuw.py:413, print_timings.py): with stage timing on, the step
 * records a HIP event on its stream at each stage boundary (no extra
 * synchronisation) and mw_sim_stage_times fills MW_NTIMINGS doubles, seconds,
 * at the reference's indices: 1 Stokes assembly, 2 Stokes solve, 3 nodal
 * pressure, 4 energy assembly, 5 energy solve, 6 global quantities, 7 heat flux
 * and Nu, 11 nodal strain rate, 12 pressure normalisation, 13 advection, 15 particle sort, T
 * interpolation and material model, 16 locate, 17 elemental rho/eta, 18 nodal
 * projections, 19 time step, 21 fields on the quadrature points; the others
 * 0.  Off by default. */
enum { MW_NTIMINGS = 40 };
int mw_sim_stage_timing(void* sim, int on);
int mw_sim_stage_times(void* sim, double* seconds_host);

/* Reference-format output of a simulation, written by a host thread off the
 * stepping path.  Replaces the reference's per-step writers: the time series
 * opened in open_files.py:10-66 and written at meeuuw.py:1201-1203, 2148,
 * 2308, 2518-2525 (vrms, dt, Nu, T_avrg, kinetic_energy, total_mass, tvd,
 * wag, delta_wag_tvd, eta_avrg, stats_particle, stats_eta_e .ascii),
 * output_solution_to_vtu.py (solution_NNNN.vtu), output_swarm_to_vtu.py
 * (SWARM/swarm_NNNN.vtu) and write_in_pvd_files.py (solution.pvd, swarm.pvd),
 * with the arrays the timestep holds.
 *   mw_output_create: files in `dir` (created if missing); a solution every
 *     every_solution steps and a swarm every every_swarm steps (0: never);
 *     material_names: comma-separated names of the material-weight arrays of
 *     the swarm files (one per material; NULL: "material 0", ...).
 *   mw_output_step: after mw_sim_step, with the MW_NSTATS statistics that
 *     step returned; last != 0 also writes the solution and swarm of this step
 *     (the reference's istep == nstep-1).  Due fields are copied on the
 *     simulation's stream into a device staging slot, then to pinned host
 *     memory on a second stream; the files are formatted by the writer thread.
 *     Blocks only while both staging slots are still being written out.
 *   mw_output_flush: wait until every queued step is on disk.
 *   mw_output_destroy: flush, close the .pvd collections, free; call it before
 *     mw_sim_destroy of the simulation.
 * A write error is reported by the next mw_output_* call. */
int mw_output_create(void* sim, const char* dir, int every_solution, int every_swarm,
                     const char* material_names, void** out);
int mw_output_step(void* out, const double* stats_host, int last);
int mw_output_flush(void* out);
int mw_output_destroy(void* out);

/* The linear operators of the mixed-precision Stokes solve of the last step
 * (solver 1), applied on the simulation's stream to device arrays:
 *   MW_OP_STOKES32   y = A x, the Stokes matrix in FP32 (N = 2 nn_V + nn_P values;
 *                    velocity dofs 2*node+comp first, then pressure)
 *   MW_OP_PREC32     y = P^-1 x, the block upper-triangular preconditioner
 *                    (V-cycle for the velocity block, lumped M/eta for the Schur complement)
 *   MW_OP_RESTRICT32 y = R x_u: fine velocity (2 nn_V) -> once-coarsened Q2 grid
 *                    (2 (nelx+1)(nelz+1)), transpose of the prolongation, Dirichlet masked
 *   MW_OP_PROLONG32  y = P x: once-coarsened grid -> fine velocity (overwrites y)
 *   MW_OP_MASK32     y = x with the fixed dofs of the once-coarsened grid zeroed
 *   MW_OP_PREC32_SPLIT  MW_OP_PREC32 with the Schur block's Chebyshev steps and
 *                    the block front in separate launches (same results)
 * No reference counterpart (the reference solves with SuperLU). */
enum {
  MW_OP_STOKES32 = 0,
  MW_OP_PREC32,
  MW_OP_RESTRICT32,
  MW_OP_PROLONG32,
  MW_OP_MASK32,
  MW_OP_PREC32_SPLIT
};
int mw_sim_operator(void* sim, int op, const float* x, float* y);

/* Batch of independent timesteps (no reference counterpart: meeuuw runs one
 * model per process).  n instances, each with its own configuration (mesh
 * size, aspect ratio, Ra, viscosity contrast, solver ...), each a
 * device-resident simulation on its own stream; mw_batch_step runs `steps`
 * timesteps of every instance with `threads` host threads taking instances
 * in turn, so the kernels of different instances overlap on the device.
 * An instance's results are those of the same configuration run alone.
 * mw_batch_sim gives the handle of instance i for the mw_sim_* calls (set
 * its temperature and particles, read its fields); the batch owns it.
 * stats_host (n * MW_NSTATS, may be NULL) receives each instance's last-step
 * statistics, status_host (n, may be NULL) each instance's status; the call
 * fails if any instance failed, naming the first. */
int mw_batch_create(int n, const mw_sim_config* cfgs, int threads, void** batch);
int mw_batch_sim(void* batch, int i, void** sim);
int mw_batch_step(void* batch, int steps, double* stats_host, int* status_host);
int mw_batch_destroy(void* batch);

/* Convolutional network of the learned Schur-complement correction (no
 * reference counterpart: meeuuw has no learned solver component).  Four 3x3
 * convolutions 4 -> 32 -> 32 -> 32 -> 1 channels with ReLU between, on an H x W
 * grid (the pressure nodes, H = nelz+1, W = nelx+1), NHWC, zero padding;
 * bfloat16 weights and activations, FP32 accumulation.  Parameters are packed
 * FP32 (MW_CNN_NPARAM): W1 [9][32][4], b1 [32], W2 [9][32][32], b2, W3, b3,
 * W4 [9][1][32], b4 [1], tap t = 3*(dy+1) + (dx+1).  X0 and C are device
 * arrays of batch*H*W*4 and batch*H*W floats; backward() writes the gradient
 * of sum(dC * C) over the last forward pass; adam() applies one Adam step
 * (beta1 0.9, beta2 0.999, eps 1e-8) with that gradient. */
#define MW_CNN_NPARAM 19969
int mw_cnn_create(int H, int W, int max_batch, void** net);
int mw_cnn_destroy(void* net);
int mw_cnn_set_params(void* net, const float* params_host, void* stream);
int mw_cnn_get_params(void* net, float* params_host, void* stream);
int mw_cnn_get_grads(void* net, float* grads_host, void* stream);
int mw_cnn_forward(void* net, int batch, const float* X0, float* C, void* stream);
int mw_cnn_backward(void* net, int batch, const float* dC, void* stream);
int mw_cnn_adam(void* net, double lr, void* stream);

/* Average duration of `reps` back-to-back launches of one kernel of the last
 * step's solver at its real shape, timed with HIP events on the simulation's
 * stream: which = 0, the finest-level FP32 Chebyshev smoothing step (reads x,
 * b, 1/diag, d, the fixed-dof mask and the element viscosities; writes d and x).
 * The smoother state is overwritten.  No reference counterpart. */
int mw_sim_time_kernel(void* sim, int which, int reps, double* avg_us);

/* Live kernel timing with HIP events on the launching stream, for the
 * roofline figures: classes 0 assembled Stokes operator, 1 finest-level
 * Chebyshev smoother, 2 Stokes element assembly, 3 assembled energy operator,
 * 4 FGMRES Gram-Schmidt update, 5 RK advection.  mw_profile_read fills
 * 2*MW_NPROF doubles (launch count, total ms) per class.  No reference
 * counterpart (the reference times stages with wall clocks, print_timings.py). */
enum { MW_NPROF = 6 };
int mw_profile_enable(int on);
int mw_profile_read(double* out_host, int reset);

#ifdef __cplusplus
}
#endif

#endif /* MEEUUW_HIP_H */
