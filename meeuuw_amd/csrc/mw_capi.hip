// extern "C" entry points: argument checks and dispatch to the kernels.
#include <string>

#include "meeuuw_hip.h"
#include "mw_capi_impl.h"
#include "mw_common.h"

namespace mw {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

}  // namespace mw

using namespace mw;

#define ST(s) ((hipStream_t)(s))

extern "C" {

int mw_version(void) { return 100; }

const char* mw_last_error(void) { return g_err.c_str(); }

int mw_local_to_global_V(int nel, const int32_t* icon_V, int32_t* l2g, void* stream) {
  return local_to_global(nel, icon_V, l2g, ST(stream));
}

int mw_coo_stokes(int nel, const int32_t* l2g, const int32_t* icon_P, int Nfem_V, int32_t* II,
                  int32_t* JJ, void* stream) {
  return coo_stokes(nel, l2g, icon_P, Nfem_V, II, JJ, ST(stream));
}

int mw_coo_K(int nel, const int32_t* l2g, int32_t* II, int32_t* JJ, void* stream) {
  return coo_outer(nel, 18, 18, l2g, l2g, II, JJ, ST(stream));
}

int mw_coo_G(int nel, const int32_t* l2g, const int32_t* icon_P, int32_t* II_G, int32_t* JJ_G,
             int32_t* II_GT, int32_t* JJ_GT, void* stream) {
  int rc = coo_outer(nel, 18, 4, l2g, icon_P, II_G, JJ_G, ST(stream));
  if (rc) return rc;
  return coo_outer(nel, 4, 18, icon_P, l2g, II_GT, JJ_GT, ST(stream));
}

int mw_coo_scalar(int nel, int m, const int32_t* icon, int32_t* II, int32_t* JJ, void* stream) {
  return coo_outer(nel, m, m, icon, icon, II, JJ, ST(stream));
}

size_t mw_incidence_workspace_bytes(int nn) {
  if (nn <= 0) return 0;
  return sizeof(int) * 2 * ((size_t)nn + 1) + 256 + incidence_temp_bytes(nn);
}

int mw_incidence_build(int nel, int m, const int32_t* icon, int nn, int32_t* ptr, int32_t* list,
                       void* work, void* stream) {
  return incidence_build(nel, m, icon, nn, ptr, list, work, ST(stream));
}

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
                           double* VV_H, double* work, void* stream) {
  StokesArgs a{};
  a.nel = nel; a.Nfem_V = Nfem_V; a.Nfem_P = Nfem_P; a.blocks = blocks;
  a.icon_P = icon_P; a.l2g = l2g; a.rhoq = rhoq; a.etaq = etaq; a.JxWq = JxWq;
  a.gxq = gxq; a.gyq = gyq; a.N_V = N_V; a.N_P = N_P; a.dNdr_V = dNdr_V; a.dNdt_V = dNdt_V;
  a.dNdr_P = dNdr_P; a.dNdt_P = dNdt_P; a.j00 = jcbi00q; a.j01 = jcbi01q; a.j10 = jcbi10q;
  a.j11 = jcbi11q; a.eta_e = eta_e; a.eta_ref = eta_ref; a.L_ref = L_ref;
  a.bc_fix = bc_fix_V; a.bc_val = bc_val_V; a.inc_v_ptr = inc_v_ptr; a.inc_v_list = inc_v_list;
  a.inc_p_ptr = inc_p_ptr; a.inc_p_list = inc_p_list; a.VV = VV_Stokes; a.rhs_f = rhs_f;
  a.rhs_h = rhs_h; a.VV_K = VV_K; a.VV_G = VV_G; a.VV_GT = VV_GT; a.VV_M = VV_M;
  a.VV_Me = VV_M_eta; a.VV_H = VV_H; a.work = work;
  if (!(L_ref != 0.0)) return fail(MW_EINVAL, "build_matrix_stokes: L_ref must be nonzero");
  return build_matrix_stokes(a, ST(stream));
}

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
    double* VV_M, double* VV_M_eta, double* VV_H, double* work, void* stream) {
  StokesArgs a{};
  a.nel = nel; a.Nfem_V = Nfem_V; a.Nfem_P = Nfem_P; a.blocks = blocks;
  a.icon_P = icon_P; a.l2g = l2g; a.rhoq = rhoq; a.etaq = etaq; a.JxWq = JxWq;
  a.gxq = gxq; a.gyq = gyq; a.N_V = N_V; a.N_P = N_P; a.dNdr_V = dNdr_V; a.dNdt_V = dNdt_V;
  a.dNdr_P = dNdr_P; a.dNdt_P = dNdt_P; a.j00 = jcbi00q; a.j01 = jcbi01q; a.j10 = jcbi10q;
  a.j11 = jcbi11q; a.eta_e = eta_e; a.eta_ref = eta_ref; a.L_ref = L_ref;
  a.bc_fix = bc_fix_V; a.bc_val = bc_val_V; a.inc_v_ptr = inc_v_ptr; a.inc_v_list = inc_v_list;
  a.inc_p_ptr = inc_p_ptr; a.inc_p_list = inc_p_list; a.VV = VV_Stokes; a.rhs_f = rhs_f;
  a.rhs_h = rhs_h; a.VV_K = VV_K; a.VV_G = VV_G; a.VV_GT = VV_GT; a.VV_M = VV_M;
  a.VV_Me = VV_M_eta; a.VV_H = VV_H; a.work = work;
  // the free-slip rotation belongs to the curved geometries only (the box's
  // free slip is a Dirichlet condition on the normal component)
  a.bot_el = bot_element; a.top_el = top_element;
  a.bot_fs = curved && bot_free_slip; a.top_fs = curved && top_free_slip;
  a.theta = (a.bot_fs || a.top_fs) ? theta_V : nullptr;
  a.axi = axisymmetric != 0; a.xq = xq;
  if (!(L_ref != 0.0)) return fail(MW_EINVAL, "build_matrix_stokes: L_ref must be nonzero");
  if ((a.bot_fs || a.top_fs) && !theta_V)
    return fail(MW_EINVAL, "build_matrix_stokes: curved free slip needs theta_V");
  if (a.axi && !xq) return fail(MW_EINVAL, "build_matrix_stokes: axisymmetric needs xq");
  return build_matrix_stokes(a, ST(stream));
}

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
                           double* rhs, double* work, void* stream) {
  EnergyArgs a;
  a.nel = nel; a.Nfem_T = Nfem_T; a.eba = eba; a.T = T; a.icon_T = icon_T; a.etaq = etaq;
  a.Tq = Tq; a.uq = uq; a.wq = wq; a.hcondq = hcondq; a.hcapaq = hcapaq; a.alphaq = alphaq;
  a.hprodq = hprodq; a.exxq = exxq; a.ezzq = ezzq; a.exzq = exzq; a.JxWq = JxWq; a.N_T = N_T;
  a.dNdr_T = dNdr_T; a.dNds_T = dNds_T; a.j00 = jcbi00q; a.j01 = jcbi01q; a.j10 = jcbi10q;
  a.j11 = jcbi11q; a.bc_fix = bc_fix_T; a.bc_val = bc_val_T; a.dt = dt; a.rho0 = rho0;
  a.inc_ptr = inc_ptr; a.inc_list = inc_list; a.VV = VV_T; a.rhs = rhs; a.work = work;
  return build_matrix_energy(a, ST(stream));
}

int mw_locate_particles_box(int np, const uint8_t* active, const double* x, const double* z,
                            double hx, double hz, const double* x_V, const double* z_V,
                            const int32_t* icon_V, int nel, int nelx, double* r, double* t,
                            int32_t* iel, int32_t* flags, void* stream) {
  if (!flags) return fail(MW_EINVAL, "locate_particles_box: flags is null");
  MW_HIP(hipMemsetAsync(flags, 0, 3 * sizeof(int32_t), ST(stream)));
  return locate_box(np, active, x, z, hx, hz, x_V, z_V, icon_V, nel, nelx, r, t, iel, flags,
                    ST(stream));
}

int mw_interpolate_fields(int np, int nf, const uint8_t* active, const double* r,
                          const double* t, const int32_t* iel, const double* const* phi_host,
                          const int32_t* icon, int nel, double* const* out_host, void* stream) {
  if (!phi_host || !out_host) return fail(MW_EINVAL, "interpolate_fields: null field list");
  return interpolate_fields(np, nf, active, r, t, iel, phi_host, icon, nel, out_host, ST(stream));
}

int mw_advect_particles_box(int rk, double dt, int np, double* x, double* z, uint8_t* active,
                            const double* u, const double* w, double Lx, double Lz, double hx,
                            double hz, int nelx, int nel, const int32_t* icon_V,
                            const double* x_V, const double* z_V, double* su, double* sw,
                            int32_t* flags, void* stream) {
  if (!flags) return fail(MW_EINVAL, "advect_particles_box: flags is null");
  MW_HIP(hipMemsetAsync(flags, 0, 2 * sizeof(int32_t), ST(stream)));
  return advect_box(rk, dt, np, x, z, active, u, w, Lx, Lz, hx, hz, nelx, nel, icon_V, x_V, z_V,
                    su, sw, flags, ST(stream));
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t mw_sort_workspace_bytes(int np, int nel) {
  if (np <= 0 || nel <= 0) return 0;
  return sort_work_bytes(np, nel);
}

int mw_sort_particles_by_element(int np, int nel, const uint8_t* active, const int32_t* iel,
                                 int32_t* order, int32_t* seg, void* work, void* stream) {
  if (np <= 0 || nel <= 0 || !active || !iel || !order || !seg || !work)
    return fail(MW_EINVAL, "sort_particles_by_element: bad arguments");
  return sort_by_element(np, nel, active, iel, order, seg, work, ST(stream));
}

int mw_project_particles_on_elements(int nel, const int32_t* seg, const int32_t* order,
                                     const double* rho, const double* eta, int averaging,
                                     double* rho_e, double* eta_e, int32_t* count, void* stream) {
  if (!seg || !order || !rho || !eta || !rho_e || !eta_e || !count)
    return fail(MW_EINVAL, "project_particles_on_elements: null pointer argument");
  return project_on_elements(nel, seg, order, rho, eta, averaging, rho_e, eta_e, count,
                             ST(stream));
}

int mw_project_particle_fields_on_nodes(int type, int nel, int nn, int nf, const int32_t* seg,
                                        const int32_t* order, const double* const* phi_host,
                                        const int32_t* averaging_host, const int32_t* icon_V,
                                        const double* r, const double* t,
                                        const int32_t* inc_ptr, const int32_t* inc_list,
                                        double* work, double* const* out_host, void* stream) {
  if (!phi_host || !out_host || !averaging_host)
    return fail(MW_EINVAL, "project_particle_fields_on_nodes: null field list");
  return project_on_nodes(type, nel, nn, nf, seg, order, phi_host, averaging_host, icon_V, r, t,
                          inc_ptr, inc_list, work, out_host, ST(stream));
}

int mw_nodal_to_qpoints(int nf, const double* const* phi_host, int nel, int m, const double* N,
                        const int32_t* icon, double* const* phiq_host, void* stream) {
  if (!phi_host || !phiq_host) return fail(MW_EINVAL, "nodal_to_qpoints: null field list");
  return nodal_to_qpoints(nf, phi_host, nel, m, N, icon, phiq_host, ST(stream));
}

int mw_pressure_offset(int mode, const uint8_t* top_element, int nel, int nelx,
                       const double* N_P, const double* JxWq, const double* p,
                       const int32_t* icon_P, double volume, double* out, double* work,
                       void* stream) {
  return pressure_offset(mode, top_element, nel, nelx, N_P, JxWq, p, icon_P, volume, out, work,
                         ST(stream));
}

int mw_pressure_offset_axisymmetric(int mode, const uint8_t* top_element, int nel,
                                    const double* N_P, const double* JxWq, const double* p,
                                    const int32_t* icon_P, const double* theta_P,
                                    const double* xq, double* out, double* work, void* stream) {
  return pressure_offset_axi(mode, top_element, nel, N_P, JxWq, p, icon_P, theta_P, xq, out, work,
                             ST(stream));
}

int mw_nodal_pressure(int nel, int nn, const int32_t* icon_P, const double* p,
                      const double* N_P_n, const int32_t* inc_ptr, const int32_t* inc_list,
                      double* q, double* work, void* stream) {
  return nodal_pressure(nel, nn, nullptr, icon_P, p, N_P_n, inc_ptr, inc_list, work, q,
                        ST(stream));
}

int mw_nodal_heat_flux(int nel, int nn, const int32_t* icon_T, const double* T,
                       const double* hcond_n, const double* dNdr_n, const double* dNds_n,
                       const double* jcbi00, const double* jcbi01, const double* jcbi10,
                       const double* jcbi11, const int32_t* inc_ptr, const int32_t* inc_list,
                       double* dTdx, double* dTdy, double* qx, double* qy, double* work,
                       void* stream) {
  return nodal_heat_flux(icon_T, T, hcond_n, nn, nel, dNdr_n, dNds_n, jcbi00, jcbi01, jcbi10,
                         jcbi11, inc_ptr, inc_list, work, dTdx, dTdy, qx, qy, ST(stream));
}

int mw_global_quantities(int nel, const double* uq, const double* wq, const double* Tq,
                         const double* rhoq, const double* hcapaq, const double* etaq,
                         const double* exxq, const double* ezzq, const double* exzq,
                         double volume, const double* JxWq, const double* gxq,
                         const double* gzq, double* out, double* work, void* stream) {
  return global_quantities(nel, uq, wq, Tq, rhoq, hcapaq, etaq, exxq, ezzq, exzq, volume, JxWq,
                           gxq, gzq, out, work, ST(stream));
}

int mw_compute_Nu(double Lx, double Lz, const double* x_V, const double* z_V, int nel,
                  const uint8_t* left, const uint8_t* right, const uint8_t* top,
                  const uint8_t* bottom, const int32_t* icon_V, const double* T,
                  const double* dTdx, const double* dTdz, const double* qc, const double* qw,
                  const double* hcond_n, double* out, double* work, void* stream) {
  return compute_Nu(Lx, Lz, x_V, z_V, nel, left, right, top, bottom, icon_V, T, dTdx, dTdz, qc,
                    qw, hcond_n, out, work, ST(stream));
}

}  // extern "C"
