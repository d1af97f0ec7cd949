/*
 * CPU FP64 oracle for the meeuuw per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * reference algorithms; it is the checker for the HIP kernels and the timed
 * CPU baseline.  Nothing in the shipped package may link, load or call it.
 *
 * Layout conventions follow the reference exactly:
 *   icon_V  int32 (9, nel)  row-major   -> icon_V[k*nel + iel]
 *   icon_P  int32 (4, nel)
 *   l2g     int32 (18, nel)
 *   qpoint arrays f64 (nel, 9)          -> a[iel*9 + iq]
 *   N_V (9 qpoints, 9), N_P (9, 4), dNdr/dNdt likewise
 *   bool arrays are uint8
 * Compiled with -ffp-contract=off so a*b+c is never fused, as in the
 * reference's non-fastmath Numba code.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MV 9
#define MP 4
#define NDV 18
#define ORC_PI 3.14159265358979323846 /* rounds to numpy.pi */

/* ---------------------------------------------------------------------- */
/* basis functions: basis_functions.py:20-138 (same written expressions)   */
/* ---------------------------------------------------------------------- */
void orc_basis_V(double r, double s, double *N)
{
    N[0] = 0.5 * r * (r - 1.0) * 0.5 * s * (s - 1.0);
    N[1] = 0.5 * r * (r + 1.0) * 0.5 * s * (s - 1.0);
    N[2] = 0.5 * r * (r + 1.0) * 0.5 * s * (s + 1.0);
    N[3] = 0.5 * r * (r - 1.0) * 0.5 * s * (s + 1.0);
    N[4] = (1.0 - r * r) * 0.5 * s * (s - 1.0);
    N[5] = 0.5 * r * (r + 1.0) * (1.0 - s * s);
    N[6] = (1.0 - r * r) * 0.5 * s * (s + 1.0);
    N[7] = 0.5 * r * (r - 1.0) * (1.0 - s * s);
    N[8] = (1.0 - r * r) * (1.0 - s * s);
}

void orc_basis_V_dr(double r, double s, double *d)
{
    d[0] = 0.5 * (2.0 * r - 1.0) * 0.5 * s * (s - 1);
    d[1] = 0.5 * (2.0 * r + 1.0) * 0.5 * s * (s - 1);
    d[2] = 0.5 * (2.0 * r + 1.0) * 0.5 * s * (s + 1);
    d[3] = 0.5 * (2.0 * r - 1.0) * 0.5 * s * (s + 1);
    d[4] = (-2.0 * r) * 0.5 * s * (s - 1);
    d[5] = 0.5 * (2.0 * r + 1.0) * (1.0 - s * s);
    d[6] = (-2.0 * r) * 0.5 * s * (s + 1);
    d[7] = 0.5 * (2.0 * r - 1.0) * (1.0 - s * s);
    d[8] = (-2.0 * r) * (1.0 - s * s);
}

void orc_basis_V_dt(double r, double s, double *d)
{
    d[0] = 0.5 * r * (r - 1.0) * 0.5 * (2.0 * s - 1.0);
    d[1] = 0.5 * r * (r + 1.0) * 0.5 * (2.0 * s - 1.0);
    d[2] = 0.5 * r * (r + 1.0) * 0.5 * (2.0 * s + 1.0);
    d[3] = 0.5 * r * (r - 1.0) * 0.5 * (2.0 * s + 1.0);
    d[4] = (1.0 - r * r) * 0.5 * (2.0 * s - 1.0);
    d[5] = 0.5 * r * (r + 1.0) * (-2.0 * s);
    d[6] = (1.0 - r * r) * 0.5 * (2.0 * s + 1.0);
    d[7] = 0.5 * r * (r - 1.0) * (-2.0 * s);
    d[8] = (1.0 - r * r) * (-2.0 * s);
}

void orc_basis_P(double r, double s, double *N)
{
    N[0] = 0.25 * (1 - r) * (1 - s);
    N[1] = 0.25 * (1 + r) * (1 - s);
    N[2] = 0.25 * (1 + r) * (1 + s);
    N[3] = 0.25 * (1 - r) * (1 + s);
}

void orc_basis_P_dr(double r, double s, double *d)
{
    (void)r;
    d[0] = -0.25 * (1 - s);
    d[1] = 0.25 * (1 - s);
    d[2] = 0.25 * (1 + s);
    d[3] = -0.25 * (1 + s);
}

void orc_basis_P_dt(double r, double s, double *d)
{
    (void)s;
    d[0] = -0.25 * (1 - r);
    d[1] = -0.25 * (1 + r);
    d[2] = 0.25 * (1 + r);
    d[3] = 0.25 * (1 - r);
}

static double dot9(const double *a, const double *b)
{
    double s = 0.0;
    for (int k = 0; k < 9; k++) s += a[k] * b[k];
    return s;
}

/* ---------------------------------------------------------------------- */
/* mesh: build_velocity_mesh.py:110-146, build_velocity_connectivity.py:61-94,
 *       build_pressure_connectivity.py:15-25, build_pressure_mesh.py (box) */
/* ---------------------------------------------------------------------- */
void orc_mesh_box(int nelx, int nelz, double Lx, double Lz,
                  double *x_V, double *z_V, int *icon_V, int *icon_P,
                  double *x_P, double *z_P)
{
    const double hx = Lx / nelx, hz = Lz / nelz;
    const int nnx = 2 * nelx + 1;
    const int nel = nelx * nelz;
    int c = 0;
    for (int j = 0; j < 2 * nelz + 1; j++)
        for (int i = 0; i < 2 * nelx + 1; i++) {
            x_V[c] = i * hx / 2;
            z_V[c] = j * hz / 2;
            c++;
        }
    c = 0;
    for (int j = 0; j < nelz; j++)
        for (int i = 0; i < nelx; i++) {
            icon_V[0 * nel + c] = i * 2 + 1 + j * 2 * nnx - 1;
            icon_V[1 * nel + c] = i * 2 + 3 + j * 2 * nnx - 1;
            icon_V[2 * nel + c] = i * 2 + 3 + j * 2 * nnx + nnx * 2 - 1;
            icon_V[3 * nel + c] = i * 2 + 1 + j * 2 * nnx + nnx * 2 - 1;
            icon_V[4 * nel + c] = i * 2 + 2 + j * 2 * nnx - 1;
            icon_V[5 * nel + c] = i * 2 + 3 + j * 2 * nnx + nnx - 1;
            icon_V[6 * nel + c] = i * 2 + 2 + j * 2 * nnx + nnx * 2 - 1;
            icon_V[7 * nel + c] = i * 2 + 1 + j * 2 * nnx + nnx - 1;
            icon_V[8 * nel + c] = i * 2 + 2 + j * 2 * nnx + nnx - 1;
            icon_P[0 * nel + c] = i + j * (nelx + 1);
            icon_P[1 * nel + c] = i + 1 + j * (nelx + 1);
            icon_P[2 * nel + c] = i + 1 + (j + 1) * (nelx + 1);
            icon_P[3 * nel + c] = i + (j + 1) * (nelx + 1);
            c++;
        }
    c = 0;
    for (int j = 0; j < nelz + 1; j++)
        for (int i = 0; i < nelx + 1; i++) {
            x_P[c] = i * hx;
            z_P[c] = j * hz;
            c++;
        }
}

/* ---------------------------------------------------------------------- */
/* reference-element tables and per-element Jacobians:
 * basis_functions_setup_q.py:47-79 (qpoint order cq = iq*3 + jq,
 * r = q[iq], t = q[jq]); basis_functions_setup_Vnodes.py (node tables).   */
/* ---------------------------------------------------------------------- */
void orc_qpoint_tables(const double *qcoords, const double *qweights, int nqd,
                       double *rq, double *tq, double *wq,
                       double *N_V, double *dNdr_V, double *dNdt_V,
                       double *N_P, double *dNdr_P, double *dNdt_P)
{
    int cq = 0;
    for (int iq = 0; iq < nqd; iq++)
        for (int jq = 0; jq < nqd; jq++) {
            rq[cq] = qcoords[iq];
            tq[cq] = qcoords[jq];
            wq[cq] = qweights[iq] * qweights[jq];
            orc_basis_V(rq[cq], tq[cq], N_V + cq * MV);
            orc_basis_V_dr(rq[cq], tq[cq], dNdr_V + cq * MV);
            orc_basis_V_dt(rq[cq], tq[cq], dNdt_V + cq * MV);
            orc_basis_P(rq[cq], tq[cq], N_P + cq * MP);
            orc_basis_P_dr(rq[cq], tq[cq], dNdr_P + cq * MP);
            orc_basis_P_dt(rq[cq], tq[cq], dNdt_P + cq * MP);
            cq++;
        }
}

/* Jacobian inverse and JxW at nq points per element whose reference
 * derivative tables are given (qpoints or the 9 V nodes). */
void orc_jacobians(int nel, int nq, const double *dNdr, const double *dNdt,
                   const double *w, const double *x_V, const double *z_V,
                   const int *icon_V, double *JxW, double *j00, double *j01,
                   double *j10, double *j11)
{
    for (int iel = 0; iel < nel; iel++) {
        double xe[9], ze[9];
        for (int k = 0; k < 9; k++) {
            xe[k] = x_V[icon_V[k * nel + iel]];
            ze[k] = z_V[icon_V[k * nel + iel]];
        }
        for (int q = 0; q < nq; q++) {
            double a = dot9(dNdr + q * 9, xe), b = dot9(dNdr + q * 9, ze);
            double c = dot9(dNdt + q * 9, xe), d = dot9(dNdt + q * 9, ze);
            double det = a * d - b * c;
            if (JxW) JxW[iel * nq + q] = det * (w ? w[q] : 1.0);
            j00[iel * nq + q] = d / det;
            j01[iel * nq + q] = -b / det;
            j10[iel * nq + q] = -c / det;
            j11[iel * nq + q] = a / det;
        }
    }
}

/* ---------------------------------------------------------------------- */
/* COO index arrays: meeuuw.py:874-1015                                    */
/* ---------------------------------------------------------------------- */
void orc_local_to_global_V(int nel, const int *icon_V, int *l2g)
{
    for (int iel = 0; iel < nel; iel++)
        for (int k1 = 0; k1 < MV; k1++)
            for (int i1 = 0; i1 < 2; i1++)
                l2g[(2 * k1 + i1) * nel + iel] = 2 * icon_V[k1 * nel + iel] + i1;
}

void orc_coo_stokes(int nel, const int *l2g, const int *icon_P, int Nfem_V,
                    int *II, int *JJ)
{
    long c = 0;
    for (int iel = 0; iel < nel; iel++)
        for (int ikk = 0; ikk < NDV; ikk++) {
            int m1 = l2g[ikk * nel + iel];
            for (int jkk = 0; jkk < NDV; jkk++) {
                II[c] = m1;
                JJ[c] = l2g[jkk * nel + iel];
                c++;
            }
            for (int jkk = 0; jkk < MP; jkk++) {
                int m2 = icon_P[jkk * nel + iel] + Nfem_V;
                II[c] = m1; JJ[c] = m2; c++;
                II[c] = m2; JJ[c] = m1; c++;
            }
        }
}

void orc_coo_K(int nel, const int *l2g, int *II, int *JJ)
{
    long c = 0;
    for (int iel = 0; iel < nel; iel++)
        for (int ikk = 0; ikk < NDV; ikk++)
            for (int jkk = 0; jkk < NDV; jkk++) {
                II[c] = l2g[ikk * nel + iel];
                JJ[c] = l2g[jkk * nel + iel];
                c++;
            }
}

void orc_coo_G(int nel, const int *l2g, const int *icon_P, int *II, int *JJ,
               int *IIT, int *JJT)
{
    long c = 0;
    for (int iel = 0; iel < nel; iel++)
        for (int ikk = 0; ikk < NDV; ikk++)
            for (int jkk = 0; jkk < MP; jkk++) {
                II[c] = l2g[ikk * nel + iel];
                JJ[c] = icon_P[jkk * nel + iel];
                c++;
            }
    c = 0;
    for (int iel = 0; iel < nel; iel++)
        for (int jkk = 0; jkk < MP; jkk++)
            for (int ikk = 0; ikk < NDV; ikk++) {
                IIT[c] = icon_P[jkk * nel + iel];
                JJT[c] = l2g[ikk * nel + iel];
                c++;
            }
}

/* scalar connectivity COO (MP: m=4 with icon_P; T: m=9 with icon_T) */
void orc_coo_scalar(int nel, int m, const int *icon, int *II, int *JJ)
{
    long c = 0;
    for (int iel = 0; iel < nel; iel++)
        for (int ikk = 0; ikk < m; ikk++)
            for (int jkk = 0; jkk < m; jkk++) {
                II[c] = icon[ikk * nel + iel];
                JJ[c] = icon[jkk * nel + iel];
                c++;
            }
}

/* ---------------------------------------------------------------------- */
/* build_matrix_stokes.py:99-276 (plane strain)                            */
/* ---------------------------------------------------------------------- */
/* Free slip on a curved boundary node, build_matrix_stokes.py:166-182: the
 * element system is rotated to (normal, tangential) components at local node
 * i, K <- RR K RR^T, G <- RR G, f <- RR f with RR the identity except
 * RR[n][n] = RR[t][t] = cos(theta), RR[n][t] = sin(theta),
 * RR[t][n] = -sin(theta), then the normal dof n is eliminated (row and column
 * zeroed, diagonal kept; its G row and f entry zeroed).  Each product is the
 * sum of the two nonzero terms of RR's row or column, in increasing column
 * order, as the dense products reduce to. */
static void orc_free_slip_rotate(double K[NDV][NDV], double G[NDV][MP], double f[NDV], int i,
                                 double theta)
{
    const int n = 2 * i, t = 2 * i + 1;
    const double c = cos(theta), s = sin(theta);
    for (int a = 0; a < NDV; a++) { /* K_el.dot(RR.T) */
        const double kn = K[a][n], kt = K[a][t];
        K[a][n] = kn * c + kt * s;
        K[a][t] = kn * (-s) + kt * c;
    }
    for (int b = 0; b < NDV; b++) { /* RR.dot(.) */
        const double rn = K[n][b], rt = K[t][b];
        K[n][b] = c * rn + s * rt;
        K[t][b] = (-s) * rn + c * rt;
    }
    for (int j = 0; j < MP; j++) {
        const double gn = G[n][j], gt = G[t][j];
        G[n][j] = c * gn + s * gt;
        G[t][j] = (-s) * gn + c * gt;
    }
    const double fn = f[n], ft = f[t];
    f[n] = c * fn + s * ft;
    f[t] = (-s) * fn + c * ft;
    const double Kref = K[n][n];
    for (int a = 0; a < NDV; a++) { K[n][a] = 0; K[a][n] = 0; }
    K[n][n] = Kref;
    for (int j = 0; j < MP; j++) G[n][j] = 0;
    f[n] = 0;
}

void orc_build_matrix_stokes(
    int nel, const int *icon_V, const int *icon_P, const double *rhoq,
    const double *etaq, const double *JxWq, const int *l2g, const double *gxq,
    const double *gyq, const double *N_V, const double *N_P,
    const double *dNdr_V, const double *dNdt_V, const double *dNdr_P,
    const double *dNdt_P, const double *j00q, const double *j01q,
    const double *j10q, const double *j11q, const double *eta_e,
    double eta_ref, double L_ref, const uint8_t *bc_fix_V,
    const double *bc_val_V, int Nfem_V, int Nfem_P, int blocks,
    double *VV_Stokes, double *rhs_f, double *rhs_h, double *VV_K,
    double *VV_G, double *VV_GT, double *VV_M, double *VV_M_eta, double *VV_H,
    int curved, const double *theta_V, const uint8_t *bot_element,
    const uint8_t *top_element, int bot_free_slip, int top_free_slip, int axi,
    const double *xq)
{
    static const int top_nodes[3] = {2, 3, 6}, bot_nodes[3] = {0, 1, 4};
    memset(rhs_f, 0, sizeof(double) * Nfem_V);
    memset(rhs_h, 0, sizeof(double) * Nfem_P);
    long cS = 0, cK = 0, cG = 0, cM = 0;
    for (int iel = 0; iel < nel; iel++) {
        double K[NDV][NDV], G[NDV][MP], f[NDV], h[MP], M[MP][MP], H[MP][NDV];
        memset(K, 0, sizeof K); memset(G, 0, sizeof G); memset(f, 0, sizeof f);
        memset(h, 0, sizeof h); memset(M, 0, sizeof M); memset(H, 0, sizeof H);
        for (int iq = 0; axi && iq < 9; iq++) {
            /* axisymmetric branch, build_matrix_stokes.py:108-125: B rows
             * (du/dx, u/x, dw/dz, du/dz + dw/dx), C = diag(2, 2, 2, 1), every
             * term weighted by 2 pi x; no M or H blocks */
            const int e = iel * 9 + iq;
            double dNdx[MV], dNdz[MV];
            for (int i = 0; i < MV; i++) {
                dNdx[i] = j00q[e] * dNdr_V[iq * MV + i] + j01q[e] * dNdt_V[iq * MV + i];
                dNdz[i] = j10q[e] * dNdr_V[iq * MV + i] + j11q[e] * dNdt_V[iq * MV + i];
            }
            const double coeffq = 2 * ORC_PI * xq[e];
            double B[4][NDV];
            memset(B, 0, sizeof B);
            for (int i = 0; i < MV; i++) {
                B[0][2 * i] = dNdx[i];
                B[1][2 * i] = N_V[iq * MV + i] / xq[e];
                B[2][2 * i + 1] = dNdz[i];
                B[3][2 * i] = dNdz[i];
                B[3][2 * i + 1] = dNdx[i];
            }
            const double ew = etaq[e], jw = JxWq[e];
            for (int i = 0; i < NDV; i++)
                for (int j = 0; j < NDV; j++) {
                    double btcb = B[0][i] * (2.0 * B[0][j]) + B[1][i] * (2.0 * B[1][j]) +
                                  B[2][i] * (2.0 * B[2][j]) + B[3][i] * (1.0 * B[3][j]);
                    K[i][j] += btcb * ew * jw * coeffq;
                }
            for (int i = 0; i < MV; i++) {
                f[2 * i] += N_V[iq * MV + i] * jw * coeffq * rhoq[e] * gxq[e];
                f[2 * i + 1] += N_V[iq * MV + i] * jw * coeffq * rhoq[e] * gyq[e];
            }
            /* N_mat rows 0-2 = N_P, row 3 = 0 */
            for (int i = 0; i < NDV; i++)
                for (int j = 0; j < MP; j++) {
                    const double np_ = N_P[iq * MP + j];
                    double bn = B[0][i] * np_ + B[1][i] * np_ + B[2][i] * np_ + B[3][i] * 0.0;
                    G[i][j] -= bn * jw * coeffq;
                }
        }
        for (int iq = 0; !axi && iq < 9; iq++) {
            const int e = iel * 9 + iq;
            double dNdx[MV], dNdz[MV], dPdx[MP], dPdz[MP];
            for (int i = 0; i < MV; i++) {
                dNdx[i] = j00q[e] * dNdr_V[iq * MV + i] + j01q[e] * dNdt_V[iq * MV + i];
                dNdz[i] = j10q[e] * dNdr_V[iq * MV + i] + j11q[e] * dNdt_V[iq * MV + i];
            }
            for (int i = 0; i < MP; i++) {
                dPdx[i] = j00q[e] * dNdr_P[iq * MP + i] + j01q[e] * dNdt_P[iq * MP + i];
                dPdz[i] = j10q[e] * dNdr_P[iq * MP + i] + j11q[e] * dNdt_P[iq * MP + i];
            }
            double B[3][NDV];
            memset(B, 0, sizeof B);
            for (int i = 0; i < MV; i++) {
                B[0][2 * i] = dNdx[i];
                B[1][2 * i + 1] = dNdz[i];
                B[2][2 * i] = dNdz[i];
                B[2][2 * i + 1] = dNdx[i];
            }
            const double ew = etaq[e], jw = JxWq[e];
            /* K_el += B^T C B * eta * JxW with C = diag(2,2,1) */
            for (int i = 0; i < NDV; i++)
                for (int j = 0; j < NDV; j++) {
                    double btcb = B[0][i] * (2.0 * B[0][j]) + B[1][i] * (2.0 * B[1][j]) +
                                  B[2][i] * (1.0 * B[2][j]);
                    K[i][j] += btcb * ew * jw;
                }
            for (int i = 0; i < MV; i++) {
                f[2 * i] += N_V[iq * MV + i] * jw * rhoq[e] * gxq[e];
                f[2 * i + 1] += N_V[iq * MV + i] * jw * rhoq[e] * gyq[e];
            }
            /* G_el -= B^T N_mat JxW, N_mat rows 0,1 = N_P, row 2 = 0 */
            for (int i = 0; i < NDV; i++)
                for (int j = 0; j < MP; j++) {
                    double bn = B[0][i] * N_P[iq * MP + j] + B[1][i] * N_P[iq * MP + j] +
                                B[2][i] * 0.0;
                    G[i][j] -= bn * jw;
                }
            for (int i = 0; i < MP; i++)
                for (int j = 0; j < MP; j++)
                    M[i][j] += N_P[iq * MP + i] * N_P[iq * MP + j] * jw;
            /* H_el += aa_mat @ bb_mat * JxW */
            for (int i = 0; i < MP; i++)
                for (int k = 0; k < MV; k++) {
                    H[i][2 * k] += (dPdx[i] * N_V[iq * MV + k] + dPdz[i] * 0.0) * jw;
                    H[i][2 * k + 1] += (dPdx[i] * 0.0 + dPdz[i] * N_V[iq * MV + k]) * jw;
                }
        }
        const double gs = eta_ref / L_ref;
        for (int i = 0; i < NDV; i++)
            for (int j = 0; j < MP; j++) G[i][j] *= gs;

        /* curved free slip, top then bottom (build_matrix_stokes.py:163-203);
         * a node with either dof fixed is left alone */
        if (curved) {
            for (int side = 0; side < 2; side++) {
                const int on = side == 0 ? (top_free_slip && top_element[iel])
                                         : (bot_free_slip && bot_element[iel]);
                if (!on) continue;
                for (int k = 0; k < 3; k++) {
                    const int i = side == 0 ? top_nodes[k] : bot_nodes[k];
                    const int inode = icon_V[i * nel + iel];
                    if (!bc_fix_V[2 * inode] && !bc_fix_V[2 * inode + 1])
                        orc_free_slip_rotate(K, G, f, i, theta_V[inode]);
                }
            }
        }

        /* Dirichlet, progressive per fixed dof (build_matrix_stokes.py:206-217) */
        for (int ikk = 0; ikk < NDV; ikk++) {
            int m1 = l2g[ikk * nel + iel];
            if (bc_fix_V[m1]) {
                double Kref = K[ikk][ikk];
                double v = bc_val_V[m1];
                for (int jkk = 0; jkk < NDV; jkk++) f[jkk] -= K[jkk][ikk] * v;
                for (int jkk = 0; jkk < NDV; jkk++) { K[ikk][jkk] = 0; K[jkk][ikk] = 0; }
                K[ikk][ikk] = Kref;
                f[ikk] = Kref * v;
                for (int j = 0; j < MP; j++) h[j] -= G[ikk][j] * v;
                for (int j = 0; j < MP; j++) G[ikk][j] = 0;
            }
        }
        for (int i = 0; i < NDV; i++) rhs_f[l2g[i * nel + iel]] += f[i];
        for (int i = 0; i < MP; i++) rhs_h[icon_P[i * nel + iel]] += h[i];

        if (blocks) {
            for (int i = 0; i < NDV; i++)
                for (int j = 0; j < NDV; j++) VV_K[cK++] = K[i][j];
            for (int i = 0; i < NDV; i++)
                for (int j = 0; j < MP; j++) VV_G[cG + i * MP + j] = G[i][j];
            for (int i = 0; i < MP; i++)
                for (int j = 0; j < NDV; j++) {
                    VV_GT[cG + i * NDV + j] = G[j][i];
                    VV_H[cG + i * NDV + j] = H[i][j];
                }
            cG += NDV * MP;
            for (int i = 0; i < MP; i++)
                for (int j = 0; j < MP; j++) {
                    VV_M[cM] = M[i][j];
                    VV_M_eta[cM] = M[i][j] / eta_e[iel];
                    cM++;
                }
        } else {
            for (int i = 0; i < NDV; i++) {
                for (int j = 0; j < NDV; j++) VV_Stokes[cS++] = K[i][j];
                for (int j = 0; j < MP; j++) {
                    VV_Stokes[cS++] = G[i][j];
                    VV_Stokes[cS++] = G[i][j];
                }
            }
        }
    }
}

/* ---------------------------------------------------------------------- */
/* build_matrix_energy.py:55-139                                           */
/* ---------------------------------------------------------------------- */
void orc_build_matrix_energy(
    int nel, int Nfem_T, const double *T, const int *icon_T, const double *etaq,
    const double *Tq, const double *uq, const double *wq, const double *hcondq,
    const double *hcapaq, const double *alphaq, const double *hprodq,
    const double *exxq, const double *ezzq, const double *exzq,
    const double *JxWq, const double *N_T, const double *dNdr_T,
    const double *dNds_T, const double *j00q, const double *j01q,
    const double *j10q, const double *j11q, const uint8_t *bc_fix_T,
    const double *bc_val_T, double dt, int eba, double rho0, double *VV_T,
    double *rhs)
{
    memset(rhs, 0, sizeof(double) * Nfem_T);
    long c = 0;
    for (int iel = 0; iel < nel; iel++) {
        double b[9] = {0}, A[9][9], Ka[9][9], Kd[9][9], MM[9][9], Tv[9];
        memset(A, 0, sizeof A); memset(Ka, 0, sizeof Ka);
        memset(Kd, 0, sizeof Kd); memset(MM, 0, sizeof MM);
        for (int k = 0; k < 9; k++) Tv[k] = T[icon_T[k * nel + iel]];
        for (int iq = 0; iq < 9; iq++) {
            const int e = iel * 9 + iq;
            const double *N = N_T + iq * 9;
            double dNdx[9], dNdy[9];
            for (int i = 0; i < 9; i++) {
                dNdx[i] = j00q[e] * dNdr_T[iq * 9 + i] + j01q[e] * dNds_T[iq * 9 + i];
                dNdy[i] = j10q[e] * dNdr_T[iq * 9 + i] + j11q[e] * dNds_T[iq * 9 + i];
            }
            const double u = uq[e], w = wq[e];
            for (int i = 0; i < 9; i++)
                for (int j = 0; j < 9; j++) {
                    MM[i][j] += N[i] * N[j] * rho0 * hcapaq[e] * JxWq[e];
                    Kd[i][j] += (dNdx[i] * dNdx[j] + dNdy[i] * dNdy[j]) * hcondq[e] * JxWq[e];
                    double vb = u * dNdx[j] + w * dNdy[j];
                    Ka[i][j] += N[i] * vb * rho0 * hcapaq[e] * JxWq[e];
                }
            for (int i = 0; i < 9; i++) b[i] += N[i] * hprodq[e] * JxWq[e];
            if (eba) {
                double ex = exxq[e], ez = ezzq[e], exz = exzq[e];
                /* 2/3 exx**2 + 2/3 ezz**2 - 2/3 exx ezz + 2 exz**2, powers first */
                double sr = 2.0 / 3.0 * (ex * ex) + 2.0 / 3.0 * (ez * ez) - 2.0 / 3.0 * ex * ez +
                            2 * (exz * exz);
                for (int i = 0; i < 9; i++) b[i] += N[i] * 2 * etaq[e] * sr * JxWq[e];
                for (int i = 0; i < 9; i++)
                    b[i] += N[i] * alphaq[e] * Tq[e] * (-w * rho0 * 10) * JxWq[e];
            }
        }
        for (int i = 0; i < 9; i++) b[i] *= dt;
        for (int i = 0; i < 9; i++)
            for (int j = 0; j < 9; j++) A[i][j] += MM[i][j] + (Ka[i][j] + Kd[i][j]) * dt * 0.5;
        for (int i = 0; i < 9; i++) {
            double s = 0.0;
            for (int j = 0; j < 9; j++) s += (MM[i][j] - (Ka[i][j] + Kd[i][j]) * dt * 0.5) * Tv[j];
            b[i] += s;
        }
        for (int k1 = 0; k1 < 9; k1++) {
            int m1 = icon_T[k1 * nel + iel];
            if (bc_fix_T[m1]) {
                double Aref = A[k1][k1];
                for (int k2 = 0; k2 < 9; k2++) {
                    b[k2] -= A[k2][k1] * bc_val_T[m1];
                    A[k1][k2] = 0;
                    A[k2][k1] = 0;
                }
                A[k1][k1] = Aref;
                b[k1] = Aref * bc_val_T[m1];
            }
        }
        for (int i = 0; i < 9; i++) {
            for (int j = 0; j < 9; j++) VV_T[c++] = A[i][j];
            rhs[icon_T[i * nel + iel]] += b[i];
        }
    }
}

/* ---------------------------------------------------------------------- */
/* particles: pic_functions.py:17-292, locate_particles.py:33-64           */
/* ---------------------------------------------------------------------- */

/* locate_particles___box (pic_functions.py:129-136).  Returns the number of
 * active particles whose element index falls outside [0, nel): the reference
 * reads out of bounds there; the oracle reports it instead. */
int orc_locate_box(int np_, const uint8_t *active, const double *x,
                   const double *z, double hx, double hz, const double *x_V,
                   const double *z_V, const int *icon_V, int nel, int nelx,
                   double *r, double *t, int *iel_out)
{
    int bad = 0;
    for (int ip = 0; ip < np_; ip++) {
        r[ip] = 0.0; t[ip] = 0.0; iel_out[ip] = 0;
        if (!active[ip]) continue;
        int ielx = (int)(x[ip] / hx);
        int ielz = (int)(z[ip] / hz);
        int iel = nelx * ielz + ielx;
        if (iel < 0 || iel >= nel) { bad++; iel_out[ip] = iel; r[ip] = NAN; t[ip] = NAN; continue; }
        r[ip] = ((x[ip] - x_V[icon_V[iel]]) / hx - 0.5) * 2;
        t[ip] = ((z[ip] - z_V[icon_V[iel]]) / hz - 0.5) * 2;
        iel_out[ip] = iel;
    }
    return bad;
}

/* interpolate_field_on_particles (pic_functions.py:85-101) */
void orc_interpolate_field(int np_, const uint8_t *active, const double *r,
                           const double *t, const int *iel, const double *phi,
                           const int *icon, int nel, double *out)
{
    for (int ip = 0; ip < np_; ip++) {
        out[ip] = 0.0;
        if (!active[ip]) continue;
        double N[9], v[9];
        orc_basis_V(r[ip], t[ip], N);
        for (int k = 0; k < 9; k++) v[k] = phi[icon[k * nel + iel[ip]]];
        out[ip] = dot9(N, v);
    }
}

/* interpolate_vel_on_pt___box (pic_functions.py:17-37); returns 0 when the
 * element index leaves the mesh (reference: out-of-bounds read). */
static int vel_on_pt(double xp, double zp, const double *u, const double *w,
                     double hx, double hz, int nelx, int nel, const int *icon,
                     const double *x_V, const double *z_V, double *up,
                     double *wp)
{
    int ielx = (int)(xp / hx);
    int ielz = (int)(zp / hz);
    int iel = nelx * ielz + ielx;
    if (iel < 0 || iel >= nel) { *up = NAN; *wp = NAN; return 0; }
    double rp = ((xp - x_V[icon[iel]]) / hx - 0.5) * 2;
    double tp = ((zp - z_V[icon[iel]]) / hz - 0.5) * 2;
    double N[9], uv[9], wv[9];
    orc_basis_V(rp, tp, N);
    for (int k = 0; k < 9; k++) {
        uv[k] = u[icon[k * nel + iel]];
        wv[k] = w[icon[k * nel + iel]];
    }
    *up = dot9(N, uv);
    *wp = dot9(N, wv);
    return 1;
}

/* advect_particles___box (pic_functions.py:179-292): mutates x, z, active. */
int orc_advect_box(int rk, double dt, int np_, double *x, double *z,
                   uint8_t *active, const double *u, const double *w,
                   double Lx, double Lz, double hx, double hz, int nelx,
                   int nel, const int *icon, const double *x_V,
                   const double *z_V, double *su, double *sw)
{
    int bad = 0;
    for (int ip = 0; ip < np_; ip++) {
        su[ip] = 0.0; sw[ip] = 0.0;
        if (!active[ip]) continue;
        if (rk == 1) {
            double uu, ww;
            bad += !vel_on_pt(x[ip], z[ip], u, w, hx, hz, nelx, nel, icon, x_V, z_V, &uu, &ww);
            su[ip] = uu; sw[ip] = ww;
            x[ip] += uu * dt;
            z[ip] += ww * dt;
            if (x[ip] < 0 || x[ip] > Lx || z[ip] < 0 || z[ip] > Lz) active[ip] = 0;
        } else if (rk == 2) {
            double xA = x[ip], zA = z[ip], uA, wA, uB, wB;
            bad += !vel_on_pt(xA, zA, u, w, hx, hz, nelx, nel, icon, x_V, z_V, &uA, &wA);
            double xB = xA + uA * dt / 2.0, yB = zA + wA * dt / 2.0;
            if (xB < 0 || xB > Lx || yB < 0 || yB > Lz) { active[ip] = 0; }
            else {
                bad += !vel_on_pt(xB, yB, u, w, hx, hz, nelx, nel, icon, x_V, z_V, &uB, &wB);
                x[ip] = xA + uB * dt;
                z[ip] = zA + wB * dt;
                su[ip] = uB; sw[ip] = wB;
            }
        } else if (rk == 4) {
            double xA = x[ip], zA = z[ip], uA, wA, uB, wB, uC, wC, uD, wD;
            bad += !vel_on_pt(xA, zA, u, w, hx, hz, nelx, nel, icon, x_V, z_V, &uA, &wA);
            double xB = xA + uA * dt / 2.0, yB = zA + wA * dt / 2.0;
            if (xB < 0 || xB > Lx || yB < 0 || yB > Lz) { active[ip] = 0; continue; }
            bad += !vel_on_pt(xB, yB, u, w, hx, hz, nelx, nel, icon, x_V, z_V, &uB, &wB);
            double xC = xA + uB * dt / 2.0, yC = zA + wB * dt / 2.0;
            if (xC < 0 || xC > Lx || yC < 0 || yC > Lz) { active[ip] = 0; continue; }
            bad += !vel_on_pt(xC, yC, u, w, hx, hz, nelx, nel, icon, x_V, z_V, &uC, &wC);
            double xD = xA + uC * dt, yD = zA + wC * dt;
            if (xD < 0 || xD > Lx || yD < 0 || yD > Lz) { active[ip] = 0; continue; }
            bad += !vel_on_pt(xD, yD, u, w, hx, hz, nelx, nel, icon, x_V, z_V, &uD, &wD);
            su[ip] = (uA + 2 * uB + 2 * uC + uD) / 6;
            sw[ip] = (wA + 2 * wB + 2 * wC + wD) / 6;
            x[ip] = xA + su[ip] * dt;
            z[ip] = zA + sw[ip] * dt;
        }
    }
    for (int ip = 0; ip < np_; ip++)
        if (!active[ip]) { x[ip] = 0; z[ip] = 0; }
    return bad;
}

/* averaging codes: 0 arithmetic, 1 geometric, 2 harmonic */

/* project_particles_on_elements (pic_functions.py:602-650) */
void orc_project_on_elements(int nel, int np_, const uint8_t *active,
                             const double *rho, const double *eta,
                             const int *iel, int averaging, double *rho_e,
                             double *eta_e, int *cnt)
{
    for (int e = 0; e < nel; e++) { rho_e[e] = 0; eta_e[e] = 0; cnt[e] = 0; }
    for (int ip = 0; ip < np_; ip++)
        if (active[ip]) { rho_e[iel[ip]] += rho[ip]; cnt[iel[ip]] += 1; }
    for (int e = 0; e < nel; e++) rho_e[e] /= (double)cnt[e];
    for (int ip = 0; ip < np_; ip++) {
        if (!active[ip]) continue;
        if (averaging == 0) eta_e[iel[ip]] += eta[ip];
        else if (averaging == 1) eta_e[iel[ip]] += log10(eta[ip]);
        else eta_e[iel[ip]] += 1.0 / eta[ip];
    }
    for (int e = 0; e < nel; e++) {
        if (averaging == 0) eta_e[e] /= (double)cnt[e];
        else if (averaging == 1) eta_e[e] = pow(10.0, eta_e[e] / (double)cnt[e]);
        else eta_e[e] = (double)cnt[e] / eta_e[e];
    }
}

/* project_particle_field_on_nodes_{1,2,3,4} (pic_functions.py:658-1053) */
void orc_project_on_nodes(int type, int nel, int nn_V, int np_,
                          const uint8_t *active, const double *phi,
                          const int *icon, const int *iel, const double *r,
                          const double *t, int averaging, double *out)
{
    double *cnt = (double *)calloc(nn_V, sizeof(double));
    for (int n = 0; n < nn_V; n++) out[n] = 0.0;
    for (int ip = 0; ip < np_; ip++) {
        if (!active[ip]) continue;
        const int e = iel[ip];
        double v = phi[ip];
        if (averaging == 1) v = log10(phi[ip]);
        else if (averaging == 2) v = 1.0 / phi[ip];
        const double rr = r[ip], tt = t[ip];
        double N[4];
        int use[4] = {1, 1, 1, 1};
        switch (type) {
        case 1:
            N[0] = N[1] = N[2] = N[3] = 1.0;
            break;
        case 2:
            N[0] = 0.25 * (1 - rr) * (1 - tt);
            N[1] = 0.25 * (1 + rr) * (1 - tt);
            N[2] = 0.25 * (1 + rr) * (1 + tt);
            N[3] = 0.25 * (1 - rr) * (1 + tt);
            break;
        case 3:
            N[0] = N[1] = N[2] = N[3] = 1.0;
            use[0] = (rr <= 0 && tt <= 0); use[1] = (rr >= 0 && tt <= 0);
            use[2] = (rr >= 0 && tt >= 0); use[3] = (rr <= 0 && tt >= 0);
            break;
        default:
            use[0] = (rr <= 0 && tt <= 0); use[1] = (rr >= 0 && tt <= 0);
            use[2] = (rr >= 0 && tt >= 0); use[3] = (rr <= 0 && tt >= 0);
            N[0] = -rr * -tt;
            N[1] = (1 - rr) * -tt;
            N[2] = (1 - rr) * (1 - tt);
            N[3] = -rr * (1 - tt);
            break;
        }
        for (int k = 0; k < 4; k++) {
            if (!use[k]) continue;
            int n = icon[k * nel + e];
            if (type == 1 || type == 3) { out[n] += v; cnt[n] += 1; }
            else { out[n] += v * N[k]; cnt[n] += 1 * N[k]; }
        }
    }
    for (int n = 0; n < nn_V; n++) {
        if (averaging == 0) out[n] = out[n] / cnt[n];
        else if (averaging == 1) out[n] = pow(10.0, out[n] / cnt[n]);
        else out[n] = cnt[n] / out[n];
    }
    for (int e = 0; e < nel; e++) {
        const int *ic = icon;
        int n0 = ic[0 * nel + e], n1 = ic[1 * nel + e], n2 = ic[2 * nel + e], n3 = ic[3 * nel + e];
        out[ic[4 * nel + e]] = 0.50 * (out[n0] + out[n1]);
        out[ic[5 * nel + e]] = 0.50 * (out[n1] + out[n2]);
        out[ic[6 * nel + e]] = 0.50 * (out[n2] + out[n3]);
        out[ic[7 * nel + e]] = 0.50 * (out[n0] + out[n3]);
        out[ic[8 * nel + e]] = 0.25 * (out[n0] + out[n1] + out[n2] + out[n3]);
    }
    free(cnt);
}

/* ---------------------------------------------------------------------- */
/* nodal -> qpoint: project_nodal_field_onto_qpoints.py:11-62              */
/* m = 4 (P variant, any icon with >= 4 rows) or 9 (V/T variants)          */
/* ---------------------------------------------------------------------- */
void orc_nodal_to_qpoints(const double *phi, int nel, int m, const double *N,
                          const int *icon, double *phiq)
{
    for (int iel = 0; iel < nel; iel++)
        for (int iq = 0; iq < 9; iq++) {
            double s = 0.0;
            for (int k = 0; k < m; k++) s += N[iq * m + k] * phi[icon[k * nel + iel]];
            phiq[iel * 9 + iq] = s;
        }
}

/* ---------------------------------------------------------------------- */
/* compute_pressure_offset.py:29-61                                        */
/* mode 0 none, 1 surface, 2 volume; axi: the axisymmetric branches (33-39, */
/* 49-55: the surface integral of p sin(colatitude) over the top edges and  */
/* the 2 pi x weighted volume integral, neither divided by anything)        */
/* ---------------------------------------------------------------------- */
double orc_pressure_offset(int mode, const uint8_t *top_element, int nel,
                           int nelx, const double *N_P, const double *JxWq,
                           const double *p, const int *icon_P, double volume,
                           int axi, const double *theta_P, const double *xq)
{
    double off = 0.0;
    if (axi && mode == 1) {
        for (int e = 0; e < nel; e++)
            if (top_element[e]) {
                const int n2 = icon_P[2 * nel + e], n3 = icon_P[3 * nel + e];
                const double pmean = 0.5 * (p[n2] + p[n3]);
                const double theta_mean = ORC_PI / 2 - 0.5 * (theta_P[n2] + theta_P[n3]);
                const double dtheta = theta_P[n3] - theta_P[n2];
                off += 0.5 * pmean * sin(theta_mean) * dtheta;
            }
    } else if (axi && mode == 2) {
        for (int e = 0; e < nel; e++)
            for (int iq = 0; iq < 9; iq++) {
                double s = 0.0;
                for (int k = 0; k < 4; k++) s += N_P[iq * 4 + k] * p[icon_P[k * nel + e]];
                off += s * JxWq[e * 9 + iq] * 2 * ORC_PI * xq[e * 9 + iq];
            }
    } else if (mode == 1) {
        for (int e = 0; e < nel; e++)
            if (top_element[e]) off += 0.5 * (p[icon_P[2 * nel + e]] + p[icon_P[3 * nel + e]]);
        off /= nelx;
    } else if (mode == 2) {
        for (int e = 0; e < nel; e++)
            for (int iq = 0; iq < 9; iq++) {
                double s = 0.0;
                for (int k = 0; k < 4; k++) s += N_P[iq * 4 + k] * p[icon_P[k * nel + e]];
                off += s * JxWq[e * 9 + iq];
            }
        off /= volume;
    }
    return off;
}

/* compute_nodal_pressure.py:7-21 */
void orc_nodal_pressure(int nel, int nn_V, const int *icon_V, const int *icon_P,
                        const double *p, const double *N_P_n, double *q)
{
    int *cnt = (int *)calloc(nn_V, sizeof(int));
    for (int n = 0; n < nn_V; n++) q[n] = 0.0;
    for (int e = 0; e < nel; e++)
        for (int k = 0; k < 9; k++) {
            double s = 0.0;
            for (int j = 0; j < 4; j++) s += N_P_n[k * 4 + j] * p[icon_P[j * nel + e]];
            q[icon_V[k * nel + e]] += s;
            cnt[icon_V[k * nel + e]] += 1;
        }
    for (int n = 0; n < nn_V; n++) q[n] /= (double)cnt[n];
    free(cnt);
}

/* compute_nodal_heat_flux.py:22-38 */
void orc_nodal_heat_flux(const int *icon_T, const double *T, const double *hcond_n,
                         int nn_T, int nel, const double *dNdr_n,
                         const double *dNds_n, const double *j00, const double *j01,
                         const double *j10, const double *j11, double *dTdx,
                         double *dTdy, double *qx, double *qy)
{
    double *cnt = (double *)calloc(nn_T, sizeof(double));
    for (int n = 0; n < nn_T; n++) dTdx[n] = dTdy[n] = qx[n] = qy[n] = 0.0;
    for (int e = 0; e < nel; e++) {
        double Tv[9];
        for (int k = 0; k < 9; k++) Tv[k] = T[icon_T[k * nel + e]];
        for (int i = 0; i < 9; i++) {
            int n = icon_T[i * nel + e];
            double dNdx[9], dNdy[9];
            for (int k = 0; k < 9; k++) {
                dNdx[k] = j00[e * 9 + i] * dNdr_n[i * 9 + k] + j01[e * 9 + i] * dNds_n[i * 9 + k];
                dNdy[k] = j10[e * 9 + i] * dNdr_n[i * 9 + k] + j11[e * 9 + i] * dNds_n[i * 9 + k];
            }
            double gx = dot9(dNdx, Tv), gy = dot9(dNdy, Tv);
            dTdx[n] -= gx;
            dTdy[n] -= gy;
            qx[n] -= hcond_n[n] * gx;
            qy[n] -= hcond_n[n] * gy;
            cnt[n] += 1;
        }
    }
    for (int n = 0; n < nn_T; n++) {
        qx[n] /= cnt[n]; qy[n] /= cnt[n]; dTdx[n] /= cnt[n]; dTdy[n] /= cnt[n];
    }
    free(cnt);
}

/* compute_global_quantities (postprocessors.py:13-83).
 * out = [vrms, EK, WAG, TVD, GPE, ITE, TM, Tavrg, eta_avrg] */
void orc_global_quantities(int nel, const double *uq, const double *wq,
                           const double *Tq, const double *rhoq,
                           const double *hcapaq, const double *etaq,
                           const double *exxq, const double *ezzq,
                           const double *exzq, double volume, const double *JxWq,
                           const double *gxq, const double *gzq, double *out)
{
    double TM = 0, EK = 0, WAG = 0, TVD = 0, ITE = 0, vrms = 0, Tavrg = 0, num = 0, den = 0;
    for (int e = 0; e < nel * 9; e++) {
        double exx = exxq[e], ezz = ezzq[e], exz = exzq[e], jw = JxWq[e];
        double eII = exx * exx + ezz * ezz + 2 * exz * exz;
        double u2 = uq[e] * uq[e] + wq[e] * wq[e];
        TM += rhoq[e] * jw;
        EK += 0.5 * rhoq[e] * u2 * jw;
        WAG -= rhoq[e] * (uq[e] * gxq[e] + wq[e] * gzq[e]) * jw;
        TVD += 2 * etaq[e] * (2 * exx * exx / 3 + 2 * ezz * ezz / 3 - 2 * exx * ezz / 3 + 2 * exz * exz) * jw;
        ITE += rhoq[e] * hcapaq[e] * Tq[e] * jw;
        vrms += u2 * jw;
        Tavrg += Tq[e] * jw;
        num += etaq[e] * eII * jw;
        den += eII * jw;
    }
    out[0] = sqrt(vrms / volume);
    out[1] = EK; out[2] = WAG; out[3] = TVD; out[4] = 0.0; out[5] = ITE; out[6] = TM;
    out[7] = Tavrg / volume;
    out[8] = den > 0 ? num / den : 0.0;
}

/* compute_Nu (postprocessors.py:152-248), box only.
 * out = [T_left, T_right, T_bottom, T_top, dTdx_left, dTdx_right,
 *        dTdz_bottom, dTdz_top, qx_left, qx_right, qz_bottom, qz_top, Nu] */
void orc_compute_Nu(double Lx, double Lz, const double *x_V, const double *z_V,
                    int nel, const uint8_t *left, const uint8_t *right,
                    const uint8_t *top, const uint8_t *bottom, const int *icon,
                    const double *T, const double *dTdx, const double *dTdz,
                    const double *qc, const double *qw, const double *hcond_n,
                    double *out)
{
    double o[12] = {0};
    for (int e = 0; e < nel; e++) {
        double Tv[9], hv[9], gx[9], gz[9];
        for (int k = 0; k < 9; k++) {
            int n = icon[k * nel + e];
            Tv[k] = T[n]; hv[k] = hcond_n[n]; gx[k] = dTdx[n]; gz[k] = dTdz[n];
        }
        double N[9];
        if (left[e]) {
            double jc = (z_V[icon[3 * nel + e]] - z_V[icon[0 * nel + e]]) / 2;
            for (int iq = 0; iq < 3; iq++) {
                orc_basis_V(-1, qc[iq], N);
                double Tq = dot9(N, Tv), hq = dot9(N, hv), dq = dot9(N, gx);
                o[0] += Tq * jc * qw[iq];
                o[4] += dq * jc * qw[iq] * -1;
                o[8] -= hq * dq * jc * qw[iq] * -1;
            }
        }
        if (right[e]) {
            double jc = (z_V[icon[2 * nel + e]] - z_V[icon[1 * nel + e]]) / 2;
            for (int iq = 0; iq < 3; iq++) {
                orc_basis_V(+1, qc[iq], N);
                double Tq = dot9(N, Tv), hq = dot9(N, hv), dq = dot9(N, gx);
                o[1] += Tq * jc * qw[iq];
                o[5] += dq * jc * qw[iq] * +1;
                o[9] -= hq * dq * jc * qw[iq] * +1;
            }
        }
        if (top[e]) {
            double jc = (x_V[icon[2 * nel + e]] - x_V[icon[3 * nel + e]]) / 2;
            for (int iq = 0; iq < 3; iq++) {
                orc_basis_V(qc[iq], +1, N);
                double Tq = dot9(N, Tv), hq = dot9(N, hv), dq = dot9(N, gz);
                o[3] += Tq * jc * qw[iq];
                o[7] += dq * jc * qw[iq] * +1;
                o[11] -= hq * dq * jc * qw[iq] * +1;
            }
        }
        if (bottom[e]) {
            double jc = (x_V[icon[1 * nel + e]] - x_V[icon[0 * nel + e]]) / 2;
            for (int iq = 0; iq < 3; iq++) {
                orc_basis_V(qc[iq], -1, N);
                double Tq = dot9(N, Tv), hq = dot9(N, hv), dq = dot9(N, gz);
                o[2] += Tq * jc * qw[iq];
                o[6] += dq * jc * qw[iq] * -1;
                o[10] -= hq * dq * jc * qw[iq] * -1;
            }
        }
    }
    o[0] /= Lz; o[1] /= Lz; o[2] /= Lx; o[3] /= Lx;
    o[4] /= Lz; o[5] /= Lz; o[6] /= Lx; o[7] /= Lx;
    for (int k = 0; k < 12; k++) out[k] = o[k];
    out[12] = fabs(o[7]) / (o[2] - o[3]) * Lz;
}

/* particle positions from reduced coordinates (swarm_coordinates_setup.py:71-86) */
void orc_place_particles(int nel, int ppe, const double *r, const double *t,
                         const double *x_V, const double *z_V, const int *icon,
                         double *x, double *z, int *iel)
{
    for (int e = 0; e < nel; e++)
        for (int k = 0; k < ppe; k++) {
            long ip = (long)e * ppe + k;
            double N[9], xv[9], zv[9];
            orc_basis_V(r[ip], t[ip], N);
            for (int m = 0; m < 9; m++) {
                xv[m] = x_V[icon[m * nel + e]];
                zv[m] = z_V[icon[m * nel + e]];
            }
            x[ip] = dot9(N, xv);
            z[ip] = dot9(N, zv);
            iel[ip] = e;
        }
}

/* ------------------------------------------------------------------------
 * Curved geometries (quarter, half, eighth, annulus) and least squares.
 * ------------------------------------------------------------------------ */

/* np.pi / 2 as a double */
#define HALF_PI 1.5707963267948966

/* locate_particles___annulus (pic_functions.py:142-173); returns the number of
 * active particles whose element index leaves the mesh (reference: no check). */
int orc_locate_annulus(int np_, const uint8_t *active, const double *rad,
                       const double *theta, double hrad, double htheta,
                       const double *rad_V, const double *theta_V,
                       const int *icon_V, int nel, int nelx, double Rinner,
                       double *r, double *t, int *iel_out)
{
    int bad = 0;
    for (int ip = 0; ip < np_; ip++) {
        r[ip] = 0.0; t[ip] = 0.0; iel_out[ip] = 0;
        if (!active[ip]) continue;
        int ielx = (int)((HALF_PI - theta[ip]) / htheta);
        int ielz = (int)((rad[ip] - Rinner) / hrad);
        int iel = nelx * ielz + ielx;
        if (iel < 0 || iel >= nel) { bad++; iel_out[ip] = iel; r[ip] = NAN; t[ip] = NAN; continue; }
        r[ip] = ((-theta[ip] + theta_V[icon_V[iel]]) / htheta - 0.5) * 2;
        t[ip] = ((rad[ip] - rad_V[icon_V[iel]]) / hrad - 0.5) * 2;
        iel_out[ip] = iel;
    }
    return bad;
}

/* interpolate_vel_on_pt___annulus (pic_functions.py:40-63); returns 0 when
 * the element index leaves the mesh */
int orc_vel_on_pt_annulus(double xp, double yp, const double *u, const double *w,
                          double hrad, double htheta, int nelx, int nel,
                          const int *icon, const double *rad_V,
                          const double *theta_V, double Rinner, double *up,
                          double *wp, int *iel_out)
{
    double radp = sqrt(xp * xp + yp * yp);
    double thetap = HALF_PI - atan2(xp, yp);
    int ielx = (int)((HALF_PI - thetap) / htheta);
    int ielz = (int)((radp - Rinner) / hrad);
    int iel = nelx * ielz + ielx;
    *iel_out = iel;
    if (iel < 0 || iel >= nel) { *up = NAN; *wp = NAN; return 0; }
    double rp = ((-thetap + theta_V[icon[iel]]) / htheta - 0.5) * 2;
    double tp = ((radp - rad_V[icon[iel]]) / hrad - 0.5) * 2;
    double N[9], uv[9], wv[9];
    orc_basis_V(rp, tp, N);
    for (int k = 0; k < 9; k++) {
        uv[k] = u[icon[k * nel + iel]];
        wv[k] = w[icon[k * nel + iel]];
    }
    *up = dot9(N, uv);
    *wp = dot9(N, wv);
    return 1;
}

/* interpolate_vel_on_pt___box (pic_functions.py:17-37) at many points */
int orc_vel_on_pts_box(int n, const double *xp, const double *zp, const double *u,
                       const double *w, double hx, double hz, int nelx, int nel,
                       const int *icon, const double *x_V, const double *z_V,
                       double *up, double *wp, int *iel)
{
    int bad = 0;
    for (int i = 0; i < n; i++) {
        int ielx = (int)(xp[i] / hx), ielz = (int)(zp[i] / hz);
        iel[i] = nelx * ielz + ielx;
        bad += !vel_on_pt(xp[i], zp[i], u, w, hx, hz, nelx, nel, icon, x_V, z_V, up + i, wp + i);
    }
    return bad;
}

/* deactivation test of each curved variant after a move */
static int curved_outside(int variant, double x, double z, double rad, double Rinner,
                          double Router)
{
    switch (variant) {
    case 0: return x < 0 || z < 0 || rad < Rinner || rad > Router;       /* quarter */
    case 1: return x < 0 || rad < Rinner || rad > Router;                /* half */
    default: return x < 0 || rad < Rinner || rad > Router || z < x;      /* eighth, annulus */
    }
}

/* advect_particles___quarter/half/eighth/annulus (pic_functions.py:298-596);
 * variant 0 quarter, 1 half, 2 eighth, 3 annulus.  RK1 for all, RK2 for
 * quarter and half; other orders leave the particles where they are (as the
 * reference).  Mutates x, z, rad, theta, active. */
int orc_advect_curved(int variant, int rk, double dt, int np_, double *x, double *z,
                      double *rad, double *theta, uint8_t *active, const double *u,
                      const double *w, double Rinner, double Router, double hrad,
                      double htheta, int nelx, int nel, const int *icon,
                      const double *rad_V, const double *theta_V, double *su,
                      double *sw)
{
    int bad = 0, iel;
    for (int ip = 0; ip < np_; ip++) {
        su[ip] = 0.0; sw[ip] = 0.0;
        if (!active[ip]) continue;
        if (rk == 1) {
            double uu, ww;
            bad += !orc_vel_on_pt_annulus(x[ip], z[ip], u, w, hrad, htheta, nelx, nel, icon,
                                          rad_V, theta_V, Rinner, &uu, &ww, &iel);
            su[ip] = uu; sw[ip] = ww;
            x[ip] += uu * dt;
            z[ip] += ww * dt;
            rad[ip] = sqrt(x[ip] * x[ip] + z[ip] * z[ip]);
            theta[ip] = HALF_PI - atan2(x[ip], z[ip]);
            if (curved_outside(variant, x[ip], z[ip], rad[ip], Rinner, Router)) active[ip] = 0;
        } else if (rk == 2 && variant <= 1) {
            double xA = x[ip], zA = z[ip], uA, wA, uB, wB;
            bad += !orc_vel_on_pt_annulus(xA, zA, u, w, hrad, htheta, nelx, nel, icon, rad_V,
                                          theta_V, Rinner, &uA, &wA, &iel);
            double xB = xA + uA * dt / 2.0, yB = zA + wA * dt / 2.0;
            double rB = sqrt(xA * xA + zA * zA);  /* the reference measures A, not B */
            int outB = variant == 0 ? (xB < 0 || yB < 0 || rB < Rinner || rB > Router)
                                    : (xB < 0 || rB < Rinner || rB > Router);
            if (outB) { active[ip] = 0; continue; }
            bad += !orc_vel_on_pt_annulus(xB, yB, u, w, hrad, htheta, nelx, nel, icon, rad_V,
                                          theta_V, Rinner, &uB, &wB, &iel);
            x[ip] = xA + uB * dt;
            z[ip] = zA + wB * dt;
            su[ip] = uB; sw[ip] = wB;
            rad[ip] = sqrt(x[ip] * x[ip] + z[ip] * z[ip]);
            theta[ip] = HALF_PI - atan2(x[ip], z[ip]);
            if (curved_outside(variant, x[ip], z[ip], rad[ip], Rinner, Router)) active[ip] = 0;
        }
    }
    for (int ip = 0; ip < np_; ip++)
        if (!active[ip]) { x[ip] = 0; z[ip] = 0; }
    return bad;
}

/* Gaussian elimination with partial pivoting (the LU of LAPACK gesv, which
 * np.linalg.solve calls; equal up to rounding), n <= 4 */
static void small_solve(int n, double A[4][4], double *b)
{
    for (int k = 0; k < n; k++) {
        int p = k;
        for (int i = k + 1; i < n; i++)
            if (fabs(A[i][k]) > fabs(A[p][k])) p = i;
        if (p != k) {
            for (int j = 0; j < n; j++) { double s = A[k][j]; A[k][j] = A[p][j]; A[p][j] = s; }
            double s = b[k]; b[k] = b[p]; b[p] = s;
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

/* compute_ls_coefficients_P1 (nb = 3) and _Q1 (nb = 4) (pic_functions.py:1097-1285):
 * every particle with swarm_iel == iel counts (no activity test, as the
 * reference), in increasing particle order.  coef_rho/eta: nel x nb;
 * mm: nel x 4 (rho min, rho max, eta min, eta max). */
void orc_ls_coefficients(int nb, int nel, const double *x_e, const double *z_e, int np_,
                         const double *x, const double *z, const int *iel,
                         const double *rho, const double *eta, double *coef_rho,
                         double *coef_eta, double *mm)
{
    for (int e = 0; e < nel; e++) {
        double A[4][4] = {{0}}, br[4] = {0}, be[4] = {0};
        double rmin = INFINITY, rmax = -INFINITY, emin = INFINITY, emax = -INFINITY;
        int cnt = 0;
        for (int ip = 0; ip < np_; ip++) {
            if (iel[ip] != e) continue;
            double sx = x[ip] - x_e[e], sz = z[ip] - z_e[e];
            cnt++;
            A[0][1] += sx; A[0][2] += sz;
            A[1][1] += sx * sx; A[1][2] += sx * sz; A[2][2] += sz * sz;
            if (nb == 4) {
                A[0][3] += sx * sz; A[1][3] += sx * sx * sz;
                A[2][3] += sx * sz * sz; A[3][3] += sx * sx * sz * sz;
            }
            be[0] += eta[ip]; be[1] += eta[ip] * sx; be[2] += eta[ip] * sz;
            br[0] += rho[ip]; br[1] += rho[ip] * sx; br[2] += rho[ip] * sz;
            if (nb == 4) { be[3] += eta[ip] * sx * sz; br[3] += rho[ip] * sx * sz; }
            if (rho[ip] < rmin) rmin = rho[ip];
            if (rho[ip] > rmax) rmax = rho[ip];
            if (eta[ip] < emin) emin = eta[ip];
            if (eta[ip] > emax) emax = eta[ip];
        }
        A[0][0] = cnt;
        for (int i = 0; i < nb; i++)
            for (int j = 0; j < i; j++) A[i][j] = A[j][i];
        double A2[4][4];
        for (int i = 0; i < 4; i++) for (int j = 0; j < 4; j++) A2[i][j] = A[i][j];
        small_solve(nb, A, be);
        small_solve(nb, A2, br);
        for (int k = 0; k < nb; k++) { coef_eta[e * nb + k] = be[k]; coef_rho[e * nb + k] = br[k]; }
        mm[4 * e + 0] = rmin; mm[4 * e + 1] = rmax; mm[4 * e + 2] = emin; mm[4 * e + 3] = emax;
    }
}

/* limiter (pic_functions.py:1067-1087) over arrays */
void orc_limiter(int n, const double *c0, double *c1, double *c2, const double *S,
                 const double *L, double h)
{
    for (int i = 0; i < n; i++) {
        double ku = fmin(c0[i] - S[i], L[i] - c0[i]);
        double kl = (fabs(c1[i]) + fabs(c2[i])) * h / 2;
        if (kl > ku) {
            double cmax = ku * 2 / h;
            c1[i] = (c1[i] < 0 ? -1.0 : 1.0) * fmin(fabs(c1[i]), cmax);
            c2[i] = (c2[i] < 0 ? -1.0 : 1.0) * fmin(fabs(c2[i]), cmax);
        }
        kl = (fabs(c1[i]) + fabs(c2[i])) * h / 2;
        if (kl > ku) {
            double cc = (kl - ku) / h;
            c1[i] = (c1[i] < 0 ? -1.0 : 1.0) * (fabs(c1[i]) - cc);
            c2[i] = (c2[i] < 0 ? -1.0 : 1.0) * (fabs(c2[i]) - cc);
        }
    }
}

/* ------------------------------------------------------------------------
 * Derived nodal and elemental fields
 * ------------------------------------------------------------------------ */

/* compute_nodal_strain_rate (compute_strain_rate.py:50-73, mode 0: u, v ->
 * exx, eyy, exy) and compute_nodal_pressure_gradient
 * (compute_nodal_pressure_gradient.py:11-31, mode 1: q -> dpdx, dpdy): the
 * derivatives of each element's field at its own nodes, summed per node in
 * element order and divided by the number of contributions */
void orc_nodal_gradients(int mode, const int *icon, const double *u, const double *v, int nn,
                         int nel, const double *dNdr_n, const double *dNds_n, const double *j00,
                         const double *j01, const double *j10, const double *j11, double *o0,
                         double *o1, double *o2)
{
    double *cnt = (double *)calloc(nn, sizeof(double));
    for (int n = 0; n < nn; n++) { o0[n] = 0; o1[n] = 0; if (o2) o2[n] = 0; }
    for (int e = 0; e < nel; e++) {
        double uv[9], vv[9];
        for (int k = 0; k < 9; k++) {
            uv[k] = u[icon[k * nel + e]];
            vv[k] = mode == 0 ? v[icon[k * nel + e]] : 0.0;
        }
        for (int i = 0; i < 9; i++) {
            int n = icon[i * nel + e];
            double dx[9], dy[9];
            for (int k = 0; k < 9; k++) {
                dx[k] = j00[e * 9 + i] * dNdr_n[i * 9 + k] + j01[e * 9 + i] * dNds_n[i * 9 + k];
                dy[k] = j10[e * 9 + i] * dNdr_n[i * 9 + k] + j11[e * 9 + i] * dNds_n[i * 9 + k];
            }
            if (mode == 0) {
                o0[n] += dot9(dx, uv);
                o1[n] += dot9(dy, vv);
                o2[n] += 0.5 * dot9(dx, vv) + 0.5 * dot9(dy, uv);
            } else {
                o0[n] += dot9(dx, uv);
                o1[n] += dot9(dy, uv);
            }
            cnt[n] += 1;
        }
    }
    for (int n = 0; n < nn; n++) {
        o0[n] /= cnt[n];
        o1[n] /= cnt[n];
        if (o2) o2[n] /= cnt[n];
    }
    free(cnt);
}

/* compute_elemental_strain_rate (compute_strain_rate.py:17-42): at the element
 * centre; the 2x2 Jacobian inverted in closed form (np.linalg.inv up to rounding) */
void orc_elemental_strain_rate(const int *icon, const double *u, const double *v, int nel,
                               const double *x_V, const double *y_V, double *exx, double *eyy,
                               double *exy)
{
    double dNdr[9], dNdt[9];
    orc_basis_V_dr(0.0, 0.0, dNdr);
    orc_basis_V_dt(0.0, 0.0, dNdt);
    for (int e = 0; e < nel; e++) {
        double xe[9], ye[9], ue[9], ve[9];
        for (int k = 0; k < 9; k++) {
            int n = icon[k * nel + e];
            xe[k] = x_V[n]; ye[k] = y_V[n]; ue[k] = u[n]; ve[k] = v[n];
        }
        double a = dot9(dNdr, xe), b = dot9(dNdr, ye), c = dot9(dNdt, xe), d = dot9(dNdt, ye);
        double det = a * d - b * c;
        double i00 = d / det, i01 = -b / det, i10 = -c / det, i11 = a / det;
        double dx[9], dy[9];
        for (int k = 0; k < 9; k++) {
            dx[k] = i00 * dNdr[k] + i01 * dNdt[k];
            dy[k] = i10 * dNdr[k] + i11 * dNdt[k];
        }
        exx[e] = dot9(dx, ue);
        eyy[e] = dot9(dy, ve);
        exy[e] = dot9(dx, ve) * 0.5 + dot9(dy, ue) * 0.5;
    }
}
