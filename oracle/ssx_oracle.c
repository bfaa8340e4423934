/*
 * ssx_oracle.c -- CPU restatement of the reference's Mohr-Coulomb constitutive operator and
 * the fixed-pattern assembly kernels around it.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product package links, loads or calls this file;
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, as the checker.
 *
 * Every routine states the reference file:line it follows (paths relative to the reference
 * checkout).  Layouts follow the reference's C kernels: per-point arrays are point-major
 * (E[p*6+i], DS[p*36+k], k = i + 6*j column-major 6x6), dphi is (n_int, n_p) per component,
 * element DOF tables are (n_elem, 3*n_p) int32.
 *
 * Pinning (what checks each routine beyond the compiled reference):
 *   mc3d_point      closed forms (elastic, apex), yield-surface consistency on the smooth face
 *                   and the edges, DS = dS/dE by central differences, objectivity, and the
 *                   reference's own C compiled from source (oracle/_ref) on stratified inputs.
 *   strain/force    brute-force B built from the reference's strain-operator rule.
 *   tangent rows    element-wise dense B^T D B (the reference's legacy kernel rule).
 */
#include <math.h>
#include <stdint.h>
#include <stddef.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* Eigenprojection of the strain for eigenvalue e_k with the other two eigenvalues a, b:
 *   P_k = (eps^2 - (a+b) eps + a b I) / ((e_k-a)(e_k-b)),   Voigt tensor form.
 * (constitutive_3D_kernel.h:19-28) */
static void proj(const double esq[6], const double et[6], double a_plus_b, double a_times_b,
                 double inv_den, double P[6])
{
    for (int i = 0; i < 6; ++i) {
        double v = esq[i] - a_plus_b * et[i];
        if (i < 3) v += a_times_b;
        P[i] = inv_den * v;
    }
}

/* d(eps^2)/d(E) in the engineering-shear column convention (constitutive_3D_kernel.h:30-57).
 * Entry (i,j) with i = row of eps^2 (Voigt, tensor shear) and j = strain column. */
static void dsq(const double et[6], double d[36])
{
    memset(d, 0, sizeof(double) * 36);
#define D(i, j) d[(i) + 6 * (j)]
    D(0, 0) = 2.0 * et[0]; D(3, 0) = et[3]; D(5, 0) = et[5];
    D(1, 1) = 2.0 * et[1]; D(3, 1) = et[3]; D(4, 1) = et[4];
    D(2, 2) = 2.0 * et[2]; D(4, 2) = et[4]; D(5, 2) = et[5];
    D(0, 3) = et[3]; D(1, 3) = et[3];
    D(3, 3) = 0.5 * (et[0] + et[1]); D(4, 3) = 0.5 * et[5]; D(5, 3) = 0.5 * et[4];
    D(1, 4) = et[4]; D(2, 4) = et[4];
    D(3, 4) = 0.5 * et[5]; D(4, 4) = 0.5 * (et[1] + et[2]); D(5, 4) = 0.5 * et[3];
    D(0, 5) = et[5]; D(2, 5) = et[5];
    D(3, 5) = 0.5 * et[4]; D(4, 5) = 0.5 * et[3]; D(5, 5) = 0.5 * (et[0] + et[2]);
#undef D
}

static double vi(int i) { return i < 3 ? 1.0 : 0.0; }
static double idf(int i, int j) { return i != j ? 0.0 : (i < 3 ? 1.0 : 0.5); }

/* Return codes: 0 elastic, 1 smooth face, 2 left edge, 3 right edge, 4 apex. */
int ssx_oracle_mc3d_point(const double E[6], double c_bar, double sin_phi, double G, double lam,
                          double S[6], double *DS)
{
    /* Associated flow: sin(psi) = sin(phi) (constitutive_3D_kernel.h:73). */
    const double sp = sin_phi, sq = sin_phi;
    double et[6] = {E[0], E[1], E[2], 0.5 * E[3], 0.5 * E[4], 0.5 * E[5]};
    double esq[6];
    /* eps^2 in Voigt order (11,22,33,12,23,13) (constitutive_3D_kernel.h:78-84). */
    esq[0] = et[0] * et[0] + et[3] * et[3] + et[5] * et[5];
    esq[1] = et[1] * et[1] + et[3] * et[3] + et[4] * et[4];
    esq[2] = et[2] * et[2] + et[4] * et[4] + et[5] * et[5];
    esq[3] = et[0] * et[3] + et[1] * et[3] + et[4] * et[5];
    esq[4] = et[3] * et[5] + et[1] * et[4] + et[2] * et[4];
    esq[5] = et[0] * et[5] + et[3] * et[4] + et[2] * et[5];

    /* Invariants and Cardano's trigonometric eigenvalues (constitutive_3D_kernel.h:86-112). */
    const double I1 = et[0] + et[1] + et[2];
    const double I2 = et[0] * et[1] + et[0] * et[2] + et[1] * et[2]
                    - et[3] * et[3] - et[4] * et[4] - et[5] * et[5];
    const double I3 = et[0] * et[1] * et[2] - et[2] * et[3] * et[3] - et[1] * et[5] * et[5]
                    - et[0] * et[4] * et[4] + 2.0 * et[3] * et[4] * et[5];
    double Q = (I1 * I1 - 3.0 * I2) / 9.0;
    if (Q < 0.0) Q = 0.0;
    const double R = (-2.0 * I1 * I1 * I1 + 9.0 * I1 * I2 - 27.0 * I3) / 54.0;
    const double rQ = sqrt(Q);
    double th = 0.0;
    if (!(Q < 1e-30)) {
        double c = R / (rQ * Q);
        c = c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);
        th = acos(c) / 3.0;
    }
    const double e1 = -2.0 * rQ * cos(th + 2.0 * M_PI / 3.0) + I1 / 3.0;
    const double e2 = -2.0 * rQ * cos(th - 2.0 * M_PI / 3.0) + I1 / 3.0;
    const double e3 = -2.0 * rQ * cos(th) + I1 / 3.0;

    /* Trial yield function (constitutive_3D_kernel.h:114). */
    const double f = 2.0 * G * ((1.0 + sp) * e1 - (1.0 - sp) * e3) + 2.0 * lam * sp * I1 - c_bar;

    if (f <= 0.0) { /* elastic (constitutive_3D_kernel.h:116-142) */
        for (int i = 0; i < 6; ++i) S[i] = (i < 3 ? lam * I1 : 0.0) + 2.0 * G * et[i];
        if (DS) {
            for (int j = 0; j < 6; ++j)
                for (int i = 0; i < 6; ++i)
                    DS[i + 6 * j] = lam * vi(i) * vi(j) + 2.0 * G * idf(i, j);
        }
        return 0;
    }

    const double g_sl = (e1 - e2) / (1.0 + sq);
    const double g_sr = (e2 - e3) / (1.0 - sq);
    const double den_s = 4.0 * lam * sp * sq + 4.0 * G * (1.0 + sp * sq);
    const double l_s = f / den_s;

    if (l_s <= g_sl && l_s <= g_sr) { /* smooth face (constitutive_3D_kernel.h:144-209) */
        const double d1 = 1.0 / ((e1 - e2) * (e1 - e3));
        const double d2 = 1.0 / ((e2 - e1) * (e2 - e3));
        const double d3 = 1.0 / ((e3 - e1) * (e3 - e2));
        double P1[6], P2[6], P3[6];
        proj(esq, et, e2 + e3, e2 * e3, d1, P1);
        proj(esq, et, e1 + e3, e1 * e3, d2, P2);
        proj(esq, et, e1 + e2, e1 * e2, d3, P3);
        const double s1 = lam * I1 + 2.0 * G * e1 - l_s * (2.0 * lam * sq + 2.0 * G * (1.0 + sq));
        const double s2 = lam * I1 + 2.0 * G * e2 - l_s * 2.0 * lam * sq;
        const double s3 = lam * I1 + 2.0 * G * e3 - l_s * (2.0 * lam * sq - 2.0 * G * (1.0 - sq));
        for (int i = 0; i < 6; ++i) S[i] = s1 * P1[i] + s2 * P2[i] + s3 * P3[i];
        if (DS) {
            double dd[36], Dp[6], Dq[6];
            dsq(et, dd);
            for (int i = 0; i < 6; ++i) {
                Dp[i] = 2.0 * G * ((1.0 + sp) * P1[i] - (1.0 - sp) * P3[i]) + 2.0 * lam * sp * vi(i);
                Dq[i] = 2.0 * G * ((1.0 + sq) * P1[i] - (1.0 - sq) * P3[i]) + 2.0 * lam * sq * vi(i);
            }
            for (int j = 0; j < 6; ++j) {
                for (int i = 0; i < 6; ++i) {
                    const int k = i + 6 * j;
                    const double o1 = P1[i] * P1[j], o2 = P2[i] * P2[j], o3 = P3[i] * P3[j];
                    /* derivatives of the three eigenprojections (constitutive_3D_kernel.h:201-203) */
                    const double dP1 = d1 * (dd[k] - idf(i, j) * (e2 + e3) - (2.0 * e1 - e2 - e3) * o1 - (e2 - e3) * (o2 - o3));
                    const double dP2 = d2 * (dd[k] - idf(i, j) * (e1 + e3) - (2.0 * e2 - e1 - e3) * o2 - (e1 - e3) * (o1 - o3));
                    const double dP3 = d3 * (dd[k] - idf(i, j) * (e1 + e2) - (2.0 * e3 - e1 - e2) * o3 - (e1 - e2) * (o1 - o2));
                    DS[k] = s1 * dP1 + s2 * dP2 + s3 * dP3 + lam * vi(i) * vi(j)
                          + 2.0 * G * (o1 + o2 + o3) - Dq[i] * Dp[j] / den_s;
                }
            }
        }
        return 1;
    }

    if (g_sl < g_sr) { /* return to the left edge (constitutive_3D_kernel.h:211-273) */
        const double den_l = 4.0 * lam * sp * sq + G * (1.0 + sp) * (1.0 + sq) + 2.0 * G * (1.0 - sp) * (1.0 - sq);
        const double l_l = (G * ((1.0 + sp) * (e1 + e2) - 2.0 * (1.0 - sp) * e3) + 2.0 * lam * sp * I1 - c_bar) / den_l;
        const double g_la = (e1 + e2 - 2.0 * e3) / (3.0 - sq);
        if (l_l >= g_sl && l_l <= g_la) {
            const double d3 = 1.0 / ((e3 - e1) * (e3 - e2));
            double P3[6], P12[6];
            proj(esq, et, e1 + e2, e1 * e2, d3, P3);
            for (int i = 0; i < 6; ++i) P12[i] = vi(i) - P3[i];
            const double s1 = lam * I1 + G * (e1 + e2) - l_l * (2.0 * lam * sq + G * (1.0 + sq));
            const double s3 = lam * I1 + 2.0 * G * e3 - l_l * (2.0 * lam * sq - 2.0 * G * (1.0 - sq));
            for (int i = 0; i < 6; ++i) S[i] = s1 * P12[i] + s3 * P3[i];
            if (DS) {
                double dd[36], Dp[6], Dq[6];
                dsq(et, dd);
                for (int i = 0; i < 6; ++i) {
                    Dp[i] = G * ((1.0 + sp) * P12[i] - 2.0 * (1.0 - sp) * P3[i]) + 2.0 * lam * sp * vi(i);
                    Dq[i] = G * ((1.0 + sq) * P12[i] - 2.0 * (1.0 - sq) * P3[i]) + 2.0 * lam * sq * vi(i);
                }
                const double s12 = e1 + e2;
                for (int j = 0; j < 6; ++j) {
                    for (int i = 0; i < 6; ++i) {
                        const int k = i + 6 * j;
                        const double dP3 = d3 * (dd[k] - idf(i, j) * s12 - (P12[i] * et[j] + et[i] * P12[j])
                                                 + s12 * P12[i] * P12[j] + (s12 - 2.0 * e3) * P3[i] * P3[j]
                                                 + e3 * (P12[i] * P3[j] + P3[i] * P12[j]));
                        DS[k] = (s3 - s1) * dP3 + lam * vi(i) * vi(j)
                              + G * (P12[i] * P12[j] + 2.0 * P3[i] * P3[j]) - Dq[i] * Dp[j] / den_l;
                    }
                }
            }
            return 2;
        }
    }

    if (g_sl > g_sr) { /* return to the right edge (constitutive_3D_kernel.h:275-338) */
        const double den_r = 4.0 * lam * sp * sq + 2.0 * G * (1.0 + sp) * (1.0 + sq) + G * (1.0 - sp) * (1.0 - sq);
        const double l_r = (G * (2.0 * (1.0 + sp) * e1 - (1.0 - sp) * (e2 + e3)) + 2.0 * lam * sp * I1 - c_bar) / den_r;
        const double g_ra = (2.0 * e1 - e2 - e3) / (3.0 + sq);
        if (l_r >= g_sr && l_r <= g_ra) {
            const double d1 = 1.0 / ((e1 - e2) * (e1 - e3));
            double P1[6], P23[6];
            proj(esq, et, e2 + e3, e2 * e3, d1, P1);
            for (int i = 0; i < 6; ++i) P23[i] = vi(i) - P1[i];
            const double s1 = lam * I1 + 2.0 * G * e1 - l_r * (2.0 * lam * sq + 2.0 * G * (1.0 + sq));
            const double s3 = lam * I1 + G * (e2 + e3) - l_r * (2.0 * lam * sq - G * (1.0 - sq));
            for (int i = 0; i < 6; ++i) S[i] = s1 * P1[i] + s3 * P23[i];
            if (DS) {
                double dd[36], Dp[6], Dq[6];
                dsq(et, dd);
                for (int i = 0; i < 6; ++i) {
                    Dp[i] = G * (2.0 * (1.0 + sp) * P1[i] - (1.0 - sp) * P23[i]) + 2.0 * lam * sp * vi(i);
                    Dq[i] = G * (2.0 * (1.0 + sq) * P1[i] - (1.0 - sq) * P23[i]) + 2.0 * lam * sq * vi(i);
                }
                const double s23 = e2 + e3;
                for (int j = 0; j < 6; ++j) {
                    for (int i = 0; i < 6; ++i) {
                        const int k = i + 6 * j;
                        const double dP1 = d1 * (dd[k] - idf(i, j) * s23 - (P23[i] * et[j] + et[i] * P23[j])
                                                 + s23 * P23[i] * P23[j] + (s23 - 2.0 * e1) * P1[i] * P1[j]
                                                 + e1 * (P23[i] * P1[j] + P1[i] * P23[j]));
                        DS[k] = (s1 - s3) * dP1 + lam * vi(i) * vi(j)
                              + G * (2.0 * P1[i] * P1[j] + P23[i] * P23[j]) - Dq[i] * Dp[j] / den_r;
                    }
                }
            }
            return 3;
        }
    }

    /* apex (constitutive_3D_kernel.h:340-350) */
    const double sa = c_bar / (2.0 * sp);
    for (int i = 0; i < 6; ++i) S[i] = i < 3 ? sa : 0.0;
    if (DS) memset(DS, 0, sizeof(double) * 36);
    return 4;
}

/* Batch over points (constitutive_3d_batch.c:9-50).  DS may be NULL; branch may be NULL. */
void ssx_oracle_constitutive_3d(int64_t n, const double *E, const double *c_bar, const double *sin_phi,
                                const double *G, const double *lam, double *S, double *DS, int32_t *branch)
{
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; ++p) {
        int b = ssx_oracle_mc3d_point(E + 6 * p, c_bar[p], sin_phi[p], G[p], lam[p], S + 6 * p,
                                      DS ? DS + 36 * p : NULL);
        if (branch) branch[p] = b;
    }
}

/* Strain gather E = B u per point (assemble_tangent_values_3d.c:216-273). */
void ssx_oracle_strain_3d(const double *dphi1, const double *dphi2, const double *dphi3, const double *u,
                          const int32_t *edofs, int64_t n_elem, int n_p, int n_q, double *E)
{
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < n_elem; ++e) {
        const int32_t *dof = edofs + (size_t)e * 3 * n_p;
        for (int q = 0; q < n_q; ++q) {
            const size_t g = (size_t)e * n_q + q;
            double acc[6] = {0, 0, 0, 0, 0, 0};
            for (int j = 0; j < n_p; ++j) {
                const double ux = u[dof[3 * j]], uy = u[dof[3 * j + 1]], uz = u[dof[3 * j + 2]];
                const double gx = dphi1[g * n_p + j], gy = dphi2[g * n_p + j], gz = dphi3[g * n_p + j];
                acc[0] += gx * ux;
                acc[1] += gy * uy;
                acc[2] += gz * uz;
                acc[3] += gy * ux + gx * uy;
                acc[4] += gz * uy + gy * uz;
                acc[5] += gz * ux + gx * uz;
            }
            for (int c = 0; c < 6; ++c) E[g * 6 + c] = acc[c];
        }
    }
}

/* Row contribution of one point: w * (B_row^T sigma) for local row alpha = 3*node+comp
 * (assemble_tangent_values_3d.c:315-321). */
static double row_bt_sigma(int comp, double gx, double gy, double gz, const double *s)
{
    if (comp == 0) return gx * s[0] + gy * s[3] + gz * s[5];
    if (comp == 1) return gx * s[3] + gy * s[1] + gz * s[4];
    return gx * s[5] + gy * s[4] + gz * s[2];
}

/* Internal force per owned row through the row-slot lists (assemble_tangent_values_3d.c:275-327).
 * Rows with no slots get 0.  Constrained rows are zeroed by the caller as the reference does
 * (fem/distributed_tangent.py:932). */
void ssx_oracle_force_3d_rows(const double *dphi1, const double *dphi2, const double *dphi3, const double *S,
                              const double *w, const int32_t *row_slot_ptr, const int32_t *slot_elem,
                              const uint8_t *slot_lrow, int64_t n_rows, int n_p, int n_q, double *F)
{
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < n_rows; ++r) {
        double acc = 0.0;
        for (int32_t s = row_slot_ptr[r]; s < row_slot_ptr[r + 1]; ++s) {
            const int e = slot_elem[s], a = slot_lrow[s];
            const int node = a / 3, comp = a % 3;
            for (int q = 0; q < n_q; ++q) {
                const size_t g = (size_t)e * n_q + q;
                acc += w[g] * row_bt_sigma(comp, dphi1[g * n_p + node], dphi2[g * n_p + node],
                                           dphi3[g * n_p + node], S + 6 * g);
            }
        }
        F[r] = acc;
    }
}

/* Columns of B for local dof (node j, comp d): the nonzero strain rows and their values. */
static void bcol(int d, double gx, double gy, double gz, double col[6])
{
    for (int s = 0; s < 6; ++s) col[s] = 0.0;
    if (d == 0) { col[0] = gx; col[3] = gy; col[5] = gz; }
    else if (d == 1) { col[1] = gy; col[3] = gx; col[4] = gz; }
    else { col[2] = gz; col[4] = gy; col[5] = gx; }
}

/* Tangent values on the fixed pattern through the row-slot lists
 * (assemble_tangent_values_3d.c:111-214): for each owned row, slot (element order), point:
 * t = w * B_row^T DS, then K[row, col(j,d)] += t . B_col(j,d).  out is NOT cleared here. */
void ssx_oracle_tangent_3d_rows(const double *dphi1, const double *dphi2, const double *dphi3, const double *DS,
                                const double *w, const int32_t *row_slot_ptr, const int32_t *slot_elem,
                                const uint8_t *slot_lrow, const int32_t *slot_pos, int64_t n_rows, int n_p,
                                int n_q, double *out)
{
    const int nld = 3 * n_p;
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < n_rows; ++r) {
        for (int32_t s = row_slot_ptr[r]; s < row_slot_ptr[r + 1]; ++s) {
            const int e = slot_elem[s], a = slot_lrow[s];
            const int node = a / 3, comp = a % 3;
            const int32_t *pos = slot_pos + (size_t)s * nld;
            for (int q = 0; q < n_q; ++q) {
                const size_t g = (size_t)e * n_q + q;
                const double *D = DS + 36 * g;
                double brow[6], t[6];
                bcol(comp, dphi1[g * n_p + node], dphi2[g * n_p + node], dphi3[g * n_p + node], brow);
                /* t_m = w * sum_s B_row[s] * D[s, m]   with D[s,m] = DS[s + 6m] */
                for (int m = 0; m < 6; ++m) {
                    double acc = 0.0;
                    for (int s2 = 0; s2 < 6; ++s2) acc += brow[s2] * D[s2 + 6 * m];
                    t[m] = w[g] * acc;
                }
                for (int j = 0; j < n_p; ++j) {
                    for (int d = 0; d < 3; ++d) {
                        const int32_t p = pos[3 * j + d];
                        if (p < 0) continue;
                        double bc[6];
                        bcol(d, dphi1[g * n_p + j], dphi2[g * n_p + j], dphi3[g * n_p + j], bc);
                        double v = 0.0;
                        for (int m = 0; m < 6; ++m) v += t[m] * bc[m];
                        out[p] += v;
                    }
                }
            }
        }
    }
}

/* Dense element matrices ke = sum_q B^T (w D) B, row-major (3n_p x 3n_p) per element
 * (assemble_tangent_values_3d.c:45-94, the legacy kernel's element stage). */
void ssx_oracle_element_matrices_3d(const double *dphi1, const double *dphi2, const double *dphi3,
                                    const double *DS, const double *w, int64_t n_elem, int n_p, int n_q,
                                    double *ke)
{
    const int nld = 3 * n_p;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < n_elem; ++e) {
        double *K = ke + (size_t)e * nld * nld;
        memset(K, 0, sizeof(double) * nld * nld);
        for (int q = 0; q < n_q; ++q) {
            const size_t g = (size_t)e * n_q + q;
            const double *D = DS + 36 * g;
            for (int a = 0; a < nld; ++a) {
                double ba[6];
                bcol(a % 3, dphi1[g * n_p + a / 3], dphi2[g * n_p + a / 3], dphi3[g * n_p + a / 3], ba);
                for (int b = 0; b < nld; ++b) {
                    double bb[6];
                    bcol(b % 3, dphi1[g * n_p + b / 3], dphi2[g * n_p + b / 3], dphi3[g * n_p + b / 3], bb);
                    double v = 0.0;
                    for (int s = 0; s < 6; ++s)
                        for (int m = 0; m < 6; ++m) v += ba[s] * D[s + 6 * m] * bb[m];
                    K[(size_t)a * nld + b] += w[g] * v;
                }
            }
        }
    }
}

/* y = A x for a CSR matrix (textbook; stands in for PETSc MatMult at linear/solver.py:3431-3451). */
void ssx_oracle_spmv_csr(int64_t n_rows, const int32_t *indptr, const int32_t *indices, const double *val,
                         const double *x, double *y)
{
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < n_rows; ++r) {
        double acc = 0.0;
        for (int32_t k = indptr[r]; k < indptr[r + 1]; ++k) acc += val[k] * x[indices[k]];
        y[r] = acc;
    }
}
