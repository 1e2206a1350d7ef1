/*
 * Oracle, ideal MHD with constrained transport (TEST INFRASTRUCTURE ONLY).
 * Restates src/physics_mhd.hpp and mara::advance_mhd_ct
 * (src/scheme_mhd_v1.cpp:265-299) in the reference's association order.
 * Build with -ffp-contract=off.
 *
 * Layout (logical order, row-major, last axis fastest):
 *   uc  [N][N][N][5]      conserved (D, S1, S2, S3, E) per cell
 *   bf1 [N+1][N][N]       x-face magnetic field (both periodic end faces stored)
 *   bf2 [N][N+1][N]       y-face
 *   bf3 [N][N][N+1]       z-face
 * Primitive tuple (rho, v1, v2, v3, p, B1, B2, B3).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline double smax(double a, double b) { return (a < b) ? b : a; }
static inline double smin(double a, double b) { return (b < a) ? b : a; }

/* euclidean dot with the 0-seeded left fold of src/core_numeric_array.hpp:177-188 */
static inline double dot3(const double* a, const double* b) { return ((0.0 + a[0] * b[0]) + a[1] * b[1]) + a[2] * b[2]; }
static inline double ls3(const double* a) { return dot3(a, a); }
static inline void unit3(int axis, double* n) { n[0] = axis == 1; n[1] = axis == 2; n[2] = axis == 3; }
/* src/core_geometric.hpp:148-155 */
static inline void cross3(const double* a, const double* b, double* c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* recover_primitive, src/physics_mhd.hpp:287-305.  1 = negative density,
 * 2 = negative pressure (the reference throws std::invalid_argument). */
int orc_mhd_recover_primitive(const double* u, const double* b, double gamma, double* p)
{
    double d = u[0];
    double S[3] = { u[1], u[2], u[3] };
    double ke = 0.5 * ls3(S) / d;
    double me = 0.5 * ls3(b);
    double pg = (u[4] - ke - me) * (gamma - 1);
    p[0] = d;
    p[1] = S[0] / d;
    p[2] = S[1] / d;
    p[3] = S[2] / d;
    p[4] = pg;
    p[5] = b[0];
    p[6] = b[1];
    p[7] = b[2];
    if (d < 0.0) return 1;
    if (pg < 0.0) return 2;
    return 0;
}

/* conserved_density :307-319 */
void orc_mhd_conserved_density(const double* p, double gamma, double* u)
{
    double d = p[0];
    const double* v = &p[1];
    const double* b = &p[5];
    double E = 0.5 * d * ls3(v) + p[4] / (gamma - 1.0) + 0.5 * ls3(b);
    u[0] = d;
    u[1] = d * v[0];
    u[2] = d * v[1];
    u[3] = d * v[2];
    u[4] = E;
}

/* flux :324-347 */
void orc_mhd_flux(const double* p, int axis, double gamma, double* f)
{
    double n[3];
    unit3(axis, n);
    const double* v = &p[1];
    const double* b = &p[5];
    double bn = dot3(b, n);
    double vn = dot3(v, n);
    double pt = p[4] + 0.5 * ls3(b);
    double u[5];
    orc_mhd_conserved_density(p, gamma, u);
    double pterm[5] = { 0.0 * 0.0, pt * n[0], pt * n[1], pt * n[2], pt * vn };
    double mterm[5] = { 0.0 * 0.0, -bn * b[0], -bn * b[1], -bn * b[2], -bn * dot3(b, v) };
    for (int k = 0; k < 5; ++k) f[k] = vn * u[k] + pterm[k] + mterm[k];
}

/* electric_field_vector :154-157 (E = -v x B), induction :349-352 (n x E) */
static void mhd_induction(const double* p, int axis, double* out)
{
    double n[3], vxb[3], E[3];
    unit3(axis, n);
    cross3(&p[1], &p[5], vxb);
    E[0] = -vxb[0]; E[1] = -vxb[1]; E[2] = -vxb[2];
    cross3(n, E, out);
}

/* outer_wavespeeds :216-246 via fast_speed_squared (alfven :205-214) */
static void mhd_outer_wavespeeds(const double* p, int axis, double gamma, double* wm, double* wp)
{
    double n[3];
    unit3(axis, n);
    const double* v = &p[1];
    const double* b = &p[5];
    double vn = dot3(n, v);
    double cs2 = gamma * p[4] / p[0];
    double ca2 = ls3(b) / p[0];
    double cf2 = cs2 + ca2;
    double can2 = pow(dot3(b, n), 2.0) / p[0];
    double cg2 = sqrt(cf2 * cf2 - 4.0 * cs2 * can2);
    double cf = sqrt(0.5 * (cf2 + cg2));
    *wm = vn - cf;
    *wp = vn + cf;
}

/* fastest_speed :248-258 */
double orc_mhd_fastest_speed(const double* p, double gamma)
{
    double m[3], q[3];
    mhd_outer_wavespeeds(p, 1, gamma, &m[0], &q[0]);
    mhd_outer_wavespeeds(p, 2, gamma, &m[1], &q[1]);
    mhd_outer_wavespeeds(p, 3, gamma, &m[2], &q[2]);
    return smax(smax(smax(fabs(m[0]), fabs(q[0])), smax(fabs(m[1]), fabs(q[1]))),
                smax(fabs(m[2]), fabs(q[2])));
}

/* with_longitudinal_field :165-176 */
static void with_longitudinal_field(const double* p, int axis, double bl, double* out)
{
    double n[3];
    unit3(axis, n);
    const double* b = &p[5];
    double bn = dot3(b, n);
    memcpy(out, p, 5 * sizeof(double));
    for (int k = 0; k < 3; ++k) out[5 + k] = (b[k] - bn * n[k]) + bl * n[k];
}

/* riemann_hlle :358-390: Godunov flux f[5] and the face electric field e[3] */
void orc_mhd_riemann_hlle(const double* pl_, const double* pr_, double bl, int axis, double gamma, double* f, double* e)
{
    double pl[8], pr[8], ul[5], ur[5], fl[5], fr[5], il[3], ir[3], n[3], hi[3], c[3];
    with_longitudinal_field(pl_, axis, bl, pl);
    with_longitudinal_field(pr_, axis, bl, pr);
    orc_mhd_conserved_density(pl, gamma, ul);
    orc_mhd_conserved_density(pr, gamma, ur);
    orc_mhd_flux(pl, axis, gamma, fl);
    orc_mhd_flux(pr, axis, gamma, fr);
    mhd_induction(pl, axis, il);
    mhd_induction(pr, axis, ir);
    double alm, alp, arm, arp;
    mhd_outer_wavespeeds(pl, axis, gamma, &alm, &alp);
    mhd_outer_wavespeeds(pr, axis, gamma, &arm, &arp);
    double ap = smax(0.0, smax(alp, arp));
    double am = smin(0.0, smin(alm, arm));
    for (int k = 0; k < 5; ++k)
        f[k] = (ap * fl[k] - am * fr[k] - (ul[k] - ur[k]) * ap * am) / (ap - am);
    for (int k = 0; k < 3; ++k)
        hi[k] = (ap * il[k] - am * ir[k] - (pl[5 + k] - pr[5 + k]) * ap * am) / (ap - am);
    unit3(axis, n);
    cross3(n, hi, c);
    e[0] = -c[0]; e[1] = -c[1]; e[2] = -c[2];
}

/* PLM, src/scheme_plm_gradient.hpp:59-67 (orc_plm in oracle.c) */
double orc_plm(double yl, double y0, double yr, double theta);

/* ------------------------------------------------------------------------ */
/* advance_mhd_ct, src/scheme_mhd_v1.cpp:265-299, periodic (local_periodic_boundary_extension
 * :185-208), PLM reconstruction (godunov_fluxes_plm :152-179), edge EMFs
 * (edge_emf_from_face :76-98), solenoidal / divergence differences
 * (src/mesh_cartesian_3d.hpp:265-301).  Returns 0 or a recover status. */

#define W(i) (((i) % N + N) % N)

int orc_mhd_advance_ct(int N, const double* uc, const double* bf1, const double* bf2, const double* bf3,
                       double dl, double cfl, double theta, double gamma,
                       double* uo, double* bo1, double* bo2, double* bo3, double* dt_out)
{
    const long NN = (long)N * N, N3 = NN * N;
#define BF1(i, j, k) bf1[((long)(i) * N + (j)) * N + (k)]
#define BF2(i, j, k) bf2[((long)(i) * (N + 1) + (j)) * N + (k)]
#define BF3(i, j, k) bf3[((long)(i) * N + (j)) * (N + 1) + (k)]
    double* P = (double*)malloc(sizeof(double) * N3 * 8);
    int status = 0;
    double vmax = -INFINITY;
    /* primitive_array :214-224 with face_to_cell (mesh_cartesian_3d.hpp:235-245) */
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                long c = (i * (long)N + j) * N + k;
                double b[3] = { 0.5 * (BF1(i, j, k) + BF1(i + 1, j, k)),
                                0.5 * (BF2(i, j, k) + BF2(i, j + 1, k)),
                                0.5 * (BF3(i, j, k) + BF3(i, j, k + 1)) };
                int st = orc_mhd_recover_primitive(&uc[c * 5], b, gamma, &P[c * 8]);
                if (st && !status) status = st;
                /* max(pc | fastest_speed): nd::max folds with std::max from the first element */
                double s = orc_mhd_fastest_speed(&P[c * 8], gamma);
                vmax = (c == 0) ? s : smax(vmax, s);
            }
    if (status) { free(P); return status; }
#define PC(i, j, k) (&P[((long)W(i) * N + W(j)) * N * 8 + (long)W(k) * 8])
    /* face data: axis a faces have N+1 entries along a, N along the others */
    long nf = (long)(N + 1) * NN;
    double* F[3];
    double* E[3];
    for (int a = 0; a < 3; ++a) { F[a] = (double*)malloc(sizeof(double) * nf * 5); E[a] = (double*)malloc(sizeof(double) * nf * 3); }
    for (int a = 0; a < 3; ++a)
        for (long idx = 0; idx < nf; ++idx) {
            int i, j, k;
            if (a == 0) { i = (int)(idx / NN); j = (int)(idx / N % N); k = (int)(idx % N); }
            else if (a == 1) { i = (int)(idx / ((N + 1) * N)); j = (int)(idx / N % (N + 1)); k = (int)(idx % N); }
            else { i = (int)(idx / ((N + 1) * N)); j = (int)(idx / (N + 1) % N); k = (int)(idx % (N + 1)); }
            /* the four cells along the axis: face between cells m-1 and m */
            const double* c[4];
            for (int s = 0; s < 4; ++s) {
                int off = s - 2;
                if (a == 0) c[s] = PC(i + off, j, k);
                else if (a == 1) c[s] = PC(i, j + off, k);
                else c[s] = PC(i, j, k + off);
            }
            double pl[8], pr[8];
            for (int q = 0; q < 8; ++q) {
                double gl = orc_plm(c[0][q], c[1][q], c[2][q], theta);
                double gr = orc_plm(c[1][q], c[2][q], c[3][q], theta);
                pl[q] = c[1][q] + 0.5 * gl;
                pr[q] = c[2][q] - 0.5 * gr;
            }
            double bl = a == 0 ? BF1(i, j, k) : (a == 1 ? BF2(i, j, k) : BF3(i, j, k));
            orc_mhd_riemann_hlle(pl, pr, bl, a + 1, gamma, &F[a][idx * 5], &E[a][idx * 3]);
        }
    /* face accessors with periodic wrap on the cell (transverse) indices */
#define E1(c, i, j, k) E[0][(((long)(i) * N + W(j)) * N + W(k)) * 3 + (c)]
#define E2(c, i, j, k) E[1][(((long)W(i) * (N + 1) + (j)) * N + W(k)) * 3 + (c)]
#define E3(c, i, j, k) E[2][(((long)W(i) * N + W(j)) * (N + 1) + (k)) * 3 + (c)]
    /* edge EMFs: ee1 [N][N+1][N+1], ee2 [N+1][N][N+1], ee3 [N+1][N+1][N] */
    double* ee1 = (double*)malloc(sizeof(double) * N * (N + 1) * (N + 1));
    double* ee2 = (double*)malloc(sizeof(double) * (N + 1) * N * (N + 1));
    double* ee3 = (double*)malloc(sizeof(double) * (N + 1) * (N + 1) * N);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j <= N; ++j)
            for (int k = 0; k <= N; ++k) {
                double a = 0.5 * (E2(0, i, j, k - 1) + E2(0, i, j, k));
                double b = 0.5 * (E3(0, i, j - 1, k) + E3(0, i, j, k));
                ee1[((long)i * (N + 1) + j) * (N + 1) + k] = 0.5 * (a + b);
            }
    for (int i = 0; i <= N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k <= N; ++k) {
                double a = 0.5 * (E3(1, i - 1, j, k) + E3(1, i, j, k));
                double b = 0.5 * (E1(1, i, j, k - 1) + E1(1, i, j, k));
                ee2[((long)i * N + j) * (N + 1) + k] = 0.5 * (a + b);
            }
    for (int i = 0; i <= N; ++i)
        for (int j = 0; j <= N; ++j)
            for (int k = 0; k < N; ++k) {
                double a = 0.5 * (E1(2, i, j - 1, k) + E1(2, i, j, k));
                double b = 0.5 * (E2(2, i - 1, j, k) + E2(2, i, j, k));
                ee3[((long)i * (N + 1) + j) * N + k] = 0.5 * (a + b);
            }
#define EE1(i, j, k) ee1[((long)(i) * (N + 1) + (j)) * (N + 1) + (k)]
#define EE2(i, j, k) ee2[((long)(i) * N + (j)) * (N + 1) + (k)]
#define EE3(i, j, k) ee3[((long)(i) * (N + 1) + (j)) * N + (k)]
    double dt = cfl * dl / vmax;
    *dt_out = dt;
    /* conserved update: uc - dc * dt / dl with dc = d1(ff1) + d2(ff2) + d3(ff3) */
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                long c = (i * (long)N + j) * N + k;
                for (int q = 0; q < 5; ++q) {
                    double d1 = F[0][(((long)(i + 1) * N + j) * N + k) * 5 + q] - F[0][(((long)i * N + j) * N + k) * 5 + q];
                    double d2 = F[1][(((long)i * (N + 1) + j + 1) * N + k) * 5 + q] - F[1][(((long)i * (N + 1) + j) * N + k) * 5 + q];
                    double d3 = F[2][(((long)i * N + j) * (N + 1) + k + 1) * 5 + q] - F[2][(((long)i * N + j) * (N + 1) + k) * 5 + q];
                    uo[c * 5 + q] = uc[c * 5 + q] - (d1 + d2 + d3) * dt / dl;
                }
            }
    /* face updates: bf - cf * dt / dl with the solenoidal difference of the edge EMFs */
    for (int i = 0; i <= N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k < N; ++k) {
                double cf = (EE3(i, j + 1, k) - EE3(i, j, k)) - (EE2(i, j, k + 1) - EE2(i, j, k));
                bo1[((long)i * N + j) * N + k] = BF1(i, j, k) - cf * dt / dl;
            }
    for (int i = 0; i < N; ++i)
        for (int j = 0; j <= N; ++j)
            for (int k = 0; k < N; ++k) {
                double cf = (EE1(i, j, k + 1) - EE1(i, j, k)) - (EE3(i + 1, j, k) - EE3(i, j, k));
                bo2[((long)i * (N + 1) + j) * N + k] = BF2(i, j, k) - cf * dt / dl;
            }
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            for (int k = 0; k <= N; ++k) {
                double cf = (EE2(i + 1, j, k) - EE2(i, j, k)) - (EE1(i, j + 1, k) - EE1(i, j, k));
                bo3[((long)i * N + j) * (N + 1) + k] = BF3(i, j, k) - cf * dt / dl;
            }
    for (int a = 0; a < 3; ++a) { free(F[a]); free(E[a]); }
    free(ee1); free(ee2); free(ee3); free(P);
    return 0;
#undef BF1
#undef BF2
#undef BF3
#undef PC
#undef E1
#undef E2
#undef E3
#undef EE1
#undef EE2
#undef EE3
}
#undef W

/* control::advance_runge_kutta order 2 (src/app_control.hpp:169-195) around
 * advance_mhd_ct with the weighted_sum of examples/mhd3d.cpp:77-86: each stage
 * computes its own CFL dt, as the reference does.  In place; returns status and
 * the two stage time steps. */
int orc_mhd_rk2(int N, double* uc, double* bf1, double* bf2, double* bf3, double dl, double cfl, double theta,
                double gamma, double* dts)
{
    const long N3 = (long)N * N * N, nf = (long)(N + 1) * N * N;
    double* u1 = (double*)malloc(sizeof(double) * N3 * 5);
    double* u2 = (double*)malloc(sizeof(double) * N3 * 5);
    double* b[6];
    for (int a = 0; a < 6; ++a) b[a] = (double*)malloc(sizeof(double) * nf);
    int st = orc_mhd_advance_ct(N, uc, bf1, bf2, bf3, dl, cfl, theta, gamma, u1, b[0], b[1], b[2], &dts[0]);
    if (!st) st = orc_mhd_advance_ct(N, u1, b[0], b[1], b[2], dl, cfl, theta, gamma, u2, b[3], b[4], b[5], &dts[1]);
    if (!st) {
        for (long z = 0; z < N3 * 5; ++z) uc[z] = uc[z] * 0.5 + u2[z] * (1.0 - 0.5);
        for (long z = 0; z < nf; ++z) {
            bf1[z] = bf1[z] * 0.5 + b[3][z] * (1.0 - 0.5);
            bf2[z] = bf2[z] * 0.5 + b[4][z] * (1.0 - 0.5);
            bf3[z] = bf3[z] * 0.5 + b[5][z] * (1.0 - 0.5);
        }
    }
    free(u1); free(u2);
    for (int a = 0; a < 6; ++a) free(b[a]);
    return st;
}

/* ------------------------------------------------------------------------ */
/* mhd_scheme_v2 array functions (src/scheme_mhd_v2.cpp:75-278) on blocks of
 * any shape, adiabatic index 5/3 (:48).  Logical layout, row-major: cell
 * tuples last ([..][8] primitives, [..][5] conserved, [..][8] Godunov data =
 * the 5 flux components then the 3 components of the face electric field);
 * face and edge arrays are scalars.  For cells (n1, n2, n3): x faces
 * (n1+1, n2, n3), x edges (n1, n2+1, n3+1), and so on by symmetry. */

static const double V2_GAMMA = 5. / 3;

#define IX3(i, j, k, s2, s3) (((long)(i) * (s2) + (j)) * (s3) + (k))

/* primitive_array :75-81 with face_to_cell (src/mesh_cartesian_3d.hpp:235-245).
 * Returns the status of the first cell in row-major order that fails. */
int orc_mhd_v2_primitive_array(int n1, int n2, int n3, const double* uc, const double* bf1, const double* bf2,
                               const double* bf3, double* pc)
{
    int status = 0;
    for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n2; ++j)
            for (int k = 0; k < n3; ++k) {
                long c = IX3(i, j, k, n2, n3);
                double b[3] = { 0.5 * (bf1[IX3(i, j, k, n2, n3)] + bf1[IX3(i + 1, j, k, n2, n3)]),
                                0.5 * (bf2[IX3(i, j, k, n2 + 1, n3)] + bf2[IX3(i, j + 1, k, n2 + 1, n3)]),
                                0.5 * (bf3[IX3(i, j, k, n2, n3 + 1)] + bf3[IX3(i, j, k + 1, n2, n3 + 1)]) };
                int st = orc_mhd_recover_primitive(&uc[c * 5], b, V2_GAMMA, &pc[c * 8]);
                if (st && !status) status = st;
            }
    return status;
}

/* godunov_fluxes :168-185: first-order face states (the two cells either side
 * of the face), HLLE with the face's own longitudinal field.  pc (m1, m2, m3);
 * bf1 / gf1 (m1-1, m2, m3), bf2 / gf2 (m1, m2-1, m3), bf3 / gf3 (m1, m2, m3-1). */
void orc_mhd_v2_godunov_fluxes(int m1, int m2, int m3, const double* pc, const double* bf1, const double* bf2,
                               const double* bf3, double* gf1, double* gf2, double* gf3)
{
    for (int a = 0; a < 3; ++a) {
        const int s1 = m1 - (a == 0), s2 = m2 - (a == 1), s3 = m3 - (a == 2);
        const double* bf = a == 0 ? bf1 : (a == 1 ? bf2 : bf3);
        double* gf = a == 0 ? gf1 : (a == 1 ? gf2 : gf3);
        for (int i = 0; i < s1; ++i)
            for (int j = 0; j < s2; ++j)
                for (int k = 0; k < s3; ++k) {
                    long f = IX3(i, j, k, s2, s3);
                    const double* pl = &pc[IX3(i, j, k, m2, m3) * 8];
                    const double* pr = &pc[IX3(i + (a == 0), j + (a == 1), k + (a == 2), m2, m3) * 8];
                    orc_mhd_riemann_hlle(pl, pr, bf[f], a + 1, V2_GAMMA, &gf[f * 8], &gf[f * 8 + 5]);
                }
    }
}

/* electromotive_forces :191-214: each edge field is the mean of two face
 * estimates, each averaged along its other transverse axis (nd::adjacent_mean,
 * src/core_ndarray_ops.hpp:84-90).  gf from a cell array (m1, m2, m3);
 * ee1 (m1, m2-1, m3-1), ee2 (m1-1, m2, m3-1), ee3 (m1-1, m2-1, m3). */
void orc_mhd_v2_electromotive_forces(int m1, int m2, int m3, const double* gf1, const double* gf2, const double* gf3,
                                     double* ee1, double* ee2, double* ee3)
{
#define G1(c, i, j, k) gf1[IX3(i, j, k, m2, m3) * 8 + (c)]
#define G2(c, i, j, k) gf2[IX3(i, j, k, m2 - 1, m3) * 8 + (c)]
#define G3(c, i, j, k) gf3[IX3(i, j, k, m2, m3 - 1) * 8 + (c)]
    for (int i = 0; i < m1; ++i)
        for (int j = 0; j < m2 - 1; ++j)
            for (int k = 0; k < m3 - 1; ++k) {
                double a = 0.5 * (G2(5, i, j, k) + G2(5, i, j, k + 1));
                double b = 0.5 * (G3(5, i, j, k) + G3(5, i, j + 1, k));
                ee1[IX3(i, j, k, m2 - 1, m3 - 1)] = 0.5 * (a + b);
            }
    for (int i = 0; i < m1 - 1; ++i)
        for (int j = 0; j < m2; ++j)
            for (int k = 0; k < m3 - 1; ++k) {
                double a = 0.5 * (G3(6, i, j, k) + G3(6, i + 1, j, k));
                double b = 0.5 * (G1(6, i, j, k) + G1(6, i, j, k + 1));
                ee2[IX3(i, j, k, m2, m3 - 1)] = 0.5 * (a + b);
            }
    for (int i = 0; i < m1 - 1; ++i)
        for (int j = 0; j < m2 - 1; ++j)
            for (int k = 0; k < m3; ++k) {
                double a = 0.5 * (G1(7, i, j, k) + G1(7, i, j + 1, k));
                double b = 0.5 * (G2(7, i, j, k) + G2(7, i + 1, j, k));
                ee3[IX3(i, j, k, m2 - 1, m3)] = 0.5 * (a + b);
            }
#undef G1
#undef G2
#undef G3
}

/* updated_conserved_density :220-241: uc (n1, n2, n3); gf extended by one on
 * the transverse axes, gf1 (n1+1, n2+2, n3+2) etc. (remove_transverse_*,
 * src/mesh_cartesian_3d.hpp:421-452); U - div(F) dt / dl with
 * divergence_difference (:290-301). */
void orc_mhd_v2_updated_conserved_density(int n1, int n2, int n3, const double* uc, const double* gf1,
                                          const double* gf2, const double* gf3, double dt, double dl, double* out)
{
    for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n2; ++j)
            for (int k = 0; k < n3; ++k) {
                long c = IX3(i, j, k, n2, n3);
                for (int q = 0; q < 5; ++q) {
                    double d1 = gf1[IX3(i + 1, j + 1, k + 1, n2 + 2, n3 + 2) * 8 + q] - gf1[IX3(i, j + 1, k + 1, n2 + 2, n3 + 2) * 8 + q];
                    double d2 = gf2[IX3(i + 1, j + 1, k + 1, n2 + 1, n3 + 2) * 8 + q] - gf2[IX3(i + 1, j, k + 1, n2 + 1, n3 + 2) * 8 + q];
                    double d3 = gf3[IX3(i + 1, j + 1, k + 1, n2 + 2, n3 + 1) * 8 + q] - gf3[IX3(i + 1, j + 1, k, n2 + 2, n3 + 1) * 8 + q];
                    out[c * 5 + q] = uc[c * 5 + q] - (d1 + d2 + d3) * dt / dl;
                }
            }
}

/* solenoidal_difference (src/mesh_cartesian_3d.hpp:265-276) of edge arrays
 * e1 (n1, n2+1, n3+1), e2 (n1+1, n2, n3+1), e3 (n1+1, n2+1, n3) read through
 * offsets (o1 on axis 0 of e1, o2 on axis 1 of e2, o3 on axis 2 of e3) with
 * the stored extents t1, t2, t3 on those axes; each face value is
 * bf - cf * dt / dl (bf NULL: 0 - ...) or, with bf == NULL and dt == 1 as the
 * curl, cf * 1 / dl. */
static void v2_faces_from_edges(int n1, int n2, int n3, const double* e1, const double* e2, const double* e3,
                                int o, const double* bf1, const double* bf2, const double* bf3, double dt, double dl,
                                int is_curl, double* b1, double* b2, double* b3)
{
    /* stored shapes: e1 (n1+2o, n2+1, n3+1), e2 (n1+1, n2+2o, n3+1), e3 (n1+1, n2+1, n3+2o) */
#define E1(i, j, k) e1[IX3((i) + o, j, k, n2 + 1, n3 + 1)]
#define E2(i, j, k) e2[IX3(i, (j) + o, k, n2 + 2 * o, n3 + 1)]
#define E3(i, j, k) e3[IX3(i, j, (k) + o, n2 + 1, n3 + 2 * o)]
    for (int i = 0; i <= n1; ++i)
        for (int j = 0; j < n2; ++j)
            for (int k = 0; k < n3; ++k) {
                long f = IX3(i, j, k, n2, n3);
                double cf = (E3(i, j + 1, k) - E3(i, j, k)) - (E2(i, j, k + 1) - E2(i, j, k));
                b1[f] = is_curl ? cf * dt / dl : bf1[f] - cf * dt / dl;
            }
    for (int i = 0; i < n1; ++i)
        for (int j = 0; j <= n2; ++j)
            for (int k = 0; k < n3; ++k) {
                long f = IX3(i, j, k, n2 + 1, n3);
                double cf = (E1(i, j, k + 1) - E1(i, j, k)) - (E3(i + 1, j, k) - E3(i, j, k));
                b2[f] = is_curl ? cf * dt / dl : bf2[f] - cf * dt / dl;
            }
    for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n2; ++j)
            for (int k = 0; k <= n3; ++k) {
                long f = IX3(i, j, k, n2, n3 + 1);
                double cf = (E2(i + 1, j, k) - E2(i, j, k)) - (E1(i, j + 1, k) - E1(i, j, k));
                b3[f] = is_curl ? cf * dt / dl : bf3[f] - cf * dt / dl;
            }
#undef E1
#undef E2
#undef E3
}

/* curl :247-258: faces of cells (n1, n2, n3) from edges e1 (n1, n2+1, n3+1),
 * e2 (n1+1, n2, n3+1), e3 (n1+1, n2+1, n3), over a unit time. */
void orc_mhd_v2_curl(int n1, int n2, int n3, const double* e1, const double* e2, const double* e3, double dl,
                     double* b1, double* b2, double* b3)
{
    v2_faces_from_edges(n1, n2, n3, e1, e2, e3, 0, NULL, NULL, NULL, 1.0, dl, 1, b1, b2, b3);
}

/* updated_magnetic_flux_density :264-278: edges as electromotive_forces
 * returns them for gf extended by one (ee1 (n1+2, n2+1, n3+1), ...), the
 * longitudinal end layers dropped (rl1..rl3 :67-69). */
void orc_mhd_v2_updated_magnetic_flux_density(int n1, int n2, int n3, const double* bf1, const double* bf2,
                                              const double* bf3, const double* ee1, const double* ee2,
                                              const double* ee3, double dt, double dl, double* b1, double* b2,
                                              double* b3)
{
    v2_faces_from_edges(n1, n2, n3, ee1, ee2, ee3, 1, bf1, bf2, bf3, dt, dl, 0, b1, b2, b3);
}

/* extend_cell_primitive_variables :126-135: the 27 blocks of a 3 x 3 x 3
 * neighbourhood (row-major over the offsets, tile_blocks
 * src/mesh_cartesian_3d.hpp:307-337) of bs^3 cells x nq components, cut to
 * (bs + 2 count)^3 around the centre block (remove_surface :410-419). */
void orc_mhd_v2_extend_cells(const double* const* blocks, int bs, int count, int nq, double* out)
{
    const int r = bs - count, m = bs + 2 * count;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j)
            for (int k = 0; k < m; ++k) {
                int I = i + r, J = j + r, K = k + r;
                const double* b = blocks[(I / bs) * 9 + (J / bs) * 3 + K / bs];
                memcpy(&out[IX3(i, j, k, m, m) * nq], &b[IX3(I % bs, J % bs, K % bs, bs, bs) * nq], sizeof(double) * nq);
            }
}

/* extend_face_magnetic_flux_density :141-162: for axis a, the 9 face arrays
 * blocks[9a .. 9a+8] of a 3 x 3 square normal to a (row-major over the two
 * transverse offsets), tiled and cut to count zones on each transverse side
 * (remove_transverse :458-464): bf1 (bs+1, bs+2c, bs+2c), etc. */
void orc_mhd_v2_extend_faces(const double* const* blocks, int bs, int count, double* bf1, double* bf2, double* bf3)
{
    const int r = bs - count, m = bs + 2 * count;
    for (int a = 0; a < 3; ++a) {
        double* out = a == 0 ? bf1 : (a == 1 ? bf2 : bf3);
        const int s1 = a == 0 ? bs + 1 : m, s2 = a == 1 ? bs + 1 : m, s3 = a == 2 ? bs + 1 : m;
        const int b1 = a == 0 ? bs + 1 : bs, b2 = a == 1 ? bs + 1 : bs, b3 = a == 2 ? bs + 1 : bs;
        for (int i = 0; i < s1; ++i)
            for (int j = 0; j < s2; ++j)
                for (int k = 0; k < s3; ++k) {
                    int I = a == 0 ? i : i + r, J = a == 1 ? j : j + r, K = a == 2 ? k : k + r;
                    /* block offsets on the axes, 0 on the normal axis (extent 1 there) */
                    int oi = I / b1, oj = J / b2, ok = K / b3;
                    int slot = a == 0 ? oj * 3 + ok : (a == 1 ? oi * 3 + ok : oi * 3 + oj);
                    const double* b = blocks[9 * a + slot];
                    out[IX3(i, j, k, s2, s3)] = b[IX3(I % b1, J % b2, K % b3, b2, b3)];
                }
    }
}
#undef IX3
