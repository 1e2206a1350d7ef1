/*
 * Oracle: CPU FP64 restatement of Mara3's Godunov block update.
 * TEST INFRASTRUCTURE ONLY -- the parity reference of the tests and the timed CPU baseline; the
 * product path never calls it.  Build with -ffp-contract=off (no FMA) so that
 * every operation rounds exactly as the reference's x86-64 build does.
 *
 * Citations are to the reference (jzrake/Mara3p) file:line.  Association
 * order follows the reference's operator chains literally, e.g. tuple
 * operators are component-wise and left-folded (src/core_numeric_tuple.hpp:165-170),
 * dot/length_squared are sums seeded with 0.0 (src/core_geometric.hpp:124-146,
 * src/core_numeric_array.hpp:177-188).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* std::max(a,b) == (a < b) ? b : a ; std::min(a,b) == (b < a) ? b : a */
static inline double smax(double a, double b) { return (a < b) ? b : a; }
static inline double smin(double a, double b) { return (b < a) ? b : a; }

/* ------------------------------------------------------------------------ */
/* PLM: src/scheme_plm_gradient.hpp:59-67 with coordinates (0,1,2)
 * (plm_gradient_t::operator(), :90-95). */
double orc_plm(double yl, double y0, double yr, double theta)
{
    double a = (y0 - yl) / (1.0 - 0.0) * theta;
    double b = (yr - yl) / (2.0 - 0.0);
    double c = (yr - y0) / (2.0 - 1.0) * theta;
    double sa = copysign(1.0, a), sb = copysign(1.0, b), sc = copysign(1.0, c);
    double m = smin(smin(fabs(a), fabs(b)), fabs(c));
    return 0.25 * fabs(sa + sb) * (sa + sc) * m;
}

/* ------------------------------------------------------------------------ */
/* Euler: src/physics_euler.hpp */

/* :142-169 recover_primitive (kinetic_energy_density :142-145, internal :147-150) */
int orc_euler_recover_primitive(const double* u, double gamma, double* p)
{
    double D = u[0], S1 = u[1], S2 = u[2], S3 = u[3], E = u[4];
    if (D < 0.0) return -1;
    double ls = ((0.0 + S1 * S1) + S2 * S2) + S3 * S3;
    double ke = 0.5 * ls / D;
    p[0] = D;
    p[1] = S1 / D;
    p[2] = S2 / D;
    p[3] = S3 / D;
    p[4] = (E - ke) * (gamma - 1.0);
    return 0;
}

/* :175-183 conserved_density */
void orc_euler_conserved_density(const double* p, double gamma, double* u)
{
    double d = p[0], v1 = p[1], v2 = p[2], v3 = p[3], pg = p[4];
    double ls = ((0.0 + v1 * v1) + v2 * v2) + v3 * v3;
    u[0] = d;
    u[1] = d * v1;
    u[2] = d * v2;
    u[3] = d * v3;
    u[4] = d * ls * 0.5 + pg / (gamma - 1.0);
}

static inline double unit_dot(int axis, double v1, double v2, double v3)
{
    /* dot(unit_vector_on(axis), v) = ((0 + n1 v1) + n2 v2) + n3 v3 */
    double n1 = axis == 1 ? 1.0 : 0.0, n2 = axis == 2 ? 1.0 : 0.0, n3 = axis == 3 ? 1.0 : 0.0;
    return ((0.0 + n1 * v1) + n2 * v2) + n3 * v3;
}

/* :189-203 flux */
void orc_euler_flux(const double* p, int axis, double gamma, double* f)
{
    double n1 = axis == 1 ? 1.0 : 0.0, n2 = axis == 2 ? 1.0 : 0.0, n3 = axis == 3 ? 1.0 : 0.0;
    double pg = p[4];
    double vn = unit_dot(axis, p[1], p[2], p[3]);
    double md = 0.0 * 0.0;
    double u[5];
    orc_euler_conserved_density(p, gamma, u);
    f[0] = vn * u[0] + md;
    f[1] = vn * u[1] + pg * n1;
    f[2] = vn * u[2] + pg * n2;
    f[3] = vn * u[3] + pg * n3;
    f[4] = vn * u[4] + pg * vn;
}

/* :118-128 outer_wavespeeds (sound_speed_squared :118-121) */
static void euler_outer_wavespeeds(const double* p, int axis, double gamma, double* am, double* ap)
{
    double vn = unit_dot(axis, p[1], p[2], p[3]);
    double cs = sqrt(gamma * p[4] / p[0]);
    *am = vn - cs;
    *ap = vn + cs;
}

/* :209-222 riemann_hlle */
void orc_euler_riemann_hlle(const double* pl, const double* pr, int axis, double gamma, double* f)
{
    double ul[5], ur[5], fl[5], fr[5], alm, alp, arm, arp;
    orc_euler_conserved_density(pl, gamma, ul);
    orc_euler_conserved_density(pr, gamma, ur);
    orc_euler_flux(pl, axis, gamma, fl);
    orc_euler_flux(pr, axis, gamma, fr);
    euler_outer_wavespeeds(pl, axis, gamma, &alm, &alp);
    euler_outer_wavespeeds(pr, axis, gamma, &arm, &arp);
    double ap = smax(0.0, smax(alp, arp));
    double am = smin(0.0, smin(alm, arm));
    for (int k = 0; k < 5; ++k)
        f[k] = (ap * fl[k] - am * fr[k] - (ul[k] - ur[k]) * ap * am) / (ap - am);
}

/* HLLC for the Newtonian Euler equations.  The reference has no such solver
 * (only SRHD HLLC, src/physics_srhd.hpp:361-507); this follows Toro,
 * "Riemann Solvers and Numerical Methods for Fluid Dynamics", 3rd ed.,
 * section 10.4, eqs. (10.37)-(10.39) with the same outer (Davis) wave-speed
 * estimates as riemann_hlle above.  Parity unpinned against the reference;
 * pinned by invariants in tests (consistency, stationary contact, upwinding). */
void orc_euler_riemann_hllc(const double* pl, const double* pr, int axis, double gamma, double* f)
{
    double ul[5], ur[5], fl[5], fr[5], alm, alp, arm, arp;
    orc_euler_conserved_density(pl, gamma, ul);
    orc_euler_conserved_density(pr, gamma, ur);
    orc_euler_flux(pl, axis, gamma, fl);
    orc_euler_flux(pr, axis, gamma, fr);
    euler_outer_wavespeeds(pl, axis, gamma, &alm, &alp);
    euler_outer_wavespeeds(pr, axis, gamma, &arm, &arp);
    double sl = smin(alm, arm);
    double sr = smax(alp, arp);
    if (0.0 <= sl) { for (int k = 0; k < 5; ++k) f[k] = fl[k]; return; }
    if (sr <= 0.0) { for (int k = 0; k < 5; ++k) f[k] = fr[k]; return; }
    double dl = pl[0], dr = pr[0];
    double vl = pl[axis], vr = pr[axis];
    double ql = dl * (sl - vl), qr = dr * (sr - vr);
    double ss = (pr[4] - pl[4] + vl * ql - vr * qr) / (ql - qr);
    const double* p = (0.0 <= ss) ? pl : pr;
    const double* u = (0.0 <= ss) ? ul : ur;
    const double* fk = (0.0 <= ss) ? fl : fr;
    double s = (0.0 <= ss) ? sl : sr;
    double vk = p[axis];
    double fac = p[0] * (s - vk) / (s - ss);
    double us[5];
    us[0] = fac;
    us[1] = fac * p[1];
    us[2] = fac * p[2];
    us[3] = fac * p[3];
    us[axis] = fac * ss;
    us[4] = fac * (u[4] / p[0] + (ss - vk) * (ss + p[4] / (p[0] * (s - vk))));
    for (int k = 0; k < 5; ++k) f[k] = fk[k] + s * (us[k] - u[k]);
}

/* ------------------------------------------------------------------------ */
/* iso2d: src/physics_iso2d.hpp */

/* :119-125 */
void orc_iso2d_recover_primitive(const double* u, double* p)
{
    p[0] = u[0];
    p[1] = u[1] / u[0];
    p[2] = u[2] / u[0];
}

/* :129-135 */
void orc_iso2d_conserved_density(const double* p, double* u)
{
    u[0] = p[0];
    u[1] = p[0] * p[1];
    u[2] = p[0] * p[2];
}

/* :139-151 flux (velocity_vector has v3 = 0, :83-89; gas_pressure :93-96) */
static void iso2d_flux(const double* p, int axis, double cs2, double* f)
{
    double n1 = axis == 1 ? 1.0 : 0.0, n2 = axis == 2 ? 1.0 : 0.0;
    double pg = p[0] * cs2;
    double vn = unit_dot(axis, p[1], p[2], 0.0);
    double md = 0.0 * 0.0;
    double u[3];
    orc_iso2d_conserved_density(p, u);
    f[0] = vn * u[0] + md;
    f[1] = vn * u[1] + pg * n1;
    f[2] = vn * u[2] + pg * n2;
}

/* :156-173 riemann_hlle (outer_wavespeeds :100-105) */
void orc_iso2d_riemann_hlle(const double* pl, const double* pr, double cs2, int axis, double* f)
{
    double ul[3], ur[3], fl[3], fr[3];
    orc_iso2d_conserved_density(pl, ul);
    orc_iso2d_conserved_density(pr, ur);
    iso2d_flux(pl, axis, cs2, fl);
    iso2d_flux(pr, axis, cs2, fr);
    double cs = sqrt(cs2);
    double vnl = unit_dot(axis, pl[1], pl[2], 0.0);
    double vnr = unit_dot(axis, pr[1], pr[2], 0.0);
    double alm = vnl - cs, alp = vnl + cs;
    double arm = vnr - cs, arp = vnr + cs;
    double ap = smax(0.0, smax(alp, arp));
    double am = smin(0.0, smin(alm, arm));
    for (int k = 0; k < 3; ++k)
        f[k] = (ap * fl[k] - am * fr[k] - (ul[k] - ur[k]) * ap * am) / (ap - am);
}

/* ------------------------------------------------------------------------ */
/* two_body: src/physics_two_body.hpp (G = 1, :69) */

int orc_two_body_state(double a, double M, double q, double e, double t, double omega_frame, double* st)
{
    /* omega :81-86 */
    double W = sqrt(1.0 * M / a / a / a);
    /* eccentric_anomaly :109-127 (Newton until |f| <= 1e-12; bounded here) */
    double Man = W * t;
    double E = Man;
    int iters = 0;
    while (fabs(E - e * sin(E) - Man) > 1e-12 && iters < 200) {
        E -= (E - e * sin(E) - Man) / (1 - e * cos(E));
        ++iters;
    }
    /* orbital_state_from_eccentric_anomaly :129-153 */
    double mu = q / (1.0 + q);
    double M1 = M * (1.0 - mu);
    double M2 = M * mu;
    double x1 = -a * mu * (e - cos(E));
    double y1 = +a * mu * (0 + sin(E)) * sqrt(1 - e * e);
    double x2 = -x1 / q;
    double y2 = -y1 / q;
    double vx1 = -a * mu * W / (1 - e * cos(E)) * sin(E);
    double vy1 = +a * mu * W / (1 - e * cos(E)) * cos(E) * sqrt(1 - e * e);
    double vx2 = -vx1 / q;
    double vy2 = -vy1 / q;
    /* rotate :417-421 -> transform :377-403 with periapse argument = angle,
     * cm position/velocity 0; angle = -omega_frame * t (minidisk_scheme.cpp:335) */
    double angle = -omega_frame * t;
    double c = cos(-angle);
    double s = sin(-angle);
    st[0] = M1;
    st[1] = +x1 * c + y1 * s + 0.0;
    st[2] = -x1 * s + y1 * c + 0.0;
    st[3] = +vx1 * c + vy1 * s + 0.0;
    st[4] = -vx1 * s + vy1 * c + 0.0;
    st[5] = M2;
    st[6] = +x2 * c + y2 * s + 0.0;
    st[7] = -x2 * s + y2 * c + 0.0;
    st[8] = +vx2 * c + vy2 * s + 0.0;
    st[9] = -vx2 * s + vy2 * c + 0.0;
    return iters;
}

/* gravitational_acceleration :208-218 (pow<2> -> std::pow(x,2), pow<3,2> -> std::pow(x,1.5),
 * src/core_dimensional_math.hpp:49-53) */
void orc_two_body_acceleration(double mass, double px, double py, double x, double y, double rs, double* a)
{
    double dx = x - px;
    double dy = y - py;
    double r2 = dx * dx + dy * dy;
    double s2 = pow(rs, 2.0);
    a[0] = -1.0 * mass / pow(r2 + s2, 1.5) * dx;
    a[1] = -1.0 * mass / pow(r2 + s2, 1.5) * dy;
}

/* potential :198-205 with a unit point mass at the origin */
static double unit_potential(double x, double y, double rs)
{
    double dx = x - 0.0;
    double dy = y - 0.0;
    double r2 = dx * dx + dy * dy;
    double s2 = pow(rs, 2.0);
    return -1.0 * 1.0 / sqrt(r2 + s2);
}

/* ------------------------------------------------------------------------ */
/* Shared 2D / 1D update skeleton.  Index algebra of
 * src/module_euler.cpp:273-297 (same in src/module_locally_isothermal.cpp:192-209
 * and problems/minidisk_scheme.cpp:310-327, godunov fluxes :237-240):
 *   gx[i][j] = plm(pe[i][j], pe[i+1][j], pe[i+2][j])          (n+2) x (n+4)
 *   gy[i][j] = plm(pe[i][j], pe[i][j+1], pe[i][j+2])          (n+4) x (n+2)
 *   fx[i][j] = riemann(pe[i+1][j+2] + .5 gx[i][j+2], pe[i+2][j+2] - .5 gx[i+1][j+2])
 *   fy[i][j] = riemann(pe[i+2][j+1] + .5 gy[i+2][j], pe[i+2][j+2] - .5 gy[i+2][j+1])
 *   dfx = fx[i+1][j] - fx[i][j];  dfy = fy[i][j+1] - fy[i][j]                      */

typedef void (*face_flux_fn)(void* ctx, const double* pl, const double* pr, int axis, int i, int j, double* f);

/* Viscous stress of the isothermal schemes (nc = 3: Sigma, vx, vy), added to
 * the Godunov flux when visc = {nu, dl} is given:
 * godunov_and_viscous_fluxes, src/module_locally_isothermal.cpp:106-174 and
 * problems/minidisk_scheme.cpp:220-287.  With the reference's face states
 * pl = (p - g/2)[1:] (upper cell) and pr = (p + g/2)[:-1] (lower cell):
 *   mu = nu * 0.5 * (sigma(pl) + sigma(pr))
 *   x faces: dx_v = 0.5 (gx[i] + gx[i+1]) / dl (longitudinal slopes of the
 *            two cells), dy_v = 0.5 (gy[i+1][j+1] + gy[i+2][j+1]) / dl;
 *            F += (0, -mu (dx_vx - dy_vy), -mu (dx_vy + dy_vx))
 *   y faces: dy_v = 0.5 (gy[i+2][j] + gy[i+2][j+1]) / dl,
 *            dx_v = 0.5 (gx[i+1][j+1] + gx[i+1][j+2]) / dl;
 *            F += (0, -mu (dx_vy + dy_vx), -(-mu (dx_vx - dy_vy)))
 * (adjacent_mean, src/core_ndarray_ops.hpp:84-90: 0.5 * (a + b)). */
static void update2d_fluxes_visc(const double* pe, int ni, int nj, int nc, double theta,
                                 face_flux_fn flux, void* ctx, const double* visc, double* dfx, double* dfy);

static void update2d_fluxes(const double* pe, int ni, int nj, int nc, double theta,
                            face_flux_fn flux, void* ctx, double* dfx, double* dfy)
{
    update2d_fluxes_visc(pe, ni, nj, nc, theta, flux, ctx, NULL, dfx, dfy);
}

static void update2d_fluxes_visc(const double* pe, int ni, int nj, int nc, double theta,
                                 face_flux_fn flux, void* ctx, const double* visc, double* dfx, double* dfy)
{
    const int ei = ni + 4, ej = nj + 4;
#define PE(i, j, k) pe[((long)(i) * ej + (j)) * nc + (k)]
    double* gx = (double*)malloc(sizeof(double) * (ni + 2) * ej * nc);
    double* gy = (double*)malloc(sizeof(double) * ei * (nj + 2) * nc);
    double* fx = (double*)malloc(sizeof(double) * (ni + 1) * nj * nc);
    double* fy = (double*)malloc(sizeof(double) * ni * (nj + 1) * nc);
    for (int i = 0; i < ni + 2; ++i)
        for (int j = 0; j < ej; ++j)
            for (int k = 0; k < nc; ++k)
                gx[((long)i * ej + j) * nc + k] = orc_plm(PE(i, j, k), PE(i + 1, j, k), PE(i + 2, j, k), theta);
    for (int i = 0; i < ei; ++i)
        for (int j = 0; j < nj + 2; ++j)
            for (int k = 0; k < nc; ++k)
                gy[((long)i * (nj + 2) + j) * nc + k] = orc_plm(PE(i, j, k), PE(i, j + 1, k), PE(i, j + 2, k), theta);
    double pl[8], pr[8];
    for (int i = 0; i < ni + 1; ++i)
        for (int j = 0; j < nj; ++j) {
            for (int k = 0; k < nc; ++k) {
                pl[k] = PE(i + 1, j + 2, k) + 0.5 * gx[((long)i * ej + (j + 2)) * nc + k];
                pr[k] = PE(i + 2, j + 2, k) - 0.5 * gx[((long)(i + 1) * ej + (j + 2)) * nc + k];
            }
            double* f = &fx[((long)i * nj + j) * nc];
            flux(ctx, pl, pr, 1, i, j, f);
            if (visc) {
                const double nu = visc[0], dl = visc[1];
                const double* ga = &gx[((long)i * ej + (j + 2)) * nc];
                const double* gb = &gx[((long)(i + 1) * ej + (j + 2)) * nc];
                const double* ta = &gy[((long)(i + 1) * (nj + 2) + (j + 1)) * nc];
                const double* tb = &gy[((long)(i + 2) * (nj + 2) + (j + 1)) * nc];
                double mu = nu * 0.5 * (pr[0] + pl[0]);
                double dx_vx = 0.5 * (ga[1] + gb[1]) / dl;
                double dy_vx = 0.5 * (ta[1] + tb[1]) / dl;
                double dx_vy = 0.5 * (ga[2] + gb[2]) / dl;
                double dy_vy = 0.5 * (ta[2] + tb[2]) / dl;
                double tauxx = mu * (dx_vx - dy_vy);
                double tauxy = mu * (dx_vy + dy_vx);
                f[0] = f[0] + 0.0;
                f[1] = f[1] + -tauxx;
                f[2] = f[2] + -tauxy;
            }
        }
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj + 1; ++j) {
            for (int k = 0; k < nc; ++k) {
                pl[k] = PE(i + 2, j + 1, k) + 0.5 * gy[((long)(i + 2) * (nj + 2) + j) * nc + k];
                pr[k] = PE(i + 2, j + 2, k) - 0.5 * gy[((long)(i + 2) * (nj + 2) + (j + 1)) * nc + k];
            }
            double* f = &fy[((long)i * (nj + 1) + j) * nc];
            flux(ctx, pl, pr, 2, i, j, f);
            if (visc) {
                const double nu = visc[0], dl = visc[1];
                const double* ga = &gy[((long)(i + 2) * (nj + 2) + j) * nc];
                const double* gb = &gy[((long)(i + 2) * (nj + 2) + (j + 1)) * nc];
                const double* ta = &gx[((long)(i + 1) * ej + (j + 1)) * nc];
                const double* tb = &gx[((long)(i + 1) * ej + (j + 2)) * nc];
                double mu = nu * 0.5 * (pr[0] + pl[0]);
                double dx_vx = 0.5 * (ta[1] + tb[1]) / dl;
                double dy_vx = 0.5 * (ga[1] + gb[1]) / dl;
                double dx_vy = 0.5 * (ta[2] + tb[2]) / dl;
                double dy_vy = 0.5 * (ga[2] + gb[2]) / dl;
                double tauyx = mu * (dx_vy + dy_vx);
                double tauyy = -mu * (dx_vx - dy_vy);
                f[0] = f[0] + 0.0;
                f[1] = f[1] + -tauyx;
                f[2] = f[2] + -tauyy;
            }
        }
    for (int i = 0; i < ni; ++i)
        for (int j = 0; j < nj; ++j)
            for (int k = 0; k < nc; ++k) {
                dfx[((long)i * nj + j) * nc + k] = fx[((long)(i + 1) * nj + j) * nc + k] - fx[((long)i * nj + j) * nc + k];
                dfy[((long)i * nj + j) * nc + k] = fy[((long)i * (nj + 1) + j + 1) * nc + k] - fy[((long)i * (nj + 1) + j) * nc + k];
            }
#undef PE
    free(gx); free(gy); free(fx); free(fy);
}

typedef struct { double gamma; int hllc; } euler_ctx;

static void euler_face(void* c, const double* pl, const double* pr, int axis, int i, int j, double* f)
{
    (void)i; (void)j;
    euler_ctx* e = (euler_ctx*)c;
    if (e->hllc) orc_euler_riemann_hllc(pl, pr, axis, e->gamma, f);
    else         orc_euler_riemann_hlle(pl, pr, axis, e->gamma, f);
}

/* src/module_euler.cpp:263-298 euler2d::updated_conserved */
int orc_euler2d_updated_conserved(const double* uc, const double* pe, int ni, int nj,
                                  double dt, double dl, double theta, double gamma, int hllc, double* out)
{
    long nz = (long)ni * nj;
    double* dfx = (double*)malloc(sizeof(double) * nz * 5);
    double* dfy = (double*)malloc(sizeof(double) * nz * 5);
    euler_ctx ctx = { gamma, hllc };
    update2d_fluxes(pe, ni, nj, 5, theta, euler_face, &ctx, dfx, dfy);
    for (long z = 0; z < nz * 5; ++z)
        out[z] = uc[z] - (dfx[z] + dfy[z]) * dt / dl;
    free(dfx); free(dfy);
    return 0;
}

/* src/module_euler.cpp:107-135 euler1d::updated_conserved, with the outflow
 * boundary condition of src/mara.cpp:123-143 applied to pe first.  The
 * reference applies the lower branch when the block is first and returns
 * early, so a single-block mesh gets only the lower outflow (callers pass the
 * flags accordingly). */
int orc_euler1d_updated_conserved(const double* uc, const double* pe, int n,
                                  double dt, double dl, double theta, double gamma,
                                  int outflow_lower, int outflow_upper, double* out)
{
    const int nc = 5;
    double* pf = (double*)malloc(sizeof(double) * (n + 4) * nc);
    memcpy(pf, pe, sizeof(double) * (n + 4) * nc);
    if (outflow_lower) {
        for (int k = 0; k < nc; ++k) { pf[0 * nc + k] = pe[2 * nc + k]; pf[1 * nc + k] = pe[2 * nc + k]; }
    } else if (outflow_upper) {
        for (int k = 0; k < nc; ++k) { pf[(n + 2) * nc + k] = pe[(n + 1) * nc + k]; pf[(n + 3) * nc + k] = pe[(n + 1) * nc + k]; }
    }
    double* gx = (double*)malloc(sizeof(double) * (n + 2) * nc);
    double* fx = (double*)malloc(sizeof(double) * (n + 1) * nc);
    for (int i = 0; i < n + 2; ++i)
        for (int k = 0; k < nc; ++k)
            gx[i * nc + k] = orc_plm(pf[i * nc + k], pf[(i + 1) * nc + k], pf[(i + 2) * nc + k], theta);
    double pl[5], pr[5];
    for (int i = 0; i < n + 1; ++i) {
        for (int k = 0; k < nc; ++k) {
            pl[k] = pf[(i + 1) * nc + k] + 0.5 * gx[i * nc + k];
            pr[k] = pf[(i + 2) * nc + k] - 0.5 * gx[(i + 1) * nc + k];
        }
        orc_euler_riemann_hlle(pl, pr, 1, gamma, &fx[i * nc]);
    }
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < nc; ++k)
            out[i * nc + k] = uc[i * nc + k] - (fx[(i + 1) * nc + k] - fx[i * nc + k]) * dt / dl;
    free(pf); free(gx); free(fx);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* locally isothermal: src/module_locally_isothermal.cpp:182-212 (godunov
 * and viscous fluxes :106-174); the temperature mapping is supplied as face
 * arrays cs2_x[(ni+1)][nj], cs2_y[ni][(nj+1)] evaluated at
 * mesh_geometry.face_coordinates (the mapping is a host std::function). */
typedef struct { const double* cs2x; const double* cs2y; int nj; } iso_face_ctx;

static void iso_face_table(void* c, const double* pl, const double* pr, int axis, int i, int j, double* f)
{
    iso_face_ctx* t = (iso_face_ctx*)c;
    double cs2 = axis == 1 ? t->cs2x[(long)i * t->nj + j] : t->cs2y[(long)i * (t->nj + 1) + j];
    orc_iso2d_riemann_hlle(pl, pr, cs2, axis, f);
}

int orc_locally_isothermal_updated_conserved(const double* uc, const double* pe, int ni, int nj,
                                  double dt, double dl, const double* cs2_x, const double* cs2_y, double nu, double* out)
{
    long nz = (long)ni * nj;
    double* dfx = (double*)malloc(sizeof(double) * nz * 3);
    double* dfy = (double*)malloc(sizeof(double) * nz * 3);
    iso_face_ctx ctx = { cs2_x, cs2_y, nj };
    const double visc[2] = { nu, dl };
    update2d_fluxes_visc(pe, ni, nj, 3, 2.0, iso_face_table, &ctx, nu == 0.0 ? NULL : visc, dfx, dfy);
    for (long z = 0; z < nz * 3; ++z)
        out[z] = uc[z] - (dfx[z] + dfy[z]) * dt / dl;
    free(dfx); free(dfy);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* minidisk: problems/minidisk_scheme.cpp */

/* vertex_positions :101-116 and nd::linspace (src/core_ndarray.hpp:583-588);
 * cell centres via adjacent_mean (src/core_ndarray_ops.hpp:84-90). */
static double md_vertex(double R, int n, int level, long b, int i)
{
    double dlev = (double)(1 << level);
    double x0 = R * (-1 + 2 * b / dlev);
    double x1 = R * (-1 + 2 * (b + 1) / dlev);
    return x0 + (x1 - x0) * (double)i / (double)(n + 1 - 1);
}

void orc_minidisk_cell_coordinates(const orc_minidisk_t* sd, int level, long bi, long bj, double* xc, double* yc)
{
    int n = sd->block_size;
    for (int i = 0; i < n; ++i) {
        xc[i] = 0.5 * (md_vertex(sd->domain_radius, n, level, bi, i) + md_vertex(sd->domain_radius, n, level, bi, i + 1));
        yc[i] = 0.5 * (md_vertex(sd->domain_radius, n, level, bj, i) + md_vertex(sd->domain_radius, n, level, bj, i + 1));
    }
}

/* sound_speed_squared :127-132 */
static double md_cs2(double x, double y, double rs, double mach)
{
    double ph = unit_potential(x, y, rs);
    return -ph / mach / mach;
}

/* initial_primitive :87-99 + iso2d::conserved_density, initial_conserved_array :182-191 */
void orc_minidisk_initial_conserved(const orc_minidisk_t* sd, int level, long bi, long bj, double* out)
{
    int n = sd->block_size;
    double* xc = (double*)malloc(sizeof(double) * n);
    double* yc = (double*)malloc(sizeof(double) * n);
    orc_minidisk_cell_coordinates(sd, level, bi, bj, xc, yc);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double x = xc[i], y = yc[j];
            double ph = unit_potential(x, y, sd->softening_length);
            double r0 = sqrt(x * x + y * y);
            double vp = sqrt(-ph) - sd->omega_frame * r0;
            double vx = vp * (-y / r0);
            double vy = vp * ( x / r0);
            double p[3] = { 1.0, vx, vy };
            orc_iso2d_conserved_density(p, &out[((long)i * n + j) * 3]);
        }
    free(xc); free(yc);
}

typedef struct { double R, rs, mach; int n, level; long bi, bj; } md_face_ctx;

static void md_face(void* c, const double* pl, const double* pr, int axis, int i, int j, double* f)
{
    md_face_ctx* m = (md_face_ctx*)c;
    /* face_coordinates :134-151: axis 0 -> (xv[i], yc[j]); axis 1 -> (xc[i], yv[j]) */
    double x, y;
    if (axis == 1) {
        x = md_vertex(m->R, m->n, m->level, m->bi, i);
        y = 0.5 * (md_vertex(m->R, m->n, m->level, m->bj, j) + md_vertex(m->R, m->n, m->level, m->bj, j + 1));
    } else {
        x = 0.5 * (md_vertex(m->R, m->n, m->level, m->bi, i) + md_vertex(m->R, m->n, m->level, m->bi, i + 1));
        y = md_vertex(m->R, m->n, m->level, m->bj, j);
    }
    orc_iso2d_riemann_hlle(pl, pr, md_cs2(x, y, m->rs, m->mach), axis, f);
}

/* updated_conserved :294-355 (theta = 2 hard-coded at :313-314; viscous
 * fluxes :243-287 when kinematic_viscosity != 0) */
int orc_minidisk_updated_conserved(const double* uc, const double* pe, const orc_minidisk_t* sd,
                                   int level, long bi, long bj, double time, double dt, double* out)
{
    const int n = sd->block_size, nc = 3;
    const long nz = (long)n * n;
    double* dfx = (double*)malloc(sizeof(double) * nz * nc);
    double* dfy = (double*)malloc(sizeof(double) * nz * nc);
    double* xc = (double*)malloc(sizeof(double) * n);
    double* yc = (double*)malloc(sizeof(double) * n);
    double* ui = (double*)malloc(sizeof(double) * nz * nc);
    md_face_ctx ctx = { sd->domain_radius, sd->softening_length, sd->mach_number, n, level, bi, bj };
    /* cell_size :78-81 */
    double dl = 2.0 * sd->domain_radius / (double)sd->block_size / (double)(1 << level);
    const double visc[2] = { sd->kinematic_viscosity, dl };
    update2d_fluxes_visc(pe, n, n, nc, 2.0, md_face, &ctx, sd->kinematic_viscosity == 0.0 ? NULL : visc, dfx, dfy);
    orc_minidisk_cell_coordinates(sd, level, bi, bj, xc, yc);
    orc_minidisk_initial_conserved(sd, level, bi, bj, ui);

    double st[10];
    orc_two_body_state(1.0, 1.0, sd->mass_ratio, sd->eccentricity, time, sd->omega_frame, st);
    double Om = sd->omega_frame;
    double s6 = pow(sd->sink_radius, 6.0);

    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            long z = (long)i * n + j;
            const double* pc = &pe[((long)(i + 2) * (n + 4) + (j + 2)) * nc];
            const double* u = &uc[z * nc];
            double x = xc[i], y = yc[j];
            double ag1[2], ag2[2];
            orc_two_body_acceleration(st[0], st[1], st[2], x, y, sd->softening_length, ag1);
            orc_two_body_acceleration(st[5], st[6], st[7], x, y, sd->softening_length, ag2);
            /* coriolis_term :67-71, centrifugal_term :73-76 */
            double cor[2] = { 2.0 * Om * pc[2], 2.0 * Om * (-pc[1]) };
            double cen[2] = { Om * Om * x, Om * Om * y };
            /* sink_rate_field :57-65 (pow<3>, pow<6> -> std::pow) */
            double d1x = x - st[1], d1y = y - st[2];
            double d2x = x - st[6], d2y = y - st[7];
            double r61 = pow((0.0 + d1x * d1x) + d1y * d1y, 3.0);
            double r62 = pow((0.0 + d2x * d2x) + d2y * d2y, 3.0);
            double sr1 = sd->sink_rate * exp(-r61 / s6);
            double sr2 = sd->sink_rate * exp(-r62 / s6);
            /* buffer_rate_field :47-55 */
            double r = sqrt((0.0 + x * x) + y * y);
            double yb = (r - sd->domain_radius) / sd->buffer_scale;
            double br = 0.5 * sd->buffer_rate * (1.0 + tanh(yb));
            double sig = pc[0];
            /* acceleration_to_source_terms (physics_iso2d.hpp:176-181): (0, s ax, s ay) */
            double sg1[3] = { 0.0 / 1.0, sig * ag1[0], sig * ag1[1] };
            double sg2[3] = { 0.0 / 1.0, sig * ag2[0], sig * ag2[1] };
            double sf1[3] = { 0.0 / 1.0, sig * cen[0], sig * cen[1] };
            double sf2[3] = { 0.0 / 1.0, sig * cor[0], sig * cor[1] };
            for (int k = 0; k < nc; ++k) {
                double ss1 = -u[k] * sr1;
                double ss2 = -u[k] * sr2;
                double sb = -(u[k] - ui[z * nc + k]) * br;
                double src = sg1[k] + sg2[k] + ss1 + ss2 + sf1[k] + sf2[k] + sb;
                out[z * nc + k] = u[k] - (dfx[z * nc + k] + dfy[z * nc + k]) * dt / dl + src * dt;
            }
        }
    free(dfx); free(dfy); free(xc); free(yc); free(ui);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Guard fill: amr::extend_block (src/mesh_amr.hpp:334-393) over same-level
 * neighbours with the periodic next_on/prev_on (src/mesh_block_index.hpp:334-335).
 * Corner blocks as the reference assembles them (:361-379, and the same in
 * problems/minidisk.cpp:87-120): c02 = prev_on(next_on(b, 0), 1) fills the
 * (x-prev, y-next) corner and c20 = next_on(prev_on(b, 0), 1) the (x-next,
 * y-prev) corner, i.e. those two corners come from the diagonally opposite
 * block.  Only viscous updates read corner zones. */
static inline long wrapb(long b, int nb) { return (b + nb) % nb; }

void orc_extend_block_2d(const double* P, int nb, int n, int nc, long bi, long bj, double* pe)
{
    const int e = n + 4;
    for (int i = 0; i < e; ++i)
        for (int j = 0; j < e; ++j) {
            int ci = i < 2 ? 0 : (i >= n + 2 ? 2 : 1);
            int cj = j < 2 ? 0 : (j >= n + 2 ? 2 : 1);
            int ii = i < 2 ? i - 2 + n : (i >= n + 2 ? i - 2 - n : i - 2);
            int jj = j < 2 ? j - 2 + n : (j >= n + 2 ? j - 2 - n : j - 2);
            int oi = ci - 1, oj = cj - 1;
            if (oi != 0 && oj != 0 && oi != oj) { oi = -oi; oj = -oj; }   /* c02 / c20 */
            long sb = wrapb(bi + oi, nb) * nb + wrapb(bj + oj, nb);
            memcpy(&pe[((long)i * e + j) * nc], &P[((sb * n + ii) * (long)n + jj) * nc], sizeof(double) * nc);
        }
}

void orc_extend_block_1d(const double* P, int nb, int n, int nc, long bi, double* pe)
{
    for (int i = 0; i < n + 4; ++i) {
        int ci = i < 2 ? 0 : (i >= n + 2 ? 2 : 1);
        int ii = i < 2 ? i - 2 + n : (i >= n + 2 ? i - 2 - n : i - 2);
        long sb = wrapb(bi + ci - 1, nb);
        memcpy(&pe[(long)i * nc], &P[(sb * n + ii) * nc], sizeof(double) * nc);
    }
}

/* ------------------------------------------------------------------------ */
/* Mesh-level steps.  One stage = recover primitives of every block (node "P",
 * src/module_euler.cpp:240-245) -> extend (node "Pe") -> updated_conserved
 * (node "U"), as euler2d::updated_solution (src/module_euler.cpp:304-323).
 * RK2 = control::advance_runge_kutta (src/app_control.hpp:169-195) with
 * weighted_sum_tree s*b + t*(1-b), b = 1/2 (src/mesh_amr.hpp:455-467). */

static int nthreads_or_default(int n) { return n > 0 ? n : 1; }

int orc_euler2d_mesh_stage(const double* u, double* out, int nb, int n, double dt, double dl, double theta,
                           double gamma, int hllc, int nthreads)
{
    const long nblk = (long)nb * nb, bz = (long)n * n * 5;
    double* P = (double*)malloc(sizeof(double) * nblk * bz);
    int status = 0;
    #pragma omp parallel for num_threads(nthreads_or_default(nthreads)) schedule(static)
    for (long z = 0; z < nblk * n * (long)n; ++z)
        if (orc_euler_recover_primitive(&u[z * 5], gamma, &P[z * 5])) status = -1;
    if (status) { free(P); return status; }
    #pragma omp parallel num_threads(nthreads_or_default(nthreads))
    {
        double* pe = (double*)malloc(sizeof(double) * (n + 4) * (n + 4) * 5);
        #pragma omp for schedule(dynamic, 1)
        for (long b = 0; b < nblk; ++b) {
            orc_extend_block_2d(P, nb, n, 5, b / nb, b % nb, pe);
            orc_euler2d_updated_conserved(&u[b * bz], pe, n, n, dt, dl, theta, gamma, hllc, &out[b * bz]);
        }
        free(pe);
    }
    free(P);
    return 0;
}

int orc_euler2d_mesh_rk2(double* u, int nb, int n, double dt, double dl, double theta, double gamma,
                         int hllc, int nsteps, int nthreads)
{
    const long tot = (long)nb * nb * n * n * 5;
    double* u1 = (double*)malloc(sizeof(double) * tot);
    double* u2 = (double*)malloc(sizeof(double) * tot);
    int st = 0;
    for (int s = 0; s < nsteps && !st; ++s) {
        st = orc_euler2d_mesh_stage(u, u1, nb, n, dt, dl, theta, gamma, hllc, nthreads);
        if (!st) st = orc_euler2d_mesh_stage(u1, u2, nb, n, dt, dl, theta, gamma, hllc, nthreads);
        if (!st) {
            #pragma omp parallel for num_threads(nthreads_or_default(nthreads)) schedule(static)
            for (long z = 0; z < tot; ++z) u[z] = u[z] * 0.5 + u2[z] * (1.0 - 0.5);
        }
    }
    free(u1); free(u2);
    return st;
}

static int euler1d_mesh_stage(const double* u, double* out, int nb, int n, double dt, double dl,
                              double theta, double gamma, int nthreads)
{
    const long bz = (long)n * 5;
    double* P = (double*)malloc(sizeof(double) * nb * bz);
    int status = 0;
    for (long z = 0; z < (long)nb * n; ++z)
        if (orc_euler_recover_primitive(&u[z * 5], gamma, &P[z * 5])) status = -1;
    if (status) { free(P); return status; }
    #pragma omp parallel num_threads(nthreads_or_default(nthreads))
    {
        double* pe = (double*)malloc(sizeof(double) * (n + 4) * 5);
        #pragma omp for schedule(dynamic, 1)
        for (long b = 0; b < nb; ++b) {
            orc_extend_block_1d(P, nb, n, 5, b, pe);
            int lo = (b == 0);
            int hi = (b == nb - 1);
            orc_euler1d_updated_conserved(&u[b * bz], pe, n, dt, dl, theta, gamma, lo, hi, &out[b * bz]);
        }
        free(pe);
    }
    free(P);
    return 0;
}

int orc_euler1d_mesh_rk2(double* u, int nb, int n, double dt, double dl, double theta, double gamma,
                         int nsteps, int nthreads)
{
    const long tot = (long)nb * n * 5;
    double* u1 = (double*)malloc(sizeof(double) * tot);
    double* u2 = (double*)malloc(sizeof(double) * tot);
    int st = 0;
    for (int s = 0; s < nsteps && !st; ++s) {
        st = euler1d_mesh_stage(u, u1, nb, n, dt, dl, theta, gamma, nthreads);
        if (!st) st = euler1d_mesh_stage(u1, u2, nb, n, dt, dl, theta, gamma, nthreads);
        if (!st) for (long z = 0; z < tot; ++z) u[z] = u[z] * 0.5 + u2[z] * (1.0 - 0.5);
    }
    free(u1); free(u2);
    return st;
}

int orc_minidisk_mesh_stage(const double* u, double* out, const orc_minidisk_t* sd, double time, double dt, int nthreads)
{
    const int nb = 1 << sd->depth, n = sd->block_size;
    const long nblk = (long)nb * nb, bz = (long)n * n * 3;
    double* P = (double*)malloc(sizeof(double) * nblk * bz);
    #pragma omp parallel for num_threads(nthreads_or_default(nthreads)) schedule(static)
    for (long z = 0; z < nblk * n * (long)n; ++z)
        orc_iso2d_recover_primitive(&u[z * 3], &P[z * 3]);
    #pragma omp parallel num_threads(nthreads_or_default(nthreads))
    {
        double* pe = (double*)malloc(sizeof(double) * (n + 4) * (n + 4) * 3);
        #pragma omp for schedule(dynamic, 1)
        for (long b = 0; b < nblk; ++b) {
            orc_extend_block_2d(P, nb, n, 3, b / nb, b % nb, pe);
            orc_minidisk_updated_conserved(&u[b * bz], pe, sd, sd->depth, b / nb, b % nb, time, dt, &out[b * bz]);
        }
        free(pe);
    }
    free(P);
    return 0;
}

int orc_minidisk_mesh_rk2(double* u, const orc_minidisk_t* sd, double t0, double dt, int nsteps, int nthreads)
{
    const int nb = 1 << sd->depth, n = sd->block_size;
    const long tot = (long)nb * nb * n * n * 3;
    double* u1 = (double*)malloc(sizeof(double) * tot);
    double* u2 = (double*)malloc(sizeof(double) * tot);
    double t = t0;
    for (int s = 0; s < nsteps; ++s) {
        orc_minidisk_mesh_stage(u, u1, sd, t, dt, nthreads);
        orc_minidisk_mesh_stage(u1, u2, sd, t + dt, dt, nthreads);
        #pragma omp parallel for num_threads(nthreads_or_default(nthreads)) schedule(static)
        for (long z = 0; z < tot; ++z) u[z] = u[z] * 0.5 + u2[z] * (1.0 - 0.5);
        t = t * 0.5 + (t + dt) * (1.0 - 0.5);
    }
    free(u1); free(u2);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Exact Riemann solver for the 1D Euler equations (Toro, ch. 4: pressure
 * function, Newton iteration for p*, sampling).  Independent of the
 * reference; used only to check Sod convergence. p = (rho, u, pressure). */
static void toro_prefun(double p, double dk, double pk, double ck, double g, double* f, double* fd)
{
    if (p <= pk) {
        double prat = p / pk;
        *f = 2.0 / (g - 1.0) * ck * (pow(prat, (g - 1.0) / (2.0 * g)) - 1.0);
        *fd = (1.0 / (dk * ck)) * pow(prat, -(g + 1.0) / (2.0 * g));
    } else {
        double ak = 2.0 / ((g + 1.0) * dk);
        double bk = (g - 1.0) / (g + 1.0) * pk;
        double qrt = sqrt(ak / (bk + p));
        *f = (p - pk) * qrt;
        *fd = (1.0 - 0.5 * (p - pk) / (bk + p)) * qrt;
    }
}

void orc_euler_exact_riemann(const double* pl, const double* pr, double g, double xi, double* out)
{
    double dl = pl[0], ul = pl[1], pL = pl[2];
    double dr = pr[0], ur = pr[1], pR = pr[2];
    double cl = sqrt(g * pL / dl), cr = sqrt(g * pR / dr);
    double p = 0.5 * (pL + pR), fl, fld, fr, frd;
    for (int it = 0; it < 100; ++it) {
        toro_prefun(p, dl, pL, cl, g, &fl, &fld);
        toro_prefun(p, dr, pR, cr, g, &fr, &frd);
        double pn = p - (fl + fr + ur - ul) / (fld + frd);
        if (pn < 1e-14) pn = 1e-14;
        if (fabs(pn - p) / (0.5 * (pn + p)) < 1e-15) { p = pn; break; }
        p = pn;
    }
    toro_prefun(p, dl, pL, cl, g, &fl, &fld);
    toro_prefun(p, dr, pR, cr, g, &fr, &frd);
    double us = 0.5 * (ul + ur) + 0.5 * (fr - fl);
    double d, u, pp;
    if (xi <= us) {
        if (p <= pL) {
            double shl = ul - cl, cml = cl * pow(p / pL, (g - 1.0) / (2.0 * g)), stl = us - cml;
            if (xi <= shl) { d = dl; u = ul; pp = pL; }
            else if (xi >= stl) { d = dl * pow(p / pL, 1.0 / g); u = us; pp = p; }
            else {
                u = 2.0 / (g + 1.0) * (cl + (g - 1.0) / 2.0 * ul + xi);
                double c = 2.0 / (g + 1.0) * (cl + (g - 1.0) / 2.0 * (ul - xi));
                d = dl * pow(c / cl, 2.0 / (g - 1.0));
                pp = pL * pow(c / cl, 2.0 * g / (g - 1.0));
            }
        } else {
            double pml = p / pL;
            double sl = ul - cl * sqrt((g + 1.0) / (2.0 * g) * pml + (g - 1.0) / (2.0 * g));
            if (xi <= sl) { d = dl; u = ul; pp = pL; }
            else { d = dl * (pml + (g - 1.0) / (g + 1.0)) / (pml * (g - 1.0) / (g + 1.0) + 1.0); u = us; pp = p; }
        }
    } else {
        if (p > pR) {
            double pmr = p / pR;
            double sr = ur + cr * sqrt((g + 1.0) / (2.0 * g) * pmr + (g - 1.0) / (2.0 * g));
            if (xi >= sr) { d = dr; u = ur; pp = pR; }
            else { d = dr * (pmr + (g - 1.0) / (g + 1.0)) / (pmr * (g - 1.0) / (g + 1.0) + 1.0); u = us; pp = p; }
        } else {
            double shr = ur + cr, cmr = cr * pow(p / pR, (g - 1.0) / (2.0 * g)), str = us + cmr;
            if (xi >= shr) { d = dr; u = ur; pp = pR; }
            else if (xi <= str) { d = dr * pow(p / pR, 1.0 / g); u = us; pp = p; }
            else {
                u = 2.0 / (g + 1.0) * (-cr + (g - 1.0) / 2.0 * ur + xi);
                double c = 2.0 / (g + 1.0) * (cr - (g - 1.0) / 2.0 * (ur - xi));
                d = dr * pow(c / cr, 2.0 / (g - 1.0));
                pp = pR * pow(c / cr, 2.0 * g / (g - 1.0));
            }
        }
    }
    out[0] = d; out[1] = u; out[2] = pp;
}
