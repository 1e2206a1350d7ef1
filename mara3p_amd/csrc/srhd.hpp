// Special-relativistic hydrodynamics, src/physics_srhd.hpp (c = 1), in FP64
// (the Newton primitive recovery stays FP64 in every precision mode).
// Same operations and order as the reference, so results match it exactly.
// Primitive (rho, u1, u2, u3, p) with u = gamma beta; conserved
// (D, S1, S2, S3, tau), tau = E - D.
#pragma once
#include "device_common.hpp"

namespace m3p {


/* gamma_beta_squared :123-129, lorentz_factor :131-139 */
__device__ __forceinline__ double gbs(const double* p) { return p[1] * p[1] + p[2] * p[2] + p[3] * p[3]; }
__device__ __forceinline__ double lorentz(const double* p) { return sqrt(1.0 + gbs(p)); }
/* enthalpy_density :164-168 */
__device__ __forceinline__ double enthalpy_density(const double* p, double g) { return p[0] * (1.0 * 1.0) + p[4] * (1.0 + 1.0 / (g - 1.0)); }

/* velocity_vector :149-158 */
__device__ __forceinline__ void velocity(const double* p, double* v)
{
    double W = lorentz(p);
    v[0] = 1.0 * p[1] / W;
    v[1] = 1.0 * p[2] / W;
    v[2] = 1.0 * p[3] / W;
}

__device__ __forceinline__ double ndot(int axis, const double* v)
{
    double n1 = axis == 1, n2 = axis == 2, n3 = axis == 3;
    return ((0.0 + n1 * v[0]) + n2 * v[1]) + n3 * v[2];
}

/* recover_primitive :224-279.  Returns 0, 3 (no convergence) or 2 (negative pressure). */
__device__ __forceinline__ int srhd_recover_primitive(const double* u, double gamma, double temperature_floor, double* out)
{
    const int newton_iter_max = 50;
    const double error_tolerance = 1e-12;
    const double c2 = 1.0 * 1.0;
    const double gm = gamma;
    double D = u[0], tau = u[4];
    double SS = ((0.0 + u[1] * u[1]) + u[2] * u[2]) + u[3] * u[3];
    int found = 0, iteration = 0;
    double W0 = 1.0, p = 0.0;
    while (iteration <= newton_iter_max) {
        double Et = tau + p + D * c2;
        double b2 = smin<double>(SS * c2 / Et / Et, 1.0 - 1e-10);
        double W2 = 1.0 / (1.0 - b2);
        double W = sqrt(W2);
        double e = (tau + D * c2 * (1.0 - W) + p * (1.0 - W2)) / (D * W);
        double d = D / W;
        double h = c2 + e + p / d;
        double a2 = gm * p / (d * h);
        double f = d * e * (gm - 1.0) - p;
        double g = b2 * a2 - 1.0;
        p -= f / g;
        if (fabs(f) < error_tolerance || (fabs(f) < error_tolerance * 1e4 && iteration == newton_iter_max)) {
            W0 = W;
            found = 1;
            break;
        }
        ++iteration;
    }
    if (temperature_floor > 0.0) p = smax<double>(p, temperature_floor * D * c2);
    if (!found) return 3;
    if (temperature_floor <= 0.0 && p <= 0.0) return 2;
    out[0] = D / W0;
    out[1] = W0 * u[1] / (tau + D * c2 + p) * 1.0;
    out[2] = W0 * u[2] / (tau + D * c2 + p) * 1.0;
    out[3] = W0 * u[3] / (tau + D * c2 + p) * 1.0;
    out[4] = p;
    return 0;
}

/* conserved_density :285-299 */
__device__ __forceinline__ void srhd_conserved_density(const double* p, double gamma, double* u)
{
    double c = 1.0;
    double W = lorentz(p);
    double D = p[0] * W;
    double h = enthalpy_density(p, gamma) / p[0];
    u[0] = D;
    u[1] = D * h * p[1] / c;
    u[2] = D * h * p[2] / c;
    u[3] = D * h * p[3] / c;
    u[4] = D * (h * W - c * c) - p[4];
}

/* flux :305-320 */
__device__ __forceinline__ void srhd_flux(const double* p, int axis, double gamma, double* f)
{
    double v[3], u[5];
    velocity(p, v);
    double n1 = axis == 1, n2 = axis == 2, n3 = axis == 3;
    double pg = p[4];
    double vn = ndot(axis, v);
    double md = 0.0 * 0.0;
    srhd_conserved_density(p, gamma, u);
    double pt[5] = { md, pg * n1, pg * n2, pg * n3, pg * vn };
    for (int k = 0; k < 5; ++k) f[k] = vn * u[k] + pt[k];
}

/* outer_wavespeeds :177-192 with sound_speed_squared :170-175 */
__device__ __forceinline__ void srhd_wavespeeds(const double* p, int axis, double gamma, double* wm, double* wp)
{
    double c2 = 1.0 * 1.0;
    double a2 = gamma * c2 * p[4] / enthalpy_density(p, gamma) / c2;
    double uu = gbs(p);
    double v[3];
    velocity(p, v);
    double vn = ndot(axis, v);
    double bn = vn / 1.0;
    double bb = uu / (1 + uu);
    double b2 = bn * bn;
    double k0 = sqrt(a2 * (1 - bb) * (1 - bb * a2 - b2 * (1 - a2))) * 1.0;
    *wm = (vn * (1 - a2) - k0) / (1 - bb * a2);
    *wp = (vn * (1 - a2) + k0) / (1 - bb * a2);
}

/* riemann_solver (Mignone & Bodo 2005), :377-496.  mode 0 = HLLE fluxes,
 * 1 = HLLC fluxes (face at rest), 2 = contact speed (in f[0]), 3 = HLLC
 * fluxes through a face moving at face_speed.  Returns 0, or 3 when the
 * contact speed is not finite / out of range (the reference throws). */
__device__ __forceinline__ int srhd_riemann(const double* pl, const double* pr, int axis, double gamma, int mode, double face_speed, double* f)
{
    double ul[5], ur[5], fl[5], fr[5], alm, alp, arm, arp;
    srhd_conserved_density(pl, gamma, ul);
    srhd_conserved_density(pr, gamma, ur);
    srhd_flux(pl, axis, gamma, fl);
    srhd_flux(pr, axis, gamma, fr);
    srhd_wavespeeds(pl, axis, gamma, &alm, &alp);
    srhd_wavespeeds(pr, axis, gamma, &arm, &arp);
    double ar = smax<double>(alp, arp);
    double al = smin<double>(alm, arm);
    if (mode == 0) {
        if (0.0 < al) { for (int k = 0; k < 5; ++k) f[k] = fl[k]; return 0; }
        if (0.0 > ar) { for (int k = 0; k < 5; ++k) f[k] = fr[k]; return 0; }
        for (int k = 0; k < 5; ++k) f[k] = (ar * fl[k] - al * fr[k] - (ul[k] - ur[k]) * ar * al) / (ar - al);
        return 0;
    }
    double uh[5], fh[5];
    for (int k = 0; k < 5; ++k) {
        uh[k] = (ar * ur[k] - al * ul[k] + (fl[k] - fr[k])) / (ar - al);
        fh[k] = (ar * fl[k] - al * fr[k] - (ul[k] - ur[k]) * ar * al) / (ar - al);
    }
    double c2 = 1.0 * 1.0;
    double ue_hll = uh[4] + uh[0] * c2;
    double fe_hll = fh[4] + fh[0] * c2;
    double um_hll = ndot(axis, &uh[1]);
    double fm_hll = ndot(axis, &fh[1]);
    double a = +fe_hll;
    double b = -fm_hll - ue_hll;
    double c = +um_hll;
    double d = b * b - 4.0 * a * c;
    double as = d < 0.0 ? 0.0 : (fabs(a) > 1e-8 ? (-b - sqrt(d)) / (2.0 * a) : -c / b) * c2;
    if (isnan(as) || as < -1e3 || as > 1e3) return 3;
    if (mode == 2) { f[0] = as; return 0; }
    double ps = -fe_hll * as / c2 + fm_hll;
    const double* u; const double* fk; const double* p; double s;
    double vf = mode == 3 ? face_speed : 0.0;
    if (mode == 1) {
        if (0.0 <= al) { for (int k = 0; k < 5; ++k) f[k] = fl[k]; return 0; }
        if (0.0 >= ar) { for (int k = 0; k < 5; ++k) f[k] = fr[k]; return 0; }
    } else {
        if (vf <= al) { for (int k = 0; k < 5; ++k) f[k] = fl[k] - vf * ul[k]; return 0; }
        if (vf >= ar) { for (int k = 0; k < 5; ++k) f[k] = fr[k] - vf * ur[k]; return 0; }
    }
    if (vf <= as) { u = ul; fk = fl; p = pl; s = al; }
    else          { u = ur; fk = fr; p = pr; s = ar; }
    /* star_state_flux, equations (16) */
    double n[3] = { double(axis == 1), double(axis == 2), double(axis == 3) };
    double S[3] = { u[1], u[2], u[3] };
    double vv[3];
    velocity(p, vv);
    double v = ((0.0 + vv[0] * n[0]) + vv[1] * n[1]) + vv[2] * n[2];
    double e = u[4] + u[0] * c2;
    double m = ((0.0 + S[0] * n[0]) + S[1] * n[1]) + S[2] * n[2];
    double nt[3] = { S[0] - m * n[0], S[1] - m * n[1], S[2] - m * n[2] };
    double es = (e * (s - v) + ps * as - p[4] * v) / (s - as);
    double ms = (m * (s - v) + ps - p[4]) / (s - as);
    double ns[3] = { nt[0] * (s - v) / (s - as), nt[1] * (s - v) / (s - as), nt[2] * (s - v) / (s - as) };
    double Ds = u[0] * (s - v) / (s - as);
    double Ss[3] = { ms * n[0] + ns[0], ms * n[1] + ns[1], ms * n[2] + ns[2] };
    double us[5] = { Ds, Ss[0], Ss[1], Ss[2], es - Ds * c2 };
    for (int k = 0; k < 5; ++k) f[k] = (fk[k] + s * (us[k] - u[k])) - vf * us[k];
    return 0;
}

__global__ void __launch_bounds__(256)
srhd_recover_array_kernel(const double* u, double* p, long nz, double gamma, double tfloor, int* err)
{
    long z = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= nz) return;
    double uu[5], pp[5] = { 0, 0, 0, 0, 0 };
    for (int q = 0; q < 5; ++q) uu[q] = u[q * nz + z];
    int st = srhd_recover_primitive(uu, gamma, tfloor, pp);
    if (st) report_error(err, st, 0, (int)z);
    for (int q = 0; q < 5; ++q) p[q * nz + z] = pp[q];
}

__global__ void __launch_bounds__(256)
srhd_conserved_array_kernel(const double* p, double* u, long nz, double gamma)
{
    long z = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= nz) return;
    double pp[5], uu[5];
    for (int q = 0; q < 5; ++q) pp[q] = p[q * nz + z];
    srhd_conserved_density(pp, gamma, uu);
    for (int q = 0; q < 5; ++q) u[q * nz + z] = uu[q];
}

__global__ void __launch_bounds__(256)
srhd_riemann_array_kernel(const double* pl, const double* pr, double* f, long nf, int axis, double gamma,
                          int mode, double face_speed, int* err)
{
    long z = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= nf) return;
    double a[5], b[5], o[5] = { 0, 0, 0, 0, 0 };
    for (int q = 0; q < 5; ++q) { a[q] = pl[q * nf + z]; b[q] = pr[q * nf + z]; }
    int st = srhd_riemann(a, b, axis, gamma, mode, face_speed, o);
    if (st) report_error(err, st, 0, (int)z);
    for (int q = 0; q < 5; ++q) f[q * nf + z] = o[q];
}

} // namespace m3p
