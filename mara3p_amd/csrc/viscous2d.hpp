// Isothermal 2D stages with a kinematic viscosity: the viscous stress of
// godunov_and_viscous_fluxes (src/module_locally_isothermal.cpp:106-174,
// problems/minidisk_scheme.cpp:220-287) added to the HLLE flux of every face.
//
// The stress at a face needs the PLM slopes of its two cells along both
// axes, i.e. a 3 x 3 block of primitives plus the cross of radius 2 around
// each zone, corners included.  One thread per zone computes the four face
// fluxes of its zone; a face shared by two zones is computed from the same
// operands in the same order by both, so the update stays conservative and
// the FP64 kernel reproduces the reference's arithmetic:
//   mu    = nu * 0.5 * (sigma_upper_state + sigma_lower_state)
//   x face: tau_xx = mu (dx vx - dy vy), tau_xy = mu (dx vy + dy vx),
//           F += (0, -tau_xx, -tau_xy)
//   y face: tau_yx = mu (dx vy + dy vx), tau_yy = -mu (dx vx - dy vy),
//           F += (0, -tau_yx, -tau_yy)
// with each derivative the mean of the two cells' full PLM slopes over dl
// (adjacent_mean, src/core_ndarray_ops.hpp:84-90).  This path is used only
// when nu != 0; the inviscid stages run the tile-march kernel.
#pragma once
#include "stage2d.hpp"

namespace m3p {

template<typename R>
struct ViscousTerms {
    // derivative from two full slopes: 0.5 (a + b) / dl
    __device__ static __forceinline__ R mean_over(R a, R b, double dl, R idl)
    {
        if constexpr (is_fp64<R>::value) return R(0.5) * (a + b) / dl;
        else return R(0.5) * (a + b) * idl;
    }
    // Adds the stress of a face of axis AXIS to f.  pl, pr: the face states
    // (lower cell + g/2, upper cell - g/2); gl*, gu*: full slopes of the lower
    // and upper cells, along (L) and across (T) the face normal.
    template<int AXIS>
    __device__ static __forceinline__ void add(R* f, const R* pl, const R* pr, const R* glL, const R* guL,
                                               const R* glT, const R* guT, double nu, double dl, R idl)
    {
        R mu = R(nu) * R(0.5) * (pr[0] + pl[0]);
        R dL_vx = mean_over(glL[1], guL[1], dl, idl);
        R dT_vx = mean_over(glT[1], guT[1], dl, idl);
        R dL_vy = mean_over(glL[2], guL[2], dl, idl);
        R dT_vy = mean_over(glT[2], guT[2], dl, idl);
        if constexpr (AXIS == 1) {          // L = x, T = y
            R tauxx = mu * (dL_vx - dT_vy);
            R tauxy = mu * (dL_vy + dT_vx);
            f[0] = f[0] + R(0);
            f[1] = f[1] + -tauxx;
            f[2] = f[2] + -tauxy;
        } else {                            // L = y, T = x
            R tauyx = mu * (dT_vy + dL_vx);
            R tauyy = -mu * (dT_vx - dL_vy);
            f[0] = f[0] + R(0);
            f[1] = f[1] + -tauyx;
            f[2] = f[2] + -tauyy;
        }
    }
};

template<class S, class Ld>
__global__ void __launch_bounds__(256)
iso_visc_stage_kernel(typename S::Params prm, Ld ld_arg, StageArgs a, double nu)
{
    using R = typename S::Real;
    constexpr int NC = 3;
    resolve_params(prm);
    const int n = a.n;
    const long zpb = (long)n * n;
    const long z = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= zpb * a.nblk) return;
    const int b = (int)(z / zpb);
    const int k = (int)(z % zpb / n), c = (int)(z % n);
    Ld ld = ld_arg;
    ld.bind(b);
    TileCtx tc;
    tc.b = b; tc.k0 = k; tc.c0 = c; tc.n = n;
    if (a.bcoord) { tc.bi = a.bcoord[2 * b]; tc.bj = a.bcoord[2 * b + 1]; }
    else          { tc.bi = a.bi0; tc.bj = a.bj0; }
    typename S::template TileData<1, 1> td;
    S::setup_zone(td, prm, tc);
    const double dt = a.dt_dev ? *a.dt_dev : a.dt;
    const double dl = a.dl_blk ? a.dl_blk[b] : a.dl;
    const R idl = R(1.0 / dl);
    const R theta = R(S::theta(prm));
    bool ok = true;

    // primitives P[dr + 2][dc + 2] on the 3 x 3 block and the radius-2 cross
    R P[5][5][NC];
#pragma unroll
    for (int dr = -2; dr <= 2; ++dr)
#pragma unroll
        for (int dc = -2; dc <= 2; ++dc) {
            const bool need = (dr >= -1 && dr <= 1 && dc >= -1 && dc <= 1) || dr == 0 || dc == 0;
            if (need) ok &= ld.prim(b, k + dr, c + dc, prm, P[dr + 2][dc + 2]);
        }
    // full PLM slopes along x (rows) and y (columns) at the five cells
    // (0,0), (-1,0), (1,0), (0,-1), (0,1)
    auto gx = [&] (int dr, int dc, R* g) {
#pragma unroll
        for (int m = 0; m < NC; ++m) g[m] = plm(P[dr + 1][dc + 2][m], P[dr + 2][dc + 2][m], P[dr + 3][dc + 2][m], theta);
    };
    auto gy = [&] (int dr, int dc, R* g) {
#pragma unroll
        for (int m = 0; m < NC; ++m) g[m] = plm(P[dr + 2][dc + 1][m], P[dr + 2][dc + 2][m], P[dr + 2][dc + 3][m], theta);
    };
    R gx0[NC], gxm[NC], gxp[NC], gxl[NC], gxr[NC];   // x slopes at (0,0), (-1,0), (1,0), (0,-1), (0,1)
    R gy0[NC], gym[NC], gyp[NC], gyl[NC], gyr[NC];
    gx(0, 0, gx0); gx(-1, 0, gxm); gx(1, 0, gxp); gx(0, -1, gxl); gx(0, 1, gxr);
    gy(0, 0, gy0); gy(-1, 0, gym); gy(1, 0, gyp); gy(0, -1, gyl); gy(0, 1, gyr);

    // face (lower cell lo, upper cell up) states: p_lo + 0.5 g_lo, p_up - 0.5 g_up
    R f[4][NC];
    {
        R pl[NC], pr[NC];
        // x face below: cells (-1,0) | (0,0)
#pragma unroll
        for (int m = 0; m < NC; ++m) { pl[m] = P[1][2][m] + R(0.5) * gxm[m]; pr[m] = P[2][2][m] - R(0.5) * gx0[m]; }
        S::template face<1>(pl, pr, prm, td, tc, k, c, f[0]);
        ViscousTerms<R>::template add<1>(f[0], pl, pr, gxm, gx0, gym, gy0, nu, dl, idl);
        // x face above: cells (0,0) | (1,0)
#pragma unroll
        for (int m = 0; m < NC; ++m) { pl[m] = P[2][2][m] + R(0.5) * gx0[m]; pr[m] = P[3][2][m] - R(0.5) * gxp[m]; }
        S::template face<1>(pl, pr, prm, td, tc, k + 1, c, f[1]);
        ViscousTerms<R>::template add<1>(f[1], pl, pr, gx0, gxp, gy0, gyp, nu, dl, idl);
        // y face left: cells (0,-1) | (0,0)
#pragma unroll
        for (int m = 0; m < NC; ++m) { pl[m] = P[2][1][m] + R(0.5) * gyl[m]; pr[m] = P[2][2][m] - R(0.5) * gy0[m]; }
        S::template face<2>(pl, pr, prm, td, tc, k, c, f[2]);
        ViscousTerms<R>::template add<2>(f[2], pl, pr, gyl, gy0, gxl, gx0, nu, dl, idl);
        // y face right: cells (0,0) | (0,1)
#pragma unroll
        for (int m = 0; m < NC; ++m) { pl[m] = P[2][2][m] + R(0.5) * gy0[m]; pr[m] = P[2][3][m] - R(0.5) * gyr[m]; }
        S::template face<2>(pl, pr, prm, td, tc, k, c + 1, f[3]);
        ViscousTerms<R>::template add<2>(f[3], pl, pr, gy0, gyr, gx0, gxr, nu, dl, idl);
    }
    double dfx[NC], dfy[NC], uu[NC], res[NC];
#pragma unroll
    for (int m = 0; m < NC; ++m) {
        dfx[m] = double(f[1][m]) - double(f[0][m]);
        dfy[m] = double(f[3][m]) - double(f[2][m]);
    }
    ld.cons_row(b, k, c, uu);
    S::update(prm, td, tc, k, c, uu, P[2][2], dfx, dfy, dt, dl, res);
    double* o = a.out + ((size_t)b * NC * n + k) * n + c;
    if (a.u0) {
        const double* z0 = a.u0 + ((size_t)b * NC * n + k) * n + c;
#pragma unroll
        for (int m = 0; m < NC; ++m) o[(size_t)m * n * n] = z0[(size_t)m * n * n] * a.rk_b + res[m] * (1.0 - a.rk_b);
    } else {
#pragma unroll
        for (int m = 0; m < NC; ++m) o[(size_t)m * n * n] = res[m];
    }
    if (!ok) report_error(a.err, 1, b, k * n + c);
}

} // namespace m3p
