// Fused RK2 step of the 1D work-list stage (mixed precision): both stages of
// one RK2 step in one launch, the first stage's state never stored.
//
// The two-stage form (stage1d_kernel twice) writes U1 = L(U0) for every block
// and reads it back, with its guard zones, for U = b U0 + (1 - b) L(U1)
// (modules::euler1d::updated_conserved, src/module_euler.cpp:107-135, per
// stage; outflow ends, src/mara.cpp:123-143).  Here a workgroup of TJ lanes
// owns TJ zones c0 .. c0 + TJ - 1 of one block.  It loads U0 on zones
// c0 - 4 .. c0 + TJ + 3, computes the first stage on zones c0 - 2 ..
// c0 + TJ + 1 (its own zones and the two on each side that the second stage
// reads), keeps U1's primitives in LDS and its own zones' U1 in registers,
// then computes the second stage and the RK combine of its own zones.
//
// Every U1 value is the one the two-stage path computes: a zone of the
// neighbouring block is updated with that block's cell spacing from the same
// operands (the neighbour's guard zones are this block's zones, which needs
// the prev/next links to be mutual: block pb = prev(b) has next(pb) = b, and
// likewise for next, as in every chain and ring; a table where they are not
// is reported through the error latch with code M3P_ERR_BAD_SHAPE).  At an
// outflow end the second stage's guard zones copy U1 of the edge zone, as the
// two-stage guard fill copies it, instead of being updated.  So the step
// equals two stage1d launches value for value; HBM sees U0 (and eight guard
// zones per TJ zones) and U once, instead of U0, U1, U1, U0 and U.
#pragma once
#include "stage1d.hpp"

namespace m3p {

template<int TJ>
__global__ void __launch_bounds__(TJ)
rk2_fused_1d_kernel(EulerParams prm, Stage1DArgs a)
{
    using R = float;
    constexpr int NC = 5;
    static_assert(TJ >= 8, "eight guard lanes");
    const int n = a.n;
    const int tiles = n / TJ;
    const int b = (int)(blockIdx.x / tiles);
    const int c0 = (int)(blockIdx.x % tiles) * TJ;
    const int t = threadIdx.x;
    const R theta = R(prm.theta);
    const double dt = a.dt_dev ? *a.dt_dev : a.dt;
    const int pb = a.nbr[2 * b], nb = a.nbr[2 * b + 1];
    const bool lo_edge = c0 == 0, hi_edge = c0 + TJ == n;
    const bool lo_out = lo_edge && pb < 0, hi_out = hi_edge && nb < 0;
    if ((lo_edge && pb >= 0 && a.nbr[2 * pb + 1] != b) || (hi_edge && nb >= 0 && a.nbr[2 * nb] != b)) {
        if (t == 0) report_error(a.err, 4, b, lo_edge && pb >= 0 && a.nbr[2 * pb + 1] != b ? -1 : n);
        return;
    }
    const double dl = a.dl_blk ? a.dl_blk[b] : a.dl;
    // the cell spacing of the neighbours whose edge zones this tile updates
    // in the first stage, loaded here rather than after a barrier, where
    // its latency held up the workgroup (launch 2 % faster)
    const double dl_lo = lo_edge && pb >= 0 && a.dl_blk ? a.dl_blk[pb] : dl;
    const double dl_hi = hi_edge && nb >= 0 && a.dl_blk ? a.dl_blk[nb] : dl;

    __shared__ R P0[NC][TJ + 8];   // [i] = zone c0 - 4 + i: U0 primitives
    __shared__ R F1[NC][TJ + 5];   // [i] = face c0 - 2 + i (the left face of zone c0 - 2 + i): first stage
    __shared__ R P1[NC][TJ + 4];   // [i] = zone c0 - 2 + i: U1 primitives
    __shared__ R F2[NC][TJ + 1];   // [i] = face c0 + i: second stage

    // zone i in [-4, 0) or [n, n + 4) of block b: the neighbour's zone, or
    // the edge zone at an outflow end (the stage kernel's guard fill)
    auto zone_ptr = [&] (int i) {
        int bb = b, z = i;
        if (i < 0)       { if (pb < 0) z = 0;     else { bb = pb; z = i + n; } }
        else if (i >= n) { if (nb < 0) z = n - 1; else { bb = nb; z = i - n; } }
        return a.u + (size_t)bb * NC * n + z;
    };
    // guard lane g (t < 8): zone c0 - 4 + g, or c0 + TJ + g - 4
    const int gz_zone = t < 4 ? c0 - 4 + t : c0 + TJ + (t - 4);
    const int gz_slot = t < 4 ? t : TJ + 4 + (t - 4);

    bool ok = true;
    double u0[NC], gu0[NC];
    {
        const double* src = a.u + (size_t)b * NC * n + c0 + t;
#pragma unroll
        for (int q = 0; q < NC; ++q) u0[q] = src[(size_t)q * n];
        if (t < 8) {
            const double* g = zone_ptr(gz_zone);
#pragma unroll
            for (int q = 0; q < NC; ++q) gu0[q] = g[(size_t)q * n];
        }
        R p[NC];
        ok &= Euler<R>::recover(u0, prm, p);
#pragma unroll
        for (int q = 0; q < NC; ++q) P0[q][t + 4] = p[q];
        if (t < 8) {
            ok &= Euler<R>::recover(gu0, prm, p);
#pragma unroll
            for (int q = 0; q < NC; ++q) P0[q][gz_slot] = p[q];
        }
    }
    __syncthreads();

    // first stage, faces c0 - 2 .. c0 + TJ + 2: face i from zones i .. i + 3
    // of the P0 row, with the operands and order of stage1d_kernel's faces
    auto face1 = [&] (int i) {
        R pl[NC], pr[NC], f[NC];
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            pl[q] = P0[q][i + 1] + plm_half(P0[q][i], P0[q][i + 1], P0[q][i + 2], theta);
            pr[q] = P0[q][i + 2] - plm_half(P0[q][i + 1], P0[q][i + 2], P0[q][i + 3], theta);
        }
        Euler<R>::template riemann<1>(pl, pr, prm, f);
#pragma unroll
        for (int q = 0; q < NC; ++q) F1[q][i] = f[q];
    };
    face1(t);
    if (t < 5) face1(TJ + t);
    __syncthreads();

    // first-stage update of zone c0 - 2 + j with the cell spacing of its block
    auto update1 = [&] (int j, const double* uz, double dlz, double* u1) {
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            const double dfx = double(F1[q][j + 1]) - double(F1[q][j]);
            u1[q] = uz[q] - dfx * (dt / dlz);
        }
    };
    double u1[NC];
    {
        update1(t + 2, u0, dl, u1);
        R p[NC];
        ok &= Euler<R>::recover(u1, prm, p);
#pragma unroll
        for (int q = 0; q < NC; ++q) P1[q][t + 2] = p[q];
        // outflow ends: the second stage's guard zones copy the edge zone
        if (t == 0 && lo_out) {
#pragma unroll
            for (int q = 0; q < NC; ++q) { P1[q][0] = p[q]; P1[q][1] = p[q]; }
        }
        if (t == TJ - 1 && hi_out) {
#pragma unroll
            for (int q = 0; q < NC; ++q) { P1[q][TJ + 2] = p[q]; P1[q][TJ + 3] = p[q]; }
        }
        // lanes 2, 3 (zones c0 - 2, c0 - 1) and 4, 5 (zones c0 + TJ, c0 + TJ + 1)
        // update the guard zones they loaded, unless at an outflow end
        const bool lo_x = (t == 2 || t == 3) && !lo_out;
        const bool hi_x = (t == 4 || t == 5) && !hi_out;
        if (lo_x || hi_x) {
            const int j = lo_x ? t - 2 : TJ + (t - 2);
            const double dlz = lo_x ? dl_lo : dl_hi;
            double g1[NC];
            update1(j, gu0, dlz, g1);
            ok &= Euler<R>::recover(g1, prm, p);
#pragma unroll
            for (int q = 0; q < NC; ++q) P1[q][j] = p[q];
        }
    }
    __syncthreads();

    // second stage, faces c0 .. c0 + TJ: lane t the right face of its zone
    // (P1 zones t + 1 .. t + 4), lane 0 also the left face of the tile
    auto face2 = [&] (int i) {
        R pl[NC], pr[NC], f[NC];
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            pl[q] = P1[q][i + 1] + plm_half(P1[q][i], P1[q][i + 1], P1[q][i + 2], theta);
            pr[q] = P1[q][i + 2] - plm_half(P1[q][i + 1], P1[q][i + 2], P1[q][i + 3], theta);
        }
        Euler<R>::template riemann<1>(pl, pr, prm, f);
#pragma unroll
        for (int q = 0; q < NC; ++q) F2[q][i] = f[q];
    };
    face2(t + 1);
    if (t == 0) face2(0);
    __syncthreads();

    // second-stage update and the RK combine (stage1d_kernel's arithmetic)
    const int z = c0 + t;
    double* o = a.out + (size_t)b * NC * n + z;
    double wv[NC];
#pragma unroll
    for (int q = 0; q < NC; ++q) {
        const double dfx = double(F2[q][t + 1]) - double(F2[q][t]);
        double res = u1[q] - dfx * (dt / dl);
        res = u0[q] * a.rk_b + res * (1.0 - a.rk_b);
        o[(size_t)q * n] = res;
        wv[q] = res;
    }
    if (!ok) report_error(a.err, 1, b, z);
    if (a.rate_slot) {
        const double idl = 1.0 / dl;
        workgroup_atomic_max<TJ>(Euler<double>::max_speed(wv, prm) * idl, a.rate_slot);
    }
}

} // namespace m3p
