// The fused 1D Euler stage kernel: guard fill (periodic or outflow) +
// recover + PLM + HLLE + update (+ RK combine), one pass per stage.
//   modules::euler1d::updated_conserved, src/module_euler.cpp:107-135
//   outflow boundary condition,           src/mara.cpp:123-143
//
// A workgroup of TJ threads owns TJ zones of one block.  Primitives of the TJ
// zones and four guard zones go to LDS; each lane computes the right face of
// its zone (from the half slopes at its zone and the next, the latter also
// computed by the next lane, from the same operands), and the left faces of
// TJ items at a time come from one pass of all lanes.  Every face flux is
// computed once.
#pragma once
#include "physics.hpp"

namespace m3p {

struct Stage1DArgs {
    const double* u;       // [nblk][5][n]
    const double* u0;      // RK base or nullptr
    double* out;
    const int* nbr;        // [nblk][2] = {prev, next}; -1 = outflow boundary
    double dt, dl, rk_b;
    const double* dt_dev;  // device-resident dt or nullptr
    const double* dl_blk;  // per-block cell spacing or nullptr
    int n, nblk;
    int* err;
    // extended-array mode (per-block compatibility): pe [5][n+4] primitives
    const double* pe;
    int outflow_lower, outflow_upper;
    // fused CFL reduction: atomic max of (max(|v1|, |v2|) + c_s) / dl of the
    // values written, or nullptr
    unsigned long long* rate_slot;
};

// Persistent work-list form: the grid is sized to the device and each
// workgroup walks the list of (block, tile) items with a fixed stride, so one
// launch covers any number of blocks of any mix of problems.
template<typename R, int TJ>
__global__ void __launch_bounds__(TJ)
stage1d_kernel(EulerParams prm, Stage1DArgs a)
{
    constexpr int NC = 5;
    const int n = a.n;
    const int tiles = n / TJ;
    const long items = (long)a.nblk * tiles;
    const int t = threadIdx.x;
    const R theta = R(prm.theta);
    const double dt = a.dt_dev ? *a.dt_dev : a.dt;

    // FP64 (bound by its arithmetic) computes the left faces of TJ items in
    // one pass of all lanes; the mixed kernel (bound by memory latency) lets
    // lane 0 compute its item's left face from the LDS row, which needs no
    // further loads.
    constexpr bool kGroupLeftFaces = is_fp64<R>::value;
    __shared__ R P[NC][TJ + 4];   // [i] = zone c0 - 2 + i
    __shared__ R F[NC][TJ + 1];   // [i] = face c0 + i (left face of zone c0 + i)
    __shared__ R Fl[NC][TJ];      // left face of each item of the current group

    // The lane's own zone of the next item, and for lanes 0-3 the four guard
    // zones of that item (c0 - 2, c0 - 1, c0 + TJ, c0 + TJ + 1: neighbour
    // block, or the edge zone at an outflow boundary), are fetched while the
    // current item is computed (software pipelining across the work list).
    // The second RK stage (u0 given) walks the list backwards: its first
    // reads are what the first stage wrote last (likeliest to be cached).
    const bool rev = a.u0 != nullptr;
    auto slot = [&] (long it) { return rev ? items - 1 - it : it; };
    auto guard_zone = [&] (int b, int i) {
        // zone index i in [-2, 0) or [n, n+2) of block b
        int bb = b, z = i;
        if (i < 0) {
            int pb = a.nbr[2 * b];
            if (pb < 0) z = 0; else { bb = pb; z = i + n; }
        } else if (i >= n) {
            int nb = a.nbr[2 * b + 1];
            if (nb < 0) z = n - 1; else { bb = nb; z = i - n; }
        }
        return a.u + (size_t)bb * NC * n + z;
    };
    auto fetch = [&] (long it, double* raw, double* graw) {
        it = slot(it);
        const int b = (int)(it / tiles), c0 = (int)(it % tiles) * TJ;
        const double* src = a.u + (size_t)b * NC * n + c0 + t;
#pragma unroll
        for (int q = 0; q < NC; ++q) raw[q] = src[(size_t)q * n];
        if (!a.pe && t < 4) {
            const double* g = guard_zone(b, t < 2 ? c0 - 2 + t : c0 + TJ + (t - 2));
#pragma unroll
            for (int q = 0; q < NC; ++q) graw[q] = g[(size_t)q * n];
        }
    };
    double nxt[NC] = {}, gnxt[NC] = {};
    if (blockIdx.x < items) fetch(blockIdx.x, nxt, gnxt);

    // extended-array mode (per-block compatibility): zone index i in
    // [-2, n+2) of the caller's primitive array
    auto load_pe = [&] (int i, R* p) {
        int e = i;
        if (a.outflow_lower && i < 0) e = 0;
        else if (!a.outflow_lower && a.outflow_upper && i >= n) e = n - 1;
        for (int q = 0; q < NC; ++q) p[q] = R(a.pe[(size_t)q * (n + 4) + e + 2]);
    };
    // The left face of item `it` (zones c0 - 2 .. c0 + 1), with the operands
    // and arithmetic of every other face.  A tile of TJ zones has TJ + 1
    // faces; each lane computes the right face of its zone, and the left
    // faces of a group of TJ items come from one pass of all lanes, so no
    // item needs a second, nearly idle flux pass.  Recovery errors in these
    // zones are reported by the items that own them.
    auto left_face = [&] (long it, R* f) {
        const int b = (int)(it / tiles), c0 = (int)(it % tiles) * TJ;
        R pz[4][NC];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int i = c0 - 2 + r;
            if (a.pe) {
                load_pe(i, pz[r]);
            } else {
                const double* src = i < 0 ? guard_zone(b, i) : a.u + (size_t)b * NC * n + i;
                double uu[NC];
#pragma unroll
                for (int q = 0; q < NC; ++q) uu[q] = src[(size_t)q * n];
                (void)Euler<R>::recover(uu, prm, pz[r]);
            }
        }
        R pl[NC], pr[NC];
#pragma unroll
        for (int q = 0; q < NC; ++q) {
            pl[q] = pz[1][q] + plm_half(pz[0][q], pz[1][q], pz[2][q], theta);
            pr[q] = pz[2][q] - plm_half(pz[1][q], pz[2][q], pz[3][q], theta);
        }
        Euler<R>::template riemann<1>(pl, pr, prm, f);
    };

    bool ok = true;
    double rate = 0.0;                  // fused CFL reduction
    for (long item = blockIdx.x; item < items; item += gridDim.x) {
        // position of this item in its group of TJ consecutive items of this
        // workgroup; the group's left faces are computed at its first item
        const int gpos = (int)(((item - blockIdx.x) / gridDim.x) % TJ);
        if (kGroupLeftFaces && gpos == 0) {
            const long li = item + (long)t * gridDim.x;
            if (li < items) {
                R f[NC];
                left_face(slot(li), f);
#pragma unroll
                for (int q = 0; q < NC; ++q) Fl[q][t] = f[q];
            }
        }
        const long it = slot(item);
        const int b = (int)(it / tiles);
        const int c0 = (int)(it % tiles) * TJ;
        const double dl = a.dl_blk ? a.dl_blk[b] : a.dl;
        const double idl = a.rate_slot ? 1.0 / dl : 0.0;
        double uz[NC], gz[NC], U0[NC];
#pragma unroll
        for (int q = 0; q < NC; ++q) { uz[q] = nxt[q]; gz[q] = gnxt[q]; }
        if (item + gridDim.x < items) fetch(item + gridDim.x, nxt, gnxt);
        if (a.u0) {
#pragma unroll
            for (int q = 0; q < NC; ++q) U0[q] = a.u0[(size_t)b * NC * n + (size_t)q * n + c0 + t];
        }
        R p[NC];
        if (a.pe) load_pe(c0 + t, p);
        else ok &= Euler<R>::recover(uz, prm, p);
        for (int q = 0; q < NC; ++q) P[q][t + 2] = p[q];
        if (t < 4) {
            if (a.pe) load_pe(t < 2 ? c0 - 2 + t : c0 + TJ + (t - 2), p);
            else ok &= Euler<R>::recover(gz, prm, p);
            int slot = t < 2 ? t : TJ + 2 + (t - 2);
            for (int q = 0; q < NC; ++q) P[q][slot] = p[q];
        }
        __syncthreads();
        {
            // the right face of zone c0 + t (face t + 1), from the half slopes
            // at zones c0 + t and c0 + t + 1
            R pl[NC], pr[NC], f[NC];
#pragma unroll
            for (int q = 0; q < NC; ++q) {
                pl[q] = P[q][t + 2] + plm_half(P[q][t + 1], P[q][t + 2], P[q][t + 3], theta);
                pr[q] = P[q][t + 3] - plm_half(P[q][t + 2], P[q][t + 3], P[q][t + 4], theta);
            }
            Euler<R>::template riemann<1>(pl, pr, prm, f);
#pragma unroll
            for (int q = 0; q < NC; ++q) F[q][t + 1] = f[q];
            if (t == 0) {
                if constexpr (kGroupLeftFaces) {
#pragma unroll
                    for (int q = 0; q < NC; ++q) F[q][0] = Fl[q][gpos];
                } else {
                    // the left face from the LDS row (zones c0 - 2 .. c0 + 1)
#pragma unroll
                    for (int q = 0; q < NC; ++q) {
                        pl[q] = P[q][1] + plm_half(P[q][0], P[q][1], P[q][2], theta);
                        pr[q] = P[q][2] - plm_half(P[q][1], P[q][2], P[q][3], theta);
                    }
                    Euler<R>::template riemann<1>(pl, pr, prm, f);
#pragma unroll
                    for (int q = 0; q < NC; ++q) F[q][0] = f[q];
                }
            }
        }
        __syncthreads();

        const int z = c0 + t;
        double* o = a.out + (size_t)b * NC * n + z;
        double wv[NC];
        for (int q = 0; q < NC; ++q) {
            double dfx = double(F[q][t + 1]) - double(F[q][t]);
            double res;
            if constexpr (is_fp64<R>::value) res = uz[q] - dfx * dt / dl;
            else res = uz[q] - dfx * (dt / dl);
            if (a.u0) res = U0[q] * a.rk_b + res * (1.0 - a.rk_b);
            o[(size_t)q * n] = res;
            wv[q] = res;
        }
        if (a.rate_slot) rate = fmax(rate, Euler<double>::max_speed(wv, prm) * idl);
        if (!ok) { report_error(a.err, 1, b, z); ok = true; }
        // No barrier here: the next item writes P and Fl only after this
        // item's flux barrier, which every read of them precedes, and F only
        // after its own first barrier, which follows this item's update.
    }
    if (a.rate_slot) workgroup_atomic_max<TJ>(rate, a.rate_slot);
}

} // namespace m3p
