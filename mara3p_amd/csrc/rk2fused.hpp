// Fused RK2 step on a uniform mesh of blocks: both stages of one RK2 step in
// one launch, the first stage's state never stored.
//
// The reference's step (control::advance_runge_kutta, src/app_control.hpp:
// 169-195, over the per-block update) materialises U1 = L(U0) for the whole
// mesh and then forms U = b U0 + (1 - b) L(U1).  Here a workgroup owns H rows
// x the full width of one block and marches down its rows with the first
// stage three rows ahead of the second: every U1 row is computed in
// registers, exchanged across lanes through LDS rows, and consumed by the
// second stage a few iterations later.  The second stage of rows
// [k0, k0 + H) needs U1 on rows [k0 - 2, k0 + H + 2) of the tile's columns
// (the first stage runs on H + 4 rows, two of them on each side owned by the
// blocks before and after on axis 0) and on the two columns at each side of
// the block (a pre-pass computes those zones' first stage one zone at a
// time, with the arithmetic of the march).  Every zone's U1 is computed with
// the operands, geometry and order of the two-stage kernels, so the step
// equals two stage2d launches value for value; HBM sees U0 (plus the halo
// rows) and U once instead of U0, U1, U1, U0 and U.
//
// Zones outside the block are read the way the block that owns their row
// reads them in its own stage: rows through the axis-0 neighbours, then
// columns through that row block's axis-1 neighbours (the first stage needs
// no corner zones of any block, so the reference's corner convention does
// not enter).
#pragma once
#include "stage2d.hpp"

namespace m3p {

template<class S, int H, int TJ>
__global__ void __launch_bounds__(TJ) __attribute__((amdgpu_waves_per_eu(S::kFusedWavesPerEU, 8)))
rk2_fused_kernel(typename S::Params p1, typename S::Params p2, MeshLoader<S> ld, StageArgs a)
{
    constexpr int NC = S::NC;
    constexpr int HE = H + 4;                  // first-stage rows k0 - 2 .. k0 + H + 1
    using R = typename S::Real;
    resolve_params(p1);
    resolve_params(p2);

    const int n = a.n;                         // == TJ
    const int tiles_i = n / H;
    const int w = xcd_remap(blockIdx.x, gridDim.x);
    const int b = w / tiles_i;
    const int k0 = (w % tiles_i) * H;
    const int t = threadIdx.x;
    const int c = t;
    const double dt = a.dt_dev ? *a.dt_dev : a.dt;
    const double dl = a.dl;
    const int* nbr = ld.nbr;
    const int xm = nbr[4 * b + 0], xp = nbr[4 * b + 1], ym = nbr[4 * b + 2], yp = nbr[4 * b + 3];
    ld.bind(b);

    // Conserved state of zone (r, cc) seen from block Z: the row first
    // (through Z's axis-0 neighbours), then the column through the row
    // block's axis-1 neighbours.  |offsets| <= n.
    auto zone_in = [&] (int Z, int r, int cc) -> const double* {
        int ob = Z;
        if (r < 0)       { ob = nbr[4 * Z + 0]; r += n; }
        else if (r >= n) { ob = nbr[4 * Z + 1]; r -= n; }
        int blk = ob;
        if (cc < 0)       { blk = nbr[4 * ob + 2]; cc += n; }
        else if (cc >= n) { blk = nbr[4 * ob + 3]; cc -= n; }
        return ld.u + ((size_t)blk * NC * n + r) * n + cc;
    };
    auto cons_of = [&] (const double* z, double* uu) {
#pragma unroll
        for (int m = 0; m < NC; ++m) uu[m] = z[(size_t)m * n * n];
    };

    TileCtx tc;
    tc.b = b; tc.n = n; tc.k0 = k0 - 2; tc.c0 = 0;
    tc.bi = a.bcoord[2 * b]; tc.bj = a.bcoord[2 * b + 1];

    __shared__ typename S::template TileData<HE, TJ> td;   // rows k0 - 2 .. k0 + H + 1
    __shared__ R e1[7][NC][HE];    // U0 primitives at columns -2, -1, 0, n - 2, n - 1, n, n + 1
    __shared__ R g1e[NC][HE];      // first stage: half slope at column -1
    __shared__ R f1e[NC][HE];      // first stage: y flux at face n
    __shared__ R e2[7][NC][H];     // U1 primitives at the same columns, rows k0 .. k0 + H - 1
    __shared__ R g2e[NC][H];
    __shared__ R f2e[NC][H];
    __shared__ R rowP1[NC][TJ + 2], rowG1[NC][TJ + 1], rowF1[NC][TJ + 1];
    __shared__ R rowP2[NC][TJ + 2], rowG2[NC][TJ + 1], rowF2[NC][TJ + 1];
    // U1 of the last three first-stage rows (row j in slot s % 3), each lane
    // its own column: the second stage reads row j - 3 from the slot the
    // first stage then overwrites, in program order within the lane
    __shared__ double u1q[3][NC][TJ];

    S::setup_rows(td, p1, tc, a.bcoord[2 * xm], a.bcoord[2 * xp]);
    const R theta = R(S::theta(p1));
    bool ok = true;
    auto edge_col = [&] (int q) {
        return q == 0 ? -2 : q == 1 ? -1 : q == 2 ? 0 : q == 3 ? n - 2 : q == 4 ? n - 1 : q == 5 ? n : n + 1;
    };

    // ---- pre-pass: U0 primitives at the edge columns of rows k0-2 .. k0+H+1
    for (int idx = t; idx < 7 * HE; idx += TJ) {
        const int r = idx / 7, q = idx % 7;
        double uu[NC];
        R p[NC];
        cons_of(zone_in(b, k0 - 2 + r, edge_col(q)), uu);
        ok &= S::recover(uu, p1, p);
#pragma unroll
        for (int m = 0; m < NC; ++m) e1[q][m][r] = p[m];
    }
    // ---- pre-pass: first stage of the zones at the edge columns of rows
    // k0 .. k0+H-1, one zone per lane (the march's arithmetic), as primitives
    for (int idx = t; idx < 7 * H; idx += TJ) {
        const int r = idx / 7, q = idx % 7;
        const int rl = k0 + r;
        int Z = b, cl = edge_col(q);
        if (cl < 0)       { Z = ym; cl += n; }
        else if (cl >= n) { Z = yp; cl -= n; }
        TileCtx zc;
        zc.b = Z; zc.n = n; zc.k0 = rl; zc.c0 = cl;
        zc.bi = a.bcoord[2 * Z]; zc.bj = a.bcoord[2 * Z + 1];
        typename S::template TileData<1, 1> zt;
        S::setup_zone(zt, p1, zc);
        R P[5][NC], Q[5][NC];
        double u0z[NC];
#pragma unroll
        for (int d = 0; d < 5; ++d) {
            double uu[NC];
            cons_of(zone_in(Z, rl + d - 2, cl), uu);
            ok &= S::recover(uu, p1, P[d]);
            if (d == 2) {
#pragma unroll
                for (int m = 0; m < NC; ++m) u0z[m] = uu[m];
            }
        }
#pragma unroll
        for (int d = 0; d < 5; ++d) {
            if (d == 2) {
#pragma unroll
                for (int m = 0; m < NC; ++m) Q[2][m] = P[2][m];
                continue;
            }
            double uu[NC];
            cons_of(zone_in(Z, rl, cl + d - 2), uu);
            ok &= S::recover(uu, p1, Q[d]);
        }
        R gx[3][NC], gy[3][NC];
#pragma unroll
        for (int m = 0; m < NC; ++m) {
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                gx[i][m] = plm_half(P[i][m], P[i + 1][m], P[i + 2][m], theta);
                gy[i][m] = plm_half(Q[i][m], Q[i + 1][m], Q[i + 2][m], theta);
            }
        }
        R pl[NC], pr[NC], FxL[NC], FxH[NC], FyL[NC], FyH[NC];
#pragma unroll
        for (int m = 0; m < NC; ++m) { pl[m] = P[1][m] + gx[0][m]; pr[m] = P[2][m] - gx[1][m]; }
        S::template face<1>(pl, pr, p1, zt, zc, rl, cl, FxL);
#pragma unroll
        for (int m = 0; m < NC; ++m) { pl[m] = P[2][m] + gx[1][m]; pr[m] = P[3][m] - gx[2][m]; }
        S::template face<1>(pl, pr, p1, zt, zc, rl + 1, cl, FxH);
#pragma unroll
        for (int m = 0; m < NC; ++m) { pl[m] = Q[1][m] + gy[0][m]; pr[m] = Q[2][m] - gy[1][m]; }
        S::template face<2>(pl, pr, p1, zt, zc, rl, cl, FyL);
#pragma unroll
        for (int m = 0; m < NC; ++m) { pl[m] = Q[2][m] + gy[1][m]; pr[m] = Q[3][m] - gy[2][m]; }
        S::template face<2>(pl, pr, p1, zt, zc, rl, cl + 1, FyH);
        double dfx[NC], dfy[NC], u1[NC];
#pragma unroll
        for (int m = 0; m < NC; ++m) {
            dfx[m] = double(FxH[m]) - double(FxL[m]);
            dfy[m] = double(FyH[m]) - double(FyL[m]);
        }
        S::update(p1, zt, zc, rl, cl, u0z, P[2], dfx, dfy, dt, dl, u1);
        R p[NC];
        ok &= S::recover(u1, p2, p);
#pragma unroll
        for (int m = 0; m < NC; ++m) e2[q][m][r] = p[m];
    }
    __syncthreads();
    // edge slopes (column -1) and edge fluxes (face n) of both stages
    for (int r = t; r < HE + H; r += TJ) {
        const bool first = r < HE;
        const int rr = first ? r : r - HE;
        R v[7][NC];
#pragma unroll
        for (int q = 0; q < 7; ++q)
#pragma unroll
            for (int m = 0; m < NC; ++m) v[q][m] = first ? e1[q][m][rr] : e2[q][m][rr];
        R pl[NC], pr[NC], f[NC], g[NC];
#pragma unroll
        for (int m = 0; m < NC; ++m) {
            g[m] = plm_half(v[0][m], v[1][m], v[2][m], theta);
            const R gA = plm_half(v[3][m], v[4][m], v[5][m], theta);
            const R gB = plm_half(v[4][m], v[5][m], v[6][m], theta);
            pl[m] = v[4][m] + gA;
            pr[m] = v[5][m] - gB;
        }
        if (first) S::template face<2>(pl, pr, p1, td, tc, k0 - 2 + rr, n, f);
        else       S::template face<2>(pl, pr, p2, td, tc, k0 + rr, n, f);
#pragma unroll
        for (int m = 0; m < NC; ++m) {
            if (first) { g1e[m][rr] = g[m]; f1e[m][rr] = f[m]; }
            else       { g2e[m][rr] = g[m]; f2e[m][rr] = f[m]; }
        }
    }
    __syncthreads();

    // ---- first stage: warm-up at row k0 - 2 ------------------------------
    // Fb1: x flux at the face above row j, kept in FP32 (the update takes
    // its FP64 value, which is exact); U0 of rows j and k is re-read for the
    // two updates (a cache hit: row j was fetched three rows earlier), the
    // loads issued after the second barrier of the step so that they are in
    // flight during the y fluxes (4 % faster than loading at the updates;
    // loading at the top of the step needs 175 VGPRs and runs slower)
    R Pa1[NC], Pb1[NC], ga1[NC], Fb1[NC];
    double Unx[NC];
    {
        R Pm2[NC], Pm1[NC], gm[NC], pl[NC], pr[NC], f[NC];
        ok &= ld.prim_row(b, k0 - 4, c, p1, Pm2);
        ok &= ld.prim_row(b, k0 - 3, c, p1, Pm1);
        ok &= ld.prim_row(b, k0 - 2, c, p1, Pa1);
        ok &= ld.prim_row(b, k0 - 1, c, p1, Pb1);
#pragma unroll
        for (int m = 0; m < NC; ++m) {
            gm[m] = plm_half(Pm2[m], Pm1[m], Pa1[m], theta);
            ga1[m] = plm_half(Pm1[m], Pa1[m], Pb1[m], theta);
            pl[m] = Pm1[m] + gm[m];
            pr[m] = Pa1[m] - ga1[m];
        }
        S::template face<1>(pl, pr, p1, td, tc, k0 - 2, c, Fb1);
        ld.fetch_row(b, k0, c, Unx);
    }

    // second-stage window (filled after the first four first-stage rows)
    R Pa2[NC] = {}, Pb2[NC] = {}, ga2[NC] = {}, Pn2[NC] = {}, Pw0[NC] = {}, Pw1[NC] = {}, Fb2[NC] = {};
    const double wb = 1.0 - a.rk_b;

    // March step s: first-stage row j = k0 - 2 + s (DO1, s < H + 4), second-
    // stage row k = j - 3 (DO2, s >= 5), sharing the three workgroup barriers
    // of a row.  One loop body for every step (wave-uniform flags): the
    // first stage's rows k0 - 2 .. k0 + 1 seed the second stage's warm-up,
    // done after step 3.  Unrolled by three, most window shifts become
    // register renames (158 VGPRs, three waves per SIMD, no spills; the U1
    // queue lives in LDS, which kept 18 registers out of the march).
#pragma unroll 3
    for (int s = 0; s <= HE; ++s) {
        const bool DO1 = s < HE;
        const bool DO2 = s >= 5;
        const int j = k0 - 2 + s;
        const int k = j - 3;
        const int kl = k - k0;
        R Pc1[NC], gb1[NC], Fa1[NC], gy1[NC], Fy1[NC];
        R gb2[NC], Fa2[NC], gy2[NC], Fy2[NC];
        double U0j[NC], U0k[NC];
        // phase A: rows ahead, x slopes and x fluxes; LDS rows
        if (DO1) {
            double raw[NC];
#pragma unroll
            for (int m = 0; m < NC; ++m) raw[m] = Unx[m];
            if (s + 1 < HE) ld.fetch_row(b, j + 3, c, Unx);
            ok &= S::recover(raw, p1, Pc1);
            R pl[NC], pr[NC];
#pragma unroll
            for (int m = 0; m < NC; ++m) {
                gb1[m] = plm_half(Pa1[m], Pb1[m], Pc1[m], theta);
                pl[m] = Pa1[m] + ga1[m];
                pr[m] = Pb1[m] - gb1[m];
            }
            S::template face<1>(pl, pr, p1, td, tc, j + 1, c, Fa1);
#pragma unroll
            for (int m = 0; m < NC; ++m) rowP1[m][t + 1] = Pa1[m];
            if (t == 0) {
#pragma unroll
                for (int m = 0; m < NC; ++m) { rowP1[m][0] = e1[1][m][s]; rowG1[m][0] = g1e[m][s]; }
            }
            if (t == TJ - 1) {
#pragma unroll
                for (int m = 0; m < NC; ++m) rowP1[m][TJ + 1] = e1[5][m][s];
            }
        }
        if (DO2) {
            R pl[NC], pr[NC];
#pragma unroll
            for (int m = 0; m < NC; ++m) {
                gb2[m] = plm_half(Pa2[m], Pb2[m], Pn2[m], theta);
                pl[m] = Pa2[m] + ga2[m];
                pr[m] = Pb2[m] - gb2[m];
            }
            S::template face<1>(pl, pr, p2, td, tc, k + 1, c, Fa2);
#pragma unroll
            for (int m = 0; m < NC; ++m) rowP2[m][t + 1] = Pa2[m];
            if (t == 0) {
#pragma unroll
                for (int m = 0; m < NC; ++m) { rowP2[m][0] = e2[1][m][kl]; rowG2[m][0] = g2e[m][kl]; }
            }
            if (t == TJ - 1) {
#pragma unroll
                for (int m = 0; m < NC; ++m) rowP2[m][TJ + 1] = e2[5][m][kl];
            }
        }
        __syncthreads();
        // phase B: y half slopes
        if (DO1) {
#pragma unroll
            for (int m = 0; m < NC; ++m) {
                gy1[m] = plm_half(rowP1[m][t], Pa1[m], rowP1[m][t + 2], theta);
                rowG1[m][t + 1] = gy1[m];
            }
        }
        if (DO2) {
#pragma unroll
            for (int m = 0; m < NC; ++m) {
                gy2[m] = plm_half(rowP2[m][t], Pa2[m], rowP2[m][t + 2], theta);
                rowG2[m][t + 1] = gy2[m];
            }
        }
        __syncthreads();
        // the RK base rows of the two updates, in flight during phase C
        if (DO1) cons_of(ld.row(b, j) + c, U0j);
        if (DO2) cons_of(ld.row(b, k) + c, U0k);
        // phase C: y fluxes at the lower face of each column
        if (DO1) {
            R pl[NC], pr[NC];
#pragma unroll
            for (int m = 0; m < NC; ++m) { pl[m] = rowP1[m][t] + rowG1[m][t]; pr[m] = Pa1[m] - gy1[m]; }
            S::template face<2>(pl, pr, p1, td, tc, j, c, Fy1);
#pragma unroll
            for (int m = 0; m < NC; ++m) rowF1[m][t] = Fy1[m];
            if (t == TJ - 1) {
#pragma unroll
                for (int m = 0; m < NC; ++m) rowF1[m][TJ] = f1e[m][s];
            }
        }
        if (DO2) {
            R pl[NC], pr[NC];
#pragma unroll
            for (int m = 0; m < NC; ++m) { pl[m] = rowP2[m][t] + rowG2[m][t]; pr[m] = Pa2[m] - gy2[m]; }
            S::template face<2>(pl, pr, p2, td, tc, k, c, Fy2);
#pragma unroll
            for (int m = 0; m < NC; ++m) rowF2[m][t] = Fy2[m];
            if (t == TJ - 1) {
#pragma unroll
                for (int m = 0; m < NC; ++m) rowF2[m][TJ] = f2e[m][kl];
            }
        }
        __syncthreads();
        // phase D: updates
        if (DO2) {
            double dfx[NC], dfy[NC], res[NC];
#pragma unroll
            for (int m = 0; m < NC; ++m) {
                dfx[m] = double(Fa2[m]) - double(Fb2[m]);
                dfy[m] = double(rowF2[m][t + 1]) - double(Fy2[m]);
            }
            double u1k[NC];
#pragma unroll
            for (int m = 0; m < NC; ++m) u1k[m] = u1q[s % 3][m][t];
            S::update(p2, td, tc, k, c, u1k, Pa2, dfx, dfy, dt, dl, res);
            double* o = a.out + ((size_t)b * NC * n + k) * n + c;
#pragma unroll
            for (int m = 0; m < NC; ++m) {
                double wv;
                if constexpr (is_fp64<R>::value) wv = U0k[m] * a.rk_b + res[m] * wb;
                else                             wv = fma(U0k[m], a.rk_b, res[m] * wb);
                __builtin_nontemporal_store(wv, &o[(size_t)m * n * n]);
            }
#pragma unroll
            for (int m = 0; m < NC; ++m) { Pa2[m] = Pb2[m]; Pb2[m] = Pn2[m]; ga2[m] = gb2[m]; Fb2[m] = Fa2[m]; }
        }
        if (DO1) {
            double dfx[NC], dfy[NC], u1[NC];
#pragma unroll
            for (int m = 0; m < NC; ++m) {
                dfx[m] = double(Fa1[m]) - double(Fb1[m]);
                dfy[m] = double(rowF1[m][t + 1]) - double(Fy1[m]);
            }
            S::update(p1, td, tc, j, c, U0j, Pa1, dfx, dfy, dt, dl, u1);
            // U1 row j: its primitives feed the second stage's next row, its
            // conserved state the second stage's update three rows later
            ok &= S::recover(u1, p2, Pn2);
#pragma unroll
            for (int m = 0; m < NC; ++m) {
                u1q[s % 3][m][t] = u1[m];
                Pa1[m] = Pb1[m]; Pb1[m] = Pc1[m]; ga1[m] = gb1[m]; Fb1[m] = Fa1[m];
            }
        }
        if (s == 0) {
#pragma unroll
            for (int m = 0; m < NC; ++m) Pw0[m] = Pn2[m];
        } else if (s == 1) {
#pragma unroll
            for (int m = 0; m < NC; ++m) Pw1[m] = Pn2[m];
        } else if (s == 2) {
#pragma unroll
            for (int m = 0; m < NC; ++m) Pa2[m] = Pn2[m];
        } else if (s == 3) {
            // second-stage warm-up: slopes at k0 - 1 and k0, flux at the face k0
            R gm[NC], pl[NC], pr[NC];
#pragma unroll
            for (int m = 0; m < NC; ++m) {
                Pb2[m] = Pn2[m];
                gm[m] = plm_half(Pw0[m], Pw1[m], Pa2[m], theta);
                ga2[m] = plm_half(Pw1[m], Pa2[m], Pb2[m], theta);
                pl[m] = Pw1[m] + gm[m];
                pr[m] = Pa2[m] - ga2[m];
            }
            S::template face<1>(pl, pr, p2, td, tc, k0, c, Fb2);
        }
    }
    if (!ok) report_error(a.err, 1, b, 0);
}

} // namespace m3p
