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
// rows) and U once instead of U0, U1, U1, U0 and U.  Instantiated for the
// minidisk scheme (problems/minidisk_scheme.cpp:294-355 per stage) and the
// euler2d scheme (euler2d::updated_solution, src/module_euler.cpp:304-323);
// a scheme with a fused CFL reduction (kCfl) also reduces the CFL rate of U
// in the same launch, as its second-stage kernel does.
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

    // block coordinates (schemes with geometry); without them every block
    // takes (bi0, bj0)
    auto bi_of = [&] (int Z) { return a.bcoord ? a.bcoord[2 * Z] : a.bi0; };
    auto bj_of = [&] (int Z) { return a.bcoord ? a.bcoord[2 * Z + 1] : a.bj0; };
    TileCtx tc;
    tc.b = b; tc.n = n; tc.k0 = k0 - 2; tc.c0 = 0;
    tc.bi = bi_of(b); tc.bj = bj_of(b);
    double rate = 0.0;                                    // fused CFL reduction (S::kCfl)
    const double idl = a.rate_slot ? 1.0 / dl : 0.0;

    __shared__ typename S::template TileData<HE, TJ> td;   // rows k0 - 2 .. k0 + H + 1
    __shared__ R e1[6][NC][HE];    // U0 primitives at columns -2, -1, n - 2, n - 1, n, n + 1
    __shared__ R f1e[NC][HE];      // first stage: y flux at face n
    __shared__ R e2[6][NC][H];     // U1 primitives at the same columns, rows k0 .. k0 + H - 1
    __shared__ R f2e[NC][H];
    // a step's row of each stage: rowP [t + 2] = column t, [0] and [1] =
    // columns -2 and -1, [TJ + 2] = column n (pre-pass); rowF [t] = y flux at
    // face t, [TJ] = face n (pre-pass)
    __shared__ R rowP1[NC][TJ + 3], rowF1[NC][TJ + 1];
    __shared__ R rowP2[NC][TJ + 3], rowF2[NC][TJ + 1];
    // U1 of the last three first-stage rows (row j in slot s % 3), each lane
    // its own column: the second stage reads row j - 3 from the slot the
    // first stage then overwrites, in program order within the lane
    __shared__ double u1q[3][NC][TJ];

    S::setup_rows(td, p1, tc, bi_of(xm), bi_of(xp));
    const R theta = R(S::theta(p1));
    bool ok = true;
    auto edge_col = [&] (int q) {
        return q == 0 ? -2 : q == 1 ? -1 : n - 4 + q;
    };

    // ---- pre-pass: U0 primitives at the edge columns of rows k0-2 .. k0+H+1
    for (int idx = t; idx < 6 * HE; idx += TJ) {
        const int r = idx / 6, q = idx % 6;
        double uu[NC];
        R p[NC];
        cons_of(zone_in(b, k0 - 2 + r, edge_col(q)), uu);
        ok &= S::recover(uu, p1, p);
#pragma unroll
        for (int m = 0; m < NC; ++m) e1[q][m][r] = p[m];
    }
    // ---- pre-pass: first stage of the zones at the edge columns of rows
    // k0 .. k0+H-1, one zone per lane (the march's arithmetic), as primitives
    for (int idx = t; idx < 6 * H; idx += TJ) {
        const int r = idx / 6, q = idx % 6;
        const int rl = k0 + r;
        int Z = b, cl = edge_col(q);
        if (cl < 0)       { Z = ym; cl += n; }
        else if (cl >= n) { Z = yp; cl -= n; }
        TileCtx zc;
        zc.b = Z; zc.n = n; zc.k0 = rl; zc.c0 = cl;
        zc.bi = bi_of(Z); zc.bj = bj_of(Z);
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
    // edge fluxes (face n) of both stages
    for (int r = t; r < HE + H; r += TJ) {
        const bool first = r < HE;
        const int rr = first ? r : r - HE;
        R v[4][NC];
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int m = 0; m < NC; ++m) v[q][m] = first ? e1[q + 2][m][rr] : e2[q + 2][m][rr];
        R pl[NC], pr[NC], f[NC];
#pragma unroll
        for (int m = 0; m < NC; ++m) {
            const R gA = plm_half(v[0][m], v[1][m], v[2][m], theta);
            const R gB = plm_half(v[1][m], v[2][m], v[3][m], theta);
            pl[m] = v[1][m] + gA;
            pr[m] = v[2][m] - gB;
        }
        if (first) S::template face<2>(pl, pr, p1, td, tc, k0 - 2 + rr, n, f);
        else       S::template face<2>(pl, pr, p2, td, tc, k0 + rr, n, f);
#pragma unroll
        for (int m = 0; m < NC; ++m) {
            if (first) f1e[m][rr] = f[m];
            else       f2e[m][rr] = f[m];
        }
    }
    __syncthreads();

    // ---- the march -------------------------------------------------------
    // Step s handles first-stage row j = k0 - 2 + s (s < H + 4) and second-
    // stage row k = j - 3 (s >= 5); the two share the two workgroup barriers
    // of a step.  Every window of the march is three register sets
    // used in rotation, ROT = (s - 5) mod 3 (a compile-time constant), so the
    // steady loop, unrolled by three, shifts no window by moves:
    //   first stage:  P1[ROT], P1[ROT+1], P1[ROT+2] = primitives of rows j,
    //                 j + 1, j + 2 (the last recovered in the step); G1[ROT],
    //                 G1[ROT+1] = half slopes at rows j, j + 1; F1[ROT],
    //                 F1[ROT+1] = x fluxes at faces j, j + 1; UN[ROT] = U0 of
    //                 row j + 2, row j + 3 fetched into UN[ROT+1];
    //   second stage: P2[ROT], P2[ROT+1], P2[ROT+2] = U1 primitives of rows
    //                 k, k + 1, k + 2; G2, F2 as G1, F1 for rows k, k + 1.
    // The first stage's new U1 row (j) lands in P2[ROT], the slot of row k,
    // dead by then.  The x fluxes stay in FP32 (the update takes their FP64
    // values, which are exact).  U0 of rows j and k is re-read for the two
    // updates (a cache hit: row j was fetched three rows earlier), issued
    // after the step's first barrier so that the loads are in flight during
    // the y fluxes.  Each lane computes the y half slope of its left
    // neighbour as well as its own, from the same LDS operands, so no barrier
    // separates the y slopes from the y fluxes.
    R P1[3][NC], G1[3][NC], F1[3][NC], P2[3][NC], G2[3][NC], F2[3][NC];
    double UN[3][NC];
    const double wb = 1.0 - a.rk_b;

    auto step = [&] (auto do1_c, auto do2_c, auto rot_c, auto fetch_c, int s) {
        constexpr bool DO1 = decltype(do1_c)::value, DO2 = decltype(do2_c)::value;
        constexpr bool FETCH = decltype(fetch_c)::value;
        constexpr int A = decltype(rot_c)::value, B = (A + 1) % 3, C = (A + 2) % 3;
        constexpr int slot = C;                // s % 3, since s = 5 + ROT (mod 3)
        const int j = k0 - 2 + s;
        const int k = j - 3;
        const int kl = k - k0;
        R gy1[NC], Fy1[NC], gy2[NC], Fy2[NC];
        double U0j[NC], U0k[NC];
        // phase A: rows ahead, x slopes and x fluxes; LDS rows
        if constexpr (DO1) {
            if constexpr (FETCH) ld.fetch_row(b, j + 3, c, UN[B]);
            ok &= S::recover(UN[A], p1, P1[C]);
            R pl[NC], pr[NC];
#pragma unroll
            for (int m = 0; m < NC; ++m) {
                G1[B][m] = plm_half(P1[A][m], P1[B][m], P1[C][m], theta);
                pl[m] = P1[A][m] + G1[A][m];
                pr[m] = P1[B][m] - G1[B][m];
            }
            S::template face<1>(pl, pr, p1, td, tc, j + 1, c, F1[B]);
#pragma unroll
            for (int m = 0; m < NC; ++m) rowP1[m][t + 2] = P1[A][m];
            if (t == 0) {
#pragma unroll
                for (int m = 0; m < NC; ++m) { rowP1[m][0] = e1[0][m][s]; rowP1[m][1] = e1[1][m][s]; }
            }
            if (t == TJ - 1) {
#pragma unroll
                for (int m = 0; m < NC; ++m) rowP1[m][TJ + 2] = e1[4][m][s];
            }
        }
        if constexpr (DO2) {
            R pl[NC], pr[NC];
#pragma unroll
            for (int m = 0; m < NC; ++m) {
                G2[B][m] = plm_half(P2[A][m], P2[B][m], P2[C][m], theta);
                pl[m] = P2[A][m] + G2[A][m];
                pr[m] = P2[B][m] - G2[B][m];
            }
            S::template face<1>(pl, pr, p2, td, tc, k + 1, c, F2[B]);
#pragma unroll
            for (int m = 0; m < NC; ++m) rowP2[m][t + 2] = P2[A][m];
            if (t == 0) {
#pragma unroll
                for (int m = 0; m < NC; ++m) { rowP2[m][0] = e2[0][m][kl]; rowP2[m][1] = e2[1][m][kl]; }
            }
            if (t == TJ - 1) {
#pragma unroll
                for (int m = 0; m < NC; ++m) rowP2[m][TJ + 2] = e2[4][m][kl];
            }
        }
        __syncthreads();
        if constexpr (DO1) cons_of(ld.row(b, j) + c, U0j);
        if constexpr (DO2) cons_of(ld.row(b, k) + c, U0k);
        // phase B: y half slopes (own column and the left neighbour's), y
        // fluxes at the lower face of each column
        if constexpr (DO1) {
            R pl[NC], pr[NC];
#pragma unroll
            for (int m = 0; m < NC; ++m) {
                gy1[m] = plm_half(rowP1[m][t + 1], P1[A][m], rowP1[m][t + 3], theta);
                const R gl = plm_half(rowP1[m][t], rowP1[m][t + 1], P1[A][m], theta);
                pl[m] = rowP1[m][t + 1] + gl;
                pr[m] = P1[A][m] - gy1[m];
            }
            S::template face<2>(pl, pr, p1, td, tc, j, c, Fy1);
#pragma unroll
            for (int m = 0; m < NC; ++m) rowF1[m][t] = Fy1[m];
            if (t == TJ - 1) {
#pragma unroll
                for (int m = 0; m < NC; ++m) rowF1[m][TJ] = f1e[m][s];
            }
        }
        if constexpr (DO2) {
            R pl[NC], pr[NC];
#pragma unroll
            for (int m = 0; m < NC; ++m) {
                gy2[m] = plm_half(rowP2[m][t + 1], P2[A][m], rowP2[m][t + 3], theta);
                const R gl = plm_half(rowP2[m][t], rowP2[m][t + 1], P2[A][m], theta);
                pl[m] = rowP2[m][t + 1] + gl;
                pr[m] = P2[A][m] - gy2[m];
            }
            S::template face<2>(pl, pr, p2, td, tc, k, c, Fy2);
#pragma unroll
            for (int m = 0; m < NC; ++m) rowF2[m][t] = Fy2[m];
            if (t == TJ - 1) {
#pragma unroll
                for (int m = 0; m < NC; ++m) rowF2[m][TJ] = f2e[m][kl];
            }
        }
        __syncthreads();
        // phase C: updates; the second stage reads row k's U1 from the queue
        // slot before the first stage puts row j there
        if constexpr (DO2) {
            double dfx[NC], dfy[NC], res[NC], u1k[NC];
#pragma unroll
            for (int m = 0; m < NC; ++m) {
                dfx[m] = double(F2[B][m]) - double(F2[A][m]);
                dfy[m] = double(rowF2[m][t + 1]) - double(Fy2[m]);
                u1k[m] = u1q[slot][m][t];
            }
            S::update(p2, td, tc, k, c, u1k, P2[A], dfx, dfy, dt, dl, res);
            double* o = a.out + ((size_t)b * NC * n + k) * n + c;
            double wv[NC];
#pragma unroll
            for (int m = 0; m < NC; ++m) {
                if constexpr (is_fp64<R>::value) wv[m] = U0k[m] * a.rk_b + res[m] * wb;
                else                             wv[m] = fma(U0k[m], a.rk_b, res[m] * wb);
                __builtin_nontemporal_store(wv[m], &o[(size_t)m * n * n]);
            }
            if constexpr (S::kCfl) {
                if (a.rate_slot) rate = fmax(rate, S::cfl_speed(wv, p2) * idl);
            }
        }
        if constexpr (DO1) {
            double dfx[NC], dfy[NC], u1[NC];
#pragma unroll
            for (int m = 0; m < NC; ++m) {
                dfx[m] = double(F1[B][m]) - double(F1[A][m]);
                dfy[m] = double(rowF1[m][t + 1]) - double(Fy1[m]);
            }
            S::update(p1, td, tc, j, c, U0j, P1[A], dfx, dfy, dt, dl, u1);
            ok &= S::recover(u1, p2, P2[A]);
#pragma unroll
            for (int m = 0; m < NC; ++m) u1q[slot][m][t] = u1[m];
        }
    };
    using T_ = std::true_type;
    using F_ = std::false_type;
    using R0 = std::integral_constant<int, 0>;
    using R1 = std::integral_constant<int, 1>;
    using R2 = std::integral_constant<int, 2>;

    // first-stage warm-up at row k0 - 2, into the slots of step 0 (ROT 1)
    {
        R Pm2[NC], Pm1[NC], gm[NC], pl[NC], pr[NC];
        ok &= ld.prim_row(b, k0 - 4, c, p1, Pm2);
        ok &= ld.prim_row(b, k0 - 3, c, p1, Pm1);
        ok &= ld.prim_row(b, k0 - 2, c, p1, P1[1]);
        ok &= ld.prim_row(b, k0 - 1, c, p1, P1[2]);
#pragma unroll
        for (int m = 0; m < NC; ++m) {
            gm[m] = plm_half(Pm2[m], Pm1[m], P1[1][m], theta);
            G1[1][m] = plm_half(Pm1[m], P1[1][m], P1[2][m], theta);
            pl[m] = Pm1[m] + gm[m];
            pr[m] = P1[1][m] - G1[1][m];
        }
        S::template face<1>(pl, pr, p1, td, tc, k0 - 2, c, F1[1]);
        ld.fetch_row(b, k0, c, UN[1]);
    }
    // steps 0 .. 4: first stage only; the U1 rows k0 - 2 and k0 - 1 (steps
    // 0 and 1) seed the second stage's warm-up, rows k0 .. k0 + 2 (steps
    // 2 .. 4) land in P2[0..2], the slots of step 5 (ROT 0)
    R Pw0[NC], Pw1[NC];
    step(T_{}, F_{}, R1{}, T_{}, 0);
#pragma unroll
    for (int m = 0; m < NC; ++m) Pw0[m] = P2[1][m];
    step(T_{}, F_{}, R2{}, T_{}, 1);
#pragma unroll
    for (int m = 0; m < NC; ++m) Pw1[m] = P2[2][m];
    step(T_{}, F_{}, R0{}, T_{}, 2);
    step(T_{}, F_{}, R1{}, T_{}, 3);
    {
        // second-stage warm-up: slopes at k0 - 1 and k0, flux at the face k0
        R gm[NC], pl[NC], pr[NC];
#pragma unroll
        for (int m = 0; m < NC; ++m) {
            gm[m] = plm_half(Pw0[m], Pw1[m], P2[0][m], theta);
            G2[0][m] = plm_half(Pw1[m], P2[0][m], P2[1][m], theta);
            pl[m] = Pw1[m] + gm[m];
            pr[m] = P2[0][m] - G2[0][m];
        }
        S::template face<1>(pl, pr, p2, td, tc, k0, c, F2[0]);
    }
    step(T_{}, F_{}, R2{}, T_{}, 4);
    // steady march, both stages; the last first-stage step (s = H + 3)
    // fetches nothing
    constexpr int NS = HE - 5;                 // steps with both stages
    constexpr int NL = (NS - 1) / 3;           // unrolled iterations
    constexpr int NR = NS - 3 * NL;            // 1 .. 3 steps after them
    int s = 5;
    for (int it = 0; it < NL; ++it, s += 3) {
        step(T_{}, T_{}, R0{}, T_{}, s);
        step(T_{}, T_{}, R1{}, T_{}, s + 1);
        step(T_{}, T_{}, R2{}, T_{}, s + 2);
    }
    if constexpr (NR == 1) {
        step(T_{}, T_{}, R0{}, F_{}, s);
    } else if constexpr (NR == 2) {
        step(T_{}, T_{}, R0{}, T_{}, s);
        step(T_{}, T_{}, R1{}, F_{}, s + 1);
    } else {
        step(T_{}, T_{}, R0{}, T_{}, s);
        step(T_{}, T_{}, R1{}, T_{}, s + 1);
        step(T_{}, T_{}, R2{}, F_{}, s + 2);
    }
    // the second stage's last row
    step(F_{}, T_{}, std::integral_constant<int, NR % 3>{}, F_{}, HE);
    if (!ok) report_error(a.err, 1, b, 0);
    if constexpr (S::kCfl) {
        if (a.rate_slot) workgroup_atomic_max<TJ>(rate, a.rate_slot);
    }
}

} // namespace m3p
