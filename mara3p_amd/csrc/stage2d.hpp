// The fused 2D stage kernel: guard fill + recover + PLM + Riemann + source
// terms + conserved update (+ RK combine) in one pass per RK stage.
//
// Work decomposition ("tile march").  A workgroup of TJ threads owns TJ
// adjacent columns (the contiguous axis) of one block and marches down H
// rows.  Each thread keeps a three-row register window of primitives for the
// x (row-axis) stencil, so every x face flux is computed once and carried to
// the next row.  The y (column-axis) stencil goes through three LDS rows
// (primitives, slopes, fluxes); the two halo columns on each side and the
// tile's right-edge y flux for all H rows are computed once, up front, by a
// halo pre-pass, so no lane does redundant per-row work.
//
// Guard zones are never materialised: a zone outside the block is read
// directly from the neighbouring block through the neighbour table (the
// periodic next_on/prev_on wrap of src/mesh_block_index.hpp:334-335 when the
// table is built for a uniform mesh), which is amr::extend_block
// (src/mesh_amr.hpp:360-393) fused into the load.
#pragma once
#include "schemes2d.hpp"

namespace m3p {

// Conserved state of a whole mesh: [nblk][NC][n][n], neighbour table
// [nblk][4] = {prev on axis 0, next on axis 0, prev on axis 1, next on axis 1}.
template<class S>
struct MeshLoader {
    const double* u;
    const int* nbr;
    int n;
    // The four neighbours of the block being updated, read once per
    // workgroup by bind() (wave-uniform, kept in scalar registers).
    int bxm = 0, bxp = 0, bym = 0, byp = 0;

    __device__ __forceinline__ void bind(int b)
    {
        bxm = nbr[4 * b + 0];
        bxp = nbr[4 * b + 1];
        bym = nbr[4 * b + 2];
        byp = nbr[4 * b + 3];
    }
    // Pointer to zone (r, c) of the bound block b; rows and columns one
    // block outside [0, n) are taken from the neighbouring block.  Corners
    // take two hops through the table and follow the reference's
    // extend_block (src/mesh_amr.hpp:360-393, problems/minidisk.cpp:87-120):
    // its corner blocks c02 = prev_on(next_on(b, 0), 1) and c20 =
    // next_on(prev_on(b, 0), 1) sit at the (x-prev, y-next) and (x-next,
    // y-prev) corners, i.e. those two corners read the diagonally opposite
    // neighbour.  Only the viscous kernels read corners.
    __device__ __forceinline__ const double* zone(int b, int r, int c) const
    {
        // The four neighbour indices are copied to locals first: selecting
        // between the loader's own fields keeps the loader in scratch memory.
        const int xm = bxm, xp = bxp, ym = bym, yp = byp;
        if (r < 0) {
            r += n;
            if (c < 0)       { b = nbr[4 * xm + 2]; c += n; }
            else if (c >= n) { b = nbr[4 * xp + 2]; c -= n; }
            else b = xm;
        } else if (r >= n) {
            r -= n;
            if (c < 0)       { b = nbr[4 * xm + 3]; c += n; }
            else if (c >= n) { b = nbr[4 * xp + 3]; c -= n; }
            else b = xp;
        } else if (c < 0)  { b = ym; c += n; }
        else if (c >= n)   { b = yp; c -= n; }
        return u + ((size_t)b * S::NC * n + r) * n + c;
    }
    // Start of row r of the bound block b (wave-uniform); columns must lie in [0, n).
    __device__ __forceinline__ const double* row(int b, int r) const
    {
        if (r < 0)        { b = bxm; r += n; }
        else if (r >= n)  { b = bxp; r -= n; }
        return u + ((size_t)b * S::NC * n + r) * n;
    }
    __device__ __forceinline__ void cons_at(const double* z, double* uu) const
    {
#pragma unroll
        for (int q = 0; q < S::NC; ++q) uu[q] = z[(size_t)q * n * n];
    }
    __device__ __forceinline__ void cons(int b, int r, int c, double* uu) const { cons_at(zone(b, r, c), uu); }
    // Returns false when the recovery flagged the zone (negative density).
    __device__ __forceinline__ bool prim(int b, int r, int c, const typename S::Params& prm,
                                         typename S::Real* p) const
    {
        double uu[S::NC];
        cons(b, r, c, uu);
        return S::recover(uu, prm, p);
    }
    // Interior-column variants used by the march (c in [0, n)).
    __device__ __forceinline__ void cons_row(int b, int r, int c, double* uu) const { cons_at(row(b, r) + c, uu); }
    // Conserved state and primitives of zone (r, c) in one load.
    __device__ __forceinline__ bool load_row(int b, int r, int c, const typename S::Params& prm,
                                             double* uu, typename S::Real* p) const
    {
        cons_at(row(b, r) + c, uu);
        return S::recover(uu, prm, p);
    }
    __device__ __forceinline__ bool prim_row(int b, int r, int c, const typename S::Params& prm,
                                             typename S::Real* p) const
    {
        double uu[S::NC];
        cons_at(row(b, r) + c, uu);
        return S::recover(uu, prm, p);
    }
    // Split load for software pipelining: fetch the raw conserved state of a
    // row zone one iteration ahead, recover it when it is needed.
    static constexpr bool kPrefetch = true;
    static constexpr bool kUnroll = true;    // march unrolled by the scheme's kMarchUnroll
    static constexpr bool kEdgeFluxes = false;   // no edge-flux register (uniform meshes)
    __device__ __forceinline__ void fetch_row(int b, int r, int c, double* raw) const { cons_row(b, r, c, raw); }
    __device__ __forceinline__ bool finish_row(int, int, int, const typename S::Params& prm, const double* raw,
                                               double* uu, typename S::Real* p) const
    {
#pragma unroll
        for (int q = 0; q < S::NC; ++q) uu[q] = raw[q];
        return S::recover(uu, prm, p);
    }
};

// One block given as the reference's per-block arguments: uc [NC][n][n]
// conserved and pe [NC][n+4][n+4] primitives with two guard zones.
template<class S>
struct ExtendedLoader {
    const double* uc;
    const double* pe;
    int n;
    static constexpr bool kPrefetch = false;
    static constexpr bool kUnroll = false;
    static constexpr bool kEdgeFluxes = false;
    __device__ __forceinline__ void bind(int) {}

    // Block b of a list of blocks: uc + b NC n^2, pe + b NC (n+4)^2.
    __device__ __forceinline__ void cons(int b, int r, int c, double* uu) const
    {
        const double* base = uc + (size_t)b * S::NC * n * n;
#pragma unroll
        for (int q = 0; q < S::NC; ++q) uu[q] = base[((size_t)q * n + r) * n + c];
    }
    __device__ __forceinline__ bool prim(int b, int r, int c, const typename S::Params&,
                                         typename S::Real* p) const
    {
        const int e = n + 4;
        const double* base = pe + (size_t)b * S::NC * e * e;
#pragma unroll
        for (int q = 0; q < S::NC; ++q) p[q] = typename S::Real(base[((size_t)q * e + (r + 2)) * e + (c + 2)]);
        return true;
    }
    __device__ __forceinline__ void cons_row(int b, int r, int c, double* uu) const { cons(b, r, c, uu); }
    __device__ __forceinline__ bool prim_row(int b, int r, int c, const typename S::Params& prm,
                                             typename S::Real* p) const
    {
        return prim(b, r, c, prm, p);
    }
    // Primitives from pe; the conserved state from uc for interior rows only
    // (guard rows are never updated).
    __device__ __forceinline__ bool load_row(int b, int r, int c, const typename S::Params& prm,
                                             double* uu, typename S::Real* p) const
    {
        if (r >= 0 && r < n) cons(b, r, c, uu);
        else for (int q = 0; q < S::NC; ++q) uu[q] = 0.0;
        return prim(b, r, c, prm, p);
    }
};

// Leaves of a 2:1-balanced quad-tree, [nleaf][NC][n][n].  Guard zones are
// read through a per-leaf recipe [nleaf][4][8] for the directions -x, +x,
// -y, +y: {kind, leaf, child00, child10, child01, child11, orthant_i,
// orthant_j}.  amr::get_or_create_block (src/mesh_amr.hpp:300-328) fused into
// the load:
//   kind 0  same-level leaf: its zone
//   kind 1  coarser leaf (refine_block = upsample + split_block): the parent
//           zone ((o n + i) / 2, (o n + j) / 2), piecewise-constant injection
//   kind 2  four finer leaves (coarsen_blocks = tile_blocks + downsample):
//           0.25 (((a00 + a01) + a10) + a11) of the four primitives
// The tile march reads guard rows at interior columns and guard columns at
// interior rows only, never a corner.
template<class S>
struct TreeLoader {
    using R = typename S::Real;
    const double* u;
    const int* recipes;
    int n;
    __device__ __forceinline__ void bind(int) {}

    __device__ __forceinline__ bool prim_of(int leaf, int r, int c, const typename S::Params& prm, R* p) const
    {
        double uu[S::NC];
        const double* z = u + ((size_t)leaf * S::NC * n + r) * n + c;
#pragma unroll
        for (int q = 0; q < S::NC; ++q) uu[q] = z[(size_t)q * n * n];
        return S::recover(uu, prm, p);
    }
    // Zone (i, j) of the virtual neighbour block in direction d of leaf b.
    __device__ __forceinline__ bool guard(int b, int d, int i, int j, const typename S::Params& prm, R* p) const
    {
        const int* rec = recipes + ((size_t)b * 4 + d) * 8;
        const int kind = rec[0];
        if (kind == 0) return prim_of(rec[1], i, j, prm, p);
        if (kind == 1) return prim_of(rec[1], (rec[6] * n + i) >> 1, (rec[7] * n + j) >> 1, prm, p);
        const int I = 2 * i, J = 2 * j;           // I, I+1 (J, J+1) lie in one child: n is even
        const int ci = I >= n, cj = J >= n;
        const int child = rec[2 + ci + 2 * cj];
        const int li = I - ci * n, lj = J - cj * n;
        R a[S::NC], b1[S::NC], c1[S::NC], e[S::NC];
        bool ok = prim_of(child, li, lj, prm, a);
        ok &= prim_of(child, li, lj + 1, prm, b1);
        ok &= prim_of(child, li + 1, lj, prm, c1);
        ok &= prim_of(child, li + 1, lj + 1, prm, e);
#pragma unroll
        for (int q = 0; q < S::NC; ++q) p[q] = R(0.25) * (((a[q] + b1[q]) + c1[q]) + e[q]);
        return ok;
    }
    // Interior row r, any column (guard columns come from the -y / +y recipes).
    __device__ __forceinline__ bool prim(int b, int r, int c, const typename S::Params& prm, R* p) const
    {
        if (c < 0)  return guard(b, 2, r, c + n, prm, p);
        if (c >= n) return guard(b, 3, r, c - n, prm, p);
        return prim_of(b, r, c, prm, p);
    }
    // Interior column c, any row (guard rows come from the -x / +x recipes).
    __device__ __forceinline__ bool prim_row(int b, int r, int c, const typename S::Params& prm, R* p) const
    {
        if (r < 0)  return guard(b, 0, r + n, c, prm, p);
        if (r >= n) return guard(b, 1, r - n, c, prm, p);
        return prim_of(b, r, c, prm, p);
    }
    __device__ __forceinline__ void cons_row(int b, int r, int c, double* uu) const
    {
        const double* z = u + ((size_t)b * S::NC * n + r) * n + c;
#pragma unroll
        for (int q = 0; q < S::NC; ++q) uu[q] = z[(size_t)q * n * n];
    }
    // The conserved state is only kept for interior rows (guard rows are never updated).
    __device__ __forceinline__ bool load_row(int b, int r, int c, const typename S::Params& prm,
                                             double* uu, R* p) const
    {
        if (r >= 0 && r < n) {
            cons_row(b, r, c, uu);
            return S::recover(uu, prm, p);
        }
#pragma unroll
        for (int q = 0; q < S::NC; ++q) uu[q] = 0.0;
        return prim_row(b, r, c, prm, p);
    }
    // Software pipelining as MeshLoader for interior rows; a guard row is
    // assembled when it is needed.
    static constexpr bool kPrefetch = true;
    static constexpr bool kUnroll = false;
    static constexpr bool kEdgeFluxes = true;    // a.flux_reg may be set (conservative quad-tree mode)
    __device__ __forceinline__ void fetch_row(int b, int r, int c, double* raw) const
    {
        if (r >= 0 && r < n) cons_row(b, r, c, raw);
    }
    __device__ __forceinline__ bool finish_row(int b, int r, int c, const typename S::Params& prm, const double* raw,
                                               double* uu, R* p) const
    {
        if (r >= 0 && r < n) {
#pragma unroll
            for (int q = 0; q < S::NC; ++q) uu[q] = raw[q];
            return S::recover(uu, prm, p);
        }
#pragma unroll
        for (int q = 0; q < S::NC; ++q) uu[q] = 0.0;
        return r < 0 ? guard(b, 0, r + n, c, prm, p) : guard(b, 1, r - n, c, prm, p);
    }
};

struct StageArgs {
    const double* u0;      // RK base state (same layout as out) or nullptr
    double* out;           // [nblk][NC][n][n]
    const double* dt_dev;  // device-resident dt (graph replay) or nullptr
    double dt;
    double dl;
    const double* dl_blk;  // per-block cell spacing (quad-tree levels) or nullptr for dl
    double rk_b;           // out = u0 * rk_b + L(u) * (1 - rk_b) when u0 != nullptr
    const int* bcoord;     // [nblk][2] block coordinates, or nullptr (single block at bi0,bj0)
    int bi0, bj0;
    int n;                 // block size
    int nblk;
    int* err;
    // fused CFL reduction (schemes with kCfl): atomic max of the rate
    // max(|v1|, |v2|) + c_s over dl of the values written, or nullptr
    unsigned long long* rate_slot;
    // the face fluxes on the four edges of every block, [nblk][4][NC][n]
    // (edges -x, +x, -y, +y; index along the edge), for the flux correction
    // at quad-tree level jumps, or nullptr
    double* flux_reg;
};

// Parameters that a stage kernel takes as they are; a scheme whose
// parameters may point at device memory overloads this (found by
// argument-dependent lookup at the kernel's instantiation).
template<class P> __device__ __forceinline__ void resolve_params(P&) {}

// Edge face flux of block b into the flux register (edge e, position i).
template<int NC, typename R>
__device__ __forceinline__ void store_edge_flux(double* reg, int b, int e, int n, int i, const R* f)
{
#pragma unroll
    for (int m = 0; m < NC; ++m) reg[(((size_t)b * 4 + e) * NC + m) * n + i] = double(f[m]);
}

// KEEPU: keep each zone's conserved state in a three-row register window from
// its load to its update (no second read of the state); otherwise re-load it
// at the update (fewer registers).  FIRST: a stage without a base state
// (a.u0 == nullptr is then assumed); it has fewer loads in flight per row, so
// a scheme may ask it to fetch more rows ahead (S::kPrefetchRows).
template<class S, class Ld, int H, int TJ, bool KEEPU, bool FIRST = false>
__global__ void __launch_bounds__(TJ) __attribute__((amdgpu_waves_per_eu(S::kWavesPerEU, 8)))
stage2d_kernel(typename S::Params prm, Ld ld_arg, StageArgs a)
{
    constexpr int NC = S::NC;
    using R = typename S::Real;
    resolve_params(prm);

    const int n = a.n;
    const int tiles_i = n / H, tiles_j = n / TJ;
    const int per_block = tiles_i * tiles_j;
    // RK stages with a base state (the second stage) visit the tiles in the
    // reverse order of the first: the tiles a stage reads first are the ones
    // the previous stage wrote last, which are the likeliest to still be in
    // the memory-side cache.
    const int w0 = xcd_remap(blockIdx.x, gridDim.x);
    const int w = a.u0 ? (int)gridDim.x - 1 - w0 : w0;
    TileCtx tc;
    tc.b = w / per_block;
    const int rem = w % per_block;
    tc.k0 = (rem / tiles_j) * H;
    tc.c0 = (rem % tiles_j) * TJ;
    tc.n = n;
    if (a.bcoord) { tc.bi = a.bcoord[2 * tc.b]; tc.bj = a.bcoord[2 * tc.b + 1]; }
    else          { tc.bi = a.bi0; tc.bj = a.bj0; }

    const int t = threadIdx.x;
    const int c = tc.c0 + t;
    const int b = tc.b;
    const double dt = a.dt_dev ? *a.dt_dev : a.dt;
    const double dl = a.dl_blk ? a.dl_blk[b] : a.dl;
    Ld ld = ld_arg;
    ld.bind(b);
    double rate = 0.0;                                    // fused CFL reduction
    const double idl = a.rate_slot ? 1.0 / dl : 0.0;

    __shared__ typename S::template TileData<H, TJ> td;
    __shared__ R hraw[7][NC][H];     // pre-pass primitives at 7 edge columns
    __shared__ R hG[NC][H];          // y half slope at column c0-1
    __shared__ R hF[NC][H];          // y flux at face c0+TJ
    __shared__ R rowP[NC][TJ + 2];   // [t+1] = column c0+t; [0], [TJ+1] from hraw
    __shared__ R rowG[NC][TJ + 1];   // [t+1] = half slope at column c0+t; [0] from hG
    __shared__ R rowF[NC][TJ + 1];   // [t] = y flux at face c0+t; [TJ] from hF

    S::setup(td, prm, tc);
    bool ok = true;

    // ---- halo pre-pass --------------------------------------------------
    for (int idx = t; idx < 7 * H; idx += TJ) {
        const int r = idx / 7, q = idx % 7;
        const int cols[7] = { tc.c0 - 2, tc.c0 - 1, tc.c0, tc.c0 + TJ - 2, tc.c0 + TJ - 1, tc.c0 + TJ, tc.c0 + TJ + 1 };
        R p[NC];
        ok &= ld.prim(b, tc.k0 + r, cols[q], prm, p);
#pragma unroll
        for (int m = 0; m < NC; ++m) hraw[q][m][r] = p[m];
    }
    __syncthreads();
    const R theta = R(S::theta(prm));
    for (int r = t; r < H; r += TJ) {
        R pl[NC], pr[NC], f[NC];
#pragma unroll
        for (int m = 0; m < NC; ++m) {
            hG[m][r] = plm_half(hraw[0][m][r], hraw[1][m][r], hraw[2][m][r], theta);
            R g1 = plm_half(hraw[3][m][r], hraw[4][m][r], hraw[5][m][r], theta);
            R g2 = plm_half(hraw[4][m][r], hraw[5][m][r], hraw[6][m][r], theta);
            pl[m] = hraw[4][m][r] + g1;
            pr[m] = hraw[5][m][r] - g2;
        }
        S::template face<2>(pl, pr, prm, td, tc, tc.k0 + r, tc.c0 + TJ, f);
#pragma unroll
        for (int m = 0; m < NC; ++m) hF[m][r] = f[m];
        if (Ld::kEdgeFluxes && a.flux_reg && tc.c0 + TJ == n) store_edge_flux<NC>(a.flux_reg, b, 3, n, tc.k0 + r, f);
    }

    // ---- march ------------------------------------------------------------
    // Tiles in even tile rows march down (increasing row index), odd ones up,
    // so two vertically adjacent tiles reach the rows they share at about the
    // same time (both at their start or both at their end) and the second
    // read of those halo rows hits L2.  Every slope, face state and face flux
    // is computed from the same operands in the same order in both
    // directions, so the results do not depend on the direction.
    //
    // Edge lanes read their out-of-tile neighbours from the pre-pass arrays
    // through a selected address (no divergent branch).  Loaders that
    // support it are software-pipelined: the raw state of the row three
    // ahead and the RK base of the current row are fetched at the top of an
    // iteration, so their latency overlaps the iteration's arithmetic (a
    // scheme with kLateBase loads the RK base at the update instead, holding
    // fewer registers across the row).
    auto march = [&] (auto dir) {
        constexpr int D = decltype(dir)::value;
        const int kfirst = D > 0 ? tc.k0 : tc.k0 + H - 1;

        // warm-up: rows kfirst - 2D .. kfirst + D, flux at the face behind kfirst
        R Pa[NC], Pb[NC], ga[NC];   // current row, next row, half slope at the current row
        double Fb[NC];              // x flux at the face behind the current row
        double Ua[NC], Ub[NC];
        {
            R Pm2[NC], Pm1[NC], gm[NC], pl[NC], pr[NC];
            ok &= ld.prim_row(b, kfirst - 2 * D, c, prm, Pm2);
            ok &= ld.prim_row(b, kfirst - D, c, prm, Pm1);
            if constexpr (KEEPU) {
                ok &= ld.load_row(b, kfirst, c, prm, Ua, Pa);
                ok &= ld.load_row(b, kfirst + D, c, prm, Ub, Pb);
            } else {
                ok &= ld.prim_row(b, kfirst, c, prm, Pa);
                ok &= ld.prim_row(b, kfirst + D, c, prm, Pb);
            }
#pragma unroll
            for (int m = 0; m < NC; ++m) {
                if constexpr (D > 0) {
                    gm[m] = plm_half(Pm2[m], Pm1[m], Pa[m], theta);
                    ga[m] = plm_half(Pm1[m], Pa[m], Pb[m], theta);
                    pl[m] = Pm1[m] + gm[m];
                    pr[m] = Pa[m] - ga[m];
                } else {
                    gm[m] = plm_half(Pa[m], Pm1[m], Pm2[m], theta);
                    ga[m] = plm_half(Pb[m], Pa[m], Pm1[m], theta);
                    pl[m] = Pa[m] + ga[m];
                    pr[m] = Pm1[m] - gm[m];
                }
            }
            R f[NC];
            const int fw = D > 0 ? kfirst : kfirst + 1;
            S::template face<1>(pl, pr, prm, td, tc, fw, c, f);
#pragma unroll
            for (int m = 0; m < NC; ++m) Fb[m] = double(f[m]);
            if (Ld::kEdgeFluxes && a.flux_reg && (fw == 0 || fw == n)) store_edge_flux<NC>(a.flux_reg, b, fw == 0 ? 0 : 1, n, c, f);
        }
        __syncthreads();

        // Prefetched raw rows: Unx[i] = row k + (2 + i) D of the iteration
        // that follows, kAhead rows in flight.
        constexpr int kAhead = FIRST ? S::kPrefetchRows : 1;
        static_assert(kAhead >= 1 && kAhead <= H, "rows ahead within the tile");
        double Unx[kAhead][NC] = {};
        if constexpr (Ld::kPrefetch) {
#pragma unroll
            for (int i = 0; i < kAhead; ++i) ld.fetch_row(b, kfirst + (2 + i) * D, c, Unx[i]);
        }
        constexpr int kRowsPerIter = Ld::kUnroll ? S::kMarchUnroll : 1;
#pragma unroll kRowsPerIter
        for (int kk = 0; kk < H; ++kk) {
            const int k = kfirst + D * kk;
            const int kl = k - tc.k0;
            R Pc[NC], gb[NC], pl[NC], pr[NC], Fa[NC], gy[NC], Fy[NC];
            double Uc[NC], U0[NC];
            if constexpr (Ld::kPrefetch) {
                double raw[NC];
#pragma unroll
                for (int m = 0; m < NC; ++m) raw[m] = Unx[0][m];
#pragma unroll
                for (int i = 0; i + 1 < kAhead; ++i)
#pragma unroll
                    for (int m = 0; m < NC; ++m) Unx[i][m] = Unx[i + 1][m];
                // the last row a tile reads is kfirst + (H + 1) D
                if (kk + kAhead < H) ld.fetch_row(b, k + (2 + kAhead) * D, c, Unx[kAhead - 1]);
                if constexpr (!S::kLateBase) {
                    if (!FIRST && a.u0) {
                        const double* z0 = a.u0 + ((size_t)b * NC * n + k) * n + c;
#pragma unroll
                        for (int m = 0; m < NC; ++m) U0[m] = __builtin_nontemporal_load(&z0[(size_t)m * n * n]);
                    }
                }
                ok &= ld.finish_row(b, k + 2 * D, c, prm, raw, Uc, Pc);
            } else {
                if constexpr (KEEPU) ok &= ld.load_row(b, k + 2 * D, c, prm, Uc, Pc);
                else                 ok &= ld.prim_row(b, k + 2 * D, c, prm, Pc);
            }
            // slope at row k + D, flux at the face between rows k and k + D
#pragma unroll
            for (int m = 0; m < NC; ++m) {
                if constexpr (D > 0) {
                    gb[m] = plm_half(Pa[m], Pb[m], Pc[m], theta);
                    pl[m] = Pa[m] + ga[m];
                    pr[m] = Pb[m] - gb[m];
                } else {
                    gb[m] = plm_half(Pc[m], Pb[m], Pa[m], theta);
                    pl[m] = Pb[m] + gb[m];
                    pr[m] = Pa[m] - ga[m];
                }
            }
            const int fk = D > 0 ? k + 1 : k;
            S::template face<1>(pl, pr, prm, td, tc, fk, c, Fa);
            if (Ld::kEdgeFluxes && a.flux_reg && (fk == 0 || fk == n)) store_edge_flux<NC>(a.flux_reg, b, fk == 0 ? 0 : 1, n, c, Fa);

            // The two edge lanes also store the out-of-tile neighbours of this
            // row (from the pre-pass) at the ends of the LDS rows, so every
            // lane reads its neighbours at the same offsets, with no select.
#pragma unroll
            for (int m = 0; m < NC; ++m) rowP[m][t + 1] = Pa[m];
            if (t == 0) {
#pragma unroll
                for (int m = 0; m < NC; ++m) { rowP[m][0] = hraw[1][m][kl]; rowG[m][0] = hG[m][kl]; }
            }
            if (t == TJ - 1) {
#pragma unroll
                for (int m = 0; m < NC; ++m) rowP[m][TJ + 1] = hraw[5][m][kl];
            }
            __syncthreads();
#pragma unroll
            for (int m = 0; m < NC; ++m) {
                gy[m] = plm_half(rowP[m][t], Pa[m], rowP[m][t + 2], theta);
                rowG[m][t + 1] = gy[m];
            }
            __syncthreads();
#pragma unroll
            for (int m = 0; m < NC; ++m) {
                pl[m] = rowP[m][t] + rowG[m][t];
                pr[m] = Pa[m] - gy[m];
            }
            S::template face<2>(pl, pr, prm, td, tc, k, c, Fy);
#pragma unroll
            for (int m = 0; m < NC; ++m) rowF[m][t] = Fy[m];
            if (Ld::kEdgeFluxes && a.flux_reg && c == 0) store_edge_flux<NC>(a.flux_reg, b, 2, n, k, Fy);
            if (t == TJ - 1) {
#pragma unroll
                for (int m = 0; m < NC; ++m) rowF[m][TJ] = hF[m][kl];
            }
            // The update's loads (the zone's state where the kernel does not
            // keep it in registers, the RK base where the loader has no
            // pipeline) are issued before the barrier, which orders memory
            // and so would hold them after it, so that their latency overlaps
            // the wait (FP64 KH and Sedov steps 0.6 and 0.9 % faster).  A
            // scheme with kLateBase keeps its RK base load at the update: held
            // across the barrier it spills 20 registers.
            double uu[NC];
            if constexpr (!KEEPU) ld.cons_row(b, k, c, uu);
            if constexpr (!Ld::kPrefetch && !S::kLateBase) {
                if (!FIRST && a.u0) {
                    const double* z0 = a.u0 + ((size_t)b * NC * n + k) * n + c;
#pragma unroll
                    for (int m = 0; m < NC; ++m) U0[m] = __builtin_nontemporal_load(&z0[(size_t)m * n * n]);
                }
            }
            __syncthreads();

            double dfx[NC], dfy[NC], res[NC], fa[NC];
#pragma unroll
            for (int m = 0; m < NC; ++m) {
                fa[m] = double(Fa[m]);
                dfx[m] = D > 0 ? fa[m] - Fb[m] : Fb[m] - fa[m];
                dfy[m] = double(rowF[m][t + 1]) - double(Fy[m]);
            }
            if constexpr (KEEPU) {
#pragma unroll
                for (int m = 0; m < NC; ++m) uu[m] = Ua[m];
            }
            S::update(prm, td, tc, k, c, uu, Pa, dfx, dfy, dt, dl, res);
            // The RK base and the output are touched once per launch:
            // non-temporal loads and stores keep them from displacing the
            // halo rows that neighbouring tiles share in L2.
            double* o = a.out + ((size_t)b * NC * n + k) * n + c;
            double wv[NC];
            if (!FIRST && a.u0) {
                if constexpr (S::kLateBase) {
                    const double* z0 = a.u0 + ((size_t)b * NC * n + k) * n + c;
#pragma unroll
                    for (int m = 0; m < NC; ++m) U0[m] = __builtin_nontemporal_load(&z0[(size_t)m * n * n]);
                }
                if constexpr (is_fp64<R>::value) {
#pragma unroll
                    for (int m = 0; m < NC; ++m) wv[m] = U0[m] * a.rk_b + res[m] * (1.0 - a.rk_b);
                } else {
                    const double wb = 1.0 - a.rk_b;
#pragma unroll
                    for (int m = 0; m < NC; ++m) wv[m] = fma(U0[m], a.rk_b, res[m] * wb);
                }
            } else {
#pragma unroll
                for (int m = 0; m < NC; ++m) wv[m] = res[m];
            }
#pragma unroll
            for (int m = 0; m < NC; ++m) __builtin_nontemporal_store(wv[m], &o[(size_t)m * n * n]);
            if constexpr (S::kCfl) {
                if (a.rate_slot) rate = fmax(rate, S::cfl_speed(wv, prm) * idl);
            }
#pragma unroll
            for (int m = 0; m < NC; ++m) { Pa[m] = Pb[m]; Pb[m] = Pc[m]; ga[m] = gb[m]; Fb[m] = fa[m]; }
            if constexpr (KEEPU) {
#pragma unroll
                for (int m = 0; m < NC; ++m) { Ua[m] = Ub[m]; Ub[m] = Uc[m]; }
            }
        }
    };
    if (((rem / tiles_j) & 1) == 0) march(std::integral_constant<int, 1>{});
    else                            march(std::integral_constant<int, -1>{});
    if (!ok) report_error(a.err, 1, b, 0);
    if constexpr (S::kCfl) {
        if (a.rate_slot) workgroup_atomic_max<TJ>(rate, a.rate_slot);
    }
}

} // namespace m3p
