// Guard fill across quad-tree level jumps for a list of leaf blocks:
// amr::extend_block (src/mesh_amr.hpp:360-393) over amr::get_or_create_block
// (:300-328), writing the extended primitive block pe [nleaf][5][n+4][n+4].
//
// Per leaf and face direction d (-x, +x, -y, +y) a recipe row
//   [kind, leaf, c00, c10, c01, c11, orthant_i, orthant_j]
// says where the neighbour's zones come from:
//   SAME   (0)  leaf `leaf` at the same level: copy
//   COARSE (1)  parent leaf `leaf`: refine_block = upsample + split_block,
//               zone (i, j) of the virtual neighbour is parent
//               [(oi n + i) / 2][(oj n + j) / 2]  (injection)
//   FINE   (2)  children leaves c00..c11: coarsen_blocks = tile_blocks +
//               downsample, 0.25 * (((T[2i][2j] + T[2i][2j+1]) + T[2i+1][2j]) + T[2i+1][2j+1])
// All copies are of primitives (the reference extends the primitive tree,
// src/module_euler.cpp:316-318).  Corner guard zones are not read by the
// Euler update and are left untouched.
#pragma once
#include "physics.hpp"

namespace m3p {

__device__ __forceinline__ void amr_prim(const double* u, long zpb, int leaf, int i, int j, int n,
                                         const EulerParams& prm, double* p, bool& ok)
{
    const double* z = u + (size_t)leaf * 5 * zpb + (size_t)i * n + j;
    double uu[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) uu[q] = z[q * zpb];
    ok &= Euler<double>::recover(uu, prm, p);
}

// Zone (i, j) of the virtual block that sits in direction d of leaf b.
__device__ __forceinline__ void amr_neighbour_zone(const double* u, const int* rec, int n, int i, int j,
                                                   const EulerParams& prm, double* p, bool& ok)
{
    const long zpb = (long)n * n;
    const int kind = rec[0];
    if (kind == 0) {
        amr_prim(u, zpb, rec[1], i, j, n, prm, p, ok);
    } else if (kind == 1) {
        amr_prim(u, zpb, rec[1], (rec[6] * n + i) >> 1, (rec[7] * n + j) >> 1, n, prm, p, ok);
    } else {
        double a[5], b[5], c[5], e[5];
        auto tile = [&] (int I, int J, double* out) {
            const int child = rec[2 + (I / n) + 2 * (J / n)];
            amr_prim(u, zpb, child, I % n, J % n, n, prm, out, ok);
        };
        tile(2 * i, 2 * j, a);
        tile(2 * i, 2 * j + 1, b);
        tile(2 * i + 1, 2 * j, c);
        tile(2 * i + 1, 2 * j + 1, e);
#pragma unroll
        for (int q = 0; q < 5; ++q) p[q] = 0.25 * (a[q] + b[q] + c[q] + e[q]);
    }
}

// One thread per extended zone (i, j) in [0, n+4)^2 of each leaf.
__global__ void __launch_bounds__(256)
amr_guard_fill_kernel(const double* u, double* pe, const int* recipes, int nleaf, int n, EulerParams prm, int* err)
{
    const int e = n + 4;
    const long per = (long)e * e;
    const long z = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= per * nleaf) return;
    const int b = (int)(z / per);
    const int r = (int)(z % per);
    const int i = r / e - 2, j = r % e - 2;
    const bool in_i = i >= 0 && i < n, in_j = j >= 0 && j < n;
    if (!in_i && !in_j) return;   // corner guard zones: unused by the update
    double p[5];
    bool ok = true;
    if (in_i && in_j) {
        amr_prim(u, (long)n * n, b, i, j, n, prm, p, ok);
    } else {
        int d, ii = i, jj = j;
        if (i < 0)       { d = 0; ii = i + n; }
        else if (i >= n) { d = 1; ii = i - n; }
        else if (j < 0)  { d = 2; jj = j + n; }
        else             { d = 3; jj = j - n; }
        amr_neighbour_zone(u, recipes + ((size_t)b * 4 + d) * 8, n, ii, jj, prm, p, ok);
    }
    double* o = pe + (size_t)b * 5 * per + (size_t)(i + 2) * e + (j + 2);
#pragma unroll
    for (int q = 0; q < 5; ++q) o[q * per] = p[q];
    if (!ok) report_error(err, 1, b, (i * n + j));
}

// Flux correction at level jumps (the conservative quad-tree mode; the
// reference has none, src/mesh_amr.hpp:300-328, problems/minidisk.cpp:60-62).
// The stage wrote every leaf's edge face fluxes into reg [nleaf][4][5][n]
// (edges -x, +x, -y, +y).  Where a leaf's edge borders two finer leaves, the
// flux through each of its edge faces becomes the mean of the two fine face
// fluxes covering it (the fine side's fluxes, per unit length), so what
// leaves one side enters the other: the zone next to the edge changes by
// s dt / dl (F_fine_mean - F_own), s = +1 on the lower edges (-x, -y) and -1
// on the upper ones, times w = 1 - rk_b after an RK combine.  One launch per
// axis (phase 0: x edges, rows 0 and n-1; phase 1: y edges, columns 0 and
// n-1), so a corner zone takes its two corrections in a fixed order.
__global__ void __launch_bounds__(256)
amr_reflux_kernel(double* out, const double* reg, const int* recipes, const double* dl_leaf, int nleaf, int n,
                  double dt, const double* dt_dev, double w, int phase)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)nleaf * 2 * n) return;
    const int b = (int)(idx / (2 * n));
    const int e = 2 * phase + (int)(idx / n % 2);       // edge
    const int i = (int)(idx % n);                       // position along the edge
    const int* rec = recipes + ((size_t)b * 4 + e) * 8;
    if (rec[0] != 2) return;
    // the two finer leaves' faces covering face i: children on the side
    // facing this leaf (recipe children in the order (0,0), (1,0), (0,1),
    // (1,1) of their (x, y) position inside the neighbour block)
    const int half = (2 * i) / n, fi = 2 * i - half * n;
    int child, fe;
    if (e == 0)      { child = rec[2 + 1 + 2 * half]; fe = 1; }
    else if (e == 1) { child = rec[2 + 0 + 2 * half]; fe = 0; }
    else if (e == 2) { child = rec[2 + half + 2];     fe = 3; }
    else             { child = rec[2 + half];         fe = 2; }
    const int r = e == 0 ? 0 : (e == 1 ? n - 1 : i);
    const int c = e == 2 ? 0 : (e == 3 ? n - 1 : i);
    const double step = dt_dev ? *dt_dev : dt;
    const double dl = dl_leaf[b];
    const double s = (e == 0 || e == 2) ? 1.0 : -1.0;
    const size_t zpb = (size_t)n * n;
#pragma unroll
    for (int m = 0; m < 5; ++m) {
        const double* fr = reg + ((size_t)child * 4 + fe) * 5 * n + (size_t)m * n;
        const double fmean = 0.5 * (fr[fi] + fr[fi + 1]);
        const double own = reg[(((size_t)b * 4 + e) * 5 + m) * n + i];
        const double corr = s * ((fmean - own) * step / dl);
        double* z = out + ((size_t)b * 5 + m) * zpb + (size_t)r * n + c;
        *z = *z + corr * w;
    }
}

} // namespace m3p
