// mhd_scheme_v2 array functions (src/scheme_mhd_v2.cpp:75-278) as device
// kernels over one block of any shape: FP64 in the reference's association
// order, adiabatic index 5/3 (src/scheme_mhd_v2.cpp:48).
//
// Layout, row-major with the last axis fastest as the reference's arrays,
// components outermost:
//   cell tuples   [nq][s1][s2][s3]  (nq = 5 conserved, 8 primitive, or 8
//                                    Godunov data = 5 fluxes + 3 face E)
//   faces, edges  [s1][s2][s3]      scalars
// For cells (n1, n2, n3): x faces (n1+1, n2, n3), x edges (n1, n2+1, n3+1),
// and the other axes by symmetry.
//
// Every kernel is elementwise over its outputs (grid-stride; blockIdx.y picks
// the axis where a function returns three arrays) and HBM-bound: each output
// reads at most two cell tuples or four scalars.
#pragma once
#include "mhd.hpp"

namespace m3p {

constexpr double kV2Gamma = 5. / 3;

struct Ptr27 { const double* p[27]; };

__device__ __forceinline__ long v2_ix(int i, int j, int k, int s2, int s3) { return ((long)i * s2 + j) * s3 + k; }

__device__ __forceinline__ void v2_unravel(long c, int s2, int s3, int& i, int& j, int& k)
{
    k = (int)(c % s3);
    long t = c / s3;
    j = (int)(t % s2);
    i = (int)(t / s2);
}

#define M3P_GRID_STRIDE(c, total) \
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < (total); c += (long)gridDim.x * blockDim.x)

// primitive_array :75-81 with face_to_cell (src/mesh_cartesian_3d.hpp:235-245)
__global__ void __launch_bounds__(256)
mhd_v2_prim_kernel(int n1, int n2, int n3, const double* uc, const double* bf1, const double* bf2, const double* bf3,
                   double* pc, int* err)
{
    const long nc = (long)n1 * n2 * n3;
    M3P_GRID_STRIDE(c, nc) {
        int i, j, k;
        v2_unravel(c, n2, n3, i, j, k);
        double b[3] = { 0.5 * (bf1[v2_ix(i, j, k, n2, n3)] + bf1[v2_ix(i + 1, j, k, n2, n3)]),
                        0.5 * (bf2[v2_ix(i, j, k, n2 + 1, n3)] + bf2[v2_ix(i, j + 1, k, n2 + 1, n3)]),
                        0.5 * (bf3[v2_ix(i, j, k, n2, n3 + 1)] + bf3[v2_ix(i, j, k + 1, n2, n3 + 1)]) };
        double u[5], p[8];
#pragma unroll
        for (int q = 0; q < 5; ++q) u[q] = uc[q * nc + c];
        int st = mhd_recover(u, b, kV2Gamma, p);
        if (st) report_error(err, st, 0, (int)c);
#pragma unroll
        for (int q = 0; q < 8; ++q) pc[q * nc + c] = p[q];
    }
}

// godunov_fluxes :168-185: first-order face states (the cells either side of
// the face), HLLE + face electric field with the face's own normal field.
// pc (m1, m2, m3); axis-a faces and gf_a have one fewer entry along a.
__global__ void __launch_bounds__(256)
mhd_v2_godunov_kernel(int m1, int m2, int m3, const double* pc, const double* bf1, const double* bf2, const double* bf3,
                      double* gf1, double* gf2, double* gf3)
{
    const int a = blockIdx.y;
    const int s1 = m1 - (a == 0), s2 = m2 - (a == 1), s3 = m3 - (a == 2);
    const long nf = (long)s1 * s2 * s3, ncell = (long)m1 * m2 * m3;
    const double* bf = a == 0 ? bf1 : (a == 1 ? bf2 : bf3);
    double* gf = a == 0 ? gf1 : (a == 1 ? gf2 : gf3);
    M3P_GRID_STRIDE(f, nf) {
        int i, j, k;
        v2_unravel(f, s2, s3, i, j, k);
        const long cl = v2_ix(i, j, k, m2, m3);
        const long cr = v2_ix(i + (a == 0), j + (a == 1), k + (a == 2), m2, m3);
        double pl[8], pr[8], fl[5], e[3];
#pragma unroll
        for (int q = 0; q < 8; ++q) { pl[q] = pc[q * ncell + cl]; pr[q] = pc[q * ncell + cr]; }
        const double bl = bf[f];
        if (a == 0) Mhd<double>::riemann<1>(pl, pr, bl, kV2Gamma, fl, e);
        else if (a == 1) Mhd<double>::riemann<2>(pl, pr, bl, kV2Gamma, fl, e);
        else Mhd<double>::riemann<3>(pl, pr, bl, kV2Gamma, fl, e);
#pragma unroll
        for (int q = 0; q < 5; ++q) gf[q * nf + f] = fl[q];
#pragma unroll
        for (int q = 0; q < 3; ++q) gf[(5 + q) * nf + f] = e[q];
    }
}

// electromotive_forces :191-214: each edge field is the mean of two face
// estimates, each averaged along its other transverse axis (nd::adjacent_mean,
// src/core_ndarray_ops.hpp:84-90).  gf from cells (m1, m2, m3):
// ee1 (m1, m2-1, m3-1), ee2 (m1-1, m2, m3-1), ee3 (m1-1, m2-1, m3).
__global__ void __launch_bounds__(256)
mhd_v2_emf_kernel(int m1, int m2, int m3, const double* gf1, const double* gf2, const double* gf3,
                  double* ee1, double* ee2, double* ee3)
{
    const int a = blockIdx.y;
    const long n1f = (long)(m1 - 1) * m2 * m3, n2f = (long)m1 * (m2 - 1) * m3, n3f = (long)m1 * m2 * (m3 - 1);
    const int s1 = m1 - (a != 0), s2 = m2 - (a != 1), s3 = m3 - (a != 2);
    const long ne = (long)s1 * s2 * s3;
    double* ee = a == 0 ? ee1 : (a == 1 ? ee2 : ee3);
    M3P_GRID_STRIDE(c, ne) {
        int i, j, k;
        v2_unravel(c, s2, s3, i, j, k);
        double x, y;
        if (a == 0) {         // gf2[5] along axis 3, gf3[5] along axis 2
            const double* g2 = gf2 + 5 * n2f;
            const double* g3 = gf3 + 5 * n3f;
            x = 0.5 * (g2[v2_ix(i, j, k, m2 - 1, m3)] + g2[v2_ix(i, j, k + 1, m2 - 1, m3)]);
            y = 0.5 * (g3[v2_ix(i, j, k, m2, m3 - 1)] + g3[v2_ix(i, j + 1, k, m2, m3 - 1)]);
        } else if (a == 1) {  // gf3[6] along axis 1, gf1[6] along axis 3
            const double* g3 = gf3 + 6 * n3f;
            const double* g1 = gf1 + 6 * n1f;
            x = 0.5 * (g3[v2_ix(i, j, k, m2, m3 - 1)] + g3[v2_ix(i + 1, j, k, m2, m3 - 1)]);
            y = 0.5 * (g1[v2_ix(i, j, k, m2, m3)] + g1[v2_ix(i, j, k + 1, m2, m3)]);
        } else {              // gf1[7] along axis 2, gf2[7] along axis 1
            const double* g1 = gf1 + 7 * n1f;
            const double* g2 = gf2 + 7 * n2f;
            x = 0.5 * (g1[v2_ix(i, j, k, m2, m3)] + g1[v2_ix(i, j + 1, k, m2, m3)]);
            y = 0.5 * (g2[v2_ix(i, j, k, m2 - 1, m3)] + g2[v2_ix(i + 1, j, k, m2 - 1, m3)]);
        }
        ee[c] = 0.5 * (x + y);
    }
}

// updated_conserved_density :220-241: gf extended by one zone on the
// transverse axes (gf1 (n1+1, n2+2, n3+2), ...; remove_transverse_*,
// src/mesh_cartesian_3d.hpp:421-452), U - div(F) dt / dl
// (divergence_difference :290-301).
__global__ void __launch_bounds__(256)
mhd_v2_update_cells_kernel(int n1, int n2, int n3, const double* uc, const double* gf1, const double* gf2,
                           const double* gf3, double dt, double dl, double* out)
{
    const long nc = (long)n1 * n2 * n3;
    const long n1f = (long)(n1 + 1) * (n2 + 2) * (n3 + 2);
    const long n2f = (long)(n1 + 2) * (n2 + 1) * (n3 + 2);
    const long n3f = (long)(n1 + 2) * (n2 + 2) * (n3 + 1);
    M3P_GRID_STRIDE(c, nc) {
        int i, j, k;
        v2_unravel(c, n2, n3, i, j, k);
        const long a0 = v2_ix(i, j + 1, k + 1, n2 + 2, n3 + 2), a1 = v2_ix(i + 1, j + 1, k + 1, n2 + 2, n3 + 2);
        const long b0 = v2_ix(i + 1, j, k + 1, n2 + 1, n3 + 2), b1 = v2_ix(i + 1, j + 1, k + 1, n2 + 1, n3 + 2);
        const long c0 = v2_ix(i + 1, j + 1, k, n2 + 2, n3 + 1), c1 = v2_ix(i + 1, j + 1, k + 1, n2 + 2, n3 + 1);
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            double d1 = gf1[q * n1f + a1] - gf1[q * n1f + a0];
            double d2 = gf2[q * n2f + b1] - gf2[q * n2f + b0];
            double d3 = gf3[q * n3f + c1] - gf3[q * n3f + c0];
            out[q * nc + c] = uc[q * nc + c] - (d1 + d2 + d3) * dt / dl;
        }
    }
}

// Faces of cells (n1, n2, n3) from the solenoidal difference of edge arrays
// (src/mesh_cartesian_3d.hpp:265-276).  The edges are stored with o extra
// layers at each longitudinal end (o = 0: curl's e1 (n1, n2+1, n3+1), ...;
// o = 1: electromotive_forces' output for gf extended by one, read through
// rl1..rl3 of src/scheme_mhd_v2.cpp:67-69).  bf == null: curl :247-258,
// cf * dt / dl with a unit dt; else updated_magnetic_flux_density :264-278,
// bf - cf * dt / dl.
__global__ void __launch_bounds__(256)
mhd_v2_faces_kernel(int n1, int n2, int n3, int o, const double* e1, const double* e2, const double* e3,
                    const double* bf1, const double* bf2, const double* bf3, double dt, double dl,
                    double* b1, double* b2, double* b3)
{
    const int a = blockIdx.y;
    const int s1 = n1 + (a == 0), s2 = n2 + (a == 1), s3 = n3 + (a == 2);
    const long nf = (long)s1 * s2 * s3;
    const double* bf = a == 0 ? bf1 : (a == 1 ? bf2 : bf3);
    double* out = a == 0 ? b1 : (a == 1 ? b2 : b3);
    auto E1 = [&](int i, int j, int k) { return e1[v2_ix(i + o, j, k, n2 + 1, n3 + 1)]; };
    auto E2 = [&](int i, int j, int k) { return e2[v2_ix(i, j + o, k, n2 + 2 * o, n3 + 1)]; };
    auto E3 = [&](int i, int j, int k) { return e3[v2_ix(i, j, k + o, n2 + 1, n3 + 2 * o)]; };
    M3P_GRID_STRIDE(f, nf) {
        int i, j, k;
        v2_unravel(f, s2, s3, i, j, k);
        double cf;
        if (a == 0) cf = (E3(i, j + 1, k) - E3(i, j, k)) - (E2(i, j, k + 1) - E2(i, j, k));
        else if (a == 1) cf = (E1(i, j, k + 1) - E1(i, j, k)) - (E3(i + 1, j, k) - E3(i, j, k));
        else cf = (E2(i + 1, j, k) - E2(i, j, k)) - (E1(i, j + 1, k) - E1(i, j, k));
        out[f] = bf ? bf[f] - cf * dt / dl : cf * dt / dl;
    }
}

// extend_cell_primitive_variables :126-135: 27 blocks of bs^3 cells x nq
// components (row-major over the 3 x 3 x 3 offsets, tile_blocks
// src/mesh_cartesian_3d.hpp:307-337), cut to (bs + 2 count)^3 around the
// centre (remove_surface :410-419).
__global__ void __launch_bounds__(256)
mhd_v2_extend_cells_kernel(Ptr27 blk, int bs, int count, int nq, double* out)
{
    const int r = bs - count, m = bs + 2 * count;
    const long no = (long)m * m * m, nb = (long)bs * bs * bs;
    M3P_GRID_STRIDE(c, no * nq) {
        const int q = (int)(c / no);
        int i, j, k;
        v2_unravel(c % no, m, m, i, j, k);
        const int I = i + r, J = j + r, K = k + r;
        const double* b = blk.p[(I / bs) * 9 + (J / bs) * 3 + K / bs];
        out[c] = b[q * nb + v2_ix(I % bs, J % bs, K % bs, bs, bs)];
    }
}

// extend_face_magnetic_flux_density :141-162: for axis a, the face arrays
// blk.p[9a .. 9a+8] of the 3 x 3 square of blocks normal to a (row-major over
// the two transverse offsets), tiled and cut to count zones on each
// transverse side (remove_transverse src/mesh_cartesian_3d.hpp:458-464).
__global__ void __launch_bounds__(256)
mhd_v2_extend_faces_kernel(Ptr27 blk, int bs, int count, double* bf1, double* bf2, double* bf3)
{
    const int a = blockIdx.y;
    const int r = bs - count, m = bs + 2 * count;
    const int s1 = a == 0 ? bs + 1 : m, s2 = a == 1 ? bs + 1 : m, s3 = a == 2 ? bs + 1 : m;
    const int t1 = a == 0 ? bs + 1 : bs, t2 = a == 1 ? bs + 1 : bs, t3 = a == 2 ? bs + 1 : bs;
    double* out = a == 0 ? bf1 : (a == 1 ? bf2 : bf3);
    M3P_GRID_STRIDE(f, (long)s1 * s2 * s3) {
        int i, j, k;
        v2_unravel(f, s2, s3, i, j, k);
        const int I = a == 0 ? i : i + r, J = a == 1 ? j : j + r, K = a == 2 ? k : k + r;
        const int oi = I / t1, oj = J / t2, ok = K / t3;
        const int slot = a == 0 ? oj * 3 + ok : (a == 1 ? oi * 3 + ok : oi * 3 + oj);
        out[f] = blk.p[9 * a + slot][v2_ix(I % t1, J % t2, K % t3, t2, t3)];
    }
}

#undef M3P_GRID_STRIDE

} // namespace m3p
