// Ideal MHD with constrained transport on a periodic N^3 cube:
// mara::advance_mhd_ct (src/scheme_mhd_v1.cpp:265-299) as three passes.
//
//   1. mhd_prim_kernel     cell primitives from U and the face-averaged B
//                          (primitive_array :214-224, face_to_cell
//                          src/mesh_cartesian_3d.hpp:235-245) + CFL max speed
//   2. mhd_faces_kernel    PLM face states, HLLE Godunov flux and face electric
//                          field of the three faces below each cell (and
//                          mhd_boundary_faces_kernel for the upper planes)
//                          (godunov_fluxes_plm :152-179,
//                          mhd::riemann_hlle src/physics_mhd.hpp:358-390)
//   3. mhd_update_kernel   U - div(F) dt / dl (divergence_difference
//                          src/mesh_cartesian_3d.hpp:290-301) and, for the
//                          faces below each cell, B_f - curl(E) dt / dl with
//                          edge EMFs averaged from the faces
//                          (edge_emf_from_face :76-98, solenoidal_difference
//                          src/mesh_cartesian_3d.hpp:265-276)
//
// Layouts (row-major, last axis fastest, as the reference arrays):
//   U [5][N][N][N];  B1 [N+1][N][N], B2 [N][N+1][N], B3 [N][N][N+1]
//   workspace: P [8][N^3], face fluxes F_A [5][nf], face fields E_A [2][nf]
//   (nf = (N+1) N^2, the transverse E components of axis-A faces).
#pragma once
#include "device_common.hpp"

namespace m3p {

template<typename R>
struct Mhd {
    __device__ static __forceinline__ R dot3(const R* a, const R* b)
    {
        if constexpr (is_fp64<R>::value) return ((R(0) + a[0] * b[0]) + a[1] * b[1]) + a[2] * b[2];
        else return fmaf(a[2], b[2], fmaf(a[1], b[1], a[0] * b[0]));
    }

    // conserved_density :307-319
    __device__ static __forceinline__ void cons(const R* p, R gamma, R* u)
    {
        const R* v = &p[1];
        const R* b = &p[5];
        R d = p[0];
        u[0] = d;
        u[1] = d * v[0];
        u[2] = d * v[1];
        u[3] = d * v[2];
        if constexpr (is_fp64<R>::value) u[4] = R(0.5) * d * dot3(v, v) + p[4] / (gamma - R(1)) + R(0.5) * dot3(b, b);
        else u[4] = fmaf(R(0.5) * d, dot3(v, v), fmaf(R(0.5), dot3(b, b), p[4] * frcp(gamma - R(1))));
    }

    // flux :324-347 along unit axis AXIS (1, 2, 3)
    template<int AXIS>
    __device__ static __forceinline__ void flux(const R* p, const R* u, R* f)
    {
        const R* v = &p[1];
        const R* b = &p[5];
        constexpr int a = AXIS - 1;
        if constexpr (is_fp64<R>::value) {
            R bn = R(0) + b[a];
            R vn = R(0) + v[a];
            R pt = p[4] + R(0.5) * dot3(b, b);
            R bv = dot3(b, v);
            R n[3] = { R(AXIS == 1), R(AXIS == 2), R(AXIS == 3) };
            f[0] = vn * u[0] + R(0) + R(0);
            f[1] = vn * u[1] + pt * n[0] + -bn * b[0];
            f[2] = vn * u[2] + pt * n[1] + -bn * b[1];
            f[3] = vn * u[3] + pt * n[2] + -bn * b[2];
            f[4] = vn * u[4] + pt * vn + -bn * bv;
        } else {
            R bn = b[a], vn = v[a];
            R pt = fmaf(R(0.5), dot3(b, b), p[4]);
            f[0] = vn * u[0];
            f[1] = fmaf(vn, u[1], -bn * b[0]);
            f[2] = fmaf(vn, u[2], -bn * b[1]);
            f[3] = fmaf(vn, u[3], -bn * b[2]);
            f[1 + a] += pt;
            f[4] = fmaf(vn, u[4] + pt, -bn * dot3(b, v));
        }
    }

    // induction :349-352 = n x E with E = -(v x B) (electric_field_vector :154-157)
    template<int AXIS>
    __device__ static __forceinline__ void induction(const R* p, R* out)
    {
        const R* v = &p[1];
        const R* b = &p[5];
        if constexpr (is_fp64<R>::value) {
            R E[3] = { -(v[1] * b[2] - v[2] * b[1]), -(v[2] * b[0] - v[0] * b[2]), -(v[0] * b[1] - v[1] * b[0]) };
            R n[3] = { R(AXIS == 1), R(AXIS == 2), R(AXIS == 3) };
            out[0] = n[1] * E[2] - n[2] * E[1];
            out[1] = n[2] * E[0] - n[0] * E[2];
            out[2] = n[0] * E[1] - n[1] * E[0];
        } else {
            // n is the unit vector of axis a: n x E has no a component and
            // its other two are -E_c and E_b (b, c the next axes in cyclic
            // order), without the general form's products with the zero
            // components of n (the same values up to the sign of a zero)
            constexpr int a = AXIS - 1, bx = (a + 1) % 3, c = (a + 2) % 3;
            out[a] = R(0);
            out[bx] = v[a] * b[bx] - v[bx] * b[a];
            out[c] = -(v[c] * b[a] - v[a] * b[c]);
        }
    }

    // outer_wavespeeds :216-246
    template<int AXIS>
    __device__ static __forceinline__ void wavespeeds(const R* p, R gamma, R& wm, R& wp)
    {
        constexpr int a = AXIS - 1;
        const R* b = &p[5];
        R d = p[0];
        if constexpr (is_fp64<R>::value) {
            // one reciprocal of the density for the three squared speeds,
            // bn * bn for pow(bn, 2), and the Newton-refined reciprocal and
            // square roots (each within a few ulp of the reference's
            // divisions, pow and sqrt)
            R vn = R(0) + p[1 + a];
            const R id = rcp_nr(d);
            R cs2 = gamma * p[4] * id;
            R ca2 = dot3(b, b) * id;
            R cf2 = cs2 + ca2;
            R bn = R(0) + b[a];
            R can2 = (bn * bn) * id;
            R cg2 = sqrt_nr(cf2 * cf2 - R(4) * cs2 * can2);
            R cf = sqrt_nr(R(0.5) * (cf2 + cg2));
            wm = vn - cf;
            wp = vn + cf;
        } else {
            R vn = p[1 + a];
            R id = frcp(d);
            R cs2 = gamma * p[4] * id;
            R ca2 = dot3(b, b) * id;
            R cf2 = cs2 + ca2;
            R can2 = b[a] * b[a] * id;
            R cg2 = fsqrt(fmaxf(fmaf(cf2, cf2, -R(4) * cs2 * can2), R(0)));
            R cf = fsqrt(R(0.5) * (cf2 + cg2));
            wm = vn - cf;
            wp = vn + cf;
        }
    }

    // riemann_hlle :358-390 -> Godunov flux f[5], face electric field e[3]
    template<int AXIS>
    __device__ static __forceinline__ void riemann(R* pl, R* pr, R bl, R gamma, R* f, R* e)
    {
        constexpr int a = AXIS - 1;
        // with_longitudinal_field :165-176: (b - (b.n) n) + bl n
        if constexpr (is_fp64<R>::value) {
            R bnl = R(0) + pl[5 + a], bnr = R(0) + pr[5 + a];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                R nk = R(k == a);
                pl[5 + k] = (pl[5 + k] - bnl * nk) + bl * nk;
                pr[5 + k] = (pr[5 + k] - bnr * nk) + bl * nk;
            }
        } else {
            pl[5 + a] = bl;
            pr[5 + a] = bl;
        }
        R ul[5], ur[5], fl[5], fr[5], il[3], ir[3], alm, alp, arm, arp;
        cons(pl, gamma, ul);
        cons(pr, gamma, ur);
        flux<AXIS>(pl, ul, fl);
        flux<AXIS>(pr, ur, fr);
        induction<AXIS>(pl, il);
        induction<AXIS>(pr, ir);
        wavespeeds<AXIS>(pl, gamma, alm, alp);
        wavespeeds<AXIS>(pr, gamma, arm, arp);
        R ap = smax(R(0), smax(alp, arp));
        R am = smin(R(0), smin(alm, arm));
        R hi[3];
        if constexpr (is_fp64<R>::value) {
            const R inv = rcp_nr(ap - am);   // one reciprocal for the eight components
#pragma unroll
            for (int k = 0; k < 5; ++k) f[k] = (ap * fl[k] - am * fr[k] - (ul[k] - ur[k]) * ap * am) * inv;
#pragma unroll
            for (int k = 0; k < 3; ++k) hi[k] = (ap * il[k] - am * ir[k] - (pl[5 + k] - pr[5 + k]) * ap * am) * inv;
        } else {
            R inv = frcp(ap - am), apam = ap * am;
#pragma unroll
            for (int k = 0; k < 5; ++k) f[k] = fmaf(ap, fl[k], fmaf(-am, fr[k], -(ul[k] - ur[k]) * apam)) * inv;
            // -n x hll_induction takes the two transverse components only
            // (b, c the next axes after a in cyclic order: e_b = hi_c,
            // e_c = -hi_b, e_a = 0)
            constexpr int bx = (a + 1) % 3, c = (a + 2) % 3;
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                if (k == a) continue;
                hi[k] = fmaf(ap, il[k], fmaf(-am, ir[k], -(pl[5 + k] - pr[5 + k]) * apam)) * inv;
            }
            e[a] = R(0);
            e[bx] = hi[c];
            e[c] = -hi[bx];
            return;
        }
        // -n x hll_induction
        R n[3] = { R(AXIS == 1), R(AXIS == 2), R(AXIS == 3) };
        e[0] = -(n[1] * hi[2] - n[2] * hi[1]);
        e[1] = -(n[2] * hi[0] - n[0] * hi[2]);
        e[2] = -(n[0] * hi[1] - n[1] * hi[0]);
    }
};

// ---------------------------------------------------------------------------
// recover_primitive :287-305 (FP64), fastest_speed :248-258 (FP64).  NR:
// one Newton-refined reciprocal of the density in place of the four
// divisions (within a few ulp; the fused v1 step), else the reference's
// divisions (the v2 array functions, bitwise).
template<bool NR = false>
__device__ __forceinline__ int mhd_recover(const double* u, const double* b, double gamma, double* p)
{
    double d = u[0];
    double S[3] = { u[1], u[2], u[3] };
    double me = 0.5 * Mhd<double>::dot3(b, b);
    double pg;
    if constexpr (NR) {
        const double id = rcp_nr(d);
        double ke = 0.5 * Mhd<double>::dot3(S, S) * id;
        pg = (u[4] - ke - me) * (gamma - 1);
        p[0] = d; p[1] = S[0] * id; p[2] = S[1] * id; p[3] = S[2] * id; p[4] = pg;
    } else {
        double ke = 0.5 * Mhd<double>::dot3(S, S) / d;
        pg = (u[4] - ke - me) * (gamma - 1);
        p[0] = d; p[1] = S[0] / d; p[2] = S[1] / d; p[3] = S[2] / d; p[4] = pg;
    }
    p[5] = b[0]; p[6] = b[1]; p[7] = b[2];
    if (d < 0.0) return 1;
    if (pg < 0.0) return 2;
    return 0;
}

__device__ __forceinline__ double mhd_fastest_speed(const double* p, double gamma)
{
    double m1, q1, m2, q2, m3, q3;
    Mhd<double>::wavespeeds<1>(p, gamma, m1, q1);
    Mhd<double>::wavespeeds<2>(p, gamma, m2, q2);
    Mhd<double>::wavespeeds<3>(p, gamma, m3, q3);
    return smax(smax(smax(fabs(m1), fabs(q1)), smax(fabs(m2), fabs(q2))), smax(fabs(m3), fabs(q3)));
}

struct MhdArgs {
    const double* uc; const double* b1; const double* b2; const double* b3;
    const double* u0; const double* b01; const double* b02; const double* b03;   // RK base or null
    double* uo; double* bo1; double* bo2; double* bo3;
    double rk_b;
    int N;
    double dl, cfl, theta, gamma;
    double* dt;                      // device: CFL time step of this call (written by mhd_dt_kernel)
    unsigned long long* vmax;        // device: kRateSlots partial maxima of the fastest speed (ordered bits)
    int* err;
};

__device__ __forceinline__ int wrapi(int i, int N) { return i < 0 ? i + N : (i >= N ? i - N : i); }

// Grid-stride over cells with a bounded grid, so the speed reduction ends in
// one atomic per workgroup of a few thousand, spread over kRateSlots slots.
template<typename R>
__global__ void __launch_bounds__(256)
mhd_prim_kernel(MhdArgs a, R* P)
{
    const int N = a.N;
    const long N3 = (long)N * N * N;
    const long NN = (long)N * N;
    double s = 0.0;
    // FP64: each XCD sweeps a contiguous range of the grid's cells, so the
    // face fields that neighbouring rows and planes share stay in its L2
    // (step 0.9 % faster at 256^3).  Mixed keeps the hardware order, which
    // writes the top planes last for the tile pass that starts there (the
    // remap made its step 1.1 % slower).
    const long w = is_fp64<R>::value ? xcd_remap(blockIdx.x, gridDim.x) : blockIdx.x;
    for (long c = w * blockDim.x + threadIdx.x; c < N3; c += (long)gridDim.x * blockDim.x) {
        const int i = (int)(c / NN), j = (int)(c / N % N), k = (int)(c % N);
        double b[3] = { 0.5 * (a.b1[(long)i * NN + (long)j * N + k] + a.b1[(long)(i + 1) * NN + (long)j * N + k]),
                        0.5 * (a.b2[((long)i * (N + 1) + j) * N + k] + a.b2[((long)i * (N + 1) + j + 1) * N + k]),
                        0.5 * (a.b3[((long)i * N + j) * (N + 1) + k] + a.b3[((long)i * N + j) * (N + 1) + k + 1]) };
        double u[5], p[8];
#pragma unroll
        for (int q = 0; q < 5; ++q) u[q] = a.uc[q * N3 + c];
        if constexpr (is_fp64<R>::value) {
            int st = mhd_recover<true>(u, b, a.gamma, p);
            if (st) report_error(a.err, st == 1 ? 1 : 2, 0, (int)c);
#pragma unroll
            for (int q = 0; q < 8; ++q) P[q * N3 + c] = p[q];
            s = fmax(s, mhd_fastest_speed(p, a.gamma));
        } else {
            // mixed: the pressure difference in FP64 (one reciprocal of the
            // density instead of four divisions), the primitives stored and
            // the CFL speed taken in FP32
            const double d = u[0], id = 1.0 / d;
            const double S2 = fma(u[3], u[3], fma(u[2], u[2], u[1] * u[1]));
            const double B2 = fma(b[2], b[2], fma(b[1], b[1], b[0] * b[0]));
            const double pg = (u[4] - 0.5 * S2 * id - 0.5 * B2) * (a.gamma - 1.0);
            if (d < 0.0) report_error(a.err, 1, 0, (int)c);
            else if (pg < 0.0) report_error(a.err, 2, 0, (int)c);
            R pf[8] = { R(d), R(u[1] * id), R(u[2] * id), R(u[3] * id), R(pg), R(b[0]), R(b[1]), R(b[2]) };
#pragma unroll
            for (int q = 0; q < 8; ++q) P[q * N3 + c] = pf[q];
            R m1, q1, m2, q2, m3, q3;
            const R gf = R(a.gamma);
            Mhd<R>::template wavespeeds<1>(pf, gf, m1, q1);
            Mhd<R>::template wavespeeds<2>(pf, gf, m2, q2);
            Mhd<R>::template wavespeeds<3>(pf, gf, m3, q3);
            const R v = fmaxf(fmaxf(fmaxf(fabsf(m1), fabsf(q1)), fmaxf(fabsf(m2), fabsf(q2))), fmaxf(fabsf(m3), fabsf(q3)));
            s = fmax(s, double(v));
        }
    }
    workgroup_atomic_max<256>(s, a.vmax);
}

// dt from the kRateSlots partial maxima (one wave)
static __global__ void __launch_bounds__(64) mhd_dt_kernel(MhdArgs a)
{
    double vmax = wave_max(__longlong_as_double((long long)a.vmax[threadIdx.x]));
    if (threadIdx.x == 0) *a.dt = a.cfl * a.dl / vmax;
}

// All three face families in one pass: the thread of cell (i, j, k)
// computes the lower face of the cell on each axis (the x face at i, the y
// face at j, the z face at k), and the cells on the last plane of an axis
// also compute the face at N (periodic: the same cells as face 0).  The
// primitive stencils of the three axes share the cell's neighbourhood, so
// the primitives are read once from cache instead of once per axis.  Each
// face's stencil: cells (m-2, m-1, m, m+1) along the axis for face m, wrapped.
// Half the PLM slope of the middle value: the reference's 0.5 * plm_gradient
// in FP64, the three-operand median form in FP32 (the same value).
template<typename R>
__device__ __forceinline__ R mhd_half_slope(R c0, R c1, R c2, R theta)
{
    if constexpr (is_fp64<R>::value) return R(0.5) * plm(c0, c1, c2, theta);
    else return plm_half(c0, c1, c2, theta);
}

template<typename R, int A>
__device__ __forceinline__ void mhd_one_face(const MhdArgs& a, const R* P, R* F, R* E, int i, int j, int k)
{
    const int N = a.N;
    const long N3 = (long)N * N * N;
    const long nf = (long)(N + 1) * N * N;
    long f;
    if (A == 1)      f = ((long)i * N + j) * N + k;
    else if (A == 2) f = ((long)i * (N + 1) + j) * N + k;
    else             f = ((long)i * N + j) * (N + 1) + k;
    long cell[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const int o = s - 2;
        int ci = i, cj = j, ck = k;
        if (A == 1) ci = wrapi(i + o, N);
        if (A == 2) cj = wrapi(j + o, N);
        if (A == 3) ck = wrapi(k + o, N);
        cell[s] = ((long)ci * N + cj) * N + ck;
    }
    const R theta = R(a.theta);
    R pl[8], pr[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        R c0 = P[q * N3 + cell[0]], c1 = P[q * N3 + cell[1]], c2 = P[q * N3 + cell[2]], c3 = P[q * N3 + cell[3]];
        pl[q] = c1 + mhd_half_slope(c0, c1, c2, theta);
        pr[q] = c2 - mhd_half_slope(c1, c2, c3, theta);
    }
    const double* bf = A == 1 ? a.b1 : (A == 2 ? a.b2 : a.b3);
    R fl[5], e[3];
    Mhd<R>::template riemann<A>(pl, pr, R(bf[f]), R(a.gamma), fl, e);
#pragma unroll
    for (int q = 0; q < 5; ++q) F[q * nf + f] = fl[q];
    E[f] = e[A == 1 ? 1 : 0];
    E[nf + f] = e[A == 3 ? 1 : 2];
}

template<typename R>
__global__ void __launch_bounds__(256)
mhd_faces_kernel(MhdArgs a, const R* P, R* F1, R* E1, R* F2, R* E2, R* F3, R* E3)
{
    const int N = a.N;
    const long N3 = (long)N * N * N;
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N3) return;
    const int i = (int)(c / ((long)N * N)), j = (int)(c / N % N), k = (int)(c % N);
    mhd_one_face<R, 1>(a, P, F1, E1, i, j, k);
    mhd_one_face<R, 2>(a, P, F2, E2, i, j, k);
    mhd_one_face<R, 3>(a, P, F3, E3, i, j, k);
}

// The faces on the upper boundary planes (i = N, j = N, k = N), one per
// thread: 3 N^2 faces.  In the per-cell pass the cells next to them would
// run a fourth, nearly empty face pass (a quarter of all waves, for k = N).
template<typename R>
__global__ void __launch_bounds__(256)
mhd_boundary_faces_kernel(MhdArgs a, const R* P, R* F1, R* E1, R* F2, R* E2, R* F3, R* E3)
{
    const int N = a.N;
    const long NN = (long)N * N;
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= 3 * NN) return;
    const int axis = (int)(c / NN);
    const int u = (int)(c % NN / N), v = (int)(c % N);
    if (axis == 0)      mhd_one_face<R, 1>(a, P, F1, E1, N, u, v);
    else if (axis == 1) mhd_one_face<R, 2>(a, P, F2, E2, u, N, v);
    else                mhd_one_face<R, 3>(a, P, F3, E3, u, v, N);
}

// Column-tile form of the flux pass fused with the conserved update.  A
// workgroup owns TJ x TK columns (j, k) and marches along i over TI planes
// of one segment.  Each thread keeps a four-plane window of its column's
// primitives, so the x face above its cell needs one new load per plane;
// the y and z faces of the plane come from an LDS copy of the plane with a
// two-cell halo.  Per plane and thread: the y and z faces below the cell
// (the tile's upper y and z faces by the last row and by the first TJ
// threads), the x face above it (after the y and z faces, so the next
// plane's primitives it needs arrive meanwhile; FP64 needs 34 fewer
// VGPRs, and the RK2 step at 256^3 is 0.5 % (mixed) and 0.9 % (FP64)
// faster), then U of the cell from the six face fluxes
// (divergence_difference, src/mesh_cartesian_3d.hpp:290-301).  Face fluxes
// never leave the workgroup; the transverse face electric fields go to the
// E arrays for the face update.  Every face is computed from the same
// operands in the same order as mhd_one_face, so a face shared by two
// tiles or segments has one value.
// Mixed: four waves per SIMD, where the register allocator fits the kernel
// in 128 VGPRs without spills (150 when unconstrained, three waves) and the
// extra wave hides load latency (11 % faster per RK2 step at 256^3).  FP64:
// two waves per SIMD (up to 256 VGPRs), the update in the arithmetic and
// order of mhd_update_kernel.
template<typename R> struct MhdTileWaves { static constexpr int value = is_fp64<R>::value ? 2 : 4; };

template<typename R, int TJ, int TK>
__global__ void __launch_bounds__(TJ * TK) __attribute__((amdgpu_waves_per_eu(MhdTileWaves<R>::value, 8)))
mhd_tile_kernel(MhdArgs a, const R* P, R* E1, R* E2, R* E3, int TI)
{
    constexpr int NT = TJ * TK;
    constexpr int EJ = TJ + 4, EK = TK + 4;
    const int N = a.N;
    const long N3 = (long)N * N * N, nf = (long)(N + 1) * N * N;
    const int tiles_k = N / TK, tiles_j = N / TJ;
    int w = blockIdx.x;
    const int k0 = (w % tiles_k) * TK; w /= tiles_k;
    const int j0 = (w % tiles_j) * TJ; w /= tiles_j;
    // segments from the top plane down: the primitive pass wrote the high
    // planes last, so they are the likeliest to still be cached
    const int i0 = (N / TI - 1 - w) * TI;
    const int t = threadIdx.x, jl = t / TK, kl = t % TK;
    const int j = j0 + jl, k = k0 + kl;

    __shared__ R Pl[8][EJ][EK];       // primitives of plane i, (j0 - 2 .. j0 + TJ + 1) x (k0 - 2 .. k0 + TK + 1)
    __shared__ R F2s[5][TJ + 1][TK];  // y-face fluxes of plane i, faces j0 .. j0 + TJ
    __shared__ R F3s[5][TJ][TK + 1];  // z-face fluxes of plane i, faces k0 .. k0 + TK

    const double dt = *a.dt;
    const R theta = R(a.theta), gamma = R(a.gamma);
    auto cell = [&] (int ii, int jj, int kk) {
        return ((long)wrapi(ii, N) * N + wrapi(jj, N)) * N + wrapi(kk, N);
    };
    auto loadP = [&] (long c, R* p) {
#pragma unroll
        for (int q = 0; q < 8; ++q) p[q] = P[q * N3 + c];
    };
    // face states from four cells along the face normal (lower two, upper two)
    auto states = [&] (const R* c0, const R* c1, const R* c2, const R* c3, R* pl, R* pr) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            pl[q] = c1[q] + mhd_half_slope(c0[q], c1[q], c2[q], theta);
            pr[q] = c2[q] - mhd_half_slope(c1[q], c2[q], c3[q], theta);
        }
    };

    // x window: Pm = plane i-1, P0 = i, Pp = i+1, and g0 = half slope at
    // plane i (carried: the slope of the upper cell of one x face is the
    // slope of the lower cell of the next); warm-up face at i0
    R Pm[8], P0[8], Pp[8], g0[8], Fx_lo[5];
    {
        R Pm2[8], pl[8], pr[8], e[3];
        loadP(cell(i0 - 2, j, k), Pm2);
        loadP(cell(i0 - 1, j, k), Pm);
        loadP(cell(i0, j, k), P0);
        loadP(cell(i0 + 1, j, k), Pp);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            g0[q] = mhd_half_slope(Pm[q], P0[q], Pp[q], theta);
            pl[q] = Pm[q] + mhd_half_slope(Pm2[q], Pm[q], P0[q], theta);
            pr[q] = P0[q] - g0[q];
        }
        const long f = ((long)i0 * N + j) * N + k;
        Mhd<R>::template riemann<1>(pl, pr, R(a.b1[f]), gamma, Fx_lo, e);
        if (i0 == 0) { E1[f] = e[1]; E1[nf + f] = e[2]; }
    }
    const double dtdl = dt / a.dl;

    // The plane's two-cell halo ring (NH cells): position of ring cell r.
    constexpr int NH = 4 * EK + 4 * TJ;
    auto ring = [&] (int r, int& rj, int& rk) {
        if (r < 2 * EK)      { rj = r / EK; rk = r % EK; }
        else if (r < 4 * EK) { rj = TJ + 2 + (r - 2 * EK) / EK; rk = (r - 2 * EK) % EK; }
        else {
            const int s2 = r - 4 * EK;
            rj = 2 + s2 / 4;
            const int side = s2 % 4;
            rk = side < 2 ? side : TK + side;
        }
    };
    // Where every ring cell has a thread of its own (the 8 x 32 tiles): the
    // ring of the next plane is fetched while this plane computes (its load
    // latency off the plane's critical path; 127 VGPRs mixed, 254 FP64).
    constexpr bool kRingAhead = NH <= NT;
    int hj = 0, hk = 0;
    R Hn[8];
    if constexpr (kRingAhead) {
        if (t < NH) {
            ring(t, hj, hk);
            loadP(cell(i0, j0 - 2 + hj, k0 - 2 + hk), Hn);
        }
    }

    for (int il = 0; il < TI; ++il) {
        const int i = i0 + il;
        R Pn[8];
        loadP(cell(i + 2, j, k), Pn);

        // LDS plane i: interior from the window, halo ring from memory
#pragma unroll
        for (int q = 0; q < 8; ++q) Pl[q][jl + 2][kl + 2] = P0[q];
        if constexpr (kRingAhead) {
            if (t < NH) {
#pragma unroll
                for (int q = 0; q < 8; ++q) Pl[q][hj][hk] = Hn[q];
                if (il + 1 < TI) loadP(cell(i + 1, j0 - 2 + hj, k0 - 2 + hk), Hn);
            }
        } else {
            for (int r = t; r < NH; r += NT) {
                int rj, rk;
                ring(r, rj, rk);
                const long c = cell(i, j0 - 2 + rj, k0 - 2 + rk);
#pragma unroll
                for (int q = 0; q < 8; ++q) Pl[q][rj][rk] = P[q * N3 + c];
            }
        }
        // FP64: the longitudinal fields of the cell's own y and z faces are
        // loaded before the barrier, their latency off the faces' path (the
        // step 1.3 % faster at 256^3); the mixed pass loads them in the face
        // (0.2 % faster there)
        constexpr bool kPreB = is_fp64<R>::value;
        // the x face's longitudinal field (its face follows the barrier), in
        // both precisions (FP64 step 1.0 % faster, mixed unchanged)
        const R blx = R(a.b1[((long)(i + 1) * N + j) * N + k]);
        R bly = R(0), blz = R(0);
        if constexpr (kPreB) {
            bly = R(a.b2[((long)i * (N + 1) + j) * N + k]);
            blz = R(a.b3[((long)i * N + j) * (N + 1) + k]);
        }
        __syncthreads();

        // y face below the cell (face j), and the tile's upper y face by the last row
        auto yface = [&] (int jf /* local face row 0..TJ */, bool own) {
            R c0[8], c1[8], c2[8], c3[8], pl[8], pr[8], f[5], e[3];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                c0[q] = Pl[q][jf][kl + 2]; c1[q] = Pl[q][jf + 1][kl + 2];
                c2[q] = Pl[q][jf + 2][kl + 2]; c3[q] = Pl[q][jf + 3][kl + 2];
            }
            states(c0, c1, c2, c3, pl, pr);
            const long fi = ((long)i * (N + 1) + (j0 + jf)) * N + k;
            Mhd<R>::template riemann<2>(pl, pr, kPreB && own ? bly : R(a.b2[fi]), gamma, f, e);
#pragma unroll
            for (int q = 0; q < 5; ++q) F2s[q][jf][kl] = f[q];
            if (jf < TJ || j0 + jf == N) { E2[fi] = e[0]; E2[nf + fi] = e[2]; }
        };
        yface(jl, true);
        if (jl == TJ - 1) yface(TJ, false);

        // z face below the cell (face k), and the tile's upper z faces by threads 0 .. TJ-1
        auto zface = [&] (int row, int kf /* local face column 0..TK */, bool own) {
            R c0[8], c1[8], c2[8], c3[8], pl[8], pr[8], f[5], e[3];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                c0[q] = Pl[q][row + 2][kf]; c1[q] = Pl[q][row + 2][kf + 1];
                c2[q] = Pl[q][row + 2][kf + 2]; c3[q] = Pl[q][row + 2][kf + 3];
            }
            states(c0, c1, c2, c3, pl, pr);
            const long fi = ((long)i * N + (j0 + row)) * (N + 1) + (k0 + kf);
            Mhd<R>::template riemann<3>(pl, pr, kPreB && own ? blz : R(a.b3[fi]), gamma, f, e);
#pragma unroll
            for (int q = 0; q < 5; ++q) F3s[q][row][kf] = f[q];
            if (kf < TK || k0 + kf == N) { E3[fi] = e[0]; E3[nf + fi] = e[1]; }
        };
        zface(jl, kl, true);
        if (t < TJ) zface(t, TK, false);
        // x face above the cell (face i + 1)
        R Fx_hi[5];
        {
            R pl[8], pr[8], e[3];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const R gp = mhd_half_slope(P0[q], Pp[q], Pn[q], theta);
                pl[q] = P0[q] + g0[q];
                pr[q] = Pp[q] - gp;
                g0[q] = gp;
            }
            const long f = ((long)(i + 1) * N + j) * N + k;
            Mhd<R>::template riemann<1>(pl, pr, blx, gamma, Fx_hi, e);
            E1[f] = e[1];
            E1[nf + f] = e[2];
        }
        // the update's U and RK base, loaded before the barrier (which
        // orders memory, so the compiler keeps loads written after it there)
        // so that their latency overlaps the wait: the 256^3 RK2 step 5.7 %
        // faster mixed, 11.6 % FP64
        double ucv[5], u0v[5] = {};
        {
            const long c = ((long)i * N + j) * N + k;
#pragma unroll
            for (int q = 0; q < 5; ++q) ucv[q] = a.uc[q * N3 + c];
            if (a.u0) {
#pragma unroll
                for (int q = 0; q < 5; ++q) u0v[q] = a.u0[q * N3 + c];
            }
        }
        __syncthreads();

        // conserved update of cell (i, j, k)
        {
            const long c = ((long)i * N + j) * N + k;
#pragma unroll
            for (int q = 0; q < 5; ++q) {
                double d1 = double(Fx_hi[q]) - double(Fx_lo[q]);
                double d2 = double(F2s[q][jl + 1][kl]) - double(F2s[q][jl][kl]);
                double d3 = double(F3s[q][jl][kl + 1]) - double(F3s[q][jl][kl]);
                double r;
                if constexpr (is_fp64<R>::value) {
                    r = ucv[q] - (d1 + d2 + d3) * dtdl;
                    if (a.u0) r = u0v[q] * a.rk_b + r * (1.0 - a.rk_b);
                } else {
                    r = fma(-(d1 + d2 + d3), dtdl, ucv[q]);
                    if (a.u0) r = fma(u0v[q], a.rk_b, r * (1.0 - a.rk_b));
                }
                a.uo[q * N3 + c] = r;
            }
        }
        // No third barrier: the next plane writes Pl before its first
        // barrier, when every wave is past this plane's second one (its Pl
        // reads done), and F2s/F3s after its first barrier, which every wave
        // reaches only after this update's reads (bitwise equal; the mixed
        // step 0.9 % and the FP64 step 0.7 % faster at 256^3).
#pragma unroll
        for (int q = 0; q < 8; ++q) { Pm[q] = P0[q]; P0[q] = Pp[q]; Pp[q] = Pn[q]; }
#pragma unroll
        for (int q = 0; q < 5; ++q) Fx_lo[q] = Fx_hi[q];
    }
}

// Face-centred transverse electric field of axis-A faces, by component.
template<typename R>
struct FaceE {
    const R* E1; const R* E2; const R* E3;   // each [2][nf]
    int N;
    long nf;
    __device__ __forceinline__ double x_face(int c, int i, int j, int k) const   // c in {1:y, 2:z}
    {
        return double(E1[(c - 1) * nf + ((long)i * N + wrapi(j, N)) * N + wrapi(k, N)]);
    }
    __device__ __forceinline__ double y_face(int c, int i, int j, int k) const   // c in {0:x, 2:z}
    {
        return double(E2[(c == 0 ? 0 : 1) * nf + ((long)wrapi(i, N) * (N + 1) + j) * N + wrapi(k, N)]);
    }
    __device__ __forceinline__ double z_face(int c, int i, int j, int k) const   // c in {0:x, 1:y}
    {
        return double(E3[c * nf + ((long)wrapi(i, N) * N + wrapi(j, N)) * (N + 1) + k]);
    }
    // edge_emf_from_face :76-98
    // x-edge at (x cell i, y face j, z face k)
    __device__ __forceinline__ double ee1(int i, int j, int k) const
    {
        double a = 0.5 * (y_face(0, i, j, k - 1) + y_face(0, i, j, k));
        double b = 0.5 * (z_face(0, i, j - 1, k) + z_face(0, i, j, k));
        return 0.5 * (a + b);
    }
    // y-edge at (x face i, y cell j, z face k)
    __device__ __forceinline__ double ee2(int i, int j, int k) const
    {
        double a = 0.5 * (z_face(1, i - 1, j, k) + z_face(1, i, j, k));
        double b = 0.5 * (x_face(1, i, j, k - 1) + x_face(1, i, j, k));
        return 0.5 * (a + b);
    }
    // z-edge at (x face i, y face j, z cell k)
    __device__ __forceinline__ double ee3(int i, int j, int k) const
    {
        double a = 0.5 * (x_face(2, i, j - 1, k) + x_face(2, i, j, k));
        double b = 0.5 * (y_face(2, i - 1, j, k) + y_face(2, i, j, k));
        return 0.5 * (a + b);
    }
};

// One pass for the whole update: the thread of cell (i, j, k) updates the
// cell's conserved state and the lower face of the cell on each axis (plus
// the faces at N on the last planes), with the arithmetic of
// mhd_cell_update_kernel and mhd_face_update_kernel<R, A>.
template<typename R, int A>
__device__ __forceinline__ void mhd_update_face(const MhdArgs& a, const FaceE<R>& fe, double dt, int i, int j, int k)
{
    const int N = a.N;
    double cf, b, b0 = 0.0;
    long f;
    if (A == 1) {
        f = ((long)i * N + j) * N + k;
        cf = (fe.ee3(i, j + 1, k) - fe.ee3(i, j, k)) - (fe.ee2(i, j, k + 1) - fe.ee2(i, j, k));
        b = a.b1[f];
        if (a.u0) b0 = a.b01[f];
    } else if (A == 2) {
        f = ((long)i * (N + 1) + j) * N + k;
        cf = (fe.ee1(i, j, k + 1) - fe.ee1(i, j, k)) - (fe.ee3(i + 1, j, k) - fe.ee3(i, j, k));
        b = a.b2[f];
        if (a.u0) b0 = a.b02[f];
    } else {
        f = ((long)i * N + j) * (N + 1) + k;
        cf = (fe.ee2(i + 1, j, k) - fe.ee2(i, j, k)) - (fe.ee1(i, j + 1, k) - fe.ee1(i, j, k));
        b = a.b3[f];
        if (a.u0) b0 = a.b03[f];
    }
    double r = b - cf * (dt / a.dl);
    if (a.u0) r = b0 * a.rk_b + r * (1.0 - a.rk_b);
    double* out = A == 1 ? a.bo1 : (A == 2 ? a.bo2 : a.bo3);
    out[f] = r;
}

template<typename R>
__global__ void __launch_bounds__(256)
mhd_update_kernel(MhdArgs a, const R* F1, const R* F2, const R* F3, FaceE<R> fe)
{
    const int N = a.N;
    const long N3 = (long)N * N * N, nf = (long)(N + 1) * N * N;
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N3) return;
    const int i = (int)(c / ((long)N * N)), j = (int)(c / N % N), k = (int)(c % N);
    const double dt = *a.dt;
    const long x0 = ((long)i * N + j) * N + k, x1 = x0 + (long)N * N;
    const long y0 = ((long)i * (N + 1) + j) * N + k, y1 = y0 + N;
    const long z0 = ((long)i * N + j) * (N + 1) + k, z1 = z0 + 1;
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        double d1 = double(F1[q * nf + x1]) - double(F1[q * nf + x0]);
        double d2 = double(F2[q * nf + y1]) - double(F2[q * nf + y0]);
        double d3 = double(F3[q * nf + z1]) - double(F3[q * nf + z0]);
        double r = a.uc[q * N3 + c] - (d1 + d2 + d3) * (dt / a.dl);
        if (a.u0) r = a.u0[q * N3 + c] * a.rk_b + r * (1.0 - a.rk_b);
        a.uo[q * N3 + c] = r;
    }
    mhd_update_face<R, 1>(a, fe, dt, i, j, k);
    mhd_update_face<R, 2>(a, fe, dt, i, j, k);
    mhd_update_face<R, 3>(a, fe, dt, i, j, k);
    if (i == N - 1) mhd_update_face<R, 1>(a, fe, dt, N, j, k);
    if (j == N - 1) mhd_update_face<R, 2>(a, fe, dt, i, N, k);
    if (k == N - 1) mhd_update_face<R, 3>(a, fe, dt, i, j, N);
}

// The face part of mhd_update_kernel alone (the conserved update is done by
// mhd_tile_kernel).
template<typename R>
__global__ void __launch_bounds__(256)
mhd_face_b_kernel(MhdArgs a, FaceE<R> fe)
{
    const int N = a.N;
    const long N3 = (long)N * N * N;
    // cells from the bottom up: the tile pass wrote the low planes last
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= N3) return;
    const int i = (int)(c / ((long)N * N)), j = (int)(c / N % N), k = (int)(c % N);
    const double dt = *a.dt;
    mhd_update_face<R, 1>(a, fe, dt, i, j, k);
    mhd_update_face<R, 2>(a, fe, dt, i, j, k);
    mhd_update_face<R, 3>(a, fe, dt, i, j, k);
    if (i == N - 1) mhd_update_face<R, 1>(a, fe, dt, N, j, k);
    if (j == N - 1) mhd_update_face<R, 2>(a, fe, dt, i, N, k);
    if (k == N - 1) mhd_update_face<R, 3>(a, fe, dt, i, j, N);
}

// The face-field update by bricks, for cubes the column-tile pass covers: a
// workgroup owns a BI x BJ x BK brick of cells and takes each edge electric
// field it needs once (edge_emf_from_face, :76-98, into LDS), where
// mhd_face_b_kernel takes it once for each of the four faces that share the
// edge: the x edges of its BI cell planes ((BJ + 1) x (BK + 1) each), the y
// and z edges of its BI + 1 x-face planes (BJ x (BK + 1), (BJ + 1) x BK).
// Then each thread updates the three faces below its cell, and the faces at
// N on the last planes, with the operands and arithmetic of
// mhd_update_face.  A thread's face fields and RK base are loaded before the
// edges, so their latency overlaps the edge loads.
template<typename R, int BI, int BJ, int BK>
__global__ void __launch_bounds__(BI * BJ * BK)
mhd_face_b_brick_kernel(MhdArgs a, FaceE<R> fe)
{
    constexpr int NT = BI * BJ * BK;
    constexpr int N1 = BI * (BJ + 1) * (BK + 1);        // x edges
    constexpr int N2 = (BI + 1) * BJ * (BK + 1);        // y edges
    constexpr int N3 = (BI + 1) * (BJ + 1) * BK;        // z edges
    __shared__ double X1[BI][BJ + 1][BK + 1];           // ee1(i0 + p, j0 + r, k0 + c)
    __shared__ double Y2[BI + 1][BJ][BK + 1];           // ee2(i0 + p, j0 + r, k0 + c)
    __shared__ double Z3[BI + 1][BJ + 1][BK];           // ee3(i0 + p, j0 + r, k0 + c)
    const int N = a.N;
    int w = blockIdx.x;
    const int k0 = (w % (N / BK)) * BK; w /= N / BK;
    const int j0 = (w % (N / BJ)) * BJ; w /= N / BJ;
    const int i0 = w * BI;
    const int t = threadIdx.x;
    const int il = t / (BJ * BK), jl = t / BK % BJ, kl = t % BK;
    const int i = i0 + il, j = j0 + jl, k = k0 + kl;

    const long fx = ((long)i * N + j) * N + k;
    const long fy = ((long)i * (N + 1) + j) * N + k;
    const long fz = ((long)i * N + j) * (N + 1) + k;
    const double bx = a.b1[fx], by = a.b2[fy], bz = a.b3[fz];
    double b0x = 0.0, b0y = 0.0, b0z = 0.0;
    if (a.u0) { b0x = a.b01[fx]; b0y = a.b02[fy]; b0z = a.b03[fz]; }

    for (int e = t; e < N1 + N2 + N3; e += NT) {
        if (e < N1) {
            const int c = e % (BK + 1), r = e / (BK + 1) % (BJ + 1), p = e / ((BK + 1) * (BJ + 1));
            X1[p][r][c] = fe.ee1(i0 + p, j0 + r, k0 + c);
        } else if (e < N1 + N2) {
            const int f = e - N1;
            const int c = f % (BK + 1), r = f / (BK + 1) % BJ, p = f / ((BK + 1) * BJ);
            Y2[p][r][c] = fe.ee2(i0 + p, j0 + r, k0 + c);
        } else {
            const int f = e - N1 - N2;
            const int c = f % BK, r = f / BK % (BJ + 1), p = f / (BK * (BJ + 1));
            Z3[p][r][c] = fe.ee3(i0 + p, j0 + r, k0 + c);
        }
    }
    __syncthreads();

    const double dtdl = *a.dt / a.dl;
    // mhd_update_face's output: B - curl(E) dt / dl, and the RK combine
    auto put = [&] (double b, double b0, double* out, long f, double cf) {
        double r = b - cf * dtdl;
        if (a.u0) r = b0 * a.rk_b + r * (1.0 - a.rk_b);
        out[f] = r;
    };
    put(bx, b0x, a.bo1, fx, (Z3[il][jl + 1][kl] - Z3[il][jl][kl]) - (Y2[il][jl][kl + 1] - Y2[il][jl][kl]));
    put(by, b0y, a.bo2, fy, (X1[il][jl][kl + 1] - X1[il][jl][kl]) - (Z3[il + 1][jl][kl] - Z3[il][jl][kl]));
    put(bz, b0z, a.bo3, fz, (Y2[il + 1][jl][kl] - Y2[il][jl][kl]) - (X1[il][jl + 1][kl] - X1[il][jl][kl]));
    if (j == N - 1) {   // y face (i, N, k)
        const long f = ((long)i * (N + 1) + N) * N + k;
        put(a.b2[f], a.u0 ? a.b02[f] : 0.0, a.bo2, f,
            (X1[il][BJ][kl + 1] - X1[il][BJ][kl]) - (Z3[il + 1][BJ][kl] - Z3[il][BJ][kl]));
    }
    if (k == N - 1) {   // z face (i, j, N)
        const long f = ((long)i * N + j) * (N + 1) + N;
        put(a.b3[f], a.u0 ? a.b03[f] : 0.0, a.bo3, f,
            (Y2[il + 1][jl][BK] - Y2[il][jl][BK]) - (X1[il][jl + 1][BK] - X1[il][jl][BK]));
    }
    if (i == N - 1) {   // x face (N, j, k)
        const long f = ((long)N * N + j) * N + k;
        put(a.b1[f], a.u0 ? a.b01[f] : 0.0, a.bo1, f,
            (Z3[BI][jl + 1][kl] - Z3[BI][jl][kl]) - (Y2[BI][jl][kl + 1] - Y2[BI][jl][kl]));
    }
}

// max over cells of |div B| = |d1(B1) + d2(B2) + d3(B3)| (divergence_difference)
static __global__ void __launch_bounds__(256)
mhd_divb_kernel(const double* b1, const double* b2, const double* b3, int N, unsigned long long* slot)
{
    __shared__ double part[4];
    const long N3 = (long)N * N * N;
    const long NN = (long)N * N;
    double v = 0.0;
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < N3; c += (long)gridDim.x * blockDim.x) {
        const int i = (int)(c / NN), j = (int)(c / N % N), k = (int)(c % N);
        double d1 = b1[(long)(i + 1) * NN + (long)j * N + k] - b1[(long)i * NN + (long)j * N + k];
        double d2 = b2[((long)i * (N + 1) + j + 1) * N + k] - b2[((long)i * (N + 1) + j) * N + k];
        double d3 = b3[((long)i * N + j) * (N + 1) + k + 1] - b3[((long)i * N + j) * (N + 1) + k];
        v = fmax(v, fabs(d1 + d2 + d3));
    }
    v = wave_max(v);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = part[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fmax(m, part[w]);
        atomicMax(slot, (unsigned long long)__double_as_longlong(m));
    }
}

} // namespace m3p
