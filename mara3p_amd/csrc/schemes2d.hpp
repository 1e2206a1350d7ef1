// Block-update schemes for the 2D stage kernel: what happens at a face
// (Riemann flux) and at a zone (conserved update + source terms).
//
//   EulerScheme     <- modules::euler2d::updated_conserved   (src/module_euler.cpp:263-298)
//   MinidiskScheme  <- minidisk::updated_conserved           (problems/minidisk_scheme.cpp:294-355)
//   IsoTableScheme  <- locally_isothermal::updated_conserved (src/module_locally_isothermal.cpp:182-212),
//                      face sound speeds supplied as arrays (the reference's
//                      temperature mapping is a host std::function).
#pragma once
#include "physics.hpp"

namespace m3p {

// Per-tile context handed to the scheme hooks.
struct TileCtx {
    int b;      // storage block index
    int bi, bj; // block coordinates on its level
    int k0, c0; // first row / column of the tile inside the block
    int n;      // block size
};

// ---------------------------------------------------------------------------
template<typename R>
struct EulerScheme {
    static constexpr int NC = 5;
    static constexpr int kWavesPerEU = is_fp64<R>::value ? 2 : 3;   // occupancy floor (VGPR budget)
    static constexpr bool kLateBase = false;  // RK base loaded at the top of a row (pipelined loader)
    static constexpr int kMarchUnroll = 2;   // rows per march iteration with a pipelined loader
    // rows a first stage fetches ahead (pipelined loader): two made stage 1
    // 2 % faster for KH FP64 and the quad-tree mixed stage, 1 % slower for
    // the quad-tree FP64 stage
    static constexpr int kPrefetchRows = 2;
    static constexpr bool kCfl = true;       // stage kernels can fuse the CFL reduction
    // the fused RK2 step (both stages' windows of five components) at two
    // waves per SIMD: 178 VGPRs, and its 61 KB of LDS fit two workgroups per CU
    static constexpr int kFusedWavesPerEU = 2;
    using Real = R;
    using Params = EulerParams;
    __device__ static __forceinline__ double cfl_speed(const double* u, const Params& prm)
    {
        return Euler<double>::max_speed(u, prm);
    }
    template<int H, int TJ> struct TileData {};

    __device__ static __forceinline__ double theta(const Params& prm) { return prm.theta; }
    __device__ static __forceinline__ bool recover(const double* u, const Params& prm, R* p)
    {
        return Euler<R>::recover(u, prm, p);
    }
    template<int H, int TJ>
    __device__ static __forceinline__ void setup(TileData<H, TJ>&, const Params&, const TileCtx&) {}
    // no geometry: the fused RK2 step's tile rows and per-zone pre-pass
    template<int H, int TJ>
    __device__ static __forceinline__ void setup_rows(TileData<H, TJ>&, const Params&, const TileCtx&, int, int) {}
    __device__ static __forceinline__ void setup_zone(TileData<1, 1>&, const Params&, const TileCtx&) {}

    template<int AXIS, int H, int TJ>
    __device__ static __forceinline__ void face(const R* pl, const R* pr, const Params& prm,
                                                const TileData<H, TJ>&, const TileCtx&, int, int, R* f)
    {
        Euler<R>::template riemann<AXIS>(pl, pr, prm, f);
    }

    // U - ((dFx + dFy) * dt) / dl   (src/module_euler.cpp:297)
    template<int H, int TJ>
    __device__ static __forceinline__ void update(const Params&, const TileData<H, TJ>&, const TileCtx&,
                                                  int, int, const double* u, const R*,
                                                  const double* dfx, const double* dfy,
                                                  double dt, double dl, double* out)
    {
        if constexpr (is_fp64<R>::value) {
#pragma unroll
            for (int q = 0; q < NC; ++q) out[q] = u[q] - (dfx[q] + dfy[q]) * dt / dl;
        } else {
            double dtdl = dt / dl;
#pragma unroll
            for (int q = 0; q < NC; ++q) out[q] = fma(-(dfx[q] + dfy[q]), dtdl, u[q]);
        }
    }
};

// ---------------------------------------------------------------------------
struct MinidiskParams {
    double R;            // domain radius
    double rs;           // softening length
    double s2;           // pow(rs, 2) evaluated on the host (std::pow)
    double mach;
    double inv_mach2;    // 1 / mach^2
    double omega;        // frame rotation rate
    double sink_rate;
    double s6;           // pow(sink_radius, 6) evaluated on the host
    double inv_s6;       // 1 / s6
    double q_sink_cut;   // (104 s6)^(1/3): beyond it exp(-q^3 / s6) is below 1e-45
    double buffer_rate;
    double buffer_scale;
    double inv_buffer_scale;
    double m1, x1, y1;   // binary components at this stage's time
    double m2, x2, y2;
    // or, when set, the same six values {m1, x1, y1, m2, x2, y2} in device
    // memory, written by the clock kernel of a device-clocked step
    const double* orbit;
    int level;
};

// Stage kernels call this on their parameters first: a device-clocked
// minidisk stage takes the binary's state from device memory.
__device__ __forceinline__ void resolve_params(MinidiskParams& p)
{
    if (p.orbit) {
        p.m1 = p.orbit[0]; p.x1 = p.orbit[1]; p.y1 = p.orbit[2];
        p.m2 = p.orbit[3]; p.x2 = p.orbit[4]; p.y2 = p.orbit[5];
    }
}

template<typename R>
struct MinidiskScheme {
    static constexpr int NC = 3;
    // FP64 at three waves per SIMD (168 VGPRs): the first stage spills 3
    // and runs 13 % faster than at two; the second, with its RK base loaded
    // at the update (kLateBase), spills none and runs 5 % faster
    static constexpr int kWavesPerEU = is_fp64<R>::value ? 3 : 4;
    static constexpr bool kLateBase = is_fp64<R>::value;
    // three rows per iteration rotate the three-row window in registers; the
    // FP64 kernel runs faster at two (register pressure)
    static constexpr int kMarchUnroll = is_fp64<R>::value ? 2 : 3;
    // a first stage (no base state) keeps two rows of loads in flight: its
    // stage 1 runs 9 % (mixed) and 2 % (FP64) faster than with one
    static constexpr int kPrefetchRows = 2;
    // the fused RK2 step (both stages' windows) at three waves per SIMD
    static constexpr int kFusedWavesPerEU = 3;
    static constexpr bool kCfl = false;
    __device__ static __forceinline__ double cfl_speed(const double*, const MinidiskParams&) { return 0.0; }
    using Real = R;
    using Params = MinidiskParams;
    // Vertex coordinates of the tile's rows/columns in FP64 (reference formula),
    // plus FP32 vertex and centre copies for the mixed kernel.
    template<int H, int TJ> struct TileData {
        double xv[H + 1]; double yv[TJ + 1];
        float xvf[H + 1]; float xcf[H]; float yvf[TJ + 1]; float ycf[TJ];
    };

    // theta = 2 is hard-coded by the reference (problems/minidisk_scheme.cpp:313-314)
    __device__ static __forceinline__ double theta(const Params&) { return 2.0; }
    __device__ static __forceinline__ bool recover(const double* u, const Params&, R* p)
    {
        if constexpr (is_fp64<R>::value) {
            // one reciprocal of Sigma (within a few ulp of the two divisions
            // of recover_primitive, src/physics_iso2d.hpp:119-125)
            const double is = rcp_nr(u[0]);
            p[0] = u[0];
            p[1] = u[1] * is;
            p[2] = u[2] * is;
        } else {
            Iso2d<R>::recover(u, p);
        }
        return true;
    }

    // vertex_positions (problems/minidisk_scheme.cpp:101-116) + nd::linspace
    __device__ static __forceinline__ double vertex(double Rd, int n, int level, int b, int i)
    {
        double dlev = double(1 << level);
        double x0 = Rd * (-1 + double(2 * (unsigned long)b) / dlev);
        double x1 = Rd * (-1 + double(2 * (unsigned long)(b + 1)) / dlev);
        return x0 + (x1 - x0) * double(i) / double(n);
    }

    template<int H, int TJ>
    __device__ static __forceinline__ void setup(TileData<H, TJ>& td, const Params& prm, const TileCtx& tc)
    {
        for (int t = threadIdx.x; t <= H; t += blockDim.x) {
            double v = vertex(prm.R, tc.n, prm.level, tc.bi, tc.k0 + t);
            td.xv[t] = v;
            td.xvf[t] = float(v);
            if (t < H) td.xcf[t] = float(0.5 * (v + vertex(prm.R, tc.n, prm.level, tc.bi, tc.k0 + t + 1)));
        }
        for (int t = threadIdx.x; t <= TJ; t += blockDim.x) {
            double v = vertex(prm.R, tc.n, prm.level, tc.bj, tc.c0 + t);
            td.yv[t] = v;
            td.yvf[t] = float(v);
            if (t < TJ) td.ycf[t] = float(0.5 * (v + vertex(prm.R, tc.n, prm.level, tc.bj, tc.c0 + t + 1)));
        }
    }

    // TileData of the fused RK2 step, whose tile rows tc.k0 .. tc.k0 + H - 1
    // reach up to two rows into the blocks before (bi_m) and after (bi_p) on
    // axis 0: each row takes the centre, and each vertex the coordinate, of
    // the block that owns it, as that block's own stage computes them
    // (vertices 0 and n from this block's formula).
    template<int H, int TJ>
    __device__ static __forceinline__ void setup_rows(TileData<H, TJ>& td, const Params& prm, const TileCtx& tc,
                                                      int bi_m, int bi_p)
    {
        const int n = tc.n;
        for (int t = threadIdx.x; t <= H; t += blockDim.x) {
            const int v = tc.k0 + t;
            const int ob = v < 0 ? bi_m : (v > n ? bi_p : tc.bi);
            const int vl = v < 0 ? v + n : (v > n ? v - n : v);
            const double x = vertex(prm.R, n, prm.level, ob, vl);
            td.xv[t] = x;
            td.xvf[t] = float(x);
            if (t < H) {
                const int rb = v < 0 ? bi_m : (v >= n ? bi_p : tc.bi);
                const int rl = v < 0 ? v + n : (v >= n ? v - n : v);
                td.xcf[t] = float(0.5 * (vertex(prm.R, n, prm.level, rb, rl) + vertex(prm.R, n, prm.level, rb, rl + 1)));
            }
        }
        for (int t = threadIdx.x; t <= TJ; t += blockDim.x) {
            double v = vertex(prm.R, tc.n, prm.level, tc.bj, tc.c0 + t);
            td.yv[t] = v;
            td.yvf[t] = float(v);
            if (t < TJ) td.ycf[t] = float(0.5 * (v + vertex(prm.R, tc.n, prm.level, tc.bj, tc.c0 + t + 1)));
        }
    }

    // One zone's TileData (k0 = row, c0 = column): the per-zone kernels.
    __device__ static __forceinline__ void setup_zone(TileData<1, 1>& td, const Params& prm, const TileCtx& tc)
    {
#pragma unroll
        for (int t = 0; t <= 1; ++t) {
            double v = vertex(prm.R, tc.n, prm.level, tc.bi, tc.k0 + t);
            td.xv[t] = v;
            td.xvf[t] = float(v);
            double w = vertex(prm.R, tc.n, prm.level, tc.bj, tc.c0 + t);
            td.yv[t] = w;
            td.yvf[t] = float(w);
        }
        td.xcf[0] = float(0.5 * (td.xv[0] + td.xv[1]));
        td.ycf[0] = float(0.5 * (td.yv[0] + td.yv[1]));
    }

    // sound_speed_squared :127-132 at face_coordinates :134-151
    template<int AXIS, int H, int TJ>
    __device__ static __forceinline__ void face(const R* pl, const R* pr, const Params& prm,
                                                const TileData<H, TJ>& td, const TileCtx& tc, int fi, int fj, R* f)
    {
        R cs2;
        if constexpr (is_fp64<R>::value) {
            double x, y;
            if constexpr (AXIS == 1) {
                x = td.xv[fi - tc.k0];
                y = 0.5 * (td.yv[fj - tc.c0] + td.yv[fj - tc.c0 + 1]);
            } else {
                x = 0.5 * (td.xv[fi - tc.k0] + td.xv[fi - tc.k0 + 1]);
                y = td.yv[fj - tc.c0];
            }
            // -phi / M^2 with phi = -1 / sqrt(r^2 + r_s^2); the division by
            // M twice is folded into one product (within a few ulp)
            double r2 = x * x + y * y;
            cs2 = rsqrt_nr(r2 + prm.s2) * prm.inv_mach2;
            Iso2d<R>::template hlle<AXIS, true>(pl, pr, cs2, f);
            return;
        } else {
            float xf = AXIS == 1 ? td.xvf[fi - tc.k0] : td.xcf[fi - tc.k0];
            float yf = AXIS == 1 ? td.ycf[fj - tc.c0] : td.yvf[fj - tc.c0];
            float q = fmaf(xf, xf, fmaf(yf, yf, float(prm.s2)));
            cs2 = __builtin_amdgcn_rsqf(q) * float(prm.inv_mach2);
        }
        Iso2d<R>::template hlle<AXIS>(pl, pr, cs2, f);
    }

    // The source-term sum and update of problems/minidisk_scheme.cpp:329-352.
    template<int H, int TJ>
    __device__ static __forceinline__ void update(const Params& prm, const TileData<H, TJ>& td, const TileCtx& tc,
                                                  int k, int c, const double* u, const R* pc,
                                                  const double* dfx, const double* dfy,
                                                  double dt, double dl, double* out)
    {
        if constexpr (is_fp64<R>::value) {
            double x = 0.5 * (td.xv[k - tc.k0] + td.xv[k - tc.k0 + 1]);
            double y = 0.5 * (td.yv[c - tc.c0] + td.yv[c - tc.c0 + 1]);
            // two_body::gravitational_acceleration (src/physics_two_body.hpp:208-218);
            // m / pow(q, 1.5) = m rsqrt(q)^3 and pow(t, 3) = t t t, within a few ulp
            double dx1 = x - prm.x1, dy1 = y - prm.y1;
            double dx2 = x - prm.x2, dy2 = y - prm.y2;
            double r21 = dx1 * dx1 + dy1 * dy1;
            double r22 = dx2 * dx2 + dy2 * dy2;
            double q1 = r21 + prm.s2, q2 = r22 + prm.s2;
            // m / q^1.5 = m rsqrt(q)^3
            const double iq1 = rsqrt_nr(q1), iq2 = rsqrt_nr(q2);
            double g1 = -1.0 * prm.m1 * (iq1 * iq1 * iq1);
            double g2 = -1.0 * prm.m2 * (iq2 * iq2 * iq2);
            double ax1 = g1 * dx1, ay1 = g1 * dy1;
            double ax2 = g2 * dx2, ay2 = g2 * dy2;
            // coriolis_term :67-71, centrifugal_term :73-76 (from pc)
            double om = prm.omega;
            double corx = 2.0 * om * pc[2], cory = 2.0 * om * (-pc[1]);
            double cenx = om * om * x, ceny = om * om * y;
            // sink_rate_field :57-65; exp(-t) is exactly 0 for t > 746
            double t1 = (0.0 + dx1 * dx1) + dy1 * dy1;
            double t2 = (0.0 + dx2 * dx2) + dy2 * dy2;
            double r61 = t1 * t1 * t1, r62 = t2 * t2 * t2;
            double sr1 = 0.0, sr2 = 0.0;
            if (r61 < 746.0 * prm.s6) sr1 = prm.sink_rate * exp(-r61 * prm.inv_s6);
            if (r62 < 746.0 * prm.s6) sr2 = prm.sink_rate * exp(-r62 * prm.inv_s6);
            // buffer_rate_field :47-55; 1 + tanh(y) is exactly 0 for y < -20,
            // and then the initial state (initial_conserved_array :182-191)
            // does not enter
            double r = sqrt_nr((0.0 + x * x) + y * y);
            double yb = (r - prm.R) * prm.inv_buffer_scale;
            double br = 0.0;
            double ui[3] = { 1.0, 0.0, 0.0 };
            if (yb > -20.0) {
                // 0.5 (1 + tanh y) = 1 / (1 + exp(-2 y)); reciprocals of the
                // scale and of r0 in place of divisions (each within an ulp)
                br = prm.buffer_rate * rcp_nr(1.0 + exp(-2.0 * yb));
                double ph = -1.0 * rsqrt_nr((x * x + y * y) + prm.s2);
                double r0 = sqrt_nr(x * x + y * y);
                double vp = sqrt_nr(-ph) - om * r0;
                const double ir0 = rcp_nr(r0);
                ui[1] = 1.0 * (vp * (-y * ir0));
                ui[2] = 1.0 * (vp * (x * ir0));
            }
            double sig = pc[0];
            double sg1[3] = { 0.0 / 1.0, sig * ax1, sig * ay1 };
            double sg2[3] = { 0.0 / 1.0, sig * ax2, sig * ay2 };
            double sf1[3] = { 0.0 / 1.0, sig * cenx, sig * ceny };
            double sf2[3] = { 0.0 / 1.0, sig * corx, sig * cory };
            const double dtdl = dt / dl;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                double ss1 = -u[q] * sr1;
                double ss2 = -u[q] * sr2;
                double sb = -(u[q] - ui[q]) * br;
                double src = sg1[q] + sg2[q] + ss1 + ss2 + sf1[q] + sf2[q] + sb;
                out[q] = u[q] - (dfx[q] + dfy[q]) * dtdl + src * dt;
            }
        } else {
            float xf = td.xcf[k - tc.k0], yf = td.ycf[c - tc.c0];
            float s2 = float(prm.s2);
            float dx1 = xf - float(prm.x1), dy1 = yf - float(prm.y1);
            float dx2 = xf - float(prm.x2), dy2 = yf - float(prm.y2);
            float q1 = fmaf(dx1, dx1, dy1 * dy1), q2 = fmaf(dx2, dx2, dy2 * dy2);
            float i1 = __builtin_amdgcn_rsqf(q1 + s2), i2 = __builtin_amdgcn_rsqf(q2 + s2);
            float g1 = -float(prm.m1) * i1 * i1 * i1;
            float g2 = -float(prm.m2) * i2 * i2 * i2;
            float om = float(prm.omega);
            float sig = pc[0];
            float ax = fmaf(g1, dx1, fmaf(g2, dx2, fmaf(om * om, xf, 2.0f * om * pc[2])));
            float ay = fmaf(g1, dy1, fmaf(g2, dy2, fmaf(om * om, yf, -2.0f * om * pc[1])));
            // Sink and buffer terms are evaluated only where they are not
            // negligible (skipped wave-uniformly elsewhere): the sink
            // exp(-q^3 / r_sink^6) is below 1e-45 once its argument passes
            // 104, and the buffer rate is below 4e-18 of its maximum once
            // (r - R) / scale < -20.  The initial state U_init enters only the
            // buffer term.
            float sr = 0.0f;
            if (fminf(q1, q2) < float(prm.q_sink_cut)) {
                float is6 = float(prm.inv_s6);
                float t1 = q1 * q1 * q1 * is6, t2 = q2 * q2 * q2 * is6;
                sr = float(prm.sink_rate) * (__expf(-t1) + __expf(-t2));
            }
            float q0 = fmaf(xf, xf, yf * yf);
            float ir0 = __builtin_amdgcn_rsqf(q0);
            float rr = q0 * ir0;
            float yb = (rr - float(prm.R)) * float(prm.inv_buffer_scale);
            float br = 0.0f, ui[3] = { 1.0f, 0.0f, 0.0f };
            if (yb > -20.0f) {
                // 0.5 * rate * (1 + tanh(y)) = rate / (1 + exp(-2y))
                br = float(prm.buffer_rate) * __builtin_amdgcn_rcpf(1.0f + __expf(-2.0f * yb));
                float ph = __builtin_amdgcn_rsqf(q0 + s2);
                float vp = __builtin_amdgcn_sqrtf(ph) - om * rr;
                ui[1] = vp * (-yf * ir0);
                ui[2] = vp * (xf * ir0);
            }
            float sa[3] = { 0.0f, sig * ax, sig * ay };
            double dtdl = dt / dl;
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                float src = sa[q] - float(u[q]) * sr - float(u[q] - double(ui[q])) * br;
                out[q] = fma(double(src), dt, fma(-(dfx[q] + dfy[q]), dtdl, u[q]));
            }
        }
    }
};

// ---------------------------------------------------------------------------
// locally_isothermal::updated_conserved: the face sound speeds squared come
// from arrays cs2x [nblk][(n+1)][n] and cs2y [nblk][n][(n+1)] evaluated by the
// caller at the face coordinates of each block (the reference evaluates a
// host std::function per face, src/module_locally_isothermal.cpp:115-121).
struct IsoTableParams {
    const double* cs2x;
    const double* cs2y;
    int n;
};

template<typename R>
struct IsoTableScheme {
    static constexpr int NC = 3;
    static constexpr int kWavesPerEU = 2;
    static constexpr bool kLateBase = false;
    static constexpr int kMarchUnroll = 2;
    static constexpr int kPrefetchRows = 1;
    static constexpr bool kCfl = false;
    using Real = R;
    using Params = IsoTableParams;
    __device__ static __forceinline__ double cfl_speed(const double*, const Params&) { return 0.0; }
    template<int H, int TJ> struct TileData {};

    // theta = 2 is hard-coded by the reference (src/module_locally_isothermal.cpp:195-196)
    __device__ static __forceinline__ double theta(const Params&) { return 2.0; }
    __device__ static __forceinline__ bool recover(const double* u, const Params&, R* p)
    {
        Iso2d<R>::recover(u, p);
        return true;
    }
    template<int H, int TJ>
    __device__ static __forceinline__ void setup(TileData<H, TJ>&, const Params&, const TileCtx&) {}
    __device__ static __forceinline__ void setup_zone(TileData<1, 1>&, const Params&, const TileCtx&) {}

    template<int AXIS, int H, int TJ>
    __device__ static __forceinline__ void face(const R* pl, const R* pr, const Params& prm,
                                                const TileData<H, TJ>&, const TileCtx& tc, int fi, int fj, R* f)
    {
        const size_t per_block = (size_t)(prm.n + 1) * prm.n;
        R cs2 = R(AXIS == 1 ? prm.cs2x[tc.b * per_block + (size_t)fi * prm.n + fj]
                            : prm.cs2y[tc.b * per_block + (size_t)fi * (prm.n + 1) + fj]);
        Iso2d<R>::template hlle<AXIS>(pl, pr, cs2, f);
    }

    // U - ((dFx + dFy) * dt) / dl   (src/module_locally_isothermal.cpp:209)
    template<int H, int TJ>
    __device__ static __forceinline__ void update(const Params&, const TileData<H, TJ>&, const TileCtx&,
                                                  int, int, const double* u, const R*,
                                                  const double* dfx, const double* dfy,
                                                  double dt, double dl, double* out)
    {
#pragma unroll
        for (int q = 0; q < NC; ++q) out[q] = u[q] - (dfx[q] + dfy[q]) * dt / dl;
    }
};

} // namespace m3p
