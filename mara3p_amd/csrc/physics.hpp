// Pointwise physics on the device: Euler (ideal gas) and isothermal 2D.
//
// The FP64 instantiations restate the reference operation for operation
// (same association order, IEEE division and square root), so with
// -ffp-contract=off they reproduce its results exactly.  The FP32
// instantiations are the "mixed" form: same algebra, reciprocal-multiply
// and explicit FMA, accumulated into FP64 state by the caller.
#pragma once
#include "device_common.hpp"

namespace m3p {

// ---------------------------------------------------------------------------
// Euler, src/physics_euler.hpp.  Primitive (rho, v1, v2, v3, p); conserved
// (D, S1, S2, S3, E).
struct EulerParams {
    double gamma;
    double theta;
    int hllc;
};

template<typename R>
struct Euler {
    static constexpr int NC = 5;
    using Real = R;
    using Params = EulerParams;

    // recover_primitive, src/physics_euler.hpp:156-169.  Returns false on D < 0.
    __device__ static __forceinline__ bool recover(const double* u, const Params& prm, R* p)
    {
        if constexpr (is_fp64<R>::value) {
            double D = u[0], S1 = u[1], S2 = u[2], S3 = u[3], E = u[4];
            double ls = ((0.0 + S1 * S1) + S2 * S2) + S3 * S3;
            double ke = 0.5 * ls / D;
            p[0] = D;
            p[1] = S1 / D;
            p[2] = S2 / D;
            p[3] = S3 / D;
            p[4] = (E - ke) * (prm.gamma - 1.0);
            return !(D < 0.0);
        } else {
            // Kinetic energy is removed in FP64 so the pressure keeps its
            // digits when E is dominated by kinetic energy (the reciprocal
            // of D Newton-refined, within a few ulp).
            double D = u[0];
            double iD = rcp_nr(D);
            float fiD = float(iD);
            double ls = (u[1] * u[1] + u[2] * u[2]) + u[3] * u[3];
            p[0] = float(D);
            p[1] = float(u[1]) * fiD;
            p[2] = float(u[2]) * fiD;
            p[3] = float(u[3]) * fiD;
            p[4] = float((u[4] - 0.5 * ls * iD) * (prm.gamma - 1.0));
            return !(D < 0.0);
        }
    }

    // conserved_density :175-183, flux :189-203
    template<int AXIS>
    __device__ static __forceinline__ void cons_flux(const R* p, R gamma, R igm1, R* u, R* f)
    {
        R d = p[0], v1 = p[1], v2 = p[2], v3 = p[3], pg = p[4];
        if constexpr (is_fp64<R>::value) {
            R ls = ((R(0) + v1 * v1) + v2 * v2) + v3 * v3;
            u[0] = d;
            u[1] = d * v1;
            u[2] = d * v2;
            u[3] = d * v3;
            u[4] = d * ls * R(0.5) + pg / (gamma - R(1));
            R vn = R(0) + p[AXIS];
            f[0] = vn * u[0] + R(0);
            f[1] = vn * u[1] + pg * R(AXIS == 1 ? 1 : 0);
            f[2] = vn * u[2] + pg * R(AXIS == 2 ? 1 : 0);
            f[3] = vn * u[3] + pg * R(AXIS == 3 ? 1 : 0);
            f[4] = vn * u[4] + pg * vn;
        } else {
            R ls = fmaf(v3, v3, fmaf(v2, v2, v1 * v1));
            u[0] = d;
            u[1] = d * v1;
            u[2] = d * v2;
            u[3] = d * v3;
            u[4] = fmaf(d * ls, R(0.5), pg * igm1);
            R vn = p[AXIS];
            f[0] = vn * u[0];
            f[1] = vn * u[1];
            f[2] = vn * u[2];
            f[3] = vn * u[3];
            f[AXIS] += pg;
            f[4] = vn * (u[4] + pg);
        }
    }

    // outer_wavespeeds :118-128
    template<int AXIS>
    __device__ static __forceinline__ void speeds(const R* p, R gamma, R& am, R& ap)
    {
        if constexpr (is_fp64<R>::value) {
            R vn = R(0) + p[AXIS];
            R cs = sqrt(gamma * p[4] / p[0]);
            am = vn - cs;
            ap = vn + cs;
        } else {
            R vn = p[AXIS];
            R cs = fsqrt(gamma * p[4] * frcp(p[0]));
            am = vn - cs;
            ap = vn + cs;
        }
    }

    template<int AXIS>
    __device__ static __forceinline__ void state(const R* p, R gamma, R igm1, R* u, R* f, R& am, R& ap)
    {
        cons_flux<AXIS>(p, gamma, igm1, u, f);
        speeds<AXIS>(p, gamma, am, ap);
    }

    // riemann_hlle :209-222
    template<int AXIS>
    __device__ static __forceinline__ void hlle(const R* pl, const R* pr, const Params& prm, R* f)
    {
        R gamma = R(prm.gamma);
        R igm1 = frcp(gamma - R(1));
        R ul[NC], ur[NC], fl[NC], fr[NC], alm, alp, arm, arp;
        state<AXIS>(pl, gamma, igm1, ul, fl, alm, alp);
        state<AXIS>(pr, gamma, igm1, ur, fr, arm, arp);
        R ap = smax(R(0), smax(alp, arp));
        R am = smin(R(0), smin(alm, arm));
        if constexpr (is_fp64<R>::value) {
#pragma unroll
            for (int k = 0; k < NC; ++k)
                f[k] = (ap * fl[k] - am * fr[k] - (ul[k] - ur[k]) * ap * am) / (ap - am);
        } else {
            // the wave-speed weights folded in: F = wp F_L + wm F_R + wpm (U_L - U_R)
            // with wp = a+ / (a+ - a-), wm = -a- / (a+ - a-), wpm = -a+ a- / (a+ - a-)
            R inv = frcp(ap - am);
            R wp = ap * inv, wm = -am * inv;
            R wpm = wp * -am;
#pragma unroll
            for (int k = 0; k < NC; ++k)
                f[k] = fmaf(wp, fl[k], fmaf(wm, fr[k], wpm * (ul[k] - ur[k])));
        }
    }

    // HLLC (Toro 3rd ed. sec. 10.4), same wave-speed estimates as hlle.  No
    // reference counterpart (the reference's only HLLC is relativistic).
    // The wave speeds and the contact speed come from the primitives; the
    // conserved state and flux are then formed for the one side the flux
    // takes (left if 0 <= S_L, right if S_R <= 0, else the side of the
    // contact) -- the values of forming both and selecting.
    template<int AXIS>
    __device__ static __forceinline__ void hllc(const R* pl, const R* pr, const Params& prm, R* f)
    {
        R gamma = R(prm.gamma);
        R igm1 = frcp(gamma - R(1));
        R alm, alp, arm, arp;
        speeds<AXIS>(pl, gamma, alm, alp);
        speeds<AXIS>(pr, gamma, arm, arp);
        R sl = smin(alm, arm);
        R sr = smax(alp, arp);
        R vl = pl[AXIS], vr = pr[AXIS];
        R ql = pl[0] * (sl - vl), qr = pr[0] * (sr - vr);
        // unused (and possibly not finite) where the flow is supersonic
        R ss = fdiv(pr[4] - pl[4] + vl * ql - vr * qr, ql - qr);
        const bool sup_l = R(0) <= sl, sup_r = sr <= R(0);
        const bool left = sup_l || (!sup_r && R(0) <= ss);
        R p[NC], u[NC], fk[NC];
#pragma unroll
        for (int k = 0; k < NC; ++k) p[k] = left ? pl[k] : pr[k];
        cons_flux<AXIS>(p, gamma, igm1, u, fk);
        if (sup_l || sup_r) {
#pragma unroll
            for (int k = 0; k < NC; ++k) f[k] = fk[k];
            return;
        }
        R s = left ? sl : sr;
        R vk = p[AXIS];
        R fac = fdiv(p[0] * (s - vk), s - ss);
        R us[NC];
        us[0] = fac;
        us[1] = fac * p[1];
        us[2] = fac * p[2];
        us[3] = fac * p[3];
        us[AXIS] = fac * ss;
        us[4] = fac * (fdiv(u[4], p[0]) + (ss - vk) * (ss + fdiv(p[4], p[0] * (s - vk))));
#pragma unroll
        for (int k = 0; k < NC; ++k) f[k] = fk[k] + s * (us[k] - u[k]);
    }

    template<int AXIS>
    __device__ static __forceinline__ void riemann(const R* pl, const R* pr, const Params& prm, R* f)
    {
        if (prm.hllc) hllc<AXIS>(pl, pr, prm, f);
        else          hlle<AXIS>(pl, pr, prm, f);
    }

    // CFL signal speed max(|v1|, |v2|) + c_s of a conserved state (one
    // reciprocal of the density); the stage kernels that fuse the CFL
    // reduction and the separate reduction evaluate the same expression.
    __device__ static __forceinline__ double max_speed(const double* u, const Params& prm)
    {
        double D = u[0], iD = 1.0 / D;
        double v1 = u[1] * iD, v2 = u[2] * iD, v3 = u[3] * iD;
        double pg = (u[4] - 0.5 * ((v1 * u[1] + v2 * u[2]) + v3 * u[3])) * (prm.gamma - 1.0);
        double cs = sqrt(prm.gamma * pg * iD);
        return fmax(fabs(v1), fabs(v2)) + cs;
    }
};

// ---------------------------------------------------------------------------
// Isothermal 2D, src/physics_iso2d.hpp.  Primitive (Sigma, vx, vy); conserved
// (Sigma, Sx, Sy).  The face sound speed squared is an input.
template<typename R>
struct Iso2d {
    static constexpr int NC = 3;

    // recover_primitive :119-125 (no negative-density check in the reference)
    __device__ static __forceinline__ void recover(const double* u, R* p)
    {
        if constexpr (is_fp64<R>::value) {
            p[0] = u[0];
            p[1] = u[1] / u[0];
            p[2] = u[2] / u[0];
        } else {
            float s = float(u[0]);
            float is = __builtin_amdgcn_rcpf(s);
            p[0] = s;
            p[1] = float(u[1]) * is;
            p[2] = float(u[2]) * is;
        }
    }

    // riemann_hlle :156-173 with flux :139-151; RCP (FP64 only, the
    // minidisk kernel) takes the sound speed and one reciprocal of a+ - a-
    // from the Newton-refined estimates and multiplies each component by it
    // (within a few ulp of the reference's form)
    template<int AXIS, bool RCP = false>
    __device__ static __forceinline__ void hlle(const R* pl, const R* pr, R cs2, R* f)
    {
        if constexpr (is_fp64<R>::value) {
            R ul[3] = { pl[0], pl[0] * pl[1], pl[0] * pl[2] };
            R ur[3] = { pr[0], pr[0] * pr[1], pr[0] * pr[2] };
            R vnl = R(0) + pl[AXIS], vnr = R(0) + pr[AXIS];
            R pgl = pl[0] * cs2, pgr = pr[0] * cs2;
            R fl[3] = { vnl * ul[0] + R(0), vnl * ul[1] + pgl * R(AXIS == 1 ? 1 : 0), vnl * ul[2] + pgl * R(AXIS == 2 ? 1 : 0) };
            R fr[3] = { vnr * ur[0] + R(0), vnr * ur[1] + pgr * R(AXIS == 1 ? 1 : 0), vnr * ur[2] + pgr * R(AXIS == 2 ? 1 : 0) };
            R cs = RCP ? sqrt_nr(cs2) : sqrt(cs2);
            R alm = vnl - cs, alp = vnl + cs;
            R arm = vnr - cs, arp = vnr + cs;
            R ap = smax(R(0), smax(alp, arp));
            R am = smin(R(0), smin(alm, arm));
            if constexpr (RCP) {
                const R inv = rcp_nr(ap - am);
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    f[k] = (ap * fl[k] - am * fr[k] - (ul[k] - ur[k]) * ap * am) * inv;
            } else {
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    f[k] = (ap * fl[k] - am * fr[k] - (ul[k] - ur[k]) * ap * am) / (ap - am);
            }
        } else {
            // The same HLL flux with the wave-speed weights folded in:
            // wp = a+ / (a+ - a-), wm = -a- / (a+ - a-), wpm = -a+ a- / (a+ - a-),
            // F = wp F_L + wm F_R + wpm (U_L - U_R), and with F = U v_n (+ p on
            // the normal component), U = Sigma (1, v1, v2):
            // F_k = Sigma_L v_kL al + Sigma_R v_kR ar (+ cs2 (wp Sigma_L + wm Sigma_R)),
            // al = wp v_nL + wpm, ar = wm v_nR - wpm (v_0 = 1).  a+ - a- >= c_s > 0.
            R vnl = pl[AXIS], vnr = pr[AXIS];
            R cs = fsqrt(cs2);
            R ap = fmaxf(0.f, fmaxf(vnl, vnr) + cs);
            R am = fminf(0.f, fminf(vnl, vnr) - cs);
            R inv = frcp(ap - am);
            R wp = ap * inv, wm = -am * inv;
            R wpm = wp * -am;
            R al = fmaf(wp, vnl, wpm), ar = fmaf(wm, vnr, -wpm);
            R bl = pl[0] * al, br = pr[0] * ar;
            f[0] = bl + br;
            f[1] = fmaf(pl[1], bl, pr[1] * br);
            f[2] = fmaf(pl[2], bl, pr[2] * br);
            f[AXIS] = fmaf(cs2, fmaf(wp, pl[0], wm * pr[0]), f[AXIS]);
        }
    }
};

} // namespace m3p
