// Device helpers shared by every kernel.  CDNA4 (gfx950), wave64.
//
// All HIP code is compiled with -ffp-contract=off: no FMA is formed unless it
// is written explicitly, so the FP64 kernels round exactly as the reference's
// x86-64 build and every face flux is bit-identical wherever it is computed
// (which keeps fluxes shared by two tiles conservative).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace m3p {

constexpr int kWave = 64;

// std::max(a, b) == (a < b) ? b : a ; std::min(a, b) == (b < a) ? b : a
template<typename R> __device__ __forceinline__ R smax(R a, R b) { return (a < b) ? b : a; }
template<typename R> __device__ __forceinline__ R smin(R a, R b) { return (b < a) ? b : a; }

template<typename R> struct is_fp64 : std::is_same<R, double> {};

// Generalised-minmod PLM slope, src/scheme_plm_gradient.hpp:59-67 with
// coordinates (0, 1, 2).  The divisions by 1.0 and 2.0 there are exact, so
// they are written as the identical-valued forms below.
//
// The reference's 0.25 |sa + sb| (sa + sc) min(|a|, |b|, |c|) (s = sign) is
// sign(a) min|.| when a, b, c share a sign (the factors are 0.25 * 2 * 2,
// exact) and zero otherwise: the value of the median of (min(a, b, c),
// max(a, b, c), 0), which FP64 takes as max(min, min(max, 0)) -- the same
// number in every case (a zero may differ only in its sign, which no sum
// or product that follows can tell apart from the other zero's, since the
// face state adds it to a nonzero primitive or the product is zero anyway).
template<typename R>
__device__ __forceinline__ R plm(R yl, R y0, R yr, R theta)
{
    R a = (y0 - yl) * theta;
    R b = (yr - yl) * R(0.5);
    R c = (yr - y0) * theta;
    if constexpr (is_fp64<R>::value) {
        R mn = fmin(fmin(a, b), c), mx = fmax(fmax(a, b), c);
        return fmax(mn, fmin(mx, R(0)));
    } else {
        // Same value: non-zero only when a, b, c share a sign; then sign(a) * min|.|.
        R m = fminf(fminf(fabsf(a), fabsf(b)), fabsf(c));
        bool same = (signbit(a) == signbit(b)) && (signbit(a) == signbit(c));
        return same ? copysignf(m, a) : R(0);
    }
}

// Half the PLM slope, 0.5 * plm(...), used for the face states p -/+ 0.5 g.
// FP64: the reference form times 0.5 (exact).  FP32: the same value from
// three-operand min/max/median: a common sign gives sign * min|.|, else 0.
template<typename R>
__device__ __forceinline__ R plm_half(R yl, R y0, R yr, R theta)
{
    if constexpr (is_fp64<R>::value) {
        return R(0.5) * plm(yl, y0, yr, theta);
    } else {
        R ht = R(0.5) * theta;
        R a = (y0 - yl) * ht;
        R b = (yr - yl) * R(0.25);
        R c = (yr - y0) * ht;
        // median of (min, max, 0): min if all are positive, max if all are
        // negative, 0 when the signs differ
        return __builtin_amdgcn_fmed3f(fminf(fminf(a, b), c), fmaxf(fmaxf(a, b), c), R(0));
    }
}

// Division / square root: IEEE for FP64 (parity), fast for FP32 (mixed).
template<typename R> __device__ __forceinline__ R fdiv(R a, R b)
{
    if constexpr (is_fp64<R>::value) return a / b;
    else return a * __builtin_amdgcn_rcpf(b);
}
template<typename R> __device__ __forceinline__ R frcp(R b)
{
    if constexpr (is_fp64<R>::value) return R(1) / b;
    else return __builtin_amdgcn_rcpf(b);
}
template<typename R> __device__ __forceinline__ R fsqrt(R a)
{
    if constexpr (is_fp64<R>::value) return sqrt(a);
    else return __builtin_amdgcn_sqrtf(a);
}

// FP64 reciprocal, reciprocal square root and square root from the
// hardware estimates (v_rcp_f64, v_rsq_f64) refined by two Newton steps each
// (quadratic convergence from the estimate's ~2^-22 to the last bit or two):
// a few ulp instead of IEEE rounding, at a third of the instructions of the
// IEEE division and square root sequences.  For the FP64 kernels whose bar
// against the reference is 1e-12 relative (minidisk, MHD), never for the
// ones that equal it bit for bit.
__device__ __forceinline__ double rcp_nr(double b)
{
    double r = __builtin_amdgcn_rcp(b);
    r = fma(r, fma(-b, r, 1.0), r);
    return fma(r, fma(-b, r, 1.0), r);
}
__device__ __forceinline__ double rsqrt_nr(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    const double h = 0.5 * x;
    y = y * fma(-h * y, y, 1.5);
    return y * fma(-h * y, y, 1.5);
}
// sqrt(x) = x rsqrt(x) for finite x > 0; +-0, +inf and NaN as they are,
// NaN for x < 0 (the IEEE results)
__device__ __forceinline__ double sqrt_nr(double x)
{
    if (x > 0.0 && x < __builtin_inf()) return x * rsqrt_nr(x);
    return x < 0.0 ? __builtin_nan("") : x;
}

// Error slot: err[0] = status code (first writer wins), err[1] = block, err[2] = zone.
__device__ __forceinline__ void report_error(int* err, int code, int block, int zone)
{
    if (err && atomicCAS(&err[0], 0, code) == 0) {
        err[1] = block;
        err[2] = zone;
    }
}

// Max of a non-negative double over a workgroup of NT threads (NT a multiple
// of 64, or one partial wave), then one atomic max into one of kRateSlots
// slots picked by the workgroup index (non-negative doubles order like their
// bit patterns; spreading the atomics keeps thousands of workgroups that end
// together from serialising on one address).  The atomic is skipped when the
// slot already holds at least the workgroup's maximum: slots only grow, so a
// stale read can only cost an atomic, never lose one, and a launch of tens
// of thousands of small workgroups does few atomics after its first ones.
// Every thread of the workgroup must call it.
constexpr int kRateSlots = 64;

__device__ __forceinline__ void slot_atomic_max(unsigned long long* slot, double m)
{
    const unsigned long long mb = (unsigned long long)__double_as_longlong(m);
    if (mb > __atomic_load_n(slot, __ATOMIC_RELAXED)) atomicMax(slot, mb);
}

template<int NT>
__device__ __forceinline__ void workgroup_atomic_max(double v, unsigned long long* slots)
{
    unsigned long long* slot = slots + (blockIdx.x & (kRateSlots - 1));
    __shared__ double part[(NT + 63) / 64 > 1 ? (NT + 63) / 64 : 1];
    __shared__ double lanes[NT < 64 ? NT : 1];
    if constexpr (NT >= 64) {
        for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
        if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = part[0];
            for (int w = 1; w < NT / 64; ++w) m = fmax(m, part[w]);
            slot_atomic_max(slot, m);
        }
    } else {
        // a partial wave: no shuffles across inactive lanes
        lanes[threadIdx.x] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = lanes[0];
            for (int i = 1; i < NT; ++i) m = fmax(m, lanes[i]);
            slot_atomic_max(slot, m);
        }
    }
}

// Fill n words with v (clears reduction slots; a kernel, not a memset, so a
// captured graph replays it correctly).
static __global__ void fill_u64_kernel(unsigned long long* p, int n, unsigned long long v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Wave64 reductions through DPP/permute shuffles.
__device__ __forceinline__ double wave_max(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
    return v;
}
__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// XCD-aware workgroup remap.  Workgroups are dispatched round-robin over the
// 8 XCDs (each with its own L2); give each XCD a contiguous range of logical
// tiles so that tiles sharing halo rows share an L2.
__device__ __forceinline__ int xcd_remap(int w, int total)
{
    constexpr int nxcd = 8;
    if (total % nxcd != 0) return w;
    return (w % nxcd) * (total / nxcd) + (w / nxcd);
}

} // namespace m3p
