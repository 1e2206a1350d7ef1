// Definitions of the stage launchers declared in launch.hpp, included only by
// the k_*.hip units that instantiate them (M3P_INSTANTIATE_* below).
#pragma once
#include <algorithm>
#include "launch.hpp"
#include "viscous2d.hpp"

namespace m3p {

template<class S, class Ld, bool FIRST>
int launch_stage2d_variant(TileShape ts, const typename S::Params& prm, const Ld& ld, const StageArgs& a, hipStream_t s)
{
    // The mixed kernels keep the conserved state in registers between its load
    // and its update; the FP64 kernels (register-bound) re-load it.
    constexpr bool K = !is_fp64<typename S::Real>::value;
    const int n = a.n;
    auto go = [&] (auto kern) {
        const int grid = a.nblk * (n / ts.h) * (n / ts.tj);
        hipLaunchKernelGGL(kern, dim3(grid), dim3(ts.tj), 0, s, prm, ld, a);
        return launch_status("stage2d_kernel");
    };
    if (ts.tj == 256 && ts.h == 128) return go(stage2d_kernel<S, Ld, 128, 256, K, FIRST>);
    if (ts.tj == 256 && ts.h == 64) return go(stage2d_kernel<S, Ld, 64, 256, K, FIRST>);
    if (ts.tj == 256 && ts.h == 32) return go(stage2d_kernel<S, Ld, 32, 256, K, FIRST>);
    if (ts.tj == 256 && ts.h == 16) return go(stage2d_kernel<S, Ld, 16, 256, K, FIRST>);
    if (ts.tj == 128 && ts.h == 64) return go(stage2d_kernel<S, Ld, 64, 128, K, FIRST>);
    if (ts.tj == 128 && ts.h == 32) return go(stage2d_kernel<S, Ld, 32, 128, K, FIRST>);
    if (ts.tj == 64 && ts.h == 64)  return go(stage2d_kernel<S, Ld, 64, 64, K, FIRST>);
    if (ts.tj == 64 && ts.h == 32)  return go(stage2d_kernel<S, Ld, 32, 64, K, FIRST>);
    if (ts.tj == 64 && ts.h == 16)  return go(stage2d_kernel<S, Ld, 16, 64, K, FIRST>);
    if (ts.tj == 32 && ts.h == 32)  return go(stage2d_kernel<S, Ld, 32, 32, K, FIRST>);
    if (ts.tj == 16 && ts.h == 16)  return go(stage2d_kernel<S, Ld, 16, 16, K, FIRST>);
    return fail(M3P_ERR_BAD_SHAPE, "m3p: 2D block size must be 16, 32 or a multiple of 64 (got " + std::to_string(n) + ")");
}

template<class S, class Ld>
int launch_stage2d_shape(TileShape ts, const typename S::Params& prm, const Ld& ld, const StageArgs& a, hipStream_t s)
{
    // A stage without a base state has its own kernels only where the scheme
    // fetches more rows ahead there (and the loader pipelines its rows).
    if constexpr (S::kPrefetchRows > 1 && Ld::kPrefetch)
        if (!a.u0) return launch_stage2d_variant<S, Ld, true>(ts, prm, ld, a, s);
    return launch_stage2d_variant<S, Ld, false>(ts, prm, ld, a, s);
}

template<class S, class Ld>
int launch_visc(const typename S::Params& prm, const Ld& ld, const StageArgs& a, double nu, hipStream_t s)
{
    const long zones = (long)a.nblk * a.n * a.n;
    hipLaunchKernelGGL((iso_visc_stage_kernel<S, Ld>), dim3((zones + 255) / 256), dim3(256), 0, s, prm, ld, a, nu);
    return launch_status("iso_visc_stage_kernel");
}

template<typename R>
int launch_stage1d(const EulerParams& prm, const Stage1DArgs& a, hipStream_t s, int cus)
{
    const int n = a.n;
    // persistent grid: 16 (FP64) or 32 (mixed) waves per CU, somewhat above
    // what is resident at once (about 160 and 100 VGPRs: 12 and 20 waves);
    // the excess workgroups start as others finish.  A grid of exactly the
    // resident workgroups (the runtime's occupancy query) measured 1.5 %
    // faster in FP64 and 3 % slower in the mixed second stage.
    const int waves_per_cu = is_fp64<R>::value ? 16 : 32;
    auto go = [&] (auto kern, int TJ) {
        const long items = (long)a.nblk * (n / TJ);
        const int wgs_per_cu = std::max(1, waves_per_cu / std::max(1, TJ / 64));
        const int grid = (int)std::min<long>(items, (long)std::max(cus, 1) * wgs_per_cu);
        hipLaunchKernelGGL(kern, dim3(grid), dim3(TJ), 0, s, prm, a);
        return launch_status("stage1d_kernel");
    };
    switch (n) {
        case 16:  return go(stage1d_kernel<R, 16>, 16);
        case 32:  return go(stage1d_kernel<R, 32>, 32);
        case 64:  return go(stage1d_kernel<R, 64>, 64);
        case 128: return go(stage1d_kernel<R, 128>, 128);
        default:
            if (n >= 256 && n % 256 == 0) return go(stage1d_kernel<R, 256>, 256);
    }
    return fail(M3P_ERR_BAD_SHAPE, "m3p: 1D block size must be 16, 32, 64, 128 or a multiple of 256 (got " + std::to_string(n) + ")");
}

} // namespace m3p

#define M3P_INSTANTIATE_STAGE2D(S, LD) \
    template int m3p::launch_stage2d_shape<S, LD<S>>(m3p::TileShape, const typename S::Params&, const LD<S>&, \
                                                    const m3p::StageArgs&, hipStream_t);
#define M3P_INSTANTIATE_VISC(S, LD) \
    template int m3p::launch_visc<S, LD<S>>(const typename S::Params&, const LD<S>&, const m3p::StageArgs&, \
                                           double, hipStream_t);
