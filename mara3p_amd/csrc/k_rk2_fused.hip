// Fused RK2 steps of the mixed minidisk and euler2d schemes (their own unit:
// large kernels).
#include "launch.hpp"
#include "rk2fused.hpp"

namespace m3p {

template<class S>
int launch_rk2_fused(const typename S::Params& p1, const typename S::Params& p2, const MeshLoader<S>& ld,
                     const StageArgs& a, hipStream_t s)
{
    // 32 rows per workgroup: the first stage's four extra rows cost 1/8 of
    // its march, and a 256-block mesh of 256^2 gives 2048 workgroups (64
    // rows: about 55 KB of LDS, two workgroups per CU, 14 % slower)
    constexpr int H = 32;
    const int grid = a.nblk * (a.n / H);
    switch (a.n) {
        case 64:  hipLaunchKernelGGL((rk2_fused_kernel<S, H, 64>), dim3(grid), dim3(64), 0, s, p1, p2, ld, a); break;
        case 128: hipLaunchKernelGGL((rk2_fused_kernel<S, H, 128>), dim3(grid), dim3(128), 0, s, p1, p2, ld, a); break;
        case 256: hipLaunchKernelGGL((rk2_fused_kernel<S, H, 256>), dim3(grid), dim3(256), 0, s, p1, p2, ld, a); break;
        default:
            return fail(M3P_ERR_BAD_SHAPE, "m3p: the fused RK2 step takes blocks of 64, 128 or 256 zones a side (got " +
                                           std::to_string(a.n) + ")");
    }
    return launch_status("rk2_fused_kernel");
}

template int launch_rk2_fused<MinidiskScheme<float>>(const MinidiskScheme<float>::Params&,
                                                      const MinidiskScheme<float>::Params&,
                                                      const MeshLoader<MinidiskScheme<float>>&, const StageArgs&,
                                                      hipStream_t);
template int launch_rk2_fused<EulerScheme<float>>(const EulerScheme<float>::Params&,
                                                   const EulerScheme<float>::Params&,
                                                   const MeshLoader<EulerScheme<float>>&, const StageArgs&,
                                                   hipStream_t);

} // namespace m3p
