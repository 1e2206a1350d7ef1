// Fused RK2 step of the 1D work-list stage (mixed precision).
#include <climits>
#include "launch.hpp"
#include "rk2fused1d.hpp"

namespace m3p {

int launch_rk2_fused_1d(const EulerParams& prm, const Stage1DArgs& a, hipStream_t s)
{
    const int n = a.n;
    // one workgroup per (block, tile) item; the kernel is small (about 5 KB
    // of LDS at TJ = 64), so many workgroups share a CU and hide each
    // other's load latency
    auto go = [&] (auto kern, int TJ) {
        const long items = (long)a.nblk * (n / TJ);
        if (items > INT_MAX) return fail(M3P_ERR_BAD_SHAPE, "m3p: too many 1D blocks for one launch");
        hipLaunchKernelGGL(kern, dim3((unsigned)items), dim3(TJ), 0, s, prm, a);
        return launch_status("rk2_fused_1d_kernel");
    };
    switch (n) {
        case 16:  return go(rk2_fused_1d_kernel<16>, 16);
        case 32:  return go(rk2_fused_1d_kernel<32>, 32);
        case 64:  return go(rk2_fused_1d_kernel<64>, 64);
        case 128: return go(rk2_fused_1d_kernel<128>, 128);
        default:
            if (n >= 256 && n % 256 == 0) return go(rk2_fused_1d_kernel<256>, 256);
    }
    return fail(M3P_ERR_BAD_SHAPE, "m3p: 1D block size must be 16, 32, 64, 128 or a multiple of 256 (got " +
                                   std::to_string(n) + ")");
}

} // namespace m3p
