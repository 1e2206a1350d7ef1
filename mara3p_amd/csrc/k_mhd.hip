// MHD constrained-transport stage passes (both precisions).
#include <algorithm>
#include "launch.hpp"

namespace m3p {

// Cubes with N % 8 == 0 take the column-tile pass, whose face fluxes never
// leave the workgroup; other N write them, [5][nf] per face family.
static bool mhd_tiled(int N) { return N % 8 == 0 && N >= 8; }

long mhd_workspace_bytes(int N, int precision)
{
    const long N3 = (long)N * N * N, nf = (long)(N + 1) * N * N;
    const long rs = precision == M3P_FP64 ? 8 : 4;
    // primitives [8][N^3], face fields [2][nf] per family, face fluxes (per-cell
    // passes only), + alignment, the speed slots and dt
    return (8 * N3 + 3 * 2 * nf + (mhd_tiled(N) ? 0 : 3 * 5 * nf)) * rs + 1024;
}

template<typename R>
int mhd_advance(MhdArgs a, char* work, hipStream_t s)
{
    const int N = a.N;
    const long N3 = (long)N * N * N, nf = (long)(N + 1) * N * N;
    const bool tiled = mhd_tiled(N);
    R* P = reinterpret_cast<R*>(work);
    R* E1 = P + 8 * N3;
    R* E2 = E1 + 2 * nf;
    R* E3 = E2 + 2 * nf;
    R* F1 = E3 + 2 * nf;                     // per-cell passes only
    R* F2 = F1 + 5 * nf;
    R* F3 = F2 + 5 * nf;
    char* tail = reinterpret_cast<char*>(tiled ? F1 : F3 + 5 * nf);
    tail += (256 - (reinterpret_cast<uintptr_t>(tail) & 255)) & 255;
    a.vmax = reinterpret_cast<unsigned long long*>(tail);
    if (!a.dt) a.dt = reinterpret_cast<double*>(tail + 8 * kRateSlots);
    const int gc = (int)((N3 + 255) / 256);
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, s, a.vmax, kRateSlots, 0ull);
    hipLaunchKernelGGL(mhd_prim_kernel<R>, dim3(std::min(gc, 4096)), dim3(256), 0, s, a, P);
    hipLaunchKernelGGL(mhd_dt_kernel, dim3(1), dim3(kRateSlots), 0, s, a);
    FaceE<R> fe{ E1, E2, E3, N, nf };
    if (tiled) {
        // column tiles: fluxes and the conserved update in one pass, then the
        // faces.  FP64 too (236 VGPRs at two waves per SIMD, no spills): at
        // 256^3 its RK2 step takes 5.77 ms against 7.06 ms through the
        // per-cell passes below, which write every face flux and read it back.
        // 16-plane segments: 4096 workgroups at 256^3, four full rounds of
        // the 1024 resident at four waves per SIMD (32-plane segments run
        // as fast; 16 x 16 column tiles 10 % slower).  TI must divide N: an
        // N with N % 16 == 8 takes 8-plane segments.
        const int TI = N < 64 ? N : (N % 16 == 0 ? 16 : 8);
        const int segs = N / TI;
        if (N % 32 == 0) {
            const int grid = (N / 8) * (N / 32) * segs;
            hipLaunchKernelGGL((mhd_tile_kernel<R, 8, 32>), dim3(grid), dim3(256), 0, s, a, (const R*)P, E1, E2, E3, TI);
        } else {
            const int grid = (N / 8) * (N / 8) * segs;
            hipLaunchKernelGGL((mhd_tile_kernel<R, 8, 8>), dim3(grid), dim3(64), 0, s, a, (const R*)P, E1, E2, E3, TI);
        }
        if (int st = launch_status("mhd_tile_kernel")) return st;
        // The face fields: FP64 by bricks of 4 x 4 x 32 cells, each edge field
        // loaded once per brick (at 256^3 the FP64 RK2 step 5.60 -> 5.31 ms;
        // bricks of 2 x 4 x 32: 5.34 ms); mixed per cell, where the FP32 face
        // fields that the four faces of an edge re-read come from cache and
        // the bricks ran 0.8 % slower.
        if constexpr (is_fp64<R>::value) {
            if (N % 32 == 0)
                hipLaunchKernelGGL((mhd_face_b_brick_kernel<R, 4, 4, 32>), dim3((N / 4) * (N / 4) * (N / 32)), dim3(512),
                                   0, s, a, fe);
            else
                hipLaunchKernelGGL((mhd_face_b_brick_kernel<R, 2, 4, 8>), dim3((N / 2) * (N / 4) * (N / 8)), dim3(64), 0,
                                   s, a, fe);
        } else {
            hipLaunchKernelGGL(mhd_face_b_kernel<R>, dim3(gc), dim3(256), 0, s, a, fe);
        }
        return launch_status("m3p_mhd_advance_ct");
    }
    hipLaunchKernelGGL(mhd_faces_kernel<R>, dim3(gc), dim3(256), 0, s, a, (const R*)P, F1, E1, F2, E2, F3, E3);
    hipLaunchKernelGGL(mhd_boundary_faces_kernel<R>, dim3((int)((3L * N * N + 255) / 256)), dim3(256), 0, s, a,
                       (const R*)P, F1, E1, F2, E2, F3, E3);
    hipLaunchKernelGGL(mhd_update_kernel<R>, dim3(gc), dim3(256), 0, s, a, (const R*)F1, (const R*)F2, (const R*)F3, fe);
    return launch_status("m3p_mhd_advance_ct");
}

template int mhd_advance<double>(MhdArgs, char*, hipStream_t);
template int mhd_advance<float>(MhdArgs, char*, hipStream_t);

} // namespace m3p
