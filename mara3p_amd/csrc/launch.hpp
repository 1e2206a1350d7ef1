// Host-side interface between the C ABI (capi.hip) and the translation units
// that compile the stage kernels (k_*.hip).  The launchers are templates
// declared here and explicitly instantiated, one scheme family per unit, so
// the kernel variants compile in parallel; capi.hip sees only the
// declarations and never instantiates a kernel of its own.
#pragma once
#include <hip/hip_runtime.h>
#include <string>
#include "mara3p_amd.h"
#include "mhd.hpp"
#include "stage1d.hpp"
#include "stage2d.hpp"

namespace m3p {

// Failure with a message for m3p_last_error (thread-local text).
int fail(int code, const std::string& msg);
int hip_check(hipError_t e, const char* what);
int launch_status(const char* what);

// 2D stage tiles: TJ columns x H rows (the compiled variants).
struct TileShape { int tj, h; };
constexpr TileShape kTileVariants[] = { {256, 128}, {256, 64}, {256, 32}, {256, 16}, {128, 64}, {128, 32},
                                        {64, 64}, {64, 32}, {64, 16}, {32, 32}, {16, 16} };
constexpr int kNumVariants = sizeof(kTileVariants) / sizeof(kTileVariants[0]);

// One 2D stage with tile shape ts (a compiled variant).
template<class S, class Ld>
int launch_stage2d_shape(TileShape ts, const typename S::Params& prm, const Ld& ld, const StageArgs& a, hipStream_t s);

// One viscous isothermal stage (per-zone kernel).
template<class S, class Ld>
int launch_visc(const typename S::Params& prm, const Ld& ld, const StageArgs& a, double nu, hipStream_t s);

// One fused RK2 step (both stages in one launch) over a uniform mesh of
// blocks of n = 64, 128 or 256 zones a side: p1, p2 the two stages'
// parameters, a.rk_b the RK weight of the base state.
template<class S>
int launch_rk2_fused(const typename S::Params& p1, const typename S::Params& p2, const MeshLoader<S>& ld,
                     const StageArgs& a, hipStream_t s);

// One 1D work-list stage; cus = compute units of the device (grid sizing).
template<typename R>
int launch_stage1d(const EulerParams& prm, const Stage1DArgs& a, hipStream_t s, int cus);

// One fused RK2 step (both stages in one launch, mixed precision) over the
// 1D work-list: a.u the state, a.out the result, a.rk_b the RK weight of the
// base state; blocks of 16, 32, 64, 128 or a multiple of 256 zones.
int launch_rk2_fused_1d(const EulerParams& prm, const Stage1DArgs& a, hipStream_t s);

// One RK stage of MHD constrained transport; work = workspace of
// mhd_workspace_bytes(N, precision) bytes.
long mhd_workspace_bytes(int N, int precision);
template<typename R>
int mhd_advance(MhdArgs a, char* work, hipStream_t s);

} // namespace m3p
