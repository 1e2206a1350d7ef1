// 1D work-list stage kernels (both precisions).
#include "launch_impl.hpp"

template int m3p::launch_stage1d<double>(const m3p::EulerParams&, const m3p::Stage1DArgs&, hipStream_t, int);
template int m3p::launch_stage1d<float>(const m3p::EulerParams&, const m3p::Stage1DArgs&, hipStream_t, int);
