// Stage kernels of IsoTableScheme<float> (one unit per scheme and precision, compiled in parallel).
#include "launch_impl.hpp"

M3P_INSTANTIATE_STAGE2D(m3p::IsoTableScheme<float>, m3p::MeshLoader)
M3P_INSTANTIATE_STAGE2D(m3p::IsoTableScheme<float>, m3p::ExtendedLoader)
M3P_INSTANTIATE_VISC(m3p::IsoTableScheme<float>, m3p::MeshLoader)
M3P_INSTANTIATE_VISC(m3p::IsoTableScheme<float>, m3p::ExtendedLoader)
