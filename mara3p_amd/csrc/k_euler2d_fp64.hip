// Stage kernels of EulerScheme<double> (one unit per scheme and precision, compiled in parallel).
#include "launch_impl.hpp"

M3P_INSTANTIATE_STAGE2D(m3p::EulerScheme<double>, m3p::MeshLoader)
M3P_INSTANTIATE_STAGE2D(m3p::EulerScheme<double>, m3p::ExtendedLoader)
M3P_INSTANTIATE_STAGE2D(m3p::EulerScheme<double>, m3p::TreeLoader)
