// Element-wise block operations behind the reference's module helpers:
//   recover_primitive_array  src/module_euler.cpp:84-89,240-245,
//                            src/module_locally_isothermal.cpp:86-91
//   estimate_gradient        src/module_euler.cpp:95-101,251-257 (PLM along an axis)
//   weighted_sum             src/mesh_amr.hpp:455-467 (s*b + t*(1-b))
// All arrays SoA: component-major, zones row-major within a component.
#pragma once
#include "physics.hpp"

namespace m3p {

__global__ void __launch_bounds__(256)
euler_recover_array_kernel(const double* u, double* p, long nz, double gamma, int* err)
{
    long z = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= nz) return;
    double uu[5], pp[5];
    for (int q = 0; q < 5; ++q) uu[q] = u[q * nz + z];
    EulerParams prm{ gamma, 0.0, 0 };
    bool ok = Euler<double>::recover(uu, prm, pp);
    for (int q = 0; q < 5; ++q) p[q * nz + z] = pp[q];
    if (!ok) report_error(err, 1, 0, (int)z);
}

__global__ void __launch_bounds__(256)
iso2d_recover_array_kernel(const double* u, double* p, long nz)
{
    long z = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= nz) return;
    double uu[3] = { u[z], u[nz + z], u[2 * nz + z] }, pp[3];
    Iso2d<double>::recover(uu, pp);
    for (int q = 0; q < 3; ++q) p[q * nz + z] = pp[q];
}

// g[q][i][j] = plm(p[i-1], p[i], p[i+1]) along `axis` of an ni x nj component
// array; output shape (ni-2) x nj (axis 0) or ni x (nj-2) (axis 1), exactly
// nd::adjacent_zip3(axis) | plm_gradient (src/core_ndarray_ops.hpp:68-74).
__global__ void __launch_bounds__(256)
plm_gradient_2d_kernel(const double* p, double* g, int nc, int ni, int nj, int axis, double theta)
{
    const int oi = axis == 0 ? ni - 2 : ni, oj = axis == 1 ? nj - 2 : nj;
    const long nout = (long)oi * oj;
    long z = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= nout * nc) return;
    const int q = (int)(z / nout);
    const long r = z % nout;
    const int i = (int)(r / oj), j = (int)(r % oj);
    const double* c = p + (size_t)q * ni * nj;
    double yl, y0, yr;
    if (axis == 0) { yl = c[(size_t)i * nj + j]; y0 = c[(size_t)(i + 1) * nj + j]; yr = c[(size_t)(i + 2) * nj + j]; }
    else           { yl = c[(size_t)i * nj + j]; y0 = c[(size_t)i * nj + j + 1]; yr = c[(size_t)i * nj + j + 2]; }
    g[z] = plm(yl, y0, yr, theta);
}

__global__ void __launch_bounds__(256)
weighted_sum_kernel(const double* s, const double* t, double* out, long n, double b)
{
    long z = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (z < n) out[z] = s[z] * b + t[z] * (1.0 - b);
}

} // namespace m3p
