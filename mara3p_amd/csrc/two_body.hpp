// Binary orbit for the minidisk sources, evaluated once per stage: on the
// host for the stage calls that take a host time, on the device (one thread)
// for the clocked steps that a hipGraph can capture.  Restates
// two_body::orbital_state + rotate (src/physics_two_body.hpp:81-157,
// 377-421) as called at problems/minidisk_scheme.cpp:329-335 (G = 1).  The
// host build calls the C library's sqrt, sin and cos, the device build the
// device library's: sqrt is correctly rounded in both, sin and cos may differ
// in the last bit.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>

namespace m3p {

// st = {M1, x1, y1, vx1, vy1, M2, x2, y2, vx2, vy2}; returns Newton iterations.
__host__ __device__ inline int two_body_state(double a, double M, double q, double e, double t, double omega_frame,
                                              double* st)
{
    double W = sqrt(1.0 * M / a / a / a);
    double Man = W * t;
    double E = Man;
    int iters = 0;
    // The reference iterates until |f| <= 1e-12 with no bound; 200 Newton
    // steps is far beyond what any bound orbit (e < 1) needs.
    while (fabs(E - e * sin(E) - Man) > 1e-12 && iters < 200) {
        E -= (E - e * sin(E) - Man) / (1 - e * cos(E));
        ++iters;
    }
    double mu = q / (1.0 + q);
    double M1 = M * (1.0 - mu);
    double M2 = M * mu;
    double x1 = -a * mu * (e - cos(E));
    double y1 = +a * mu * (0 + sin(E)) * sqrt(1 - e * e);
    double x2 = -x1 / q;
    double y2 = -y1 / q;
    double vx1 = -a * mu * W / (1 - e * cos(E)) * sin(E);
    double vy1 = +a * mu * W / (1 - e * cos(E)) * cos(E) * sqrt(1 - e * e);
    double vx2 = -vx1 / q;
    double vy2 = -vy1 / q;
    double angle = -omega_frame * t;
    double c = cos(-angle);
    double s = sin(-angle);
    st[0] = M1;
    st[1] = +x1 * c + y1 * s + 0.0;
    st[2] = -x1 * s + y1 * c + 0.0;
    st[3] = +vx1 * c + vy1 * s + 0.0;
    st[4] = -vx1 * s + vy1 * c + 0.0;
    st[5] = M2;
    st[6] = +x2 * c + y2 * s + 0.0;
    st[7] = -x2 * s + y2 * c + 0.0;
    st[8] = +vx2 * c + vy2 * s + 0.0;
    st[9] = -vx2 * s + vy2 * c + 0.0;
    return iters;
}

} // namespace m3p
