// Harness around the reference's own hot-path functions, compiled from the
// reference sources where they lie (see oracle/Makefile).  TEST
// INFRASTRUCTURE ONLY: it exports C entry points over plain double arrays in
// the reference's logical component order so tests can compare the oracle
// restatement (and, through it, the HIP kernels) with the reference.
//
// Nothing here re-implements reference logic: each function converts plain
// arrays to the reference types, calls one reference function, and converts
// back.
#include <stdexcept>
#include <memory>
#include <string>
#include <cstring>
#include "physics_euler.hpp"
#include "physics_iso2d.hpp"
#include "physics_mhd.hpp"
#include "physics_srhd.hpp"
#include "physics_two_body.hpp"
#include "scheme_plm_gradient.hpp"
#include "core_ndarray.hpp"
#include "mesh_block_index.hpp"
#include "minidisk.hpp"

namespace {

euler::primitive_t euler_prim(const double* p)
{
    return euler::primitive(p[0], p[1], p[2], p[3], p[4]);
}

template<typename Tuple, std::size_t... Is>
void tuple_to_array(const Tuple& t, double* out, std::index_sequence<Is...>)
{
    ((out[Is] = numeric::get<Is>(t).value), ...);
}

template<typename... Ts>
void to_array(const numeric::tuple_t<Ts...>& t, double* out)
{
    tuple_to_array(t, out, std::make_index_sequence<sizeof...(Ts)>());
}

iso2d::primitive_t iso_prim(const double* p) { return iso2d::primitive(p[0], p[1], p[2]); }

iso2d::conserved_density_t iso_cons(const double* u)
{
    return iso2d::conserved_density_t{{u[0], u[1], u[2]}};
}

minidisk::solver_data_t solver_data(const double* s)
{
    // s = {block_size, depth, omega_frame, mach, nu, domain_radius, softening,
    //      eccentricity, mass_ratio, sink_radius, sink_rate, buffer_rate, buffer_scale, cfl}
    return minidisk::solver_data_t{
        int(s[0]), int(s[1]), s[2], s[3], s[4], s[5], s[6], s[7], s[8], s[9], s[10], s[11], s[12], s[13]};
}

} // namespace

extern "C" {

double ref_plm(double yl, double y0, double yr, double theta)
{
    auto z = [] (double y) { return euler::primitive(y, 0.0, 0.0, 0.0, 0.0); };
    auto g = mara::plm_gradient(theta)(std::tuple(z(yl), z(y0), z(yr)));
    return numeric::get<0>(g).value;
}

int ref_euler_recover_primitive(const double* u, double gamma, double* p)
{
    try {
        auto uu = euler::conserved_density_t{{u[0], u[1], u[2], u[3], u[4]}};
        to_array(euler::recover_primitive(uu, gamma), p);
        return 0;
    } catch (const std::invalid_argument&) {
        return -1;
    }
}

void ref_euler_conserved_density(const double* p, double gamma, double* u)
{
    to_array(euler::conserved_density(euler_prim(p), gamma), u);
}

void ref_euler_flux(const double* p, int axis, double gamma, double* f)
{
    to_array(euler::flux(euler_prim(p), geometric::unit_vector_on(axis), gamma), f);
}

void ref_euler_riemann_hlle(const double* pl, const double* pr, int axis, double gamma, double* f)
{
    to_array(euler::riemann_hlle(euler_prim(pl), euler_prim(pr), geometric::unit_vector_on(axis), gamma), f);
}

void ref_iso2d_recover_primitive(const double* u, double* p)
{
    to_array(iso2d::recover_primitive(iso_cons(u)), p);
}

void ref_iso2d_riemann_hlle(const double* pl, const double* pr, double cs2, int axis, double* f)
{
    to_array(iso2d::riemann_hlle(iso_prim(pl), iso_prim(pr), cs2, geometric::unit_vector_on(axis)), f);
}

int ref_two_body_state(double a, double M, double q, double e, double t, double omega_frame, double* st)
{
    auto elements = two_body::orbital_elements(a, M, q, e);
    auto inertial = two_body::orbital_state(elements, t);
    auto s = two_body::rotate(inertial, -dimensional::unit_rate(omega_frame) * dimensional::unit_time(t));
    const two_body::point_mass_t* c[2] = {&s.first, &s.second};
    for (int k = 0; k < 2; ++k) {
        st[5 * k + 0] = c[k]->mass.value;
        st[5 * k + 1] = c[k]->position_x.value;
        st[5 * k + 2] = c[k]->position_y.value;
        st[5 * k + 3] = c[k]->velocity_x.value;
        st[5 * k + 4] = c[k]->velocity_y.value;
    }
    return 0;
}

void ref_two_body_acceleration(double mass, double px, double py, double x, double y, double rs, double* a)
{
    auto p = two_body::point_mass_t{mass, px, py, 0.0, 0.0};
    auto g = two_body::gravitational_acceleration(p, x, y, rs);
    a[0] = g[0].value;
    a[1] = g[1].value;
}

// Recover orbital elements (a, M, q, e) from a state, as the reference KAT does.
void ref_two_body_recover_elements(const double* st, double t, double* el)
{
    auto c1 = two_body::point_mass_t{st[0], st[1], st[2], st[3], st[4]};
    auto c2 = two_body::point_mass_t{st[5], st[6], st[7], st[8], st[9]};
    auto r = two_body::recover_orbital_elements(two_body::orbital_state_t{c1, c2}, t);
    auto e = numeric::get<0>(r);
    el[0] = numeric::get<0>(e).value;
    el[1] = numeric::get<1>(e).value;
    el[2] = double(numeric::get<2>(e));
    el[3] = double(numeric::get<3>(e));
}

// minidisk::updated_conserved (problems/minidisk_scheme.cpp:294-355) on one block.
// uc: [n][n][3] conserved densities, pe: [n+4][n+4][3] primitives, out: [n][n][3].
void ref_minidisk_updated_conserved(const double* uc, const double* pe, const double* sd,
                                    int level, long bi, long bj, double time, double dt, double* out)
{
    auto data = solver_data(sd);
    const unsigned long n = data.block_size;
    auto ucarr = nd::make_array(nd::indexing([uc, n] (auto i, auto j)
    {
        return iso_cons(&uc[(i * n + j) * 3]);
    }), nd::uivec(n, n)) | nd::to_shared();
    auto pearr = nd::make_array(nd::indexing([pe, n] (auto i, auto j)
    {
        return iso_prim(&pe[(i * (n + 4) + j) * 3]);
    }), nd::uivec(n + 4, n + 4)) | nd::to_shared();
    auto block = mesh::block_index_t<2>{(unsigned long)level, numeric::array((unsigned long)bi, (unsigned long)bj)};
    auto result = minidisk::updated_conserved(ucarr, pearr, time, dt, block, data);
    for (unsigned long i = 0; i < n; ++i)
        for (unsigned long j = 0; j < n; ++j)
            to_array(result(i, j), &out[(i * n + j) * 3]);
}

void ref_minidisk_initial_conserved(const double* sd, int level, long bi, long bj, double* out)
{
    auto data = solver_data(sd);
    const unsigned long n = data.block_size;
    auto block = mesh::block_index_t<2>{(unsigned long)level, numeric::array((unsigned long)bi, (unsigned long)bj)};
    auto u = minidisk::initial_conserved_array(data, block);
    for (unsigned long i = 0; i < n; ++i)
        for (unsigned long j = 0; j < n; ++j)
            to_array(u(i, j), &out[(i * n + j) * 3]);
}

} // extern "C"
