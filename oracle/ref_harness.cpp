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
#include "scheme_mhd_v1.hpp"
#include "parallel_thread_pool.hpp"

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

// ---------------------------------------------------------------------------
// A uniform periodic minidisk mesh held in the reference's own types, stepped
// by the reference's functions on the reference's thread pool (the CPU
// baseline).  One stage has the task structure of the reference's substep
// (problems/minidisk.cpp:172-189): per block "P" =
// minidisk::recover_primitive_array, then per block "Pe" (the extended
// primitive array from the nine neighbour blocks, problems/minidisk.cpp:87-120)
// followed by "U" = minidisk::updated_conserved.  RK2 combines per block as
// amr::weighted_sum_tree (src/mesh_amr.hpp:455-467).  Every task is one
// mara::ThreadPool job (src/parallel_thread_pool.hpp:121-137); the calling
// thread waits on the futures.
struct ref_minidisk_mesh {
    minidisk::solver_data_t data;
    unsigned long nb = 0;
    std::vector<minidisk::conserved_array_t> u;
    std::unique_ptr<mara::ThreadPool> pool;
};

namespace {

minidisk::primitive_array_t extend_from(const std::vector<minidisk::primitive_array_t>& P, unsigned long nb,
                                        unsigned long bi, unsigned long bj)
{
    auto wrap = [nb] (unsigned long c, int d) { return (c + nb + d) % nb; };
    // the corner blocks as problems/minidisk.cpp:90-93 picks them
    std::array<std::array<minidisk::primitive_array_t, 3>, 3> cs;
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) cs[a][c] = P[wrap(bi, a - 1) * nb + wrap(bj, c - 1)];
    cs[0][2] = P[wrap(bi, 1) * nb + wrap(bj, -1)];
    cs[2][0] = P[wrap(bi, -1) * nb + wrap(bj, 1)];
    const auto nx = shape(cs[1][1], 0), ny = shape(cs[1][1], 1);
    return nd::make_array(nd::indexing([cs, nx, ny] (auto i, auto j)
    {
        auto ci = i < 2 ? 0 : (i >= nx + 2 ? 2 : 1);
        auto cj = j < 2 ? 0 : (j >= ny + 2 ? 2 : 1);
        auto ii = i < 2 ? i - 2 + nx : (i >= nx + 2 ? i - 2 - nx : i - 2);
        auto jj = j < 2 ? j - 2 + ny : (j >= ny + 2 ? j - 2 - ny : j - 2);
        return cs[ci][cj](ii, jj);
    }), nd::uivec(nx + 4, ny + 4)) | nd::to_shared();
}

std::vector<minidisk::conserved_array_t> mesh_stage(ref_minidisk_mesh& m, const std::vector<minidisk::conserved_array_t>& u,
                                                    double time, double dt)
{
    const unsigned long nb = m.nb, nblk = nb * nb;
    std::vector<std::future<minidisk::primitive_array_t>> fp;
    for (unsigned long b = 0; b < nblk; ++b)
        fp.push_back(m.pool->enqueue(0, [&u, b] { return minidisk::recover_primitive_array(u[b]); }));
    std::vector<minidisk::primitive_array_t> P;
    for (auto& f : fp) P.push_back(f.get());
    std::vector<std::future<minidisk::conserved_array_t>> fu;
    for (unsigned long b = 0; b < nblk; ++b)
        fu.push_back(m.pool->enqueue(0, [&m, &u, &P, b, nb, time, dt] {
            auto pe = extend_from(P, nb, b / nb, b % nb);
            auto block = mesh::block_index_t<2>{ (unsigned long)m.data.depth, numeric::array(b / nb, b % nb) };
            return minidisk::updated_conserved(u[b], pe, time, dt, block, m.data);
        }));
    std::vector<minidisk::conserved_array_t> out;
    for (auto& f : fu) out.push_back(f.get());
    return out;
}

} // namespace

extern "C" {

// u: [nb*nb][n][n][3] conserved, block b = bi * nb + bj at the solver's depth.
ref_minidisk_mesh* ref_minidisk_mesh_create(const double* u, const double* sd, int nthreads)
{
    auto m = new ref_minidisk_mesh;
    m->data = solver_data(sd);
    m->nb = 1ul << m->data.depth;
    const unsigned long n = m->data.block_size;
    for (unsigned long b = 0; b < m->nb * m->nb; ++b) {
        const double* ub = u + b * n * n * 3;
        m->u.push_back(nd::make_array(nd::indexing([ub, n] (auto i, auto j)
        {
            return iso_cons(&ub[(i * n + j) * 3]);
        }), nd::uivec(n, n)) | nd::to_shared());
    }
    m->pool = std::make_unique<mara::ThreadPool>(nthreads);
    return m;
}

// RK2 steps as control::advance_runge_kutta (src/app_control.hpp:169-195).
void ref_minidisk_mesh_rk2(ref_minidisk_mesh* m, double time, double dt, int steps)
{
    for (int s = 0; s < steps; ++s) {
        auto u1 = mesh_stage(*m, m->u, time, dt);
        auto u2 = mesh_stage(*m, u1, time + dt, dt);
        std::vector<std::future<minidisk::conserved_array_t>> fw;
        for (unsigned long b = 0; b < m->u.size(); ++b)
            fw.push_back(m->pool->enqueue(0, [m, &u2, b] {
                return (m->u[b] * 0.5 + u2[b] * (1.0 - 0.5)) | nd::to_shared();
            }));
        for (unsigned long b = 0; b < m->u.size(); ++b) m->u[b] = fw[b].get();
        time = time * 0.5 + (time + dt) * (1.0 - 0.5);
    }
}

void ref_minidisk_mesh_get(const ref_minidisk_mesh* m, double* out)
{
    const unsigned long n = m->data.block_size;
    for (unsigned long b = 0; b < m->u.size(); ++b)
        for (unsigned long i = 0; i < n; ++i)
            for (unsigned long j = 0; j < n; ++j) to_array(m->u[b](i, j), &out[((b * n + i) * n + j) * 3]);
}

void ref_minidisk_mesh_destroy(ref_minidisk_mesh* m) { delete m; }

} // extern "C"

// ---------------------------------------------------------------------------
// MHD: src/physics_mhd.hpp and mara::advance_mhd_ct (src/scheme_mhd_v1.cpp:265-299)
static mhd::primitive_t mhd_prim(const double* p)
{
    return mhd::primitive(p[0], {p[1], p[2], p[3]}, p[4], {p[5], p[6], p[7]});
}

int ref_mhd_recover_primitive(const double* u, const double* b, double gamma, double* p)
{
    try {
        auto uu = mhd::conserved_density_t{{u[0], u[1], u[2], u[3], u[4]}};
        auto bb = mhd::magnetic_field_vector_t{b[0], b[1], b[2]};
        to_array(mhd::recover_primitive(uu, bb, gamma), p);
        return 0;
    } catch (const std::invalid_argument& e) {
        return std::string(e.what()).find("density") != std::string::npos ? 1 : 2;
    }
}

void ref_mhd_conserved_density(const double* p, double gamma, double* u)
{
    to_array(mhd::conserved_density(mhd_prim(p), gamma), u);
}

void ref_mhd_flux(const double* p, int axis, double gamma, double* f)
{
    to_array(mhd::flux(mhd_prim(p), geometric::unit_vector_on(axis), gamma), f);
}

double ref_mhd_fastest_speed(const double* p, double gamma)
{
    return mhd::fastest_speed(mhd_prim(p), gamma).value;
}

void ref_mhd_riemann_hlle(const double* pl, const double* pr, double bl, int axis, double gamma, double* f, double* e)
{
    auto [flux, ef] = mhd::riemann_hlle(mhd_prim(pl), mhd_prim(pr), bl, geometric::unit_vector_on(axis), gamma);
    to_array(flux, f);
    e[0] = ef.component_1().value;
    e[1] = ef.component_2().value;
    e[2] = ef.component_3().value;
}

// One call of advance_mhd_ct with the periodic boundary extension.  Arrays in
// logical order: uc [N][N][N][5], bf1 [N+1][N][N], bf2 [N][N+1][N], bf3 [N][N][N+1].
int ref_mhd_advance_ct(int N, const double* uc, const double* bf1, const double* bf2, const double* bf3,
                       double dl, double cfl, double plm, double* uo, double* bo1, double* bo2, double* bo3, double* dt)
{
    using ulong = unsigned long;
    const ulong n = N;
    auto ua = nd::make_array(nd::indexing([uc, n] (auto i, auto j, auto k)
    {
        const double* z = &uc[((i * n + j) * n + k) * 5];
        return mhd::conserved_density_t{{z[0], z[1], z[2], z[3], z[4]}};
    }), nd::uivec(n, n, n)) | nd::to_shared();
    auto face = [] (const double* src, ulong a, ulong b, ulong c)
    {
        return nd::make_array(nd::indexing([src, b, c] (auto i, auto j, auto k)
        {
            return mhd::unit_magnetic_field(src[(i * b + j) * c + k]);
        }), nd::uivec(a, b, c)) | nd::to_shared();
    };
    auto b1 = face(bf1, n + 1, n, n);
    auto b2 = face(bf2, n, n + 1, n);
    auto b3 = face(bf3, n, n, n + 1);
    try {
        auto bc = mara::local_periodic_boundary_extension();
        auto [t, u_, b1_, b2_, b3_] = mara::advance_mhd_ct(0.0, ua, b1, b2, b3, dl, cfl, plm, *bc);
        *dt = t.value;
        for (ulong i = 0; i < n; ++i)
            for (ulong j = 0; j < n; ++j)
                for (ulong k = 0; k < n; ++k) to_array(u_(i, j, k), &uo[((i * n + j) * n + k) * 5]);
        auto out = [] (auto arr, double* dst, ulong a, ulong b, ulong c)
        {
            for (ulong i = 0; i < a; ++i)
                for (ulong j = 0; j < b; ++j)
                    for (ulong k = 0; k < c; ++k) dst[(i * b + j) * c + k] = arr(i, j, k).value;
        };
        out(b1_, bo1, n + 1, n, n);
        out(b2_, bo2, n, n + 1, n);
        out(b3_, bo3, n, n, n + 1);
        return 0;
    } catch (const std::invalid_argument& e) {
        return std::string(e.what()).find("density") != std::string::npos ? 1 : 2;
    }
}

// ---------------------------------------------------------------------------
// SRHD: src/physics_srhd.hpp
static srhd::primitive_t srhd_prim(const double* p) { return srhd::primitive(p[0], p[1], p[2], p[3], p[4]); }

int ref_srhd_recover_primitive(const double* u, double gamma, double temperature_floor, double* p)
{
    try {
        auto uu = srhd::conserved_density_t{{u[0], u[1], u[2], u[3], u[4]}};
        to_array(srhd::recover_primitive(uu, gamma, temperature_floor), p);
        return 0;
    } catch (const std::invalid_argument& e) {
        return std::string(e.what()).find("converging") != std::string::npos ? 3 : 2;
    }
}

void ref_srhd_conserved_density(const double* p, double gamma, double* u)
{
    to_array(srhd::conserved_density(srhd_prim(p), gamma), u);
}

int ref_srhd_riemann(const double* pl, const double* pr, int axis, double gamma, int mode, double face_speed, double* f)
{
    auto n = geometric::unit_vector_on(axis);
    try {
        switch (mode) {
            case 0: to_array(srhd::riemann_hlle(srhd_prim(pl), srhd_prim(pr), n, gamma), f); break;
            case 1: to_array(srhd::riemann_hllc(srhd_prim(pl), srhd_prim(pr), n, gamma), f); break;
            case 2: f[0] = srhd::riemann_solver(srhd_prim(pl), srhd_prim(pr), n, gamma,
                                                srhd::riemann_solver_mode_contact_speed_t()).value; break;
            case 3: to_array(srhd::riemann_hllc(srhd_prim(pl), srhd_prim(pr), n, face_speed, gamma), f); break;
        }
        return 0;
    } catch (const std::invalid_argument&) {
        return 3;
    }
}

} // extern "C"
