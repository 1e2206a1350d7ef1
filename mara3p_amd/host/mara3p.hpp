// C++17 host API with the reference's block-update entry points, over the
// C ABI in include/mara3p_amd.h.  Header-only; link with libmara3p_amd.so.
//
// Names, argument meaning and error behaviour follow the reference:
//   modules::euler1d / euler2d      src/module_euler.hpp:77-125,161-207
//   modules::locally_isothermal     src/module_locally_isothermal.hpp:73-117
//   minidisk                        problems/minidisk.hpp:148-170
//   mhd_scheme_v2                   src/scheme_mhd_v2.hpp:50-210
//   mesh::block_index_t, next_on / prev_on   src/mesh_block_index.hpp:53-57,334-335
//   mesh::cartesian_{1,2}d::geometry_t       src/mesh_geometry.hpp:53-101
// Dimensional quantities (unit_time, unit_length, ...) are plain doubles
// here; per-zone tuples are std::array<double, n_c> in logical component
// order (numeric::get<I>), and block arrays are row-major with the last axis
// fastest, as nd::shared_array.  Errors the reference throws as
// std::invalid_argument are thrown with the reference's message.
//
// Each per-block call stages its arrays to the device and back (the
// compatibility path).  Device-resident meshes use the whole-mesh stage entry
// points of the C ABI directly (m3p_euler2d_stage, m3p_minidisk_stage, ...).
#pragma once
#include <hip/hip_runtime.h>
#include <array>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#include "mara3p_amd.h"

namespace mara3p {

// ---------------------------------------------------------------------------
// Errors
inline void check(int status)
{
    if (status == M3P_OK) return;
    std::string msg = m3p_last_error();
    if (status == M3P_ERR_NEGATIVE_DENSITY || status == M3P_ERR_NEGATIVE_PRESSURE ||
        status == M3P_ERR_NO_CONVERGENCE || status == M3P_ERR_BAD_SHAPE || status == M3P_ERR_BAD_ARGUMENT)
        throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

inline void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// ---------------------------------------------------------------------------
// nd::shared_array-compatible block container (src/core_ndarray.hpp:187-266):
// shared immutable buffer, row-major, last axis fastest.
namespace nd {

template<unsigned long Rank>
using uivec_t = std::array<std::size_t, Rank>;

template<typename... Args>
auto uivec(Args... args) { return uivec_t<sizeof...(Args)>{ std::size_t(args)... }; }

template<typename T, unsigned long Rank>
class shared_array
{
public:
    using value_type = T;
    shared_array() = default;
    explicit shared_array(uivec_t<Rank> shape)
    : shape_(shape), data_(std::make_shared<std::vector<T>>(count(shape))) {}
    shared_array(uivec_t<Rank> shape, std::vector<T> values)
    : shape_(shape), data_(std::make_shared<std::vector<T>>(std::move(values)))
    {
        if (data_->size() != count(shape)) throw std::invalid_argument("nd::shared_array (size does not match shape)");
    }
    template<typename... Is>
    const T& operator()(Is... is) const { return (*data_)[offset(uivec_t<Rank>{ std::size_t(is)... })]; }
    template<typename... Is>
    T& operator()(Is... is) { return (*data_)[offset(uivec_t<Rank>{ std::size_t(is)... })]; }
    const T* data() const { return data_->data(); }
    T* data() { return data_->data(); }
    uivec_t<Rank> shape() const { return shape_; }
    std::size_t shape(unsigned long axis) const { return shape_[axis]; }
    std::size_t size() const { return data_ ? data_->size() : 0; }

private:
    static std::size_t count(uivec_t<Rank> s) { std::size_t n = 1; for (auto x : s) n *= x; return n; }
    std::size_t offset(uivec_t<Rank> i) const
    {
        std::size_t o = 0;
        for (unsigned long a = 0; a < Rank; ++a) o = o * shape_[a] + i[a];
        return o;
    }
    uivec_t<Rank> shape_{};
    std::shared_ptr<std::vector<T>> data_;
};

template<typename T, unsigned long R> std::size_t shape(const shared_array<T, R>& a, unsigned long axis) { return a.shape(axis); }
template<typename T, unsigned long R> std::size_t size(const shared_array<T, R>& a) { return a.size(); }

} // namespace nd

// ---------------------------------------------------------------------------
// mesh::block_index_t and the periodic next_on / prev_on
namespace mesh {

template<unsigned long Rank>
struct block_index_t
{
    unsigned long level = 0;
    std::array<unsigned long, Rank> coordinates = {};
};

template<unsigned long R>
block_index_t<R> next_on(block_index_t<R> i, std::size_t a)
{
    i.coordinates[a] = (i.coordinates[a] + (1ul << i.level) + 1) % (1ul << i.level);
    return i;
}

template<unsigned long R>
block_index_t<R> prev_on(block_index_t<R> i, std::size_t a)
{
    i.coordinates[a] = (i.coordinates[a] + (1ul << i.level) - 1) % (1ul << i.level);
    return i;
}

namespace cartesian_1d {
struct geometry_t
{
    double domain_size = 1.0;
    int block_size = 64;
    double cell_spacing(block_index_t<1> b) const { return domain_size / double(block_size) / double(1 << b.level); }
};
}

namespace cartesian_2d {
struct geometry_t
{
    double domain_size = 1.0;
    int block_size = 64;
    double cell_spacing(block_index_t<2> b) const { return domain_size / double(block_size) / double(1 << b.level); }
    // face_coordinates, src/mesh_geometry.cpp:125-146: (x, y) of the axis-oriented faces
    std::array<double, 2> face_coordinate(block_index_t<2> b, unsigned long axis, std::size_t i, std::size_t j) const
    {
        double dl = double(1 << b.level);
        auto vert = [&] (unsigned long c, std::size_t k) {
            double x0 = domain_size * (-0.5 + 1.0 * c / dl);
            double x1 = domain_size * (-0.5 + 1.0 * (c + 1) / dl);
            return x0 + (x1 - x0) * double(k) / double(block_size);
        };
        auto cent = [&] (unsigned long c, std::size_t k) { return 0.5 * (vert(c, k) + vert(c, k + 1)); };
        if (axis == 0) return { vert(b.coordinates[0], i), cent(b.coordinates[1], j) };
        return { cent(b.coordinates[0], i), vert(b.coordinates[1], j) };
    }
};
}

} // namespace mesh

// ---------------------------------------------------------------------------
// Device staging helpers
struct device_buffer
{
    double* ptr = nullptr;
    explicit device_buffer(std::size_t count)
    {
        hip_check(hipMalloc(&ptr, count * sizeof(double)), "hipMalloc");
    }
    ~device_buffer() { if (ptr) (void)hipFree(ptr); }
    device_buffer(const device_buffer&) = delete;
    device_buffer& operator=(const device_buffer&) = delete;
};

// Logical tuples [zones][NC] (host) -> SoA [NC][zones] (device), component by
// component, never a memcpy of the tuple layout.
template<std::size_t NC>
inline void upload_soa(const std::array<double, NC>* src, std::size_t zones, double* dst)
{
    std::vector<double> tmp(NC * zones);
    for (std::size_t z = 0; z < zones; ++z)
        for (std::size_t q = 0; q < NC; ++q) tmp[q * zones + z] = src[z][q];
    hip_check(hipMemcpy(dst, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy H2D");
}

template<std::size_t NC>
inline void download_soa(const double* src, std::size_t zones, std::array<double, NC>* dst)
{
    std::vector<double> tmp(NC * zones);
    hip_check(hipMemcpy(tmp.data(), src, tmp.size() * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    for (std::size_t z = 0; z < zones; ++z)
        for (std::size_t q = 0; q < NC; ++q) dst[z][q] = tmp[q * zones + z];
}

// ---------------------------------------------------------------------------
namespace euler {
using primitive_t = std::array<double, 5>;           // rho, v1, v2, v3, p
using conserved_density_t = std::array<double, 5>;   // D, S1, S2, S3, E
}

namespace iso2d {
using primitive_t = std::array<double, 3>;           // Sigma, vx, vy
using conserved_density_t = std::array<double, 3>;   // Sigma, Sx, Sy
}

namespace modules {

// --------------------------------------------------------------------------- euler2d
namespace euler2d {
using conserved_array_t = nd::shared_array<euler::conserved_density_t, 2>;
using primitive_array_t = nd::shared_array<euler::primitive_t, 2>;
using mesh_geometry_t   = mesh::cartesian_2d::geometry_t;

// src/module_euler.cpp:240-245
inline primitive_array_t recover_primitive_array(conserved_array_t uc, double gamma_law_index)
{
    const std::size_t nz = uc.size();
    device_buffer du(5 * nz), dp(5 * nz);
    upload_soa(uc.data(), nz, du.ptr);
    check(m3p_euler_recover_primitive_array(du.ptr, dp.ptr, long(nz), gamma_law_index, nullptr));
    primitive_array_t p(uc.shape());
    download_soa(dp.ptr, nz, p.data());
    return p;
}

// src/module_euler.cpp:251-257
inline primitive_array_t estimate_gradient(primitive_array_t pc, unsigned long axis, double plm_theta)
{
    const std::size_t ni = pc.shape(0), nj = pc.shape(1), nz = pc.size();
    if (axis > 1 || (axis == 0 ? ni : nj) < 3) throw std::invalid_argument("nd::select (selection out of range)");
    auto oshape = nd::uivec(axis == 0 ? ni - 2 : ni, axis == 1 ? nj - 2 : nj);
    const std::size_t no = oshape[0] * oshape[1];
    device_buffer dp(5 * nz), dg(5 * no);
    upload_soa(pc.data(), nz, dp.ptr);
    check(m3p_plm_gradient_2d(dp.ptr, dg.ptr, 5, int(ni), int(nj), int(axis), plm_theta, nullptr));
    primitive_array_t g(oshape);
    download_soa(dg.ptr, no, g.data());
    return g;
}

// src/module_euler.cpp:263-298 (declared src/module_euler.hpp:182-190)
inline conserved_array_t updated_conserved(conserved_array_t uc, primitive_array_t pe, double /*time*/, double dt,
                                           mesh_geometry_t mesh_geometry, mesh::block_index_t<2> block,
                                           double plm_theta, double gamma_law_index)
{
    const std::size_t n = uc.shape(0);
    if (uc.shape(1) != n || pe.shape(0) != n + 4 || pe.shape(1) != n + 4)
        throw std::invalid_argument("euler2d::updated_conserved (pe must be uc extended by 2 guard zones)");
    device_buffer du(5 * n * n), dp(5 * (n + 4) * (n + 4)), dout(5 * n * n);
    upload_soa(uc.data(), n * n, du.ptr);
    upload_soa(pe.data(), (n + 4) * (n + 4), dp.ptr);
    check(m3p_euler2d_updated_conserved(du.ptr, dp.ptr, dout.ptr, int(n), dt, mesh_geometry.cell_spacing(block),
                                        plm_theta, gamma_law_index, 0, M3P_FP64, nullptr));
    conserved_array_t out(uc.shape());
    download_soa(dout.ptr, n * n, out.data());
    return out;
}

// weighted_sum over blocks (src/module_euler.cpp:353-360 -> src/mesh_amr.hpp:455-467)
inline conserved_array_t weighted_sum(conserved_array_t s, conserved_array_t t, double b)
{
    const std::size_t nz = s.size();
    if (t.size() != nz) throw std::invalid_argument("euler2d::weighted_sum (shape mismatch)");
    device_buffer ds(5 * nz), dt_(5 * nz), dout(5 * nz);
    upload_soa(s.data(), nz, ds.ptr);
    upload_soa(t.data(), nz, dt_.ptr);
    check(m3p_weighted_sum(ds.ptr, dt_.ptr, dout.ptr, long(5 * nz), b, nullptr));
    conserved_array_t out(s.shape());
    download_soa(dout.ptr, nz, out.data());
    return out;
}
} // namespace euler2d

// --------------------------------------------------------------------------- euler1d
namespace euler1d {
using conserved_array_t = nd::shared_array<euler::conserved_density_t, 1>;
using primitive_array_t = nd::shared_array<euler::primitive_t, 1>;
using mesh_geometry_t   = mesh::cartesian_1d::geometry_t;
// The reference's boundary_condition_t is a std::function applied to pe; the
// outflow condition (src/mara.cpp:123-143) is the one the module supports
// natively here: it is selected by block position, as in the reference.
enum class boundary_condition_t { periodic, outflow };

inline primitive_array_t recover_primitive_array(conserved_array_t uc, double gamma_law_index)
{
    const std::size_t nz = uc.size();
    device_buffer du(5 * nz), dp(5 * nz);
    upload_soa(uc.data(), nz, du.ptr);
    check(m3p_euler_recover_primitive_array(du.ptr, dp.ptr, long(nz), gamma_law_index, nullptr));
    primitive_array_t p(uc.shape());
    download_soa(dp.ptr, nz, p.data());
    return p;
}

// src/module_euler.cpp:107-135 (declared src/module_euler.hpp:98-107)
inline conserved_array_t updated_conserved(conserved_array_t uc, primitive_array_t pe, double /*time*/, double dt,
                                           mesh_geometry_t mesh_geometry, boundary_condition_t bc,
                                           mesh::block_index_t<1> block, double plm_theta, double gamma_law_index)
{
    const std::size_t n = uc.shape(0);
    if (pe.shape(0) != n + 4) throw std::invalid_argument("euler1d::updated_conserved (pe must be uc extended by 2 guard zones)");
    const bool outflow = bc == boundary_condition_t::outflow;
    const int lower = outflow && block.coordinates[0] == 0;
    const int upper = outflow && !lower && block.coordinates[0] == (1ul << block.level) - 1;
    device_buffer du(5 * n), dp(5 * (n + 4)), dout(5 * n);
    upload_soa(uc.data(), n, du.ptr);
    upload_soa(pe.data(), n + 4, dp.ptr);
    check(m3p_euler1d_updated_conserved(du.ptr, dp.ptr, dout.ptr, int(n), dt, mesh_geometry.cell_spacing(block),
                                        plm_theta, gamma_law_index, lower, upper, M3P_FP64, nullptr));
    conserved_array_t out(uc.shape());
    download_soa(dout.ptr, n, out.data());
    return out;
}
} // namespace euler1d

// --------------------------------------------------------------------------- locally_isothermal
namespace locally_isothermal {
using conserved_array_t     = nd::shared_array<iso2d::conserved_density_t, 2>;
using primitive_array_t     = nd::shared_array<iso2d::primitive_t, 2>;
using mesh_geometry_t       = mesh::cartesian_2d::geometry_t;
using coords_t              = std::array<double, 2>;
using temperature_mapping_t = std::function<double(coords_t)>;

inline primitive_array_t recover_primitive_array(conserved_array_t uc)
{
    const std::size_t nz = uc.size();
    device_buffer du(3 * nz), dp(3 * nz);
    upload_soa(uc.data(), nz, du.ptr);
    check(m3p_iso2d_recover_primitive_array(du.ptr, dp.ptr, long(nz), nullptr));
    primitive_array_t p(uc.shape());
    download_soa(dp.ptr, nz, p.data());
    return p;
}

// src/module_locally_isothermal.cpp:182-212.  The temperature mapping is a
// host callback in the reference too; it is evaluated here at the face
// coordinates once per call and handed to the kernel as arrays.  Like the
// reference, plm_theta is ignored (theta = 2 is used, :195-196).
inline conserved_array_t updated_conserved(conserved_array_t uc, primitive_array_t pe, double /*time*/, double dt,
                                           mesh_geometry_t mesh_geometry, mesh::block_index_t<2> block,
                                           double kinematic_viscosity, temperature_mapping_t temperature,
                                           double /*plm_theta*/)
{
    const std::size_t n = uc.shape(0);
    if (pe.shape(0) != n + 4 || pe.shape(1) != n + 4)
        throw std::invalid_argument("locally_isothermal::updated_conserved (pe must be uc extended by 2 guard zones)");
    if (kinematic_viscosity < 0.0)
        throw std::invalid_argument("locally_isothermal::updated_conserved (negative kinematic viscosity)");
    std::vector<double> cx((n + 1) * n), cy(n * (n + 1));
    for (std::size_t i = 0; i <= n; ++i)
        for (std::size_t j = 0; j < n; ++j) cx[i * n + j] = temperature(mesh_geometry.face_coordinate(block, 0, i, j));
    for (std::size_t i = 0; i < n; ++i)
        for (std::size_t j = 0; j <= n; ++j) cy[i * (n + 1) + j] = temperature(mesh_geometry.face_coordinate(block, 1, i, j));
    device_buffer du(3 * n * n), dp(3 * (n + 4) * (n + 4)), dout(3 * n * n), dcx(cx.size()), dcy(cy.size());
    upload_soa(uc.data(), n * n, du.ptr);
    upload_soa(pe.data(), (n + 4) * (n + 4), dp.ptr);
    hip_check(hipMemcpy(dcx.ptr, cx.data(), cx.size() * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy H2D");
    hip_check(hipMemcpy(dcy.ptr, cy.data(), cy.size() * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy H2D");
    check(m3p_locally_isothermal_updated_conserved(du.ptr, dp.ptr, dout.ptr, int(n), dt, mesh_geometry.cell_spacing(block),
                                                   dcx.ptr, dcy.ptr, kinematic_viscosity, M3P_FP64, nullptr));
    conserved_array_t out(uc.shape());
    download_soa(dout.ptr, n * n, out.data());
    return out;
}
} // namespace locally_isothermal

} // namespace modules

// --------------------------------------------------------------------------- minidisk
namespace minidisk {
using conserved_array_t = nd::shared_array<iso2d::conserved_density_t, 2>;
using primitive_array_t = nd::shared_array<iso2d::primitive_t, 2>;
using solver_data_t = m3p_minidisk_data;

// problems/minidisk_scheme.cpp:294-355 (declared problems/minidisk.hpp:164-170)
inline conserved_array_t updated_conserved(conserved_array_t uc, primitive_array_t pe, double time, double dt,
                                           mesh::block_index_t<2> block, solver_data_t solver_data)
{
    const std::size_t n = uc.shape(0);
    if (n != std::size_t(solver_data.block_size) || pe.shape(0) != n + 4 || pe.shape(1) != n + 4)
        throw std::invalid_argument("minidisk::updated_conserved (array shapes do not match block_size)");
    device_buffer du(3 * n * n), dp(3 * (n + 4) * (n + 4)), dout(3 * n * n);
    upload_soa(uc.data(), n * n, du.ptr);
    upload_soa(pe.data(), (n + 4) * (n + 4), dp.ptr);
    check(m3p_minidisk_updated_conserved(du.ptr, dp.ptr, dout.ptr, int(block.level), int(block.coordinates[0]),
                                         int(block.coordinates[1]), time, dt, &solver_data, M3P_FP64, nullptr));
    conserved_array_t out(uc.shape());
    download_soa(dout.ptr, n * n, out.data());
    return out;
}
} // namespace minidisk

// --------------------------------------------------------------------------- mhd_scheme_v2
namespace mhd {
using primitive_t = std::array<double, 8>;           // rho, v1, v2, v3, p, B1, B2, B3
using conserved_density_t = std::array<double, 5>;   // D, S1, S2, S3, E
using godunov_data_t = std::array<double, 8>;        // 5 flux components, then the face electric field
}

// src/scheme_mhd_v2.hpp:50-210 (bodies src/scheme_mhd_v2.cpp:75-278): FP64,
// adiabatic index 5/3.  Face and edge arrays hold the field values as doubles.
namespace mhd_scheme_v2 {
using cell_primitive_variables_t   = nd::shared_array<mhd::primitive_t, 3>;
using cell_conserved_density_t     = nd::shared_array<mhd::conserved_density_t, 3>;
using face_godunov_data_t          = std::array<nd::shared_array<mhd::godunov_data_t, 3>, 3>;
using face_magnetic_flux_density_t = std::array<nd::shared_array<double, 3>, 3>;
using edge_electromotive_density_t = std::array<nd::shared_array<double, 3>, 3>;

namespace detail {
using shape3 = nd::uivec_t<3>;
inline std::size_t count(shape3 s) { return s[0] * s[1] * s[2]; }

inline std::unique_ptr<device_buffer> up(const nd::shared_array<double, 3>& a)
{
    auto d = std::make_unique<device_buffer>(a.size());
    hip_check(hipMemcpy(d->ptr, a.data(), a.size() * sizeof(double), hipMemcpyHostToDevice), "hipMemcpy H2D");
    return d;
}

template<std::size_t NC>
inline std::unique_ptr<device_buffer> up(const nd::shared_array<std::array<double, NC>, 3>& a)
{
    auto d = std::make_unique<device_buffer>(a.size() * NC);
    upload_soa(a.data(), a.size(), d->ptr);
    return d;
}

inline nd::shared_array<double, 3> down(const device_buffer& d, shape3 s)
{
    nd::shared_array<double, 3> a(s);
    hip_check(hipMemcpy(a.data(), d.ptr, count(s) * sizeof(double), hipMemcpyDeviceToHost), "hipMemcpy D2H");
    return a;
}

template<std::size_t NC>
inline nd::shared_array<std::array<double, NC>, 3> down_tuples(const device_buffer& d, shape3 s)
{
    nd::shared_array<std::array<double, NC>, 3> a(s);
    download_soa(d.ptr, count(s), a.data());
    return a;
}

inline shape3 grow(shape3 s, int d0, int d1, int d2) { return { s[0] + d0, s[1] + d1, s[2] + d2 }; }
inline shape3 face_shape(shape3 n, int a) { return grow(n, a == 0, a == 1, a == 2); }
inline shape3 edge_shape(shape3 n, int a) { return grow(n, a != 0, a != 1, a != 2); }

inline void require(bool ok, const char* what)
{
    if (!ok) throw std::invalid_argument(std::string(what) + " (array shapes do not match)");
}
} // namespace detail

// primitive_array, src/scheme_mhd_v2.cpp:75-81
inline cell_primitive_variables_t primitive_array(cell_conserved_density_t uc, face_magnetic_flux_density_t bf)
{
    using namespace detail;
    const shape3 n = uc.shape();
    for (int a = 0; a < 3; ++a) require(bf[a].shape() == face_shape(n, a), "mhd_scheme_v2::primitive_array");
    auto du = up(uc);
    auto b1 = up(bf[0]), b2 = up(bf[1]), b3 = up(bf[2]);
    device_buffer dp(8 * count(n));
    check(m3p_mhd_v2_primitive_array(du->ptr, b1->ptr, b2->ptr, b3->ptr, int(n[0]), int(n[1]), int(n[2]), dp.ptr, nullptr));
    return down_tuples<8>(dp, n);
}

// godunov_fluxes, :168-185 (pc extended by one zone on all axes, bf on the transverse axes)
inline face_godunov_data_t godunov_fluxes(cell_primitive_variables_t pc, face_magnetic_flux_density_t bf)
{
    using namespace detail;
    const shape3 m = pc.shape();
    for (int a = 0; a < 3; ++a)
        require(bf[a].shape() == grow(m, -(a == 0), -(a == 1), -(a == 2)), "mhd_scheme_v2::godunov_fluxes");
    auto dp = up(pc);
    auto b1 = up(bf[0]), b2 = up(bf[1]), b3 = up(bf[2]);
    device_buffer g1(8 * bf[0].size()), g2(8 * bf[1].size()), g3(8 * bf[2].size());
    check(m3p_mhd_v2_godunov_fluxes(dp->ptr, b1->ptr, b2->ptr, b3->ptr, int(m[0]), int(m[1]), int(m[2]),
                                    g1.ptr, g2.ptr, g3.ptr, nullptr));
    return { down_tuples<8>(g1, bf[0].shape()), down_tuples<8>(g2, bf[1].shape()), down_tuples<8>(g3, bf[2].shape()) };
}

// electromotive_forces, :191-214
inline edge_electromotive_density_t electromotive_forces(face_godunov_data_t gf)
{
    using namespace detail;
    const shape3 m = { gf[1].shape(0), gf[0].shape(1), gf[0].shape(2) };
    for (int a = 0; a < 3; ++a)
        require(gf[a].shape() == grow(m, -(a == 0), -(a == 1), -(a == 2)), "mhd_scheme_v2::electromotive_forces");
    auto g1 = up(gf[0]), g2 = up(gf[1]), g3 = up(gf[2]);
    std::array<shape3, 3> es = { grow(m, 0, -1, -1), grow(m, -1, 0, -1), grow(m, -1, -1, 0) };
    device_buffer e1(count(es[0])), e2(count(es[1])), e3(count(es[2]));
    check(m3p_mhd_v2_electromotive_forces(g1->ptr, g2->ptr, g3->ptr, int(m[0]), int(m[1]), int(m[2]),
                                          e1.ptr, e2.ptr, e3.ptr, nullptr));
    return { down(e1, es[0]), down(e2, es[1]), down(e3, es[2]) };
}

// updated_conserved_density, :220-241 (gf extended by one zone on the transverse axes)
inline cell_conserved_density_t updated_conserved_density(cell_conserved_density_t uc, face_godunov_data_t gf,
                                                          double dt, double dl)
{
    using namespace detail;
    const shape3 n = uc.shape();
    for (int a = 0; a < 3; ++a)
        require(gf[a].shape() == grow(n, a == 0 ? 1 : 2, a == 1 ? 1 : 2, a == 2 ? 1 : 2),
                "mhd_scheme_v2::updated_conserved_density");
    auto du = up(uc);
    auto g1 = up(gf[0]), g2 = up(gf[1]), g3 = up(gf[2]);
    device_buffer dout(5 * count(n));
    check(m3p_mhd_v2_updated_conserved_density(du->ptr, g1->ptr, g2->ptr, g3->ptr, int(n[0]), int(n[1]), int(n[2]),
                                               dt, dl, dout.ptr, nullptr));
    return down_tuples<5>(dout, n);
}

// curl, :247-258
inline face_magnetic_flux_density_t curl(edge_electromotive_density_t ee, double dl)
{
    using namespace detail;
    const shape3 n = { ee[0].shape(0), ee[1].shape(1), ee[2].shape(2) };
    for (int a = 0; a < 3; ++a) require(ee[a].shape() == edge_shape(n, a), "mhd_scheme_v2::curl");
    auto e1 = up(ee[0]), e2 = up(ee[1]), e3 = up(ee[2]);
    device_buffer b1(count(face_shape(n, 0))), b2(count(face_shape(n, 1))), b3(count(face_shape(n, 2)));
    check(m3p_mhd_v2_curl(e1->ptr, e2->ptr, e3->ptr, int(n[0]), int(n[1]), int(n[2]), dl, b1.ptr, b2.ptr, b3.ptr, nullptr));
    return { down(b1, face_shape(n, 0)), down(b2, face_shape(n, 1)), down(b3, face_shape(n, 2)) };
}

// updated_magnetic_flux_density, :264-278 (ee as electromotive_forces returns it
// for Godunov data extended by one zone)
inline face_magnetic_flux_density_t updated_magnetic_flux_density(face_magnetic_flux_density_t bf,
                                                                  edge_electromotive_density_t ee, double dt, double dl)
{
    using namespace detail;
    const shape3 n = { bf[1].shape(0), bf[0].shape(1), bf[0].shape(2) };
    for (int a = 0; a < 3; ++a) {
        require(bf[a].shape() == face_shape(n, a), "mhd_scheme_v2::updated_magnetic_flux_density");
        require(ee[a].shape() == grow(edge_shape(n, a), 2 * (a == 0), 2 * (a == 1), 2 * (a == 2)),
                "mhd_scheme_v2::updated_magnetic_flux_density");
    }
    auto b1 = up(bf[0]), b2 = up(bf[1]), b3 = up(bf[2]);
    auto e1 = up(ee[0]), e2 = up(ee[1]), e3 = up(ee[2]);
    device_buffer o1(bf[0].size()), o2(bf[1].size()), o3(bf[2].size());
    check(m3p_mhd_v2_updated_magnetic_flux_density(b1->ptr, b2->ptr, b3->ptr, e1->ptr, e2->ptr, e3->ptr,
                                                   int(n[0]), int(n[1]), int(n[2]), dt, dl, o1.ptr, o2.ptr, o3.ptr, nullptr));
    return { down(o1, bf[0].shape()), down(o2, bf[1].shape()), down(o3, bf[2].shape()) };
}

// extend_cell_primitive_variables, :126-135 (27 blocks in the order of mesh::neighbors_27)
inline cell_primitive_variables_t extend_cell_primitive_variables(std::vector<cell_primitive_variables_t> pc_vector,
                                                                  std::size_t block_size, std::size_t count)
{
    using namespace detail;
    if (pc_vector.size() != 27) throw std::invalid_argument("mhd_scheme_v2::extend_cell_primitive_variables (need 27 blocks)");
    std::vector<std::unique_ptr<device_buffer>> blocks;
    const double* ptrs[27];
    for (std::size_t m = 0; m < 27; ++m) {
        require(pc_vector[m].shape() == shape3{ block_size, block_size, block_size }, "mhd_scheme_v2::extend_cell_primitive_variables");
        blocks.push_back(up(pc_vector[m]));
        ptrs[m] = blocks.back()->ptr;
    }
    const std::size_t e = block_size + 2 * count;
    device_buffer out(8 * e * e * e);
    check(m3p_mhd_v2_extend_cell_primitive_variables(ptrs, int(block_size), int(count), out.ptr, nullptr));
    return down_tuples<8>(out, { e, e, e });
}

// extend_face_magnetic_flux_density, :141-162 (27 face triplets, three runs of 9
// in the order of mesh::neighbors_9 on axes 0, 1, 2)
inline face_magnetic_flux_density_t extend_face_magnetic_flux_density(std::vector<face_magnetic_flux_density_t> bf_vector,
                                                                      std::size_t block_size, std::size_t count)
{
    using namespace detail;
    if (bf_vector.size() != 27) throw std::invalid_argument("mhd_scheme_v2::extend_face_magnetic_flux_density (need 27 blocks)");
    const shape3 n = { block_size, block_size, block_size };
    std::vector<std::unique_ptr<device_buffer>> blocks;
    const double* ptrs[27];
    for (std::size_t m = 0; m < 27; ++m) {
        const int a = int(m / 9);
        require(bf_vector[m][a].shape() == face_shape(n, a), "mhd_scheme_v2::extend_face_magnetic_flux_density");
        blocks.push_back(up(bf_vector[m][a]));
        ptrs[m] = blocks.back()->ptr;
    }
    const std::size_t e = block_size + 2 * count;
    std::array<shape3, 3> os = { shape3{ block_size + 1, e, e }, shape3{ e, block_size + 1, e }, shape3{ e, e, block_size + 1 } };
    device_buffer o1(detail::count(os[0])), o2(detail::count(os[1])), o3(detail::count(os[2]));
    check(m3p_mhd_v2_extend_face_magnetic_flux_density(ptrs, int(block_size), int(count), o1.ptr, o2.ptr, o3.ptr, nullptr));
    return { down(o1, os[0]), down(o2, os[1]), down(o3, os[2]) };
}
} // namespace mhd_scheme_v2

} // namespace mara3p
