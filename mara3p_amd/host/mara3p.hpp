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
// compatibility path) on the stream given as its last, optional argument
// (default: the null stream).  The per-block functions may be called from
// any number of threads at once, as the reference's thread pool calls
// updated_conserved (src/parallel_computable.cpp:248-296): the staging
// buffers belong to the calling thread and the library reports each call's
// kernel errors to that call.  The whole-mesh functions (updated_solution,
// updated_conserved_tree) upload a tree once and run one fused stage launch
// over all its blocks.  Device-resident meshes use the whole-mesh stage
// entry points of the C ABI directly (m3p_euler2d_stage, m3p_minidisk_stage,
// ...).
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
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

// Ordering by level, then coordinates (keys of the block trees below).
template<unsigned long R>
bool operator<(const block_index_t<R>& a, const block_index_t<R>& b)
{
    return std::tie(a.level, a.coordinates) < std::tie(b.level, b.coordinates);
}
template<unsigned long R>
bool operator==(const block_index_t<R>& a, const block_index_t<R>& b)
{
    return a.level == b.level && a.coordinates == b.coordinates;
}

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
    std::size_t cells_per_block() const { return std::size_t(block_size); }
};
}

namespace cartesian_2d {
struct geometry_t
{
    double domain_size = 1.0;
    int block_size = 64;
    double cell_spacing(block_index_t<2> b) const { return domain_size / double(block_size) / double(1 << b.level); }
    std::size_t cells_per_block() const { return std::size_t(block_size) * std::size_t(block_size); }
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
// Staging.  A call copies its arrays to the device and back through numbered
// staging slots of the calling thread (device buffers per device, host
// buffers), which grow as needed and are reused by the thread's later calls:
// concurrent callers share nothing, and a warm caller allocates nothing.
namespace detail {

struct device_staging
{
    std::vector<double*> buf;
    std::vector<std::size_t> cap;
    device_staging() = default;
    device_staging(const device_staging&) = delete;
    device_staging& operator=(const device_staging&) = delete;
    ~device_staging() { for (double* p : buf) if (p) (void)hipFree(p); }
    double* get(std::size_t slot, std::size_t count)
    {
        if (slot >= buf.size()) { buf.resize(slot + 1, nullptr); cap.resize(slot + 1, 0); }
        if (cap[slot] < count) {
            if (buf[slot]) hip_check(hipFree(buf[slot]), "hipFree");
            buf[slot] = nullptr;
            cap[slot] = 0;
            hip_check(hipMalloc(&buf[slot], count * sizeof(double)), "hipMalloc");
            cap[slot] = count;
        }
        return buf[slot];
    }
};

// Device buffer `slot` of the calling thread, at least `count` doubles.
inline double* device_slot(std::size_t slot, std::size_t count)
{
    thread_local std::map<int, device_staging> per_device;
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    return per_device[dev].get(slot, std::max<std::size_t>(count, 1));
}

// Host buffer `slot` of the calling thread, at least `count` doubles.
inline double* host_slot(std::size_t slot, std::size_t count)
{
    thread_local std::vector<std::vector<double>> bufs;
    if (slot >= bufs.size()) bufs.resize(slot + 1);
    if (bufs[slot].size() < count) bufs[slot].resize(count);
    return bufs[slot].data();
}

} // namespace detail

// Logical tuples [zones][NC] (host) -> SoA [NC][zones] in device slot `slot`,
// component by component (never a memcpy of the tuple layout).
template<std::size_t NC>
inline double* upload_soa(const std::array<double, NC>* src, std::size_t zones, std::size_t slot, hipStream_t stream)
{
    double* h = detail::host_slot(slot, NC * zones);
    for (std::size_t z = 0; z < zones; ++z)
        for (std::size_t q = 0; q < NC; ++q) h[q * zones + z] = src[z][q];
    double* d = detail::device_slot(slot, NC * zones);
    hip_check(hipMemcpyAsync(d, h, NC * zones * sizeof(double), hipMemcpyHostToDevice, stream), "hipMemcpyAsync H2D");
    return d;
}

inline double* upload(const double* src, std::size_t count, std::size_t slot, hipStream_t stream)
{
    double* h = detail::host_slot(slot, count);
    std::copy(src, src + count, h);
    double* d = detail::device_slot(slot, count);
    hip_check(hipMemcpyAsync(d, h, count * sizeof(double), hipMemcpyHostToDevice, stream), "hipMemcpyAsync H2D");
    return d;
}

// SoA [NC][zones] (device) -> logical tuples; waits for the stream.
template<std::size_t NC>
inline void download_soa(const double* src, std::size_t zones, std::array<double, NC>* dst, std::size_t slot,
                         hipStream_t stream)
{
    double* h = detail::host_slot(slot, NC * zones);
    hip_check(hipMemcpyAsync(h, src, NC * zones * sizeof(double), hipMemcpyDeviceToHost, stream), "hipMemcpyAsync D2H");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    for (std::size_t z = 0; z < zones; ++z)
        for (std::size_t q = 0; q < NC; ++q) dst[z][q] = h[q * zones + z];
}

inline void download(const double* src, std::size_t count, double* dst, std::size_t slot, hipStream_t stream)
{
    double* h = detail::host_slot(slot, count);
    hip_check(hipMemcpyAsync(h, src, count * sizeof(double), hipMemcpyDeviceToHost, stream), "hipMemcpyAsync D2H");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    std::copy(h, h + count, dst);
}

// ---------------------------------------------------------------------------
// Block trees: the leaves of a bsp::shared_tree as an ordered map from block
// index to block array (src/core_bqo_tree.hpp; the reference stores lazy
// computables there, this API the evaluated arrays), and the topology (the
// leaf indexes).
namespace bsp {
template<typename T, unsigned long Rank>
using tree_t = std::map<mesh::block_index_t<Rank>, T>;

template<typename T, unsigned long Rank>
std::vector<mesh::block_index_t<Rank>> indexes(const tree_t<T, Rank>& tree)
{
    std::vector<mesh::block_index_t<Rank>> out;
    for (const auto& kv : tree) out.push_back(kv.first);
    return out;
}
} // namespace bsp

namespace detail {

// Leaf b's neighbour table row: next_on / prev_on of each axis (periodic on
// its level), as indexes into `leaves`; all leaves must share one level.
template<unsigned long R>
std::vector<int> uniform_neighbours(const std::vector<mesh::block_index_t<R>>& leaves)
{
    std::map<mesh::block_index_t<R>, int> at;
    for (std::size_t k = 0; k < leaves.size(); ++k) at[leaves[k]] = int(k);
    std::vector<int> nbr(2 * R * leaves.size());
    for (std::size_t k = 0; k < leaves.size(); ++k) {
        if (leaves[k].level != leaves[0].level)
            throw std::invalid_argument("mara3p: this update needs a tree whose leaves share one level");
        for (unsigned long a = 0; a < R; ++a) {
            auto lo = at.find(mesh::prev_on(leaves[k], a)), hi = at.find(mesh::next_on(leaves[k], a));
            if (lo == at.end() || hi == at.end())
                throw std::invalid_argument("mara3p: the tree does not cover its level (a neighbour block is missing)");
            nbr[2 * R * k + 2 * a] = lo->second;
            nbr[2 * R * k + 2 * a + 1] = hi->second;
        }
    }
    return nbr;
}

// Guard-fill recipes of a 2:1-balanced quad-tree for m3p_amr_euler2d_stage:
// per leaf and direction (-x, +x, -y, +y), the next_on / prev_on neighbour
// as amr::get_or_create_block (src/mesh_amr.hpp:300-328) finds it -- a leaf
// of the same level (kind 0), the leaf containing it one level up (kind 1,
// with the orthant), or four leaves one level down (kind 2, children in the
// order (0,0), (1,0), (0,1), (1,1)).
inline std::vector<int> quadtree_recipes(const std::vector<mesh::block_index_t<2>>& leaves)
{
    using idx = mesh::block_index_t<2>;
    std::map<idx, int> at;
    for (std::size_t k = 0; k < leaves.size(); ++k) at[leaves[k]] = int(k);
    std::vector<int> rec(leaves.size() * 4 * 8, -1);
    for (std::size_t k = 0; k < leaves.size(); ++k)
        for (int d = 0; d < 4; ++d) {
            const unsigned long axis = d / 2;
            const idx nb = d % 2 ? mesh::next_on(leaves[k], axis) : mesh::prev_on(leaves[k], axis);
            int* r = &rec[(k * 4 + d) * 8];
            auto same = at.find(nb);
            if (same != at.end()) { r[0] = 0; r[1] = same->second; continue; }
            if (nb.level > 0) {
                idx parent{ nb.level - 1, { nb.coordinates[0] >> 1, nb.coordinates[1] >> 1 } };
                auto up = at.find(parent);
                if (up != at.end()) {
                    r[0] = 1; r[1] = up->second;
                    r[6] = int(nb.coordinates[0] & 1); r[7] = int(nb.coordinates[1] & 1);
                    continue;
                }
            }
            r[0] = 2;
            for (int c = 0; c < 4; ++c) {
                idx child{ nb.level + 1, { 2 * nb.coordinates[0] + (c & 1), 2 * nb.coordinates[1] + (c >> 1) } };
                auto down = at.find(child);
                if (down == at.end()) throw std::invalid_argument("mara3p: the quad-tree is not 2:1 balanced");
                r[2 + c] = down->second;
            }
        }
    return rec;
}

// Device copy of an int table (caller's slot, reinterpreted storage).
inline const int* upload_ints(const std::vector<int>& v, std::size_t slot, hipStream_t stream)
{
    const std::size_t words = (v.size() * sizeof(int) + sizeof(double) - 1) / sizeof(double);
    double* d = device_slot(slot, words);
    double* h = host_slot(slot, words);
    std::memcpy(h, v.data(), v.size() * sizeof(int));
    hip_check(hipMemcpyAsync(d, h, v.size() * sizeof(int), hipMemcpyHostToDevice, stream), "hipMemcpyAsync H2D");
    return reinterpret_cast<const int*>(d);
}

// Uploads the blocks of a tree as one [nblk][NC][zones] SoA array.
template<std::size_t NC, unsigned long R, typename A>
double* upload_tree(const bsp::tree_t<A, R>& tree, std::size_t zones, std::size_t slot, hipStream_t stream)
{
    const std::size_t nblk = tree.size();
    double* h = host_slot(slot, nblk * NC * zones);
    std::size_t b = 0;
    for (const auto& kv : tree) {
        if (kv.second.size() != zones) throw std::invalid_argument("mara3p: tree blocks must all have block_size zones per axis");
        const std::array<double, NC>* src = kv.second.data();
        for (std::size_t z = 0; z < zones; ++z)
            for (std::size_t q = 0; q < NC; ++q) h[(b * NC + q) * zones + z] = src[z][q];
        ++b;
    }
    double* d = device_slot(slot, nblk * NC * zones);
    hip_check(hipMemcpyAsync(d, h, nblk * NC * zones * sizeof(double), hipMemcpyHostToDevice, stream), "hipMemcpyAsync H2D");
    return d;
}

// The inverse: a tree with the keys of `like` from a device SoA array.
template<std::size_t NC, unsigned long R, typename A>
bsp::tree_t<A, R> download_tree(const double* d, const bsp::tree_t<A, R>& like, std::size_t zones, std::size_t slot,
                                hipStream_t stream)
{
    const std::size_t nblk = like.size();
    double* h = host_slot(slot, nblk * NC * zones);
    hip_check(hipMemcpyAsync(h, d, nblk * NC * zones * sizeof(double), hipMemcpyDeviceToHost, stream), "hipMemcpyAsync D2H");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    bsp::tree_t<A, R> out;
    std::size_t b = 0;
    for (const auto& kv : like) {
        A a(kv.second.shape());
        std::array<double, NC>* dst = a.data();
        for (std::size_t z = 0; z < zones; ++z)
            for (std::size_t q = 0; q < NC; ++q) dst[z][q] = h[(b * NC + q) * zones + z];
        out.emplace(kv.first, std::move(a));
        ++b;
    }
    return out;
}

// Waits for the stream and reports the kernel errors latched on it.
inline void check_stream(hipStream_t stream)
{
    int code = 0, block = 0, zone = 0;
    check(m3p_check_errors(stream, &code, &block, &zone));
    if (code == M3P_ERR_NEGATIVE_DENSITY) throw std::invalid_argument("euler::recover_primitive (negative density)");
    if (code) throw std::runtime_error("mara3p: kernel error " + std::to_string(code));
}

} // namespace detail

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
using conserved_tree_t  = bsp::tree_t<conserved_array_t, 2>;
using mesh_topology_t   = std::vector<mesh::block_index_t<2>>;

// src/module_euler.hpp:181-186
struct solution_t
{
    double iteration = 0.0;
    double time = 0.0;
    conserved_tree_t conserved;
};

// src/module_euler.cpp:240-245
inline primitive_array_t recover_primitive_array(conserved_array_t uc, double gamma_law_index, hipStream_t stream = nullptr)
{
    const std::size_t nz = uc.size();
    double* du = upload_soa(uc.data(), nz, 0, stream);
    double* dp = detail::device_slot(1, 5 * nz);
    check(m3p_euler_recover_primitive_array(du, dp, long(nz), gamma_law_index, stream));
    primitive_array_t p(uc.shape());
    download_soa(dp, nz, p.data(), 1, stream);
    return p;
}

// src/module_euler.cpp:251-257
inline primitive_array_t estimate_gradient(primitive_array_t pc, unsigned long axis, double plm_theta,
                                           hipStream_t stream = nullptr)
{
    const std::size_t ni = pc.shape(0), nj = pc.shape(1), nz = pc.size();
    if (axis > 1 || (axis == 0 ? ni : nj) < 3) throw std::invalid_argument("nd::select (selection out of range)");
    auto oshape = nd::uivec(axis == 0 ? ni - 2 : ni, axis == 1 ? nj - 2 : nj);
    const std::size_t no = oshape[0] * oshape[1];
    double* dp = upload_soa(pc.data(), nz, 0, stream);
    double* dg = detail::device_slot(1, 5 * no);
    check(m3p_plm_gradient_2d(dp, dg, 5, int(ni), int(nj), int(axis), plm_theta, stream));
    primitive_array_t g(oshape);
    download_soa(dg, no, g.data(), 1, stream);
    return g;
}

// src/module_euler.cpp:263-298 (declared src/module_euler.hpp:182-190)
inline conserved_array_t updated_conserved(conserved_array_t uc, primitive_array_t pe, double /*time*/, double dt,
                                           mesh_geometry_t mesh_geometry, mesh::block_index_t<2> block,
                                           double plm_theta, double gamma_law_index, hipStream_t stream = nullptr)
{
    const std::size_t n = uc.shape(0);
    if (uc.shape(1) != n || pe.shape(0) != n + 4 || pe.shape(1) != n + 4)
        throw std::invalid_argument("euler2d::updated_conserved (pe must be uc extended by 2 guard zones)");
    double* du = upload_soa(uc.data(), n * n, 0, stream);
    double* dp = upload_soa(pe.data(), (n + 4) * (n + 4), 1, stream);
    double* dout = detail::device_slot(2, 5 * n * n);
    check(m3p_euler2d_updated_conserved(du, dp, dout, int(n), dt, mesh_geometry.cell_spacing(block),
                                        plm_theta, gamma_law_index, 0, M3P_FP64, stream));
    conserved_array_t out(uc.shape());
    download_soa(dout, n * n, out.data(), 2, stream);
    return out;
}

// updated_solution, src/module_euler.cpp:304-323: one forward-Euler update of
// every block (guard zones from the neighbours, amr::extend_block over
// amr::get_or_create_block), iteration + 1, time + dt.  A tree whose leaves
// share one level runs the uniform-mesh stage; a 2:1-balanced tree with
// level jumps runs the quad-tree stage (injection from coarser and 2 x 2
// means of finer neighbours, as the reference).
inline solution_t updated_solution(solution_t solution, double dt, mesh_geometry_t mesh_geometry,
                                   double plm_theta, double gamma_law_index, hipStream_t stream = nullptr)
{
    const auto leaves = bsp::indexes(solution.conserved);
    if (leaves.empty()) throw std::invalid_argument("euler2d::updated_solution (empty tree)");
    const std::size_t n = std::size_t(mesh_geometry.block_size), zones = n * n;
    bool uniform = true;
    for (const auto& b : leaves) uniform &= b.level == leaves[0].level;
    double* du = detail::upload_tree<5>(solution.conserved, zones, 0, stream);
    double* dout = detail::device_slot(1, leaves.size() * 5 * zones);
    if (uniform) {
        const int* nbr = detail::upload_ints(detail::uniform_neighbours(leaves), 2, stream);
        check(m3p_euler2d_stage(du, nullptr, dout, nbr, int(leaves.size()), int(n), dt, nullptr,
                                mesh_geometry.cell_spacing(leaves[0]), 0.5, plm_theta, gamma_law_index, 0, M3P_FP64, stream));
    } else {
        const int* rec = detail::upload_ints(detail::quadtree_recipes(leaves), 2, stream);
        std::vector<double> dl(leaves.size());
        for (std::size_t k = 0; k < leaves.size(); ++k) dl[k] = mesh_geometry.cell_spacing(leaves[k]);
        double* ddl = upload(dl.data(), dl.size(), 3, stream);
        check(m3p_amr_euler2d_stage(du, nullptr, dout, nullptr, rec, ddl, int(leaves.size()), int(n), dt, nullptr, 0.5,
                                    plm_theta, gamma_law_index, 0, M3P_FP64, stream));
    }
    detail::check_stream(stream);
    return { solution.iteration + 1, solution.time + dt,
             detail::download_tree<5>(dout, solution.conserved, zones, 4, stream) };
}

// src/module_euler.cpp:329-337
inline double smallest_cell_size(mesh_topology_t mesh_topology, mesh_geometry_t mesh_geometry)
{
    double dl = std::numeric_limits<double>::max();
    for (const auto& b : mesh_topology) dl = std::min(dl, mesh_geometry.cell_spacing(b));
    return dl;
}

// src/module_euler.cpp:343-348
inline std::size_t total_zones(mesh_topology_t mesh_topology, mesh_geometry_t mesh_geometry)
{
    return mesh_topology.size() * mesh_geometry.cells_per_block();
}

// weighted_sum over blocks (src/module_euler.cpp:353-360 -> src/mesh_amr.hpp:455-467)
inline conserved_array_t weighted_sum(conserved_array_t s, conserved_array_t t, double b, hipStream_t stream = nullptr)
{
    const std::size_t nz = s.size();
    if (t.size() != nz) throw std::invalid_argument("euler2d::weighted_sum (shape mismatch)");
    double* ds = upload_soa(s.data(), nz, 0, stream);
    double* dt_ = upload_soa(t.data(), nz, 1, stream);
    double* dout = detail::device_slot(2, 5 * nz);
    check(m3p_weighted_sum(ds, dt_, dout, long(5 * nz), b, stream));
    conserved_array_t out(s.shape());
    download_soa(dout, nz, out.data(), 2, stream);
    return out;
}

// weighted_sum of solutions, src/module_euler.cpp:353-360: iteration, time
// and every block as s * b + t * (1 - b).
inline solution_t weighted_sum(solution_t s, solution_t t, double b, hipStream_t stream = nullptr)
{
    if (bsp::indexes(s.conserved) != bsp::indexes(t.conserved))
        throw std::invalid_argument("euler2d::weighted_sum (the trees differ)");
    solution_t out{ s.iteration * b + t.iteration * (1 - b), s.time * b + t.time * (1 - b), {} };
    for (const auto& kv : s.conserved) out.conserved.emplace(kv.first, weighted_sum(kv.second, t.conserved.at(kv.first), b, stream));
    return out;
}
} // namespace euler2d

// --------------------------------------------------------------------------- euler1d
namespace euler1d {
using conserved_array_t = nd::shared_array<euler::conserved_density_t, 1>;
using primitive_array_t = nd::shared_array<euler::primitive_t, 1>;
using mesh_geometry_t   = mesh::cartesian_1d::geometry_t;
using conserved_tree_t  = bsp::tree_t<conserved_array_t, 1>;
using mesh_topology_t   = std::vector<mesh::block_index_t<1>>;
// The reference's boundary_condition_t is a std::function applied to pe; the
// outflow condition (src/mara.cpp:123-143) is the one the module supports
// natively here: it is selected by block position, as in the reference.
enum class boundary_condition_t { periodic, outflow };

// src/module_euler.hpp:66-71
struct solution_t
{
    double iteration = 0.0;
    double time = 0.0;
    conserved_tree_t conserved;
};

inline primitive_array_t recover_primitive_array(conserved_array_t uc, double gamma_law_index, hipStream_t stream = nullptr)
{
    const std::size_t nz = uc.size();
    double* du = upload_soa(uc.data(), nz, 0, stream);
    double* dp = detail::device_slot(1, 5 * nz);
    check(m3p_euler_recover_primitive_array(du, dp, long(nz), gamma_law_index, stream));
    primitive_array_t p(uc.shape());
    download_soa(dp, nz, p.data(), 1, stream);
    return p;
}

// src/module_euler.cpp:107-135 (declared src/module_euler.hpp:98-107)
inline conserved_array_t updated_conserved(conserved_array_t uc, primitive_array_t pe, double /*time*/, double dt,
                                           mesh_geometry_t mesh_geometry, boundary_condition_t bc,
                                           mesh::block_index_t<1> block, double plm_theta, double gamma_law_index,
                                           hipStream_t stream = nullptr)
{
    const std::size_t n = uc.shape(0);
    if (pe.shape(0) != n + 4) throw std::invalid_argument("euler1d::updated_conserved (pe must be uc extended by 2 guard zones)");
    const bool outflow = bc == boundary_condition_t::outflow;
    const int lower = outflow && block.coordinates[0] == 0;
    const int upper = outflow && !lower && block.coordinates[0] == (1ul << block.level) - 1;
    double* du = upload_soa(uc.data(), n, 0, stream);
    double* dp = upload_soa(pe.data(), n + 4, 1, stream);
    double* dout = detail::device_slot(2, 5 * n);
    check(m3p_euler1d_updated_conserved(du, dp, dout, int(n), dt, mesh_geometry.cell_spacing(block),
                                        plm_theta, gamma_law_index, lower, upper, M3P_FP64, stream));
    conserved_array_t out(uc.shape());
    download_soa(dout, n, out.data(), 2, stream);
    return out;
}

// updated_solution, src/module_euler.cpp:141-161: one forward-Euler update of
// every block of a tree whose leaves share one level, the chain's ends
// closed by the boundary condition (outflow: the first and last block copy
// their edge zone into the guard zones, src/mara.cpp:123-143; periodic: the
// ends wrap), iteration + 1, time + dt.
inline solution_t updated_solution(solution_t solution, double dt, mesh_geometry_t mesh_geometry,
                                   boundary_condition_t bc, double plm_theta, double gamma_law_index,
                                   hipStream_t stream = nullptr)
{
    const auto leaves = bsp::indexes(solution.conserved);
    if (leaves.empty()) throw std::invalid_argument("euler1d::updated_solution (empty tree)");
    const std::size_t n = std::size_t(mesh_geometry.block_size);
    std::vector<int> nbr = detail::uniform_neighbours(leaves);
    if (bc == boundary_condition_t::outflow) {
        nbr.front() = -1;    // prev of the first block
        nbr.back() = -1;     // next of the last block
    }
    double* du = detail::upload_tree<5>(solution.conserved, n, 0, stream);
    double* dout = detail::device_slot(1, leaves.size() * 5 * n);
    const int* dn = detail::upload_ints(nbr, 2, stream);
    check(m3p_euler1d_stage(du, nullptr, dout, dn, int(leaves.size()), int(n), dt, mesh_geometry.cell_spacing(leaves[0]),
                            0.5, plm_theta, gamma_law_index, M3P_FP64, stream));
    detail::check_stream(stream);
    return { solution.iteration + 1, solution.time + dt,
             detail::download_tree<5>(dout, solution.conserved, n, 3, stream) };
}

// src/module_euler.cpp:166-175
inline double smallest_cell_size(mesh_topology_t mesh_topology, mesh_geometry_t mesh_geometry)
{
    double dl = std::numeric_limits<double>::max();
    for (const auto& b : mesh_topology) dl = std::min(dl, mesh_geometry.cell_spacing(b));
    return dl;
}

// src/module_euler.cpp:181-186
inline std::size_t total_zones(mesh_topology_t mesh_topology, mesh_geometry_t mesh_geometry)
{
    return mesh_topology.size() * mesh_geometry.cells_per_block();
}

// weighted_sum of solutions, src/module_euler.cpp:191-198
inline solution_t weighted_sum(solution_t s, solution_t t, double b, hipStream_t stream = nullptr)
{
    if (bsp::indexes(s.conserved) != bsp::indexes(t.conserved))
        throw std::invalid_argument("euler1d::weighted_sum (the trees differ)");
    solution_t out{ s.iteration * b + t.iteration * (1 - b), s.time * b + t.time * (1 - b), {} };
    for (const auto& kv : s.conserved) {
        const conserved_array_t& x = kv.second;
        const conserved_array_t& y = t.conserved.at(kv.first);
        const std::size_t nz = x.size();
        double* ds = upload_soa(x.data(), nz, 0, stream);
        double* dt_ = upload_soa(y.data(), nz, 1, stream);
        double* dout = detail::device_slot(2, 5 * nz);
        check(m3p_weighted_sum(ds, dt_, dout, long(5 * nz), b, stream));
        conserved_array_t w(x.shape());
        download_soa(dout, nz, w.data(), 2, stream);
        out.conserved.emplace(kv.first, std::move(w));
    }
    return out;
}
} // namespace euler1d

// --------------------------------------------------------------------------- locally_isothermal
namespace locally_isothermal {
using conserved_array_t     = nd::shared_array<iso2d::conserved_density_t, 2>;
using primitive_array_t     = nd::shared_array<iso2d::primitive_t, 2>;
using mesh_geometry_t       = mesh::cartesian_2d::geometry_t;
using conserved_tree_t      = bsp::tree_t<conserved_array_t, 2>;
using mesh_topology_t       = std::vector<mesh::block_index_t<2>>;
using coords_t              = std::array<double, 2>;
using temperature_mapping_t = std::function<double(coords_t)>;

inline primitive_array_t recover_primitive_array(conserved_array_t uc, hipStream_t stream = nullptr)
{
    const std::size_t nz = uc.size();
    double* du = upload_soa(uc.data(), nz, 0, stream);
    double* dp = detail::device_slot(1, 3 * nz);
    check(m3p_iso2d_recover_primitive_array(du, dp, long(nz), stream));
    primitive_array_t p(uc.shape());
    download_soa(dp, nz, p.data(), 1, stream);
    return p;
}

namespace detail_ {
// The temperature mapping at the faces of block b (the reference evaluates it
// per face, src/module_locally_isothermal.cpp:115-121): cx [(n+1)][n] on the
// x faces, cy [n][(n+1)] on the y faces.
inline void face_temperatures(const mesh_geometry_t& g, mesh::block_index_t<2> b, const temperature_mapping_t& T,
                              double* cx, double* cy)
{
    const std::size_t n = std::size_t(g.block_size);
    for (std::size_t i = 0; i <= n; ++i)
        for (std::size_t j = 0; j < n; ++j) cx[i * n + j] = T(g.face_coordinate(b, 0, i, j));
    for (std::size_t i = 0; i < n; ++i)
        for (std::size_t j = 0; j <= n; ++j) cy[i * (n + 1) + j] = T(g.face_coordinate(b, 1, i, j));
}
} // namespace detail_

// src/module_locally_isothermal.cpp:182-212.  The temperature mapping is a
// host callback in the reference too; it is evaluated here at the face
// coordinates once per call and handed to the kernel as arrays.  Like the
// reference, plm_theta is ignored (theta = 2 is used, :195-196).
inline conserved_array_t updated_conserved(conserved_array_t uc, primitive_array_t pe, double /*time*/, double dt,
                                           mesh_geometry_t mesh_geometry, mesh::block_index_t<2> block,
                                           double kinematic_viscosity, temperature_mapping_t temperature,
                                           double /*plm_theta*/, hipStream_t stream = nullptr)
{
    const std::size_t n = uc.shape(0);
    if (pe.shape(0) != n + 4 || pe.shape(1) != n + 4)
        throw std::invalid_argument("locally_isothermal::updated_conserved (pe must be uc extended by 2 guard zones)");
    if (kinematic_viscosity < 0.0)
        throw std::invalid_argument("locally_isothermal::updated_conserved (negative kinematic viscosity)");
    mesh_geometry_t g = mesh_geometry;
    g.block_size = int(n);
    std::vector<double> cx((n + 1) * n), cy(n * (n + 1));
    detail_::face_temperatures(g, block, temperature, cx.data(), cy.data());
    double* du = upload_soa(uc.data(), n * n, 0, stream);
    double* dp = upload_soa(pe.data(), (n + 4) * (n + 4), 1, stream);
    double* dcx = upload(cx.data(), cx.size(), 3, stream);
    double* dcy = upload(cy.data(), cy.size(), 4, stream);
    double* dout = detail::device_slot(2, 3 * n * n);
    check(m3p_locally_isothermal_updated_conserved(du, dp, dout, int(n), dt, mesh_geometry.cell_spacing(block),
                                                   dcx, dcy, kinematic_viscosity, M3P_FP64, stream));
    conserved_array_t out(uc.shape());
    download_soa(dout, n * n, out.data(), 2, stream);
    return out;
}

// updated_conserved_tree, src/module_locally_isothermal.cpp:218-238: every
// block of a tree whose leaves share one level, updated in one stage launch
// (guard zones from the neighbours, periodic on the level).
inline conserved_tree_t updated_conserved_tree(conserved_tree_t conserved, double /*time*/, double dt,
                                               mesh_geometry_t mesh_geometry, temperature_mapping_t temperature,
                                               double kinematic_viscosity, double /*plm_theta*/,
                                               hipStream_t stream = nullptr)
{
    const auto leaves = bsp::indexes(conserved);
    if (leaves.empty()) throw std::invalid_argument("locally_isothermal::updated_conserved_tree (empty tree)");
    if (kinematic_viscosity < 0.0)
        throw std::invalid_argument("locally_isothermal::updated_conserved_tree (negative kinematic viscosity)");
    const std::size_t n = std::size_t(mesh_geometry.block_size), zones = n * n, faces = (n + 1) * n;
    const std::vector<int> nbr = detail::uniform_neighbours(leaves);
    std::vector<double> cx(leaves.size() * faces), cy(leaves.size() * faces);
    for (std::size_t k = 0; k < leaves.size(); ++k)
        detail_::face_temperatures(mesh_geometry, leaves[k], temperature, &cx[k * faces], &cy[k * faces]);
    double* du = detail::upload_tree<3>(conserved, zones, 0, stream);
    double* dout = detail::device_slot(1, leaves.size() * 3 * zones);
    const int* dn = detail::upload_ints(nbr, 2, stream);
    double* dcx = upload(cx.data(), cx.size(), 3, stream);
    double* dcy = upload(cy.data(), cy.size(), 4, stream);
    check(m3p_locally_isothermal_stage(du, nullptr, dout, dn, int(leaves.size()), int(n), dt, nullptr,
                                       mesh_geometry.cell_spacing(leaves[0]), 0.5, dcx, dcy, kinematic_viscosity,
                                       M3P_FP64, stream));
    detail::check_stream(stream);
    return detail::download_tree<3>(dout, conserved, zones, 5, stream);
}

// src/module_locally_isothermal.cpp:244-252
inline double smallest_cell_size(mesh_topology_t mesh_topology, mesh_geometry_t mesh_geometry)
{
    double dl = std::numeric_limits<double>::max();
    for (const auto& b : mesh_topology) dl = std::min(dl, mesh_geometry.cell_spacing(b));
    return dl;
}

// src/module_locally_isothermal.cpp:258-263
inline std::size_t total_zones(mesh_topology_t mesh_topology, mesh_geometry_t mesh_geometry)
{
    return mesh_topology.size() * mesh_geometry.cells_per_block();
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
                                           mesh::block_index_t<2> block, solver_data_t solver_data,
                                           hipStream_t stream = nullptr)
{
    const std::size_t n = uc.shape(0);
    if (n != std::size_t(solver_data.block_size) || pe.shape(0) != n + 4 || pe.shape(1) != n + 4)
        throw std::invalid_argument("minidisk::updated_conserved (array shapes do not match block_size)");
    double* du = upload_soa(uc.data(), n * n, 0, stream);
    double* dp = upload_soa(pe.data(), (n + 4) * (n + 4), 1, stream);
    double* dout = detail::device_slot(2, 3 * n * n);
    check(m3p_minidisk_updated_conserved(du, dp, dout, int(block.level), int(block.coordinates[0]),
                                         int(block.coordinates[1]), time, dt, &solver_data, M3P_FP64, stream));
    conserved_array_t out(uc.shape());
    download_soa(dout, n * n, out.data(), 2, stream);
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

// Staging through the calling thread's slots (mara3p::detail), one slot per
// array of a call.
inline double* up(const nd::shared_array<double, 3>& a, std::size_t slot, hipStream_t stream)
{
    return upload(a.data(), a.size(), slot, stream);
}

template<std::size_t NC>
inline double* up(const nd::shared_array<std::array<double, NC>, 3>& a, std::size_t slot, hipStream_t stream)
{
    return upload_soa(a.data(), a.size(), slot, stream);
}

inline double* out(std::size_t slot, std::size_t count) { return mara3p::detail::device_slot(slot, count); }

inline nd::shared_array<double, 3> down(const double* d, shape3 s, std::size_t slot, hipStream_t stream)
{
    nd::shared_array<double, 3> a(s);
    download(d, count(s), a.data(), slot, stream);
    return a;
}

template<std::size_t NC>
inline nd::shared_array<std::array<double, NC>, 3> down_tuples(const double* d, shape3 s, std::size_t slot, hipStream_t stream)
{
    nd::shared_array<std::array<double, NC>, 3> a(s);
    download_soa(d, count(s), a.data(), slot, stream);
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
inline cell_primitive_variables_t primitive_array(cell_conserved_density_t uc, face_magnetic_flux_density_t bf,
                                                  hipStream_t stream = nullptr)
{
    using namespace detail;
    const shape3 n = uc.shape();
    for (int a = 0; a < 3; ++a) require(bf[a].shape() == face_shape(n, a), "mhd_scheme_v2::primitive_array");
    double* du = up(uc, 0, stream);
    double* b1 = up(bf[0], 1, stream);
    double* b2 = up(bf[1], 2, stream);
    double* b3 = up(bf[2], 3, stream);
    double* dp = out(4, 8 * count(n));
    check(m3p_mhd_v2_primitive_array(du, b1, b2, b3, int(n[0]), int(n[1]), int(n[2]), dp, stream));
    return down_tuples<8>(dp, n, 4, stream);
}

// godunov_fluxes, :168-185 (pc extended by one zone on all axes, bf on the transverse axes)
inline face_godunov_data_t godunov_fluxes(cell_primitive_variables_t pc, face_magnetic_flux_density_t bf,
                                          hipStream_t stream = nullptr)
{
    using namespace detail;
    const shape3 m = pc.shape();
    for (int a = 0; a < 3; ++a)
        require(bf[a].shape() == grow(m, -(a == 0), -(a == 1), -(a == 2)), "mhd_scheme_v2::godunov_fluxes");
    double* dp = up(pc, 0, stream);
    double* b1 = up(bf[0], 1, stream);
    double* b2 = up(bf[1], 2, stream);
    double* b3 = up(bf[2], 3, stream);
    double* g1 = out(4, 8 * bf[0].size());
    double* g2 = out(5, 8 * bf[1].size());
    double* g3 = out(6, 8 * bf[2].size());
    check(m3p_mhd_v2_godunov_fluxes(dp, b1, b2, b3, int(m[0]), int(m[1]), int(m[2]), g1, g2, g3, stream));
    return { down_tuples<8>(g1, bf[0].shape(), 4, stream), down_tuples<8>(g2, bf[1].shape(), 5, stream),
             down_tuples<8>(g3, bf[2].shape(), 6, stream) };
}

// electromotive_forces, :191-214
inline edge_electromotive_density_t electromotive_forces(face_godunov_data_t gf, hipStream_t stream = nullptr)
{
    using namespace detail;
    const shape3 m = { gf[1].shape(0), gf[0].shape(1), gf[0].shape(2) };
    for (int a = 0; a < 3; ++a)
        require(gf[a].shape() == grow(m, -(a == 0), -(a == 1), -(a == 2)), "mhd_scheme_v2::electromotive_forces");
    double* g1 = up(gf[0], 0, stream);
    double* g2 = up(gf[1], 1, stream);
    double* g3 = up(gf[2], 2, stream);
    std::array<shape3, 3> es = { grow(m, 0, -1, -1), grow(m, -1, 0, -1), grow(m, -1, -1, 0) };
    double* e1 = out(3, count(es[0]));
    double* e2 = out(4, count(es[1]));
    double* e3 = out(5, count(es[2]));
    check(m3p_mhd_v2_electromotive_forces(g1, g2, g3, int(m[0]), int(m[1]), int(m[2]), e1, e2, e3, stream));
    return { down(e1, es[0], 3, stream), down(e2, es[1], 4, stream), down(e3, es[2], 5, stream) };
}

// updated_conserved_density, :220-241 (gf extended by one zone on the transverse axes)
inline cell_conserved_density_t updated_conserved_density(cell_conserved_density_t uc, face_godunov_data_t gf,
                                                          double dt, double dl, hipStream_t stream = nullptr)
{
    using namespace detail;
    const shape3 n = uc.shape();
    for (int a = 0; a < 3; ++a)
        require(gf[a].shape() == grow(n, a == 0 ? 1 : 2, a == 1 ? 1 : 2, a == 2 ? 1 : 2),
                "mhd_scheme_v2::updated_conserved_density");
    double* du = up(uc, 0, stream);
    double* g1 = up(gf[0], 1, stream);
    double* g2 = up(gf[1], 2, stream);
    double* g3 = up(gf[2], 3, stream);
    double* dout = out(4, 5 * count(n));
    check(m3p_mhd_v2_updated_conserved_density(du, g1, g2, g3, int(n[0]), int(n[1]), int(n[2]), dt, dl, dout, stream));
    return down_tuples<5>(dout, n, 4, stream);
}

// curl, :247-258
inline face_magnetic_flux_density_t curl(edge_electromotive_density_t ee, double dl, hipStream_t stream = nullptr)
{
    using namespace detail;
    const shape3 n = { ee[0].shape(0), ee[1].shape(1), ee[2].shape(2) };
    for (int a = 0; a < 3; ++a) require(ee[a].shape() == edge_shape(n, a), "mhd_scheme_v2::curl");
    double* e1 = up(ee[0], 0, stream);
    double* e2 = up(ee[1], 1, stream);
    double* e3 = up(ee[2], 2, stream);
    double* b1 = out(3, count(face_shape(n, 0)));
    double* b2 = out(4, count(face_shape(n, 1)));
    double* b3 = out(5, count(face_shape(n, 2)));
    check(m3p_mhd_v2_curl(e1, e2, e3, int(n[0]), int(n[1]), int(n[2]), dl, b1, b2, b3, stream));
    return { down(b1, face_shape(n, 0), 3, stream), down(b2, face_shape(n, 1), 4, stream),
             down(b3, face_shape(n, 2), 5, stream) };
}

// updated_magnetic_flux_density, :264-278 (ee as electromotive_forces returns it
// for Godunov data extended by one zone)
inline face_magnetic_flux_density_t updated_magnetic_flux_density(face_magnetic_flux_density_t bf,
                                                                  edge_electromotive_density_t ee, double dt, double dl,
                                                                  hipStream_t stream = nullptr)
{
    using namespace detail;
    const shape3 n = { bf[1].shape(0), bf[0].shape(1), bf[0].shape(2) };
    for (int a = 0; a < 3; ++a) {
        require(bf[a].shape() == face_shape(n, a), "mhd_scheme_v2::updated_magnetic_flux_density");
        require(ee[a].shape() == grow(edge_shape(n, a), 2 * (a == 0), 2 * (a == 1), 2 * (a == 2)),
                "mhd_scheme_v2::updated_magnetic_flux_density");
    }
    double* b1 = up(bf[0], 0, stream);
    double* b2 = up(bf[1], 1, stream);
    double* b3 = up(bf[2], 2, stream);
    double* e1 = up(ee[0], 3, stream);
    double* e2 = up(ee[1], 4, stream);
    double* e3 = up(ee[2], 5, stream);
    double* o1 = out(6, bf[0].size());
    double* o2 = out(7, bf[1].size());
    double* o3 = out(8, bf[2].size());
    check(m3p_mhd_v2_updated_magnetic_flux_density(b1, b2, b3, e1, e2, e3, int(n[0]), int(n[1]), int(n[2]), dt, dl,
                                                   o1, o2, o3, stream));
    return { down(o1, bf[0].shape(), 6, stream), down(o2, bf[1].shape(), 7, stream), down(o3, bf[2].shape(), 8, stream) };
}

// extend_cell_primitive_variables, :126-135 (27 blocks in the order of mesh::neighbors_27)
inline cell_primitive_variables_t extend_cell_primitive_variables(std::vector<cell_primitive_variables_t> pc_vector,
                                                                  std::size_t block_size, std::size_t count,
                                                                  hipStream_t stream = nullptr)
{
    using namespace detail;
    if (pc_vector.size() != 27) throw std::invalid_argument("mhd_scheme_v2::extend_cell_primitive_variables (need 27 blocks)");
    const double* ptrs[27];
    for (std::size_t m = 0; m < 27; ++m) {
        require(pc_vector[m].shape() == shape3{ block_size, block_size, block_size }, "mhd_scheme_v2::extend_cell_primitive_variables");
        ptrs[m] = up(pc_vector[m], m, stream);
    }
    const std::size_t e = block_size + 2 * count;
    double* o = out(27, 8 * e * e * e);
    check(m3p_mhd_v2_extend_cell_primitive_variables(ptrs, int(block_size), int(count), o, stream));
    return down_tuples<8>(o, { e, e, e }, 27, stream);
}

// extend_face_magnetic_flux_density, :141-162 (27 face triplets, three runs of 9
// in the order of mesh::neighbors_9 on axes 0, 1, 2)
inline face_magnetic_flux_density_t extend_face_magnetic_flux_density(std::vector<face_magnetic_flux_density_t> bf_vector,
                                                                      std::size_t block_size, std::size_t count,
                                                                      hipStream_t stream = nullptr)
{
    using namespace detail;
    if (bf_vector.size() != 27) throw std::invalid_argument("mhd_scheme_v2::extend_face_magnetic_flux_density (need 27 blocks)");
    const shape3 n = { block_size, block_size, block_size };
    const double* ptrs[27];
    for (std::size_t m = 0; m < 27; ++m) {
        const int a = int(m / 9);
        require(bf_vector[m][a].shape() == face_shape(n, a), "mhd_scheme_v2::extend_face_magnetic_flux_density");
        ptrs[m] = up(bf_vector[m][a], m, stream);
    }
    const std::size_t e = block_size + 2 * count;
    std::array<shape3, 3> os = { shape3{ block_size + 1, e, e }, shape3{ e, block_size + 1, e }, shape3{ e, e, block_size + 1 } };
    double* o1 = out(27, detail::count(os[0]));
    double* o2 = out(28, detail::count(os[1]));
    double* o3 = out(29, detail::count(os[2]));
    check(m3p_mhd_v2_extend_face_magnetic_flux_density(ptrs, int(block_size), int(count), o1, o2, o3, stream));
    return { down(o1, os[0], 27, stream), down(o2, os[1], 28, stream), down(o3, os[2], 29, stream) };
}
} // namespace mhd_scheme_v2

} // namespace mara3p
