// C++ host API (mara3p_amd/host/mara3p.hpp) against the oracle, called the way
// a reference caller would call modules::euler2d::updated_conserved and
// friends.  Built by mara3p_amd/build.py; run by tests/test_gpu_host_api.py.
// Exit status 0 and "ALL PASS" on success.  With "--dump-amr FILE" it instead
// runs euler2d::updated_solution on a quad-tree with level jumps and writes
// the tree, the input and the output to FILE for the Python AMR oracle.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <thread>
#include "mara3p.hpp"
#include "../../oracle/oracle.h"

using namespace mara3p;

static int failures = 0;

#define EXPECT(cond, what) do { if (!(cond)) { std::printf("FAIL: %s\n", what); ++failures; } else std::printf("ok: %s\n", what); } while (0)

template<std::size_t NC, unsigned long R>
static std::vector<double> flat(const nd::shared_array<std::array<double, NC>, R>& a)
{
    std::vector<double> v(a.size() * NC);
    for (std::size_t z = 0; z < a.size(); ++z)
        for (std::size_t q = 0; q < NC; ++q) v[z * NC + q] = a.data()[z][q];
    return v;
}

static void test_euler2d(std::mt19937& rng)
{
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const std::size_t n = 64;
    modules::euler2d::primitive_array_t pe(nd::uivec(n + 4, n + 4));
    for (std::size_t i = 0; i < n + 4; ++i)
        for (std::size_t j = 0; j < n + 4; ++j)
            pe(i, j) = { 1.0 + 0.3 * std::sin(0.2 * i + 0.1 * j) + 0.01 * U(rng), 0.3 * U(rng), -0.2 * U(rng), 0.05 * U(rng),
                         1.0 + 0.2 * std::cos(0.15 * i) + 0.01 * U(rng) };
    modules::euler2d::conserved_array_t uc(nd::uivec(n, n));
    for (std::size_t i = 0; i < n; ++i)
        for (std::size_t j = 0; j < n; ++j) orc_euler_conserved_density(pe(i + 2, j + 2).data(), 1.4, uc(i, j).data());

    modules::euler2d::mesh_geometry_t geom{ 1.0, int(n) };
    mesh::block_index_t<2> block{ 1, { 1, 0 } };
    auto out = modules::euler2d::updated_conserved(uc, pe, 0.0, 1e-3, geom, block, 1.5, 1.4);
    std::vector<double> ref(n * n * 5);
    orc_euler2d_updated_conserved(flat(uc).data(), flat(pe).data(), int(n), int(n), 1e-3, geom.cell_spacing(block),
                                  1.5, 1.4, 0, ref.data());
    EXPECT(flat(out) == ref, "euler2d::updated_conserved equals the oracle exactly");

    auto p = modules::euler2d::recover_primitive_array(uc, 1.4);
    bool same = true;
    for (std::size_t z = 0; z < uc.size(); ++z) {
        double pp[5];
        orc_euler_recover_primitive(uc.data()[z].data(), 1.4, pp);
        for (int q = 0; q < 5; ++q) same &= pp[q] == p.data()[z][q];
    }
    EXPECT(same, "euler2d::recover_primitive_array equals the oracle exactly");

    auto g = modules::euler2d::estimate_gradient(pe, 1, 1.5);
    same = g.shape(0) == n + 4 && g.shape(1) == n + 2;
    for (std::size_t i = 0; same && i < n + 4; ++i)
        for (std::size_t j = 0; j < n + 2; ++j)
            for (int q = 0; q < 5; ++q)
                same &= g(i, j)[q] == orc_plm(pe(i, j)[q], pe(i, j + 1)[q], pe(i, j + 2)[q], 1.5);
    EXPECT(same, "euler2d::estimate_gradient equals PLM along axis 1");

    auto w = modules::euler2d::weighted_sum(uc, out, 0.5);
    same = true;
    for (std::size_t z = 0; z < uc.size(); ++z)
        for (int q = 0; q < 5; ++q) same &= w.data()[z][q] == uc.data()[z][q] * 0.5 + out.data()[z][q] * (1.0 - 0.5);
    EXPECT(same, "euler2d::weighted_sum is s*b + t*(1-b)");

    auto bad = uc;
    bad = modules::euler2d::conserved_array_t(uc.shape(), std::vector<euler::conserved_density_t>(uc.data(), uc.data() + uc.size()));
    bad(3, 5)[0] = -1.0;
    bool threw = false;
    try { modules::euler2d::recover_primitive_array(bad, 1.4); }
    catch (const std::invalid_argument& e) { threw = std::string(e.what()).find("negative density") != std::string::npos; }
    EXPECT(threw, "negative density throws std::invalid_argument like euler::recover_primitive");
}

static void test_euler1d(std::mt19937& rng)
{
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const std::size_t n = 64;
    modules::euler1d::primitive_array_t pe(nd::uivec(n + 4));
    for (std::size_t i = 0; i < n + 4; ++i) pe(i) = { 1.0 + 0.5 * U(rng), 0.3 * U(rng), 0.0, 0.0, 1.0 + 0.5 * U(rng) };
    modules::euler1d::conserved_array_t uc(nd::uivec(n));
    for (std::size_t i = 0; i < n; ++i) orc_euler_conserved_density(pe(i + 2).data(), 1.4, uc(i).data());
    modules::euler1d::mesh_geometry_t geom{ 1.0, int(n) };
    for (unsigned long c = 0; c < 4; ++c) {
        mesh::block_index_t<1> block{ 2, { c } };
        auto out = modules::euler1d::updated_conserved(uc, pe, 0.0, 1e-3, geom, modules::euler1d::boundary_condition_t::outflow,
                                                       block, 1.5, 1.4);
        std::vector<double> ref(n * 5);
        orc_euler1d_updated_conserved(flat(uc).data(), flat(pe).data(), int(n), 1e-3, geom.cell_spacing(block), 1.5, 1.4,
                                      c == 0, c == 3, ref.data());
        EXPECT(flat(out) == ref, "euler1d::updated_conserved (outflow by block position) equals the oracle exactly");
    }
}

static void test_locally_isothermal(std::mt19937& rng)
{
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const std::size_t n = 32;
    modules::locally_isothermal::primitive_array_t pe(nd::uivec(n + 4, n + 4));
    for (std::size_t i = 0; i < n + 4; ++i)
        for (std::size_t j = 0; j < n + 4; ++j) pe(i, j) = { 1.0 + 0.5 * U(rng), 0.4 * U(rng) - 0.2, 0.4 * U(rng) - 0.2 };
    modules::locally_isothermal::conserved_array_t uc(nd::uivec(n, n));
    for (std::size_t i = 0; i < n; ++i)
        for (std::size_t j = 0; j < n; ++j) orc_iso2d_conserved_density(pe(i + 2, j + 2).data(), uc(i, j).data());
    modules::locally_isothermal::mesh_geometry_t geom{ 1.0, int(n) };
    mesh::block_index_t<2> block{ 1, { 0, 1 } };
    auto temperature = [] (std::array<double, 2> x) { return 0.05 * (1.0 + x[0] * x[0] + 0.5 * x[1]); };
    auto out = modules::locally_isothermal::updated_conserved(uc, pe, 0.0, 1e-3, geom, block, 0.0, temperature, 1.5);
    std::vector<double> cx((n + 1) * n), cy(n * (n + 1));
    for (std::size_t i = 0; i <= n; ++i)
        for (std::size_t j = 0; j < n; ++j) cx[i * n + j] = temperature(geom.face_coordinate(block, 0, i, j));
    for (std::size_t i = 0; i < n; ++i)
        for (std::size_t j = 0; j <= n; ++j) cy[i * (n + 1) + j] = temperature(geom.face_coordinate(block, 1, i, j));
    std::vector<double> ref(n * n * 3);
    orc_locally_isothermal_updated_conserved(flat(uc).data(), flat(pe).data(), int(n), int(n), 1e-3,
                                             geom.cell_spacing(block), cx.data(), cy.data(), 0.0, ref.data());
    EXPECT(flat(out) == ref, "locally_isothermal::updated_conserved equals the oracle exactly");
}

static void test_minidisk(std::mt19937& rng)
{
    std::normal_distribution<double> N(0.0, 1.0);
    minidisk::solver_data_t sd{ 64, 2, 1.0, 10.0, 0.0, 8.0, 0.05, 0.0, 1.0, 0.05, 1e3, 1e2, 0.2, 0.3 };
    const std::size_t n = 64;
    orc_minidisk_t od{ 64, 2, 1.0, 10.0, 0.0, 8.0, 0.05, 0.0, 1.0, 0.05, 1e3, 1e2, 0.2, 0.3 };
    mesh::block_index_t<2> block{ 2, { 1, 2 } };
    minidisk::conserved_array_t uc(nd::uivec(n, n));
    orc_minidisk_initial_conserved(&od, 2, 1, 2, &uc.data()[0][0]);
    minidisk::primitive_array_t pe(nd::uivec(n + 4, n + 4));
    for (std::size_t z = 0; z < pe.size(); ++z) pe.data()[z] = { 1.0 + 0.1 * std::fabs(N(rng)), 0.3 * N(rng), 0.3 * N(rng) };
    auto out = minidisk::updated_conserved(uc, pe, 0.25, 1e-3, block, sd);
    std::vector<double> ref(n * n * 3);
    orc_minidisk_updated_conserved(flat(uc).data(), flat(pe).data(), &od, 2, 1, 2, 0.25, 1e-3, ref.data());
    auto got = flat(out);
    double scale[3] = { 0, 0, 0 }, err = 0;
    for (std::size_t z = 0; z < n * n; ++z)
        for (int q = 0; q < 3; ++q) scale[q] = std::fmax(scale[q], std::fabs(ref[z * 3 + q]));
    for (std::size_t z = 0; z < n * n; ++z)
        for (int q = 0; q < 3; ++q) err = std::fmax(err, std::fabs(got[z * 3 + q] - ref[z * 3 + q]) / scale[q]);
    std::printf("minidisk max relative error %.3e\n", err);
    EXPECT(err <= 1e-12, "minidisk::updated_conserved within 1e-12 of the oracle");
}

// mhd_scheme_v2 through the host API on a block of unequal extents against the
// oracle's restatement (src/scheme_mhd_v2.cpp:75-278)
static void test_mhd_v2(std::mt19937& rng)
{
    using namespace mhd_scheme_v2;
    std::uniform_real_distribution<double> U(0.5, 2.0);
    std::normal_distribution<double> N(0.0, 0.5);
    const std::size_t n1 = 4, n2 = 3, n3 = 5, m1 = n1 + 2, m2 = n2 + 2, m3 = n3 + 2;
    auto scalar = [&] (std::size_t a, std::size_t b, std::size_t c) {
        nd::shared_array<double, 3> x(nd::uivec(a, b, c));
        for (std::size_t z = 0; z < x.size(); ++z) x.data()[z] = N(rng);
        return x;
    };
    auto prim = [&] (std::size_t a, std::size_t b, std::size_t c) {
        cell_primitive_variables_t x(nd::uivec(a, b, c));
        for (std::size_t z = 0; z < x.size(); ++z)
            x.data()[z] = { U(rng), N(rng), N(rng), N(rng), U(rng), N(rng), N(rng), N(rng) };
        return x;
    };
    // primitive_array
    face_magnetic_flux_density_t bf = { scalar(n1 + 1, n2, n3), scalar(n1, n2 + 1, n3), scalar(n1, n2, n3 + 1) };
    auto p0 = prim(n1, n2, n3);
    cell_conserved_density_t uc(nd::uivec(n1, n2, n3));
    for (std::size_t z = 0; z < uc.size(); ++z) orc_mhd_conserved_density(p0.data()[z].data(), 5.0 / 3.0, uc.data()[z].data());
    auto pc = primitive_array(uc, bf);
    std::vector<double> ref(8 * n1 * n2 * n3);
    int st = orc_mhd_v2_primitive_array(int(n1), int(n2), int(n3), flat(uc).data(), bf[0].data(), bf[1].data(),
                                        bf[2].data(), ref.data());
    EXPECT(st == 0 && flat(pc) == ref, "mhd_scheme_v2::primitive_array equals the oracle exactly");
    // godunov_fluxes on an extended block
    auto pce = prim(m1, m2, m3);
    face_magnetic_flux_density_t bfe = { scalar(m1 - 1, m2, m3), scalar(m1, m2 - 1, m3), scalar(m1, m2, m3 - 1) };
    auto gf = godunov_fluxes(pce, bfe);
    std::vector<double> r1(8 * bfe[0].size()), r2(8 * bfe[1].size()), r3(8 * bfe[2].size());
    orc_mhd_v2_godunov_fluxes(int(m1), int(m2), int(m3), flat(pce).data(), bfe[0].data(), bfe[1].data(), bfe[2].data(),
                              r1.data(), r2.data(), r3.data());
    double err = 0, scale = 0;
    for (auto [g, r] : { std::pair{ &gf[0], &r1 }, std::pair{ &gf[1], &r2 }, std::pair{ &gf[2], &r3 } }) {
        auto v = flat(*g);
        for (std::size_t z = 0; z < v.size(); ++z) { err = std::fmax(err, std::fabs(v[z] - (*r)[z])); scale = std::fmax(scale, std::fabs((*r)[z])); }
    }
    std::printf("mhd_scheme_v2::godunov_fluxes max relative error %.3e\n", err / scale);
    EXPECT(err <= 1e-12 * scale, "mhd_scheme_v2::godunov_fluxes within 1e-12 of the oracle");
    // electromotive_forces and updated_conserved_density, fed the same Godunov data
    auto ee = electromotive_forces(gf);
    std::vector<double> e1(ee[0].size()), e2(ee[1].size()), e3(ee[2].size());
    auto f1 = flat(gf[0]), f2 = flat(gf[1]), f3 = flat(gf[2]);
    orc_mhd_v2_electromotive_forces(int(m1), int(m2), int(m3), f1.data(), f2.data(), f3.data(), e1.data(), e2.data(), e3.data());
    bool same = true;
    for (std::size_t z = 0; z < e1.size(); ++z) same &= ee[0].data()[z] == e1[z];
    for (std::size_t z = 0; z < e2.size(); ++z) same &= ee[1].data()[z] == e2[z];
    for (std::size_t z = 0; z < e3.size(); ++z) same &= ee[2].data()[z] == e3[z];
    EXPECT(same, "mhd_scheme_v2::electromotive_forces equals the oracle exactly");
    auto uo = updated_conserved_density(uc, gf, 0.01, 0.2);
    std::vector<double> ru(5 * uc.size());
    orc_mhd_v2_updated_conserved_density(int(n1), int(n2), int(n3), flat(uc).data(), f1.data(), f2.data(), f3.data(),
                                         0.01, 0.2, ru.data());
    EXPECT(flat(uo) == ru, "mhd_scheme_v2::updated_conserved_density equals the oracle exactly");
    auto bo = updated_magnetic_flux_density(bf, ee, 0.01, 0.2);
    std::vector<double> b1(bf[0].size()), b2(bf[1].size()), b3(bf[2].size());
    orc_mhd_v2_updated_magnetic_flux_density(int(n1), int(n2), int(n3), bf[0].data(), bf[1].data(), bf[2].data(),
                                             e1.data(), e2.data(), e3.data(), 0.01, 0.2, b1.data(), b2.data(), b3.data());
    same = true;
    for (std::size_t z = 0; z < b1.size(); ++z) same &= bo[0].data()[z] == b1[z];
    for (std::size_t z = 0; z < b2.size(); ++z) same &= bo[1].data()[z] == b2[z];
    for (std::size_t z = 0; z < b3.size(); ++z) same &= bo[2].data()[z] == b3[z];
    EXPECT(same, "mhd_scheme_v2::updated_magnetic_flux_density equals the oracle exactly");
    // extend_cell_primitive_variables over 27 distinct blocks
    std::vector<cell_primitive_variables_t> blocks;
    std::vector<std::vector<double>> flats;
    flats.reserve(27);
    const double* ptrs[27];
    for (int m = 0; m < 27; ++m) { blocks.push_back(prim(3, 3, 3)); flats.push_back(flat(blocks.back())); ptrs[m] = flats.back().data(); }
    auto ext = extend_cell_primitive_variables(blocks, 3, 2);
    std::vector<double> rx(8 * 7 * 7 * 7);
    orc_mhd_v2_extend_cells(ptrs, 3, 2, 8, rx.data());
    EXPECT(flat(ext) == rx, "mhd_scheme_v2::extend_cell_primitive_variables equals the oracle exactly");
}


// The reference calls the per-block functions concurrently, one block per
// thread-pool worker (src/parallel_computable.cpp:248-296).  8 threads, each
// on its own stream, make 32 calls each (recover_primitive_array then
// updated_conserved of an euler2d block, or a minidisk block); one call of
// one thread recovers a block holding a negative density.  Only that call
// may throw, with the reference's message, and every other result must equal
// the single-threaded oracle (euler2d exactly, minidisk within 1e-12).
static void test_concurrent_callers(std::mt19937& rng)
{
    std::uniform_real_distribution<double> U(0.0, 1.0);
    std::normal_distribution<double> N(0.0, 1.0);
    const std::size_t n = 32;
    const int kBlocks = 4, kThreads = 8, kCalls = 32, kBadThread = 3, kBadCall = 7;
    std::vector<modules::euler2d::conserved_array_t> uc;
    std::vector<modules::euler2d::primitive_array_t> pe;
    std::vector<std::vector<double>> ref_out, ref_p;
    modules::euler2d::mesh_geometry_t geom{ 1.0, int(n) };
    mesh::block_index_t<2> block{ 2, { 1, 3 } };
    for (int k = 0; k < kBlocks; ++k) {
        modules::euler2d::primitive_array_t p(nd::uivec(n + 4, n + 4));
        for (std::size_t z = 0; z < p.size(); ++z)
            p.data()[z] = { 1.0 + 0.5 * U(rng), 0.3 * N(rng), 0.3 * N(rng), 0.1 * N(rng), 1.0 + 0.5 * U(rng) };
        modules::euler2d::conserved_array_t u(nd::uivec(n, n));
        for (std::size_t i = 0; i < n; ++i)
            for (std::size_t j = 0; j < n; ++j) orc_euler_conserved_density(p(i + 2, j + 2).data(), 1.4, u(i, j).data());
        std::vector<double> r(n * n * 5), rp(n * n * 5);
        orc_euler2d_updated_conserved(flat(u).data(), flat(p).data(), int(n), int(n), 1e-3, geom.cell_spacing(block),
                                      1.5, 1.4, 0, r.data());
        for (std::size_t z = 0; z < u.size(); ++z) orc_euler_recover_primitive(u.data()[z].data(), 1.4, &rp[5 * z]);
        uc.push_back(u); pe.push_back(p); ref_out.push_back(r); ref_p.push_back(rp);
    }
    auto poisoned = modules::euler2d::conserved_array_t(uc[0].shape(), std::vector<euler::conserved_density_t>(uc[0].data(), uc[0].data() + uc[0].size()));
    poisoned(5, 9)[0] = -1.0;

    minidisk::solver_data_t sd{ int(n), 2, 1.0, 10.0, 0.0, 8.0, 0.05, 0.0, 1.0, 0.05, 1e3, 1e2, 0.2, 0.3 };
    orc_minidisk_t od{ int(n), 2, 1.0, 10.0, 0.0, 8.0, 0.05, 0.0, 1.0, 0.05, 1e3, 1e2, 0.2, 0.3 };
    mesh::block_index_t<2> mblock{ 2, { 2, 1 } };
    minidisk::conserved_array_t muc(nd::uivec(n, n));
    orc_minidisk_initial_conserved(&od, 2, 2, 1, &muc.data()[0][0]);
    minidisk::primitive_array_t mpe(nd::uivec(n + 4, n + 4));
    for (std::size_t z = 0; z < mpe.size(); ++z) mpe.data()[z] = { 1.0 + 0.1 * std::fabs(N(rng)), 0.3 * N(rng), 0.3 * N(rng) };
    std::vector<double> mref(n * n * 3);
    orc_minidisk_updated_conserved(flat(muc).data(), flat(mpe).data(), &od, 2, 2, 1, 0.25, 1e-3, mref.data());
    double mscale[3] = { 0, 0, 0 };
    for (std::size_t z = 0; z < n * n; ++z)
        for (int q = 0; q < 3; ++q) mscale[q] = std::fmax(mscale[q], std::fabs(mref[z * 3 + q]));

    std::atomic<int> mismatches{ 0 }, unexpected{ 0 }, planted_thrown{ 0 }, calls_done{ 0 };
    std::vector<std::thread> pool;
    for (int t = 0; t < kThreads; ++t)
        pool.emplace_back([&, t] {
            hipStream_t s = nullptr;
            if (hipStreamCreate(&s) != hipSuccess) { ++unexpected; return; }
            for (int c = 0; c < kCalls; ++c) {
                const int k = (t + c) % kBlocks;
                try {
                    if (t == kBadThread && c == kBadCall) {
                        modules::euler2d::recover_primitive_array(poisoned, 1.4, s);
                        ++unexpected;                       // did not throw
                        continue;
                    }
                    if ((t + c) % 2 == 0) {
                        auto p = modules::euler2d::recover_primitive_array(uc[k], 1.4, s);
                        if (flat(p) != ref_p[k]) ++mismatches;
                        auto o = modules::euler2d::updated_conserved(uc[k], pe[k], 0.0, 1e-3, geom, block, 1.5, 1.4, s);
                        if (flat(o) != ref_out[k]) ++mismatches;
                    } else {
                        auto o = flat(minidisk::updated_conserved(muc, mpe, 0.25, 1e-3, mblock, sd, s));
                        double err = 0;
                        for (std::size_t z = 0; z < n * n; ++z)
                            for (int q = 0; q < 3; ++q) err = std::fmax(err, std::fabs(o[z * 3 + q] - mref[z * 3 + q]) / mscale[q]);
                        if (!(err <= 1e-12)) ++mismatches;
                    }
                    ++calls_done;
                } catch (const std::invalid_argument& e) {
                    if (t == kBadThread && c == kBadCall && std::string(e.what()) == "euler::recover_primitive (negative density)")
                        ++planted_thrown;
                    else
                        ++unexpected;
                } catch (...) {
                    ++unexpected;
                }
            }
            (void)hipStreamDestroy(s);
        });
    for (auto& th : pool) th.join();
    std::printf("concurrent callers: %d calls, %d mismatches, %d unexpected, planted error thrown %d\n",
                calls_done.load(), mismatches.load(), unexpected.load(), planted_thrown.load());
    EXPECT(planted_thrown == 1 && unexpected == 0, "8 threads on 8 streams: only the planted negative density throws");
    EXPECT(mismatches == 0 && calls_done == kThreads * kCalls - 1, "8 threads on 8 streams: every other result equals the oracle");

    // the same on the null stream, shared by all threads
    std::atomic<int> null_bad{ 0 }, null_thrown{ 0 };
    pool.clear();
    for (int t = 0; t < kThreads; ++t)
        pool.emplace_back([&, t] {
            for (int c = 0; c < 8; ++c) {
                try {
                    if (t == kBadThread && c == 2) { modules::euler2d::recover_primitive_array(poisoned, 1.4); ++null_bad; continue; }
                    auto o = modules::euler2d::updated_conserved(uc[c % kBlocks], pe[c % kBlocks], 0.0, 1e-3, geom, block, 1.5, 1.4);
                    if (flat(o) != ref_out[c % kBlocks]) ++null_bad;
                    auto p = modules::euler2d::recover_primitive_array(uc[c % kBlocks], 1.4);
                    if (flat(p) != ref_p[c % kBlocks]) ++null_bad;
                } catch (const std::invalid_argument& e) {
                    if (t == kBadThread && c == 2) ++null_thrown; else ++null_bad;
                } catch (...) { ++null_bad; }
            }
        });
    for (auto& th : pool) th.join();
    EXPECT(null_bad == 0 && null_thrown == 1, "8 threads on the null stream: results exact, only the planted call throws");
}

// updated_solution / updated_conserved_tree over whole trees against the
// oracle's per-block updates with the guard fill of amr::extend_block.
static void test_tree_functions(std::mt19937& rng)
{
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const int nb = 4;
    const std::size_t n = 16;
    // euler2d: uniform level-2 tree of 4 x 4 blocks
    {
        modules::euler2d::mesh_geometry_t geom{ 1.0, int(n) };
        modules::euler2d::solution_t sol{ 3.0, 0.125, {} };
        std::vector<double> all(nb * nb * n * n * 5);
        for (int b = 0; b < nb * nb; ++b) {
            modules::euler2d::conserved_array_t u(nd::uivec(n, n));
            for (std::size_t z = 0; z < u.size(); ++z) {
                double p[5] = { 1.0 + 0.5 * U(rng), 0.3 * U(rng) - 0.15, 0.3 * U(rng) - 0.15, 0.0, 1.0 + 0.5 * U(rng) };
                orc_euler_conserved_density(p, 1.4, u.data()[z].data());
            }
            std::memcpy(&all[b * n * n * 5], u.data(), n * n * 5 * sizeof(double));
            sol.conserved.emplace(mesh::block_index_t<2>{ 2, { (unsigned long)(b / nb), (unsigned long)(b % nb) } }, u);
        }
        const double dt = 1e-3;
        auto next = modules::euler2d::updated_solution(sol, dt, geom, 1.5, 1.4);
        std::vector<double> ref(all.size());
        orc_euler2d_mesh_stage(all.data(), ref.data(), nb, int(n), dt, geom.cell_spacing({ 2, { 0, 0 } }), 1.5, 1.4, 0, 1);
        std::vector<double> got;
        for (const auto& kv : next.conserved) { auto f = flat(kv.second); got.insert(got.end(), f.begin(), f.end()); }
        EXPECT(got == ref && next.iteration == 4.0 && next.time == 0.125 + dt,
               "euler2d::updated_solution (uniform tree) equals the oracle's mesh stage exactly");
        auto topo = bsp::indexes(sol.conserved);
        EXPECT(modules::euler2d::total_zones(topo, geom) == std::size_t(nb * nb) * n * n, "euler2d::total_zones");
        EXPECT(modules::euler2d::smallest_cell_size(topo, geom) == geom.cell_spacing({ 2, { 0, 0 } }), "euler2d::smallest_cell_size");
        auto half = modules::euler2d::weighted_sum(sol, next, 0.5);
        EXPECT(half.iteration == 3.5 && half.time == 0.125 * 0.5 + (0.125 + dt) * 0.5, "euler2d::weighted_sum of solutions");
        // a level jump: a uniform state is a fixed point of the quad-tree stage
        modules::euler2d::solution_t flat_sol{ 0.0, 0.0, {} };
        euler::conserved_density_t u0;
        double p0[5] = { 1.3, 0.2, -0.1, 0.0, 0.9 };
        orc_euler_conserved_density(p0, 1.4, u0.data());
        for (unsigned long i = 0; i < 2; ++i)
            for (unsigned long j = 0; j < 2; ++j) {
                if (i == 0 && j == 0) {
                    for (unsigned long a = 0; a < 2; ++a)
                        for (unsigned long c = 0; c < 2; ++c)
                            flat_sol.conserved.emplace(mesh::block_index_t<2>{ 2, { a, c } },
                                modules::euler2d::conserved_array_t(nd::uivec(n, n), std::vector<euler::conserved_density_t>(n * n, u0)));
                } else {
                    flat_sol.conserved.emplace(mesh::block_index_t<2>{ 1, { i, j } },
                        modules::euler2d::conserved_array_t(nd::uivec(n, n), std::vector<euler::conserved_density_t>(n * n, u0)));
                }
            }
        auto moved = modules::euler2d::updated_solution(flat_sol, 1e-3, geom, 1.5, 1.4);
        bool fixed = true;
        for (const auto& kv : moved.conserved)
            for (std::size_t z = 0; z < n * n; ++z) fixed &= kv.second.data()[z] == u0;
        EXPECT(fixed, "euler2d::updated_solution (tree with a level jump): a uniform state is a fixed point");
    }
    // euler1d: chain of 8 blocks, outflow ends
    {
        const int nb1 = 8;
        modules::euler1d::mesh_geometry_t geom{ 1.0, int(n) };
        modules::euler1d::solution_t sol{ 0.0, 0.0, {} };
        std::vector<double> all(nb1 * n * 5), P(nb1 * n * 5);
        for (int b = 0; b < nb1; ++b) {
            modules::euler1d::conserved_array_t u(nd::uivec(n));
            for (std::size_t z = 0; z < n; ++z) {
                double p[5] = { 1.0 + 0.5 * U(rng), 0.3 * U(rng) - 0.15, 0.0, 0.0, 1.0 + 0.5 * U(rng) };
                orc_euler_conserved_density(p, 1.4, u.data()[z].data());
                std::memcpy(&P[(b * n + z) * 5], p, sizeof(p));
            }
            std::memcpy(&all[b * n * 5], u.data(), n * 5 * sizeof(double));
            sol.conserved.emplace(mesh::block_index_t<1>{ 3, { (unsigned long)b } }, u);
        }
        const double dt = 1e-3, dl = geom.cell_spacing({ 3, { 0 } });
        auto next = modules::euler1d::updated_solution(sol, dt, geom, modules::euler1d::boundary_condition_t::outflow, 1.5, 1.4);
        std::vector<double> ref(all.size()), pe((n + 4) * 5);
        for (int b = 0; b < nb1; ++b) {
            std::vector<double> Pr(all.size());
            for (std::size_t z = 0; z < nb1 * n; ++z) orc_euler_recover_primitive(&all[z * 5], 1.4, &Pr[z * 5]);
            orc_extend_block_1d(Pr.data(), nb1, int(n), 5, b, pe.data());
            orc_euler1d_updated_conserved(&all[b * n * 5], pe.data(), int(n), dt, dl, 1.5, 1.4, b == 0, b == nb1 - 1, &ref[b * n * 5]);
        }
        std::vector<double> got;
        for (const auto& kv : next.conserved) { auto f = flat(kv.second); got.insert(got.end(), f.begin(), f.end()); }
        EXPECT(got == ref, "euler1d::updated_solution (outflow chain) equals the oracle's per-block updates exactly");
        EXPECT(modules::euler1d::total_zones(bsp::indexes(sol.conserved), geom) == nb1 * n, "euler1d::total_zones");
    }
    // locally_isothermal: uniform level-2 tree, temperature mapping at the faces
    {
        modules::locally_isothermal::mesh_geometry_t geom{ 1.0, int(n) };
        modules::locally_isothermal::conserved_tree_t tree;
        std::vector<double> all(nb * nb * n * n * 3), Pall(all.size());
        for (int b = 0; b < nb * nb; ++b) {
            modules::locally_isothermal::conserved_array_t u(nd::uivec(n, n));
            for (std::size_t z = 0; z < u.size(); ++z) {
                double p[3] = { 1.0 + 0.5 * U(rng), 0.4 * U(rng) - 0.2, 0.4 * U(rng) - 0.2 };
                orc_iso2d_conserved_density(p, u.data()[z].data());
            }
            std::memcpy(&all[b * n * n * 3], u.data(), n * n * 3 * sizeof(double));
            tree.emplace(mesh::block_index_t<2>{ 2, { (unsigned long)(b / nb), (unsigned long)(b % nb) } }, u);
        }
        for (std::size_t z = 0; z < all.size() / 3; ++z) orc_iso2d_recover_primitive(&all[z * 3], &Pall[z * 3]);
        auto temperature = [] (std::array<double, 2> x) { return 0.05 * (1.0 + x[0] * x[0] + 0.5 * x[1]); };
        const double dt = 1e-3;
        for (double nu : { 0.0, 0.01 }) {
            auto next = modules::locally_isothermal::updated_conserved_tree(tree, 0.0, dt, geom, temperature, nu, 2.0);
            std::vector<double> ref(all.size()), pe((n + 4) * (n + 4) * 3), cx((n + 1) * n), cy(n * (n + 1));
            for (int b = 0; b < nb * nb; ++b) {
                mesh::block_index_t<2> blk{ 2, { (unsigned long)(b / nb), (unsigned long)(b % nb) } };
                orc_extend_block_2d(Pall.data(), nb, int(n), 3, b / nb, b % nb, pe.data());
                for (std::size_t i = 0; i <= n; ++i)
                    for (std::size_t j = 0; j < n; ++j) cx[i * n + j] = temperature(geom.face_coordinate(blk, 0, i, j));
                for (std::size_t i = 0; i < n; ++i)
                    for (std::size_t j = 0; j <= n; ++j) cy[i * (n + 1) + j] = temperature(geom.face_coordinate(blk, 1, i, j));
                orc_locally_isothermal_updated_conserved(&all[b * n * n * 3], pe.data(), int(n), int(n), dt,
                                                         geom.cell_spacing(blk), cx.data(), cy.data(), nu, &ref[b * n * n * 3]);
            }
            std::vector<double> got;
            for (const auto& kv : next) { auto f = flat(kv.second); got.insert(got.end(), f.begin(), f.end()); }
            EXPECT(got == ref, nu == 0.0 ? "locally_isothermal::updated_conserved_tree equals the oracle exactly"
                                         : "locally_isothermal::updated_conserved_tree (viscous) equals the oracle exactly");
        }
    }
}

// euler2d::updated_solution on a 2:1-balanced quad-tree with level jumps:
// the leaves, input and output go to a file for the Python AMR oracle.
static int dump_amr(const char* path)
{
    std::mt19937 rng(99);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const std::size_t n = 16;
    // level 2, except the level-1 block (1, 1) over the level-2 blocks
    // i, j in {2, 3}, and the level-2 block (0, 0) split into four level-3
    // blocks: level jumps both ways, 2:1 balanced, periodic
    std::vector<mesh::block_index_t<2>> leaves;
    leaves.push_back({ 1, { 1, 1 } });
    for (unsigned long i = 0; i < 4; ++i)
        for (unsigned long j = 0; j < 4; ++j)
            if (!(i >= 2 && j >= 2) && !(i == 0 && j == 0)) leaves.push_back({ 2, { i, j } });
    for (unsigned long i = 0; i < 2; ++i)
        for (unsigned long j = 0; j < 2; ++j) leaves.push_back({ 3, { i, j } });
    modules::euler2d::solution_t sol{ 0.0, 0.0, {} };
    for (const auto& b : leaves) {
        modules::euler2d::conserved_array_t u(nd::uivec(n, n));
        for (std::size_t z = 0; z < u.size(); ++z) {
            double p[5] = { 1.0 + 0.5 * U(rng), 0.3 * U(rng) - 0.15, 0.3 * U(rng) - 0.15, 0.05 * U(rng), 1.0 + 0.5 * U(rng) };
            orc_euler_conserved_density(p, 1.4, u.data()[z].data());
        }
        sol.conserved.emplace(b, u);
    }
    modules::euler2d::mesh_geometry_t geom{ 1.0, int(n) };
    auto next = modules::euler2d::updated_solution(sol, 1e-3, geom, 1.5, 1.4);
    FILE* f = std::fopen(path, "wb");
    if (!f) return 1;
    const long head[2] = { long(sol.conserved.size()), long(n) };
    std::fwrite(head, sizeof(long), 2, f);
    for (const auto& kv : sol.conserved) {
        const long idx[3] = { long(kv.first.level), long(kv.first.coordinates[0]), long(kv.first.coordinates[1]) };
        std::fwrite(idx, sizeof(long), 3, f);
    }
    for (const auto& kv : sol.conserved) std::fwrite(kv.second.data(), sizeof(double), n * n * 5, f);
    for (const auto& kv : next.conserved) std::fwrite(kv.second.data(), sizeof(double), n * n * 5, f);
    std::fclose(f);
    std::printf("wrote %zu leaves\n", sol.conserved.size());
    return 0;
}

int main(int argc, char** argv)
{
    if (argc == 3 && std::strcmp(argv[1], "--dump-amr") == 0) return dump_amr(argv[2]);
    std::mt19937 rng(1234);
    test_euler2d(rng);
    test_euler1d(rng);
    test_locally_isothermal(rng);
    test_minidisk(rng);
    test_mhd_v2(rng);
    test_concurrent_callers(rng);
    test_tree_functions(rng);
    if (failures) { std::printf("%d FAILURES\n", failures); return 1; }
    std::printf("ALL PASS\n");
    return 0;
}
