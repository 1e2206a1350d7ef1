// C ABI of the MI355X Godunov block update (declarations: include/mara3p_amd.h).
// The stage kernels are compiled in the k_*.hip units (launch.hpp); this unit
// holds the entry points, the library state and the small kernels.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>
#include "launch.hpp"
#include "amr.hpp"
#include "elementwise.hpp"
#include "mhd_v2.hpp"
#include "srhd.hpp"
#include "two_body.hpp"

namespace m3p {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

int hip_check(hipError_t e, const char* what)
{
    if (e == hipSuccess) return M3P_OK;
    return fail(M3P_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

int launch_status(const char* what)
{
    return hip_check(hipGetLastError(), what);
}

} // namespace m3p

using namespace m3p;

namespace {

// ---------------------------------------------------------------------------
// Library state.  Kernels report errors (negative density, ...) into a latch
// of four ints, and the reductions keep partial results in a scratch buffer.
// Both are per stream: an asynchronous call uses the latch and scratch of its
// stream, so callers on different streams never see each other's errors and
// never share reduction slots, and calls on one stream are ordered by it.
// The blocking per-block calls instead lease a latch of their own from a
// per-device pool for the duration of the call, so they are safe from any
// number of host threads on any stream, the null stream included.
constexpr int kRedWords = 4096;
// reduction scratch: words [0, 2048) totals partials, then single slots
constexpr int kSlotDivB = 3000;
constexpr int kSlotCflRate = 3008;           // kRateSlots words

struct StreamState {
    int* err = nullptr;                  // int[4]: code, block, zone
    unsigned long long* red = nullptr;   // kRedWords words
    int cus = 0;                         // compute units of the device
};

struct DeviceState {
    int cus = 0;
    std::unordered_map<hipStream_t, StreamState> streams;
    std::vector<int*> free_latches;
};

constexpr int kMaxDevices = 64;
DeviceState g_dev[kMaxDevices];
std::mutex g_mutex;                          // guards g_dev

int current_device(int* dev)
{
    if (hipGetDevice(dev) != hipSuccess || *dev < 0 || *dev >= kMaxDevices)
        return fail(M3P_ERR_HIP, "m3p: no current HIP device");
    return M3P_OK;
}

// A zeroed latch on the current device (synchronous copy, no memset node).
int new_latch(int** out)
{
    const int zero[4] = { 0, 0, 0, 0 };
    if (int st = hip_check(hipMalloc(out, sizeof(zero)), "hipMalloc(error latch)")) return st;
    return hip_check(hipMemcpy(*out, zero, sizeof(zero), hipMemcpyHostToDevice), "hipMemcpy(error latch)");
}

int device_cus(DeviceState& d, int dev)
{
    if (d.cus == 0)
        if (int st = hip_check(hipDeviceGetAttribute(&d.cus, hipDeviceAttributeMultiprocessorCount, dev),
                               "hipDeviceGetAttribute")) return st;
    return M3P_OK;
}

// The state of (current device, stream), created on first use.  Creating it
// allocates, which a stream under graph capture must not do: the first call
// on a stream has to come before its capture begins (m3p_graph_begin does it).
int stream_state(void* stream, StreamState** out)
{
    hipStream_t s = (hipStream_t)stream;
    int dev;
    if (int st = current_device(&dev)) return st;
    std::lock_guard<std::mutex> lock(g_mutex);
    DeviceState& d = g_dev[dev];
    auto it = d.streams.find(s);
    if (it != d.streams.end()) { *out = &it->second; return M3P_OK; }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (s && hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone)
        return fail(M3P_ERR_BAD_ARGUMENT, "m3p: the first call on a stream must come before its graph capture "
                                          "(begin captures with m3p_graph_begin)");
    if (int st = device_cus(d, dev)) return st;
    StreamState ss;
    ss.cus = d.cus;
    if (int st = new_latch(&ss.err)) return st;
    if (int st = hip_check(hipMalloc(&ss.red, kRedWords * sizeof(unsigned long long)), "hipMalloc(reduction)")) {
        (void)hipFree(ss.err);
        return st;
    }
    *out = &(d.streams[s] = ss);
    return M3P_OK;
}

// The latch of one blocking call: leased from the device's pool, read after
// the call's kernels have finished, cleared if set, and given back.
struct Lease {
    int* latch = nullptr;
    int cus = 0;
    int dev = -1;
    bool pending = false;      // kernels may still write the latch
    hipStream_t stream = nullptr;

    int acquire(void* s)
    {
        stream = (hipStream_t)s;
        if (int st = current_device(&dev)) return st;
        {
            std::lock_guard<std::mutex> lock(g_mutex);
            DeviceState& d = g_dev[dev];
            if (int st = device_cus(d, dev)) return st;
            cus = d.cus;
            if (!d.free_latches.empty()) {
                latch = d.free_latches.back();
                d.free_latches.pop_back();
            }
        }
        if (!latch)
            if (int st = new_latch(&latch)) { latch = nullptr; return st; }
        pending = true;
        return M3P_OK;
    }
    // Wait for the call's kernels; the latch's code, block and zone.
    int wait(int* code, int* block, int* zone)
    {
        int h[4] = { 0, 0, 0, 0 };
        if (int st = hip_check(hipMemcpyAsync(h, latch, sizeof(h), hipMemcpyDeviceToHost, stream), "hipMemcpyAsync(error latch)")) return st;
        if (int st = hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize")) return st;
        pending = false;
        if (h[0] != 0) {
            const int zero[4] = { 0, 0, 0, 0 };
            if (int st = hip_check(hipMemcpy(latch, zero, sizeof(zero), hipMemcpyHostToDevice), "hipMemcpy(error latch)")) return st;
        }
        if (code) *code = h[0];
        if (block) *block = h[1];
        if (zone) *zone = h[2];
        return M3P_OK;
    }
    // wait(), with a latched negative density reported as the reference throws it.
    int finish()
    {
        int code = 0, block = 0, zone = 0;
        if (int st = wait(&code, &block, &zone)) return st;
        if (code == M3P_ERR_NEGATIVE_DENSITY) return fail(code, "euler::recover_primitive (negative density)");
        if (code) return fail(code, "m3p: kernel error " + std::to_string(code));
        return M3P_OK;
    }
    ~Lease()
    {
        if (!latch) return;
        if (pending) {
            // a launch failed after an earlier one of the call was queued:
            // wait for it and clear what it may have written
            if (wait(nullptr, nullptr, nullptr) != M3P_OK) return;     // unusable: not returned to the pool
            const int zero[4] = { 0, 0, 0, 0 };
            if (hipMemcpy(latch, zero, sizeof(zero), hipMemcpyHostToDevice) != hipSuccess) return;
        }
        std::lock_guard<std::mutex> lock(g_mutex);
        g_dev[dev].free_latches.push_back(latch);
    }
};

// ---------------------------------------------------------------------------
// 2D launch dispatch.  The variant used for each (scheme kind, precision) is
// a per-process setting (an index into kTileVariants) chosen by
// m3p_set_stage2d_tile or m3p_autotune_*; it is read and written atomically.
constexpr int kNumKinds = 3;
constexpr int tile_index(int tj, int h)
{
    for (int i = 0; i < kNumVariants; ++i)
        if (kTileVariants[i].tj == tj && kTileVariants[i].h == h) return i;
    return -1;
}
// [kind][precision]: defaults from sweeps of the benchmark meshes; the tree
// kind runs 64^2 leaves, where 64 x 16 tiles give the most workgroups.
std::atomic<int> g_tile[kNumKinds][2] = { { tile_index(256, 16), tile_index(256, 16) },
                                          { tile_index(256, 16), tile_index(256, 16) },
                                          { tile_index(64, 16), tile_index(64, 16) } };

bool tile_fits(TileShape v, int n) { return n % v.tj == 0 && n % v.h == 0 && v.h <= n; }

TileShape pick_tile(int kind, int precision, int n)
{
    TileShape want = kTileVariants[g_tile[kind][precision].load(std::memory_order_relaxed)];
    if (tile_fits(want, n)) return want;
    // widest fitting tile, then the fewest rows (most workgroups)
    TileShape best{0, 0};
    for (auto v : kTileVariants)
        if (tile_fits(v, n) && (v.tj > best.tj || (v.tj == best.tj && v.h < best.h))) best = v;
    return best;
}

template<class S, class Ld>
int launch_stage2d(int kind, const typename S::Params& prm, const Ld& ld, const StageArgs& a, hipStream_t s)
{
    const int prec = is_fp64<typename S::Real>::value ? M3P_FP64 : M3P_MIXED;
    const TileShape ts = pick_tile(kind, prec, a.n);
    if (ts.tj == 0) return fail(M3P_ERR_BAD_SHAPE, "m3p: 2D block size must be 16, 32 or a multiple of 64 (got " + std::to_string(a.n) + ")");
    return launch_stage2d_shape<S, Ld>(ts, prm, ld, a, s);
}

// Time every tile variant that fits on the caller's arrays and keep the fastest.
template<class S, class Ld>
int autotune_stage2d(int kind, const typename S::Params& prm, const Ld& ld, const StageArgs& a, hipStream_t s,
                     int reps, int* best_tj, int* best_h, float* best_ms)
{
    const int prec = is_fp64<typename S::Real>::value ? M3P_FP64 : M3P_MIXED;
    hipEvent_t e0, e1;
    if (int st = hip_check(hipEventCreate(&e0), "hipEventCreate")) return st;
    if (int st = hip_check(hipEventCreate(&e1), "hipEventCreate")) { (void)hipEventDestroy(e0); return st; }
    int best = -1;
    float bms = 1e30f;
    int st = M3P_OK;
    for (int i = 0; i < kNumVariants && !st; ++i) {
        const TileShape v = kTileVariants[i];
        if (!tile_fits(v, a.n)) continue;
        if ((st = launch_stage2d_shape<S, Ld>(v, prm, ld, a, s))) break;       // warm-up
        (void)hipEventRecord(e0, s);
        for (int r = 0; r < reps && !st; ++r) st = launch_stage2d_shape<S, Ld>(v, prm, ld, a, s);
        if (st) break;
        (void)hipEventRecord(e1, s);
        if ((st = hip_check(hipEventSynchronize(e1), "hipEventSynchronize"))) break;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        ms /= reps;
        if (ms < bms) { bms = ms; best = i; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st) return st;
    if (best < 0) return fail(M3P_ERR_BAD_SHAPE, "m3p: no tile variant fits block size " + std::to_string(a.n));
    g_tile[kind][prec].store(best, std::memory_order_relaxed);
    if (best_tj) *best_tj = kTileVariants[best].tj;
    if (best_h) *best_h = kTileVariants[best].h;
    if (best_ms) *best_ms = bms;
    return M3P_OK;
}

MinidiskParams minidisk_params(const m3p_minidisk_data* d, double time)
{
    MinidiskParams p;
    p.R = d->domain_radius;
    p.rs = d->softening_length;
    p.s2 = std::pow(d->softening_length, 2.0);
    p.mach = d->mach_number;
    p.inv_mach2 = 1.0 / (d->mach_number * d->mach_number);
    p.omega = d->omega_frame;
    p.sink_rate = d->sink_rate;
    p.s6 = std::pow(d->sink_radius, 6.0);
    p.inv_s6 = 1.0 / p.s6;
    p.q_sink_cut = std::cbrt(104.0 * p.s6);
    p.buffer_rate = d->buffer_rate;
    p.buffer_scale = d->buffer_scale;
    p.inv_buffer_scale = 1.0 / d->buffer_scale;
    double st[10];
    two_body_state(1.0, 1.0, d->mass_ratio, d->eccentricity, time, d->omega_frame, st);
    p.m1 = st[0]; p.x1 = st[1]; p.y1 = st[2];
    p.m2 = st[5]; p.x2 = st[6]; p.y2 = st[7];
    p.orbit = nullptr;
    p.level = d->depth;
    return p;
}

int check_minidisk_data(const m3p_minidisk_data* d)
{
    if (!d) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: minidisk data is NULL");
    if (!(d->kinematic_viscosity >= 0.0)) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: negative kinematic viscosity");
    if (d->depth < 0 || d->depth > 20) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: bad minidisk depth");
    return M3P_OK;
}

// Reductions
// CFL rate: max over zones of (max(|v1|, |v2|) + c_s) / dl_b, with the
// reciprocal 1 / dl_b of the zone's block (dl_blk) or of the mesh (dl).  The
// stage kernels with a fused reduction compute the same rate of the values
// they write.  Positive doubles order like their bit patterns, so the
// cross-workgroup atomic max is exact and deterministic.
__global__ void __launch_bounds__(256)
euler_max_rate_kernel(const double* u, int nblk, long zpb, const double* dl_blk, double dl, EulerParams prm,
                      unsigned long long* slots)
{
    double m = 0.0;
    const long total = (long)nblk * zpb;
    for (long z = (long)blockIdx.x * blockDim.x + threadIdx.x; z < total; z += (long)gridDim.x * blockDim.x) {
        const long b = z / zpb, i = z % zpb;
        const double* base = u + b * 5 * zpb + i;
        double uu[5] = { base[0], base[zpb], base[2 * zpb], base[3 * zpb], base[4 * zpb] };
        const double idl = 1.0 / (dl_blk ? dl_blk[b] : dl);
        m = fmax(m, Euler<double>::max_speed(uu, prm) * idl);
    }
    workgroup_atomic_max<256>(m, slots);
}

// dt = cfl / max rate (over the kRateSlots partial maxima; one wave)
__global__ void __launch_bounds__(64) cfl_from_rate_kernel(const unsigned long long* slots, double cfl, double* dt_out)
{
    double m = __longlong_as_double((long long)slots[threadIdx.x]);
    m = wave_max(m);
    if (threadIdx.x == 0) *dt_out = cfl / m;
}

__global__ void copy_u64_as_double_kernel(const unsigned long long* slot, double* out)
{
    *out = __longlong_as_double((long long)*slot);
}

// Two-level deterministic sum: per-workgroup partials (fixed order), then
// one workgroup adds the partials in index order.
__global__ void __launch_bounds__(256)
totals_partial_kernel(const double* u, int nblk, int nc, long zpb, double* partial)
{
    // grid: (nparts, nc)
    __shared__ double part[4];
    const int q = blockIdx.y;
    const long total = (long)nblk * zpb;
    double s = 0.0;
    for (long z = (long)blockIdx.x * blockDim.x + threadIdx.x; z < total; z += (long)gridDim.x * blockDim.x) {
        const long b = z / zpb, i = z % zpb;
        s += u[(b * nc + q) * zpb + i];
    }
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double v = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) v += part[w];
        partial[(size_t)q * gridDim.x + blockIdx.x] = v;
    }
}

__global__ void __launch_bounds__(64)
totals_final_kernel(const double* partial, int nparts, double* totals)
{
    const int q = blockIdx.x;
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += 64) s += partial[(size_t)q * nparts + i];
    s = wave_sum(s);
    if (threadIdx.x == 0) totals[q] = s;
}


} // namespace

// ---------------------------------------------------------------------------
// Static scratch of m3p_mhd_advance_ct_noalloc: the MHD workspace for callers
// that cannot allocate.  One copy per GPU, held until the process ends.  Host
// code never passes the array's name to a kernel (on the host it names a
// stand-in, not device memory): each device's address comes from
// hipGetSymbolAddress, taken once per device with that device current.
constexpr long kScratchBytes = 48l << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

void* g_scratch_ptr[kMaxDevices] = {};

int scratch_address(void** out)
{
    int dev;
    if (int st = current_device(&dev)) return st;
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        if (int st = hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)")) return st;
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return M3P_OK;
}

// One minidisk step's clock (one thread): dt from the reference's rule at
// t = clock[0] (problems/minidisk.cpp:202-206: smallest_cell_size / vmax *
// cfl * min(1, 0.01 + t / 0.05), evaluated left to right), the binary's
// state at t and at t + dt (the second stage's time) into orbit[0..5] and
// orbit[6..11], and the time after the step into clock[0] (the RK2
// combination t * 0.5 + (t + dt) * 0.5 of the solution time, or t + dt).
__global__ void minidisk_clock_kernel(double dx_over_vmax, double cfl, double q, double e, double omega,
                                      int rk_order, double* clock, double* orbit)
{
    if (threadIdx.x != 0) return;
    const double t = clock[0];
    const double dt = dx_over_vmax * (cfl * fmin(1.0, 0.01 + t / 0.05));
    double st[10];
    for (int k = 0; k < 2; ++k) {
        two_body_state(1.0, 1.0, q, e, k == 0 ? t : t + dt, omega, st);
        double* o = orbit + 6 * k;
        o[0] = st[0]; o[1] = st[1]; o[2] = st[2];
        o[3] = st[5]; o[4] = st[6]; o[5] = st[7];
    }
    clock[1] = dt;
    clock[0] = rk_order == 2 ? t * 0.5 + (t + dt) * (1.0 - 0.5) : t + dt;
}

__global__ void clear_status_kernel(int* status)
{
    if (threadIdx.x < 4) status[threadIdx.x] = 0;
}

} // namespace

// ===========================================================================
extern "C" {

int m3p_version(void) { return 1; }

const char* m3p_last_error(void) { return g_last_error.c_str(); }

int m3p_check_errors(void* stream, int* code, int* block, int* zone)
{
    StreamState* ds;
    if (int st = stream_state(stream, &ds)) return st;
    hipStream_t s = (hipStream_t)stream;
    int h[4] = { 0, 0, 0, 0 };
    if (int st = hip_check(hipMemcpyAsync(h, ds->err, sizeof(h), hipMemcpyDeviceToHost, s), "hipMemcpyAsync(error latch)")) return st;
    if (int st = hip_check(hipStreamSynchronize(s), "hipStreamSynchronize")) return st;
    if (h[0] != 0) {
        const int zero[4] = { 0, 0, 0, 0 };
        if (int st = hip_check(hipMemcpy(ds->err, zero, sizeof(zero), hipMemcpyHostToDevice), "hipMemcpy(error latch)")) return st;
    }
    if (code) *code = h[0];
    if (block) *block = h[1];
    if (zone) *zone = h[2];
    return M3P_OK;
}

int m3p_euler2d_updated_conserved(const double* uc, const double* pe, double* out, int n,
                                  double dt, double dl, double plm_theta, double gamma_law_index,
                                  int hllc, int precision, void* stream)
{
    Lease lease;
    if (int st = lease.acquire(stream)) return st;
    if (!uc || !pe || !out) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: NULL array");
    EulerParams prm{ gamma_law_index, plm_theta, hllc };
    StageArgs a{};
    a.out = out; a.dt = dt; a.dl = dl; a.rk_b = 0.5; a.n = n; a.nblk = 1; a.err = lease.latch;
    hipStream_t s = (hipStream_t)stream;
    int st;
    if (precision == M3P_FP64)
        st = launch_stage2d<EulerScheme<double>>(M3P_KIND_EULER2D, prm, ExtendedLoader<EulerScheme<double>>{uc, pe, n}, a, s);
    else
        st = launch_stage2d<EulerScheme<float>>(M3P_KIND_EULER2D, prm, ExtendedLoader<EulerScheme<float>>{uc, pe, n}, a, s);
    if (st) return st;
    return lease.finish();
}

int m3p_euler1d_updated_conserved(const double* uc, const double* pe, double* out, int n,
                                  double dt, double dl, double plm_theta, double gamma_law_index,
                                  int outflow_lower, int outflow_upper, int precision, void* stream)
{
    Lease lease;
    if (int st = lease.acquire(stream)) return st;
    if (!uc || !pe || !out) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: NULL array");
    EulerParams prm{ gamma_law_index, plm_theta, 0 };
    Stage1DArgs a{};
    a.u = uc; a.out = out; a.dt = dt; a.dl = dl; a.rk_b = 0.5; a.n = n; a.nblk = 1; a.err = lease.latch;
    a.pe = pe; a.outflow_lower = outflow_lower; a.outflow_upper = outflow_upper;
    hipStream_t s = (hipStream_t)stream;
    int st = precision == M3P_FP64 ? launch_stage1d<double>(prm, a, s, lease.cus) : launch_stage1d<float>(prm, a, s, lease.cus);
    if (st) return st;
    return lease.finish();
}

int m3p_minidisk_updated_conserved(const double* uc, const double* pe, double* out,
                                   int level, int bi, int bj, double time, double dt,
                                   const m3p_minidisk_data* data, int precision, void* stream)
{
    Lease lease;
    if (int st = lease.acquire(stream)) return st;
    if (int st = check_minidisk_data(data)) return st;
    if (!uc || !pe || !out) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: NULL array");
    const int n = data->block_size;
    MinidiskParams prm = minidisk_params(data, time);
    prm.level = level;
    StageArgs a{};
    a.out = out; a.dt = dt; a.n = n; a.nblk = 1; a.err = lease.latch;
    a.bi0 = bi; a.bj0 = bj;
    // cell_size, problems/minidisk_scheme.cpp:78-81
    a.dl = 2.0 * data->domain_radius / double(data->block_size) / double(1 << level);
    hipStream_t s = (hipStream_t)stream;
    const double nu = data->kinematic_viscosity;
    int st;
    if (nu != 0.0)
        st = precision == M3P_FP64
           ? launch_visc<MinidiskScheme<double>>(prm, ExtendedLoader<MinidiskScheme<double>>{uc, pe, n}, a, nu, s)
           : launch_visc<MinidiskScheme<float>>(prm, ExtendedLoader<MinidiskScheme<float>>{uc, pe, n}, a, nu, s);
    else if (precision == M3P_FP64)
        st = launch_stage2d<MinidiskScheme<double>>(M3P_KIND_MINIDISK, prm, ExtendedLoader<MinidiskScheme<double>>{uc, pe, n}, a, s);
    else
        st = launch_stage2d<MinidiskScheme<float>>(M3P_KIND_MINIDISK, prm, ExtendedLoader<MinidiskScheme<float>>{uc, pe, n}, a, s);
    if (st) return st;
    return lease.finish();
}

static int euler2d_stage(const double* u, const double* u0, double* out, const int* nbr,
                         int nblk, int n, double dt, const double* dt_dev, double dl, double rk_b,
                         double plm_theta, double gamma_law_index, int hllc, int precision, void* stream,
                         unsigned long long* rate_slot)
{
    StreamState* ds;
    if (int st = stream_state(stream, &ds)) return st;
    if (!u || !out || !nbr || nblk <= 0) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: bad euler2d stage arguments");
    if (u == out || (u0 && u0 == out)) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: stage output must not alias its inputs");
    EulerParams prm{ gamma_law_index, plm_theta, hllc };
    StageArgs a{};
    a.u0 = u0; a.out = out; a.dt_dev = dt_dev; a.dt = dt; a.dl = dl; a.rk_b = rk_b;
    a.n = n; a.nblk = nblk; a.err = ds->err; a.rate_slot = rate_slot;
    hipStream_t s = (hipStream_t)stream;
    if (precision == M3P_FP64)
        return launch_stage2d<EulerScheme<double>>(M3P_KIND_EULER2D, prm, MeshLoader<EulerScheme<double>>{u, nbr, n}, a, s);
    return launch_stage2d<EulerScheme<float>>(M3P_KIND_EULER2D, prm, MeshLoader<EulerScheme<float>>{u, nbr, n}, a, s);
}

int m3p_euler2d_stage(const double* u, const double* u0, double* out, const int* nbr,
                      int nblk, int n, double dt, const double* dt_dev, double dl, double rk_b,
                      double plm_theta, double gamma_law_index, int hllc, int precision, void* stream)
{
    return euler2d_stage(u, u0, out, nbr, nblk, n, dt, dt_dev, dl, rk_b, plm_theta, gamma_law_index, hllc,
                         precision, stream, nullptr);
}

static int minidisk_stage(const double* u, const double* u0, double* out, const int* nbr, const int* bcoord,
                          int nblk, double time, double dt, const double* orbit, const double* dt_dev, double rk_b,
                          const m3p_minidisk_data* data, int precision, void* stream)
{
    StreamState* ds;
    if (int st = stream_state(stream, &ds)) return st;
    if (int st = check_minidisk_data(data)) return st;
    if (!u || !out || !nbr || !bcoord || nblk <= 0) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: bad minidisk stage arguments");
    if (u == out || (u0 && u0 == out)) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: stage output must not alias its inputs");
    const int n = data->block_size;
    MinidiskParams prm = minidisk_params(data, time);
    prm.orbit = orbit;
    StageArgs a{};
    a.u0 = u0; a.out = out; a.dt = dt; a.dt_dev = dt_dev; a.rk_b = rk_b; a.bcoord = bcoord;
    a.n = n; a.nblk = nblk; a.err = ds->err;
    a.dl = 2.0 * data->domain_radius / double(data->block_size) / double(1 << data->depth);
    hipStream_t s = (hipStream_t)stream;
    const double nu = data->kinematic_viscosity;
    if (nu != 0.0)
        return precision == M3P_FP64
             ? launch_visc<MinidiskScheme<double>>(prm, MeshLoader<MinidiskScheme<double>>{u, nbr, n}, a, nu, s)
             : launch_visc<MinidiskScheme<float>>(prm, MeshLoader<MinidiskScheme<float>>{u, nbr, n}, a, nu, s);
    if (precision == M3P_FP64)
        return launch_stage2d<MinidiskScheme<double>>(M3P_KIND_MINIDISK, prm, MeshLoader<MinidiskScheme<double>>{u, nbr, n}, a, s);
    return launch_stage2d<MinidiskScheme<float>>(M3P_KIND_MINIDISK, prm, MeshLoader<MinidiskScheme<float>>{u, nbr, n}, a, s);
}

int m3p_minidisk_stage(const double* u, const double* u0, double* out, const int* nbr, const int* bcoord,
                       int nblk, double time, double dt, double rk_b,
                       const m3p_minidisk_data* data, int precision, void* stream)
{
    return minidisk_stage(u, u0, out, nbr, bcoord, nblk, time, dt, nullptr, nullptr, rk_b, data, precision, stream);
}

int m3p_minidisk_stage_dev(const double* u, const double* u0, double* out, const int* nbr, const int* bcoord,
                           int nblk, const double* orbit, const double* dt_dev, double rk_b,
                           const m3p_minidisk_data* data, int precision, void* stream)
{
    if (!orbit || !dt_dev) return fail(M3P_ERR_BAD_ARGUMENT, "m3p_minidisk_stage_dev: orbit and dt_dev are required");
    return minidisk_stage(u, u0, out, nbr, bcoord, nblk, 0.0, 0.0, orbit, dt_dev, rk_b, data, precision, stream);
}

static int minidisk_rk2_step(const double* u, double* out, const int* nbr, const int* bcoord, int nblk,
                             double time, double dt, const double* orbit, const double* dt_dev, double rk_b,
                             const m3p_minidisk_data* data, int precision, void* stream)
{
    StreamState* ds;
    if (int st = stream_state(stream, &ds)) return st;
    if (int st = check_minidisk_data(data)) return st;
    if (!u || !out || !nbr || !bcoord || nblk <= 0) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: bad minidisk RK2 step arguments");
    if (u == out) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: RK2 step output must not alias its input");
    if (precision != M3P_MIXED)
        return fail(M3P_ERR_BAD_ARGUMENT, "m3p: the fused RK2 step is compiled for the mixed precision mode only");
    if (data->kinematic_viscosity != 0.0)
        return fail(M3P_ERR_BAD_ARGUMENT, "m3p: the fused RK2 step has no viscous stress (use the stage entry points)");
    const int n = data->block_size;
    MinidiskParams p1 = minidisk_params(data, time), p2 = minidisk_params(data, time + dt);
    p1.orbit = orbit;
    p2.orbit = orbit ? orbit + 6 : nullptr;
    StageArgs a{};
    a.out = out; a.dt = dt; a.dt_dev = dt_dev; a.rk_b = rk_b; a.bcoord = bcoord;
    a.n = n; a.nblk = nblk; a.err = ds->err;
    a.dl = 2.0 * data->domain_radius / double(data->block_size) / double(1 << data->depth);
    return launch_rk2_fused<MinidiskScheme<float>>(p1, p2, MeshLoader<MinidiskScheme<float>>{u, nbr, n}, a,
                                                   (hipStream_t)stream);
}

int m3p_minidisk_rk2_step(const double* u, double* out, const int* nbr, const int* bcoord, int nblk,
                          double time, double dt, double rk_b, const m3p_minidisk_data* data, int precision,
                          void* stream)
{
    return minidisk_rk2_step(u, out, nbr, bcoord, nblk, time, dt, nullptr, nullptr, rk_b, data, precision, stream);
}

int m3p_minidisk_rk2_step_dev(const double* u, double* out, const int* nbr, const int* bcoord, int nblk,
                              const double* orbit, const double* dt_dev, double rk_b,
                              const m3p_minidisk_data* data, int precision, void* stream)
{
    if (!orbit || !dt_dev) return fail(M3P_ERR_BAD_ARGUMENT, "m3p_minidisk_rk2_step_dev: orbit and dt_dev are required");
    return minidisk_rk2_step(u, out, nbr, bcoord, nblk, 0.0, 0.0, orbit, dt_dev, rk_b, data, precision, stream);
}

static int euler2d_rk2_step(const double* u, double* out, const int* nbr, int nblk, int n, double dt,
                            const double* dt_dev, double dl, double rk_b, double plm_theta, double gamma_law_index,
                            int hllc, int precision, void* stream, unsigned long long* rate_slot)
{
    StreamState* ds;
    if (int st = stream_state(stream, &ds)) return st;
    if (!u || !out || !nbr || nblk <= 0) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: bad euler2d RK2 step arguments");
    if (u == out) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: RK2 step output must not alias its input");
    if (precision != M3P_MIXED)
        return fail(M3P_ERR_BAD_ARGUMENT, "m3p: the fused RK2 step is compiled for the mixed precision mode only");
    EulerParams prm{ gamma_law_index, plm_theta, hllc };
    StageArgs a{};
    a.out = out; a.dt = dt; a.dt_dev = dt_dev; a.dl = dl; a.rk_b = rk_b;
    a.n = n; a.nblk = nblk; a.err = ds->err; a.rate_slot = rate_slot;
    return launch_rk2_fused<EulerScheme<float>>(prm, prm, MeshLoader<EulerScheme<float>>{u, nbr, n}, a,
                                                (hipStream_t)stream);
}

int m3p_euler2d_rk2_step(const double* u, double* out, const int* nbr, int nblk, int n, double dt,
                         const double* dt_dev, double dl, double rk_b, double plm_theta, double gamma_law_index,
                         int hllc, int precision, void* stream)
{
    return euler2d_rk2_step(u, out, nbr, nblk, n, dt, dt_dev, dl, rk_b, plm_theta, gamma_law_index, hllc, precision,
                            stream, nullptr);
}

int m3p_minidisk_clock(const m3p_minidisk_data* data, double* clock, double* orbit, int rk_order, void* stream)
{
    StreamState* ds;
    if (int st = stream_state(stream, &ds)) return st;
    if (int st = check_minidisk_data(data)) return st;
    if (!clock || !orbit || (rk_order != 1 && rk_order != 2))
        return fail(M3P_ERR_BAD_ARGUMENT, "m3p_minidisk_clock: bad arguments");
    // smallest_cell_size / vmax on the host, in the reference's order
    // (problems/minidisk_scheme.cpp:122-125, problems/minidisk.cpp:204-206)
    const double dx = 2.0 * data->domain_radius / double(data->block_size) / double(1 << data->depth);
    const double vmax = std::sqrt(1.0 * 0.5 / data->softening_length);
    hipLaunchKernelGGL(minidisk_clock_kernel, dim3(1), dim3(64), 0, (hipStream_t)stream,
                       dx / vmax, data->cfl, data->mass_ratio, data->eccentricity, data->omega_frame, rk_order,
                       clock, orbit);
    return launch_status("m3p_minidisk_clock");
}

int m3p_euler1d_stage(const double* u, const double* u0, double* out, const int* nbr,
                      int nblk, int n, double dt, double dl, double rk_b,
                      double plm_theta, double gamma_law_index, int precision, void* stream)
{
    StreamState* ds;
    if (int st = stream_state(stream, &ds)) return st;
    if (!u || !out || !nbr || nblk <= 0) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: bad euler1d stage arguments");
    if (u == out || (u0 && u0 == out)) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: stage output must not alias its inputs");
    EulerParams prm{ gamma_law_index, plm_theta, 0 };
    Stage1DArgs a{};
    a.u = u; a.u0 = u0; a.out = out; a.nbr = nbr; a.dt = dt; a.dl = dl; a.rk_b = rk_b;
    a.n = n; a.nblk = nblk; a.err = ds->err;
    hipStream_t s = (hipStream_t)stream;
    return precision == M3P_FP64 ? launch_stage1d<double>(prm, a, s, ds->cus) : launch_stage1d<float>(prm, a, s, ds->cus);
}

int m3p_euler_cfl_dt(const double* u, int nblk, long zones_per_block, double gamma_law_index,
                     double cfl, double dl, double* dt_out, void* stream)
{
    StreamState* ds;
    if (int st = stream_state(stream, &ds)) return st;
    if (!u || !dt_out || nblk <= 0 || zones_per_block <= 0) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: bad cfl arguments");
    hipStream_t s = (hipStream_t)stream;
    EulerParams prm{ gamma_law_index, 0.0, 0 };
    unsigned long long* slot = ds->red + kSlotCflRate;
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, s, slot, kRateSlots, 0ull);
    const long total = (long)nblk * zones_per_block;
    int grid = (int)std::min<long>((total + 255) / 256, 2048);
    hipLaunchKernelGGL(euler_max_rate_kernel, dim3(grid), dim3(256), 0, s, u, nblk, zones_per_block,
                       (const double*)nullptr, dl, prm, slot);
    hipLaunchKernelGGL(cfl_from_rate_kernel, dim3(1), dim3(kRateSlots), 0, s, slot, cfl, dt_out);
    return launch_status("m3p_euler_cfl_dt");
}

int m3p_conserved_totals(const double* u, int nblk, int nc, long zones_per_block, double* totals, void* stream)
{
    StreamState* ds;
    if (int st = stream_state(stream, &ds)) return st;
    if (!u || !totals || nblk <= 0 || nc <= 0 || nc > 8 || zones_per_block <= 0)
        return fail(M3P_ERR_BAD_ARGUMENT, "m3p: bad totals arguments");
    hipStream_t s = (hipStream_t)stream;
    const int nparts = 256;     // nc <= 8 components x 256 partials fit in the scratch
    double* partial = reinterpret_cast<double*>(ds->red);
    hipLaunchKernelGGL(totals_partial_kernel, dim3(nparts, nc), dim3(256), 0, s, u, nblk, nc, zones_per_block, partial);
    hipLaunchKernelGGL(totals_final_kernel, dim3(nc), dim3(64), 0, s, partial, nparts, totals);
    return launch_status("m3p_conserved_totals");
}

int m3p_euler_recover_primitive_array(const double* u, double* p, long nzones, double gamma_law_index, void* stream)
{
    Lease lease;
    if (int st = lease.acquire(stream)) return st;
    if (!u || !p || nzones <= 0) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: bad recover arguments");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(euler_recover_array_kernel, dim3((nzones + 255) / 256), dim3(256), 0, s, u, p, nzones, gamma_law_index, lease.latch);
    if (int st = launch_status("euler_recover_array_kernel")) return st;
    return lease.finish();
}

int m3p_iso2d_recover_primitive_array(const double* u, double* p, long nzones, void* stream)
{
    StreamState* ds;
    if (int st = stream_state(stream, &ds)) return st;
    if (!u || !p || nzones <= 0) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: bad recover arguments");
    hipLaunchKernelGGL(iso2d_recover_array_kernel, dim3((nzones + 255) / 256), dim3(256), 0, (hipStream_t)stream, u, p, nzones);
    return launch_status("iso2d_recover_array_kernel");
}

int m3p_plm_gradient_2d(const double* p, double* g, int nc, int ni, int nj, int axis, double plm_theta, void* stream)
{
    StreamState* ds;
    if (int st = stream_state(stream, &ds)) return st;
    if (!p || !g || nc <= 0 || axis < 0 || axis > 1 || (axis == 0 ? ni : nj) < 3 || ni <= 0 || nj <= 0)
        return fail(M3P_ERR_BAD_SHAPE, "m3p: bad gradient shape");
    long nout = (long)(axis == 0 ? ni - 2 : ni) * (axis == 1 ? nj - 2 : nj) * nc;
    hipLaunchKernelGGL(plm_gradient_2d_kernel, dim3((nout + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       p, g, nc, ni, nj, axis, plm_theta);
    return launch_status("plm_gradient_2d_kernel");
}

int m3p_weighted_sum(const double* s, const double* t, double* out, long n, double b, void* stream)
{
    StreamState* ds;
    if (int st = stream_state(stream, &ds)) return st;
    if (!s || !t || !out || n <= 0) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: bad weighted_sum arguments");
    hipLaunchKernelGGL(weighted_sum_kernel, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, s, t, out, n, b);
    return launch_status("weighted_sum_kernel");
}

long m3p_mhd_workspace_bytes(int N, int precision)
{
    if (N <= 0) return -1;
    return mhd_workspace_bytes(N, precision);
}

int m3p_mhd_advance_ct(const double* uc, const double* bf1, const double* bf2, const double* bf3,
                       const double* u0, const double* b01, const double* b02, const double* b03, double rk_b,
                       int N, double dl, double cfl_number, double plm_theta, double gamma_law_index,
                       double* uo, double* bo1, double* bo2, double* bo3, double* dt,
                       void* work, long work_bytes, int precision, void* stream)
{
    StreamState* ds;
    if (int st = stream_state(stream, &ds)) return st;
    if (!uc || !bf1 || !bf2 || !bf3 || !uo || !bo1 || !bo2 || !bo3 || !work || N < 2)
        return fail(M3P_ERR_BAD_ARGUMENT, "m3p_mhd_advance_ct: bad arguments");
    if (u0 && (!b01 || !b02 || !b03)) return fail(M3P_ERR_BAD_ARGUMENT, "m3p_mhd_advance_ct: RK base needs all four arrays");
    if (uo == uc || bo1 == bf1 || bo2 == bf2 || bo3 == bf3)
        return fail(M3P_ERR_BAD_ARGUMENT, "m3p_mhd_advance_ct: outputs must not alias inputs");
    if (work_bytes < mhd_workspace_bytes(N, precision))
        return fail(M3P_ERR_BAD_ARGUMENT, "m3p_mhd_advance_ct: workspace too small (see m3p_mhd_workspace_bytes)");
    MhdArgs a{};
    a.uc = uc; a.b1 = bf1; a.b2 = bf2; a.b3 = bf3;
    a.u0 = u0; a.b01 = b01; a.b02 = b02; a.b03 = b03; a.rk_b = rk_b;
    a.uo = uo; a.bo1 = bo1; a.bo2 = bo2; a.bo3 = bo3;
    a.N = N; a.dl = dl; a.cfl = cfl_number; a.theta = plm_theta; a.gamma = gamma_law_index;
    a.dt = dt; a.err = ds->err;
    hipStream_t s = (hipStream_t)stream;
    if (precision == M3P_FP64) return mhd_advance<double>(a, (char*)work, s);
    return mhd_advance<float>(a, (char*)work, s);
}

int m3p_mhd_advance_ct_noalloc(const double* uc, const double* bf1, const double* bf2, const double* bf3,
                               const double* u0, const double* b01, const double* b02, const double* b03, double rk_b,
                               int N, double dl, double cfl_number, double plm_theta, double gamma_law_index,
                               double* uo, double* bo1, double* bo2, double* bo3, double* dt, int* status,
                               int precision, void* stream)
{
    const char* what = "m3p_mhd_advance_ct_noalloc";
    if (!uc || !bf1 || !bf2 || !bf3 || !uo || !bo1 || !bo2 || !bo3 || !status || N < 2)
        return fail(M3P_ERR_BAD_ARGUMENT, std::string(what) + ": bad arguments");
    if (u0 && (!b01 || !b02 || !b03)) return fail(M3P_ERR_BAD_ARGUMENT, std::string(what) + ": RK base needs all four arrays");
    if (uo == uc || bo1 == bf1 || bo2 == bf2 || bo3 == bf3)
        return fail(M3P_ERR_BAD_ARGUMENT, std::string(what) + ": outputs must not alias inputs");
    const long need = mhd_workspace_bytes(N, precision);
    if (need > kScratchBytes) {
        int nmax = N;
        while (nmax > 2 && mhd_workspace_bytes(nmax, precision) > kScratchBytes) --nmax;
        return fail(M3P_ERR_BAD_ARGUMENT, std::string(what) + ": N = " + std::to_string(N) + " needs " +
                    std::to_string(need) + " bytes of workspace, more than the " + std::to_string(kScratchBytes) +
                    "-byte static scratch (largest N in this precision: " + std::to_string(nmax) + ")");
    }
    void* work = nullptr;
    if (int st = scratch_address(&work)) return st;
    MhdArgs a{};
    a.uc = uc; a.b1 = bf1; a.b2 = bf2; a.b3 = bf3;
    a.u0 = u0; a.b01 = b01; a.b02 = b02; a.b03 = b03; a.rk_b = rk_b;
    a.uo = uo; a.bo1 = bo1; a.bo2 = bo2; a.bo3 = bo3;
    a.N = N; a.dl = dl; a.cfl = cfl_number; a.theta = plm_theta; a.gamma = gamma_law_index;
    a.dt = dt; a.err = status;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(clear_status_kernel, dim3(1), dim3(64), 0, s, status);
    if (int st = launch_status("clear_status_kernel")) return st;
    if (precision == M3P_FP64) return mhd_advance<double>(a, (char*)work, s);
    return mhd_advance<float>(a, (char*)work, s);
}

int m3p_mhd_max_div_b(const double* bf1, const double* bf2, const double* bf3, int N, double* out, void* stream)
{
    StreamState* ds;
    if (int st = stream_state(stream, &ds)) return st;
    if (!bf1 || !bf2 || !bf3 || !out || N < 1) return fail(M3P_ERR_BAD_ARGUMENT, "m3p_mhd_max_div_b: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    const long N3 = (long)N * N * N;
    unsigned long long* slot = ds->red + kSlotDivB;
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, s, slot, 1, 0ull);
    hipLaunchKernelGGL(mhd_divb_kernel, dim3((int)std::min<long>((N3 + 255) / 256, 4096)), dim3(256), 0, s, bf1, bf2, bf3, N, slot);
    hipLaunchKernelGGL(copy_u64_as_double_kernel, dim3(1), dim3(1), 0, s, slot, out);
    return launch_status("m3p_mhd_max_div_b");
}

int m3p_srhd_recover_primitive_array(const double* u, double* p, long nzones, double gamma_law_index,
                                     double temperature_floor, void* stream)
{
    Lease lease;
    if (int st = lease.acquire(stream)) return st;
    if (!u || !p || nzones <= 0) return fail(M3P_ERR_BAD_ARGUMENT, "m3p_srhd_recover_primitive_array: bad arguments");
    hipLaunchKernelGGL(srhd_recover_array_kernel, dim3((nzones + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       u, p, nzones, gamma_law_index, temperature_floor, lease.latch);
    if (int st = launch_status("srhd_recover_array_kernel")) return st;
    int code = 0, block = 0, zone = 0;
    if (int st = lease.wait(&code, &block, &zone)) return st;
    if (code == M3P_ERR_NO_CONVERGENCE) return fail(code, "srhd::recover_primitive (root finder not converging)");
    if (code == M3P_ERR_NEGATIVE_PRESSURE) return fail(code, "srhd::recover_primitive (negative pressure)");
    return code ? fail(code, "m3p: kernel error") : M3P_OK;
}

int m3p_srhd_conserved_density_array(const double* p, double* u, long nzones, double gamma_law_index, void* stream)
{
    StreamState* ds;
    if (int st = stream_state(stream, &ds)) return st;
    if (!u || !p || nzones <= 0) return fail(M3P_ERR_BAD_ARGUMENT, "m3p_srhd_conserved_density_array: bad arguments");
    hipLaunchKernelGGL(srhd_conserved_array_kernel, dim3((nzones + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       p, u, nzones, gamma_law_index);
    return launch_status("srhd_conserved_array_kernel");
}

int m3p_srhd_riemann_array(const double* pl, const double* pr, double* f, long nfaces, int axis,
                           double gamma_law_index, int mode, double face_speed, void* stream)
{
    Lease lease;
    if (int st = lease.acquire(stream)) return st;
    if (!pl || !pr || !f || nfaces <= 0 || axis < 1 || axis > 3 || mode < 0 || mode > 3)
        return fail(M3P_ERR_BAD_ARGUMENT, "m3p_srhd_riemann_array: bad arguments");
    hipLaunchKernelGGL(srhd_riemann_array_kernel, dim3((nfaces + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       pl, pr, f, nfaces, axis, gamma_law_index, mode, face_speed, lease.latch);
    if (int st = launch_status("srhd_riemann_array_kernel")) return st;
    int code = 0, block = 0, zone = 0;
    if (int st = lease.wait(&code, &block, &zone)) return st;
    if (code) return fail(code, "srhd::riemann_solver (nan contact speed)");
    return M3P_OK;
}

static int euler1d_stage_blocks(const double* u, const double* u0, double* out, const int* nbr, const double* dl_blk,
                                int nblk, int n, double dt, const double* dt_dev, double rk_b,
                                double plm_theta, double gamma_law_index, int precision, void* stream,
                                unsigned long long* rate_slot)
{
    StreamState* ds;
    if (int st = stream_state(stream, &ds)) return st;
    if (!u || !out || !nbr || !dl_blk || nblk <= 0) return fail(M3P_ERR_BAD_ARGUMENT, "m3p_euler1d_stage_blocks: bad arguments");
    if (u == out || (u0 && u0 == out)) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: stage output must not alias its inputs");
    EulerParams prm{ gamma_law_index, plm_theta, 0 };
    Stage1DArgs a{};
    a.u = u; a.u0 = u0; a.out = out; a.nbr = nbr; a.dt = dt; a.dt_dev = dt_dev; a.dl_blk = dl_blk; a.rk_b = rk_b;
    a.n = n; a.nblk = nblk; a.err = ds->err; a.rate_slot = rate_slot;
    hipStream_t s = (hipStream_t)stream;
    return precision == M3P_FP64 ? launch_stage1d<double>(prm, a, s, ds->cus) : launch_stage1d<float>(prm, a, s, ds->cus);
}

int m3p_euler1d_stage_blocks(const double* u, const double* u0, double* out, const int* nbr, const double* dl_blk,
                             int nblk, int n, double dt, const double* dt_dev, double rk_b,
                             double plm_theta, double gamma_law_index, int precision, void* stream)
{
    return euler1d_stage_blocks(u, u0, out, nbr, dl_blk, nblk, n, dt, dt_dev, rk_b, plm_theta, gamma_law_index,
                                precision, stream, nullptr);
}

static int euler1d_rk2_step_blocks(const double* u, double* out, const int* nbr, const double* dl_blk, int nblk,
                                   int n, double dt, const double* dt_dev, double rk_b, double plm_theta,
                                   double gamma_law_index, int precision, void* stream, unsigned long long* rate_slot)
{
    StreamState* ds;
    if (int st = stream_state(stream, &ds)) return st;
    if (!u || !out || !nbr || !dl_blk || nblk <= 0)
        return fail(M3P_ERR_BAD_ARGUMENT, "m3p_euler1d_rk2_step_blocks: bad arguments");
    if (u == out) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: RK2 step output must not alias its input");
    if (precision != M3P_MIXED)
        return fail(M3P_ERR_BAD_ARGUMENT, "m3p: the fused RK2 step is compiled for the mixed precision mode only");
    EulerParams prm{ gamma_law_index, plm_theta, 0 };
    Stage1DArgs a{};
    a.u = u; a.out = out; a.nbr = nbr; a.dt = dt; a.dt_dev = dt_dev; a.dl_blk = dl_blk; a.rk_b = rk_b;
    a.n = n; a.nblk = nblk; a.err = ds->err; a.rate_slot = rate_slot;
    return launch_rk2_fused_1d(prm, a, (hipStream_t)stream);
}

int m3p_euler1d_rk2_step_blocks(const double* u, double* out, const int* nbr, const double* dl_blk, int nblk, int n,
                                double dt, const double* dt_dev, double rk_b, double plm_theta,
                                double gamma_law_index, int precision, void* stream)
{
    return euler1d_rk2_step_blocks(u, out, nbr, dl_blk, nblk, n, dt, dt_dev, rk_b, plm_theta, gamma_law_index,
                                   precision, stream, nullptr);
}

static int amr_euler2d_stage(const double* u, const double* u0, double* out, double* pe_work, const int* recipes,
                             const double* dl_leaf, int nleaf, int n, double dt, const double* dt_dev, double rk_b,
                             double plm_theta, double gamma_law_index, int hllc, int precision, void* stream,
                             unsigned long long* rate_slot, double* flux_reg = nullptr)
{
    StreamState* ds;
    if (int st = stream_state(stream, &ds)) return st;
    if (!u || !out || !recipes || !dl_leaf || nleaf <= 0) return fail(M3P_ERR_BAD_ARGUMENT, "m3p_amr_euler2d_stage: bad arguments");
    if (u == out || (u0 && u0 == out)) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: stage output must not alias its inputs");
    if (n < 16 || n % 2) return fail(M3P_ERR_BAD_SHAPE, "m3p_amr_euler2d_stage: leaf size must be an even number >= 16");
    EulerParams prm{ gamma_law_index, plm_theta, hllc };
    hipStream_t s = (hipStream_t)stream;
    StageArgs a{};
    a.u0 = u0; a.out = out; a.dt_dev = dt_dev; a.dt = dt; a.dl_blk = dl_leaf; a.rk_b = rk_b;
    a.n = n; a.nblk = nleaf; a.err = ds->err; a.rate_slot = rate_slot; a.flux_reg = flux_reg;
    if (!pe_work) {
        // fused: guard zones read through the recipes inside the stage kernel
        int st = precision == M3P_FP64
               ? launch_stage2d<EulerScheme<double>>(M3P_KIND_AMR_EULER2D, prm, TreeLoader<EulerScheme<double>>{u, recipes, n}, a, s)
               : launch_stage2d<EulerScheme<float>>(M3P_KIND_AMR_EULER2D, prm, TreeLoader<EulerScheme<float>>{u, recipes, n}, a, s);
        if (st || !flux_reg) return st;
        const long threads = (long)nleaf * 2 * n;
        const double w = u0 ? 1.0 - rk_b : 1.0;
        for (int phase = 0; phase < 2; ++phase)
            hipLaunchKernelGGL(amr_reflux_kernel, dim3((threads + 255) / 256), dim3(256), 0, s, out, (const double*)flux_reg,
                               recipes, dl_leaf, nleaf, n, dt, dt_dev, w, phase);
        return launch_status("amr_reflux_kernel");
    }
    // two passes: extended primitive blocks into pe_work, then the update
    const long ext = (long)(n + 4) * (n + 4) * nleaf;
    hipLaunchKernelGGL(amr_guard_fill_kernel, dim3((ext + 255) / 256), dim3(256), 0, s, u, pe_work, recipes, nleaf, n, prm, ds->err);
    if (int st = launch_status("amr_guard_fill_kernel")) return st;
    if (precision == M3P_FP64)
        return launch_stage2d<EulerScheme<double>>(M3P_KIND_AMR_EULER2D, prm, ExtendedLoader<EulerScheme<double>>{u, pe_work, n}, a, s);
    return launch_stage2d<EulerScheme<float>>(M3P_KIND_AMR_EULER2D, prm, ExtendedLoader<EulerScheme<float>>{u, pe_work, n}, a, s);
}

int m3p_amr_euler2d_stage(const double* u, const double* u0, double* out, double* pe_work, const int* recipes,
                          const double* dl_leaf, int nleaf, int n, double dt, const double* dt_dev, double rk_b,
                          double plm_theta, double gamma_law_index, int hllc, int precision, void* stream)
{
    return amr_euler2d_stage(u, u0, out, pe_work, recipes, dl_leaf, nleaf, n, dt, dt_dev, rk_b, plm_theta,
                             gamma_law_index, hllc, precision, stream, nullptr);
}

int m3p_amr_euler2d_stage_conservative(const double* u, const double* u0, double* out, double* flux_work,
                                       const int* recipes, const double* dl_leaf, int nleaf, int n, double dt,
                                       const double* dt_dev, double rk_b, double plm_theta, double gamma_law_index,
                                       int hllc, int precision, void* stream)
{
    if (!flux_work) return fail(M3P_ERR_BAD_ARGUMENT, "m3p_amr_euler2d_stage_conservative: NULL flux workspace");
    return amr_euler2d_stage(u, u0, out, nullptr, recipes, dl_leaf, nleaf, n, dt, dt_dev, rk_b, plm_theta,
                             gamma_law_index, hllc, precision, stream, nullptr, flux_work);
}

// Stages with the CFL reduction of their output fused in: the rate slot is
// cleared before the stage and turned into dt_next after it (stream order,
// so dt_next may be the dt_dev the stage reads).
static int cfl_stage(double cfl, double* dt_next, void* stream, const std::function<int(unsigned long long*)>& stage)
{
    StreamState* ds;
    if (int st = stream_state(stream, &ds)) return st;
    if (!dt_next || !(cfl > 0.0)) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: a *_stage_cfl call needs cfl > 0 and dt_next");
    hipStream_t s = (hipStream_t)stream;
    unsigned long long* slot = ds->red + kSlotCflRate;
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, s, slot, kRateSlots, 0ull);
    if (int st = launch_status("fill_u64_kernel")) return st;
    if (int st = stage(slot)) return st;
    hipLaunchKernelGGL(cfl_from_rate_kernel, dim3(1), dim3(kRateSlots), 0, s, slot, cfl, dt_next);
    return launch_status("cfl_from_rate_kernel");
}

int m3p_euler2d_stage_cfl(const double* u, const double* u0, double* out, const int* nbr,
                          int nblk, int n, double dt, const double* dt_dev, double dl, double rk_b,
                          double plm_theta, double gamma_law_index, int hllc, double cfl, double* dt_next,
                          int precision, void* stream)
{
    return cfl_stage(cfl, dt_next, stream, [&] (unsigned long long* slot) {
        return euler2d_stage(u, u0, out, nbr, nblk, n, dt, dt_dev, dl, rk_b, plm_theta, gamma_law_index, hllc,
                             precision, stream, slot);
    });
}

int m3p_euler2d_rk2_step_cfl(const double* u, double* out, const int* nbr, int nblk, int n, double dt,
                             const double* dt_dev, double dl, double rk_b, double plm_theta, double gamma_law_index,
                             int hllc, double cfl, double* dt_next, int precision, void* stream)
{
    return cfl_stage(cfl, dt_next, stream, [&] (unsigned long long* slot) {
        return euler2d_rk2_step(u, out, nbr, nblk, n, dt, dt_dev, dl, rk_b, plm_theta, gamma_law_index, hllc,
                                precision, stream, slot);
    });
}

int m3p_euler1d_rk2_step_blocks_cfl(const double* u, double* out, const int* nbr, const double* dl_blk, int nblk,
                                    int n, double dt, const double* dt_dev, double rk_b, double plm_theta,
                                    double gamma_law_index, double cfl, double* dt_next, int precision, void* stream)
{
    return cfl_stage(cfl, dt_next, stream, [&] (unsigned long long* slot) {
        return euler1d_rk2_step_blocks(u, out, nbr, dl_blk, nblk, n, dt, dt_dev, rk_b, plm_theta, gamma_law_index,
                                       precision, stream, slot);
    });
}

int m3p_euler1d_stage_blocks_cfl(const double* u, const double* u0, double* out, const int* nbr, const double* dl_blk,
                                 int nblk, int n, double dt, const double* dt_dev, double rk_b,
                                 double plm_theta, double gamma_law_index, double cfl, double* dt_next,
                                 int precision, void* stream)
{
    return cfl_stage(cfl, dt_next, stream, [&] (unsigned long long* slot) {
        return euler1d_stage_blocks(u, u0, out, nbr, dl_blk, nblk, n, dt, dt_dev, rk_b, plm_theta, gamma_law_index,
                                    precision, stream, slot);
    });
}

int m3p_amr_euler2d_stage_cfl(const double* u, const double* u0, double* out, double* pe_work, const int* recipes,
                              const double* dl_leaf, int nleaf, int n, double dt, const double* dt_dev, double rk_b,
                              double plm_theta, double gamma_law_index, int hllc, double cfl, double* dt_next,
                              int precision, void* stream)
{
    return cfl_stage(cfl, dt_next, stream, [&] (unsigned long long* slot) {
        return amr_euler2d_stage(u, u0, out, pe_work, recipes, dl_leaf, nleaf, n, dt, dt_dev, rk_b, plm_theta,
                                 gamma_law_index, hllc, precision, stream, slot);
    });
}

int m3p_euler_cfl_dt_blocks(const double* u, int nblk, long zones_per_block, const double* dl_blk,
                            double gamma_law_index, double cfl, double* dt_out, void* stream)
{
    StreamState* ds;
    if (int st = stream_state(stream, &ds)) return st;
    if (!u || !dl_blk || !dt_out || nblk <= 0 || zones_per_block <= 0) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: bad cfl arguments");
    hipStream_t s = (hipStream_t)stream;
    EulerParams prm{ gamma_law_index, 0.0, 0 };
    unsigned long long* slot = ds->red + kSlotCflRate;
    hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, s, slot, kRateSlots, 0ull);
    const long total = (long)nblk * zones_per_block;
    int grid = (int)std::min<long>((total + 255) / 256, 2048);
    hipLaunchKernelGGL(euler_max_rate_kernel, dim3(grid), dim3(256), 0, s, u, nblk, zones_per_block, dl_blk, 0.0, prm, slot);
    hipLaunchKernelGGL(cfl_from_rate_kernel, dim3(1), dim3(kRateSlots), 0, s, slot, cfl, dt_out);
    return launch_status("m3p_euler_cfl_dt_blocks");
}

// hipGraph capture of any sequence of asynchronous m3p calls on one stream.
int m3p_graph_begin(void* stream)
{
    StreamState* ds;
    if (int st = stream_state(stream, &ds)) return st;      // library state exists before capture starts
    return hip_check(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
}

int m3p_graph_end(void* stream, void** graph_exec)
{
    if (!graph_exec) return fail(M3P_ERR_BAD_ARGUMENT, "m3p_graph_end: NULL handle");
    hipGraph_t g = nullptr;
    if (int st = hip_check(hipStreamEndCapture((hipStream_t)stream, &g), "hipStreamEndCapture")) return st;
    hipGraphExec_t e = nullptr;
    int st = hip_check(hipGraphInstantiate(&e, g, nullptr, nullptr, 0), "hipGraphInstantiate");
    (void)hipGraphDestroy(g);
    if (st) return st;
    *graph_exec = (void*)e;
    return M3P_OK;
}

int m3p_graph_launch(void* graph_exec, void* stream)
{
    if (!graph_exec) return fail(M3P_ERR_BAD_ARGUMENT, "m3p_graph_launch: NULL handle");
    return hip_check(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream), "hipGraphLaunch");
}

int m3p_graph_destroy(void* graph_exec)
{
    if (!graph_exec) return M3P_OK;
    return hip_check(hipGraphExecDestroy((hipGraphExec_t)graph_exec), "hipGraphExecDestroy");
}

int m3p_set_stage2d_tile(int kind, int precision, int tj, int h)
{
    if (kind < 0 || kind >= kNumKinds || precision < 0 || precision > 1) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: bad kind/precision");
    const int i = tile_index(tj, h);
    if (i >= 0) { g_tile[kind][precision].store(i, std::memory_order_relaxed); return M3P_OK; }
    return fail(M3P_ERR_BAD_ARGUMENT, "m3p: tile " + std::to_string(tj) + "x" + std::to_string(h) + " is not a compiled variant");
}

int m3p_get_stage2d_tile(int kind, int precision, int* tj, int* h)
{
    if (kind < 0 || kind >= kNumKinds || precision < 0 || precision > 1) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: bad kind/precision");
    const TileShape v = kTileVariants[g_tile[kind][precision].load(std::memory_order_relaxed)];
    if (tj) *tj = v.tj;
    if (h) *h = v.h;
    return M3P_OK;
}

int m3p_autotune_euler2d(const double* u, double* out, const int* nbr, int nblk, int n, double dl,
                         double gamma_law_index, int precision, int reps, int* tj, int* h, float* ms, void* stream)
{
    StreamState* ds;
    if (int st = stream_state(stream, &ds)) return st;
    if (!u || !out || !nbr || nblk <= 0 || u == out) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: bad autotune arguments");
    EulerParams prm{ gamma_law_index, 1.5, 0 };
    StageArgs a{};
    a.out = out; a.dt = 1e-6 * dl; a.dl = dl; a.rk_b = 0.5; a.n = n; a.nblk = nblk; a.err = ds->err;
    hipStream_t s = (hipStream_t)stream;
    if (precision == M3P_FP64)
        return autotune_stage2d<EulerScheme<double>>(M3P_KIND_EULER2D, prm, MeshLoader<EulerScheme<double>>{u, nbr, n}, a, s, reps, tj, h, ms);
    return autotune_stage2d<EulerScheme<float>>(M3P_KIND_EULER2D, prm, MeshLoader<EulerScheme<float>>{u, nbr, n}, a, s, reps, tj, h, ms);
}

int m3p_autotune_minidisk(const double* u, double* out, const int* nbr, const int* bcoord, int nblk,
                          const m3p_minidisk_data* data, int precision, int reps, int* tj, int* h, float* ms, void* stream)
{
    StreamState* ds;
    if (int st = stream_state(stream, &ds)) return st;
    if (int st = check_minidisk_data(data)) return st;
    if (!u || !out || !nbr || !bcoord || nblk <= 0 || u == out) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: bad autotune arguments");
    const int n = data->block_size;
    MinidiskParams prm = minidisk_params(data, 0.0);
    StageArgs a{};
    a.out = out; a.dt = 1e-6; a.rk_b = 0.5; a.bcoord = bcoord; a.n = n; a.nblk = nblk; a.err = ds->err;
    a.dl = 2.0 * data->domain_radius / double(data->block_size) / double(1 << data->depth);
    hipStream_t s = (hipStream_t)stream;
    if (precision == M3P_FP64)
        return autotune_stage2d<MinidiskScheme<double>>(M3P_KIND_MINIDISK, prm, MeshLoader<MinidiskScheme<double>>{u, nbr, n}, a, s, reps, tj, h, ms);
    return autotune_stage2d<MinidiskScheme<float>>(M3P_KIND_MINIDISK, prm, MeshLoader<MinidiskScheme<float>>{u, nbr, n}, a, s, reps, tj, h, ms);
}

int m3p_locally_isothermal_updated_conserved(const double* uc, const double* pe, double* out, int n,
                                   double dt, double dl, const double* cs2_x, const double* cs2_y,
                                   double kinematic_viscosity, int precision, void* stream)
{
    Lease lease;
    if (int st = lease.acquire(stream)) return st;
    if (!uc || !pe || !out || !cs2_x || !cs2_y) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: NULL array");
    if (!(kinematic_viscosity >= 0.0)) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: negative kinematic viscosity");
    IsoTableParams prm{ cs2_x, cs2_y, n };
    StageArgs a{};
    a.out = out; a.dt = dt; a.dl = dl; a.rk_b = 0.5; a.n = n; a.nblk = 1; a.err = lease.latch;
    hipStream_t s = (hipStream_t)stream;
    int st;
    if (kinematic_viscosity != 0.0)
        st = precision == M3P_FP64
           ? launch_visc<IsoTableScheme<double>>(prm, ExtendedLoader<IsoTableScheme<double>>{uc, pe, n}, a, kinematic_viscosity, s)
           : launch_visc<IsoTableScheme<float>>(prm, ExtendedLoader<IsoTableScheme<float>>{uc, pe, n}, a, kinematic_viscosity, s);
    else if (precision == M3P_FP64)
        st = launch_stage2d<IsoTableScheme<double>>(M3P_KIND_MINIDISK, prm, ExtendedLoader<IsoTableScheme<double>>{uc, pe, n}, a, s);
    else
        st = launch_stage2d<IsoTableScheme<float>>(M3P_KIND_MINIDISK, prm, ExtendedLoader<IsoTableScheme<float>>{uc, pe, n}, a, s);
    if (st) return st;
    return lease.finish();
}

int m3p_locally_isothermal_stage(const double* u, const double* u0, double* out, const int* nbr,
                                 int nblk, int n, double dt, const double* dt_dev, double dl, double rk_b,
                                 const double* cs2_x, const double* cs2_y, double kinematic_viscosity,
                                 int precision, void* stream)
{
    StreamState* ds;
    if (int st = stream_state(stream, &ds)) return st;
    if (!u || !out || !nbr || !cs2_x || !cs2_y || nblk <= 0)
        return fail(M3P_ERR_BAD_ARGUMENT, "m3p: bad locally isothermal stage arguments");
    if (u == out || (u0 && u0 == out)) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: stage output must not alias its inputs");
    if (!(kinematic_viscosity >= 0.0)) return fail(M3P_ERR_BAD_ARGUMENT, "m3p: negative kinematic viscosity");
    IsoTableParams prm{ cs2_x, cs2_y, n };
    StageArgs a{};
    a.u0 = u0; a.out = out; a.dt_dev = dt_dev; a.dt = dt; a.dl = dl; a.rk_b = rk_b;
    a.n = n; a.nblk = nblk; a.err = ds->err;
    hipStream_t s = (hipStream_t)stream;
    if (kinematic_viscosity != 0.0)
        return precision == M3P_FP64
             ? launch_visc<IsoTableScheme<double>>(prm, MeshLoader<IsoTableScheme<double>>{u, nbr, n}, a, kinematic_viscosity, s)
             : launch_visc<IsoTableScheme<float>>(prm, MeshLoader<IsoTableScheme<float>>{u, nbr, n}, a, kinematic_viscosity, s);
    if (precision == M3P_FP64)
        return launch_stage2d<IsoTableScheme<double>>(M3P_KIND_MINIDISK, prm, MeshLoader<IsoTableScheme<double>>{u, nbr, n}, a, s);
    return launch_stage2d<IsoTableScheme<float>>(M3P_KIND_MINIDISK, prm, MeshLoader<IsoTableScheme<float>>{u, nbr, n}, a, s);
}

// ---------------------------------------------------------------------------
// mhd_scheme_v2 array functions
static dim3 v2_grid(long n, int axes)
{
    return dim3((unsigned)std::max<long>(1, std::min<long>((n + 255) / 256, 16384)), (unsigned)axes);
}

static bool v2_shape_ok(int a, int b, int c, int least) { return a >= least && b >= least && c >= least; }

int m3p_mhd_v2_primitive_array(const double* uc, const double* bf1, const double* bf2, const double* bf3,
                               int n1, int n2, int n3, double* pc, void* stream)
{
    Lease lease;
    if (int st = lease.acquire(stream)) return st;
    if (!uc || !bf1 || !bf2 || !bf3 || !pc) return fail(M3P_ERR_BAD_ARGUMENT, "m3p_mhd_v2_primitive_array: NULL array");
    if (!v2_shape_ok(n1, n2, n3, 1)) return fail(M3P_ERR_BAD_SHAPE, "m3p_mhd_v2_primitive_array: empty block");
    hipLaunchKernelGGL(mhd_v2_prim_kernel, v2_grid((long)n1 * n2 * n3, 1), dim3(256), 0, (hipStream_t)stream,
                       n1, n2, n3, uc, bf1, bf2, bf3, pc, lease.latch);
    if (int st = launch_status("mhd_v2_prim_kernel")) return st;
    int code = 0, block = 0, zone = 0;
    if (int st = lease.wait(&code, &block, &zone)) return st;
    if (code == M3P_ERR_NEGATIVE_DENSITY) return fail(code, "mhd::recover_primitive (negative density)");
    if (code == M3P_ERR_NEGATIVE_PRESSURE) return fail(code, "mhd::recover_primitive (negative pressure)");
    return code ? fail(code, "m3p: kernel error " + std::to_string(code)) : M3P_OK;
}

int m3p_mhd_v2_godunov_fluxes(const double* pc, const double* bf1, const double* bf2, const double* bf3,
                              int m1, int m2, int m3, double* gf1, double* gf2, double* gf3, void* stream)
{
    if (!pc || !bf1 || !bf2 || !bf3 || !gf1 || !gf2 || !gf3)
        return fail(M3P_ERR_BAD_ARGUMENT, "m3p_mhd_v2_godunov_fluxes: NULL array");
    if (!v2_shape_ok(m1, m2, m3, 2)) return fail(M3P_ERR_BAD_SHAPE, "m3p_mhd_v2_godunov_fluxes: cells must be >= 2 per axis");
    const long nmax = (long)m1 * m2 * m3;
    hipLaunchKernelGGL(mhd_v2_godunov_kernel, v2_grid(nmax, 3), dim3(256), 0, (hipStream_t)stream,
                       m1, m2, m3, pc, bf1, bf2, bf3, gf1, gf2, gf3);
    return launch_status("mhd_v2_godunov_kernel");
}

int m3p_mhd_v2_electromotive_forces(const double* gf1, const double* gf2, const double* gf3, int m1, int m2, int m3,
                                    double* ee1, double* ee2, double* ee3, void* stream)
{
    if (!gf1 || !gf2 || !gf3 || !ee1 || !ee2 || !ee3)
        return fail(M3P_ERR_BAD_ARGUMENT, "m3p_mhd_v2_electromotive_forces: NULL array");
    if (!v2_shape_ok(m1, m2, m3, 2)) return fail(M3P_ERR_BAD_SHAPE, "m3p_mhd_v2_electromotive_forces: cells must be >= 2 per axis");
    hipLaunchKernelGGL(mhd_v2_emf_kernel, v2_grid((long)m1 * m2 * m3, 3), dim3(256), 0, (hipStream_t)stream,
                       m1, m2, m3, gf1, gf2, gf3, ee1, ee2, ee3);
    return launch_status("mhd_v2_emf_kernel");
}

int m3p_mhd_v2_updated_conserved_density(const double* uc, const double* gf1, const double* gf2, const double* gf3,
                                         int n1, int n2, int n3, double dt, double dl, double* out, void* stream)
{
    if (!uc || !gf1 || !gf2 || !gf3 || !out) return fail(M3P_ERR_BAD_ARGUMENT, "m3p_mhd_v2_updated_conserved_density: NULL array");
    if (out == uc) return fail(M3P_ERR_BAD_ARGUMENT, "m3p_mhd_v2_updated_conserved_density: out must not alias uc");
    if (!v2_shape_ok(n1, n2, n3, 1)) return fail(M3P_ERR_BAD_SHAPE, "m3p_mhd_v2_updated_conserved_density: empty block");
    hipLaunchKernelGGL(mhd_v2_update_cells_kernel, v2_grid((long)n1 * n2 * n3, 1), dim3(256), 0, (hipStream_t)stream,
                       n1, n2, n3, uc, gf1, gf2, gf3, dt, dl, out);
    return launch_status("mhd_v2_update_cells_kernel");
}

int m3p_mhd_v2_curl(const double* ee1, const double* ee2, const double* ee3, int n1, int n2, int n3, double dl,
                    double* bf1, double* bf2, double* bf3, void* stream)
{
    if (!ee1 || !ee2 || !ee3 || !bf1 || !bf2 || !bf3) return fail(M3P_ERR_BAD_ARGUMENT, "m3p_mhd_v2_curl: NULL array");
    if (!v2_shape_ok(n1, n2, n3, 1)) return fail(M3P_ERR_BAD_SHAPE, "m3p_mhd_v2_curl: empty block");
    hipLaunchKernelGGL(mhd_v2_faces_kernel, v2_grid((long)(n1 + 1) * (n2 + 1) * (n3 + 1), 3), dim3(256), 0,
                       (hipStream_t)stream, n1, n2, n3, 0, ee1, ee2, ee3, (const double*)nullptr,
                       (const double*)nullptr, (const double*)nullptr, 1.0, dl, bf1, bf2, bf3);
    return launch_status("mhd_v2_faces_kernel");
}

int m3p_mhd_v2_updated_magnetic_flux_density(const double* bf1, const double* bf2, const double* bf3,
                                             const double* ee1, const double* ee2, const double* ee3,
                                             int n1, int n2, int n3, double dt, double dl,
                                             double* bo1, double* bo2, double* bo3, void* stream)
{
    if (!bf1 || !bf2 || !bf3 || !ee1 || !ee2 || !ee3 || !bo1 || !bo2 || !bo3)
        return fail(M3P_ERR_BAD_ARGUMENT, "m3p_mhd_v2_updated_magnetic_flux_density: NULL array");
    if (bo1 == bf1 || bo2 == bf2 || bo3 == bf3)
        return fail(M3P_ERR_BAD_ARGUMENT, "m3p_mhd_v2_updated_magnetic_flux_density: outputs must not alias inputs");
    if (!v2_shape_ok(n1, n2, n3, 1)) return fail(M3P_ERR_BAD_SHAPE, "m3p_mhd_v2_updated_magnetic_flux_density: empty block");
    hipLaunchKernelGGL(mhd_v2_faces_kernel, v2_grid((long)(n1 + 1) * (n2 + 1) * (n3 + 1), 3), dim3(256), 0,
                       (hipStream_t)stream, n1, n2, n3, 1, ee1, ee2, ee3, bf1, bf2, bf3, dt, dl, bo1, bo2, bo3);
    return launch_status("mhd_v2_faces_kernel");
}

static int v2_blocks(const double* const* src, Ptr27& dst, const char* what)
{
    if (!src) return fail(M3P_ERR_BAD_ARGUMENT, std::string(what) + ": NULL block list");
    for (int m = 0; m < 27; ++m) {
        if (!src[m]) return fail(M3P_ERR_BAD_ARGUMENT, std::string(what) + ": NULL block pointer");
        dst.p[m] = src[m];
    }
    return M3P_OK;
}

int m3p_mhd_v2_extend_cell_primitive_variables(const double* const* pc_blocks, int block_size, int count,
                                               double* out, void* stream)
{
    const char* what = "m3p_mhd_v2_extend_cell_primitive_variables";
    Ptr27 blk;
    if (int st = v2_blocks(pc_blocks, blk, what)) return st;
    if (!out) return fail(M3P_ERR_BAD_ARGUMENT, std::string(what) + ": NULL output");
    if (block_size < 1 || count < 0 || count > block_size)
        return fail(M3P_ERR_BAD_SHAPE, std::string(what) + ": need 0 <= count <= block_size");
    const long m = block_size + 2L * count;
    hipLaunchKernelGGL(mhd_v2_extend_cells_kernel, v2_grid(m * m * m * 8, 1), dim3(256), 0, (hipStream_t)stream,
                       blk, block_size, count, 8, out);
    return launch_status("mhd_v2_extend_cells_kernel");
}

int m3p_mhd_v2_extend_face_magnetic_flux_density(const double* const* bf_blocks, int block_size, int count,
                                                 double* bf1, double* bf2, double* bf3, void* stream)
{
    const char* what = "m3p_mhd_v2_extend_face_magnetic_flux_density";
    Ptr27 blk;
    if (int st = v2_blocks(bf_blocks, blk, what)) return st;
    if (!bf1 || !bf2 || !bf3) return fail(M3P_ERR_BAD_ARGUMENT, std::string(what) + ": NULL output");
    if (block_size < 1 || count < 0 || count > block_size)
        return fail(M3P_ERR_BAD_SHAPE, std::string(what) + ": need 0 <= count <= block_size");
    const long m = block_size + 2L * count;
    hipLaunchKernelGGL(mhd_v2_extend_faces_kernel, v2_grid((block_size + 1L) * m * m, 3), dim3(256), 0,
                       (hipStream_t)stream, blk, block_size, count, bf1, bf2, bf3);
    return launch_status("mhd_v2_extend_faces_kernel");
}

} // extern "C"
