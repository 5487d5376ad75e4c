// foort_gpu.hip -- gfx950 kernels and the C ABI of include/foort_gpu.h.
//
// Build (see foort_amd/build.py): hipcc --offload-arch=gfx950 -O3 -std=c++17
// -ffp-contract=off -Xclang -target-feature -Xclang -packed-fp32-ops -fPIC -shared.
// Contraction must stay off: the finite-difference path reproduces the CPU
// oracle's roundings exactly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "foort_gpu.h"
#include "foort_kernels.hpp"

#pragma clang fp contract(off)

// Work memory of foort_gpu_trace_device_noalloc, for callers that cannot allocate: the batch
// counters (the first 256 bytes) and the lane slots of up to kScratchLanes resident lanes of
// the persistent schedule (144 bytes each), 18 MiB in all.  Every GPU the library runs on holds
// its own copy until the process ends; host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device (scratch_address below), never through
// the array's name.
namespace {
constexpr uint32_t kScratchLanes = 1u << 17;
constexpr size_t kScratchCounterBytes = 256;
constexpr size_t kScratchBytes =
    kScratchCounterBytes + size_t(kScratchLanes) * (foort::kSlotD * 8 + 8 + foort::kSlotU * 4);
} // namespace
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

using foort::DevResults;
using foort::GeoState;
using foort::KernelPair;
using foort::LaneSlots;
using foort::RunConst;

namespace {

// Reset the lane slots and counters of the persistent schedule (a kernel, not
// a memset, so the sequence can be captured in a graph).
__global__ void __launch_bounds__(256) reset_lanes(foort::LaneSlots S, unsigned long long *total_steps)
{
    const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < S.L)
        S.pix[l] = -1;
    if (l == 0) {
        *S.next = 0;
        *S.done = 0;
        *total_steps = 0;
        total_steps[foort::kCntFuture] = 0;
        total_steps[foort::kCntOutward] = 0;
        total_steps[foort::kCntNonfinite] = 0;
    }
}

// The initial velocities of a batch's pixels that the reference warns about
// (FOORT/ViewScreen.cpp:112-116), counted in a pass of their own (the integration kernels carry
// no code for it): one ballot and one atomic per wavefront.
__global__ void __launch_bounds__(256) count_ic_warnings(const RunConst R, uint64_t side, uint64_t first,
                                                         uint64_t count, const uint64_t *si, unsigned long long *c)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    bool fut = false, outw = false;
    if (i < count) {
        double ux, uy;
        foort::pixel_unit(side, first, si, i, ux, uy);
        GeoState g;
        foort::geo_init<true>(R, ux, uy, g);
        fut = g.vel[0] > 0;
        outw = g.vel[1] > 0;
    }
    const unsigned long long bf = __ballot(fut), bo = __ballot(outw);
    if ((threadIdx.x & 63) == 0) {
        if (bf)
            atomicAdd(c + foort::kCntFuture, static_cast<unsigned long long>(__popcll(bf)));
        if (bo)
            atomicAdd(c + foort::kCntOutward, static_cast<unsigned long long>(__popcll(bo)));
    }
}

// Geodesics whose final position or velocity is NaN or infinite (x - x is 0 only for a finite x),
// from the result arrays.
__global__ void __launch_bounds__(256) count_nonfinite(const double *pos, const double *vel, uint64_t count,
                                                       unsigned long long *c)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    bool bad = false;
    if (i < count) {
        double z = 0;
        for (int k = 0; k < 4; ++k)
            z += (pos[4 * i + k] - pos[4 * i + k]) + (vel[4 * i + k] - vel[4 * i + k]);
        bad = z != 0;
    }
    const unsigned long long b = __ballot(bad);
    if ((threadIdx.x & 63) == 0 && b)
        atomicAdd(c + foort::kCntNonfinite, static_cast<unsigned long long>(__popcll(b)));
}

// the warning counters of a batch (see kCntFuture) into the caller's statistics
void set_warnings(foort_stats *stats, const unsigned long long *c)
{
    stats->future_pointing = c[foort::kCntFuture];
    stats->outward_pointing = c[foort::kCntOutward];
    stats->nonfinite = c[foort::kCntNonfinite];
    stats->graph_launches = 0;
}

__global__ void __launch_bounds__(64) zero_counters(unsigned long long *p, int n)
{
    if (static_cast<int>(threadIdx.x) < n)
        p[threadIdx.x] = 0;
}

// After a queue launch: lanes that found the ring empty advanced the claim counter past
// the pixels pushed so far without taking any; later pushes start at `count`.
__global__ void __launch_bounds__(64) clamp_claims(unsigned long long *next, unsigned long long count)
{
    if (threadIdx.x == 0 && *next > count)
        *next = count;
}

thread_local std::string g_open_error;

uint32_t clamp_u32(uint64_t v)
{
    return v > 0xffffffffull ? 0xffffffffu : static_cast<uint32_t>(v);
}

// ---------------------------------------------------------------------------
// Host-side set-up
// ---------------------------------------------------------------------------

// Metric constructors, FOORT/Metric.cpp:111-128, :205-215, :244-271, :378-407, :490-509
bool init_metric(const foort_run_desc &d, foort::MetricParams &M, std::string &err)
{
    const double *P = d.metric_params;
    std::memset(&M, 0, sizeof(M));
    M.rlog = d.rlog_scale ? 1 : 0;
    M.has_horizon = 1;
    switch (d.metric) {
    case FOORT_METRIC_KERR:
        M.a = P[0];
        M.horizon = 1 + std::sqrt(1 - P[0] * P[0]);
        break;
    case FOORT_METRIC_FLATSPACE:
        M.has_horizon = 0;
        M.rlog = 0;
        break;
    case FOORT_METRIC_RASHEEDLARSEN: {
        const double mP = P[0], aP = P[1], pP = P[2], qP = P[3];
        M.a = aP / ((pP + qP) / 4.0);
        M.m = mP / ((pP + qP) / 4.0);
        M.p = pP / ((pP + qP) / 4.0);
        M.q = qP / ((pP + qP) / 4.0);
        // the constant square roots of FOORT/Metric.cpp:288-298 (and :333-343), in this unit's
        // contraction-free arithmetic
        M.rl_root_pq = std::sqrt(M.p * M.q);
        M.rl_root_pm = std::sqrt((M.p * M.p - 4. * M.m * M.m) * (M.q * M.q - 4. * M.m * M.m));
        M.horizon = (mP + std::sqrt(mP * mP - aP * aP)) / ((pP + qP) / 4.0);
        break;
    }
    case FOORT_METRIC_JOHANNSEN:
        M.a = P[0];
        M.a13 = P[1];
        M.a22 = P[2];
        M.a52 = P[3];
        M.eps3 = P[4];
        M.horizon = 1 + std::sqrt(1 - P[0] * P[0]);
        break;
    case FOORT_METRIC_MANKONOVIKOV:
        M.a = P[0];
        M.alpha3 = P[1];
        M.mn_alpha = (P[0] == 0.0) ? 0.0 : (-1. + std::sqrt(1. - P[0] * P[0])) / P[0];
        M.mn_k = std::sqrt(1. - P[0] * P[0]);
        M.horizon = 1 + std::sqrt(1. - P[0] * P[0]);
        break;
    default:
        err = "unknown metric id " + std::to_string(d.metric);
        return false;
    }
    return true;
}

double inner(const double g[4][4], const double *a, const double *b)
{
    double ret = 0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            ret += a[i] * g[i][j] * b[j];
    return ret;
}

void expand(const double *g5, double g[4][4])
{
    std::memset(g, 0, sizeof(double) * 16);
    g[0][0] = g5[0];
    g[0][3] = g5[1];
    g[3][0] = g5[1];
    g[1][1] = g5[2];
    g[2][2] = g5[3];
    g[3][3] = g5[4];
}

// ViewScreen::ConstructVielbein, FOORT/ViewScreen.cpp:10-73: Gram-Schmidt on
// the rows of g^uu in the order t, phi, r, theta.  The theta row keeps the
// reference's form: its e_r projection is a discarded statement and its e_phi
// projection multiplies the still-zero theta row (FOORT/ViewScreen.cpp:56-58).
void build_camera(int metric, RunConst &R, const double cam[4])
{
    double s, c;
    fm_sincos(cam[2], &s, &c);
    double g5[5], gdd[4][4], guu[4][4];
    foort::metric_dd_rt(metric, R.M, R.cam_u, s, c, g5);
    expand(g5, gdd);
    foort::metric_uu_rt(metric, R.M, R.cam_u, s, c, g5);
    expand(g5, guu);
    double V[4][4] = {};
    double sg[4] = {0, 0, 0, 0};
    double n;
    // e_t
    for (int i = 0; i < 4; ++i)
        V[0][i] = guu[0][i];
    n = inner(gdd, V[0], V[0]);
    sg[0] = (n < 0) ? -1 : 1;
    {
        const double t = 1 / std::sqrt(sg[0] * n);
        for (int i = 0; i < 4; ++i)
            V[0][i] = (V[0][i] * sg[0]) * t;
    }
    // e_phi
    {
        const double c0 = sg[0] * inner(gdd, guu[3], V[0]);
        for (int i = 0; i < 4; ++i)
            V[3][i] = guu[3][i] - V[0][i] * c0;
        n = inner(gdd, V[3], V[3]);
        sg[3] = (n < 0) ? -1 : 1;
        const double sc = sg[3] * cam[1] * fm_sin(cam[2]);
        const double t = 1 / std::sqrt(sg[3] * n);
        for (int i = 0; i < 4; ++i)
            V[3][i] = (V[3][i] * sc) * t;
    }
    // e_r
    {
        const double c0 = sg[0] * inner(gdd, guu[1], V[0]);
        const double c3 = sg[3] * inner(gdd, guu[1], V[3]);
        for (int i = 0; i < 4; ++i)
            V[1][i] = (guu[1][i] - V[0][i] * c0) - V[3][i] * c3;
        n = inner(gdd, V[1], V[1]);
        sg[1] = (n < 0) ? -1 : 1;
        const double t = 1 / std::sqrt(sg[1] * n);
        for (int i = 0; i < 4; ++i)
            V[1][i] = (V[1][i] * sg[1]) * t;
    }
    // e_theta
    {
        const double c0 = sg[0] * inner(gdd, guu[2], V[0]);
        const double c3 = sg[3] * inner(gdd, guu[2], V[3]);
        double row[4];
        for (int i = 0; i < 4; ++i)
            row[i] = (guu[2][i] - V[0][i] * c0) - V[2][i] * c3;
        for (int i = 0; i < 4; ++i)
            V[2][i] = row[i];
        n = inner(gdd, V[2], V[2]);
        sg[2] = (n < 0) ? -1 : 1;
        const double sc = sg[2] * cam[1];
        const double t = 1 / std::sqrt(sg[2] * n);
        for (int i = 0; i < 4; ++i)
            V[2][i] = (V[2][i] * sc) * t;
    }
    std::memcpy(R.V, V, sizeof(V));
    for (int i = 0; i < 4; ++i)
        R.g0[i] = gdd[0][i];
}

} // namespace

struct foort_gpu_ctx {
    foort_run_desc desc;
    RunConst rc;
    KernelPair kernels{};
    int device = 0;
    int num_cu = 0;
    std::string err;
    // grow-only device staging for the host entry points
    void *d_buf = nullptr;
    size_t d_cap = 0;
    unsigned long long *d_counters = nullptr; // [0] steps, [1] next pixel, [2] pixels done
    unsigned long long *d_steps = nullptr;
    // device copy of rc, in the same allocation (the resident kernels' refills read the camera
    // from it); written at open and by every configure that changes rc
    RunConst *d_rc = nullptr;
    // persistent schedule
    LaneSlots slots{};
    void *d_slots = nullptr;
    uint64_t budget = 0;          // steps per launch (UINT64_MAX: no limit)
    int clock_khz = 0;            // wall-clock rate (for the launch time limit)
    uint64_t last_launches = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // captured batches (use_graph): the context's own stream, pinned host staging with the
    // layout of d_buf, and the executable graph that every batch's capture updates
    hipStream_t stream = nullptr;
    void *h_buf = nullptr;
    size_t h_cap = 0;
    hipGraphExec_t graph = nullptr;
    uint64_t graph_updates = 0;   // batches whose capture updated the executable graph in place
    bool last_graph = false;      // the last host-pointer batch ran as a graph launch
    // the device address of `scratch` on this context's device (NULL, with the reason in
    // scratch_err, if it could not be found)
    char *scratch = nullptr;
    std::string scratch_err;
};

namespace {

int fail(foort_gpu_ctx *ctx, int code, const std::string &msg)
{
    if (ctx)
        ctx->err = msg;
    else
        g_open_error = msg;
    return code;
}

int hip_fail(foort_gpu_ctx *ctx, hipError_t e, const char *what)
{
    return fail(ctx, FOORT_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(ctx, expr)                         \
    do {                                           \
        hipError_t e_ = (expr);                    \
        if (e_ != hipSuccess)                      \
            return hip_fail((ctx), e_, #expr);     \
    } while (0)

int ensure_buffer(foort_gpu_ctx *ctx, size_t bytes)
{
    if (bytes <= ctx->d_cap)
        return FOORT_OK;
    bytes = std::max(bytes, ctx->d_cap + ctx->d_cap / 2); // grow geometrically
    if (ctx->d_buf)
        (void)hipFree(ctx->d_buf);
    ctx->d_buf = nullptr;
    ctx->d_cap = 0;
    hipError_t e = hipMalloc(&ctx->d_buf, bytes);
    if (e != hipSuccess)
        return fail(ctx, FOORT_ENOMEM, std::string("hipMalloc of staging buffer: ") + hipGetErrorString(e));
    ctx->d_cap = bytes;
    return FOORT_OK;
}

int ensure_pinned(foort_gpu_ctx *ctx, size_t bytes)
{
    if (bytes <= ctx->h_cap)
        return FOORT_OK;
    bytes = std::max(bytes, ctx->h_cap + ctx->h_cap / 2);
    if (ctx->h_buf)
        (void)hipHostFree(ctx->h_buf);
    ctx->h_buf = nullptr;
    ctx->h_cap = 0;
    hipError_t e = hipHostMalloc(&ctx->h_buf, bytes, hipHostMallocDefault);
    if (e != hipSuccess)
        return fail(ctx, FOORT_ENOMEM, std::string("hipHostMalloc of staging buffer: ") + hipGetErrorString(e));
    ctx->h_cap = bytes;
    return FOORT_OK;
}

void drop_graph(foort_gpu_ctx *ctx)
{
    if (ctx->graph)
        (void)hipGraphExecDestroy(ctx->graph);
    ctx->graph = nullptr;
}

// the persistent schedule's options, from the descriptor (the autotuner changes them)
void apply_lane_options(foort_gpu_ctx *ctx)
{
    ctx->budget = ctx->desc.steps_per_launch ? ctx->desc.steps_per_launch : UINT64_MAX;
    ctx->slots.refill = ctx->desc.refill_lanes ? std::min(64u, ctx->desc.refill_lanes) : 1u;
    ctx->slots.interval = static_cast<uint32_t>(
        std::min<uint64_t>(ctx->budget, ctx->desc.refill_interval ? ctx->desc.refill_interval : 32u));
    ctx->slots.ticks = static_cast<uint64_t>(ctx->desc.launch_us) * static_cast<uint64_t>(ctx->clock_khz) / 1000u;
    if (ctx->desc.launch_us && ctx->slots.ticks == 0)
        ctx->slots.ticks = 1;
    // the four-lane tail: finite differences, a batch that is one launch (set in launch())
    ctx->slots.tail = 0;
}

// geodesics left to the tail kernel (0: none), for L resident lanes
uint64_t tail_threshold(const foort_gpu_ctx *ctx, uint32_t L)
{
    if (!ctx->kernels.tail || ctx->desc.tail_geodesics == UINT32_MAX)
        return 0;
    return ctx->desc.tail_geodesics ? ctx->desc.tail_geodesics : std::max<uint32_t>(1, L / 8);
}

// Resident lanes of the persistent schedule: lanes_per_cu, or as many as the persistent
// kernel's occupancy allows, on every CU; a multiple of 256 (the workgroup size).
int lane_count(foort_gpu_ctx *ctx, uint64_t &L)
{
    uint32_t lanes_per_cu = ctx->desc.lanes_per_cu;
    if (lanes_per_cu == 0) {
        int blocks = 0;
        HIP_TRY(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(
                         &blocks, reinterpret_cast<const void *>(ctx->kernels.persistent), 256, 0));
        lanes_per_cu = static_cast<uint32_t>(std::max(1, blocks)) * 256u;
    }
    L = static_cast<uint64_t>(ctx->num_cu) * lanes_per_cu;
    L = (L + 255) / 256 * 256;
    if (L == 0 || L > (1u << 24))
        return fail(ctx, FOORT_EINVAL, "persistent lane count out of range");
    return FOORT_OK;
}

// offset of the device copy of the run constants in a context's counter allocation
constexpr size_t kRunConstOffset = 256;

// bytes of lane state per resident lane: kSlotD doubles, the pixel index, kSlotU words
constexpr size_t kLaneBytes = foort::kSlotD * 8 + 8 + foort::kSlotU * 4;

// Point the lane slots of S at L lanes of state starting at `base` (256-byte aligned), with
// the claim and done counters of the batch counters c.
void place_slots(LaneSlots &S, char *base, uint64_t L, unsigned long long *c, const RunConst *rc)
{
    S.rc = rc;
    S.d = reinterpret_cast<double *>(base);
    S.pix = reinterpret_cast<long long *>(base + L * foort::kSlotD * 8);
    S.u = reinterpret_cast<uint32_t *>(base + L * (foort::kSlotD * 8 + 8));
    S.next = c + 1;
    S.done = c + 2;
    S.L = static_cast<uint32_t>(L);
}

int setup_persistent(foort_gpu_ctx *ctx)
{
    apply_lane_options(ctx);
    if (ctx->d_slots)
        return FOORT_OK;
    uint64_t L = 0;
    int rc = lane_count(ctx, L);
    if (rc)
        return rc;
    hipError_t e = hipMalloc(&ctx->d_slots, L * kLaneBytes);
    if (e != hipSuccess)
        return fail(ctx, FOORT_ENOMEM, std::string("hipMalloc of lane slots: ") + hipGetErrorString(e));
    place_slots(ctx->slots, static_cast<char *>(ctx->d_slots), L, ctx->d_counters, ctx->d_rc);
    return FOORT_OK;
}

// The device memory a batch's kernels use besides its inputs and results: the batch
// counters (steps, claimed and finished pixels, warnings; kNumCounters) and, for the
// persistent schedule, the lane slots and the schedule's options.
struct Work {
    unsigned long long *counters;
    LaneSlots slots;
};

// The context's own work memory (the persistent schedule's lane slots are allocated on
// first use).
int ctx_work(foort_gpu_ctx *ctx, Work &w)
{
    if (ctx->desc.schedule == FOORT_SCHEDULE_PERSISTENT) {
        int rc = setup_persistent(ctx);
        if (rc)
            return rc;
    }
    w.counters = ctx->d_counters;
    w.slots = ctx->slots;
    return FOORT_OK;
}

// a batch is one kernel launch (and needs no host decisions between launches)
bool single_launch(const foort_gpu_ctx *ctx)
{
    return ctx->desc.schedule == FOORT_SCHEDULE_ONESHOT ||
           (ctx->desc.steps_per_launch == 0 && ctx->desc.launch_us == 0);
}

// The address of `scratch` on each device, looked up once per device with that device current.
std::mutex g_scratch_mu;
std::map<int, char *> g_scratch_at;

hipError_t scratch_address(int device, char **out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch_at.find(device);
    if (it != g_scratch_at.end()) {
        *out = it->second;
        return hipSuccess;
    }
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return e != hipSuccess ? e : hipErrorInvalidSymbol;
    g_scratch_at[device] = static_cast<char *>(p);
    *out = static_cast<char *>(p);
    return hipSuccess;
}

// The work memory in `scratch` (no allocation): the counters, and for the persistent schedule
// the lane slots of at most kScratchLanes lanes -- the occupancy's lane count capped there, or
// lanes_per_cu on every CU, which must fit.
int scratch_work(foort_gpu_ctx *ctx, Work &w)
{
    if (!ctx->scratch)
        return fail(ctx, FOORT_EDEVICE, "no-allocation work memory unavailable: " + ctx->scratch_err);
    int dev = -1;
    HIP_TRY(ctx, hipGetDevice(&dev));
    if (dev != ctx->device)
        return fail(ctx, FOORT_EINVAL, "the current device is not the context's");
    w.counters = reinterpret_cast<unsigned long long *>(ctx->scratch);
    w.slots = ctx->slots;
    if (ctx->desc.schedule != FOORT_SCHEDULE_PERSISTENT)
        return FOORT_OK;
    apply_lane_options(ctx);
    w.slots = ctx->slots;
    uint64_t L = 0;
    int rc = lane_count(ctx, L);
    if (rc)
        return rc;
    if (L > kScratchLanes) {
        if (ctx->desc.lanes_per_cu)
            return fail(ctx, FOORT_EINVAL,
                        "lanes_per_cu gives " + std::to_string(L) + " resident lanes; the no-allocation entry point "
                        "holds at most " + std::to_string(kScratchLanes));
        L = kScratchLanes;
    }
    place_slots(w.slots, ctx->scratch + kScratchCounterBytes, L, w.counters, ctx->d_rc);
    return FOORT_OK;
}

// Queue the kernels of one batch on `stream`, with the work memory w.  A batch that is one
// launch is only queued; otherwise the persistent schedule relaunches until the device counts
// every pixel done.
int launch_trace(foort_gpu_ctx *ctx, const Work &w, uint64_t side, uint64_t first, uint64_t count,
                 const uint64_t *si, const DevResults &out, hipStream_t stream)
{
    ctx->last_launches = 0;
    if (ctx->desc.schedule == FOORT_SCHEDULE_ONESHOT) {
        hipLaunchKernelGGL(zero_counters, dim3(1), dim3(64), 0, stream, w.counters, foort::kNumCounters);
        HIP_TRY(ctx, hipGetLastError());
        if (count == 0)
            return FOORT_OK;
        const int block = ctx->desc.block_size > 0 ? ctx->desc.block_size : 256;
        if (block > 256 || block % 64 != 0)
            return fail(ctx, FOORT_EINVAL, "block_size must be a multiple of 64 and at most 256");
        const uint64_t grid = (count + block - 1) / block;
        if (grid > 0x7fffffffull)
            return fail(ctx, FOORT_EINVAL, "batch too large");
        hipLaunchKernelGGL(ctx->kernels.oneshot, dim3(static_cast<unsigned>(grid)), dim3(block), 0, stream, ctx->rc,
                           side, first, count, si, out, w.counters, w.slots, uint64_t(0));
        HIP_TRY(ctx, hipGetLastError());
        ctx->last_launches = 1;
        return FOORT_OK;
    }
    const LaneSlots &S = w.slots;
    hipLaunchKernelGGL(reset_lanes, dim3(S.L / 256), dim3(256), 0, stream, S, w.counters);
    HIP_TRY(ctx, hipGetLastError());
    if (count == 0)
        return FOORT_OK;
    if (single_launch(ctx)) {
        // no step or time limit: the launch runs until every pixel is done, or until the last
        // few are left to the four-lane tail kernel
        LaneSlots St = S;
        St.tail = tail_threshold(ctx, S.L);
        hipLaunchKernelGGL(ctx->kernels.persistent, dim3(S.L / 256), dim3(256), 0, stream, ctx->rc, side, first,
                           count, si, out, w.counters, St, ctx->budget);
        HIP_TRY(ctx, hipGetLastError());
        ctx->last_launches = 1;
        if (St.tail) {
            hipLaunchKernelGGL(ctx->kernels.tail, dim3(S.L / 64), dim3(256), 0, stream, ctx->rc, side, first, count,
                               si, out, w.counters, St, ctx->budget);
            HIP_TRY(ctx, hipGetLastError());
            ctx->last_launches = 2;
        }
        return FOORT_OK;
    }
    // every launch advances each busy lane by at least min(budget, interval) steps (or to the end
    // of its geodesic); no geodesic is longer than step_cap, so the launches are bounded
    const uint64_t per_launch = S.ticks ? S.interval : std::min<uint64_t>(ctx->budget, UINT32_MAX);
    const uint64_t max_launches =
        (count / S.L + 2) * (static_cast<uint64_t>(ctx->rc.step_cap) / per_launch + 2) + 16;
    unsigned long long done = 0;
    uint64_t n = 0;
    while (done < count) {
        if (n >= max_launches)
            return fail(ctx, FOORT_EDEVICE, "persistent schedule did not finish within its launch bound");
        hipLaunchKernelGGL(ctx->kernels.persistent, dim3(S.L / 256), dim3(256), 0, stream, ctx->rc, side, first,
                           count, si, out, w.counters, S, ctx->budget);
        HIP_TRY(ctx, hipGetLastError());
        ++n;
        HIP_TRY(ctx, hipMemcpyAsync(&done, S.done, sizeof(done), hipMemcpyDeviceToHost, stream));
        HIP_TRY(ctx, hipStreamSynchronize(stream));
    }
    ctx->last_launches = n;
    return FOORT_OK;
}

// A batch's integration, then the passes that count its warnings (and its non-finite results
// when the final positions and velocities are kept).
int launch(foort_gpu_ctx *ctx, const Work &w, uint64_t side, uint64_t first, uint64_t count, const uint64_t *si,
           const DevResults &out, hipStream_t stream)
{
    int rc = launch_trace(ctx, w, side, first, count, si, out, stream);
    if (rc || count == 0)
        return rc;
    const unsigned grid = static_cast<unsigned>((count + 255) / 256);
    hipLaunchKernelGGL(count_ic_warnings, dim3(grid), dim3(256), 0, stream, ctx->rc, side, first, count, si,
                       w.counters);
    HIP_TRY(ctx, hipGetLastError());
    if (out.pos && out.vel) {
        hipLaunchKernelGGL(count_nonfinite, dim3(grid), dim3(256), 0, stream, out.pos, out.vel, count, w.counters);
        HIP_TRY(ctx, hipGetLastError());
    }
    return FOORT_OK;
}

int check_batch(foort_gpu_ctx *ctx, uint64_t side, uint64_t first, uint64_t count, bool grid)
{
    if (side < 2)
        return fail(ctx, FOORT_EINVAL, "row_column_size must be at least 2");
    if (grid && (first > side * side || count > side * side - first))
        return fail(ctx, FOORT_EINVAL, "grid indices exceed row_column_size^2");
    return FOORT_OK;
}

// Host-pointer entry points: stage through the context buffer, launch, copy back.  With
// use_graph, a batch that is one launch goes through pinned host staging and runs as one HIP
// graph (index upload, counter reset, kernels, result copies): each batch is captured and
// updates the context's executable graph in place, which is then launched.
int trace_host(foort_gpu_ctx *ctx, uint64_t side, uint64_t first, uint64_t count, const uint64_t *si,
               const foort_geodesic_results *out, foort_stats *stats)
{
    if (!out)
        return fail(ctx, FOORT_EINVAL, "results pointer is NULL");
    // layout of the staging buffers: [screen index | term | quadrant | passes | steps | closest | lambda | pos | vel]
    auto align = [](size_t v) { return (v + 255) & ~size_t(255); };
    struct Field {
        void *host;
        size_t off, bytes;
    };
    size_t off = 0;
    const size_t o_si = off;
    off = align(off + (si ? count * 16 : 0));
    Field f[8] = {{out->term, 0, count * 4},         {out->quadrant, 0, count * 4},
                  {out->equat_passes, 0, count * 4}, {out->steps, 0, count * 4},
                  {out->closest_radius, 0, count * 8}, {out->lambda, 0, count * 8},
                  {out->final_pos, 0, count * 32},   {out->final_vel, 0, count * 32}};
    for (Field &x : f) {
        x.off = off;
        off = align(off + x.bytes);
    }
    int rc = ensure_buffer(ctx, off);
    if (rc)
        return rc;
    char *base = static_cast<char *>(ctx->d_buf);
    auto dptr = [&](const Field &x) -> void * { return x.host ? base + x.off : nullptr; };
    DevResults d{};
    d.term = static_cast<int32_t *>(dptr(f[0]));
    d.quadrant = static_cast<int32_t *>(dptr(f[1]));
    d.passes = static_cast<int32_t *>(dptr(f[2]));
    d.steps = static_cast<uint32_t *>(dptr(f[3]));
    d.closest = static_cast<double *>(dptr(f[4]));
    d.lambda = static_cast<double *>(dptr(f[5]));
    d.pos = static_cast<double *>(dptr(f[6]));
    d.vel = static_cast<double *>(dptr(f[7]));
    const uint64_t *d_si = si ? reinterpret_cast<const uint64_t *>(base + o_si) : nullptr;

    // the counters and lane slots (allocated here on first use: never inside a capture)
    Work w{};
    rc = ctx_work(ctx, w);
    if (rc)
        return rc;
    const bool graph = ctx->desc.use_graph && single_launch(ctx) && count > 0;
    if (!graph) {
        // everything on the context's own stream, so contexts on other host threads run concurrently
        if (si)
            HIP_TRY(ctx, hipMemcpyAsync(base + o_si, si, count * 16, hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        rc = launch(ctx, w, side, first, count, d_si, d, ctx->stream);
        if (rc)
            return rc;
        HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
        for (const Field &x : f)
            if (x.host)
                HIP_TRY(ctx, hipMemcpyAsync(x.host, base + x.off, x.bytes, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        ctx->last_graph = false;
    } else {
        // The batch -- index upload, counter reset, kernels, warning counts, result copies -- is
        // captured from the context's stream and run as one graph.  The executable graph of the
        // previous batch is updated in place with this capture (hipGraphExecUpdate: the same
        // nodes with new sizes, grids and arguments), so the batches of an adaptive mesh, whose
        // sizes never repeat, each run as a launch of the same executable graph; it is built anew
        // only when the nodes change (another schedule, other result fields).
        rc = ensure_pinned(ctx, off);
        if (rc)
            return rc;
        char *hb = static_cast<char *>(ctx->h_buf);
        HIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        auto capture = [&]() -> int {
            if (si)
                HIP_TRY(ctx, hipMemcpyAsync(base + o_si, hb + o_si, count * 16, hipMemcpyHostToDevice, ctx->stream));
            int r = launch(ctx, w, side, first, count, d_si, d, ctx->stream);
            if (r)
                return r;
            for (const Field &x : f)
                if (x.host)
                    HIP_TRY(ctx, hipMemcpyAsync(hb + x.off, base + x.off, x.bytes, hipMemcpyDeviceToHost, ctx->stream));
            return FOORT_OK;
        };
        rc = capture();
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (rc || e != hipSuccess) {
            if (g)
                (void)hipGraphDestroy(g);
            return rc ? rc : hip_fail(ctx, e, "hipStreamEndCapture");
        }
        bool updated = false;
        if (ctx->graph) {
            hipGraphNode_t bad = nullptr;
            hipGraphExecUpdateResult res = hipGraphExecUpdateError;
            e = hipGraphExecUpdate(ctx->graph, g, &bad, &res);
            updated = e == hipSuccess && res == hipGraphExecUpdateSuccess;
            if (!updated) {
                (void)hipGetLastError(); // a refused update is not an error: the graph is built anew
                drop_graph(ctx);
            }
        }
        if (!updated) {
            e = hipGraphInstantiate(&ctx->graph, g, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                ctx->graph = nullptr;
                (void)hipGraphDestroy(g);
                return hip_fail(ctx, e, "hipGraphInstantiate");
            }
        } else {
            ++ctx->graph_updates;
        }
        (void)hipGraphDestroy(g);
        if (si)
            std::memcpy(hb + o_si, si, count * 16);
        HIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
        HIP_TRY(ctx, hipGraphLaunch(ctx->graph, ctx->stream));
        HIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
        HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
        for (const Field &x : f)
            if (x.host)
                std::memcpy(x.host, hb + x.off, x.bytes);
        ctx->last_graph = true;
    }
    if (stats) {
        unsigned long long c[foort::kNumCounters] = {};
        HIP_TRY(ctx, hipMemcpyAsync(c, w.counters, sizeof c, hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        float ms = 0;
        HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        stats->geodesics = count;
        stats->steps = c[0];
        stats->launches = ctx->last_launches;
        stats->device_ms = ms;
        set_warnings(stats, c);
        stats->graph_launches = ctx->last_graph ? 1 : 0;
    }
    return FOORT_OK;
}

// Validate a run description and derive what the kernels take: the run constants
// and the kernels.  Returns FOORT_OK, or FOORT_EINVAL with the reason in err.
int build_run(const foort_run_desc &d, RunConst &R, KernelPair &kernels, std::string &err)
{
    if (d.integrator != FOORT_INTEGRATOR_RK4 && d.integrator != FOORT_INTEGRATOR_VERLET) {
        err = "unknown integrator";
        return FOORT_EINVAL;
    }
    if (d.christoffel != FOORT_CHRISTOFFEL_FD && d.christoffel != FOORT_CHRISTOFFEL_ANALYTIC) {
        err = "unknown Christoffel policy";
        return FOORT_EINVAL;
    }
    if (d.precision != FOORT_PRECISION_FP64 && d.precision != FOORT_PRECISION_MIXED) {
        err = "unknown precision";
        return FOORT_EINVAL;
    }
    if (d.precision == FOORT_PRECISION_MIXED && d.christoffel != FOORT_CHRISTOFFEL_ANALYTIC) {
        err = "mixed precision needs the analytic Christoffel policy";
        return FOORT_EINVAL;
    }
    if (d.schedule != FOORT_SCHEDULE_ONESHOT && d.schedule != FOORT_SCHEDULE_PERSISTENT) {
        err = "unknown schedule";
        return FOORT_EINVAL;
    }
    if (!(d.derivative_h > 0)) {
        err = "DerivativeH must be positive";
        return FOORT_EINVAL;
    }
    if (d.step_cap > 0x7fffffffull) {
        err = "step_cap above 2^31-1";
        return FOORT_EINVAL;
    }
    if (!(d.camera_position[1] > 0)) {
        err = "camera radius must be positive";
        return FOORT_EINVAL;
    }

    std::memset(&R, 0, sizeof(R));
    if (!init_metric(d, R.M, err))
        return FOORT_EINVAL;
    R.eps = d.step_size;
    R.minstep = d.smallest_step;
    R.h_fd = d.derivative_h;
    R.inv2h = 1 / (2 * d.derivative_h); // FOORT/Metric.cpp:30: / (2 * h) is * (1 / (2 * h))
    R.verlet_tol = d.verlet_tolerance;
    R.terms = d.term_flags & 0xfu;
    if (!R.M.has_horizon) // Config turns Horizon off for metrics without one (FOORT/Config.cpp:652-658)
        R.terms &= ~FOORT_TERMFLAG_HORIZON;
    R.diags = d.diag_flags & (FOORT_DIAGFLAG_FOURCOLORSCREEN | FOORT_DIAGFLAG_EQUATORIALPASSES | FOORT_DIAGFLAG_CLOSESTRADIUS);
    R.h_eps = d.horizon_epsilon;
    R.sphere_R = d.sphere_radius;
    R.theta_eps = d.theta_epsilon;
    R.h_n = clamp_u32(d.horizon_update_n);
    R.s_n = clamp_u32(d.sphere_update_n);
    R.t_n = clamp_u32(d.timeout_update_n);
    R.th_n = clamp_u32(d.theta_update_n);
    R.max_steps = clamp_u32(d.timeout_max_steps);
    R.step_cap = d.step_cap ? static_cast<uint32_t>(d.step_cap) : 0x7fffffffu;
    R.ep_thr = d.ep_threshold;
    R.ep_n = clamp_u32(d.ep_update.update_n_steps);
    R.ep_start = d.ep_update.update_start;
    R.ep_finish = d.ep_update.update_finish;
    R.cr_n = clamp_u32(d.cr_update.update_n_steps);
    R.cr_start = d.cr_update.update_start;
    R.cr_finish = d.cr_update.update_finish;
    R.cam_t = d.camera_position[0];
    R.cam_r = d.camera_position[1];
    R.cam_u = R.M.rlog ? std::log(d.camera_position[1]) : d.camera_position[1];
    R.cam_theta = d.camera_position[2];
    R.cam_phi = d.camera_position[3];
    R.sin_theta0 = fm_sin(d.camera_position[2]);
    R.size_x = d.screen_size[0];
    R.size_y = d.screen_size[1];
    R.center_x = d.screen_center[0];
    R.center_y = d.screen_center[1];
    R.mixed_u = R.M.rlog ? std::log(d.mixed_switch_radius) : d.mixed_switch_radius;
    R.mixed_pole = d.mixed_pole_margin;
    build_camera(d.metric, R, d.camera_position);
    // kernels without update counters when every termination and diagnostic updates every step
    // (UpdateFrequency 1; for a termination 0 means the same, FOORT/Terminations.cpp:71-90)
    auto term_every = [&R](uint32_t flag, uint32_t n) { return !(R.terms & flag) || n <= 1; };
    const bool every = term_every(FOORT_TERMFLAG_HORIZON, R.h_n) && term_every(FOORT_TERMFLAG_BOUNDARYSPHERE, R.s_n) &&
                       term_every(FOORT_TERMFLAG_TIMEOUT, R.t_n) && term_every(FOORT_TERMFLAG_THETASINGULARITY, R.th_n) &&
                       (!(R.diags & FOORT_DIAGFLAG_EQUATORIALPASSES) || R.ep_n == 1) &&
                       (!(R.diags & FOORT_DIAGFLAG_CLOSESTRADIUS) || R.cr_n == 1);
    kernels = d.christoffel == FOORT_CHRISTOFFEL_FD
                       ? foort::pick_policy<foort::RhsFD>(d.metric, d.integrator, every)
                       : foort::pick_analytic(d.metric, d.integrator, static_cast<int>(d.precision), every);
    return FOORT_OK;
}

// Validate and apply a run description.  Nothing of ctx changes unless it is valid.
int configure(foort_gpu_ctx *ctx, const foort_run_desc &d)
{
    RunConst R;
    KernelPair kernels{};
    std::string err;
    if (build_run(d, R, kernels, err) != FOORT_OK)
        return fail(ctx, FOORT_EINVAL, err);
    if (ctx->d_rc && std::memcmp(&ctx->rc, &R, sizeof R) != 0) {
        // kernels of earlier batches may still read the device copy (on any stream)
        HIP_TRY(ctx, hipDeviceSynchronize());
        if (hipMemcpy(ctx->d_rc, &R, sizeof R, hipMemcpyHostToDevice) != hipSuccess) {
            // leave the copy as the previous description has it
            (void)hipMemcpy(ctx->d_rc, &ctx->rc, sizeof R, hipMemcpyHostToDevice);
            return fail(ctx, FOORT_EDEVICE, "writing the device copy of the run constants failed");
        }
    }
    if (ctx->kernels.persistent != kernels.persistent && ctx->d_slots) {
        // the lane count follows the persistent kernel's occupancy: size the slots again
        (void)hipFree(ctx->d_slots);
        ctx->d_slots = nullptr;
    }
    drop_graph(ctx);
    ctx->desc = d;
    ctx->rc = R;
    ctx->kernels = kernels;
    ctx->budget = d.steps_per_launch ? d.steps_per_launch : UINT64_MAX;
    return FOORT_OK;
}

// What must be equal for requests to share a queue: everything the step loop
// reads except the metric parameters (the camera and screen only start geodesics).
std::string group_key(const foort_run_desc &d, const RunConst &R, const KernelPair &k)
{
    std::string key;
    auto add = [&key](const void *p, size_t n) { key.append(static_cast<const char *>(p), n); };
    add(&d.metric, sizeof d.metric);
    add(&k.queue, sizeof k.queue);
    const double dv[] = {R.eps, R.minstep, R.h_fd, R.inv2h, R.verlet_tol, R.h_eps, R.sphere_R, R.theta_eps,
                         R.ep_thr, R.mixed_u, R.mixed_pole};
    add(dv, sizeof dv);
    const uint32_t uv[] = {R.terms, R.diags, R.h_n, R.s_n, R.t_n, R.th_n, R.max_steps, R.step_cap, R.ep_n, R.cr_n,
                           static_cast<uint32_t>(R.ep_start), static_cast<uint32_t>(R.ep_finish),
                           static_cast<uint32_t>(R.cr_start), static_cast<uint32_t>(R.cr_finish),
                           static_cast<uint32_t>(R.M.rlog), static_cast<uint32_t>(R.M.has_horizon)};
    add(uv, sizeof uv);
    return key;
}

// foort_gpu_trace_device and its no-allocation variant: the same kernels, with the context's
// work memory or with `scratch`.
int trace_device(foort_gpu_ctx *ctx, uint64_t side, uint64_t first, uint64_t count, const uint64_t *screen_index,
                 const foort_geodesic_results *out, void *stream, foort_stats *stats, bool noalloc)
{
    if (!ctx)
        return fail(nullptr, FOORT_EINVAL, "NULL context");
    if (!out)
        return fail(ctx, FOORT_EINVAL, "results pointer is NULL");
    int rc = check_batch(ctx, side, first, count, screen_index == nullptr);
    if (rc)
        return rc;
    Work w{};
    rc = noalloc ? scratch_work(ctx, w) : ctx_work(ctx, w);
    if (rc)
        return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    DevResults d{out->term, out->quadrant, out->equat_passes, out->closest_radius,
                 out->final_pos, out->final_vel, out->lambda, out->steps};
    if (stats)
        HIP_TRY(ctx, hipEventRecord(ctx->ev0, s));
    rc = launch(ctx, w, side, first, count, screen_index, d, s);
    if (rc)
        return rc;
    if (stats) {
        HIP_TRY(ctx, hipEventRecord(ctx->ev1, s));
        HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
        unsigned long long c[foort::kNumCounters] = {};
        HIP_TRY(ctx, hipMemcpy(c, w.counters, sizeof c, hipMemcpyDeviceToHost));
        float ms = 0;
        HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        stats->geodesics = count;
        stats->steps = c[0];
        stats->launches = ctx->last_launches;
        stats->device_ms = ms;
        set_warnings(stats, c);
    }
    return FOORT_OK;
}

} // namespace

struct foort_gpu_queue {
    foort_gpu_ctx *ctx = nullptr; // the group's kernels, stream, counters and lane state
    std::string key;
    uint64_t cap = 0, mask = 0;   // ring capacity (a power of two)
    uint64_t count = 0;           // pixels pushed so far (a monotonic position)
    std::map<uint64_t, uint64_t> unread; // pushed ranges whose results were not read yet: first -> count
    uint32_t max_req = 0;
    std::vector<char> used;
    RunConst *d_reqc = nullptr;
    uint64_t *d_req_side = nullptr;
    uint32_t *d_req_of = nullptr;
    uint64_t *d_si = nullptr;
    unsigned long long *d_req_done = nullptr;
    int32_t *d_term = nullptr, *d_quad = nullptr, *d_pass = nullptr;
    double *d_close = nullptr;
    void *d_block = nullptr;      // one allocation holding all of the above
    bool lanes_reset = false;
};

extern "C" {

void foort_desc_init(foort_run_desc *d)
{
    if (!d)
        return;
    std::memset(d, 0, sizeof(*d));
    d->metric = FOORT_METRIC_KERR;
    d->metric_params[0] = 0.5;
    d->integrator = FOORT_INTEGRATOR_RK4;
    d->christoffel = FOORT_CHRISTOFFEL_FD;
    d->step_size = 0.03;
    d->smallest_step = 1e-12;
    d->derivative_h = 1e-7;
    d->verlet_tolerance = 0.001;
    d->term_flags = FOORT_TERMFLAG_BOUNDARYSPHERE | FOORT_TERMFLAG_TIMEOUT;
    d->precision = FOORT_PRECISION_FP64;
    d->horizon_epsilon = 0.01;
    d->horizon_update_n = 1;
    d->sphere_radius = 1000;
    d->sphere_update_n = 1;
    d->timeout_max_steps = 10000;
    d->timeout_update_n = 1;
    d->theta_epsilon = 1e-5;
    d->theta_update_n = 1;
    d->diag_flags = FOORT_DIAGFLAG_FOURCOLORSCREEN;
    d->value_diag = FOORT_DIAGFLAG_FOURCOLORSCREEN;
    d->ep_threshold = 0.01;
    d->ep_update = foort_update_frequency{1, 1, 1};
    d->cr_update = foort_update_frequency{1, 1, 1};
    d->camera_position[1] = 1000;
    d->camera_position[2] = foort::kPi / 2.0;
    d->screen_size[0] = 10;
    d->screen_size[1] = 10;
    d->schedule = FOORT_SCHEDULE_ONESHOT;
    d->mixed_switch_radius = 10;
    d->mixed_pole_margin = 0.1;
    d->gp_update = foort_update_frequency{1, 0, 1};
}

int foort_desc_size(void)
{
    return static_cast<int>(sizeof(foort_run_desc));
}

int foort_gpu_open(const foort_run_desc *desc, foort_gpu_ctx **out)
{
    if (!desc || !out)
        return fail(nullptr, FOORT_EINVAL, "NULL descriptor or context pointer");
    *out = nullptr;
    foort_gpu_ctx *ctx = new foort_gpu_ctx();
    int rc = configure(ctx, *desc);
    if (rc) {
        g_open_error = ctx->err;
        delete ctx;
        return rc;
    }

    hipError_t e = hipGetDevice(&ctx->device);
    if (e == hipSuccess)
        e = hipDeviceGetAttribute(&ctx->num_cu, hipDeviceAttributeMultiprocessorCount, ctx->device);
    if (e == hipSuccess)
        e = hipDeviceGetAttribute(&ctx->clock_khz, hipDeviceAttributeWallClockRate, ctx->device);
    if (e == hipSuccess)
        e = hipMalloc(reinterpret_cast<void **>(&ctx->d_counters), kRunConstOffset + sizeof(RunConst));
    if (e == hipSuccess) {
        ctx->d_steps = ctx->d_counters;
        ctx->d_rc = reinterpret_cast<RunConst *>(reinterpret_cast<char *>(ctx->d_counters) + kRunConstOffset);
        e = hipMemcpy(ctx->d_rc, &ctx->rc, sizeof(RunConst), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess)
        e = hipEventCreate(&ctx->ev0);
    if (e == hipSuccess)
        e = hipEventCreate(&ctx->ev1);
    if (e == hipSuccess)
        e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) {
        // the no-allocation entry point's work memory on this device (only that entry point
        // needs it, so a failure here is reported there)
        const hipError_t es = scratch_address(ctx->device, &ctx->scratch);
        if (es != hipSuccess) {
            ctx->scratch = nullptr;
            ctx->scratch_err = std::string("hipGetSymbolAddress: ") + hipGetErrorString(es);
        }
    }
    if (e != hipSuccess) {
        std::string msg = std::string("HIP initialisation: ") + hipGetErrorString(e);
        foort_gpu_close(ctx);
        return fail(nullptr, FOORT_EDEVICE, msg);
    }
    *out = ctx;
    return FOORT_OK;
}

int foort_gpu_configure(foort_gpu_ctx *ctx, const foort_run_desc *desc)
{
    if (!ctx || !desc)
        return fail(ctx, FOORT_EINVAL, "NULL context or descriptor");
    return configure(ctx, *desc);
}

void foort_gpu_close(foort_gpu_ctx *ctx)
{
    if (!ctx)
        return;
    drop_graph(ctx);
    if (ctx->stream)
        (void)hipStreamDestroy(ctx->stream);
    if (ctx->h_buf)
        (void)hipHostFree(ctx->h_buf);
    if (ctx->d_buf)
        (void)hipFree(ctx->d_buf);
    if (ctx->d_counters)
        (void)hipFree(ctx->d_counters);
    if (ctx->d_slots)
        (void)hipFree(ctx->d_slots);
    if (ctx->ev0)
        (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1)
        (void)hipEventDestroy(ctx->ev1);
    delete ctx;
}

const char *foort_gpu_last_error(const foort_gpu_ctx *ctx)
{
    return ctx ? ctx->err.c_str() : g_open_error.c_str();
}

int foort_gpu_trace_grid(foort_gpu_ctx *ctx, uint64_t side, uint64_t first, uint64_t count,
                         const foort_geodesic_results *out, foort_stats *stats)
{
    if (!ctx)
        return fail(nullptr, FOORT_EINVAL, "NULL context");
    int rc = check_batch(ctx, side, first, count, true);
    if (rc)
        return rc;
    return trace_host(ctx, side, first, count, nullptr, out, stats);
}

int foort_gpu_trace_pixels(foort_gpu_ctx *ctx, uint64_t side, uint64_t count, const uint64_t *screen_index,
                           const foort_geodesic_results *out, foort_stats *stats)
{
    if (!ctx)
        return fail(nullptr, FOORT_EINVAL, "NULL context");
    if (count && !screen_index)
        return fail(ctx, FOORT_EINVAL, "screen_index is NULL");
    int rc = check_batch(ctx, side, 0, count, false);
    if (rc)
        return rc;
    for (uint64_t i = 0; i < 2 * count; ++i)
        if (screen_index[i] >= side)
            return fail(ctx, FOORT_EINVAL, "screen index outside the row_column_size grid");
    return trace_host(ctx, side, 0, count, screen_index, out, stats);
}

int foort_gpu_trace_device(foort_gpu_ctx *ctx, uint64_t side, uint64_t first, uint64_t count,
                           const uint64_t *screen_index, const foort_geodesic_results *out, void *stream,
                           foort_stats *stats)
{
    return trace_device(ctx, side, first, count, screen_index, out, stream, stats, false);
}

int foort_gpu_trace_device_noalloc(foort_gpu_ctx *ctx, uint64_t side, uint64_t first, uint64_t count,
                                   const uint64_t *screen_index, const foort_geodesic_results *out, void *stream,
                                   foort_stats *stats)
{
    return trace_device(ctx, side, first, count, screen_index, out, stream, stats, true);
}

int foort_gpu_autotune(foort_gpu_ctx *ctx, uint64_t side, uint64_t count, const uint64_t *screen_index,
                       uint32_t flags, foort_run_desc *chosen, foort_stats *best)
{
    if (!ctx)
        return fail(nullptr, FOORT_EINVAL, "NULL context");
    if (count == 0)
        return fail(ctx, FOORT_EINVAL, "the autotuner needs a non-empty sample");
    int rc = check_batch(ctx, side, 0, count, screen_index == nullptr);
    if (rc)
        return rc;
    if (screen_index)
        for (uint64_t i = 0; i < 2 * count; ++i)
            if (screen_index[i] >= side)
                return fail(ctx, FOORT_EINVAL, "screen index outside the row_column_size grid");
    // candidates: block sizes of the one-shot schedule; steps between refills and the refill
    // threshold of the persistent one (each batch one launch, the fastest for throughput)
    struct Cand {
        int32_t schedule, block;
        uint32_t interval, refill;
    };
    std::vector<Cand> cands;
    const bool both = (flags & FOORT_TUNE_SCHEDULE) != 0;
    if (both || ctx->desc.schedule == FOORT_SCHEDULE_ONESHOT)
        for (int32_t b : {64, 128, 256})
            cands.push_back({FOORT_SCHEDULE_ONESHOT, b, 0, 0});
    if (both || ctx->desc.schedule == FOORT_SCHEDULE_PERSISTENT)
        for (uint32_t iv : {8u, 32u, 128u})
            for (uint32_t rf : {1u, 16u, 64u})
                cands.push_back({FOORT_SCHEDULE_PERSISTENT, 0, iv, rf});
    const foort_run_desc saved = ctx->desc;
    const foort_geodesic_results none{};
    auto run = [&](const Cand &c, foort_stats &st) -> int {
        ctx->desc.schedule = c.schedule;
        ctx->desc.block_size = c.block;
        ctx->desc.refill_interval = c.interval;
        ctx->desc.refill_lanes = c.refill;
        ctx->desc.steps_per_launch = 0;
        ctx->desc.launch_us = 0;
        ctx->desc.use_graph = 0;
        apply_lane_options(ctx);
        return trace_host(ctx, side, 0, count, screen_index, &none, &st);
    };
    foort_stats st{}, top{};
    size_t pick = 0;
    double top_ms = 0;
    rc = run(cands[0], st); // warm-up (code loading, first allocations)
    for (size_t k = 0; rc == FOORT_OK && k < cands.size(); ++k) {
        double ms = 1e300;
        // short samples are timed twice, keeping the faster run
        for (int rep = 0; rep < 2 && rc == FOORT_OK; ++rep) {
            rc = run(cands[k], st);
            if (rc == FOORT_OK && st.device_ms < ms)
                ms = st.device_ms;
            if (ms > 50.0)
                break;
        }
        if (rc == FOORT_OK && (k == 0 || ms < top_ms)) {
            top_ms = ms;
            top = st;
            top.device_ms = ms;
            pick = k;
        }
    }
    ctx->desc = saved;
    if (rc != FOORT_OK) {
        apply_lane_options(ctx);
        return rc;
    }
    ctx->desc.schedule = cands[pick].schedule;
    ctx->desc.block_size = cands[pick].block;
    if (cands[pick].schedule == FOORT_SCHEDULE_PERSISTENT) {
        ctx->desc.refill_interval = cands[pick].interval;
        ctx->desc.refill_lanes = cands[pick].refill;
    }
    apply_lane_options(ctx);
    drop_graph(ctx);
    if (chosen)
        *chosen = ctx->desc;
    if (best)
        *best = top;
    return FOORT_OK;
}

int foort_gpu_queue_open(const foort_run_desc *group, uint64_t capacity, uint32_t max_requests,
                         foort_gpu_queue **out)
{
    if (!group || !out)
        return fail(nullptr, FOORT_EINVAL, "NULL argument");
    *out = nullptr;
    if (capacity < 256 || capacity > (uint64_t(1) << 32) || max_requests == 0 || max_requests > (1u << 20))
        return fail(nullptr, FOORT_EINVAL, "queue capacity must be 256 .. 2^32 pixels and 1 .. 2^20 requests");
    foort_run_desc d = *group;
    d.schedule = FOORT_SCHEDULE_PERSISTENT;
    d.steps_per_launch = 0;
    d.use_graph = 0;
    foort_gpu_ctx *ctx = nullptr;
    int rc = foort_gpu_open(&d, &ctx);
    if (rc)
        return rc;
    auto *q = new foort_gpu_queue();
    q->ctx = ctx;
    q->key = group_key(d, ctx->rc, ctx->kernels);
    uint64_t cap = 256;
    while (cap < capacity)
        cap *= 2;
    q->cap = cap;
    q->mask = cap - 1;
    q->max_req = max_requests;
    q->used.assign(max_requests, 0);
    auto align = [](size_t v) { return (v + 255) & ~size_t(255); };
    size_t off = 0, o_reqc = off;
    off = align(off + max_requests * sizeof(RunConst));
    size_t o_side = off;
    off = align(off + max_requests * 8);
    size_t o_done = off;
    off = align(off + max_requests * 16);
    size_t o_of = off;
    off = align(off + cap * 4);
    size_t o_si = off;
    off = align(off + cap * 16);
    size_t o_term = off;
    off = align(off + cap * 4);
    size_t o_quad = off;
    off = align(off + cap * 4);
    size_t o_pass = off;
    off = align(off + cap * 4);
    size_t o_close = off;
    off = align(off + cap * 8);
    hipError_t e = hipMalloc(&q->d_block, off);
    if (e != hipSuccess) {
        foort_gpu_queue_close(q);
        return fail(nullptr, FOORT_ENOMEM, std::string("hipMalloc of the queue: ") + hipGetErrorString(e));
    }
    char *b = static_cast<char *>(q->d_block);
    q->d_reqc = reinterpret_cast<RunConst *>(b + o_reqc);
    q->d_req_side = reinterpret_cast<uint64_t *>(b + o_side);
    q->d_req_done = reinterpret_cast<unsigned long long *>(b + o_done);
    q->d_req_of = reinterpret_cast<uint32_t *>(b + o_of);
    q->d_si = reinterpret_cast<uint64_t *>(b + o_si);
    q->d_term = reinterpret_cast<int32_t *>(b + o_term);
    q->d_quad = reinterpret_cast<int32_t *>(b + o_quad);
    q->d_pass = reinterpret_cast<int32_t *>(b + o_pass);
    q->d_close = reinterpret_cast<double *>(b + o_close);
    *out = q;
    return FOORT_OK;
}

void foort_gpu_queue_close(foort_gpu_queue *q)
{
    if (!q)
        return;
    if (q->d_block)
        (void)hipFree(q->d_block);
    foort_gpu_close(q->ctx);
    delete q;
}

const char *foort_gpu_queue_last_error(const foort_gpu_queue *q)
{
    return q ? q->ctx->err.c_str() : g_open_error.c_str();
}

int foort_gpu_queue_add(foort_gpu_queue *q, const foort_run_desc *desc, uint64_t side, uint32_t *slot)
{
    if (!q || !desc || !slot)
        return fail(nullptr, FOORT_EINVAL, "NULL argument");
    if (side < 2)
        return fail(q->ctx, FOORT_EINVAL, "row_column_size must be at least 2");
    RunConst R;
    KernelPair k{};
    std::string err;
    if (build_run(*desc, R, k, err) != FOORT_OK)
        return fail(q->ctx, FOORT_EINVAL, err);
    if (group_key(*desc, R, k) != q->key)
        return fail(q->ctx, FOORT_EINVAL, "request does not match the queue's group (only the metric parameters, "
                                          "the camera and the screen may differ)");
    uint32_t s = 0;
    while (s < q->max_req && q->used[s])
        ++s;
    if (s == q->max_req)
        return fail(q->ctx, FOORT_ENOMEM, "no free request slot");
    const unsigned long long zero[2] = {0, 0};
    hipStream_t st = q->ctx->stream;
    HIP_TRY(q->ctx, hipMemcpyAsync(q->d_reqc + s, &R, sizeof R, hipMemcpyHostToDevice, st));
    HIP_TRY(q->ctx, hipMemcpyAsync(q->d_req_side + s, &side, 8, hipMemcpyHostToDevice, st));
    HIP_TRY(q->ctx, hipMemcpyAsync(q->d_req_done + 2 * s, zero, 16, hipMemcpyHostToDevice, st));
    HIP_TRY(q->ctx, hipStreamSynchronize(st));
    q->used[s] = 1;
    *slot = s;
    return FOORT_OK;
}

int foort_gpu_queue_remove(foort_gpu_queue *q, uint32_t slot)
{
    if (!q || slot >= q->max_req || !q->used[slot])
        return fail(q ? q->ctx : nullptr, FOORT_EINVAL, "no such request slot");
    q->used[slot] = 0;
    return FOORT_OK;
}

int foort_gpu_queue_push(foort_gpu_queue *q, uint32_t slot, uint64_t n, const uint64_t *screen_index,
                         uint64_t *first)
{
    if (!q || (n && !screen_index) || !first)
        return fail(q ? q->ctx : nullptr, FOORT_EINVAL, "NULL argument");
    if (slot >= q->max_req || !q->used[slot])
        return fail(q->ctx, FOORT_EINVAL, "no such request slot");
    const uint64_t tail = q->unread.empty() ? q->count : q->unread.begin()->first;
    if (q->count + n - tail > q->cap)
        return fail(q->ctx, FOORT_ENOMEM, "queue full: read the results of earlier batches first");
    *first = q->count;
    if (n == 0)
        return FOORT_OK;
    std::vector<uint32_t> of(n, slot);
    hipStream_t st = q->ctx->stream;
    // the ring may wrap: at most two contiguous pieces
    uint64_t done = 0;
    while (done < n) {
        const uint64_t at = (q->count + done) & q->mask;
        const uint64_t len = std::min(n - done, q->cap - at);
        HIP_TRY(q->ctx, hipMemcpyAsync(q->d_si + 2 * at, screen_index + 2 * done, len * 16, hipMemcpyHostToDevice, st));
        HIP_TRY(q->ctx, hipMemcpyAsync(q->d_req_of + at, of.data() + done, len * 4, hipMemcpyHostToDevice, st));
        done += len;
    }
    HIP_TRY(q->ctx, hipStreamSynchronize(st));
    q->unread[q->count] = n;
    q->count += n;
    return FOORT_OK;
}

int foort_gpu_queue_run(foort_gpu_queue *q, uint32_t launch_us, foort_stats *stats)
{
    if (!q)
        return fail(nullptr, FOORT_EINVAL, "NULL queue");
    foort_gpu_ctx *ctx = q->ctx;
    ctx->desc.launch_us = launch_us;
    int rc = setup_persistent(ctx);
    if (rc)
        return rc;
    const LaneSlots &S = ctx->slots;
    hipStream_t st = ctx->stream;
    if (!q->lanes_reset) {
        hipLaunchKernelGGL(reset_lanes, dim3(S.L / 256), dim3(256), 0, st, S, ctx->d_steps);
        HIP_TRY(ctx, hipGetLastError());
        q->lanes_reset = true;
    }
    hipLaunchKernelGGL(zero_counters, dim3(1), dim3(64), 0, st, ctx->d_steps, 1);
    HIP_TRY(ctx, hipGetLastError());
    hipLaunchKernelGGL(zero_counters, dim3(1), dim3(64), 0, st, ctx->d_counters + foort::kCntFuture, 3);
    HIP_TRY(ctx, hipGetLastError());
    foort::QueueArgs Q{q->d_reqc, q->d_req_side, q->d_req_of, q->d_si, q->d_req_done, q->count, q->mask};
    DevResults d{q->d_term, q->d_quad, q->d_pass, q->d_close, nullptr, nullptr, nullptr, nullptr};
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
    hipLaunchKernelGGL(ctx->kernels.queue, dim3(S.L / 256), dim3(256), 0, st, ctx->rc, Q, d, ctx->d_steps, S,
                       UINT64_MAX);
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
    hipLaunchKernelGGL(clamp_claims, dim3(1), dim3(64), 0, st, S.next, static_cast<unsigned long long>(q->count));
    HIP_TRY(ctx, hipGetLastError());
    unsigned long long c[foort::kNumCounters] = {};
    HIP_TRY(ctx, hipMemcpyAsync(c, ctx->d_counters, sizeof c, hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
    if (stats) {
        float ms = 0;
        HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        stats->geodesics = c[2];
        stats->steps = c[0];
        stats->launches = 1;
        stats->device_ms = ms;
        set_warnings(stats, c);
    }
    return FOORT_OK;
}

int foort_gpu_queue_progress(foort_gpu_queue *q, uint64_t *done_steps, uint64_t *pushed, uint64_t *finished)
{
    if (!q)
        return fail(nullptr, FOORT_EINVAL, "NULL queue");
    foort_gpu_ctx *ctx = q->ctx;
    hipStream_t st = ctx->stream;
    unsigned long long c[3] = {0, 0, 0};
    if (done_steps)
        HIP_TRY(ctx, hipMemcpyAsync(done_steps, q->d_req_done, q->max_req * 16, hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipMemcpyAsync(c, ctx->d_counters, sizeof c, hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
    if (pushed)
        *pushed = q->count;
    if (finished)
        *finished = q->lanes_reset ? c[2] : 0;
    return FOORT_OK;
}

int foort_gpu_queue_results(foort_gpu_queue *q, uint64_t first, uint64_t n, const foort_geodesic_results *out)
{
    if (!q || !out)
        return fail(q ? q->ctx : nullptr, FOORT_EINVAL, "NULL argument");
    auto it = q->unread.find(first);
    if (it == q->unread.end() || it->second != n)
        return fail(q->ctx, FOORT_EINVAL, "not a pushed batch whose results are unread");
    if (out->final_pos || out->final_vel || out->lambda || out->steps)
        return fail(q->ctx, FOORT_EINVAL, "a queue keeps termination and the three diagnostics only");
    hipStream_t st = q->ctx->stream;
    struct F {
        void *host;
        const void *dev;
        size_t size;
    } f[4] = {{out->term, q->d_term, 4}, {out->quadrant, q->d_quad, 4}, {out->equat_passes, q->d_pass, 4},
              {out->closest_radius, q->d_close, 8}};
    uint64_t done = 0;
    while (done < n) {
        const uint64_t at = (first + done) & q->mask;
        const uint64_t len = std::min(n - done, q->cap - at);
        for (const F &x : f)
            if (x.host)
                HIP_TRY(q->ctx, hipMemcpyAsync(static_cast<char *>(x.host) + done * x.size,
                                               static_cast<const char *>(x.dev) + at * x.size, len * x.size,
                                               hipMemcpyDeviceToHost, st));
        done += len;
    }
    HIP_TRY(q->ctx, hipStreamSynchronize(st));
    q->unread.erase(it);
    return FOORT_OK;
}

// GeodesicPosition: a counting pass (saves per pixel, and every other result), the kept
// points per pixel and their offsets on the host, then writing passes over chunks of pixels
// whose points fit the device buffer.
int foort_gpu_trace_positions(foort_gpu_ctx *ctx, uint64_t side, uint64_t count, const uint64_t *screen_index,
                              const foort_geodesic_results *out, foort_positions *pos, foort_stats *stats)
{
    if (!ctx)
        return fail(nullptr, FOORT_EINVAL, "NULL context");
    if (!out || !pos)
        return fail(ctx, FOORT_EINVAL, "results or positions pointer is NULL");
    pos->offsets = nullptr;
    pos->points = nullptr;
    pos->total = 0;
    if (!(ctx->desc.diag_flags & FOORT_DIAGFLAG_GEODESICPOSITION))
        return fail(ctx, FOORT_EINVAL, "GeodesicPosition is not on in the context's description");
    int rc = check_batch(ctx, side, 0, count, screen_index == nullptr);
    if (rc)
        return rc;
    if (screen_index)
        for (uint64_t i = 0; i < 2 * count; ++i)
            if (screen_index[i] >= side)
                return fail(ctx, FOORT_EINVAL, "screen index outside the row_column_size grid");
    if (count > (uint64_t(1) << 31))
        return fail(ctx, FOORT_EINVAL, "batch too large");
    // points per writing pass: 2^24 (512 MiB), or FOORT_GP_CHUNK_POINTS (for tests of the chunking)
    uint64_t kChunkPoints = uint64_t(1) << 24;
    if (const char *e = std::getenv("FOORT_GP_CHUNK_POINTS"))
        kChunkPoints = std::max<uint64_t>(1, std::strtoull(e, nullptr, 10));
    const foort_update_frequency &f = ctx->desc.gp_update;
    hipStream_t st = ctx->stream;

    // device: [si | term | quadrant | passes | steps | steps of the writing pass | closest | lambda |
    //          pos | vel | saved | seen | chunk offsets]
    auto align = [](size_t v) { return (v + 255) & ~size_t(255); };
    size_t off = 0;
    size_t o[13];
    const size_t bytes[13] = {screen_index ? count * 16 : 0, count * 4, count * 4, count * 4, count * 4, count * 4,
                              count * 8, count * 8, count * 32, count * 32, count * 8, count * 8, count * 8};
    for (int k = 0; k < 13; ++k) {
        o[k] = off;
        off = align(off + bytes[k]);
    }
    struct Dev {
        void *p = nullptr;
        ~Dev()
        {
            if (p)
                (void)hipFree(p);
        }
    } blk, pts;
    HIP_TRY(ctx, hipMalloc(&blk.p, std::max<size_t>(off, 256)));
    char *b = static_cast<char *>(blk.p);
    const uint64_t *d_si = screen_index ? reinterpret_cast<const uint64_t *>(b + o[0]) : nullptr;
    DevResults d{reinterpret_cast<int32_t *>(b + o[1]), reinterpret_cast<int32_t *>(b + o[2]),
                 reinterpret_cast<int32_t *>(b + o[3]), reinterpret_cast<double *>(b + o[6]),
                 reinterpret_cast<double *>(b + o[8]), reinterpret_cast<double *>(b + o[9]),
                 reinterpret_cast<double *>(b + o[7]), reinterpret_cast<uint32_t *>(b + o[4])};
    uint32_t *d_steps2 = reinterpret_cast<uint32_t *>(b + o[5]);
    uint64_t *d_saved = reinterpret_cast<uint64_t *>(b + o[10]);
    uint64_t *d_seen = reinterpret_cast<uint64_t *>(b + o[11]);
    uint64_t *d_off = reinterpret_cast<uint64_t *>(b + o[12]);
    if (screen_index)
        HIP_TRY(ctx, hipMemcpyAsync(b + o[0], screen_index, count * 16, hipMemcpyHostToDevice, st));

    foort::GPArgs A{f.update_n_steps, f.update_start, f.update_finish, ctx->desc.gp_output_steps,
                    d_saved, nullptr, nullptr, nullptr};
    hipLaunchKernelGGL(zero_counters, dim3(1), dim3(64), 0, st, ctx->d_counters, foort::kNumCounters);
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
    if (count)
        hipLaunchKernelGGL(ctx->kernels.gp, dim3(static_cast<unsigned>((count + 255) / 256)), dim3(256), 0, st,
                           ctx->rc, A, side, uint64_t(0), count, d_si, d, ctx->d_steps);
    HIP_TRY(ctx, hipGetLastError());
    if (count) {
        const unsigned grid = static_cast<unsigned>((count + 255) / 256);
        hipLaunchKernelGGL(count_ic_warnings, dim3(grid), dim3(256), 0, st, ctx->rc, side, uint64_t(0), count,
                           d_si, ctx->d_counters);
        HIP_TRY(ctx, hipGetLastError());
        if (d.pos && d.vel) {
            hipLaunchKernelGGL(count_nonfinite, dim3(grid), dim3(256), 0, st, d.pos, d.vel, count, ctx->d_counters);
            HIP_TRY(ctx, hipGetLastError());
        }
    }
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
    unsigned long long cnt[foort::kNumCounters] = {}; // of the counting pass
    HIP_TRY(ctx, hipMemcpyAsync(cnt, ctx->d_counters, sizeof cnt, hipMemcpyDeviceToHost, st));
    unsigned long long &steps = cnt[0];
    std::vector<uint64_t> saved(count), seen(count), offs(count);
    std::vector<uint32_t> steps1(count), steps2(count);
    HIP_TRY(ctx, hipMemcpyAsync(saved.data(), d_saved, count * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipMemcpyAsync(steps1.data(), d.steps, count * 4, hipMemcpyDeviceToHost, st));
    const struct {
        void *host;
        const void *dev;
        size_t size;
    } res[8] = {{out->term, d.term, 4},        {out->quadrant, d.quadrant, 4}, {out->equat_passes, d.passes, 4},
                {out->closest_radius, d.closest, 8}, {out->final_pos, d.pos, 32}, {out->final_vel, d.vel, 32},
                {out->lambda, d.lambda, 8},    {out->steps, d.steps, 4}};
    for (const auto &x : res)
        if (x.host)
            HIP_TRY(ctx, hipMemcpyAsync(x.host, x.dev, count * x.size, hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
    float ms = 0;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));

    // kept points per pixel (the reference's decimation) and their offsets
    const uint64_t keep = ctx->desc.gp_output_steps;
    auto kept = [keep](uint64_t S) { return (keep > 0 && keep < S) ? (S - 1) / (S / keep) + 1 : S; };
    pos->offsets = static_cast<uint64_t *>(std::malloc((count + 1) * sizeof(uint64_t)));
    if (!pos->offsets)
        return fail(ctx, FOORT_ENOMEM, "out of host memory for the position offsets");
    uint64_t total = 0, most = 0;
    for (uint64_t i = 0; i < count; ++i) {
        pos->offsets[i] = total;
        total += kept(saved[i]);
        most = std::max(most, kept(saved[i]));
    }
    pos->offsets[count] = total;
    pos->total = total;
    if (total) {
        pos->points = static_cast<double *>(std::malloc(total * 4 * sizeof(double)));
        if (!pos->points) {
            foort_positions_free(pos);
            return fail(ctx, FOORT_ENOMEM, "out of host memory for the positions");
        }
    }
    const uint64_t cap = std::min(total, std::max(kChunkPoints, most));
    if (cap)
        HIP_TRY(ctx, hipMalloc(&pts.p, cap * 32));
    uint64_t launches = 1;
    // nothing saved anywhere: no writing pass (the kernel takes a NULL points buffer for counting)
    for (uint64_t a = 0; total > 0 && a < count;) {
        uint64_t e = a, n_pts = 0;
        while (e < count && (e == a || n_pts + kept(saved[e]) <= cap)) {
            offs[e] = n_pts;
            n_pts += kept(saved[e]);
            ++e;
        }
        HIP_TRY(ctx, hipMemcpyAsync(d_off + a, offs.data() + a, (e - a) * 8, hipMemcpyHostToDevice, st));
        foort::GPArgs W{f.update_n_steps, f.update_start, f.update_finish, keep, d_saved + a, d_off + a,
                        static_cast<double *>(pts.p), d_seen + a};
        DevResults w{};
        w.steps = d_steps2 + a;
        HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
        hipLaunchKernelGGL(ctx->kernels.gp, dim3(static_cast<unsigned>((e - a + 255) / 256)), dim3(256), 0, st,
                           ctx->rc, W, side, a, e - a, d_si ? d_si + 2 * a : nullptr, w, ctx->d_steps);
        HIP_TRY(ctx, hipGetLastError());
        HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
        if (n_pts)
            HIP_TRY(ctx, hipMemcpyAsync(pos->points + 4 * pos->offsets[a], pts.p, n_pts * 32, hipMemcpyDeviceToHost,
                                        st));
        HIP_TRY(ctx, hipStreamSynchronize(st));
        float wms = 0;
        HIP_TRY(ctx, hipEventElapsedTime(&wms, ctx->ev0, ctx->ev1));
        ms += wms;
        ++launches;
        a = e;
    }
    if (total > 0) {
        HIP_TRY(ctx, hipMemcpyAsync(seen.data(), d_seen, count * 8, hipMemcpyDeviceToHost, st));
        HIP_TRY(ctx, hipMemcpyAsync(steps2.data(), d_steps2, count * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(ctx, hipStreamSynchronize(st));
    }
    for (uint64_t i = 0; total > 0 && i < count; ++i)
        if (seen[i] != saved[i] || steps1[i] != steps2[i]) {
            foort_positions_free(pos);
            return fail(ctx, FOORT_EDEVICE, "the writing pass of GeodesicPosition did not repeat the counting pass");
        }
    if (stats) {
        stats->geodesics = count;
        stats->steps = steps;
        stats->launches = count ? launches : 0;
        stats->device_ms = ms;
        set_warnings(stats, cnt);
    }
    return FOORT_OK;
}

void foort_positions_free(foort_positions *pos)
{
    if (!pos)
        return;
    std::free(pos->offsets);
    std::free(pos->points);
    pos->offsets = nullptr;
    pos->points = nullptr;
    pos->total = 0;
}

int foort_gpu_device_alloc(size_t bytes, void **ptr)
{
    if (!ptr)
        return fail(nullptr, FOORT_EINVAL, "NULL pointer");
    hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
    if (e != hipSuccess)
        return fail(nullptr, FOORT_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    return FOORT_OK;
}

int foort_gpu_device_free(void *ptr)
{
    hipError_t e = hipFree(ptr);
    return e == hipSuccess ? FOORT_OK : hip_fail(nullptr, e, "hipFree");
}

int foort_gpu_copy_to_device(void *dst, const void *src, size_t bytes)
{
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    return e == hipSuccess ? FOORT_OK : hip_fail(nullptr, e, "hipMemcpy to device");
}

int foort_gpu_copy_to_host(void *dst, const void *src, size_t bytes)
{
    hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
    return e == hipSuccess ? FOORT_OK : hip_fail(nullptr, e, "hipMemcpy to host");
}

int foort_gpu_synchronize(void)
{
    hipError_t e = hipDeviceSynchronize();
    return e == hipSuccess ? FOORT_OK : hip_fail(nullptr, e, "hipDeviceSynchronize");
}

} // extern "C"

namespace {

__global__ void check_division(const double *a, const double *b, double *q_path, double *q_div, uint64_t n)
{
    const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
    if (i < n) {
        q_path[i] = foort::fm_div(a[i], b[i]);
        q_div[i] = a[i] / b[i];
    }
}

} // namespace

extern "C" {

int foort_gpu_check_division(uint64_t n, const double *a, const double *b, double *q_path, double *q_div)
{
    if (n == 0 || n > (uint64_t(1) << 26) || !a || !b || !q_path || !q_div)
        return FOORT_EINVAL;
    const size_t bytes = n * sizeof(double);
    double *d = nullptr;
    hipError_t e = hipMalloc(&d, 4 * bytes);
    if (e != hipSuccess)
        return hip_fail(nullptr, e, "hipMalloc");
    e = hipMemcpy(d, a, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(d + n, b, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        check_division<<<static_cast<unsigned>((n + 255) / 256), 256>>>(d, d + n, d + 2 * n, d + 3 * n, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpy(q_path, d + 2 * n, bytes, hipMemcpyDeviceToHost);
    if (e == hipSuccess)
        e = hipMemcpy(q_div, d + 3 * n, bytes, hipMemcpyDeviceToHost);
    const hipError_t f = hipFree(d);
    if (e == hipSuccess)
        e = f;
    return e == hipSuccess ? FOORT_OK : hip_fail(nullptr, e, "foort_gpu_check_division");
}

int foort_gpu_camera(const foort_gpu_ctx *ctx, double vielbein[16], double row0[4])
{
    if (!ctx || !vielbein || !row0)
        return FOORT_EINVAL;
    std::memcpy(vielbein, ctx->rc.V, sizeof(double) * 16);
    std::memcpy(row0, ctx->rc.g0, sizeof(double) * 4);
    return FOORT_OK;
}

} // extern "C"
