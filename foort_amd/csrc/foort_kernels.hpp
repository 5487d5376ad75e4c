// foort_kernels.hpp -- kernel templates over a right-hand-side policy and an
// integrator.  Instantiated by foort_gpu.hip (finite differences, contraction
// off: bit-exact with the CPU oracle) and foort_gpu_an.hip (analytic and
// mixed-precision right-hand sides).
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "foort_path.hpp"

// Occupancy targets (waves per SIMD) of the two schedules; they bound the
// registers the compiler may use.  Chosen by measurement on MI355X.
#ifndef FOORT_WPE_ONESHOT
#define FOORT_WPE_ONESHOT 2
#endif
#ifndef FOORT_WPE_PERSIST
#define FOORT_WPE_PERSIST 2
#endif
// One-shot kernels of the closed-form formulations: three waves per SIMD hide the latency of
// their dependent chains better than two.  Measured on MI355X on a 1024x1024 Kerr image:
// binary64 3.3% faster at three waves (slower at four); mixed 1.5% faster at three than at
// four (where it spilled 100 B/lane to scratch memory; none at three).
#ifndef FOORT_WPE_ONESHOT_AN64
#define FOORT_WPE_ONESHOT_AN64 3
#endif
#ifndef FOORT_WPE_ONESHOT_MIXED
#define FOORT_WPE_ONESHOT_MIXED 3
#endif
// Resident kernels of the mixed formulation: three waves per SIMD (a few VGPRs spilled).
// Measured on MI355X on a 1024x1024 Kerr image: the persistent kernel 3% faster than at two
// waves, no faster at four (which spills more than the one-shot kernel does).
#ifndef FOORT_WPE_PERSIST_MIXED
#define FOORT_WPE_PERSIST_MIXED 3
#endif
// Resident kernels of the closed-form binary64 formulation: three waves per SIMD (no spills).
// Measured on MI355X on a 1024x1024 Kerr image: the persistent kernel 5.4% faster than at two.
#ifndef FOORT_WPE_PERSIST_AN64
#define FOORT_WPE_PERSIST_AN64 3
#endif
// Manko-Novikov, Rasheed-Larsen and (except its finite-difference one-shot kernel) Johannsen:
// their potentials with derivatives need more registers than two waves per SIMD leave, and
// spills to scratch memory stall every step, worst for a lone wave (the time-outs that end an
// image).  Measured on MI355X: one wave per SIMD halves the time of Manko-Novikov steps and
// takes about a sixth off Rasheed-Larsen's.  With one wave per SIMD the accumulation registers
// take the spills instead, and no kernel uses scratch memory.
#ifndef FOORT_WPE_HEAVY
#define FOORT_WPE_HEAVY 1
#endif

namespace foort {

// occupancy target of a kernel: `resident` for the persistent and queue kernels (their lanes
// keep a pixel index, launch counters and the wall clock next to the geodesic state) and the
// GeodesicPosition kernel (save counters and the decimation's slot arithmetic)
template <class RHS>
constexpr int waves_per_eu(bool resident)
{
    if (RHS::metric == FOORT_METRIC_MANKONOVIKOV || RHS::metric == FOORT_METRIC_RASHEEDLARSEN)
        return FOORT_WPE_HEAVY;
    if (RHS::metric == FOORT_METRIC_JOHANNSEN)
        return (RHS::exact && !resident) ? FOORT_WPE_ONESHOT : FOORT_WPE_HEAVY;
    if (resident)
        return RHS::split_rk4 ? FOORT_WPE_PERSIST_MIXED : (RHS::exact ? FOORT_WPE_PERSIST : FOORT_WPE_PERSIST_AN64);
    if (RHS::exact)
        return FOORT_WPE_ONESHOT;
    return RHS::split_rk4 ? FOORT_WPE_ONESHOT_MIXED : FOORT_WPE_ONESHOT_AN64;
}

struct DevResults {
    int32_t *term;
    int32_t *quadrant;
    int32_t *passes;
    double *closest;
    double *pos;
    double *vel;
    double *lambda;
    uint32_t *steps;
};

// Per-lane state of the persistent schedule, structure of arrays over L lanes,
// kept in HBM between launches.
constexpr int kSlotD = 11; // pos[4] vel[4] lambda prevtheta closest
constexpr int kSlotU = 12; // term steps hc bc tc thc tsteps epc crc quadrant passes request
struct LaneSlots {
    double *d;                 // kSlotD * L
    uint32_t *u;               // kSlotU * L
    long long *pix;            // L; -1 = idle
    unsigned long long *next;  // next pixel of the batch to hand out
    unsigned long long *done;  // pixels finished
    uint32_t L;
    uint32_t refill;           // refill idle lanes once at least this many of a wavefront are idle
    uint32_t interval;         // steps between refills (the length of the inner step loop)
    uint64_t ticks;            // wall-clock ticks after which a launch ends (0 = no time limit)
    uint64_t tail;             // leave once every pixel is claimed and at most this many are unfinished,
                               // for the four-lane tail kernel (0 = never)
    const RunConst *rc;        // device copy of the launch's run constants (refills read the camera)
};

// A ring of pixels from many requests that share one kernel and every run
// constant except the metric parameters and the camera (a render server's
// queue): request slot r has its own run constants reqc[r] (camera, screen,
// metric parameters; the rest equal to the group's) and grid side req_side[r];
// pixel k (a monotonic counter; storage index k & mask) belongs to request
// req_of[k & mask] and has screen index si[2 (k & mask)], si[2 (k & mask) + 1].
// req_done[2 r] counts the finished pixels of request r, req_done[2 r + 1] their steps.
struct QueueArgs {
    const RunConst *reqc;
    const uint64_t *req_side;
    const uint32_t *req_of;
    const uint64_t *si;
    unsigned long long *req_done;
    uint64_t count;            // pixels queued so far
    uint64_t mask;             // ring capacity - 1 (a power of two)
};

// The batch counters (one array; the kernels take its address as total_steps): [0] steps,
// [1] next pixel and [2] pixels done of the persistent schedule, then the warnings of
// ViewScreen::SetNewInitialConditions (FOORT/ViewScreen.cpp:112-116) -- [3] initial velocity
// pointing to the future, [4] pointing outward -- and [5] geodesics ending non-finite, which
// passes of their own count (the integration kernels carry no code for them).
constexpr int kCntFuture = 3, kCntOutward = 4, kCntNonfinite = 5, kNumCounters = 6;

__device__ __forceinline__ void store_result(const DevResults &out, uint64_t i, const GeoState &g)
{
    if (out.term)
        out.term[i] = g.term;
    if (out.quadrant)
        out.quadrant[i] = g.quadrant;
    if (out.passes)
        out.passes[i] = g.passes;
    if (out.closest)
        out.closest[i] = g.closest;
    if (out.lambda)
        out.lambda[i] = g.lambda;
    if (out.steps)
        out.steps[i] = g.steps;
    if (out.pos) {
        double2 *p = reinterpret_cast<double2 *>(out.pos + 4 * i);
        p[0] = make_double2(g.pos[0], g.pos[1]);
        p[1] = make_double2(g.pos[2], g.pos[3]);
    }
    if (out.vel) {
        double2 *p = reinterpret_cast<double2 *>(out.vel + 4 * i);
        p[0] = make_double2(g.vel[0], g.vel[1]);
        p[1] = make_double2(g.vel[2], g.vel[3]);
    }
}

// Unit screen point of a pixel, as the meshes compute it (FOORT/Mesh.cpp:42-49, :204-205)
__device__ __forceinline__ void pixel_unit(uint64_t side, uint64_t first, const uint64_t *si, uint64_t i,
                                           double &ux, double &uy)
{
    uint64_t row, col;
    if (si) {
        row = si[2 * i];
        col = si[2 * i + 1];
    } else {
        const uint64_t index = first + i;
        row = index / side;
        col = index - row * side;
    }
    const double den = static_cast<double>(side - 1);
    ux = static_cast<double>(row) * 1.0 / den;
    uy = static_cast<double>(col) * 1.0 / den;
}

// The termination code, read afresh after a run of steps whose lanes leave at different
// iterations.  Without it the compiler (ROCm 7.2, gfx950) may reuse, after the run, the lane
// mask of the run's own exit test `term != CONTINUE` as computed on the wavefront's last
// iteration: lanes that left on an earlier iteration are inactive there and read 0, so a
// later `if (term != CONTINUE) break` misses them and they take one more step.  The empty asm
// with a read-write register operand makes the compiler treat term as changed here, so every
// test after the run compares it again with all the run's lanes active.
__device__ __forceinline__ void term_after_run(int32_t &term)
{
    asm volatile("" : "+v"(term));
}

// The run constants for a refill's initial conditions (camera frame, screen), read from the
// context's device copy where a lane takes a new pixel.  The empty asm hides where the pointer
// points, so the compiler cannot hoist the loads out of the resident kernel's loop: the camera
// constants, held in scalar registers across the step loop (as the by-value argument's are),
// were spilled to vector-register lanes and reloaded by vector instructions inside it.
__device__ __forceinline__ const RunConst &refill_consts(const RunConst *rc)
{
    asm volatile("" : "+s"(rc));
    return *rc;
}

// The lane index at which a resident kernel parks its lane state after the step loop.  The
// empty asm makes the compiler form the slot addresses again there instead of keeping the
// 64-bit addresses it formed when the state was loaded alive across the loop: that register
// pair was one more than the occupancy target leaves and went to scratch memory.
__device__ __forceinline__ uint32_t park_index(uint32_t l)
{
    asm volatile("" : "+v"(l));
    return l;
}

// The number of lanes below this one whose bit is set in a wavefront mask (v_mbcnt: no
// per-lane mask kept across the resident kernels' loops).
__device__ __forceinline__ uint32_t lanes_below(unsigned long long mask)
{
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// A uniform value the compiler must take as new where it is used: the resident kernels' refills
// divide pixel indices by the grid side, and the division's set-up, hoisted out of the step
// loop, held registers across it that went to scratch memory.
__device__ __forceinline__ uint64_t opaque_u64(uint64_t x)
{
    asm volatile("" : "+s"(x));
    return x;
}

__device__ __forceinline__ void add_steps(unsigned long long *total, unsigned long long mine)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        mine += __shfl_xor(mine, off);
    if ((threadIdx.x & 63) == 0)
        atomicAdd(total, mine);
}

// One lane per pixel, integrated to termination: the drop-in for the
// per-thread loop of FOORT/Main.cpp:373-400.
template <class RHS, int INTEG, bool EVERY>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(waves_per_eu<RHS>(false), waves_per_eu<RHS>(false))))
trace_oneshot(const RunConst R, uint64_t side, uint64_t first, uint64_t count,
                                                     const uint64_t *__restrict__ si, DevResults out,
                                                     unsigned long long *__restrict__ total_steps, LaneSlots,
                                                     uint64_t)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    unsigned long long mine = 0;
    if (i < count) {
        double ux, uy;
        pixel_unit(side, first, si, i, ux, uy);
        GeoState g;
        geo_init<EVERY>(R, ux, uy, g);
        // runs of steps with a test for NaN between them.  The finite-difference kernels leave
        // the loop as a wavefront, the analytic ones lane by lane (each form measured faster on
        // MI355X for its kernels: 0.9% and 2.2%); both read the termination afresh after a run.
        if constexpr (RHS::exact) {
            do {
                for (uint32_t k = 0; k < kNanTestSteps && g.term == FOORT_TERM_CONTINUE; ++k)
                    geo_update<RHS, INTEG, EVERY, true>(R, g);
                term_after_run(g.term);
                geo_skip_nan<EVERY>(R, g);
            } while (__any(g.term == FOORT_TERM_CONTINUE));
        } else {
            for (;;) {
                for (uint32_t k = 0; k < kNanTestSteps && g.term == FOORT_TERM_CONTINUE; ++k)
                    geo_update<RHS, INTEG, EVERY, true>(R, g);
                term_after_run(g.term);
                geo_skip_nan<EVERY>(R, g);
                if (g.term != FOORT_TERM_CONTINUE)
                    break;
            }
        }
        store_result(out, i, g);
        mine = g.steps;
    }
    add_steps(total_steps, mine);
}

// The update-frequency counters are kept only where some update frequency is not 1 (EVERY
// false): an EVERY kernel never reads them, so they take no registers across its step loop.
template <bool EVERY>
__device__ __forceinline__ void slot_load(const LaneSlots &S, uint32_t l, GeoState &g)
{
    const uint32_t L = S.L;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        g.pos[k] = S.d[k * L + l];
        g.vel[k] = S.d[(4 + k) * L + l];
    }
    g.lambda = S.d[8 * L + l];
    g.prevtheta = S.d[9 * L + l];
    g.closest = S.d[10 * L + l];
    g.term = static_cast<int32_t>(S.u[0 * L + l]);
    g.steps = S.u[1 * L + l];
    g.tsteps = S.u[6 * L + l];
    if constexpr (!EVERY) {
        g.hc = S.u[2 * L + l];
        g.bc = S.u[3 * L + l];
        g.tc = S.u[4 * L + l];
        g.thc = S.u[5 * L + l];
        g.epc = S.u[7 * L + l];
        g.crc = S.u[8 * L + l];
    }
    g.quadrant = static_cast<int32_t>(S.u[9 * L + l]);
    g.passes = static_cast<int32_t>(S.u[10 * L + l]);
}

template <bool EVERY>
__device__ __forceinline__ void slot_save(const LaneSlots &S, uint32_t l, const GeoState &g)
{
    const uint32_t L = S.L;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        S.d[k * L + l] = g.pos[k];
        S.d[(4 + k) * L + l] = g.vel[k];
    }
    S.d[8 * L + l] = g.lambda;
    S.d[9 * L + l] = g.prevtheta;
    S.d[10 * L + l] = g.closest;
    S.u[0 * L + l] = static_cast<uint32_t>(g.term);
    S.u[1 * L + l] = g.steps;
    S.u[6 * L + l] = g.tsteps;
    if constexpr (!EVERY) {
        S.u[2 * L + l] = g.hc;
        S.u[3 * L + l] = g.bc;
        S.u[4 * L + l] = g.tc;
        S.u[5 * L + l] = g.thc;
        S.u[7 * L + l] = g.epc;
        S.u[8 * L + l] = g.crc;
    }
    S.u[9 * L + l] = static_cast<uint32_t>(g.quadrant);
    S.u[10 * L + l] = static_cast<uint32_t>(g.passes);
}

// Persistent work-stealing schedule.  L resident lanes run the one-shot
// kernel's step loop for `interval` steps at a time; between those runs, lanes
// whose geodesic terminated write their result and, once at least `refill`
// lanes of the wavefront are idle, take the next pixels of the batch, the
// lanes of a wavefront claiming them with one atomic (ballot + prefix count).
// A launch advances every busy lane by at most `budget` steps, or stops once
// `ticks` of the wall clock have passed (checked between step runs, so all
// wavefronts stop together), and parks the lane state in HBM; a wavefront
// leaves as soon as none of its lanes has work, or, with a tail threshold, once
// every pixel is claimed and at most `tail` geodesics are unfinished (trace_tail
// then finishes them, four lanes each).  The host relaunches until `done`
// reaches the batch size.  The steps counted are those of the geodesics that
// finish, summed per wavefront as they finish.
template <class RHS, int INTEG, bool EVERY>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(waves_per_eu<RHS>(true), waves_per_eu<RHS>(true))))
trace_persistent(const RunConst R, uint64_t side, uint64_t first,
                                                        uint64_t count, const uint64_t *__restrict__ si,
                                                        DevResults out, unsigned long long *__restrict__ total_steps,
                                                        LaneSlots S, uint64_t budget)
{
    const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; // grid covers exactly S.L lanes
    const uint32_t lane_id = threadIdx.x & 63;
    GeoState g;
    long long pix = S.pix[l];
    if (pix >= 0)
        slot_load<EVERY>(S, l, g);
    else
        g.term = -1; // no geodesic: the step loop below skips the lane
    bool exhausted = false;
    uint64_t it = 0;
    const uint64_t t0 = S.ticks ? wall_clock64() : 0;
    for (;;) {
        // refill: the idle lanes of the wavefront claim consecutive pixels with one atomic
        const unsigned long long need = __ballot(pix < 0 && !exhausted);
        if (need && (static_cast<uint32_t>(__popcll(need)) >= S.refill || !__any(pix >= 0))) {
            const int leader = __ffsll(static_cast<long long>(need)) - 1;
            unsigned long long base = 0;
            if (static_cast<int>(lane_id) == leader)
                base = atomicAdd(S.next, static_cast<unsigned long long>(__popcll(need)));
            base = __shfl(base, leader);
            if (pix < 0 && !exhausted) {
                const unsigned long long k = base + lanes_below(need);
                if (k < count) {
                    pix = static_cast<long long>(k);
                    double ux, uy;
                    pixel_unit(opaque_u64(side), first, si, k, ux, uy);
                    geo_init<EVERY>(refill_consts(S.rc), ux, uy, g);
                } else {
                    exhausted = true;
                }
            }
        }
        if (!__any(pix >= 0) || it >= budget ||
            (S.ticks && static_cast<uint64_t>(wall_clock64()) - t0 >= S.ticks))
            break;
        // the batch's last geodesics go to the tail kernel: once every pixel is claimed and at
        // most `tail` are unfinished, every wavefront parks its lanes
        if (S.tail && __atomic_load_n(S.next, __ATOMIC_RELAXED) >= count &&
            __atomic_load_n(S.done, __ATOMIC_RELAXED) + S.tail >= count)
            break;
        // up to `interval` steps: the step loop of the one-shot kernel
        const uint32_t n = budget - it < S.interval ? static_cast<uint32_t>(budget - it) : S.interval;
        for (uint32_t k = 0; k < n && g.term == FOORT_TERM_CONTINUE; ++k)
            geo_update<RHS, INTEG, EVERY>(R, g);
        term_after_run(g.term);
        geo_skip_nan<EVERY>(R, g);
        it += n;
        // lanes whose geodesic ended write it out and become idle
        // (their steps are summed over the wavefront here rather than kept per lane: one
        // register pair less across the step loop)
        const bool finished = pix >= 0 && g.term != FOORT_TERM_CONTINUE;
        if (finished) {
            store_result(out, static_cast<uint64_t>(pix), g);
            pix = -1;
        }
        const unsigned long long fin = __ballot(finished);
        if (fin) {
            add_steps(total_steps, finished ? g.steps : 0ull);
            if (static_cast<int>(lane_id) == __ffsll(static_cast<long long>(fin)) - 1)
                atomicAdd(S.done, static_cast<unsigned long long>(__popcll(fin)));
        }
    }
    const uint32_t lp = park_index(l);
    S.pix[lp] = pix;
    if (pix >= 0)
        slot_save<EVERY>(S, lp, g);
}

// Persistent schedule over a request queue (QueueArgs): as trace_persistent, with
// each lane carrying its pixel's request -- the camera when the geodesic starts,
// the metric parameters on every step -- and counting finished pixels per request.
// Results go to storage index k & mask of the ring.
template <class RHS, int INTEG, bool EVERY>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(waves_per_eu<RHS>(true), waves_per_eu<RHS>(true))))
trace_queue(const RunConst R, const QueueArgs Q, DevResults out, unsigned long long *__restrict__ total_steps,
            LaneSlots S, uint64_t budget)
{
    const uint32_t l = blockIdx.x * blockDim.x + threadIdx.x; // grid covers exactly S.L lanes
    const uint32_t lane_id = threadIdx.x & 63;
    GeoState g;
    MetricParams M = R.M;
    uint32_t req = 0;
    long long pix = S.pix[l];
    if (pix >= 0) {
        slot_load<EVERY>(S, l, g);
        req = S.u[11 * S.L + l];
        M = Q.reqc[req].M;
    } else {
        g.term = -1; // no geodesic: the step loop below skips the lane
    }
    bool exhausted = false;
    uint64_t it = 0;
    const uint64_t t0 = S.ticks ? wall_clock64() : 0;
    for (;;) {
        const unsigned long long need = __ballot(pix < 0 && !exhausted);
        if (need && (static_cast<uint32_t>(__popcll(need)) >= S.refill || !__any(pix >= 0))) {
            const int leader = __ffsll(static_cast<long long>(need)) - 1;
            unsigned long long base = 0;
            if (static_cast<int>(lane_id) == leader)
                base = atomicAdd(S.next, static_cast<unsigned long long>(__popcll(need)));
            base = __shfl(base, leader);
            if (pix < 0 && !exhausted) {
                const unsigned long long k = base + lanes_below(need);
                if (k < Q.count) {
                    pix = static_cast<long long>(k);
                    const uint64_t at = k & Q.mask;
                    req = Q.req_of[at];
                    const RunConst &Rq = Q.reqc[req];
                    M = Rq.M;
                    double ux, uy;
                    pixel_unit(Q.req_side[req], 0, Q.si, at, ux, uy);
                    geo_init<EVERY>(Rq, ux, uy, g);
                } else {
                    exhausted = true;
                }
            }
        }
        if (!__any(pix >= 0) || it >= budget ||
            (S.ticks && static_cast<uint64_t>(wall_clock64()) - t0 >= S.ticks))
            break;
        const uint32_t n = budget - it < S.interval ? static_cast<uint32_t>(budget - it) : S.interval;
        RunConst Rl = R;
        Rl.M = M;
        for (uint32_t k = 0; k < n && g.term == FOORT_TERM_CONTINUE; ++k)
            geo_update<RHS, INTEG, EVERY>(Rl, g);
        term_after_run(g.term);
        geo_skip_nan<EVERY>(Rl, g);
        it += n;
        const bool finished = pix >= 0 && g.term != FOORT_TERM_CONTINUE;
        if (finished) {
            store_result(out, static_cast<uint64_t>(pix) & Q.mask, g);
            atomicAdd(&Q.req_done[2 * req], 1ull);
            atomicAdd(&Q.req_done[2 * req + 1], static_cast<unsigned long long>(g.steps));
            pix = -1;
        }
        const unsigned long long fin = __ballot(finished);
        if (fin) {
            add_steps(total_steps, finished ? g.steps : 0ull);
            if (static_cast<int>(lane_id) == __ffsll(static_cast<long long>(fin)) - 1)
                atomicAdd(S.done, static_cast<unsigned long long>(__popcll(fin)));
        }
    }
    const uint32_t lp = park_index(l);
    S.pix[lp] = pix;
    if (pix >= 0) {
        slot_save<EVERY>(S, lp, g);
        S.u[11 * S.L + lp] = req;
    }
}

// GeodesicPosition (FOORT/Diagnostics.cpp:239-284).  The saved positions of a geodesic are
// the positions at which the diagnostic's DecideUpdate (FOORT/Diagnostics.cpp:113-140) is true,
// at Reset and after every step.  At termination the reference keeps save i when
// i % J == 0, J = S / keep (S saves, 0 < keep < S), and puts save S - 1 in place of the last
// kept one if it is not kept already: K = (S - 1) / J + 1 points, save i going to slot i / J
// and save S - 1 to slot K - 1.  The counting pass (points == NULL) stores S per pixel; the
// writing pass re-integrates with the same code and stores each kept save at its slot.
struct GPArgs {
    uint64_t n;                // UpdateFrequency: UpdateNSteps
    int32_t start, finish;     //                  UpdateStart, UpdateFinish
    uint64_t keep;             // OutputSteps (0 = all)
    uint64_t *saved;           // counting pass: S per pixel (written); writing pass: S (read)
    const uint64_t *offset;    // writing pass: first slot of each pixel in points
    double *points;            // writing pass: 4 doubles per kept point; NULL when counting
    uint64_t *seen;            // writing pass: saves seen (equal to S when the passes agree)
};

template <class RHS, int INTEG>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(waves_per_eu<RHS>(true), waves_per_eu<RHS>(true))))
trace_gp(const RunConst R, const GPArgs A, uint64_t side, uint64_t first, uint64_t count,
         const uint64_t *__restrict__ si, DevResults out, unsigned long long *__restrict__ total_steps)
{
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    unsigned long long mine = 0;
    if (i < count) {
        double ux, uy;
        pixel_unit(side, first, si, i, ux, uy);
        GeoState g;
        geo_init<false>(R, ux, uy, g);
        const bool write = A.points != nullptr;
        const uint64_t S = write ? A.saved[i] : 0;
        uint64_t J = 0, K = S;
        if (write && A.keep > 0 && A.keep < S) {
            J = S / A.keep;
            K = (S - 1) / J + 1;
        }
        double *const base = write ? A.points + 4 * A.offset[i] : nullptr;
        uint64_t cnt = 0, s = 0;
        for (;;) {
            // GeodesicPositionDiagnostic::UpdateData after Reset and after every Update
            bool save;
            if (A.n > 0) {
                save = ++cnt >= A.n;
                if (save)
                    cnt = 0;
            } else {
                save = (A.start && g.lambda == 0.0) || (A.finish && g.term != FOORT_TERM_CONTINUE);
            }
            if (save) {
                if (write && s < S) {
                    const uint64_t slot = J == 0 ? s : (s % J == 0 ? s / J : (s == S - 1 ? K - 1 : K));
                    if (slot < K) {
                        double2 *p = reinterpret_cast<double2 *>(base + 4 * slot);
                        p[0] = make_double2(g.pos[0], g.pos[1]);
                        p[1] = make_double2(g.pos[2], g.pos[3]);
                    }
                }
                ++s;
            }
            if (g.term != FOORT_TERM_CONTINUE)
                break;
            geo_update<RHS, INTEG, false>(R, g);
        }
        store_result(out, i, g);
        if (write)
            A.seen[i] = s;
        else
            A.saved[i] = s;
        mine = g.steps;
    }
    add_steps(total_steps, mine);
}

// ---------------------------------------------------------------------------
// Four lanes per geodesic for the end of a batch.  When few geodesics are left
// they run alone in their wavefronts, each step as long as its instruction stream.
// A quad of lanes holds one geodesic (every lane the same state, so every
// decision is the same); in the right-hand side each lane evaluates the metric at
// one of the four finite-difference points and the quad exchanges the values
// (DPP quad broadcasts).  Every arithmetic operation is the one rhs_fd performs,
// so the results are bit for bit those of one lane.
// ---------------------------------------------------------------------------
template <int J>
__device__ __forceinline__ double quad_bcast(double x)
{
    const long long b = __builtin_bit_cast(long long, x);
    int lo = static_cast<int>(b), hi = static_cast<int>(b >> 32);
    lo = __builtin_amdgcn_mov_dpp(lo, J * 0x55, 0xF, 0xF, false); // quad_perm [J, J, J, J]
    hi = __builtin_amdgcn_mov_dpp(hi, J * 0x55, 0xF, 0xF, false);
    return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// rhs_fd with lane q of the quad at (u + h, th), (u - h, th), (u, th + h), (u, th - h)
template <int METRIC>
__device__ __forceinline__ void rhs_fd_quad(const RunConst &R, const double *p, const double *v, double *out)
{
    const int q = static_cast<int>(threadIdx.x & 3);
    const double u = p[1], th = p[2], h = R.h_fd;
    const double uq = q == 0 ? u + h : (q == 1 ? u - h : u);
    const double tq = q == 2 ? th + h : (q == 3 ? th - h : th);
    double s, c, g[5];
    {
        // as rhs_fd: theta's sin and cos, those of theta +- h shifted from them
        double sh, sl, ch, cl, s1, c1;
        const int n = fm_sincos_red(th, &sh, &sl, &ch, &cl);
        fm_quadrant(n, sh, ch, &s, &c);
        fm_sincos_near(n, sh, sl, ch, cl, th, tq, &s1, &c1);
        s = q >= 2 ? s1 : s;
        c = q >= 2 ? c1 : c;
    }
    metric_dd<METRIC, double>(R.M, uq, s, c, g);
    double d1[5], d2[5], uu[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        d1[k] = (quad_bcast<0>(g[k]) - quad_bcast<1>(g[k])) * R.inv2h;
        d2[k] = (quad_bcast<2>(g[k]) - quad_bcast<3>(g[k])) * R.inv2h;
    }
    metric_uu<METRIC, double>(R.M, u, quad_bcast<0>(s), quad_bcast<0>(c), uu);
    geo_contract<double>(d1, d2, uu, v, out);
}

template <int METRIC>
struct RhsFDQuad {
    static constexpr int metric = METRIC;
    static constexpr bool exact = true;
    static constexpr bool split_rk4 = false;
    static constexpr bool stage_sc = false;
    __device__ __forceinline__ static void eval(const RunConst &R, const double *p, const double *v, double *out)
    {
        rhs_fd_quad<METRIC>(R, p, v, out);
    }
};

// The quad version of a policy, where there is one
template <class RHS>
struct QuadOf {
    using type = void;
};
template <int METRIC>
struct QuadOf<RhsFD<METRIC>> {
    using type = RhsFDQuad<METRIC>;
};

// Finish the geodesics parked in the lane slots, one quad of lanes each (grid: 4 L threads).
template <class RHSQ, int INTEG, bool EVERY>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(waves_per_eu<RHSQ>(true), waves_per_eu<RHSQ>(true))))
trace_tail(const RunConst R, uint64_t, uint64_t, uint64_t, const uint64_t *, DevResults out,
           unsigned long long *__restrict__ total_steps, LaneSlots S, uint64_t)
{
    const uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t l = t >> 2; // a quad per lane slot: the quad's lanes leave together
    if (l >= S.L)
        return;
    const long long pix = S.pix[l];
    if (pix < 0)
        return;
    GeoState g;
    slot_load<EVERY>(S, static_cast<uint32_t>(l), g);
    // the quad's lanes hold the same state and leave together (the right-hand side exchanges
    // values between them); the wavefront leaves when every quad is done
    do {
        for (uint32_t k = 0; k < kNanTestSteps && g.term == FOORT_TERM_CONTINUE; ++k)
            geo_update<RHSQ, INTEG, EVERY>(R, g);
        term_after_run(g.term);
        geo_skip_nan<EVERY>(R, g);
    } while (__any(g.term == FOORT_TERM_CONTINUE));
    if ((t & 3) == 0) {
        store_result(out, static_cast<uint64_t>(pix), g);
        S.pix[l] = -1;
        atomicAdd(total_steps, static_cast<unsigned long long>(g.steps));
        atomicAdd(S.done, 1ull);
    }
}

using GPKernelFn = void (*)(const RunConst, const GPArgs, uint64_t, uint64_t, uint64_t, const uint64_t *, DevResults,
                            unsigned long long *);

using KernelFn = void (*)(const RunConst, uint64_t, uint64_t, uint64_t, const uint64_t *, DevResults,
                          unsigned long long *, LaneSlots, uint64_t);
using QueueKernelFn = void (*)(const RunConst, const QueueArgs, DevResults, unsigned long long *, LaneSlots,
                               uint64_t);

struct KernelPair {
    KernelFn oneshot;
    KernelFn persistent;
    QueueKernelFn queue;
    GPKernelFn gp;             // GeodesicPosition (all update frequencies general)
    KernelFn tail;             // four lanes per geodesic (finite differences), or nullptr
};

template <class RHS, int INTEG, bool EVERY>
constexpr KernelFn tail_kernel()
{
    if constexpr (std::is_void_v<typename QuadOf<RHS>::type>)
        return nullptr;
    else
        return trace_tail<typename QuadOf<RHS>::type, INTEG, EVERY>;
}

// `every`: every termination and diagnostic updates on every step (the kernels
// without update counters)
template <class RHS, int INTEG>
KernelPair kernel_pair(bool every)
{
    return every ? KernelPair{trace_oneshot<RHS, INTEG, true>, trace_persistent<RHS, INTEG, true>,
                              trace_queue<RHS, INTEG, true>, trace_gp<RHS, INTEG>, tail_kernel<RHS, INTEG, true>()}
                 : KernelPair{trace_oneshot<RHS, INTEG, false>, trace_persistent<RHS, INTEG, false>,
                              trace_queue<RHS, INTEG, false>, trace_gp<RHS, INTEG>,
                              tail_kernel<RHS, INTEG, false>()};
}

template <template <int> class POLICY, int M>
KernelPair pick_integ(int integ, bool every)
{
    return integ == FOORT_INTEGRATOR_VERLET ? kernel_pair<POLICY<M>, FOORT_INTEGRATOR_VERLET>(every)
                                            : kernel_pair<POLICY<M>, FOORT_INTEGRATOR_RK4>(every);
}

template <template <int> class POLICY>
KernelPair pick_policy(int metric, int integ, bool every)
{
    switch (metric) {
    case FOORT_METRIC_KERR: return pick_integ<POLICY, FOORT_METRIC_KERR>(integ, every);
    case FOORT_METRIC_FLATSPACE: return pick_integ<POLICY, FOORT_METRIC_FLATSPACE>(integ, every);
    case FOORT_METRIC_RASHEEDLARSEN: return pick_integ<POLICY, FOORT_METRIC_RASHEEDLARSEN>(integ, every);
    case FOORT_METRIC_JOHANNSEN: return pick_integ<POLICY, FOORT_METRIC_JOHANNSEN>(integ, every);
    default: return pick_integ<POLICY, FOORT_METRIC_MANKONOVIKOV>(integ, every);
    }
}

// defined in foort_gpu_an.hip (analytic binary64 and mixed right-hand sides)
KernelPair pick_analytic(int metric, int integ, int precision, bool every);

} // namespace foort
