/*
 * foort_gpu.h -- C ABI of the MI355X (gfx950) per-pixel null-geodesic
 * integrator.  It drops in for the OpenMP block of the reference driver,
 * FOORT/Main.cpp:320-414: for every pixel index of one mesh iteration that
 * block calls ViewScreen::SetNewInitialConditions (FOORT/ViewScreen.cpp:75-119),
 * Geodesic::Reset (FOORT/Geodesic.cpp:33-58), Geodesic::Update until the
 * termination condition is no longer Term::Continue (FOORT/Geodesic.cpp:61-88),
 * then hands the diagnostics to Mesh::GeodesicFinished and
 * GeodesicOutputHandler::NewGeodesicOutput.  One foort_gpu_trace_* call does
 * the same for a whole batch of pixels on the GPU and returns, per pixel, the
 * termination, the diagnostic values and the final state.
 *
 * Plain C types only; no exceptions cross the boundary; every entry point
 * returns FOORT_OK (0) or a negative FOORT_E* code and records a message that
 * foort_gpu_last_error() returns.  A context is used by one host thread at a
 * time; several contexts may run concurrently: the host-pointer entry points run
 * on the context's own stream.
 */
#ifndef FOORT_GPU_H
#define FOORT_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ---- */
#define FOORT_OK 0
#define FOORT_EINVAL (-1)    /* invalid descriptor or argument */
#define FOORT_EDEVICE (-2)   /* HIP runtime error */
#define FOORT_ENOMEM (-3)    /* device allocation failed */

/* ---- metrics: the Metric descendants of FOORT/Metric.h:78-189 ---- */
#define FOORT_METRIC_KERR 0          /* params {a}                          FOORT/Metric.cpp:111-197 */
#define FOORT_METRIC_FLATSPACE 1     /* params {}                           FOORT/Metric.cpp:205-235 */
#define FOORT_METRIC_RASHEEDLARSEN 2 /* params {m, a, p, q}                 FOORT/Metric.cpp:244-369 */
#define FOORT_METRIC_JOHANNSEN 3     /* params {a, alpha13, alpha22, alpha52, epsilon3}  FOORT/Metric.cpp:378-481 */
#define FOORT_METRIC_MANKONOVIKOV 4  /* params {a, alpha3}                  FOORT/Metric.cpp:490-633 */

/* ---- integrators: FOORT/Integrators.h:54-62 ---- */
#define FOORT_INTEGRATOR_RK4 0
#define FOORT_INTEGRATOR_VERLET 1

/* ---- Christoffel policy ---- */
#define FOORT_CHRISTOFFEL_FD 0       /* central differences of g_dd, FOORT/Metric.cpp:14-55 (the reference) */
#define FOORT_CHRISTOFFEL_ANALYTIC 1 /* exact derivatives of the same g_dd (forward-mode differentiation) */

/* ---- arithmetic of the per-step right-hand side ---- */
#define FOORT_PRECISION_FP64 0
#define FOORT_PRECISION_MIXED 1      /* binary32 right-hand side where r > mixed_switch_radius
                                        and theta is away from the poles; state, step size and
                                        the update stay binary64 (needs FOORT_CHRISTOFFEL_ANALYTIC) */

/* ---- work distribution ---- */
#define FOORT_SCHEDULE_ONESHOT 0     /* one lane per pixel, integrated to termination */
#define FOORT_SCHEDULE_PERSISTENT 1  /* resident lanes refill from a pixel counter as geodesics end */

/* ---- termination bitflags: TermBitflag, FOORT/Terminations.h:29-33 ---- */
#define FOORT_TERMFLAG_BOUNDARYSPHERE 0x1u
#define FOORT_TERMFLAG_TIMEOUT 0x2u
#define FOORT_TERMFLAG_HORIZON 0x4u
#define FOORT_TERMFLAG_THETASINGULARITY 0x8u

/* ---- termination conditions: enum class Term, FOORT/Terminations.h:48-67 ---- */
#define FOORT_TERM_CONTINUE 0
#define FOORT_TERM_HORIZON 1
#define FOORT_TERM_BOUNDARYSPHERE 2
#define FOORT_TERM_TIMEOUT 3
#define FOORT_TERM_THETASINGULARITY 4
#define FOORT_TERM_STEPCAP 5         /* no counterpart: the step_cap safety ceiling was reached */

/* ---- diagnostic bitflags: DiagBitflag, FOORT/Diagnostics.h:33-37 ---- */
#define FOORT_DIAGFLAG_GEODESICPOSITION 0x1u
#define FOORT_DIAGFLAG_FOURCOLORSCREEN 0x2u
#define FOORT_DIAGFLAG_EQUATORIALPASSES 0x4u
#define FOORT_DIAGFLAG_CLOSESTRADIUS 0x8u

/* UpdateFrequency, FOORT/Diagnostics.h:50-55 */
typedef struct foort_update_frequency {
    uint64_t update_n_steps;
    int32_t update_start;
    int32_t update_finish;
} foort_update_frequency;

/*
 * Everything the reference reads from its configuration file for one run
 * (Config::GetMetric ... Config::GetGeodesicIntegrator, FOORT/Config.cpp:123-1071),
 * plus execution options.  foort_desc_init() fills the reference defaults.
 */
typedef struct foort_run_desc {
    /* Metric */
    int32_t metric;              /* FOORT_METRIC_* */
    int32_t rlog_scale;          /* Metric.RLogScale: the radial coordinate is u = log r */
    double metric_params[6];     /* constructor arguments, in the order listed above */
    /* Integrator */
    int32_t integrator;          /* FOORT_INTEGRATOR_* */
    int32_t christoffel;         /* FOORT_CHRISTOFFEL_* */
    double step_size;            /* Integrator.StepSize (epsilon, default 0.03) */
    double smallest_step;        /* Integrator.SmallestPossibleStepsize (1e-12) */
    double derivative_h;         /* Integrator.DerivativeH (1e-7) */
    double verlet_tolerance;     /* Integrator.VerletVelocityTolerance (0.001) */
    /* Terminations */
    uint32_t term_flags;         /* FOORT_TERMFLAG_* */
    uint32_t precision;          /* FOORT_PRECISION_* */
    double horizon_epsilon;      /* Horizon.Epsilon_Horizon (0.01) */
    uint64_t horizon_update_n;   /* Horizon.UpdateFrequency (1) */
    double sphere_radius;        /* BoundarySphere.SphereRadius (1000) */
    uint64_t sphere_update_n;
    uint64_t timeout_max_steps;  /* TimeOut.MaxSteps (10000) */
    uint64_t timeout_update_n;
    double theta_epsilon;        /* ThetaSingularity.Epsilon (1e-5) */
    uint64_t theta_update_n;
    /* Diagnostics */
    uint32_t diag_flags;         /* FOORT_DIAGFLAG_* */
    uint32_t value_diag;         /* the UseForMesh diagnostic (one FOORT_DIAGFLAG_*) */
    double ep_threshold;         /* EquatorialPasses.Threshold (0.01) */
    foort_update_frequency ep_update;  /* (1, start, finish) */
    foort_update_frequency cr_update;  /* ClosestRadius (1, start, finish) */
    /* ViewScreen */
    double camera_position[4];   /* (t, r, theta, phi); r is the true radius even with RLogScale */
    double screen_size[2];       /* ScreenSize {x, y} (10, 10) */
    double screen_center[2];     /* ScreenCenter {x, y} (0, 0) */
    /* Execution (no counterpart in the reference) */
    uint64_t step_cap;           /* per-geodesic ceiling on integration steps; 0 = 2^31-1 */
    int32_t schedule;            /* FOORT_SCHEDULE_* */
    int32_t block_size;          /* threads per workgroup; 0 = tuned default */
    uint32_t steps_per_launch;   /* persistent schedule: at most this many steps per kernel launch;
                                    0 = no step limit */
    uint32_t lanes_per_cu;       /* persistent schedule: resident lanes per CU; 0 = default */
    uint32_t refill_lanes;       /* persistent schedule: a wavefront refills its idle lanes once at
                                    least this many are idle (1..64; 0 = default) */
    uint32_t refill_interval;    /* persistent schedule: steps a lane runs between refills (0 = default) */
    double mixed_switch_radius;  /* FOORT_PRECISION_MIXED: binary32 steps only where r > this (10) */
    double mixed_pole_margin;    /* ... and where min(theta, pi - theta) > this (0.1) */
    uint32_t launch_us;          /* persistent schedule: end each kernel launch after about this many
                                    microseconds (0 = no time limit; with both limits 0 a batch is one launch) */
    uint32_t use_graph;          /* host entry points: run each batch that is one launch (one-shot, or
                                    persistent without limits) as one HIP graph -- index upload, kernels and
                                    result copies -- captured per batch into the context's executable graph,
                                    which is updated in place (hipGraphExecUpdate) and launched */
    /* GeodesicPosition (FOORT/Diagnostics.h:331-337): its points come from foort_gpu_trace_positions */
    foort_update_frequency gp_update;  /* (1, false, true) */
    uint64_t gp_output_steps;    /* OutputSteps: points kept per geodesic (0 = all) */
    uint32_t tail_geodesics;     /* persistent schedule, finite differences, a batch that is one launch:
                                    once every pixel is started and at most this many are unfinished,
                                    finish them with four lanes each (0 = an eighth of the resident
                                    lanes; UINT32_MAX = never) */
    uint32_t reserved;
} foort_run_desc;

/*
 * Per-geodesic results of one batch, structure of arrays, `count` entries
 * each (final_pos and final_vel: 4 doubles per geodesic, t r theta phi).
 * Any pointer may be NULL and is then not written.  For the *_device entry
 * points all pointers are device pointers.
 */
typedef struct foort_geodesic_results {
    int32_t *term;            /* Term at termination (FOORT_TERM_*) */
    int32_t *quadrant;        /* FourColorScreen value, FOORT/Diagnostics.cpp:159-200 */
    int32_t *equat_passes;    /* EquatorialPasses value, FOORT/Diagnostics.cpp:351-380 */
    double *closest_radius;   /* ClosestRadius value, FOORT/Diagnostics.cpp:427-446 */
    double *final_pos;        /* Geodesic::getCurrentPos at termination */
    double *final_vel;        /* Geodesic::getCurrentVel */
    double *lambda;           /* Geodesic::getCurrentLambda */
    uint32_t *steps;          /* number of Geodesic::Update calls */
} foort_geodesic_results;

typedef struct foort_stats {
    uint64_t geodesics;       /* geodesics integrated by the call */
    uint64_t steps;           /* sum of their integration steps */
    uint64_t launches;        /* kernel launches */
    double device_ms;         /* GPU time of the integration kernels */
    /* the per-pixel warnings of ViewScreen::SetNewInitialConditions (FOORT/ViewScreen.cpp:112-116),
       counted: initial velocities pointing to the future (v^t > 0) and outward (v^r > 0)
       (the request queue's runs count none) */
    uint64_t future_pointing;
    uint64_t outward_pointing;
    uint64_t nonfinite;       /* geodesics that ended with a NaN or infinite position or velocity,
                                 counted when the call returns final positions and velocities */
    uint64_t graph_launches;  /* batches run as one captured HIP graph (use_graph) */
} foort_stats;

typedef struct foort_gpu_ctx foort_gpu_ctx;

/* Fill *d with the reference defaults (FOORT/Config.cpp:123-1155, FOORT/Integrators.h:31-57). */
void foort_desc_init(foort_run_desc *d);

/* sizeof(foort_run_desc) as compiled into the library (lets a binding check its layout). */
int foort_desc_size(void);

/* Validate *desc, set up the camera frame (FOORT/ViewScreen.cpp:10-73) on the
 * current HIP device and return a context.  Replaces the object construction
 * of FOORT/Main.cpp:205-256. */
int foort_gpu_open(const foort_run_desc *desc, foort_gpu_ctx **ctx);

/* Give a context another run description (no counterpart in the reference): the
 * constants and kernels change, its device buffers, stream and lane state stay allocated,
 * so a server can move one context from request to request.  When the run constants change,
 * it first waits for the device's work (hipDeviceSynchronize), since batches in flight may
 * still read the context's device copy of them.  On failure (FOORT_EINVAL, or FOORT_EDEVICE
 * if that copy cannot be written) the context keeps its previous description. */
int foort_gpu_configure(foort_gpu_ctx *ctx, const foort_run_desc *desc);
void foort_gpu_close(foort_gpu_ctx *ctx);

/* Message of the last failure on ctx (or of the last failed open when ctx is NULL). */
const char *foort_gpu_last_error(const foort_gpu_ctx *ctx);

/* SimpleSquareMesh batch (FOORT/Mesh.cpp:31-50): pixel indices first .. first+count-1 of a
 * row_column_size x row_column_size grid, row = index / size, column = index % size.
 * Host result pointers; synchronous. */
int foort_gpu_trace_grid(foort_gpu_ctx *ctx, uint64_t row_column_size, uint64_t first,
                         uint64_t count, const foort_geodesic_results *out, foort_stats *stats);

/* Explicit pixel list, as the subdivision meshes queue them (FOORT/Mesh.cpp:199-206, :666-673):
 * screen_index holds (row, column) pairs, unit screen coordinates are index / (size - 1).
 * Host pointers; synchronous. */
int foort_gpu_trace_pixels(foort_gpu_ctx *ctx, uint64_t row_column_size, uint64_t count,
                           const uint64_t *screen_index, const foort_geodesic_results *out,
                           foort_stats *stats);

/* Device-resident variant: screen_index (or NULL for grid indices first .. first+count-1)
 * and every result pointer are device memory; work is queued on `stream` (a hipStream_t,
 * NULL for the default stream) and the call returns without waiting.  stats, if given,
 * is filled when the work has completed (the call then synchronizes the stream). */
int foort_gpu_trace_device(foort_gpu_ctx *ctx, uint64_t row_column_size, uint64_t first,
                           uint64_t count, const uint64_t *screen_index,
                           const foort_geodesic_results *out, void *stream, foort_stats *stats);

/* No counterpart in the reference: foort_gpu_trace_device for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, kernels and results; the call allocates nothing: the batch
 * counters and the persistent schedule's lane slots live in a static device array of 18 MiB
 * per GPU (held until the process ends) instead of the context's allocations.  The persistent
 * schedule runs at most 131072 resident lanes there (the occupancy's lane count is capped at
 * that; a lanes_per_cu that asks for more fails with FOORT_EINVAL).  The persistent kernels'
 * refills read the context's device copy of its run constants, allocated by foort_gpu_open.
 * Call it with the context's device current.  Every call on one GPU shares that array and leaves its contents
 * behind (each call clears what must start at zero): calls that overlap on one GPU, from
 * other threads or on other streams, race and must not be made. */
int foort_gpu_trace_device_noalloc(foort_gpu_ctx *ctx, uint64_t row_column_size, uint64_t first,
                                   uint64_t count, const uint64_t *screen_index,
                                   const foort_geodesic_results *out, void *stream, foort_stats *stats);

/* Autotuner (no counterpart in the reference): time a sample batch -- screen_index as for
 * foort_gpu_trace_pixels (host memory), or NULL for grid indices 0 .. count-1 -- under
 * candidate execution options and keep the fastest in the context: the one-shot block size
 * (64, 128, 256), the persistent schedule's steps between refills (8, 32, 128) and refill
 * threshold (1, 16, 64), each batch as one launch; with FOORT_TUNE_SCHEDULE in flags also
 * the schedule, else the context's own.  Launch limits (steps_per_launch, launch_us) are kept.
 * *chosen (if not NULL) receives the context's descriptor afterwards, *best the timing of the
 * chosen options on the sample. */
#define FOORT_TUNE_SCHEDULE 0x1u
int foort_gpu_autotune(foort_gpu_ctx *ctx, uint64_t row_column_size, uint64_t count,
                       const uint64_t *screen_index, uint32_t flags, foort_run_desc *chosen,
                       foort_stats *best);

/* Request queue (no counterpart in the reference; for the render server): one device-resident
 * ring of pixels shared by many requests whose run descriptions differ only in the metric
 * parameters, the camera and the screen (the group).  Each pixel carries its request slot;
 * a persistent kernel refills its lanes from the ring, so lanes of every request stay busy
 * together and short requests finish while long ones still run.
 *   open:     group = a descriptor of the group (its metric, integrator, precision, Christoffel
 *             method, terminations and diagnostics); capacity = ring size in pixels (rounded up
 *             to a power of two); max_requests = request slots.
 *   add:      register a request (its own descriptor, which must match the group, and its
 *             screen side); *slot receives its slot.  remove frees the slot: only once every
 *             pixel pushed for it is done.
 *   push:     append n (row, column) pairs of screen_index (host memory) for a slot; *first
 *             receives the ring position of the first.  Fails with FOORT_ENOMEM when results
 *             of earlier batches still occupy the room the batch needs.
 *   run:      one launch over the ring, for at most launch_us microseconds of wall clock
 *             (0: until every pushed pixel is done); lanes keep unfinished geodesics between
 *             runs.  stats->steps counts the steps of the geodesics finished by this run.
 *   progress: per slot, pixels done and their steps (done_steps[2*slot], [2*slot+1], for
 *             max_requests slots, since the slot was added); pixels pushed and done in all.
 *   results:  copy a pushed batch's results (term, quadrant, equat_passes, closest_radius;
 *             final state pointers must be NULL) to host memory and free its room; only after
 *             progress shows the batch done.
 * Host memory in and out; every call is synchronous on the queue's own stream.  One thread
 * at a time may use a queue. */
typedef struct foort_gpu_queue foort_gpu_queue;
int foort_gpu_queue_open(const foort_run_desc *group, uint64_t capacity, uint32_t max_requests,
                         foort_gpu_queue **out);
void foort_gpu_queue_close(foort_gpu_queue *q);
const char *foort_gpu_queue_last_error(const foort_gpu_queue *q);
int foort_gpu_queue_add(foort_gpu_queue *q, const foort_run_desc *desc, uint64_t row_column_size,
                        uint32_t *slot);
int foort_gpu_queue_remove(foort_gpu_queue *q, uint32_t slot);
int foort_gpu_queue_push(foort_gpu_queue *q, uint32_t slot, uint64_t n, const uint64_t *screen_index,
                         uint64_t *first);
int foort_gpu_queue_run(foort_gpu_queue *q, uint32_t launch_us, foort_stats *stats);
int foort_gpu_queue_progress(foort_gpu_queue *q, uint64_t *done_steps, uint64_t *pushed,
                             uint64_t *finished);
int foort_gpu_queue_results(foort_gpu_queue *q, uint64_t first, uint64_t n,
                            const foort_geodesic_results *out);

/* GeodesicPosition diagnostic (FOORT/Diagnostics.cpp:239-335): the positions a geodesic saves
 * (at Reset and after every step, as the UpdateFrequency gating decides), decimated at
 * termination to about gp_output_steps points.  For each pixel i the kept points (t, r or log r,
 * theta, phi) are points[4 * offsets[i] .. 4 * offsets[i + 1]); offsets has count + 1 entries.
 * The library allocates offsets and points (malloc); release them with foort_positions_free. */
typedef struct foort_positions {
    uint64_t *offsets;
    double *points;
    uint64_t total;              /* offsets[count]: points kept in all */
} foort_positions;

/* Integrate a batch with GeodesicPosition: screen_index as for foort_gpu_trace_pixels, or NULL
 * for grid indices 0 .. count-1.  Fills *out (host pointers, as trace_pixels) and *pos.  Each
 * geodesic is integrated twice (once to count its saved points, once to store the kept ones);
 * the two passes run the same code and must agree step for step, else FOORT_EDEVICE.  The
 * writing passes cover chunks of pixels whose points fit 2^24 points of device memory
 * (FOORT_GP_CHUNK_POINTS in the environment sets another size). */
int foort_gpu_trace_positions(foort_gpu_ctx *ctx, uint64_t row_column_size, uint64_t count,
                              const uint64_t *screen_index, const foort_geodesic_results *out,
                              foort_positions *pos, foort_stats *stats);
void foort_positions_free(foort_positions *pos);

/* Device memory helpers for callers that keep batches resident on the GPU
 * (no counterpart in the reference): allocate, free, copy, and wait for all
 * work queued on the device. */
int foort_gpu_device_alloc(size_t bytes, void **ptr);
int foort_gpu_device_free(void *ptr);
int foort_gpu_copy_to_device(void *dst, const void *src, size_t bytes);
int foort_gpu_copy_to_host(void *dst, const void *src, size_t bytes);
int foort_gpu_synchronize(void);

/* The camera frame the context uses: vielbein[4][4] (row a = frame vector e_a,
 * FOORT/ViewScreen.h:88) and row 0 of g_dd at the camera (FOORT/ViewScreen.h:86). */
int foort_gpu_camera(const foort_gpu_ctx *ctx, double vielbein[16], double metric_dd_row0[4]);

/* Self-check of the path's binary64 division on the device (no counterpart in the
 * reference; it tests this library's own arithmetic): q_path[i] = a[i] / b[i] as the
 * integration kernels divide (fm_div.hpp), q_div[i] = the compiler's a[i] / b[i].  Host
 * arrays of n <= 2^26 elements; the two results are bit-identical unless an operand or the
 * quotient lies near the ends of the exponent range. */
int foort_gpu_check_division(uint64_t n, const double *a, const double *b, double *q_path, double *q_div);

#ifdef __cplusplus
}
#endif

#endif /* FOORT_GPU_H */
