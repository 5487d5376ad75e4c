/*
 * foort_host.h -- C ABI of the host side of the drop-in: the configuration
 * file, the meshes that choose which pixels to integrate, the output files
 * and the driver loop around the GPU batch call of foort_gpu.h.
 *
 * Together with foort_gpu.h this replaces the reference program
 * (FOORT/Main.cpp, FOORT/Config.cpp, FOORT/Mesh.cpp, FOORT/InputOutput.cpp,
 * FOORT/Utilities.cpp): `foort <file.cfg>` reads the same configuration
 * files, integrates on the GPU and writes the same output files.  The
 * entries below are the pieces a caller can use on their own (a server that
 * drives a mesh itself, a test that checks one piece).
 *
 * Plain C types; no exceptions cross the boundary; entry points return
 * FOORT_OK (0) or a negative FOORT_E* code (foort_gpu.h) and record a message
 * that foort_host_last_error() returns (per thread).
 */
#ifndef FOORT_HOST_H
#define FOORT_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "foort_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failed host call on this thread. */
const char *foort_host_last_error(void);

/* ---- meshes: the Mesh descendants of FOORT/Mesh.h:68-423 ---- */
#define FOORT_MESH_SIMPLESQUARE 0        /* SimpleSquareMesh, FOORT/Mesh.cpp:25-73 */
#define FOORT_MESH_INPUTCERTAINPIXELS 1  /* InputCertainPixelsMesh, FOORT/Mesh.cpp:81-185 (pixel list given) */
#define FOORT_MESH_SQUARESUBDIVISION 2   /* SquareSubdivisionMesh, FOORT/Mesh.cpp:193-652 */
#define FOORT_MESH_SQUARESUBDIVISIONV2 3 /* SquareSubdivisionMeshV2, FOORT/Mesh.cpp:660-1187 */

/* Constructor arguments of the meshes (FOORT/Config.cpp:873-977). */
typedef struct foort_mesh_desc {
    int32_t type;                 /* FOORT_MESH_* */
    uint32_t value_diag;          /* diagnostic whose distance function weighs pixels (FOORT_DIAGFLAG_*) */
    uint64_t total_pixels;        /* SimpleSquareMesh / InputCertainPixelsMesh: TotalPixels (side^2) */
    uint64_t initial_pixels;      /* subdivision meshes: InitialPixels */
    uint64_t max_pixels;          /* MaxPixels (0 = no limit) */
    uint64_t iteration_pixels;    /* IterationPixels */
    int32_t max_subdivide;        /* MaxSubdivide (>= 1) */
    int32_t initial_subdivision_to_final; /* InitialSubdivisionToFinal */
    const uint64_t *pixels;       /* InputCertainPixelsMesh: (row, column) pairs ... */
    uint64_t pixel_count;         /* ... and their number */
} foort_mesh_desc;

typedef struct foort_mesh foort_mesh;

int foort_mesh_create(const foort_mesh_desc *desc, foort_mesh **mesh);
void foort_mesh_destroy(foort_mesh *mesh);
/* Mesh::getCurNrGeodesics, FOORT/Mesh.h:38 */
uint64_t foort_mesh_count(const foort_mesh *mesh);
/* side of the pixel grid; unit screen coordinates are index / (side - 1) */
uint64_t foort_mesh_row_column_size(const foort_mesh *mesh);
/* Mesh::getNewInitConds for indices 0 .. count-1 (FOORT/Mesh.h:41): (row, column) pairs */
int foort_mesh_screen_indices(const foort_mesh *mesh, uint64_t *screen_index);
/* Mesh::GeodesicFinished for indices 0 .. count-1 (FOORT/Mesh.h:44): the value of the
 * value diagnostic of every geodesic of the current iteration */
int foort_mesh_finished(foort_mesh *mesh, const double *values);
/* The same with two values per geodesic, (theta, phi), for a GeodesicPosition mesh value
 * (FOORT/Diagnostics.cpp:303-308). */
int foort_mesh_finished_pairs(foort_mesh *mesh, const double *pairs);
/* Mesh::EndCurrentLoop, FOORT/Mesh.h:49 */
int foort_mesh_end_loop(foort_mesh *mesh);
/* Mesh::IsFinished, FOORT/Mesh.h:52 */
int foort_mesh_is_finished(const foort_mesh *mesh);
/* Mesh::getFullDescriptionStr into buf (NUL-terminated, truncated to cap); returns the full length */
int foort_mesh_description(const foort_mesh *mesh, char *buf, size_t cap);

/* ---- configuration: Config::* of FOORT/Config.cpp:94-1155 ---- */
typedef struct foort_config foort_config;

/* Read a configuration file (path) or text (text, when path is NULL).  Syntax errors fail
 * with FOORT_EINVAL; settings that are missing or of the wrong type fall back to the
 * reference's defaults with the reference's warnings, as Config::* does. */
int foort_config_read(const char *path, const char *text, foort_config **cfg);
void foort_config_free(foort_config *cfg);
/* The run description (metric, integrator, terminations, diagnostics, camera) the
 * configuration selects, with the execution options of its optional GPU group. */
int foort_config_run_desc(const foort_config *cfg, foort_run_desc *desc);
int foort_config_mesh_desc(const foort_config *cfg, foort_mesh_desc *desc);
/* Utilities::GetFirstLineInfoString (FOORT/Utilities.cpp:52-85) for this configuration */
int foort_config_first_line(const foort_config *cfg, char *buf, size_t cap);
/* Names of the output files' diagnostics in output order (Utilities::GetDiagNameStrings,
 * FOORT/Utilities.cpp:38-50), separated by single spaces */
int foort_config_diag_names(const foort_config *cfg, char *buf, size_t cap);

/* ---- output files: GeodesicOutputHandler, FOORT/InputOutput.h:62-133 ---- */
typedef struct foort_output foort_output;

/* prefix "" writes to the console; timestamp "" omits it; per_file 0 = no limit */
int foort_output_create(const char *prefix, const char *timestamp, const char *extension,
                        const char *const *diag_names, int n_diags, uint64_t to_cache,
                        uint64_t per_file, const char *first_line, foort_output **out);
void foort_output_destroy(foort_output *out);
/* GeodesicOutputHandler::PrepareForOutput */
int foort_output_prepare(foort_output *out, uint64_t count);
/* GeodesicOutputHandler::NewGeodesicOutput for one geodesic: its screen index and one
 * NUL-terminated text per diagnostic, the list ending with a NULL pointer */
int foort_output_geodesic(foort_output *out, uint64_t index, uint64_t row, uint64_t column,
                          const char *const *texts);
/* GeodesicOutputHandler::OutputFinished */
int foort_output_finished(foort_output *out);

/* ---- the program: FOORT/Main.cpp:158-433 ---- */

/* A batch tracer integrates `count` pixels of a row_column_size grid (screen_index as
 * (row, column) pairs) and fills the host arrays of *out; returns 0 on success.  The
 * program uses the GPU (foort_gpu_trace_pixels) when none is given. */
typedef int (*foort_trace_fn)(void *user, const foort_run_desc *desc, uint64_t row_column_size,
                              uint64_t count, const uint64_t *screen_index,
                              const foort_geodesic_results *out);

/* A batch tracer for runs with GeodesicPosition: as foort_trace_fn, and also the kept positions
 * (see foort_positions in foort_gpu.h): offsets[0 .. count] always, points (4 doubles each) only
 * when offsets[count] <= capacity (the program then calls again with room enough). */
typedef int (*foort_positions_fn)(void *user, const foort_run_desc *desc, uint64_t row_column_size,
                                  uint64_t count, const uint64_t *screen_index,
                                  const foort_geodesic_results *out, uint64_t *offsets, double *points,
                                  uint64_t capacity);

typedef struct foort_program_options {
    const char *output_prefix;    /* replaces Output.FilePrefix when not NULL */
    int32_t christoffel;          /* -1: from the configuration; else FOORT_CHRISTOFFEL_* */
    int32_t precision;            /* -1 or FOORT_PRECISION_* */
    int32_t schedule;             /* -1 or FOORT_SCHEDULE_* */
    int32_t screen_output_level;  /* -1: from the configuration; else 0..5 */
    foort_trace_fn tracer;        /* NULL: the GPU */
    void *tracer_user;
    foort_stats *stats;           /* if not NULL: totals over all iterations */
    foort_positions_fn positions_tracer; /* with `tracer`, for runs with GeodesicPosition */
} foort_program_options;

void foort_program_options_init(foort_program_options *opts);
/* Run the whole program on a configuration file: returns FOORT_OK after writing the output. */
int foort_program_run(const char *cfg_path, const foort_program_options *opts);

/* ---- render server (no counterpart in the reference) ----
 * Requests -- each a run description and its mesh (a grid, an adaptive subdivision mesh or
 * a pixel list) -- run concurrently on one GPU by continuous batching.  Requests whose descriptions
 * differ only in metric parameters, camera and screen share a group: one device request
 * queue (foort_gpu_queue_*, see foort_gpu.h) whose persistent kernel integrates the pixels
 * of all of them together, in time slices of launch_us.  Each request keeps one iteration
 * batch on the device at a time; mesh work runs on `workers` host threads meanwhile.  Each
 * group's ring holds the largest power of two of pixels whose 40 bytes each (indices,
 * request slot and results) fit in hbm_budget_bytes; a batch larger than the ring fails
 * its request with FOORT_ENOMEM.  At most 1024 requests of a group are on the device at
 * once; later ones wait for a slot.  The groups' queues run concurrently only on distinct
 * hardware queues: give the process at least one per group and one more (HIP reads
 * GPU_MAX_HW_QUEUES, default 4, when it initialises); with too few, two groups can share one
 * and their launches run one after the other.  A request with GeodesicPosition (whose saved
 * points the queue kernel does not keep) runs whole on one of the worker threads, with a
 * context of its own and foort_gpu_trace_positions for each iteration, as the program does. */
typedef struct foort_request {
    foort_run_desc desc;
    foort_mesh_desc mesh;         /* any mesh; an InputCertainPixelsMesh list is copied at submission
                                     and every (row, column) must lie on the screen */
    uint64_t tag;                 /* the caller's identifier, returned with the result */
    /* NULL: no files.  Else the request writes the reference's output files, as the program
     * does for a configuration with Output.FilePrefix = output_prefix, FileExtension =
     * output_extension (NULL: none), TimeStamp = false and FirstLineInfo = true: one file per
     * diagnostic, "<prefix>_<name>[.<extension>]", the run description on its first line and
     * every mesh iteration's geodesics (FOORT/InputOutput.cpp:129-332).  Both strings are
     * copied at submission. */
    const char *output_prefix;
    const char *output_extension;
} foort_request;

typedef struct foort_request_result {
    uint64_t tag;
    int32_t status;               /* FOORT_OK or the failing call's code */
    uint32_t iterations;          /* mesh iterations run */
    uint64_t pixels, steps;       /* geodesics integrated and their steps */
    double submit_s, start_s, end_s; /* seconds since the server was created */
    double kernel_ms;             /* its share of the GPU time of the launches it took part in
                                     (by the steps of its geodesics that finished in each) */
    uint64_t footprint_bytes;     /* ring bytes of its largest batch */
} foort_request_result;

typedef struct foort_server_config {
    uint32_t workers;             /* host threads for mesh work; default 4 */
    uint32_t launch_us;           /* time slice of a queue launch; default 2000 */
    uint64_t hbm_budget_bytes;    /* ring size per group; default 4 GiB */
} foort_server_config;

typedef struct foort_server foort_server;

void foort_server_config_init(foort_server_config *cfg);
int foort_server_create(const foort_server_config *cfg, foort_server **srv);
/* queue a request (returns at once; thread-safe) */
int foort_server_submit(foort_server *srv, const foort_request *req);
/* move up to cap finished results to out, waiting up to timeout_s for the first; returns their number */
int foort_server_results(foort_server *srv, foort_request_result *out, int cap, double timeout_s);
/* wait until every submitted request has finished */
int foort_server_drain(foort_server *srv);
/* finish the queued requests and stop the workers */
void foort_server_destroy(foort_server *srv);

#ifdef __cplusplus
}
#endif

#endif /* FOORT_HOST_H */
