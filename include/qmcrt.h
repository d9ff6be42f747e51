/*
 * qmcrt.h — C ABI of the MI355X-native quasi-Monte-Carlo heliostat-field ray tracer.
 *
 * Plain C, no exceptions across the boundary: every entry point returns a qmc_status
 * (0 = ok, < 0 = error; the message of the last error on the calling thread is given by
 * qmc_last_error()). Pointers are plain host pointers unless a name says "device".
 *
 * Two object kinds:
 *   qmc_scene  host-side scene: configuration + receivers + grids + heliostats, loaded from the
 *              .scn / .json formats or built field by field, then processed (aim normals, world
 *              vertices, grid cell lists). Needs no GPU.
 *   qmc_ctx    one HIP device + one HIP stream + the device copy of one processed scene
 *              (sub-heliostat quads, cell lists, heliostat records, sample pools, facet offsets,
 *              flux maps). Calls on one ctx are serialised onto its stream; entry points that
 *              return host data synchronise, the others are asynchronous until qmc_sync().
 *
 * Each declaration names the reference interface it replaces (file:line in
 * JachyLikeCoding/SolarEnergy_Chier).
 */
#ifndef QMCRT_H
#define QMCRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int qmc_status;
#define QMC_OK               0
#define QMC_E_INVALID_ARG   -1   /* null pointer, bad enum, inconsistent sizes            */
#define QMC_E_RANGE         -2   /* heliostat / grid / receiver index out of range        */
#define QMC_E_HIP           -3   /* a HIP runtime call failed (message has the HIP error) */
#define QMC_E_OOM           -4   /* device or host allocation failed                      */
#define QMC_E_UNSUPPORTED   -5   /* receiver/grid/heliostat type not provided             */
#define QMC_E_IO            -6   /* file cannot be opened                                 */
#define QMC_E_PARSE         -7   /* malformed .scn / .json / task file                    */
#define QMC_E_STATE         -8   /* call order violated (e.g. trace before upload)        */

typedef struct qmc_scene qmc_scene;
typedef struct qmc_ctx qmc_ctx;

/* ------------------------------------------------------------------------------------------
 * Descriptors (plain data; the same numbers the .scn / .json formats carry)
 * ------------------------------------------------------------------------------------------ */

/* SceneConfiguration fields and defaults: Scene/SceneConfiguration.h:37,41-51 (json keys
 * Scene/SceneConfiguration.cpp:54-63). */
typedef struct qmc_config {
    float sun_dir[3];                      /* "sun_dir" (direction light travels)  */
    float dni;                             /* "dni"                                 */
    float csr;                             /* "csr"                                 */
    int32_t num_sunshape_groups;           /* "num_of_sunshape_groups"              */
    int32_t num_sunshape_lights_per_group; /* "num_per_sunshape_group"  (L)         */
    int32_t inverse_transform_sampling_groups;
    float receiver_pixel_length;           /* "receiver_pixel_length"               */
    float helio_disturb_std;               /* "helio_disturb_std" (slope error σ)   */
    float helio_pixel_length;              /* "helio_pixel_length" (micro-facet)    */
    float helio_reflected_rate;            /* "helio_reflected_rate" (ρ)            */
} qmc_config;

/* Recv block of the .scn format: Scene/SceneLoader.cpp:104-127.
 * res_x/res_y = 0 keeps the reference resolution rule (Scene/DataStructure/Receiver/
 * RectangleReceiver.cu:98-101, CylinderReceiver.cu:12-15); > 0 overrides it (extension).
 * tilt_deg tilts a rectangle receiver about its horizontal axis before the reference's yaw
 * rotation (extension; 0 keeps the reference geometry exactly). */
typedef struct qmc_receiver_desc {
    int32_t type;          /* 0 rectangle, 1 cylinder */
    float pos[3];
    float size[3];         /* rectangle: width, height, depth; cylinder: radius, height, - */
    float normal[3];
    int32_t face;
    int32_t res_x, res_y;
    float tilt_deg;
} qmc_receiver_desc;

/* Grid block: Scene/SceneLoader.cpp:130-159. */
typedef struct qmc_grid_desc {
    int32_t type;              /* 0 rectangle grid */
    float pos[3];
    float size[3];
    float interval[3];
    int32_t start_heliostat;   /* first heliostat index of this grid                    */
    int32_t n_heliostats;
    int32_t heliostat_type;    /* 0 flat RectangleHelio, 1 FocusFlatRectangleHelio      */
    int32_t receiver;          /* belonging receiver index                              */
} qmc_grid_desc;

/* helio line plus the sticky gap / matrix / surface_property / subhelio modifiers:
 * Scene/SceneLoader.cpp:161-198,280-333. */
typedef struct qmc_heliostat_desc {
    float pos[3];
    float size[3];             /* width (x), thickness (y), height (z) */
    int32_t rows, cols;        /* "matrix": facet rows x cols (focus-flat) / sub-heliostats */
    float gap[2];              /* "gap" x, y */
    float surface_property[6]; /* [0] focal length (<0: slant range), [1] read method      */
    int32_t grid;              /* belonging grid index */
    int32_t sub_count_override;/* 0: type default; >0: sub-heliostat count used by the grid
                                  cell lists (reference Heliostat::setNumberOfSubHelio,
                                  Scene/DataStructure/Heliostat/Heliostat.cuh:26-28) */
    float aim_offset[2];       /* extension: aim point = receiver focus centre + offset in the
                                  receiver surface ([0] across: rectangle width / cylinder
                                  tangent, [1] up: rectangle height / cylinder axis), metres;
                                  {0, 0} keeps the reference aim exactly */
    const float *local_centers;/* read method 1: rows*cols*3 floats, else NULL */
    const float *local_normals;/* read method 1: rows*cols*3 floats, else NULL */
} qmc_heliostat_desc;

typedef struct qmc_scene_counts {
    int32_t n_receivers, n_grids, n_heliostats;
    int32_t grid_num_declared;   /* "ngrid" of the ground line */
    float ground_length, ground_width;
} qmc_scene_counts;

/* Derived receiver state after processing: Receiver.cuh:257-266, RectangleReceiver.cuh:47-49. */
typedef struct qmc_receiver_info {
    int32_t type;
    int32_t res_x, res_y;        /* columns, rows */
    float pixel_length;          /* 1/ppm (reference) or width/res_x (override) */
    float factor_area;           /* receiver pixel area used by the energy factor */
    float vertices[4][3];        /* rectangle world vertices (zero for cylinder) */
    float normal[3];             /* normalised rectangle normal */
    float focus_center[3];       /* rectangle aim point */
} qmc_receiver_info;

/* Derived heliostat state after processing: Heliostat.cuh:72-84. */
typedef struct qmc_heliostat_info {
    float pos[3], size[3];
    float normal[3];
    float aim[3];                /* receiver focus centre used for the aim normal */
    float vertices[4][3];        /* world vertices (Heliostat.cu:25-40) */
    float focal_length;          /* focus-flat: after the slant-range reset */
    int32_t type;                /* 0 flat, 1 focus-flat */
    int32_t grid;
    int32_t n_micro;             /* micro-facets F */
    int32_t sub_rows, sub_cols;  /* micro-facets per facet: rows, cols */
    int32_t sub_begin;           /* first grid-relative sub-heliostat id */
    int32_t sub_count;           /* number of sub-heliostats (blocking quads) */
} qmc_heliostat_info;

typedef struct qmc_grid_info {
    int32_t nx, ny, nz;
    int32_t n_match;             /* entries of the cell item list */
    int32_t n_sub;               /* sub-heliostats in this grid */
} qmc_grid_info;

/* Per-ray record of one traced ray (parity / diagnostics). */
typedef struct qmc_ray_record {
    int32_t status;    /* QMC_RAY_* */
    int32_t index;     /* shadowed/blocked: grid-relative sub-heliostat id; deposited: pixel */
    float energy;      /* deposited energy (0 otherwise) */
    int32_t flags;     /* QMC_RAYF_* */
} qmc_ray_record;
#define QMC_RAY_SHADOWED   0
#define QMC_RAY_BLOCKED    1
#define QMC_RAY_MISSED     2
#define QMC_RAY_DEPOSITED  3
#define QMC_RAYF_EDGE_CLAMPED 1  /* u or v was exactly 1: pixel index clamped to res-1 */

/* Algorithmic work counters of a trace (reference control flow, no on-chip reuse). */
typedef struct qmc_trace_stats {
    uint64_t rays;           /* traced rays (facets x L) */
    uint64_t unshadowed;     /* rays reaching the reflection step */
    uint64_t unblocked;      /* rays reaching the receiver test */
    uint64_t deposits;       /* rays adding energy to the flux map */
    uint64_t cells;          /* DDA cells visited (both traversals) */
    uint64_t candidates;     /* cell-list entries read */
    uint64_t tests;          /* non-own candidates tested (quad vertices read) */
    uint64_t facets;         /* micro-facets generated */
    uint64_t edge_rays;      /* rays whose pixel index was clamped */
} qmc_trace_stats;

/* Sample-pool generation parameters (replaces cuRAND XORWOW seeded with time(NULL),
 * Tool/RandomNumberGenerator/RandomGenerator.cu:10-13). */
typedef struct qmc_pool_params {
    uint64_t seed_sunshape;      /* scrambles the Sobol (theta, phi) pairs */
    uint64_t seed_perturbation;  /* counter-based Gaussian slope errors     */
    int32_t sunshape_model;      /* QMC_SUNSHAPE_*; 0 = the reference's sampler */
    int32_t reserved_;           /* 0 */
} qmc_pool_params;

/* Sun-shape samplers. REFERENCE restates samplelights::linear_interpolation
 * (Scene/SceneProcessor.cu:61-81): core angles land in [0, 4.65] mrad, circumsolar ones in
 * [9.3, 10.7] mrad, with nothing between. BUIE is the physical variant: theta drawn from Buie's
 * profile (core cos(0.326t)/cos(0.308t) up to 4.65 mrad, circumsolar e^k t^g up to 43.6 mrad,
 * k and g from the CSR as the reference computes them) weighted by sin(theta) over the sky. */
#define QMC_SUNSHAPE_REFERENCE 0
#define QMC_SUNSHAPE_BUIE      1

/* ------------------------------------------------------------------------------------------
 * Errors
 * ------------------------------------------------------------------------------------------ */
const char *qmc_last_error(void);          /* thread-local; "" when none */
const char *qmc_status_string(qmc_status s);

/* ------------------------------------------------------------------------------------------
 * Host scene (no GPU needed)
 * ------------------------------------------------------------------------------------------ */
/* SolarScene singleton / ctor: Scene/SolarScene.cpp:10-24 (here an explicit object). */
qmc_status qmc_scene_create(qmc_scene **out);
/* SolarScene::clear: Scene/SolarScene.cpp:31-46. */
qmc_status qmc_scene_destroy(qmc_scene *scene);
/* SceneConfiguration::loadConfiguration: Scene/SceneConfiguration.cpp:39-69.
 * *n_fields gets the number of the 10 fields found; missing file -> QMC_E_IO,
 * malformed json -> QMC_E_PARSE. */
qmc_status qmc_scene_load_config(qmc_scene *scene, const char *json_path, int32_t *n_fields);
qmc_status qmc_scene_set_config(qmc_scene *scene, const qmc_config *cfg);
qmc_status qmc_scene_get_config(const qmc_scene *scene, qmc_config *cfg);
/* Defaults of SceneConfiguration (Scene/SceneConfiguration.h:37,41-51). */
qmc_status qmc_config_default(qmc_config *cfg);
/* SceneLoader::SceneFileRead: Scene/SceneLoader.cpp:243-353. On a malformed scene the
 * scene is left empty and QMC_E_PARSE is returned (the reference returns false). */
qmc_status qmc_scene_load_file(qmc_scene *scene, const char *scn_path);
qmc_status qmc_scene_load_text(qmc_scene *scene, const char *scn_text);
/* Programmatic construction (the reference tests build scenes field by field,
 * e.g. Tests/DataStructureTests/RectangleGridTest.cpp:29-71). */
qmc_status qmc_scene_add_receiver(qmc_scene *scene, const qmc_receiver_desc *d);
qmc_status qmc_scene_add_grid(qmc_scene *scene, const qmc_grid_desc *d);
qmc_status qmc_scene_add_heliostat(qmc_scene *scene, const qmc_heliostat_desc *d);
/* SceneProcessor::set_grid_receiver_heliostat_content: Scene/SceneProcessor.cpp:13-39
 * (receiver init, grid cell lists, aim normals, world vertices, micro-facet layout). */
qmc_status qmc_scene_process(qmc_scene *scene);
qmc_status qmc_scene_get_counts(const qmc_scene *scene, qmc_scene_counts *out);
qmc_status qmc_scene_get_receiver_desc(const qmc_scene *scene, int32_t r, qmc_receiver_desc *out);
qmc_status qmc_scene_get_grid_desc(const qmc_scene *scene, int32_t g, qmc_grid_desc *out);
/* local_centers/local_normals of the returned desc point into the scene (valid until destroy). */
qmc_status qmc_scene_get_heliostat_desc(const qmc_scene *scene, int32_t h, qmc_heliostat_desc *out);
qmc_status qmc_scene_get_receiver_info(const qmc_scene *scene, int32_t r, qmc_receiver_info *out);
qmc_status qmc_scene_get_heliostat_info(const qmc_scene *scene, int32_t h, qmc_heliostat_info *out);
/* Grid cell lists: RectangleGrid::CGridHelioMatch, Scene/DataStructure/Grid/RectangleGrid.cu:62-117.
 * cell_start gets nx*ny*nz+1 ints, match gets n_match ints (either may be NULL). */
qmc_status qmc_scene_get_grid_info(const qmc_scene *scene, int32_t g, qmc_grid_info *out);
qmc_status qmc_scene_get_grid_cells(const qmc_scene *scene, int32_t g, int32_t *cell_start,
                                    int32_t *match);
/* Sub-heliostat vertex array of one grid, 3 vertices (9 floats) per sub-heliostat:
 * QuasiMonteCarloRayTracer::setFlatRectangleHeliostatVertexs,
 * QuasiMonteCarloRayTracing/QuasiMonteCarloRayTracer.cpp:182-209. *n gets the vertex count. */
qmc_status qmc_scene_get_subheliostat_vertices(const qmc_scene *scene, int32_t g, float *out,
                                               int32_t capacity_vertices, int32_t *n);
/* Energy factor DNI*(pl_h/pl_r)^2*rho/L of one heliostat:
 * QuasiMonteCarloRayTracing/QuasiMonteCarloRayTracer.cpp:58-60,71-73. */
qmc_status qmc_scene_get_factor(const qmc_scene *scene, int32_t h, float *factor);
/* QuasiMonteCarloRayTracer::checkValidHeliostatIndex (QuasiMonteCarloRayTracer.cpp:96-127):
 * QMC_E_RANGE with the reference's message when the heliostat/grid/receiver index is invalid. */
qmc_status qmc_scene_check_heliostat(const qmc_scene *scene, int32_t h);
/* TaskLoader::loadRayTracingHeliostatIndex: RayTracingPipeline/TaskLoader.cpp:114-204.
 * Fills ids (capacity cap) with the distinct indices in file order; *n gets their count.
 * *clean = 0 when the reference would return false (repeats or extra entries), 1 otherwise.
 * Malformed / short files -> QMC_E_PARSE, missing file -> QMC_E_IO. */
qmc_status qmc_task_load(const qmc_scene *scene, const char *path, int32_t *ids, int32_t cap,
                         int32_t *n, int32_t *clean);

/* ------------------------------------------------------------------------------------------
 * Device context
 * ------------------------------------------------------------------------------------------ */
qmc_status qmc_create(int32_t device, qmc_ctx **out);
qmc_status qmc_destroy(qmc_ctx *ctx);
/* The HIP stream (hipStream_t) every launch of this ctx goes to. */
qmc_status qmc_get_stream(qmc_ctx *ctx, void **stream);
qmc_status qmc_sync(qmc_ctx *ctx);
/* Upload a processed scene (quads, cell lists, heliostat records, zeroed flux maps). */
qmc_status qmc_upload(qmc_ctx *ctx, const qmc_scene *scene);

/* Sun-shape and slope-error sample pools of P = groups*L entries.
 * Generation replaces SceneProcessor::set_samplelights / set_perturbation,
 * Scene/SceneProcessor.cu:147-271 (Buie inverse CDF, angle2xyz, normalize(gx,1,gz)). */
qmc_status qmc_pool_generate(qmc_ctx *ctx, const qmc_pool_params *params);
/* Caller-supplied pools (P*3 floats each, xyz interleaved). */
qmc_status qmc_pool_set(qmc_ctx *ctx, const float *dirs, const float *perts, int32_t P, int32_t L);
qmc_status qmc_pool_get(qmc_ctx *ctx, float *dirs, float *perts, int32_t P);
/* Buie CDF table of the current configuration: (theta, cdf) pairs and slopes.
 * samplelights::parameters_generate, Scene/SceneProcessor.cu:23-44. */
qmc_status qmc_sunshape_table(const qmc_scene *scene, float *cdf_xy /* 2*(n+1) */,
                              float *slopes /* n */, float *abc_i /* A,B,C,I,gamma */);
/* Table of the physical Buie sampler (QMC_SUNSHAPE_BUIE) of the current configuration: bin edges
 * theta[n+1] (mrad), cumulative probability cdf[n+1] (cdf[0] = 0, cdf[n] = 1) and the unnormalised
 * sky density phi(t)*sin(t) at each bin's two ends (dens_lo[n], dens_hi[n]). Null arrays are not
 * written; *n receives the bin count. No counterpart in the reference. */
qmc_status qmc_sunshape_table_buie(const qmc_scene *scene, double *theta, double *cdf, double *dens_lo,
                                   double *dens_hi, int32_t *n);
/* Transform (theta-uniforms, phi-uniforms) and (gx, gz) Gaussians into pools on the device:
 * samplelights::linear_interpolation / map_angle2xyz / map_permuation,
 * Scene/SceneProcessor.cu:61-103. Inputs are P floats each. */
qmc_status qmc_pool_from_uniforms(qmc_ctx *ctx, const float *u_theta, const float *u_phi,
                                  const float *g_x, const float *g_z);

/* Micro-facet group offsets, F+1 per heliostat (the +1 covers the next-facet read at m=F-1,
 * QuasiMonteCarloRayTracing/RectangleReceiverRectangleGridRayTracing.cu:26-27).
 * Replaces Heliostat::generateDeviceMicrohelioGroup, Heliostat.cu:57-62 + map_float2int,
 * Tool/RandomNumberGenerator/RandomGenerator.cu:20-25: offset = int(u*P), u in (0,1]. */
qmc_status qmc_offsets_generate(qmc_ctx *ctx, uint64_t seed);
/* Per-call resampling of the group offsets, as the reference draws fresh offsets for its heliostat on
 * every rayTracing call (QuasiMonteCarloRayTracer.cpp:135-161 -> Heliostat::generateDeviceMicrohelioGroup,
 * Heliostat.cu:57-62). enable = 1: from now on every trace (qmc_trace, qmc_trace_records*,
 * qmc_trace_stats_run) first regenerates the offsets of the heliostats it traces, as
 * qmc_offsets_generate would with seed + k for the k-th such regeneration (k = 0, 1, ... counted from
 * this call); a heliostat listed twice in one call is traced twice, each time with fresh offsets.
 * qmc_offsets_get returns the offsets of the last trace. enable = 0 (the default): traces use the
 * offsets generated or set before, unchanged. */
qmc_status qmc_offsets_resample(qmc_ctx *ctx, int32_t enable, uint64_t seed);
qmc_status qmc_offsets_set(qmc_ctx *ctx, int32_t heliostat, const int32_t *offsets, int32_t n);
qmc_status qmc_offsets_get(qmc_ctx *ctx, int32_t heliostat, int32_t *offsets, int32_t n);
/* Micro-facet origins and normals of one heliostat (F*3 floats each):
 * RectangleHelio::CGetDiscreteMicroHelioOriginsAndNormals, RectangleHelio.cu:84-116;
 * FocusFlatRectangleHelio::CGetDiscreteMicroHelioOriginsAndNormals, FocusFlatRectangleHelio.cu:68-116. */
qmc_status qmc_facets_get(qmc_ctx *ctx, int32_t heliostat, float *origins, float *normals,
                          int32_t F);

/* Trace heliostats into their receivers' flux maps (accumulates; never clears):
 * QuasiMonteCarloRayTracer::rayTracing, QuasiMonteCarloRayTracer.cpp:15-91, and the launchers
 * RectangleReceiverRectangleGridRayTracing / CylinderReceiverRectangleGridRayTracing,
 * RectangleReceiverRectangleGridRayTracing.cu:45-57, CylinderReceiverRectangleGridRayTracing.cu:44-57.
 * Heliostats of several grids are traced grid by grid; a heliostat listed twice is traced twice.
 * Group offsets: the reference draws fresh ones for its heliostat on every rayTracing call
 * (QuasiMonteCarloRayTracer.cpp:135-161); here that needs qmc_offsets_resample(ctx, 1, seed). With it
 * off (the default) every trace reuses the offsets generated or set before, so tracing a heliostat
 * twice repeats the same rays. Asynchronous. */
qmc_status qmc_trace(qmc_ctx *ctx, const int32_t *heliostat_ids, int32_t n);
/* The same trace as qmc_trace, allocating no device memory. No counterpart in the reference: for callers
 * that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). The list, per-heliostat candidate lists, workgroup bands and per-wave masks live in a
 * file-scope device array of 128 MiB per device (held until the process ends); a call whose wave masks
 * do not fit in it is split into launches over fewer heliostats. Limits: n <= 65536 entries, and the
 * scene's candidate lists (576 B per heliostat) plus one band per CU must fit (QMC_E_RANGE otherwise).
 * Calls that overlap on one device (contexts on different streams) share the array and race. */
qmc_status qmc_trace_noalloc(qmc_ctx *ctx, const int32_t *heliostat_ids, int32_t n);
/* The map path (qmc_trace, qmc_trace_noalloc, and records with QMC_KERNEL_MAP_DECISIONS) decides each
 * shadow and blocking traversal from the ray's tests of its wave's candidates and a walk-free proof;
 * the traversals these leave open (none to a few per 10^4 rays on the fields measured) defer their ray
 * to an exact re-trace after the launch, through a queue. No counterpart in the reference.
 * qmc_set_defer_capacity caps the queue at `rays` entries per launch (0: sized from the launch, one
 * entry per 256 rays and at least 65536; qmc_trace_noalloc holds at most 2^20). A launch that defers
 * more than its queue holds re-traces every ray of the launch instead and outputs those it deferred:
 * the same results, slower. qmc_get_defer_count waits for the context's stream and returns how many
 * rays the last such launch deferred. */
qmc_status qmc_set_defer_capacity(qmc_ctx *ctx, int64_t rays);
qmc_status qmc_get_defer_count(qmc_ctx *ctx, int64_t *rays);
/* Which trace kernel a record trace runs (qmc_trace_records_many): AUTO = the one qmc_trace would run
 * for the same list, PER_RAY = one lane per ray (k_trace), BANDED = the persistent LDS-banded kernel
 * (k_trace_band), whatever the list's size. */
#define QMC_KERNEL_AUTO    0
#define QMC_KERNEL_PER_RAY 1
#define QMC_KERNEL_BANDED  2
/* Or-ed into the kernel choice: records of the map path's own decisions (the path qmc_trace runs).
 * Without it a shadowed or blocked record names the traversal's first hit, as the reference's
 * collision test returns it (rectGridDDA.cu:31-119), and the lookup walk that finds it always runs;
 * with it, a walk proven to reach the cells of a quad the ray hits is skipped, and the record names
 * that quad (the same status, pixel and energy; another hit quad than the first only when the ray
 * hits two). */
#define QMC_KERNEL_MAP_DECISIONS 4
/* Rates (G operations/s) of the atomic adds the tracer deposits with, measured on the ctx's device:
 * float and 64-bit device-scope adds scattered over a 65,536-entry map, float LDS adds scattered over
 * a 64 KB tile. A capability probe (about 0.1 s) to state the trace's deposit rates against; no
 * counterpart in the reference. Synchronous. */
typedef struct qmc_atomic_rates {
    double global_f32, global_f64, lds_f32;
} qmc_atomic_rates;
qmc_status qmc_probe_atomic_rates(qmc_ctx *ctx, qmc_atomic_rates *rates);
/* Check of the device's short square-root and reciprocal-square-root sequences (used by the tracer's
 * normalisations for arguments in [2^-8, 2^8]) against the general correctly rounded sequences, for
 * every float in that range: `checked` receives the number of arguments, `mismatches[0]` those whose
 * short sqrtf differs from sqrtf, `mismatches[1]` those whose short 1.0f / sqrtf differs, and
 * `mismatches[2]` those where sqrtf or 1.0f / sqrtf differs from the double-precision computation
 * rounded to float. No counterpart in the reference. Synchronous. */
qmc_status qmc_check_fast_sqrt(qmc_ctx *ctx, uint64_t *checked, uint64_t mismatches[3]);
/* Host check of the cylinder map's column lookup for a map `width` columns wide (1 to 32768): for the
 * float cosines in [-1, 1] — every one with stride 1; with a larger stride every stride-th one and the
 * 513 centred on each breakpoint, where the column changes — and both signs of the hit's z offset, the
 * column from the breakpoint table and the acos estimate (the trace's path) against the column from
 * the double arc cosine. `checked` receives the number of (cosine, sign) pairs, `mismatches` those
 * that differ, `max_acos_error` the estimate's largest error in radians. Runs on host threads, needs
 * no GPU; no counterpart in the reference. */
qmc_status qmc_check_cylinder_columns(int32_t width, int32_t stride, uint64_t *checked, uint64_t *mismatches,
                                      double *max_acos_error);

/* Traversal-culling flags the trace uses for the listed heliostats under the current pools and
 * scene: bit 0 set = the shadow traversal can meet another quad, bit 1 set = the blocking traversal
 * can; a clear bit means every ray's traversal is proven to miss and the trace skips it (results are
 * unchanged). No counterpart in the reference (its DDA always runs). Synchronous. */
qmc_status qmc_cull_flags(qmc_ctx *ctx, const int32_t *heliostat_ids, int32_t n, uint8_t *flags);
/* Same trace with one record per ray of one heliostat copied to host (F*L records). Synchronous. */
qmc_status qmc_trace_records(qmc_ctx *ctx, int32_t heliostat, qmc_ray_record *records, int64_t n);
/* One trace of n distinct heliostats (one launch per grid, as qmc_trace) with the chosen kernel
 * (QMC_KERNEL_*), writing every ray's record: heliostat ids[i]'s F*L records, indexed by ray id
 * (micro-facet * L + ray), follow those of ids[0..i-1]; n_records = sum of F*L. Flux is accumulated
 * as by qmc_trace. Synchronous. Parity checks of each kernel's own indexing against the host oracle:
 * map_raytracing, RectangleReceiverRectangleGridRayTracing.cu:6-41. */
qmc_status qmc_trace_records_many(qmc_ctx *ctx, const int32_t *heliostat_ids, int32_t n, int32_t kernel,
                                  qmc_ray_record *records, int64_t n_records);
/* Trace and count the algorithmic work (instrumented kernel; flux is accumulated too). Synchronous. */
qmc_status qmc_trace_stats_run(qmc_ctx *ctx, const int32_t *heliostat_ids, int32_t n,
                               qmc_trace_stats *stats);

/* Receiver::Cclean_image_content, Scene/DataStructure/Receiver/Receiver.cu:33-45. */
qmc_status qmc_flux_clear(qmc_ctx *ctx, int32_t receiver);
qmc_status qmc_flux_read(qmc_ctx *ctx, int32_t receiver, float *out, int64_t n);
qmc_status qmc_flux_write(qmc_ctx *ctx, int32_t receiver, const float *in, int64_t n);
qmc_status qmc_flux_device_ptr(qmc_ctx *ctx, int32_t receiver, float **device_ptr, int32_t *w,
                               int32_t *h);

/* Trimmed-mean smoother, in place on a device image of w*h floats:
 * ImageSmoother::image_smooth, Tool/Smoother/ImageSmoother.cu:68-113. */
qmc_status qmc_smooth(qmc_ctx *ctx, float *device_image, int32_t w, int32_t h, int32_t radius,
                      float trimmed_ratio);
/* Convenience: smooth a receiver's flux map in place. */
qmc_status qmc_smooth_receiver(qmc_ctx *ctx, int32_t receiver, int32_t radius, float trimmed_ratio);

/* ------------------------------------------------------------------------------------------
 * Re-configuration and batched traces (sun positions x aim strategies)
 * ------------------------------------------------------------------------------------------ */
/* Set the aim-point offsets of every heliostat (n = 2 * heliostats floats; see
 * qmc_heliostat_desc.aim_offset). The scene must be processed again afterwards. No reference
 * counterpart: the reference always aims at the receiver focus centre
 * (Scene/SceneProcessor.cpp:33-35). */
qmc_status qmc_scene_set_aim_offsets(qmc_scene *scene, const float *offsets, int32_t n);
/* Refresh the sun- and aim-dependent device state (heliostat frames and energy factors, blocking
 * quads, sun frame) of a re-processed scene with the same layout, keeping pools, offsets and flux
 * maps. QMC_E_STATE if the layout changed (use qmc_upload then). This is the plain path's
 * equivalent of re-running SceneProcessor::set_grid_receiver_heliostat_content for a new sun
 * (Scene/SceneProcessor.cpp:13-39). */
qmc_status qmc_update(qmc_ctx *ctx, const qmc_scene *scene);

/* One instance of a batched trace: a sun direction, its DNI and an aim-point set. */
typedef struct qmc_instance {
    float sun_dir[3];
    float dni;
    int32_t aim_set;   /* index into the table set by qmc_batch_set_aims */
    int32_t reserved;
} qmc_instance;

/* Launch schedule of a batched trace (chosen by the autotuner when NULL is passed). */
typedef struct qmc_schedule {
    int32_t block_threads;      /* 256, 512 or 1024 */
    int32_t workgroups_per_cu;  /* persistent workgroups per compute unit (capped at what is resident) */
    int32_t lds_flux_tile;      /* 1: accumulate each instance's map in LDS; 0: atomics into its HBM map */
    int32_t facet_tiles;        /* 1: a wave takes a 2-D tile of micro-facets; 0: a row (facet order);
                                 * 2: facet-shared, each wave computes 64 micro-facets' origins and group
                                 * offsets once and traces their rays in steps (2 to 64 rays per facet, a
                                 * power of two) */
} qmc_schedule;

/* Aim-point sets for batched traces: offsets holds n_sets * heliostats * 2 floats (per set, per
 * heliostat: across, up), turned into aim points with the scene's receiver geometry. No reference
 * counterpart (batching is an extension). */
qmc_status qmc_batch_set_aims(qmc_ctx *ctx, const qmc_scene *scene, const float *offsets, int32_t n_sets);
/* Trace the heliostats of one grid for n instances in one persistent launch; instance i writes its
 * own flux map (res_x*res_y floats) at d_maps + i*res_x*res_y (device pointer), or into a ctx buffer
 * read back with qmc_batch_read when d_maps is NULL. Maps are overwritten, not accumulated. Each
 * instance equals re-processing the scene with that sun, DNI and aim set and calling qmc_trace
 * (QuasiMonteCarloRayTracer::rayTracing per heliostat, QuasiMonteCarloRayTracer.cpp:15-91) up to
 * float summation order. Flat heliostats only. Asynchronous. */
qmc_status qmc_trace_batched(qmc_ctx *ctx, int32_t grid, const qmc_instance *instances, int32_t n,
                             const qmc_schedule *schedule, float *d_maps);
qmc_status qmc_batch_read(qmc_ctx *ctx, int32_t first, int32_t count, float *out);
/* Bytes of per-instance device data (frames, factors, quads, candidate lists) a batched trace keeps at
 * once: a sweep needing more is traced in chunks of instances, one launch each. 0 restores the default
 * (2 GiB). No reference counterpart. */
qmc_status qmc_batch_set_work_budget(qmc_ctx *ctx, uint64_t bytes);
/* Time the candidate schedules on (a prefix of) the instances and return the fastest; results are
 * cached per (device, field shape) for the process and, when QMC_TUNE_CACHE names a file, across
 * processes. */
qmc_status qmc_batch_autotune(qmc_ctx *ctx, int32_t grid, const qmc_instance *instances, int32_t n,
                              qmc_schedule *out);

/* ------------------------------------------------------------------------------------------
 * Differentiable trace (no reference counterpart: the reference renders flux maps only; the
 * per-ray model is map_raytracing, RectangleReceiverRectangleGridRayTracing.cu:6-41, with a
 * bilinear instead of a nearest-pixel deposit). Flat heliostats, rectangle receivers.
 * ------------------------------------------------------------------------------------------ */
/* Forward of every heliostat of a grid with full visibility (shadowing, blocking, receiver
 * bounds, decided exactly as qmc_trace decides them); each visible ray splats its energy
 * bilinearly at (v*res_x-0.5, u*res_y-0.5) into d_flux (device, overwritten) and sets its bit of
 * the ctx's visibility mask. The splat position and energy are computed in double. */
qmc_status qmc_adjoint_forward(qmc_ctx *ctx, int32_t grid, float *d_flux);
/* The same model with the mask of the last qmc_adjoint_forward frozen (no visibility tests), for
 * the current sun / aim parameters (after qmc_update), accumulated in double and copied to
 * flux_out (host, capacity n >= res_x*res_y doubles): finite differences with common random
 * numbers. Synchronous. */
qmc_status qmc_adjoint_forward_masked(qmc_ctx *ctx, int32_t grid, double *flux_out, int64_t n);
/* loss = sum (F - target)^2 + lambda * sum max(0, F - peak)^2 over n pixels of d_flux; dL/dF goes
 * to d_grad (device, may be NULL: kept in the ctx for qmc_adjoint_backward). Synchronous. */
qmc_status qmc_flux_loss(qmc_ctx *ctx, const float *d_flux, int32_t n, float target, float peak, float lambda,
                         float *d_grad, double *loss);
/* Adjoint of the masked model: given dL/dflux (device, NULL = the one of the last qmc_flux_loss),
 * writes d(loss)/d(aim offset) for every heliostat (2 floats each: across, up; zero outside the
 * grid) and d(loss)/d(sun_dir) (3 floats). Rays are regenerated from their sample indices and
 * recomputed; visibility is the mask of the last qmc_adjoint_forward. Synchronous. */
qmc_status qmc_adjoint_backward(qmc_ctx *ctx, const qmc_scene *scene, int32_t grid, const float *d_dL_dflux,
                                float *grad_aim, float *grad_sun);

/* ------------------------------------------------------------------------------------------
 * Output (RayTracingPipeline, RayTracingPipeline.cpp:145-175)
 * ------------------------------------------------------------------------------------------ */
typedef struct qmc_image_stats {
    float max_value;   /* running float maximum of the written values >= 1e-5 */
    float sum;         /* float sum of those values in file order */
    int32_t max_row;   /* file row / column of the last occurrence of max_value (-1 if none) */
    int32_t max_col;
} qmc_image_stats;
/* ImageSaver::saveText (Tool/ImageSaver/ImageSaver.cpp:13-57, defaults precision 2): writes an
 * height x width row-major map as text, rows bottom-up, values < 1e-5 as "0", others with `precision`
 * decimals, comma-separated. Host only; stats may be NULL. */
qmc_status qmc_image_save_text(const char *path, int32_t height, int32_t width, const float *data,
                               int32_t precision, qmc_image_stats *stats);

/* HIP-event timing on the ctx stream (slots 0..15). */
qmc_status qmc_event_record(qmc_ctx *ctx, int32_t slot);
qmc_status qmc_event_elapsed_ms(qmc_ctx *ctx, int32_t slot_begin, int32_t slot_end, float *ms);
/* Instrumentation (no counterpart in the reference): from now on every plain trace launch records
 * event slot `before` just before its trace kernel (k_trace or k_trace_band) and `after` just after
 * it, so that kernel's own duration can be read with qmc_event_elapsed_ms apart from the culling
 * kernels and folds around it. (-1, -1) turns this off. */
qmc_status qmc_kernel_events(qmc_ctx *ctx, int32_t before, int32_t after);

#ifdef __cplusplus
}
#endif
#endif /* QMCRT_H */
