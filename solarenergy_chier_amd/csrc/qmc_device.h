// qmc_device.h — device-side tables and kernel argument blocks of the gfx950 tracer.
//
// HBM layout (one qmc_ctx):
//   grid      cell_start[nx*ny*nz+1], match[n_match]          (CSR of grid-relative sub ids)
//             cells[nx*ny*nz] int4: count, first three ids      (one 16 B load per visited cell)
//             quads[3*n_sub] float4: A, E1=B-A, E2=C-A         (48 B per blocking quad)
//   helios    DevHelio[n_heliostats]                           (wave-uniform, scalar-loaded)
//   facets    DevFacet[sum rows*cols of focus-flat heliostats]
//   pools     pool[P], pert[P] float4 (x, y, z, 0)             (16 B aligned gathers)
//   offsets   int32 group starts, F+1 per heliostat, concatenated
//   flux      float res_x*res_y per receiver
#pragma once

#include <stdint.h>

#include "../../include/qmcrt.h"
#include "qmc_math.h"

namespace qmc {

struct DevFrame {
    float ux, uy, uz, nx, ny, nz, vx, vy, vz;
    int identity;
};

struct DevHelio {
    DevFrame fr;        // local2world frame of the heliostat normal (flat: origin and slope error)
    DevFrame fr_focus;  // focus-variant frame of the heliostat normal (focus-flat origins)
    float pos[3];
    float size[3];
    float gap[2];
    float len_row, len_col;  // facet extent along z (rows) and x (cols)
    float pl;           // micro-facet pixel length
    float factor;       // DNI*(pl_h/pl_r)^2*rho/L
    int32_t rows, cols, sub_rows, sub_cols;
    int32_t F;          // micro-facets
    int32_t type;       // 0 flat, 1 focus-flat
    int32_t sub_begin, sub_count;
    int32_t facet_base; // first DevFacet of a focus-flat heliostat
    int32_t grid, receiver;
    int32_t fac_rule;   // 0: factor = DNI*x*x*rho/L (x = pl_h/pl_r); 1: DNI*x*rho/L (x = pl_h^2/pixel area)
    float fac_x;
    int32_t pad_;
    int64_t offsets_base;
    // divisions by per-heliostat constants: micro-facets per row (cols*sub_cols), per-facet
    // rows/cols, and the facet-tile counts across a row for 4- and 8-wide tiles
    FastDiv fd_cpr, fd_sub_rows, fd_sub_cols, fd_ntx4, fd_ntx8;
};

struct DevFacet {
    DevFrame fr_local;  // focus-variant frame of the facet's local normal
    DevFrame fr_pert;   // local2world frame of the facet's world normal (slope-error rotation)
    float center[3];    // facet centre in the heliostat frame
    float normal[3];    // facet world normal
};

struct DevReceiver {
    int32_t type;       // 0 rectangle, 1 cylinder
    int32_t res_x, res_y;
    int32_t pad_;
    // rectangle: A = v0, E1 = v1 - v0, E2 = v3 - v0
    float A[3], E1[3], E2[3];
    // cylinder
    float pos[3];
    float radius, height;
    // cylinder: the map's column breakpoints (cylinder_column_table, 2 res_x + 4 floats; null: the
    // column from the double arc cosine)
    const float *colt;
};

struct DevGrid {
    float pos[3];
    float interval[3];
    int32_t nx, ny, nz;
    int32_t pad_;
    const int4 *cells;         // per cell: candidate count and the first three candidate ids
    const int32_t *cell_start;
    const int32_t *match;
    const float4 *quads;
    const int4 *boxes;         // per sub-heliostat id: the cells listing it, (x0, y0, z0) then (x1, y1, z1)
    const ProofFrame *frame;   // the grid's frame for the walk-free proof (dda_visit_proof), in HBM
    const float4 *pbox;        // per sub-heliostat id: its cell box as the proof reads it, mid then lim
};

// Buie sun-shape table on the device (Scene/SceneProcessor.cu:23-44,237-239)
struct BuieArgs {
    const float2 *cdf;   // n+1 (theta, cumulative)
    const float *slope;  // n
    int n;
    float A, B, C, I, gamma;
    // physical sampler (model 1): bin edges [mrad], cumulative probability, end densities
    int model, n_phys;
    const double *ph_theta, *ph_cdf, *ph_lo, *ph_hi;
};

struct StatsAcc {
    unsigned long long rays, unshadowed, unblocked, deposits, cells, candidates, tests, facets, edge;
};

// Where a heliostat's traced rays start and how far their reflections turn, for traversal culling:
// every micro-facet origin lies in the box c +- hu +- hn +- hv (world, metres), and every facet
// normal is within `cant` radians of the heliostat normal n (0 for a flat heliostat).
struct DevCullShape {
    double c[3], hu[3], hn[3], hv[3], n[3];
    double cant;
};

// Quads q in [begin, begin + count) of a grid (one heliostat's sub-heliostats, or one quad alone),
// with a ball (centre c, radius r, metres) holding each quad's own bounding ball: the culling scan
// tests the group's ball before its quads.
struct QuadGroup {
    double c[3];
    double r;
    int32_t begin, count;
};

// Candidate quads of a heliostat's two traversals (index 0 shadow, 1 blocking), from the culling scan:
// n[k] quads of the grid may meet the cone; 0 means the traversal can only miss. With n[k] <=
// kCullList the candidates are listed (in no particular order) and a ray that hits none of them
// cannot hit anything in its traversal; above that only the count is kept.
constexpr int kCullList = 6;
struct CullList {
    int32_t n[2];
    int32_t ids[2][kCullList];
};

// The same candidates with room for kCullWide ids, for the plain trace: each wave of slots then keeps,
// in a bit mask over these ids, the candidates that its own rays may meet (k_cull_waves).
constexpr int kCullWide = 64;
struct WideList {
    int32_t n[2];
    int32_t ids[2][kCullWide];
};

// Largest number of candidates a ray tests itself; a wave with more runs the traversal instead.
constexpr int kRayTests = 8;

struct CullArgs {
    const DevHelio *helios;
    const DevCullShape *shapes;
    const int32_t *list;
    const float4 *quads;  // the grid's blocking quads (A, E1, E2)
    const QuadGroup *groups;
    int32_t n_groups;
    DevFrame sun;         // frame of -sun_dir: shadow rays leave along its n axis
    double tan_sun;       // cone of shadow rays (negative: no culling)
    double ang_block;     // half-angle of the blocking cone of a flat heliostat (negative: no culling)
    double margin;
    CullList *lists;      // per heliostat id
    WideList *wide;       // per heliostat id (nullable: collect for `lists` only)
    unsigned *zero;       // nullable: a counter the launch's trace kernel starts from, set to 0 here
};

// Per-wave candidate masks of the plain trace: wave w of list entry li (slots [64w, 64w + 64) of
// heliostat list[li]) gets masks[2 * (li * wstride + w) + k], bit i set when ray origins of that
// wave and directions of traversal k may meet wide[h].ids[k][i].
struct WaveArgs {
    const DevHelio *helios;
    const DevFacet *facets;
    const int32_t *list;
    const WideList *wide;
    const float4 *quads;
    DevFrame sun;
    double tan_sun, tan_block;  // cone tangents (negative: that traversal is not culled)
    double margin;
    int32_t n_list, L, wstride;
    FastDiv fdL;
    uint64_t *masks;
};

struct TraceArgs {
    DevGrid grid;
    DevReceiver recv;
    const DevHelio *helios;
    const DevFacet *facets;
    const float4 *pool;        // sun-shape samples in world space (rotated by the sun frame)
    const float4 *pert;
    const int32_t *offsets;
    const int32_t *list;       // heliostat ids of this launch
    float *flux;
    double *acc;               // the launch's deposits (double; folded into flux after the launch)
    qmc_ray_record *records;   // nullable: ray r of list entry li goes to records[rec_base[li] + r]
    const long long *rec_base; // with records: first record of each list entry
    StatsAcc *stats;           // nullable
    const WideList *wide;      // per heliostat id: candidate quads of each traversal (null: no culling)
    const uint64_t *wmask;     // per-wave candidate masks (WaveArgs), with wide
    int32_t wstride;           // waves per list entry in wmask
    int32_t P, L;
    FastDiv fdP, fdL;
    int32_t n_list;
    int32_t blocks_per_helio;  // grid.x extent per heliostat
    // map path: rays whose traversal outcome the walk-free proof leaves open, as (list entry << 32) |
    // slot, for the deferred re-trace (k_trace_deferred); defer_n counts them (zeroed by k_cull before
    // the launch). defer_cap = 0: no queue, such rays run the full traversal in place.
    unsigned long long *defer;
    unsigned *defer_n;
    int32_t defer_cap;
};

}  // namespace qmc
