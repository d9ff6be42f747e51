// qmc_ctx.h — the device context behind qmc_ctx and the host helpers shared by the .hip files.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "qmc_device.h"
#include "scene_host.h"

namespace qmc {

struct GridBuf {
    DevGrid g{};
    int4 *cells = nullptr;
    int32_t *cell_start = nullptr;
    int32_t *match = nullptr;
    float4 *quads = nullptr;
    int4 *boxes = nullptr;        // cells listing each sub-heliostat: lo, hi
    ProofFrame *frame = nullptr;  // the grid's frame for the walk-free proof
    float4 *pbox = nullptr;       // each sub-heliostat's cell box for the proof: mid, lim
    QuadGroup *groups = nullptr;  // culling scan groups of the quads
    int n_sub = 0;
    int n_groups = 0;
    int receiver = 0;
    int htype = 0;                // heliostat type of every heliostat of the grid (0 flat, 1 focus-flat)
};

// One launch's share of a trace call: list entries [begin, begin + count) of pass `pass`, all of grid
// `grid`, with at most max_slots ray slots per heliostat.
struct TraceSpan {
    int pass, grid, begin, count;
    long long max_slots;
    bool sectors;  // entries dealt from 64 azimuth sectors
};

struct RecvBuf {
    DevReceiver r{};
    float *flux = nullptr;
    float *colt = nullptr;  // cylinder: the column breakpoints r.colt points to
    double *acc = nullptr;  // a launch's deposits in double, folded into flux after it (zero between launches)
    int w = 0, h = 0;
};

}  // namespace qmc

struct qmc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[16] = {};
    int kev_before = -1, kev_after = -1;  // event slots around each plain trace kernel (-1: none)
    bool uploaded = false;
    qmc_config cfg{};
    std::vector<qmc::GridBuf> grids;
    std::vector<qmc::RecvBuf> recvs;
    std::vector<qmc::DevHelio> h_helios;
    std::vector<double> h_aim64;  // aim points in double, 3 per heliostat (differentiable model)
    qmc::DevHelio *d_helios = nullptr;
    qmc::DevFacet *d_facets = nullptr;
    int32_t *d_offsets = nullptr;
    int64_t offsets_total = 0;
    float4 *d_pool = nullptr;
    float4 *d_pert = nullptr;
    // the sun-shape pool rotated by the sun frame (what every traced ray starts from), recomputed by
    // the next plain trace after the pool or the sun changes
    float4 *d_pool_world = nullptr;
    bool pool_world_ready = false;
    int P = 0, L = 0;
    bool pool_ready = false, offsets_ready = false;
    int32_t *d_list = nullptr;
    int list_cap = 0;
    qmc::StatsAcc *d_stats = nullptr;
    qmc_ray_record *d_rec = nullptr;
    size_t rec_cap = 0;
    float *d_scratch = nullptr;
    size_t scratch_cap = 0;
    qmc::DevFrame sun{};
    float2 *d_cdf = nullptr;
    float *d_slope = nullptr;
    double *d_buie_phys = nullptr;  // physical sun-shape table: theta, cdf (n+1 each), lo, hi (n each)
    qmc::BuieArgs buie{};
    // batched traces (sun x aim instances)
    float *d_aims = nullptr;       // aim points, n_aim_sets * n_heliostats * 3
    int n_aim_sets = 0;
    float *d_batch_maps = nullptr; // ctx-owned per-instance flux maps
    size_t batch_cap = 0;          // floats
    void *d_batch_inst = nullptr;  // instance table
    size_t batch_inst_cap = 0;
    char *d_batch_work = nullptr;  // per-instance frames, factors and quads of a chunk
    size_t batch_work_cap = 0;     // bytes
    size_t batch_work_budget = 0;  // bytes of per-instance data per chunk (0: the 2 GiB default)
    bool batch_maps_caller = false; // the last batched trace wrote into the caller's buffer
    int32_t batch_n = 0;           // instances of the last batched trace
    int32_t batch_npix = 0;
    int32_t batch_grid = 0;
    int num_cus = 0;
    size_t max_lds = 0, lds_per_cu = 0;
    float *d_band_partials = nullptr;  // per-workgroup bands of the persistent plain trace
    size_t band_cap = 0;
    // traversal culling: tangents of the sample pools' cone half-angles (sun shape; sun shape plus
    // twice the slope error), negative when culling does not apply, and the candidate lists per heliostat
    double cull_tan_sun = -1.0, cull_tan_block = -1.0, cull_ang_block = -1.0;
    qmc::CullList *d_cull = nullptr;
    qmc::DevCullShape *d_cull_shapes = nullptr;
    size_t cull_cap = 0;
    qmc::WideList *d_wide = nullptr;  // wide candidate lists per heliostat (plain trace)
    size_t wide_cap = 0;
    uint64_t *d_wmask = nullptr;      // per-wave candidate masks of the last plain trace launch
    size_t wmask_cap = 0;
    unsigned long long *d_defer = nullptr;  // rays the map path defers to k_trace_deferred
    size_t defer_cap = 0;
    unsigned *d_defer_n = nullptr;
    size_t defer_limit = 0;              // qmc_set_defer_capacity (0: the launch's own sizing)
    const unsigned *last_defer_n = nullptr;  // the counter of the last deferring launch
    long long *d_rec_base = nullptr;  // first record of each list entry of a record trace
    size_t rec_base_cap = 0;
    // per-call resampling of the micro-facet group offsets: when on, every pass of a trace call first
    // regenerates the offsets of the heliostats it traces with seed resample_seed + resample_next++
    bool resample = false;
    uint64_t resample_seed = 0, resample_next = 0;
    // host bookkeeping of a trace call (reserved at upload, so that a call within the reserved sizes
    // allocates no host memory either)
    std::vector<int32_t> h_pass, h_occ, h_order, h_sorted, h_tmp;
    std::vector<long long> h_sorted_base;
    // each grid's heliostats in azimuth order about its receiver's axis (grid g from h_az_begin[g]),
    // and a zeroed mark per heliostat (set and cleared within a call)
    std::vector<int32_t> h_az_order, h_az_begin;
    std::vector<uint8_t> h_mark;
    std::vector<qmc::TraceSpan> h_spans;
    // adjoint
    void *adj = nullptr;
};

namespace qmc {

#define HIP_TRY(expr)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail(QMC_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));                 \
    } while (0)

template <class T>
inline qmc_status dev_alloc(T *&p, size_t count) {
    if (p) {
        (void)hipFree(p);
        p = nullptr;
    }
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
    if (e != hipSuccess) {
        p = nullptr;
        return fail(e == hipErrorOutOfMemory ? QMC_E_OOM : QMC_E_HIP,
                    std::string("hipMalloc(") + std::to_string(count * sizeof(T)) + " B): " + hipGetErrorString(e));
    }
    return QMC_OK;
}

template <class T>
inline void dev_free(T *&p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

inline DevFrame to_dev(const Frame &f) {
    DevFrame d;
    d.ux = f.u.x; d.uy = f.u.y; d.uz = f.u.z;
    d.nx = f.n.x; d.ny = f.n.y; d.nz = f.n.z;
    d.vx = f.v.x; d.vy = f.v.y; d.vz = f.v.z;
    d.identity = f.identity;
    return d;
}

void qmc_adjoint_release(qmc_ctx *c);  // qmc_adjoint.hip
// Traversal culling (qmc_cull.hip), on the ctx stream:
// candidate lists (CullList, and WideList when `wide` is set; indexed by heliostat id) of the listed
// heliostats of a grid, into the given buffers (one entry per heliostat id; no allocation);
qmc_status launch_cull_into(qmc_ctx *c, int grid, const int32_t *d_list, int n, CullList *lists, WideList *wide,
                            unsigned *zero = nullptr);
// the same into the ctx-owned buffers, grown to the scene's heliostat count first;
qmc_status launch_cull(qmc_ctx *c, int grid, const int32_t *d_list, int n, const CullList **lists,
                       const WideList **wide);
// waves per list entry of a launch over max_slots ray slots per heliostat (one mask pair each);
inline long long wave_stride(long long max_slots) { return (max_slots + 63) / 64; }
// per-wave candidate masks of a plain trace launch over waves of max_slots slots per list entry, after
// launch_cull with wide lists (a.wide), into masks (2 * n * wave_stride(max_slots) entries; no
// allocation); sets a.wmask and a.wstride.
qmc_status launch_cull_waves_into(qmc_ctx *c, int grid, const int32_t *d_list, int n, long long max_slots,
                                  uint64_t *masks, TraceArgs &a);
// Device buffers a plain trace uses besides the scene, pools, offsets and maps: the call's heliostat
// list, per-heliostat candidate lists, per-wave candidate masks and the banded trace's per-workgroup
// bands. Either ctx-owned and grown to the call (qmc_trace), or fixed (trace_fixed: a launch whose
// masks do not fit is split into launches over fewer heliostats, and no buffer grows).
struct TraceBufs {
    int32_t *list = nullptr;  // one entry per list entry of the call
    CullList *cull = nullptr; // one per heliostat id of the scene
    WideList *wide = nullptr; // one per heliostat id of the scene
    uint64_t *wmask = nullptr;
    size_t wmask_cap = 0;     // uint64 entries
    float *band = nullptr;
    size_t band_cap = 0;      // floats
    unsigned long long *defer = nullptr;  // deferred rays of the map path (k_trace_deferred) and their count
    unsigned *defer_n = nullptr;
    size_t defer_cap = 0;     // entries (0: none; the trace kernels then walk every traversal themselves)
    bool fixed = false;
};
// A plain trace (as qmc_trace) into fixed buffers: allocates no device memory.
qmc_status trace_fixed(qmc_ctx *c, const int32_t *ids, int n, const TraceBufs &b);
// the culling scan's quad groups of grid g (after its quads and the heliostat records are known).
qmc_status upload_groups(qmc_ctx *c, int g, const std::vector<float4> &quads);

inline void free_scene(qmc_ctx *c) {
    for (auto &g : c->grids) {
        dev_free(g.cells);
        dev_free(g.cell_start);
        dev_free(g.match);
        dev_free(g.quads);
        dev_free(g.boxes);
        dev_free(g.frame);
        dev_free(g.pbox);
        dev_free(g.groups);
    }
    c->grids.clear();
    for (auto &r : c->recvs) {
        dev_free(r.flux);
        dev_free(r.acc);
        dev_free(r.colt);
    }
    c->recvs.clear();
    dev_free(c->d_helios);
    dev_free(c->d_facets);
    dev_free(c->d_offsets);
    dev_free(c->d_aims);  // aim points belong to the scene they were computed for
    c->n_aim_sets = 0;
    c->batch_n = 0;
    qmc_adjoint_release(c);  // visibility masks belong to the uploaded scene
    c->h_helios.clear();
    c->h_aim64.clear();
    c->uploaded = false;
    c->offsets_ready = false;
}

// CU count and LDS limits of the ctx's device, queried once (all three together).
inline qmc_status device_limits(qmc_ctx *c) {
    if (c->num_cus > 0 && c->max_lds > 0 && c->lds_per_cu > 0) return QMC_OK;
    int cus = 0, per_block = 0, per_cu = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    HIP_TRY(hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
    HIP_TRY(hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, c->device));
    c->num_cus = cus;
    c->max_lds = (size_t)per_block;
    c->lds_per_cu = (size_t)per_cu;
    return QMC_OK;
}

inline qmc_status set_device(qmc_ctx *c) {
    HIP_TRY(hipSetDevice(c->device));
    return QMC_OK;
}

// Sobol direction numbers of the first two dimensions (dimension 0: van der Corput; dimension 1:
// primitive polynomial x + 1 with m_1 = 1).
inline void sobol_directions(uint32_t v0[32], uint32_t v1[32]) {
    uint32_t m[33];
    for (int k = 0; k < 32; ++k) v0[k] = 1u << (31 - k);
    m[1] = 1;
    for (int k = 2; k <= 32; ++k) m[k] = (m[k - 1] << 1) ^ m[k - 1];
    for (int k = 1; k <= 32; ++k) v1[k - 1] = m[k] << (32 - k);
}

}  // namespace qmc
