// qmc_device.hip — gfx950 kernels of the heliostat-field tracer and the device half of the C ABI.
//
// Kernels (each restates a reference kernel; the arithmetic order is the reference's so that the
// per-ray results are bit-identical to the host oracle):
//   k_trace<RECV, REC, STATS>   map_raytracing, QuasiMonteCarloRayTracing/
//                               RectangleReceiverRectangleGridRayTracing.cu:6-41 and
//                               CylinderReceiverRectangleGridRayTracing.cu:6-40, with the 3D-DDA of
//                               Grid/rectGridDDA.cu:31-119 and the deposits of Receiver/*.cu.
//                               One launch covers every heliostat of a task list (grid.y), micro-facets
//                               are generated in registers instead of being materialised.
//   k_trace_band<RECV, BS>      the same rays for large launches: persistent, deposits into an LDS band of
//                               map rows (k_band_fold adds the workgroups' bands).
//   k_pool_sunshape / k_pool_perturbation / k_pool_from_uniforms
//                               samplelights::linear_interpolation, map_angle2xyz, map_permuation
//                               (Scene/SceneProcessor.cu:61-103) fed by a scrambled Sobol sequence and a
//                               counter-based Gaussian instead of cuRAND XORWOW.
//   k_offsets                   map_float2int (Tool/RandomNumberGenerator/RandomGenerator.cu:20-25).
//   k_facets                    micro-facet kernels of RectangleHelio.cu:54-80 and
//                               FocusFlatRectangleHelio.cu:12-64.
// The traversal culling that the plain trace launches first is in qmc_cull.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "qmc_ctx.h"
#include "qmc_trace_core.h"

using namespace qmc;

// ---------------------------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------------------------
namespace {

// Outcome of one ray: record fields, the deposit's pixel, and what the work counters need.
struct RayOut {
    int status, index, flags, row, col;
    float energy;
    bool unshadowed, unblocked, first_of_facet;
    bool replayed;  // kTmReplay: a traversal that the deferring pass left open was decided here
};

// A ray the map path leaves to the deferred re-trace (RayOut.status only; never written to a record).
constexpr int kRayDeferred = -1;

// How trace_ray decides a traversal (plan from traversal_plan):
//   kTmExact  — the traversal's own walk: the lookup walk (dda_known) for plan 1 with at most two hits,
//               the full walk (dda_collision) otherwise; a shadowed or blocked record names the first
//               hit.
//   kTmDefer  — the map path with culling: the plan tests every masked candidate; plan 0 has no hit;
//               plan 1 is decided by the walk-free proof (known_hit_proven) where it holds; every other
//               traversal is deferred: the ray is appended to the launch's queue (while it has room)
//               and k_trace_deferred traces it again. The kernel carries no walk at all.
//   kTmReplay — k_trace_deferred: the same decisions as kTmDefer, with each traversal kTmDefer would
//               defer decided by its own walk instead (and RayOut.replayed set).
// Statuses, pixels and energies are the same in every mode.
constexpr int kTmExact = 0, kTmDefer = 1, kTmReplay = 2;

// A traversal's outcome: 0 no hit, 1 hit (hit set), 2 deferred.
template <bool STATS, int TM>
__device__ __forceinline__ int traversal_outcome(const TraceArgs &a, int plan, const RayHits &rh, V3 o, V3 d,
                                                 int own_lo, int own_hi, int li, long long slot, int &hit,
                                                 bool &replayed, RayStats &st) {
    if (plan == 0) return 0;
    if (TM != kTmExact && plan == 1 && known_hit_proven(rh, o, d, a.grid, hit)) return 1;
    if (TM == kTmDefer) {
        // ballot compaction: the wave's lanes that defer here take consecutive queue entries, reserved
        // by one atomic of the lowest such lane (its result read back by every lane), so the deferred
        // rays end up dense for the re-trace
        const uint64_t act = __ballot(1);
        const unsigned rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(act >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)act, 0u));
        unsigned base = 0;
        if (rank == 0) base = atomicAdd(a.defer_n, (unsigned)__builtin_popcountll(act));
        const unsigned k = (unsigned)__builtin_amdgcn_readlane((int)base, __builtin_ctzll(act)) + rank;
        if (k < (unsigned)a.defer_cap) a.defer[k] = ((unsigned long long)(unsigned)li << 32) | (unsigned long long)slot;
        return 2;
    }
    if (TM == kTmReplay) replayed = true;
    if (plan == 1 && rh.n <= 2) return dda_known<STATS>(rh, o, d, a.grid, own_lo, own_hi, hit, st) ? 1 : 0;
    return dda_collision<STATS>(o, d, a.grid, a.grid.quads, own_lo, own_hi, hit, st) ? 1 : 0;
}

// One ray of map_raytracing (RectangleReceiverRectangleGridRayTracing.cu:6-41,
// CylinderReceiverRectangleGridRayTracing.cu:6-40) from its traversals on, up to, not including, its
// deposit, given what its micro-facet and pool samples fix: the incident direction `dir` (the sun-shape
// sample in world space), the facet's origin, slope-error frame pf and normal, and the slope-error sample
// q. The ray is held in slot `slot` of list entry li; first: the ray is its facet's first. TM: how
// traversals are decided (kTmExact, kTmDefer, kTmReplay); a deferred ray gets status kRayDeferred (no
// deposit, no record). U: the wave's lanes share the heliostat and the wave's candidate masks (scalar
// loads); off in k_trace_deferred. ot (nullable): the origin's terms of the rectangle receiver's test.
template <int RT, bool STATS, int TM, bool U = true>
__device__ __forceinline__ void trace_from(const TraceArgs &a, const DevHelio &H, const WideList *wl,
                                           const uint64_t *wm, int li, long long slot, bool first, V3 dir,
                                           V3 origin, const DevFrame &pf, float4 q, RayOut &o, RayStats &st,
                                           const QuadOriginTerms *ot = nullptr) {
    const int own_lo = H.sub_begin, own_hi = H.sub_begin + H.sub_count;
    o.index = -1;
    o.energy = 0.0f;
    o.flags = 0;
    o.row = o.col = -1;
    o.unshadowed = o.unblocked = false;
    o.first_of_facet = first;
    o.replayed = false;

    // Step 1: shadowing (RectangleReceiverRectangleGridRayTracing.cu:14-22)
    int hit = -1;
    // traversals proved hit-free for this ray are skipped, and the others look up the hits that the
    // ray's candidate tests found (never while counting work)
    const bool culling = !STATS && wl;
    RayHits rh;
    int plan = culling ? traversal_plan<U, TM != kTmExact>(wl, wm, 0, origin, dir, a.grid, rh) : 2;
    int out = traversal_outcome<STATS, TM>(a, plan, rh, origin, dir, own_lo, own_hi, li, slot, hit, o.replayed, st);
    if (out == 2) {
        o.status = kRayDeferred;
        return;
    }
    if (out) {
        o.status = QMC_RAY_SHADOWED;
        o.index = hit;
        return;
    }
    o.unshadowed = true;
    // Step 2: perturbed facet normal, reflection, blocking (:24-36)
    V3 normal = normalize(fr_apply(pf, V3{q.x, q.y, q.z}));
    dir = normalize(reflect(-dir, normal));
    plan = culling ? traversal_plan<U, TM != kTmExact>(wl, wm, 1, origin, dir, a.grid, rh) : 2;
    out = traversal_outcome<STATS, TM>(a, plan, rh, origin, dir, own_lo, own_hi, li, slot, hit, o.replayed, st);
    if (out == 2) {
        o.status = kRayDeferred;
        return;
    }
    if (out) {
        o.status = QMC_RAY_BLOCKED;
        o.index = hit;
        return;
    }
    o.unblocked = true;
    // Step 3: receiver hit (Receiver/*ReceiverIntersection.cu)
    float t, u, v;
    int col = -1;  // set by a cylinder's column lookup
    if (!receiver_hit<RT>(a.recv, origin, dir, t, u, v, ot, &col)) {
        o.status = QMC_RAY_MISSED;
        return;
    }
    o.energy = ray_energy(t, dir, normal, H.factor);
    int row = (int)(u * (float)a.recv.res_y);
    if (col < 0) col = (int)(v * (float)a.recv.res_x);
    if (row >= a.recv.res_y) { row = a.recv.res_y - 1; o.flags |= QMC_RAYF_EDGE_CLAMPED; }
    if (col >= a.recv.res_x) { col = a.recv.res_x - 1; o.flags |= QMC_RAYF_EDGE_CLAMPED; }
    o.row = row;
    o.col = col;
    o.index = row * a.recv.res_x + col;
    o.status = QMC_RAY_DEPOSITED;
}

// One ray of map_raytracing: ray id r of list entry li, held in slot `slot`. Its micro-facet m = r / L
// and sample j = r mod L give the pool indices (RectangleReceiverRectangleGridRayTracing.cu:14-16,
// 24-26: group offsets of facets m and m + 1, plus j, modulo the pool size) and the facet's origin,
// then trace_from traces it. HT: the heliostats' type (that of the grid). TM, U as for trace_from.
template <int RT, bool STATS, int TM, int HT, bool U = true>
__device__ __forceinline__ void trace_ray(const TraceArgs &a, const DevHelio &H, const WideList *wl,
                                          const uint64_t *wm, long long r, int li, long long slot, RayOut &o,
                                          RayStats &st) {
    const int L = a.L;
    const int m = (int)fast_div((uint32_t)r, a.fdL);
    const int j = (int)r - m * L;
    const int32_t *grp = a.offsets + H.offsets_base;
    const int P = a.P;
    // both pool indices and samples are read before any test, so the four loads form two
    // dependent pairs instead of a chain of four
    const int a1 = (int)fast_mod((uint32_t)(grp[m] + j), P, a.fdP);
    const int a2 = (int)fast_mod((uint32_t)(grp[fast_mod((uint32_t)(m + 1), P, a.fdP)] + j), P, a.fdP);
    const float4 s = a.pool[a1];  // the sample in world space (k_pool_world)
    const float4 q = a.pert[a2];
    const V3 dir{s.x, s.y, s.z};
    V3 origin, fnormal;
    DevFrame pf;
    facet_geometry<HT>(H, a.facets, m, origin, pf, fnormal);
    trace_from<RT, STATS, TM, U>(a, H, wl, wm, li, slot, j == 0, dir, origin, pf, q, o, st);
}

// One lane per ray slot; grid.y = heliostat of the list, grid.x = blocks of 256 slots. Deposits
// (receiver_drawing, rectangleReceiverIntersection.cu:24) are device-scope atomics into the launch's
// double accumulator, folded into the float map after the launch (k_acc_fold): a pixel can receive
// tens of thousands of nearly equal deposits, and one float accumulator would round each against a
// sum thousands of times larger (a bias near 3e-4 on the reference's example inputs); 64-bit atomics
// run at the same rate as 32-bit ones on MI355X.
// REC: 0 = no records (the map path), 1 = records naming each traversal's first hit, 2 = records of
// the map path's own decisions (a shadowed or blocked record may name another hit quad). DEFER: the map
// path with culling (kTmDefer), else every traversal walks (kTmExact).
template <int RT, int REC, bool STATS, int HT, bool DEFER>
__global__ __launch_bounds__(256) void k_trace(TraceArgs a) {
    static_assert(!DEFER || (!STATS && REC != 1), "deferral is for the map path");
    const int h = a.list[blockIdx.y];
    const DevHelio &H = a.helios[h];
    const long long slot = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long r = slot_ray<HT>(H, a.L, slot);

    __shared__ unsigned long long s_acc[9];
    if (STATS) {
        if (threadIdx.x < 9) s_acc[threadIdx.x] = 0ull;
        __syncthreads();
    }
    RayStats st{0u, 0u, 0u};
    RayOut o;
    const bool active = r >= 0;
    if (!STATS && !active) return;
    if (active) {
        // the wave's candidate masks (slots of a wave are consecutive and 64-aligned)
        const int wave = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * 4u + (threadIdx.x >> 6)));
        const uint64_t *wm = a.wide ? a.wmask + 2 * ((size_t)blockIdx.y * a.wstride + wave) : nullptr;
        trace_ray<RT, STATS, DEFER ? kTmDefer : kTmExact, HT>(a, H, a.wide ? a.wide + h : nullptr, wm, r,
                                                             (int)blockIdx.y, slot, o, st);
        if (o.status == QMC_RAY_DEPOSITED) unsafeAtomicAdd(a.acc + o.index, (double)o.energy);
        if (REC && o.status != kRayDeferred) {
            qmc_ray_record rec;
            rec.status = o.status;
            rec.index = o.index;
            rec.energy = o.energy;
            rec.flags = o.flags;
            a.records[a.rec_base[blockIdx.y] + r] = rec;
        }
    }
    if (STATS) {
        if (active) {
            atomicAdd(&s_acc[0], 1ull);
            if (o.first_of_facet) atomicAdd(&s_acc[7], 1ull);
            if (o.unshadowed) atomicAdd(&s_acc[1], 1ull);
            if (o.unblocked) atomicAdd(&s_acc[2], 1ull);
            if (o.status == QMC_RAY_DEPOSITED) atomicAdd(&s_acc[3], 1ull);
            atomicAdd(&s_acc[4], (unsigned long long)st.cells);
            atomicAdd(&s_acc[5], (unsigned long long)st.candidates);
            atomicAdd(&s_acc[6], (unsigned long long)st.tests);
            if (o.status == QMC_RAY_DEPOSITED && o.flags) atomicAdd(&s_acc[8], 1ull);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long *g = reinterpret_cast<unsigned long long *>(a.stats);
            for (int k = 0; k < 9; ++k)
                if (s_acc[k]) atomicAdd(g + k, s_acc[k]);
        }
    }
}

// Where a workgroup's LDS band lies in the map: rows [row0, row0 + rows), and cols columns from the
// workgroup's own first column c0, cyclically (cols = the map's width: every column, c0 = 0). A band
// narrower than the map (a column window) is for cylinder receivers, whose columns are azimuth: each
// workgroup then traces one contiguous run of the launch's blocks in azimuth order and centres its
// window on its heliostats' azimuth. sectors = 1: the list was dealt from 64 azimuth sectors in turn
// (n / 64 heliostats each, the last n % 64 in azimuth order), and azimuth order visits it through the
// inverse of that permutation; 0: the list is visited in its own order.
struct BandGeom {
    int32_t row0, rows, cols, sectors;
};

// List entry of the s-th heliostat in azimuth order, for a list dealt from 64 sectors of q heliostats
// (entry i = (i mod 64) q + i / 64 in azimuth order for i < 64 q); q = 0: the identity. A bijection of
// [0, n) for every 64 q <= n.
__device__ __forceinline__ int azimuth_entry(int s, int q) { return s < 64 * q ? (s % q) * 64 + s / q : s; }

// The column of a cylinder map facing heliostat H: v = azimuth / 2 pi of the heliostat seen from the
// axis, as the receiver maps a hit point's azimuth to v (CylinderReceiver.cuh:82-86).
__device__ __forceinline__ int azimuth_column(const DevHelio &H, const DevReceiver &R, int W) {
    float phi = atan2f(H.pos[2] - R.pos[2], H.pos[0] - R.pos[0]);
    if (phi < 0.0f) phi += 6.28318531f;
    return min(W - 1, max(0, (int)(phi * (0.159154943f * (float)W))));
}

// Work claiming of a column-window launch. Its T units (64-facet blocks, or 64-slot waves) in azimuth
// order fall into G runs, run k = [T k / G, T (k + 1) / G) for workgroup k, whose window is centred on the
// run's azimuth; claims[k] counts the units of run k handed out (from 0). A wave claims from its own
// run, then, once that is exhausted, from the nearest run (cyclically) that still has units, so the
// workgroups finish together however the units' costs vary over the field. Deposits of claimed units
// outside the claiming workgroup's window go to the device-scope accumulator as any other.
constexpr int kMaxWindowRuns = 2048;  // 32 runs per lane of a wave's exclusion mask

__device__ __forceinline__ long long run_begin(long long T, int G, int k) { return T * k / G; }

// The run nearest to run k with units left according to an agent-scope read of the counters, skipping
// the runs this lane's mask `spent` marks (bit r of lane l: run 64 r + l, found exhausted by a failed
// claim); -1 if there is none. Wave-uniform.
__device__ __forceinline__ int nearest_open_run(const unsigned *claims, long long T, int G, int k, unsigned spent) {
    const int lane = (int)(threadIdx.x & 63u);
    unsigned best = 0xffffffffu;  // (distance, run) as distance * G + run
    for (int r = 0; 64 * r < G; ++r) {
        const int v = 64 * r + lane;
        if (v < G && !((spent >> r) & 1u)) {
            const unsigned t = __hip_atomic_load(claims + v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((long long)t < run_begin(T, G, v + 1) - run_begin(T, G, v)) {
                int d = v > k ? v - k : k - v;
                d = min(d, G - d);
                best = min(best, (unsigned)d * (unsigned)G + (unsigned)v);
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) best = min(best, (unsigned)__shfl_xor((int)best, off));
    best = (unsigned)__builtin_amdgcn_readfirstlane((int)best);
    return best == 0xffffffffu ? -1 : (int)(best % (unsigned)G);
}

// Runs the units of a column-window launch for this wave: unit(u) for every unit u it claims, `own`
// at a time from its workgroup's run and `steal` at a time from others. Every iteration either runs
// units or marks a run exhausted for this wave, so the loop ends. A claim is a device-scope atomic
// whose result the wave waits for (microseconds on MI355X): claims of a few thousand rays keep that
// small against the work they hand out, and the smaller steals keep the workgroups' finishing times
// close (config-3 field, 8 rays per facet: 8-unit own claims and 4-unit steals, 3.24 ms per call,
// against 3.57 ms with 1-unit steals and 4.2 ms with no claims and no stealing).
template <class Unit>
__device__ __forceinline__ void claim_units(unsigned *claims, long long T, int G, int k, unsigned own, unsigned steal,
                                            Unit unit) {
    const int lane = (int)(threadIdx.x & 63u);
    unsigned spent = 0u;
    int v = k;
#pragma nounroll
    while (v >= 0) {
        const long long u0 = run_begin(T, G, v), n = run_begin(T, G, v + 1) - u0;
        const unsigned take = v == k ? own : steal;
        unsigned t = 0u;
        if (lane == 0) t = atomicAdd(claims + v, take);
        t = (unsigned)__builtin_amdgcn_readlane((int)t, 0);
        if ((long long)t < n) {
            const long long e = min(n, (long long)t + take);
#pragma nounroll
            for (long long u = u0 + t; u < u0 + e; ++u) unit(u);
        } else {
            if (lane == (v & 63)) spent |= 1u << (v >> 6);
            v = nearest_open_run(claims, T, G, k, spent);
        }
    }
}

// Units per claim of a column-window launch: about 4096 rays for a workgroup's own run and 2048 for a
// steal, given the rays of one unit.
__device__ __forceinline__ unsigned claim_size(int rays, int unit_rays) { return (unsigned)max(1, rays / unit_rays); }

// A band deposit: inside the band (rows [row0, row0 + rows); with WIN, columns [c0, c0 + cols) modulo the
// map's width W) an LDS atomic, else a device-scope atomic on the launch's double accumulator; with REC
// the ray's record too.
template <int REC, bool WIN>
__device__ __forceinline__ void band_deposit(const TraceArgs &a, float *band, const BandGeom &bg, int W, int c0,
                                             double *acc, int li, long long r, const RayOut &o) {
    if (o.status == QMC_RAY_DEPOSITED) {
        const int br = o.row - bg.row0;
        int bc = o.col;
        if (WIN) {
            bc -= c0;
            bc += bc < 0 ? W : 0;
        }
        if ((unsigned)br < (unsigned)bg.rows && (!WIN || (unsigned)bc < (unsigned)bg.cols))
            atomicAdd(band + br * (WIN ? bg.cols : W) + bc, o.energy);
        else
            unsafeAtomicAdd(acc + o.index, (double)o.energy);
    }
    if (REC && o.status != kRayDeferred) {
        qmc_ray_record rec;
        rec.status = o.status;
        rec.index = o.index;
        rec.energy = o.energy;
        rec.flags = o.flags;
        a.records[ld_uniform(a.rec_base + __builtin_amdgcn_readfirstlane(li)) + r] = rec;
    }
}

// The heliostat-type parameters of k_trace_band for flat and focus-flat heliostats in the facet-shared
// order.
constexpr int kHtFlatShared = 2, kHtFocusShared = 3;

// Persistent form of k_trace for large launches. Global float atomics run at a fixed rate (about
// 21 G/s for scattered addresses on MI355X, against 180 G/s in LDS), and a large field deposits
// hundreds of millions of rays into one map, so each workgroup keeps a band of the map (BandGeom) in LDS
// for the whole launch: deposits inside it are LDS atomics, the others device-scope atomics on the
// launch's double accumulator. At the end each workgroup stores its band, with plain stores, to its own
// slice of `partials` (and, for a column window, its first column after all the slices, followed by
// the runs' claim counters, zero at the launch), which
// k_band_fold or k_window_fold adds into the accumulator. With REC every ray also writes its record at
// records[rec_base[li] + ray id] (parity checks of this kernel's own slot, wave-mask and band indexing).
// DEFER as for k_trace.
//
// Two orders of work over the same slots:
//  * HT = 0 or 1 (the heliostats' type): workgroups take BS-slot blocks (list entry, block), one 64-slot
//    wave each;
//  * HT = kHtFlatShared or kHtFocusShared (flat heliostats with L = 2^lLs rays per facet, 16 <= L <= 64,
//    or focus-flat ones with 2 <= L <= 64: slots that are their ray ids): each wave takes blocks of 64
//    micro-facets (64 L slots), lane k computes facet k's origin, group offsets and (focus-flat)
//    canted-facet index once, and the wave then runs the block's L steps of 64 slots, each lane reading
//    its facet's values from the lane that computed them. The L rays of a facet share all of it, so
//    this is done once per facet instead of once per ray. blocks_per_helio counts such blocks.
//    (Kernels of their own: with both loops in one kernel the scalar registers of one constrain the
//    other.)
// Blocks are dealt round robin over the grid's workgroups (waves), or, with a column window, claimed
// (claim_units) from runs in azimuth order, each workgroup's window centred on its own run.
template <int RT, int BS, int REC, int HT, bool DEFER>
__global__ __launch_bounds__(BS) void k_trace_band(TraceArgs a, BandGeom bg, int blocks_per_helio, float *partials,
                                                    int lLs) {
    static_assert(!DEFER || REC != 1, "deferral is for the map path");
    extern __shared__ float band[];
    // column windows are for cylinder receivers only (RT == 1); for a rectangle receiver cols is W
    constexpr bool WIN = RT == 1;
    const int W = a.recv.res_x, nband = bg.rows * (WIN ? bg.cols : W);
    const bool win = WIN && bg.cols < W;
    for (int p = threadIdx.x; p < nband; p += BS) band[p] = 0.0f;
    // a column window: the launch's units (64-facet blocks in the facet-shared order, else 64-slot
    // waves), this workgroup's run of them and its first column (every lane computes the same values)
    constexpr bool shared_order = HT == kHtFlatShared || HT == kHtFocusShared;
    const int upw = shared_order ? blocks_per_helio : blocks_per_helio * (BS / 64);  // units per list entry
    const long long total_units = (long long)a.n_list * upw;
    int c0 = 0, sq = 0;  // sq: heliostats per azimuth sector of the list (0: the list's own order)
    unsigned *claims = nullptr;
    if (win) {
        claims = reinterpret_cast<unsigned *>(partials + (size_t)gridDim.x * nband) + gridDim.x;
        const long long run0 = run_begin(total_units, (int)gridDim.x, (int)blockIdx.x);
        const long long run1 = run_begin(total_units, (int)gridDim.x, (int)blockIdx.x + 1);
        sq = bg.sectors ? a.n_list / 64 : 0;
        if (run1 > run0) {
            const int ha = ld_uniform(a.list + azimuth_entry((int)(run0 / upw), sq));
            const int hb = ld_uniform(a.list + azimuth_entry((int)((run1 - 1) / upw), sq));
            const int ca = azimuth_column(ld_uniform(a.helios + ha), a.recv, W);
            const int cb = azimuth_column(ld_uniform(a.helios + hb), a.recv, W);
            const int arc = cb - ca + (cb < ca ? W : 0);  // from the run's first azimuth to its last
            c0 = ca + arc / 2 - bg.cols / 2;
            c0 += c0 < 0 ? W : 0;
            c0 -= c0 >= W ? W : 0;
        }
        c0 = __builtin_amdgcn_readfirstlane(c0);
    }
    __syncthreads();
    RayStats st{0u, 0u, 0u};
    constexpr int TM = DEFER ? kTmDefer : kTmExact;
    if constexpr (HT == kHtFlatShared || HT == kHtFocusShared) {
        constexpr bool focus = HT == kHtFocusShared;
        // Launch constants that every step reads — the band's geometry, the map width, the accumulator,
        // the receiver's edges or axis — and (below) a flat heliostat's slope-error frame are kept in
        // vector registers (in_vgpr): the loop's scalar registers run out
        // otherwise, and the compiler then reloads them from the kernel arguments every step, each reload
        // a scalar load the step waits for, and spills others to lanes of vector registers; the vector
        // registers have room (the LDS band holds occupancy at 4 waves per SIMD). PS10-like field:
        // 3.79-3.81 -> 3.72-3.74 ms per launch.
        BandGeom vbg = bg;
        int vW = W;
        double *vacc = a.acc;
        TraceArgs va = a;
        in_vgpr(vbg.row0);
        in_vgpr(vbg.rows);
        in_vgpr(vbg.cols);
        in_vgpr(vW);
        in_vgpr(vacc);
        for (int k = 0; k < 3; ++k) {
            in_vgpr(va.recv.E1[k]);
            in_vgpr(va.recv.E2[k]);
        }
        if (RT == 1) {  // the cylinder's axis, radius and height (config-3 field: 3.08 -> 3.05 ms per call)
            for (int k = 0; k < 3; ++k) in_vgpr(va.recv.pos[k]);
            in_vgpr(va.recv.radius);
            in_vgpr(va.recv.height);
        }
        const int L = 1 << lLs, fpi = 64 >> lLs;  // rays per facet, facets per 64-slot step
        const int lane = (int)(threadIdx.x & 63u);
        const int lfac = lane >> lLs, lray = lane & (L - 1);
        const int P = a.P;
        // one block of 64 facets of list entry li, its first facet f0
        auto facet_block = [&](int li, int f0) {
            const int h = ld_uniform(a.list + __builtin_amdgcn_readfirstlane(li));
            const DevHelio H = ld_uniform(a.helios + h);
            // this lane's facet of the block: origin (RectangleHelio.cu:64-79, FocusFlatRectangleHelio.cu:
            // 21-39,53-64), canted facet (focus-flat), group offsets and, for a rectangle receiver, the
            // origin's terms of the receiver test
            V3 fo{0.0f, 0.0f, 0.0f};
            int g0 = 0, g1 = 0, fx = 0;
            QuadOriginTerms fot{};
            if (f0 + lane < H.F) {
                const int f = f0 + lane;
                fo = focus ? focus_facet_origin(H, a.facets, f, fx) : flat_facet_origin(H, H.fr, f);
                const int32_t *grp = a.offsets + H.offsets_base;
                g0 = grp[f];
                g1 = grp[fast_mod((uint32_t)(f + 1), P, a.fdP)];
                if (RT == 0)
                    fot = quad_origin_terms(fo, V3{a.recv.A[0], a.recv.A[1], a.recv.A[2]},
                                            V3{a.recv.E1[0], a.recv.E1[1], a.recv.E1[2]},
                                            V3{a.recv.E2[0], a.recv.E2[1], a.recv.E2[2]});
            }
            const int nf = min(64, H.F - f0);  // facets of this block
            DevFrame hfr = H.fr;
            if (!focus) vgpr_frame(hfr);
            // the pool samples of step i's rays; every lane's group offsets are in [0, P] (0 past the last
            // facet), so its pool indices are in range
            auto samples = [&](int i, float4 &s, float4 &q) {
                const int src = i * fpi + lfac;
                // (g + j) mod P for g in [0, P] and j < L <= P: one conditional subtraction
                const uint32_t u1 = (uint32_t)(lane_read(g0, src) + lray), u2 = (uint32_t)(lane_read(g1, src) + lray);
                s = a.pool[min(u1, u1 - (uint32_t)P)];  // world space (k_pool_world)
                q = a.pert[min(u2, u2 - (uint32_t)P)];
            };
            const WideList *wl = a.wide ? a.wide + h : nullptr;
            // step i + 1's samples are gathered before step i is traced
            float4 s_cur, q_cur;
            samples(0, s_cur, q_cur);
#pragma nounroll
            for (int i = 0; i * fpi < nf; ++i) {
                const int src = i * fpi + lfac;  // this step's facet for this lane, within the block
                float4 s_nxt = s_cur, q_nxt = q_cur;
                if ((i + 1) * fpi < nf) samples(i + 1, s_nxt, q_nxt);
                const V3 origin{lane_read(fo.x, src), lane_read(fo.y, src), lane_read(fo.z, src)};
                const int fidx = focus ? lane_read(fx, src) : 0;
                QuadOriginTerms ot;
                if (RT == 0) {
                    ot.T = V3{lane_read(fot.T.x, src), lane_read(fot.T.y, src), lane_read(fot.T.z, src)};
                    ot.q = V3{lane_read(fot.q.x, src), lane_read(fot.q.y, src), lane_read(fot.q.z, src)};
                    ot.tq = lane_read(fot.tq, src);
                }
                if (src < nf) {
                    const long long slot = (long long)(f0 + src) * L + lray;  // = the ray id
                    const V3 dir{s_cur.x, s_cur.y, s_cur.z};
                    RayOut o;
                    const int wave = __builtin_amdgcn_readfirstlane((f0 >> 6) * L + i);  // slot >> 6
                    const uint64_t *wm = a.wide ? a.wmask + 2 * ((size_t)li * a.wstride + wave) : nullptr;
                    // the slope-error frame: the heliostat's, or the canted facet's (a small table)
                    const DevFrame pf = focus ? a.facets[fidx].fr_pert : hfr;
                    trace_from<RT, false, TM>(va, H, wl, wm, li, slot, lray == 0, dir, origin, pf, q_cur, o, st,
                                              RT == 0 ? &ot : nullptr);
                    band_deposit<REC, WIN>(a, band, vbg, vW, c0, vacc, li, slot, o);
                }
                s_cur = s_nxt;
                q_cur = q_nxt;
            }
        };
        if (win) {
            claim_units(claims, total_units, (int)gridDim.x, (int)blockIdx.x, claim_size(4096, 64 * L),
                        claim_size(2048, 64 * L), [&](long long u) {
                            const int s = (int)(u / upw);
                            facet_block(azimuth_entry(s, sq), (int)(u - (long long)s * upw) * 64);
                        });
        } else {
            const int nw = (int)gridDim.x * (BS / 64);
            const int gw = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (BS / 64) + (threadIdx.x >> 6)));
            int li = gw / blocks_per_helio;
            int bx = gw - li * blocks_per_helio;
#pragma nounroll
            while (li < a.n_list) {
                facet_block(li, __builtin_amdgcn_readfirstlane(bx) * 64);
                bx += nw;
                while (bx >= blocks_per_helio) {
                    bx -= blocks_per_helio;
                    ++li;
                }
            }
        }
    } else {
        // block (li, bx) of the launch; the heliostat id is made explicitly wave-uniform so that its record
        // is read with scalar loads, as in k_trace
        auto slot_block = [&](int li, int bx) {
            const int h = ld_uniform(a.list + __builtin_amdgcn_readfirstlane(li));
            const DevHelio H = ld_uniform(a.helios + h);
            const long long slot = (long long)__builtin_amdgcn_readfirstlane(bx) * BS + threadIdx.x;
            const long long r = slot_ray<HT>(H, a.L, slot);
            if (r >= 0) {
                RayOut o;
                const int wave = __builtin_amdgcn_readfirstlane(bx * (BS / 64) + (int)(threadIdx.x >> 6));
                const uint64_t *wm = a.wide ? a.wmask + 2 * ((size_t)li * a.wstride + wave) : nullptr;
                trace_ray<RT, false, TM, HT>(a, H, a.wide ? a.wide + h : nullptr, wm, r, li, slot, o, st);
                band_deposit<REC, WIN>(a, band, bg, W, c0, a.acc, li, r, o);
            }
        };
        if (win) {
            // one 64-slot wave of list entry li (wave index j of its slots)
            claim_units(claims, total_units, (int)gridDim.x, (int)blockIdx.x, claim_size(4096, 64),
                        claim_size(2048, 64), [&](long long u) {
                            const int s = __builtin_amdgcn_readfirstlane((int)(u / upw));
                            const int j = __builtin_amdgcn_readfirstlane((int)(u - (long long)s * upw));
                            const int li = azimuth_entry(s, sq);
                            const int h = ld_uniform(a.list + li);
                            const DevHelio H = ld_uniform(a.helios + h);
                            const long long slot = (long long)j * 64 + (threadIdx.x & 63u);
                            const long long r = slot_ray<HT>(H, a.L, slot);
                            if (r >= 0) {
                                RayOut o;
                                const uint64_t *wm =
                                    a.wide ? a.wmask + 2 * ((size_t)li * a.wstride + j) : nullptr;
                                trace_ray<RT, false, TM, HT>(a, H, a.wide ? a.wide + h : nullptr, wm, r, li, slot, o,
                                                             st);
                                band_deposit<REC, WIN>(a, band, bg, W, c0, a.acc, li, r, o);
                            }
                        });
        } else {
            // advanced by gridDim.x blocks per step
            int li = (int)(blockIdx.x / (unsigned)blocks_per_helio);
            int bx = (int)blockIdx.x - li * blocks_per_helio;
#pragma nounroll
            while (li < a.n_list) {
                slot_block(li, bx);
                bx += (int)gridDim.x;
                while (bx >= blocks_per_helio) {
                    bx -= blocks_per_helio;
                    ++li;
                }
            }
        }
    }
    __syncthreads();
    float *mine = partials + (size_t)blockIdx.x * nband;
    for (int p = threadIdx.x; p < nband; p += BS) mine[p] = band[p];
    if (win && threadIdx.x == 0) reinterpret_cast<int *>(partials + (size_t)gridDim.x * nband)[blockIdx.x] = c0;
}

// Adds the workgroups' bands (in double) into rows [row0, row0 + nband / W) of the accumulator;
// grid.y splits the workgroups into slices, each slice adding its sum with one 64-bit atomic.
__global__ __launch_bounds__(256) void k_band_fold(double *acc, const float *partials, int n_parts, int nband,
                                                   int offset) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= nband) return;
    const int per = (n_parts + gridDim.y - 1) / gridDim.y;
    const int w0 = blockIdx.y * per, w1 = min(n_parts, w0 + per);
    double s = 0.0;
    for (int w = w0; w < w1; ++w) s += (double)partials[(size_t)w * nband + p];
    if (s != 0.0) unsafeAtomicAdd(acc + offset + p, s);
}

// Sets n counters to zero (a kernel rather than a memset node, for callers that capture the launch).
__global__ __launch_bounds__(256) void k_zero_counters(unsigned *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Adds the workgroups' column windows (in double) into rows [row0, row0 + rows) of the accumulator.
// Window w holds rows x cols floats from column c0[w] cyclically; the n_parts (<= kMaxWindowRuns) first
// columns follow the windows. A workgroup of 256 threads takes 32 pixels of a row span, its 8 groups
// of 32 threads each summing one slice of the windows in window order (a window that does not cover
// the column adds nothing), and the slices' sums are added in slice order.
__global__ __launch_bounds__(256) void k_window_fold(double *acc, const float *partials, int n_parts, int rows,
                                                     int cols, int W, int row0) {
    __shared__ int s_c0[kMaxWindowRuns];
    __shared__ double s_sum[8][32];
    const size_t nband = (size_t)rows * cols;
    const int *c0 = reinterpret_cast<const int *>(partials + (size_t)n_parts * nband);
    for (int w = threadIdx.x; w < n_parts; w += blockDim.x) s_c0[w] = c0[w];
    __syncthreads();
    const int lp = (int)(threadIdx.x & 31u), slice = (int)(threadIdx.x >> 5);
    const int p = blockIdx.x * 32 + lp;
    const int per = (n_parts + 7) / 8, w0 = slice * per, w1 = min(n_parts, w0 + per);
    double s = 0.0;
    if (p < rows * W) {
        const int r = p / W, c = p - r * W;
        const float *part = partials + (size_t)r * cols;
#pragma unroll 8
        for (int w = w0; w < w1; ++w) {
            int bc = c - s_c0[w];
            bc += bc < 0 ? W : 0;
            if (bc < cols) s += (double)part[(size_t)w * nband + bc];
        }
    }
    s_sum[slice][lp] = s;
    __syncthreads();
    if (slice == 0 && p < rows * W) {
        double t = 0.0;
        for (int k = 0; k < 8; ++k) t += s_sum[k][lp];
        if (t != 0.0) acc[(size_t)row0 * W + p] += t;
    }
}

// One ray that the deferring pass (kTmDefer) left to k_trace_deferred, replayed (kTmReplay) with
// per-lane loads; its deposit goes to the launch's float64 accumulator and, with REC, its record is
// written. `all`: the ray is one of every slot of the launch, and only a ray that the deferring pass
// left open (RayOut.replayed) is output.
template <int RT, int REC, int HT>
__device__ __forceinline__ void replay_ray(const TraceArgs &a, int li, long long slot, bool all) {
    const int h = a.list[li];
    const DevHelio &H = a.helios[h];
    const long long r = slot_ray<HT>(H, a.L, slot);
    if (r < 0) return;
    const uint64_t *wm = a.wmask + 2 * ((size_t)li * a.wstride + (size_t)(slot >> 6));
    RayOut o;
    RayStats st{0u, 0u, 0u};
    trace_ray<RT, false, kTmReplay, HT, false>(a, H, a.wide + h, wm, r, li, slot, o, st);
    if (all && !o.replayed) return;
    if (o.status == QMC_RAY_DEPOSITED) unsafeAtomicAdd(a.acc + o.index, (double)o.energy);
    if (REC) {
        qmc_ray_record rec;
        rec.status = o.status;
        rec.index = o.index;
        rec.energy = o.energy;
        rec.flags = o.flags;
        a.records[a.rec_base[li] + r] = rec;
    }
}

// The rays the map path deferred, traced again; runs after the trace kernel of the same launch. A
// grid-stride loop over the queue's entries (list entry, slot) — or, when more rays were deferred than
// the queue holds, over every slot of the launch (n_list entries of wstride * 64 slots), each ray
// replayed and output only where the deferring pass left it open: slower, and exact the same way.
template <int RT, int REC, int HT>
__global__ __launch_bounds__(256) void k_trace_deferred(TraceArgs a) {
    const unsigned n = *a.defer_n;
    const long long stride = (long long)gridDim.x * blockDim.x;
    const long long k0 = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n <= (unsigned)a.defer_cap) {
        for (long long k = k0; k < (long long)n; k += stride) {
            const unsigned long long key = a.defer[k];
            replay_ray<RT, REC, HT>(a, (int)(key >> 32), (long long)(key & 0xffffffffull), false);
        }
    } else {
        const long long per = (long long)a.wstride * 64;
        const long long total = (long long)a.n_list * per;
        for (long long k = k0; k < total; k += stride) {
            const int li = (int)(k / per);
            replay_ray<RT, REC, HT>(a, li, k - (long long)li * per, true);
        }
    }
}

// Adds a launch's accumulator into the float map (one rounding per pixel) and clears it.
__global__ __launch_bounds__(256) void k_acc_fold(float *flux, double *acc, int npix) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    const double v = acc[p];
    if (v != 0.0) {
        flux[p] = (float)((double)flux[p] + v);
        acc[p] = 0.0;
    }
}

// ---------------------------------------------------------------------------------------------
// Sample pools
// ---------------------------------------------------------------------------------------------

// samplelights::max_less_index (Scene/SceneProcessor.cu:46-59)
__device__ __forceinline__ int max_less_index(const float2 *cdf, float value, int n) {
    int left = 0, right = n;
    while (left <= right) {
        int mid = (left + right) >> 1;
        if (value > cdf[mid].y)
            left = mid + 1;
        else if (value < cdf[mid].y)
            right = mid - 1;
        else
            return mid;
    }
    return right;
}

// samplelights::linear_interpolation body (Scene/SceneProcessor.cu:67-79): uniform -> theta [rad]
__device__ __forceinline__ float buie_theta(const BuieArgs &b, float U) {
    float u = U * b.A;
    if (u < b.I) {
        int id = max_less_index(b.cdf, u, b.n);
        u = (u - b.cdf[id].y) * b.slope[id] + b.cdf[id].x;
        u = atan_r(sqrtf(u / 4.65f) * 0.5016916f);
    } else {
        u = pow_r((u - b.I) * b.B + b.C, 1 / (b.gamma + 1));
        float ri2 = 0.5016916f * 0.5016916f;
        float ro2 = 1.340874f * 1.340874f;
        u = atan_r(sqrtf(ro2 + (u - 4.65f) / 4.65f * (ro2 - ri2)));
    }
    return u / 100.0f;
}

__device__ __forceinline__ float4 sunshape_dir(const BuieArgs &b, float u_theta, float u_phi) {
    float theta = b.model == 0 ? buie_theta(b, u_theta)
                               : (float)(table_inverse_cdf(b.ph_theta, b.ph_cdf, b.ph_lo, b.ph_hi, b.n_phys,
                                                           (double)u_theta) * 1e-3);
    float phi = u_phi * 2 * kMathPi;
    // global_func::angle2xyz (Util/global_function.cuh:211-216)
    float st = sin_r(theta);
    return make_float4(st * cos_r(phi), cos_r(theta), st * sin_r(phi), 0.0f);
}

__device__ __forceinline__ float4 slope_dir(float gx, float gz) {
    V3 n = normalize(V3{gx, 1.0f, gz});  // map_permuation (Scene/SceneProcessor.cu:83-93)
    return make_float4(n.x, n.y, n.z, 0.0f);
}

struct SobolArgs {
    uint32_t v0[32];
    uint32_t v1[32];
    uint32_t mask0, mask1;
};

__device__ __forceinline__ uint32_t sobol_point(const uint32_t *v, uint32_t i) {
    uint32_t x = 0;
    for (int k = 0; i; ++k, i >>= 1)
        if (i & 1u) x ^= v[k];
    return x;
}

__global__ __launch_bounds__(256) void k_pool_sunshape(float4 *pool, int P, SobolArgs s, BuieArgs b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const uint32_t x0 = sobol_point(s.v0, (uint32_t)i) ^ s.mask0;
    const uint32_t x1 = sobol_point(s.v1, (uint32_t)i) ^ s.mask1;
    pool[i] = sunshape_dir(b, u01_open0(x0), u01_open0(x1));
}

// Box-Muller on two 53-bit counter-hash uniforms; N(0, sigma) pairs (gx, gz).
__device__ __forceinline__ void gauss_pair(uint64_t key, uint32_t i, float sigma, float &gx, float &gz) {
    const uint64_t w1 = mix64(key + 2ull * i);
    const uint64_t w2 = mix64(key + 2ull * i + 1ull);
    const double d1 = (double)((w1 >> 11) + 1ull) * (1.0 / 9007199254740992.0);
    const double d2 = (double)(w2 >> 11) * (1.0 / 9007199254740992.0);
    const double rad = sqrt(-2.0 * log(d1));
    const double ang = 6.283185307179586 * d2;
    gx = (float)((double)sigma * rad * cos(ang));
    gz = (float)((double)sigma * rad * sin(ang));
}

__global__ __launch_bounds__(256) void k_pool_perturbation(float4 *pert, int P, uint64_t key, float sigma) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    float gx, gz;
    gauss_pair(key, (uint32_t)i, sigma, gx, gz);
    pert[i] = slope_dir(gx, gz);
}

__global__ __launch_bounds__(256) void k_pool_from_uniforms(float4 *pool, float4 *pert, int P, const float *ut,
                                                            const float *up, const float *gx, const float *gz,
                                                            BuieArgs b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    pool[i] = sunshape_dir(b, ut[i], up[i]);
    pert[i] = slope_dir(gx[i], gz[i]);
}

// ---------------------------------------------------------------------------------------------
// Group offsets, micro-facets
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ int32_t group_offset(uint64_t seed, int h, int m, int P) {
    const uint64_t w = mix64(mix64(seed ^ ((uint64_t)(uint32_t)h << 32)) + (uint64_t)(uint32_t)m);
    const float u = u01_open0((uint32_t)(w >> 32));
    return (int32_t)(u * (float)P);  // map_float2int: int(u*(high-low)+low), low = 0
}

__global__ __launch_bounds__(256) void k_offsets(int32_t *off, const DevHelio *helios, uint64_t seed, int P) {
    const int h = blockIdx.y;
    const DevHelio &H = helios[h];
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m > H.F) return;  // F+1 entries
    off[H.offsets_base + m] = group_offset(seed, h, m, P);
}

// The same for the heliostats of a list only (blockIdx.y = list entry): the fresh group offsets that
// each reference rayTracing call draws for its heliostat (QuasiMonteCarloRayTracer.cpp:135-161 ->
// Heliostat::generateDeviceMicrohelioGroup, Heliostat.cu:57-62).
__global__ __launch_bounds__(256) void k_offsets_list(int32_t *off, const DevHelio *helios, const int32_t *list,
                                                      uint64_t seed, int P) {
    const int h = list[blockIdx.y];
    const DevHelio &H = helios[h];
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m > H.F) return;
    off[H.offsets_base + m] = group_offset(seed, h, m, P);
}

// The sun-shape pool in world space: local2world_rotate of each sample by the sun frame, the rotation
// every ray of map_raytracing applies to its sample (RectangleReceiverRectangleGridRayTracing.cu:14-16),
// done once per pool entry with the same float operations.
__global__ __launch_bounds__(256) void k_pool_world(float4 *world, const float4 *pool, int P, DevFrame sun) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const float4 s = pool[i];
    const V3 d = fr_apply(sun, V3{s.x, s.y, s.z});
    world[i] = make_float4(d.x, d.y, d.z, 0.0f);
}

__global__ __launch_bounds__(256) void k_facets(const DevHelio *helios, const DevFacet *facets, int h, float *origins,
                                                float *normals) {
    const DevHelio &H = helios[h];
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= H.F) return;
    V3 o, n;
    DevFrame pf;
    facet_geometry(H, facets, m, o, pf, n);
    origins[3 * m] = o.x;
    origins[3 * m + 1] = o.y;
    origins[3 * m + 2] = o.z;
    normals[3 * m] = n.x;
    normals[3 * m + 1] = n.y;
    normals[3 * m + 2] = n.z;
}


}  // namespace

// ---------------------------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------------------------
extern "C" qmc_status qmc_create(int32_t device, qmc_ctx **out) {
    if (!out) return fail(QMC_E_INVALID_ARG, "qmc_create: null out");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        return fail(QMC_E_HIP, std::string("qmc_create: no HIP device available (") +
                                   (e != hipSuccess ? hipGetErrorString(e) : "0 devices") + ")");
    if (device < 0 || device >= n) return fail(QMC_E_RANGE, "qmc_create: device index out of range");
    auto *c = new (std::nothrow) qmc_ctx();
    if (!c) return fail(QMC_E_OOM, "qmc_create: allocation failed");
    c->device = device;
    qmc_status st = set_device(c);
    if (st != QMC_OK) {
        delete c;
        return st;
    }
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(QMC_E_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    for (int k = 0; k < 16; ++k) {
        e = hipEventCreate(&c->ev[k]);
        if (e != hipSuccess) {
            delete c;
            return fail(QMC_E_HIP, std::string("hipEventCreate: ") + hipGetErrorString(e));
        }
    }
    if (dev_alloc(c->d_stats, 1) != QMC_OK) {
        delete c;
        return QMC_E_OOM;
    }
    *out = c;
    return QMC_OK;
}

extern "C" qmc_status qmc_destroy(qmc_ctx *c) {
    if (!c) return QMC_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_scene(c);
    dev_free(c->d_pool);
    dev_free(c->d_pert);
    dev_free(c->d_pool_world);
    dev_free(c->d_list);
    dev_free(c->d_stats);
    dev_free(c->d_rec);
    dev_free(c->d_scratch);
    dev_free(c->d_cdf);
    dev_free(c->d_slope);
    dev_free(c->d_aims);
    dev_free(c->d_batch_maps);
    dev_free(c->d_batch_work);
    dev_free(c->d_cull);
    dev_free(c->d_cull_shapes);
    dev_free(c->d_wide);
    dev_free(c->d_wmask);
    dev_free(c->d_defer);
    dev_free(c->d_defer_n);
    dev_free(c->d_rec_base);
    dev_free(c->d_band_partials);
    dev_free(c->d_buie_phys);
    if (c->d_batch_inst) (void)hipFree(c->d_batch_inst);
    c->d_batch_inst = nullptr;
    qmc_adjoint_release(c);
    for (int k = 0; k < 16; ++k)
        if (c->ev[k]) (void)hipEventDestroy(c->ev[k]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return QMC_OK;
}

extern "C" qmc_status qmc_get_stream(qmc_ctx *c, void **stream) {
    if (!c || !stream) return fail(QMC_E_INVALID_ARG, "qmc_get_stream: null argument");
    *stream = (void *)c->stream;
    return QMC_OK;
}

extern "C" qmc_status qmc_sync(qmc_ctx *c) {
    if (!c) return fail(QMC_E_INVALID_ARG, "qmc_sync: null ctx");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return QMC_OK;
}

namespace {

// Per-heliostat device records (frames, layout, energy factor, offsets base) and focus-flat facet
// records of a processed scene; returns the total number of group offsets (F+1 per heliostat).
// Culling shape of a focus-flat heliostat (facet records [f0, f0 + nf) in the heliostat frame): the
// box, in the focus-variant heliostat frame, around every facet's four corners, and the largest angle
// between a facet's world normal and the heliostat normal.
DevCullShape focus_cull_shape(const DevHelio &H, const DevFacet *fc, int nf) {
    auto axes = [](const DevFrame &f, CD3 &u, CD3 &n, CD3 &v) {
        u = f.identity ? CD3{1, 0, 0} : CD3{f.ux, f.uy, f.uz};
        n = f.identity ? CD3{0, 1, 0} : CD3{f.nx, f.ny, f.nz};
        v = f.identity ? CD3{0, 0, 1} : CD3{f.vx, f.vy, f.vz};
    };
    CD3 lo{1e300, 1e300, 1e300}, hi{-1e300, -1e300, -1e300};
    for (int k = 0; k < nf; ++k) {
        CD3 u, n, v;
        axes(fc[k].fr_local, u, n, v);
        for (int i = 0; i < 4; ++i) {
            const double x = (i & 1 ? 0.5 : -0.5) * H.len_col, z = (i & 2 ? 0.5 : -0.5) * H.len_row;
            const CD3 p = cadd(CD3{fc[k].center[0], fc[k].center[1], fc[k].center[2]}, cadd(cmul(u, x), cmul(v, z)));
            lo = CD3{std::fmin(lo.x, p.x), std::fmin(lo.y, p.y), std::fmin(lo.z, p.z)};
            hi = CD3{std::fmax(hi.x, p.x), std::fmax(hi.y, p.y), std::fmax(hi.z, p.z)};
        }
    }
    CD3 u, n, v;
    axes(H.fr_focus, u, n, v);
    const CD3 mid = cmul(cadd(lo, hi), 0.5), half = cmul(csub(hi, lo), 0.5);
    const CD3 c = cadd(CD3{H.pos[0], H.pos[1], H.pos[2]}, cadd(cadd(cmul(u, mid.x), cmul(n, mid.y)), cmul(v, mid.z)));
    const CD3 hu = cmul(u, half.x), hn = cmul(n, half.y), hv = cmul(v, half.z);
    CD3 hu0, nh, hv0;
    axes(H.fr, hu0, nh, hv0);
    double cant = 0.0;
    for (int k = 0; k < nf; ++k) {
        const CD3 w{fc[k].normal[0], fc[k].normal[1], fc[k].normal[2]};
        cant = std::fmax(cant, std::atan2(std::sqrt(cdot(ccross(w, nh), ccross(w, nh))), cdot(w, nh)));
    }
    DevCullShape s;
    s.c[0] = c.x; s.c[1] = c.y; s.c[2] = c.z;
    s.hu[0] = hu.x; s.hu[1] = hu.y; s.hu[2] = hu.z;
    s.hn[0] = hn.x; s.hn[1] = hn.y; s.hn[2] = hn.z;
    s.hv[0] = hv.x; s.hv[1] = hv.y; s.hv[2] = hv.z;
    s.n[0] = nh.x; s.n[1] = nh.y; s.n[2] = nh.z;
    s.cant = cant + 1e-5;  // the float rounding of the traced normals
    return s;
}

int64_t pack_heliostats(const qmc_scene &sc, std::vector<DevHelio> &out, std::vector<DevFacet> &facets,
                        std::vector<DevCullShape> &shapes) {
    const qmc_scene *s = &sc;
    out.resize(s->heliostats.size());
    shapes.resize(s->heliostats.size());
    facets.clear();
    int64_t off_base = 0;
    for (size_t h = 0; h < s->heliostats.size(); ++h) {
        const HostHeliostat &hh = s->heliostats[h];
        DevHelio &d = out[h];
        std::memset(&d, 0, sizeof(d));
        d.fr = to_dev(make_frame(hh.normal, false));
        d.fr_focus = to_dev(make_frame(hh.normal, true));
        for (int k = 0; k < 3; ++k) {
            d.pos[k] = hh.d.pos[k];
            d.size[k] = hh.d.size[k];
        }
        d.gap[0] = hh.d.gap[0];
        d.gap[1] = hh.d.gap[1];
        d.len_row = hh.len_row;
        d.len_col = hh.len_col;
        d.pl = hh.pixel_length;
        d.factor = heliostat_factor(*s, (int)h);
        heliostat_factor_parts(*s, (int)h, d.fac_rule, d.fac_x);
        d.rows = hh.d.rows;
        d.cols = hh.d.cols;
        d.sub_rows = hh.sub_rows;
        d.sub_cols = hh.sub_cols;
        d.F = hh.n_micro;
        d.type = hh.type;
        d.sub_begin = hh.sub_begin;
        d.sub_count = hh.sub_count;
        d.grid = hh.d.grid;
        d.receiver = s->grids[hh.d.grid].d.receiver;
        d.offsets_base = off_base;
        off_base += (int64_t)hh.n_micro + 1;
        const int cpr = d.cols * d.sub_cols;
        d.fd_cpr = make_fastdiv((uint32_t)std::max(1, cpr));
        d.fd_sub_rows = make_fastdiv((uint32_t)std::max(1, d.sub_rows));
        d.fd_sub_cols = make_fastdiv((uint32_t)std::max(1, d.sub_cols));
        d.fd_ntx4 = make_fastdiv((uint32_t)std::max(1, (cpr + 3) / 4));
        d.fd_ntx8 = make_fastdiv((uint32_t)std::max(1, (cpr + 7) / 8));
        d.facet_base = (int)facets.size();
        if (hh.type == 1) {
            Frame fh = make_frame(hh.normal, true);
            for (size_t k = 0; k < hh.loc_normals.size(); ++k) {
                DevFacet f;
                std::memset(&f, 0, sizeof(f));
                f.fr_local = to_dev(make_frame(hh.loc_normals[k], true));
                V3 wn = frame_apply(fh, hh.loc_normals[k]);  // map_microhelio_normals
                f.fr_pert = to_dev(make_frame(wn, false));
                f.center[0] = hh.loc_centers[k].x;
                f.center[1] = hh.loc_centers[k].y;
                f.center[2] = hh.loc_centers[k].z;
                f.normal[0] = wn.x;
                f.normal[1] = wn.y;
                f.normal[2] = wn.z;
                facets.push_back(f);
            }
            shapes[h] = focus_cull_shape(d, facets.data() + d.facet_base, (int)hh.loc_normals.size());
        } else {
            shapes[h] = flat_cull_shape(d, d.fr);
        }
    }
    return off_base;
}

// Blocking quads of a grid: (A, E1, E2) = (V[3i+1], V[3i]-V[3i+1], V[3i+2]-V[3i+1]) as in
// rectGridDDA::intersect (rectGridDDA.cu:15-18).
std::vector<float4> pack_quads(const HostGrid &hg) {
    std::vector<float4> quads((size_t)hg.n_sub * 3);
    for (int i = 0; i < hg.n_sub; ++i) {
        V3 A = hg.sub_verts[3 * i + 1], B = hg.sub_verts[3 * i], C = hg.sub_verts[3 * i + 2];
        V3 E1 = B - A, E2 = C - A;
        quads[3 * i] = make_float4(A.x, A.y, A.z, 0.0f);
        quads[3 * i + 1] = make_float4(E1.x, E1.y, E1.z, 0.0f);
        quads[3 * i + 2] = make_float4(E2.x, E2.y, E2.z, 0.0f);
    }
    return quads;
}

// Each grid's heliostats in azimuth order about the axis of the grid's receiver (ties by id), grid after
// grid (h_az_order, grid g from h_az_begin[g]): kept from qmc_upload and qmc_update, so that a trace
// call only filters it (interleave_by_azimuth). Host memory allocated here, not in trace calls.
void azimuth_order(qmc_ctx *c) {
    const int nh = (int)c->h_helios.size(), ng = (int)c->grids.size();
    struct Key {
        int grid;
        double az;
        int32_t id;
    };
    std::vector<Key> key((size_t)nh);
    for (int h = 0; h < nh; ++h) {
        const DevHelio &H = c->h_helios[h];
        const DevReceiver &R = c->recvs[c->grids[H.grid].receiver].r;
        key[(size_t)h] = {H.grid, std::atan2((double)H.pos[2] - R.pos[2], (double)H.pos[0] - R.pos[0]), h};
    }
    std::sort(key.begin(), key.end(), [](const Key &x, const Key &y) {
        if (x.grid != y.grid) return x.grid < y.grid;
        return x.az != y.az ? x.az < y.az : x.id < y.id;
    });
    c->h_az_order.resize((size_t)nh);
    c->h_az_begin.assign((size_t)ng + 1, 0);
    for (int i = 0; i < nh; ++i) {
        c->h_az_order[(size_t)i] = key[(size_t)i].id;
        ++c->h_az_begin[(size_t)key[(size_t)i].grid + 1];
    }
    for (int g = 0; g < ng; ++g) c->h_az_begin[(size_t)g + 1] += c->h_az_begin[(size_t)g];
    c->h_mark.assign((size_t)nh, 0);
}

}  // namespace

extern "C" qmc_status qmc_upload(qmc_ctx *c, const qmc_scene *s) {
    if (!c || !s) return fail(QMC_E_INVALID_ARG, "qmc_upload: null argument");
    if (!s->processed) return fail(QMC_E_STATE, "qmc_upload: scene not processed (call qmc_scene_process)");
    qmc_status st = set_device(c);
    if (st) return st;
    HIP_TRY(hipStreamSynchronize(c->stream));
    free_scene(c);
    c->cfg = s->cfg;
    // sun frame: local2world_rotate(sample, -sun_dir)
    V3 sun{s->cfg.sun_dir[0], s->cfg.sun_dir[1], s->cfg.sun_dir[2]};
    c->sun = to_dev(make_frame(-sun, false));
    c->pool_world_ready = false;
    // receivers
    c->recvs.resize(s->receivers.size());
    for (size_t i = 0; i < s->receivers.size(); ++i) {
        const HostReceiver &hr = s->receivers[i];
        RecvBuf &rb = c->recvs[i];
        DevReceiver &d = rb.r;
        std::memset(&d, 0, sizeof(d));
        d.type = hr.d.type;
        d.res_x = hr.res_x;
        d.res_y = hr.res_y;
        if (d.type == 0) {
            V3 E1 = hr.vert[1] - hr.vert[0], E2 = hr.vert[3] - hr.vert[0];
            d.A[0] = hr.vert[0].x; d.A[1] = hr.vert[0].y; d.A[2] = hr.vert[0].z;
            d.E1[0] = E1.x; d.E1[1] = E1.y; d.E1[2] = E1.z;
            d.E2[0] = E2.x; d.E2[1] = E2.y; d.E2[2] = E2.z;
        } else {
            d.pos[0] = hr.d.pos[0]; d.pos[1] = hr.d.pos[1]; d.pos[2] = hr.d.pos[2];
            d.radius = hr.d.size[0];
            d.height = hr.d.size[1];
        }
        rb.w = hr.res_x;
        rb.h = hr.res_y;
        if (rb.w <= 0 || rb.h <= 0) {
            // an unprocessed receiver (no grid uses it) keeps an empty 1x1 map
            rb.w = std::max(rb.w, 1);
            rb.h = std::max(rb.h, 1);
        }
        if ((st = dev_alloc(rb.flux, (size_t)rb.w * rb.h))) return st;
        HIP_TRY(hipMemsetAsync(rb.flux, 0, sizeof(float) * (size_t)rb.w * rb.h, c->stream));
        if ((st = dev_alloc(rb.acc, (size_t)rb.w * rb.h))) return st;
        HIP_TRY(hipMemsetAsync(rb.acc, 0, sizeof(double) * (size_t)rb.w * rb.h, c->stream));
        if (d.type == 1 && d.res_x >= 1 && d.res_x <= kColumnTableMaxW) {
            std::vector<float> T(2 * (size_t)d.res_x + 4);
            cylinder_column_table(d.res_x, T.data());
            if ((st = dev_alloc(rb.colt, T.size()))) return st;
            HIP_TRY(hipMemcpy(rb.colt, T.data(), sizeof(float) * T.size(), hipMemcpyHostToDevice));
            d.colt = rb.colt;
        }
    }
    // grids
    c->grids.resize(s->grids.size());
    for (size_t gi = 0; gi < s->grids.size(); ++gi) {
        const HostGrid &hg = s->grids[gi];
        GridBuf &gb = c->grids[gi];
        if ((int)hg.sub_verts.size() != 3 * hg.n_sub)
            return fail(QMC_E_UNSUPPORTED, "qmc_upload: grid " + std::to_string(gi) +
                                               " has a sub-heliostat count override without matching quads");
        std::vector<float4> quads = pack_quads(hg);
        // cell records: the candidate count and the first three candidates of each cell in list
        // order (further candidates stay in the CSR lists)
        const size_t ncell = hg.cell_start.size() - 1;
        std::vector<int4> cells(ncell);
        for (size_t cc = 0; cc < ncell; ++cc) {
            const int b = hg.cell_start[cc], n = hg.cell_start[cc + 1] - b;
            cells[cc] = make_int4(n, n > 0 ? hg.match[b] : -1, n > 1 ? hg.match[b + 1] : -1, n > 2 ? hg.match[b + 2] : -1);
        }
        if ((st = dev_alloc(gb.cells, cells.size()))) return st;
        if (!cells.empty())
            HIP_TRY(hipMemcpyAsync(gb.cells, cells.data(), cells.size() * sizeof(int4), hipMemcpyHostToDevice,
                                   c->stream));
        const int dims[3] = {hg.nx, hg.ny, hg.nz};
        std::vector<int4> boxes(2 * (size_t)hg.n_sub);
        std::vector<float4> pbox(2 * (size_t)hg.n_sub);
        for (int q = 0; q < hg.n_sub; ++q) {
            const int *b = &hg.sub_box[6 * (size_t)q];
            boxes[2 * (size_t)q] = make_int4(b[0], b[2], b[4], 0);
            boxes[2 * (size_t)q + 1] = make_int4(b[1], b[3], b[5], 0);
            const int lo[3] = {b[0], b[2], b[4]}, hi[3] = {b[1], b[3], b[5]};
            const ProofBox pb = proof_box(hg.d.interval, dims, lo, hi);
            pbox[2 * (size_t)q] = make_float4(pb.mid[0], pb.mid[1], pb.mid[2], 0.0f);
            pbox[2 * (size_t)q + 1] = make_float4(pb.lim[0], pb.lim[1], pb.lim[2], 0.0f);
        }
        if ((st = dev_alloc(gb.boxes, boxes.size()))) return st;
        if ((st = dev_alloc(gb.pbox, pbox.size()))) return st;
        if (!boxes.empty()) {
            HIP_TRY(hipMemcpyAsync(gb.boxes, boxes.data(), boxes.size() * sizeof(int4), hipMemcpyHostToDevice,
                                   c->stream));
            HIP_TRY(hipMemcpyAsync(gb.pbox, pbox.data(), pbox.size() * sizeof(float4), hipMemcpyHostToDevice,
                                   c->stream));
        }
        if ((st = dev_alloc(gb.cell_start, hg.cell_start.size()))) return st;
        if ((st = dev_alloc(gb.match, hg.match.size()))) return st;
        if ((st = dev_alloc(gb.quads, quads.size()))) return st;
        HIP_TRY(hipMemcpyAsync(gb.cell_start, hg.cell_start.data(), hg.cell_start.size() * sizeof(int32_t),
                               hipMemcpyHostToDevice, c->stream));
        if (!hg.match.empty())
            HIP_TRY(hipMemcpyAsync(gb.match, hg.match.data(), hg.match.size() * sizeof(int32_t),
                                   hipMemcpyHostToDevice, c->stream));
        if (!quads.empty())
            HIP_TRY(hipMemcpyAsync(gb.quads, quads.data(), quads.size() * sizeof(float4), hipMemcpyHostToDevice,
                                   c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        DevGrid &g = gb.g;
        for (int k = 0; k < 3; ++k) {
            g.pos[k] = hg.d.pos[k];
            g.interval[k] = hg.d.interval[k];
        }
        g.nx = hg.nx;
        g.ny = hg.ny;
        g.nz = hg.nz;
        const ProofFrame fr = proof_frame(g.pos, g.interval, dims);
        if ((st = dev_alloc(gb.frame, 1))) return st;
        HIP_TRY(hipMemcpy(gb.frame, &fr, sizeof(fr), hipMemcpyHostToDevice));
        g.frame = gb.frame;
        g.cells = gb.cells;
        g.cell_start = gb.cell_start;
        g.match = gb.match;
        g.quads = gb.quads;
        g.boxes = gb.boxes;
        g.pbox = gb.pbox;
        gb.n_sub = hg.n_sub;
        gb.receiver = hg.d.receiver;
        gb.htype = hg.d.heliostat_type;
    }
    // heliostats and focus-flat facets
    std::vector<DevFacet> facets;
    std::vector<DevCullShape> shapes;
    int64_t off_base = pack_heliostats(*s, c->h_helios, facets, shapes);
    aim_points_f64(*s, c->h_aim64);
    c->offsets_total = off_base;
    for (size_t gi = 0; gi < s->grids.size(); ++gi)
        if ((st = upload_groups(c, (int)gi, pack_quads(s->grids[gi])))) return st;
    if ((st = dev_alloc(c->d_helios, c->h_helios.size()))) return st;
    if ((st = dev_alloc(c->d_facets, facets.size()))) return st;
    if ((st = dev_alloc(c->d_offsets, (size_t)off_base))) return st;
    if (!c->h_helios.empty())
        HIP_TRY(hipMemcpyAsync(c->d_helios, c->h_helios.data(), c->h_helios.size() * sizeof(DevHelio),
                               hipMemcpyHostToDevice, c->stream));
    if (!facets.empty())
        HIP_TRY(hipMemcpyAsync(c->d_facets, facets.data(), facets.size() * sizeof(DevFacet), hipMemcpyHostToDevice,
                               c->stream));
    if ((st = dev_alloc(c->d_cull_shapes, shapes.size()))) return st;
    if (!shapes.empty())
        HIP_TRY(hipMemcpyAsync(c->d_cull_shapes, shapes.data(), shapes.size() * sizeof(DevCullShape),
                               hipMemcpyHostToDevice, c->stream));
    if ((int)s->heliostats.size() > c->list_cap) {
        if ((st = dev_alloc(c->d_list, s->heliostats.size()))) return st;
        c->list_cap = (int)s->heliostats.size();
    }
    // host bookkeeping of trace calls of up to max(heliostats, 65536) entries (no host allocation then)
    {
        const size_t nres = std::max<size_t>(s->heliostats.size(), 65536);
        c->h_occ.reserve(s->heliostats.size());
        c->h_pass.reserve(nres);
        c->h_order.reserve(nres);
        c->h_sorted.reserve(nres);
        c->h_sorted_base.reserve(nres);
        c->h_tmp.reserve(nres);
        c->h_spans.reserve(nres);
    }
    // Buie table of this configuration
    if (s->cfg.inverse_transform_sampling_groups > 0 && s->cfg.csr > 0.0f) {
        std::vector<float> cdf, sl;
        float abcig[5];
        sunshape_params(s->cfg, cdf, sl, abcig);
        if ((st = dev_alloc(c->d_cdf, cdf.size() / 2))) return st;
        if ((st = dev_alloc(c->d_slope, sl.size()))) return st;
        HIP_TRY(hipMemcpyAsync(c->d_cdf, cdf.data(), cdf.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipMemcpyAsync(c->d_slope, sl.data(), sl.size() * sizeof(float), hipMemcpyHostToDevice, c->stream));
        c->buie.cdf = c->d_cdf;
        c->buie.slope = c->d_slope;
        c->buie.n = s->cfg.inverse_transform_sampling_groups;
        c->buie.A = abcig[0];
        c->buie.B = abcig[1];
        c->buie.C = abcig[2];
        c->buie.I = abcig[3];
        c->buie.gamma = abcig[4];
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    azimuth_order(c);
    c->uploaded = true;
    c->offsets_ready = false;
    if (c->pool_ready && c->P != s->cfg.num_sunshape_groups * s->cfg.num_sunshape_lights_per_group)
        c->pool_ready = false;
    return QMC_OK;
}

extern "C" qmc_status qmc_update(qmc_ctx *c, const qmc_scene *s) {
    if (!c || !s) return fail(QMC_E_INVALID_ARG, "qmc_update: null argument");
    if (!c->uploaded) return fail(QMC_E_STATE, "qmc_update: nothing uploaded (call qmc_upload)");
    if (!s->processed) return fail(QMC_E_STATE, "qmc_update: scene not processed (call qmc_scene_process)");
    const qmc_config &o = c->cfg, &n = s->cfg;
    bool same = s->heliostats.size() == c->h_helios.size() && s->grids.size() == c->grids.size() &&
                s->receivers.size() == c->recvs.size() &&
                o.num_sunshape_groups == n.num_sunshape_groups &&
                o.num_sunshape_lights_per_group == n.num_sunshape_lights_per_group;
    for (size_t g = 0; same && g < s->grids.size(); ++g)
        same = s->grids[g].n_sub == c->grids[g].n_sub && (int)s->grids[g].sub_verts.size() == 3 * s->grids[g].n_sub;
    for (size_t r = 0; same && r < s->receivers.size(); ++r)
        same = s->receivers[r].res_x * s->receivers[r].res_y == c->recvs[r].w * c->recvs[r].h;
    std::vector<DevHelio> helios;
    std::vector<DevFacet> facets;
    std::vector<DevCullShape> shapes;
    if (same) {
        pack_heliostats(*s, helios, facets, shapes);
        for (size_t h = 0; same && h < helios.size(); ++h)
            same = helios[h].F == c->h_helios[h].F && helios[h].type == c->h_helios[h].type &&
                   helios[h].sub_count == c->h_helios[h].sub_count && helios[h].grid == c->h_helios[h].grid;
    }
    if (!same) return fail(QMC_E_STATE, "qmc_update: scene layout differs from the uploaded one (use qmc_upload)");
    qmc_status st = set_device(c);
    if (st) return st;
    c->cfg = n;
    V3 sun{n.sun_dir[0], n.sun_dir[1], n.sun_dir[2]};
    c->sun = to_dev(make_frame(-sun, false));
    c->pool_world_ready = false;
    c->h_helios = helios;
    azimuth_order(c);
    aim_points_f64(*s, c->h_aim64);
    HIP_TRY(hipMemcpyAsync(c->d_helios, helios.data(), helios.size() * sizeof(DevHelio), hipMemcpyHostToDevice,
                           c->stream));
    if (!facets.empty())
        HIP_TRY(hipMemcpyAsync(c->d_facets, facets.data(), facets.size() * sizeof(DevFacet), hipMemcpyHostToDevice,
                               c->stream));
    if (!shapes.empty())
        HIP_TRY(hipMemcpyAsync(c->d_cull_shapes, shapes.data(), shapes.size() * sizeof(DevCullShape),
                               hipMemcpyHostToDevice, c->stream));
    for (size_t g = 0; g < s->grids.size(); ++g) {
        std::vector<float4> q = pack_quads(s->grids[g]);
        if (!q.empty())
            HIP_TRY(hipMemcpyAsync(c->grids[g].quads, q.data(), q.size() * sizeof(float4), hipMemcpyHostToDevice,
                                   c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        if ((st = upload_groups(c, (int)g, q))) return st;
    }
    // pageable sources: wait until the copies have consumed them
    HIP_TRY(hipStreamSynchronize(c->stream));
    return QMC_OK;
}

// ---------------------------------------------------------------------------------------------
// Pools
// ---------------------------------------------------------------------------------------------
static qmc_status ensure_pool(qmc_ctx *c, int P, int L) {
    if (P <= 0 || L <= 0) return fail(QMC_E_INVALID_ARG, "pool size and lights per group must be positive");
    if (P > (1 << 30) || L > P) return fail(QMC_E_RANGE, "pool size above 2^30 or fewer entries than lights per group");
    if (c->P != P) c->offsets_ready = false;  // group offsets index the pool: [0, P]
    if (c->P != P || !c->d_pool) {
        qmc_status st;
        if ((st = dev_alloc(c->d_pool, (size_t)P))) return st;
        if ((st = dev_alloc(c->d_pert, (size_t)P))) return st;
        if ((st = dev_alloc(c->d_pool_world, (size_t)P))) return st;
    }
    c->pool_world_ready = false;  // the caller writes a new pool
    c->P = P;
    c->L = L;
    return QMC_OK;
}

// Cone half-angles of the current pools for traversal culling: alpha = largest angle of a sun-shape
// sample from the local y axis, beta = the same for the slope-error samples; shadow rays lie within
// alpha of the sun axis, reflected rays within alpha + 2 beta of the ideal reflection. Computed in
// double from the device pools, widened by 1e-4 rad; culling is off beyond 80 degrees.
static qmc_status update_cull_angles(qmc_ctx *c) {
    std::vector<float4> a((size_t)c->P), b((size_t)c->P);
    HIP_TRY(hipMemcpyAsync(a.data(), c->d_pool, sizeof(float4) * a.size(), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(b.data(), c->d_pert, sizeof(float4) * b.size(), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    auto max_angle = [](const std::vector<float4> &v) {
        double worst = 0.0;
        for (const float4 &e : v) {
            const double len = std::sqrt((double)e.x * e.x + (double)e.y * e.y + (double)e.z * e.z);
            if (!(len > 0.0)) return 10.0;
            const double cs = std::fmax(-1.0, std::fmin(1.0, e.y / len));
            worst = std::fmax(worst, std::acos(cs));
        }
        return worst;
    };
    const double alpha = max_angle(a) + 1e-4, beta = max_angle(b) + 1e-4;
    c->cull_tan_sun = alpha < kCullMaxAngle ? std::tan(alpha) : -1.0;
    c->cull_ang_block = alpha + 2.0 * beta < kCullMaxAngle ? alpha + 2.0 * beta : -1.0;
    c->cull_tan_block = c->cull_ang_block >= 0.0 ? std::tan(c->cull_ang_block) : -1.0;
    return QMC_OK;
}

extern "C" qmc_status qmc_pool_generate(qmc_ctx *c, const qmc_pool_params *p) {
    if (!c || !p) return fail(QMC_E_INVALID_ARG, "qmc_pool_generate: null argument");
    if (!c->uploaded) return fail(QMC_E_STATE, "qmc_pool_generate: upload a scene first");
    qmc_status st = set_device(c);
    if (st) return st;
    const int L = c->cfg.num_sunshape_lights_per_group;
    const int P = c->cfg.num_sunshape_groups * L;
    if (!c->d_cdf) return fail(QMC_E_INVALID_ARG, "qmc_pool_generate: sun-shape table missing (csr, groups)");
    if (p->sunshape_model != QMC_SUNSHAPE_REFERENCE && p->sunshape_model != QMC_SUNSHAPE_BUIE)
        return fail(QMC_E_INVALID_ARG, "qmc_pool_generate: unknown sunshape_model " + std::to_string(p->sunshape_model));
    if ((st = ensure_pool(c, P, L))) return st;
    BuieArgs buie = c->buie;
    buie.model = p->sunshape_model;
    if (buie.model == QMC_SUNSHAPE_BUIE) {
        BuieTable t;
        buie_physical_table(c->cfg, t);
        const int nb = (int)t.dens_lo.size();
        std::vector<double> packed;
        packed.reserve(4 * (size_t)nb + 2);
        packed.insert(packed.end(), t.theta.begin(), t.theta.end());
        packed.insert(packed.end(), t.cdf.begin(), t.cdf.end());
        packed.insert(packed.end(), t.dens_lo.begin(), t.dens_lo.end());
        packed.insert(packed.end(), t.dens_hi.begin(), t.dens_hi.end());
        if ((st = dev_alloc(c->d_buie_phys, packed.size()))) return st;
        HIP_TRY(hipMemcpyAsync(c->d_buie_phys, packed.data(), packed.size() * sizeof(double), hipMemcpyHostToDevice,
                               c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));  // pageable source
        buie.n_phys = nb;
        buie.ph_theta = c->d_buie_phys;
        buie.ph_cdf = c->d_buie_phys + nb + 1;
        buie.ph_lo = c->d_buie_phys + 2 * (nb + 1);
        buie.ph_hi = c->d_buie_phys + 2 * (nb + 1) + nb;
    }
    SobolArgs sa;
    sobol_directions(sa.v0, sa.v1);
    sa.mask0 = (uint32_t)(mix64(p->seed_sunshape) >> 32);
    sa.mask1 = (uint32_t)(mix64(p->seed_sunshape ^ 0xA5A5A5A5A5A5A5A5ull) >> 32);
    const int blocks = (P + 255) / 256;
    hipLaunchKernelGGL(k_pool_sunshape, dim3(blocks), dim3(256), 0, c->stream, c->d_pool, P, sa, buie);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_pool_perturbation, dim3(blocks), dim3(256), 0, c->stream, c->d_pert, P,
                       mix64(p->seed_perturbation), c->cfg.helio_disturb_std);
    HIP_TRY(hipGetLastError());
    if ((st = update_cull_angles(c))) return st;
    c->pool_ready = true;
    return QMC_OK;
}

extern "C" qmc_status qmc_pool_from_uniforms(qmc_ctx *c, const float *ut, const float *up, const float *gx,
                                             const float *gz) {
    if (!c || !ut || !up || !gx || !gz) return fail(QMC_E_INVALID_ARG, "qmc_pool_from_uniforms: null argument");
    if (!c->uploaded || !c->d_cdf) return fail(QMC_E_STATE, "qmc_pool_from_uniforms: upload a scene first");
    qmc_status st = set_device(c);
    if (st) return st;
    const int L = c->cfg.num_sunshape_lights_per_group;
    const int P = c->cfg.num_sunshape_groups * L;
    if ((st = ensure_pool(c, P, L))) return st;
    float *tmp = nullptr;
    if ((st = dev_alloc(tmp, 4 * (size_t)P))) return st;
    const float *src[4] = {ut, up, gx, gz};
    for (int k = 0; k < 4; ++k) {
        hipError_t e = hipMemcpyAsync(tmp + (size_t)k * P, src[k], sizeof(float) * P, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) {
            dev_free(tmp);
            return fail(QMC_E_HIP, std::string("hipMemcpyAsync: ") + hipGetErrorString(e));
        }
    }
    hipLaunchKernelGGL(k_pool_from_uniforms, dim3((P + 255) / 256), dim3(256), 0, c->stream, c->d_pool, c->d_pert, P,
                       tmp, tmp + P, tmp + 2 * (size_t)P, tmp + 3 * (size_t)P, c->buie);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dev_free(tmp);
    if (e != hipSuccess) return fail(QMC_E_HIP, std::string("k_pool_from_uniforms: ") + hipGetErrorString(e));
    if ((st = update_cull_angles(c))) return st;
    c->pool_ready = true;
    return QMC_OK;
}

extern "C" qmc_status qmc_pool_set(qmc_ctx *c, const float *dirs, const float *perts, int32_t P, int32_t L) {
    if (!c || !dirs || !perts) return fail(QMC_E_INVALID_ARG, "qmc_pool_set: null argument");
    qmc_status st = set_device(c);
    if (st) return st;
    if ((st = ensure_pool(c, P, L))) return st;
    std::vector<float4> a((size_t)P), b((size_t)P);
    for (int i = 0; i < P; ++i) {
        a[i] = make_float4(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2], 0.0f);
        b[i] = make_float4(perts[3 * i], perts[3 * i + 1], perts[3 * i + 2], 0.0f);
    }
    HIP_TRY(hipMemcpyAsync(c->d_pool, a.data(), sizeof(float4) * P, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->d_pert, b.data(), sizeof(float4) * P, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if ((st = update_cull_angles(c))) return st;
    c->pool_ready = true;
    return QMC_OK;
}

extern "C" qmc_status qmc_pool_get(qmc_ctx *c, float *dirs, float *perts, int32_t P) {
    if (!c) return fail(QMC_E_INVALID_ARG, "qmc_pool_get: null ctx");
    if (!c->pool_ready) return fail(QMC_E_STATE, "qmc_pool_get: no pool");
    if (P != c->P) return fail(QMC_E_INVALID_ARG, "qmc_pool_get: P mismatch");
    qmc_status st = set_device(c);
    if (st) return st;
    std::vector<float4> a((size_t)P), b((size_t)P);
    HIP_TRY(hipMemcpyAsync(a.data(), c->d_pool, sizeof(float4) * P, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(b.data(), c->d_pert, sizeof(float4) * P, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int i = 0; i < P; ++i) {
        if (dirs) {
            dirs[3 * i] = a[i].x;
            dirs[3 * i + 1] = a[i].y;
            dirs[3 * i + 2] = a[i].z;
        }
        if (perts) {
            perts[3 * i] = b[i].x;
            perts[3 * i + 1] = b[i].y;
            perts[3 * i + 2] = b[i].z;
        }
    }
    return QMC_OK;
}

// ---------------------------------------------------------------------------------------------
// Offsets and facets
// ---------------------------------------------------------------------------------------------
static int max_F(const qmc_ctx *c) {
    int m = 0;
    for (auto &h : c->h_helios) m = std::max(m, h.F);
    return m;
}

extern "C" qmc_status qmc_offsets_generate(qmc_ctx *c, uint64_t seed) {
    if (!c) return fail(QMC_E_INVALID_ARG, "qmc_offsets_generate: null ctx");
    if (!c->uploaded) return fail(QMC_E_STATE, "qmc_offsets_generate: upload a scene first");
    if (!c->pool_ready) return fail(QMC_E_STATE, "qmc_offsets_generate: set or generate the pool first (offsets are in [0, P])");
    qmc_status st = set_device(c);
    if (st) return st;
    const int nh = (int)c->h_helios.size();
    if (nh == 0) return QMC_OK;
    if (nh > 65535) return fail(QMC_E_UNSUPPORTED, "qmc_offsets_generate: more than 65535 heliostats");
    const int bx = (max_F(c) + 1 + 255) / 256;
    hipLaunchKernelGGL(k_offsets, dim3(bx, nh), dim3(256), 0, c->stream, c->d_offsets, c->d_helios, seed, c->P);
    HIP_TRY(hipGetLastError());
    c->offsets_ready = true;
    return QMC_OK;
}

extern "C" qmc_status qmc_offsets_set(qmc_ctx *c, int32_t h, const int32_t *off, int32_t n) {
    if (!c || !off) return fail(QMC_E_INVALID_ARG, "qmc_offsets_set: null argument");
    if (!c->uploaded) return fail(QMC_E_STATE, "qmc_offsets_set: upload a scene first");
    if (h < 0 || h >= (int)c->h_helios.size()) return fail(QMC_E_RANGE, "qmc_offsets_set: heliostat out of range");
    const DevHelio &H = c->h_helios[h];
    if (n != H.F + 1) return fail(QMC_E_INVALID_ARG, "qmc_offsets_set: need F+1 = " + std::to_string(H.F + 1) + " offsets");
    if (!c->pool_ready) return fail(QMC_E_STATE, "qmc_offsets_set: set or generate the pool first (offsets are in [0, P])");
    for (int32_t i = 0; i < n; ++i)
        if (off[i] < 0 || off[i] > c->P)
            return fail(QMC_E_RANGE, "qmc_offsets_set: offset " + std::to_string(i) + " = " + std::to_string(off[i]) +
                                         " outside [0, " + std::to_string(c->P) + "]");
    qmc_status st = set_device(c);
    if (st) return st;
    HIP_TRY(hipMemcpyAsync(c->d_offsets + H.offsets_base, off, sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->offsets_ready = true;
    return QMC_OK;
}

extern "C" qmc_status qmc_offsets_get(qmc_ctx *c, int32_t h, int32_t *off, int32_t n) {
    if (!c || !off) return fail(QMC_E_INVALID_ARG, "qmc_offsets_get: null argument");
    if (!c->uploaded) return fail(QMC_E_STATE, "qmc_offsets_get: upload a scene first");
    if (h < 0 || h >= (int)c->h_helios.size()) return fail(QMC_E_RANGE, "qmc_offsets_get: heliostat out of range");
    const DevHelio &H = c->h_helios[h];
    if (n != H.F + 1) return fail(QMC_E_INVALID_ARG, "qmc_offsets_get: need F+1 offsets");
    qmc_status st = set_device(c);
    if (st) return st;
    HIP_TRY(hipMemcpyAsync(off, c->d_offsets + H.offsets_base, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return QMC_OK;
}

extern "C" qmc_status qmc_facets_get(qmc_ctx *c, int32_t h, float *origins, float *normals, int32_t F) {
    if (!c || !origins || !normals) return fail(QMC_E_INVALID_ARG, "qmc_facets_get: null argument");
    if (!c->uploaded) return fail(QMC_E_STATE, "qmc_facets_get: upload a scene first");
    if (h < 0 || h >= (int)c->h_helios.size()) return fail(QMC_E_RANGE, "qmc_facets_get: heliostat out of range");
    if (F != c->h_helios[h].F) return fail(QMC_E_INVALID_ARG, "qmc_facets_get: F mismatch");
    if (F == 0) return QMC_OK;
    qmc_status st = set_device(c);
    if (st) return st;
    float *buf = nullptr;
    if ((st = dev_alloc(buf, 6 * (size_t)F))) return st;
    hipLaunchKernelGGL(k_facets, dim3((F + 255) / 256), dim3(256), 0, c->stream, c->d_helios, c->d_facets, h, buf,
                       buf + 3 * (size_t)F);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(origins, buf, sizeof(float) * 3 * F, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(normals, buf + 3 * (size_t)F, sizeof(float) * 3 * F, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dev_free(buf);
    if (e != hipSuccess) return fail(QMC_E_HIP, std::string("qmc_facets_get: ") + hipGetErrorString(e));
    return QMC_OK;
}

// ---------------------------------------------------------------------------------------------
// Trace
// ---------------------------------------------------------------------------------------------
namespace {

constexpr int kBandBS = 1024;

// The trace kernels for a receiver type RT, record mode REC, the grid's heliostat type ht and whether
// the launch defers (a.defer_cap > 0: the map path with culling).
template <int RT, int REC, bool DEFER>
void launch_band_d(int ht, dim3 g, size_t lds, hipStream_t st, const TraceArgs &a, const BandGeom &bg, int bph,
                   float *parts, int lLs) {
    if (ht == 0 && lLs)
        hipLaunchKernelGGL((k_trace_band<RT, kBandBS, REC, kHtFlatShared, DEFER>), g, dim3(kBandBS), lds, st, a, bg, bph,
                           parts, lLs);
    else if (lLs)
        hipLaunchKernelGGL((k_trace_band<RT, kBandBS, REC, kHtFocusShared, DEFER>), g, dim3(kBandBS), lds, st, a, bg, bph,
                           parts, lLs);
    else if (ht == 0)
        hipLaunchKernelGGL((k_trace_band<RT, kBandBS, REC, 0, DEFER>), g, dim3(kBandBS), lds, st, a, bg, bph, parts,
                           0);
    else
        hipLaunchKernelGGL((k_trace_band<RT, kBandBS, REC, 1, DEFER>), g, dim3(kBandBS), lds, st, a, bg, bph, parts,
                           0);
}

template <int RT, int REC>
void launch_band(int ht, dim3 g, size_t lds, hipStream_t st, const TraceArgs &a, const BandGeom &bg, int bph,
                 float *parts, int lLs) {
    if (REC != 1 && a.defer_cap > 0) launch_band_d<RT, REC, REC != 1>(ht, g, lds, st, a, bg, bph, parts, lLs);
    else launch_band_d<RT, REC, false>(ht, g, lds, st, a, bg, bph, parts, lLs);
}

template <int RT, int REC, bool D>
hipError_t band_occupancy_d(int ht, bool shared, int &resident, size_t lds) {
    if (ht == 0 && shared)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, k_trace_band<RT, kBandBS, REC, kHtFlatShared, D>,
                                                           kBandBS, lds);
    if (shared)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, k_trace_band<RT, kBandBS, REC, kHtFocusShared, D>,
                                                           kBandBS, lds);
    return ht == 0 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, k_trace_band<RT, kBandBS, REC, 0, D>,
                                                                 kBandBS, lds)
                   : hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, k_trace_band<RT, kBandBS, REC, 1, D>,
                                                                 kBandBS, lds);
}

template <int RT, int REC>
hipError_t band_occupancy_rt(int ht, bool defer, bool shared, int &resident, size_t lds) {
    constexpr bool D = REC != 1;
    if (defer && D) return band_occupancy_d<RT, REC, D>(ht, shared, resident, lds);
    return band_occupancy_d<RT, REC, false>(ht, shared, resident, lds);
}

template <int REC>
hipError_t band_occupancy(int rt, int ht, bool defer, bool shared, int &resident, size_t lds) {
    return rt == 0 ? band_occupancy_rt<0, REC>(ht, defer, shared, resident, lds)
                   : band_occupancy_rt<1, REC>(ht, defer, shared, resident, lds);
}

template <int RT, int REC, bool STATS>
void launch_per_ray(int ht, dim3 g, dim3 blk, hipStream_t st, const TraceArgs &a) {
    constexpr bool D = !STATS && REC != 1;
    if (D && a.defer_cap > 0) {
        if (ht == 0) hipLaunchKernelGGL((k_trace<RT, REC, STATS, 0, D>), g, blk, 0, st, a);
        else hipLaunchKernelGGL((k_trace<RT, REC, STATS, 1, D>), g, blk, 0, st, a);
    } else {
        if (ht == 0) hipLaunchKernelGGL((k_trace<RT, REC, STATS, 0, false>), g, blk, 0, st, a);
        else hipLaunchKernelGGL((k_trace<RT, REC, STATS, 1, false>), g, blk, 0, st, a);
    }
}

template <int RT, int REC>
void launch_deferred_rt(int ht, dim3 g, hipStream_t st, const TraceArgs &a) {
    if (ht == 0) hipLaunchKernelGGL((k_trace_deferred<RT, REC, 0>), g, dim3(256), 0, st, a);
    else hipLaunchKernelGGL((k_trace_deferred<RT, REC, 1>), g, dim3(256), 0, st, a);
}

// The deferred re-trace of a launch whose trace kernel queued rays (a.defer_cap > 0): a grid of two
// workgroups per CU striding over the queue (empty queues cost one short launch).
qmc_status launch_deferred(qmc_ctx *c, const TraceArgs &a, int ht, int rec_mode) {
    if (a.defer_cap <= 0) return QMC_OK;
    const dim3 g((unsigned)(2 * std::max(c->num_cus, 1)));
    if (a.recv.type == 0) {
        if (rec_mode == 2) launch_deferred_rt<0, 2>(ht, g, c->stream, a);
        else launch_deferred_rt<0, 0>(ht, g, c->stream, a);
    } else {
        if (rec_mode == 2) launch_deferred_rt<1, 2>(ht, g, c->stream, a);
        else launch_deferred_rt<1, 0>(ht, g, c->stream, a);
    }
    HIP_TRY(hipGetLastError());
    return QMC_OK;
}

// The banded persistent trace when it pays (or when forced, for parity checks of this kernel): one
// workgroup per CU with the whole LDS (4 waves per SIMD). Device-scope atomics bound the one-lane-per-ray
// kernel on large fields (about 24 G/s on MI355X), so the band is used where it holds at least a quarter
// of the map's rows, or, on a cylinder, a column window at least a quarter of the azimuth (measured on the
// config-3 cylinder field: 80 of 256 rows taking 48.5% of its deposits, 6.3 -> 4.7 ms per call), and
// where every workgroup deposits many times its band (its band store and fold are then a small part of
// the launch). sectors: the launch's list was dealt from 64 azimuth sectors (interleave_by_azimuth).
qmc_status try_band_launch(qmc_ctx *c, const RecvBuf &rb, const TraceArgs &a, long long max_slots, TraceBufs &b,
                           bool force, int rec_mode, int ht, bool sectors, bool &launched) {
    launched = false;
    const int per_cu = 1;
    if (qmc_status st = device_limits(c)) return st;
    const int W = rb.w, Hh = rb.h;
    const size_t lds_budget = std::min(c->max_lds, c->lds_per_cu / (size_t)per_cu);
    const int rows = (int)std::min<size_t>((size_t)Hh, lds_budget / (sizeof(float) * (size_t)W));
    const bool rec = a.records != nullptr;
    if (rows < 1) {
        if (force) return fail(QMC_E_UNSUPPORTED, "banded trace: one map row does not fit in LDS");
        return QMC_OK;
    }
    // a cylinder map whose rows do not all fit takes a full-height column window instead where it spans
    // at least a quarter of the azimuth: the heliostats of a workgroup's run face a few degrees of it, so
    // the window catches most of their deposits (config-3 field: 160 of 512 columns, 91% of the deposits,
    // against 48% for 80 whole rows)
    BandGeom bg{(Hh - rows) / 2, rows, W, 0};  // row band: flux concentrates around the receiver centre
    if (a.recv.type == 1 && rows < Hh && c->num_cus * per_cu <= kMaxWindowRuns) {
        const int cols = (int)std::min<size_t>((size_t)W, lds_budget / (sizeof(float) * (size_t)Hh));
        if (4 * cols >= W) bg = BandGeom{0, Hh, cols, sectors ? 1 : 0};
    }
    const bool window = bg.cols < W;
    if (!force && !window && 4 * rows < Hh) return QMC_OK;
    const long long nband = (long long)bg.rows * bg.cols;
    const size_t lds = sizeof(float) * (size_t)nband;
    // flat heliostats with 16, 32 or 64 rays per micro-facet and focus-flat ones with 2 to 64 (a power of
    // two) take the facet-shared order (k_trace_band): blocks of 64 facets per wave; their slots are their
    // ray ids (slot_ray)
    const int lL = a.L == 2 ? 1 : a.L == 4 ? 2 : a.L == 8 ? 3 : a.L == 16 ? 4 : a.L == 32 ? 5 : a.L == 64 ? 6 : 0;
    const int lLs = a.P < a.L ? 0 : ht == 0 ? (lL >= 4 ? lL : 0) : lL;
    int resident = 0;
    const bool defer = a.defer_cap > 0;
    const hipError_t e = rec ? band_occupancy<1>(a.recv.type, ht, defer, lLs > 0, resident, lds)
                             : band_occupancy<0>(a.recv.type, ht, defer, lLs > 0, resident, lds);
    if (e != hipSuccess || resident < 1) {
        if (force) return fail(QMC_E_UNSUPPORTED, "banded trace: no workgroup of the banded kernel is resident");
        return QMC_OK;
    }
    const int nwg = c->num_cus * std::min(per_cu, resident);
    const long long bph = lLs ? (max_slots / a.L + 63) / 64 : (max_slots + kBandBS - 1) / kBandBS;
    const long long total = lLs ? bph * 64 * a.L * a.n_list : bph * kBandBS * a.n_list;
    if (!force && total < 8 * (long long)nwg * nband) return QMC_OK;
    // the workgroups' bands, then (column windows) their first columns and their runs' claim counters, one
    // int in a float's slot each
    const size_t need = (size_t)nwg * nband + (window ? 2 * (size_t)nwg : 0);
    if (need > b.band_cap) {
        if (b.fixed) {
            if (force) return fail(QMC_E_RANGE, "banded trace: the workgroup bands exceed the scratch");
            return QMC_OK;
        }
        HIP_TRY(hipStreamSynchronize(c->stream));  // earlier launches may still write the old bands
        qmc_status st = dev_alloc(c->d_band_partials, need);
        if (st) {
            c->band_cap = 0;
            return st;
        }
        c->band_cap = need;
        b.band = c->d_band_partials;
        b.band_cap = c->band_cap;
    }
    if (window) {
        hipLaunchKernelGGL(k_zero_counters, dim3((unsigned)((nwg + 255) / 256)), dim3(256), 0, c->stream,
                           reinterpret_cast<unsigned *>(b.band + (size_t)nwg * nband) + nwg, nwg);
        HIP_TRY(hipGetLastError());
    }
    const bool mark = !rec;
    if (mark && c->kev_before >= 0) HIP_TRY(hipEventRecord(c->ev[c->kev_before], c->stream));
    if (rec && rec_mode == 2) {
        if (a.recv.type == 0) launch_band<0, 2>(ht, dim3(nwg), lds, c->stream, a, bg, (int)bph, b.band, lLs);
        else launch_band<1, 2>(ht, dim3(nwg), lds, c->stream, a, bg, (int)bph, b.band, lLs);
    } else if (rec) {
        if (a.recv.type == 0) launch_band<0, 1>(ht, dim3(nwg), lds, c->stream, a, bg, (int)bph, b.band, lLs);
        else launch_band<1, 1>(ht, dim3(nwg), lds, c->stream, a, bg, (int)bph, b.band, lLs);
    } else {
        if (a.recv.type == 0) launch_band<0, 0>(ht, dim3(nwg), lds, c->stream, a, bg, (int)bph, b.band, lLs);
        else launch_band<1, 0>(ht, dim3(nwg), lds, c->stream, a, bg, (int)bph, b.band, lLs);
    }
    HIP_TRY(hipGetLastError());
    if (qmc_status st = launch_deferred(c, a, ht, rec_mode)) return st;
    if (mark && c->kev_after >= 0) HIP_TRY(hipEventRecord(c->ev[c->kev_after], c->stream));
    if (window)
        hipLaunchKernelGGL(k_window_fold, dim3((unsigned)(((size_t)Hh * W + 31) / 32)), dim3(256), 0, c->stream,
                           rb.acc, b.band, nwg, Hh, bg.cols, W, 0);
    else
        hipLaunchKernelGGL(k_band_fold, dim3((unsigned)((nband + 255) / 256), 8), dim3(256), 0, c->stream, rb.acc,
                           b.band, nwg, (int)nband, bg.row0 * W);
    HIP_TRY(hipGetLastError());
    launched = true;
    return QMC_OK;
}

// One launch over n list entries of one grid (d_list: their heliostat ids): culling kernels, the trace
// kernel (QMC_KERNEL_*, with QMC_KERNEL_MAP_DECISIONS: records of the map path's decisions) and the
// fold of the launch's accumulator into the map. The wave masks of the launch must fit in b (the
// caller sizes or splits launches). sectors: the list was dealt from 64 azimuth sectors
// (interleave_by_azimuth over exactly these n entries).
qmc_status launch_trace(qmc_ctx *c, int grid, const int32_t *d_list, int n, long long max_slots, qmc_ray_record *rec,
                        const long long *rec_base, StatsAcc *stats, int kernel, TraceBufs &b, bool sectors) {
    const GridBuf &gb = c->grids[grid];
    const RecvBuf &rb = c->recvs[gb.receiver];
    TraceArgs a;
    std::memset(&a, 0, sizeof(a));
    a.grid = gb.g;
    a.recv = rb.r;
    a.helios = c->d_helios;
    a.facets = c->d_facets;
    if (!c->pool_world_ready) {
        hipLaunchKernelGGL(k_pool_world, dim3((unsigned)((c->P + 255) / 256)), dim3(256), 0, c->stream,
                           c->d_pool_world, c->d_pool, c->P, c->sun);
        HIP_TRY(hipGetLastError());
        c->pool_world_ready = true;
    }
    a.pool = c->d_pool_world;
    a.pert = c->d_pert;
    a.offsets = c->d_offsets;
    a.list = d_list;
    a.flux = rb.flux;
    a.acc = rb.acc;
    a.records = rec;
    a.rec_base = rec_base;
    a.stats = stats;
    a.P = c->P;
    a.L = c->L;
    a.fdP = make_fastdiv((uint32_t)c->P);
    a.fdL = make_fastdiv((uint32_t)c->L);
    a.n_list = n;
    const long long bx = (max_slots + 255) / 256;
    // ray ids and slots are divided by multiply-high constants, exact below 2^31
    if (max_slots >= 0x7fffffffLL) return fail(QMC_E_UNSUPPORTED, "trace: 2^31 or more rays per heliostat");
    if (n > 65535) return fail(QMC_E_UNSUPPORTED, "trace: more than 65535 heliostats in one grid launch");
    a.blocks_per_helio = (int)bx;
    if (bx == 0 || n == 0) return QMC_OK;
    qmc_status st;
    const int rec_mode = !rec ? 0 : (kernel & QMC_KERNEL_MAP_DECISIONS) ? 2 : 1;
    kernel &= ~QMC_KERNEL_MAP_DECISIONS;
    if (!stats && rec_mode != 1 && b.defer_cap > 0 && b.defer && b.defer_n) {
        a.defer = b.defer;
        a.defer_n = b.defer_n;
        a.defer_cap = (int32_t)std::min<size_t>({b.defer_cap, c->defer_limit ? c->defer_limit : b.defer_cap,
                                                 (size_t)0x7fffffff});
        c->last_defer_n = b.defer_n;
    }
    if (!stats) {
        if ((st = launch_cull_into(c, grid, d_list, n, b.cull, b.wide, a.defer_n))) return st;
        a.wide = b.wide;
        if (2 * (size_t)n * (size_t)wave_stride(max_slots) > b.wmask_cap)
            return fail(QMC_E_STATE, "trace: wave masks exceed their buffer");
        if ((st = launch_cull_waves_into(c, grid, d_list, n, max_slots, b.wmask, a))) return st;
    }
    const int npix = rb.w * rb.h;
    if (!stats && kernel != QMC_KERNEL_PER_RAY) {
        bool launched = false;
        if ((st = try_band_launch(c, rb, a, max_slots, b, kernel == QMC_KERNEL_BANDED, rec_mode, gb.htype, sectors,
                                  launched)))
            return st;
        if (launched) {
            hipLaunchKernelGGL(k_acc_fold, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, c->stream, rb.flux,
                               rb.acc, npix);
            HIP_TRY(hipGetLastError());
            return QMC_OK;
        }
    }
    dim3 g((unsigned)bx, (unsigned)n), blk(256);
    const int rt = rb.r.type;
    const bool mark = !rec && !stats;
    if (mark && c->kev_before >= 0) HIP_TRY(hipEventRecord(c->ev[c->kev_before], c->stream));
    const int ht = gb.htype;
    if (rec_mode == 2) {
        if (rt == 0) launch_per_ray<0, 2, false>(ht, g, blk, c->stream, a);
        else launch_per_ray<1, 2, false>(ht, g, blk, c->stream, a);
    } else if (rec) {
        if (rt == 0) launch_per_ray<0, 1, false>(ht, g, blk, c->stream, a);
        else launch_per_ray<1, 1, false>(ht, g, blk, c->stream, a);
    } else if (stats) {
        if (rt == 0) launch_per_ray<0, 0, true>(ht, g, blk, c->stream, a);
        else launch_per_ray<1, 0, true>(ht, g, blk, c->stream, a);
    } else {
        if (rt == 0) launch_per_ray<0, 0, false>(ht, g, blk, c->stream, a);
        else launch_per_ray<1, 0, false>(ht, g, blk, c->stream, a);
    }
    HIP_TRY(hipGetLastError());
    if ((st = launch_deferred(c, a, ht, rec_mode))) return st;
    if (mark && c->kev_after >= 0) HIP_TRY(hipEventRecord(c->ev[c->kev_after], c->stream));
    hipLaunchKernelGGL(k_acc_fold, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, c->stream, rb.flux, rb.acc,
                       npix);
    HIP_TRY(hipGetLastError());
    return QMC_OK;
}

// Reorders a cylinder grid's heliostats for a plain trace. A cylinder map is lit around the azimuth
// facing each heliostat, and a trace kernel that deals heliostats round robin has only a few of them in
// flight at a time: taken one from each of 64 azimuth sectors in turn, those in flight light different
// columns, so their atomic adds meet on the same pixels less often (2650-heliostat 512x256 field: 6.92
// ms against 7.47 in layout order and 7.76 sorted by azimuth). The banded trace visits the list in
// azimuth order through the inverse permutation (azimuth_entry). ids: n distinct heliostats of grid g,
// put in azimuth order by filtering the grid's (azimuth_order) — ties in id order — then dealt.
void interleave_by_azimuth(qmc_ctx *c, int g, int32_t *ids, int n) {
    constexpr int kSectors = 64;
    std::vector<uint8_t> &mark = c->h_mark;
    std::vector<int32_t> &az = c->h_tmp;
    az.resize((size_t)n);
    for (int i = 0; i < n; ++i) mark[(size_t)ids[i]] = 1;
    int k = 0;
    for (int j = c->h_az_begin[(size_t)g]; j < c->h_az_begin[(size_t)g + 1]; ++j) {
        const int32_t h = c->h_az_order[(size_t)j];
        if (mark[(size_t)h]) {
            mark[(size_t)h] = 0;
            az[(size_t)k++] = h;
        }
    }
    if (k != n) {  // not n distinct heliostats of the grid: keep the list as it is
        for (int i = 0; i < n; ++i) mark[(size_t)ids[i]] = 0;
        return;
    }
    const int q = n / kSectors;  // heliostats per sector; the last n % kSectors follow in azimuth order
    for (int i = 0; i < n; ++i)
        ids[i] = az[(size_t)(i < kSectors * q ? (i % kSectors) * q + i / kSectors : i)];
}

// Ctx-owned buffers of a trace call of n entries (grown as needed; record bases when rec).
qmc_status ctx_trace_bufs(qmc_ctx *c, int n, bool rec, size_t wmask_need, size_t defer_need, TraceBufs &b) {
    qmc_status st;
    const size_t nh = c->h_helios.size();
    // launches queued by earlier calls may still read the buffers that are about to be replaced
    if (n > c->list_cap || c->cull_cap < nh || c->wide_cap < nh || wmask_need > c->wmask_cap ||
        (rec && (size_t)n > c->rec_base_cap) || defer_need > c->defer_cap)
        HIP_TRY(hipStreamSynchronize(c->stream));
    if (defer_need > c->defer_cap) {
        if ((st = dev_alloc(c->d_defer, defer_need))) {
            c->defer_cap = 0;
            return st;
        }
        c->defer_cap = defer_need;
    }
    if (!c->d_defer_n && (st = dev_alloc(c->d_defer_n, 1))) return st;
    if (n > c->list_cap) {
        if ((st = dev_alloc(c->d_list, (size_t)n))) {
            c->list_cap = 0;
            return st;
        }
        c->list_cap = n;
    }
    if (c->cull_cap < nh) {
        if ((st = dev_alloc(c->d_cull, nh))) {
            c->cull_cap = 0;
            return st;
        }
        c->cull_cap = nh;
    }
    if (c->wide_cap < nh) {
        if ((st = dev_alloc(c->d_wide, nh))) {
            c->wide_cap = 0;
            return st;
        }
        c->wide_cap = nh;
    }
    if (wmask_need > c->wmask_cap) {
        if ((st = dev_alloc(c->d_wmask, wmask_need))) {
            c->wmask_cap = 0;
            return st;
        }
        c->wmask_cap = wmask_need;
    }
    if (rec && (size_t)n > c->rec_base_cap) {
        if ((st = dev_alloc(c->d_rec_base, (size_t)n))) {
            c->rec_base_cap = 0;
            return st;
        }
        c->rec_base_cap = (size_t)n;
    }
    b.list = c->d_list;
    b.cull = c->d_cull;
    b.wide = c->d_wide;
    b.wmask = c->d_wmask;
    b.wmask_cap = c->wmask_cap;
    b.band = c->d_band_partials;
    b.band_cap = c->band_cap;
    b.defer = c->d_defer;
    b.defer_n = c->d_defer_n;
    b.defer_cap = c->defer_cap;
    b.fixed = false;
    return QMC_OK;
}

// Ray slots of a heliostat in a plain trace: rays, or tiled ray slots of flat heliostats (k_trace).
long long heliostat_slots(const qmc_ctx *c, const DevHelio &H) { return heliostat_slot_count(H, c->L); }

// A trace call over n list entries. Each occurrence of a heliostat in the list is one reference
// rayTracing call (QuasiMonteCarloRayTracer.cpp:15-91); occurrences are traced in passes — pass p
// holds every heliostat's (p+1)-th occurrence — so that no launch lists a heliostat twice (its
// candidate lists are written per heliostat id). Within a pass heliostats are traced grid by grid in
// task order (a plain trace of a cylinder grid interleaves them). With resampling on, each pass first
// draws fresh group offsets for its heliostats. rec_base_in (with rec): first record of each entry.
// fixed: buffers from the scratch array (b), no allocation.
qmc_status trace_common(qmc_ctx *c, const int32_t *ids, int n, qmc_ray_record *rec, const long long *rec_base_in,
                        StatsAcc *stats, int kernel, TraceBufs *fixed) {
    if (!c->uploaded) return fail(QMC_E_STATE, "trace: upload a scene first");
    if (!c->pool_ready) return fail(QMC_E_STATE, "trace: set or generate the sample pools first");
    if (!c->offsets_ready) return fail(QMC_E_STATE, "trace: set or generate the micro-facet group offsets first");
    if (n < 0 || (n > 0 && !ids)) return fail(QMC_E_INVALID_ARG, "trace: bad heliostat list");
    const int choice = kernel & ~QMC_KERNEL_MAP_DECISIONS;
    if (choice != QMC_KERNEL_AUTO && choice != QMC_KERNEL_PER_RAY && choice != QMC_KERNEL_BANDED)
        return fail(QMC_E_INVALID_ARG, "trace: unknown kernel choice " + std::to_string(kernel));
    const int nh = (int)c->h_helios.size();
    for (int i = 0; i < n; ++i)
        if (ids[i] < 0 || ids[i] >= nh)
            return fail(QMC_E_RANGE, "The ray tracing heliostat index " + std::to_string(ids[i]) +
                                         " is invalid.The heliostat index should between 0 and " + std::to_string(nh) + ".");
    if (n == 0) return QMC_OK;
    qmc_status st = set_device(c);
    if (st) return st;
    // pass of each entry: the number of earlier occurrences of its heliostat
    std::vector<int32_t> &occ = c->h_occ, &pass = c->h_pass, &order = c->h_order, &sorted = c->h_sorted;
    std::vector<long long> &sbase = c->h_sorted_base;
    std::vector<TraceSpan> &spans = c->h_spans;
    occ.assign((size_t)nh, 0);
    pass.resize((size_t)n);
    order.resize((size_t)n);
    bool in_order = true;  // every heliostat once, grids in order: the list is already in that order
    for (int i = 0; i < n; ++i) {
        pass[i] = occ[ids[i]]++;
        order[i] = i;
        in_order = in_order && pass[i] == 0 &&
                   (i == 0 || c->h_helios[ids[i]].grid >= c->h_helios[ids[i - 1]].grid);
    }
    // entries ordered by (pass, grid, task order) (std::sort: no temporary buffer)
    if (!in_order)
        std::sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
            if (pass[x] != pass[y]) return pass[x] < pass[y];
            const int gx = c->h_helios[ids[x]].grid, gy = c->h_helios[ids[y]].grid;
            return gx != gy ? gx < gy : x < y;
        });
    sorted.resize((size_t)n);
    sbase.resize(rec ? (size_t)n : 0);
    spans.clear();
    for (int i = 0; i < n; ++i) {
        sorted[i] = ids[order[i]];
        if (rec) sbase[i] = rec_base_in[order[i]];
    }
    size_t wmask_need = 0, defer_need = 0;
    // the map path defers the rays whose traversals its candidate tests and walk-free proof leave open
    // (none to a few per 10^4 rays on the fields measured; a launch that overflows its queue replays
    // all its rays instead)
    const bool deferring = !stats && (!rec || (kernel & QMC_KERNEL_MAP_DECISIONS));
    for (int i = 0; i < n;) {
        const int p = pass[order[i]], g = c->h_helios[sorted[i]].grid;
        int e = i + 1;
        while (e < n && pass[order[e]] == p && c->h_helios[sorted[e]].grid == g) ++e;
        const DevReceiver &R = c->recvs[c->grids[g].receiver].r;
        const bool sectors = !rec && !stats && R.type == 1;
        if (sectors) interleave_by_azimuth(c, g, sorted.data() + i, e - i);
        long long max_slots = 0;
        for (int k = i; k < e; ++k) max_slots = std::max(max_slots, heliostat_slots(c, c->h_helios[sorted[k]]));
        if (max_slots > 0x7fffffffLL) return fail(QMC_E_UNSUPPORTED, "trace: too many rays per heliostat");
        spans.push_back(TraceSpan{p, g, i, e - i, max_slots, sectors});
        wmask_need = std::max(wmask_need, 2 * (size_t)(e - i) * (size_t)wave_stride(max_slots));
        if (deferring) defer_need = std::max(defer_need, std::max<size_t>(65536, (size_t)(e - i) * (size_t)max_slots / 256));
        i = e;
    }
    TraceBufs b;
    if (fixed) {
        b = *fixed;
    } else if ((st = ctx_trace_bufs(c, n, rec != nullptr, stats ? 0 : wmask_need, defer_need, b))) {
        return st;
    }
    HIP_TRY(hipMemcpyAsync(b.list, sorted.data(), sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, c->stream));
    if (rec)
        HIP_TRY(hipMemcpyAsync(c->d_rec_base, sbase.data(), sizeof(long long) * (size_t)n, hipMemcpyHostToDevice,
                               c->stream));
    int fresh_pass = -1;
    for (size_t si = 0; si < spans.size(); ++si) {
        const TraceSpan &sp = spans[si];
        if (c->resample && sp.pass != fresh_pass) {
            // the pass's entries are contiguous in the list: spans si.. of this pass
            int end = sp.begin + sp.count;
            for (size_t sj = si + 1; sj < spans.size() && spans[sj].pass == sp.pass; ++sj) end = spans[sj].begin + spans[sj].count;
            int maxf = 0;
            for (int k = sp.begin; k < end; ++k) maxf = std::max(maxf, c->h_helios[sorted[k]].F);
            const uint64_t seed = c->resample_seed + c->resample_next++;
            hipLaunchKernelGGL(k_offsets_list, dim3((unsigned)((maxf + 1 + 255) / 256), (unsigned)(end - sp.begin)),
                               dim3(256), 0, c->stream, c->d_offsets, c->d_helios, b.list + sp.begin, seed, c->P);
            HIP_TRY(hipGetLastError());
            fresh_pass = sp.pass;
        }
        // a fixed mask buffer takes the span in chunks of whole heliostats
        int chunk = sp.count;
        if (b.fixed && !stats) {
            const size_t per = 2 * (size_t)wave_stride(sp.max_slots);
            const size_t fit = b.wmask_cap / per;
            if (fit == 0)
                return fail(QMC_E_RANGE, "trace: one heliostat's wave masks (" + std::to_string(per * 8) +
                                             " B) exceed the scratch");
            chunk = (int)std::min<size_t>((size_t)sp.count, fit);
        }
        for (int k = 0; k < sp.count; k += chunk) {
            const int m = std::min(chunk, sp.count - k);
            if ((st = launch_trace(c, sp.grid, b.list + sp.begin + k, m, sp.max_slots, rec,
                                   rec ? c->d_rec_base + sp.begin + k : nullptr, stats, kernel, b,
                                   sp.sectors && m == sp.count)))
                return st;
        }
    }
    return QMC_OK;
}

}  // namespace

qmc_status qmc::trace_fixed(qmc_ctx *c, const int32_t *ids, int n, const TraceBufs &b) {
    TraceBufs fb = b;
    fb.fixed = true;
    return trace_common(c, ids, n, nullptr, nullptr, nullptr, QMC_KERNEL_AUTO, &fb);
}

extern "C" qmc_status qmc_trace(qmc_ctx *c, const int32_t *ids, int32_t n) {
    if (!c) return fail(QMC_E_INVALID_ARG, "qmc_trace: null ctx");
    return trace_common(c, ids, n, nullptr, nullptr, nullptr, QMC_KERNEL_AUTO, nullptr);
}

extern "C" qmc_status qmc_offsets_resample(qmc_ctx *c, int32_t enable, uint64_t seed) {
    if (!c) return fail(QMC_E_INVALID_ARG, "qmc_offsets_resample: null ctx");
    if (enable != 0 && enable != 1) return fail(QMC_E_INVALID_ARG, "qmc_offsets_resample: enable must be 0 or 1");
    c->resample = enable == 1;
    c->resample_seed = seed;
    c->resample_next = 0;
    return QMC_OK;
}

extern "C" qmc_status qmc_set_defer_capacity(qmc_ctx *c, int64_t rays) {
    if (!c) return fail(QMC_E_INVALID_ARG, "qmc_set_defer_capacity: null ctx");
    if (rays < 0) return fail(QMC_E_INVALID_ARG, "qmc_set_defer_capacity: negative capacity");
    c->defer_limit = (size_t)rays;
    return QMC_OK;
}

extern "C" qmc_status qmc_get_defer_count(qmc_ctx *c, int64_t *rays) {
    if (!c || !rays) return fail(QMC_E_INVALID_ARG, "qmc_get_defer_count: null argument");
    *rays = 0;
    if (!c->last_defer_n) return QMC_OK;
    qmc_status st = set_device(c);
    if (st) return st;
    unsigned n = 0;
    HIP_TRY(hipMemcpyAsync(&n, c->last_defer_n, sizeof(n), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    *rays = (int64_t)n;
    return QMC_OK;
}

extern "C" qmc_status qmc_cull_flags(qmc_ctx *c, const int32_t *ids, int32_t n, uint8_t *flags) {
    if (!c || (n > 0 && (!ids || !flags))) return fail(QMC_E_INVALID_ARG, "qmc_cull_flags: null argument");
    if (!c->uploaded) return fail(QMC_E_STATE, "qmc_cull_flags: upload a scene first");
    if (!c->pool_ready) return fail(QMC_E_STATE, "qmc_cull_flags: set or generate the sample pools first");
    const int nh = (int)c->h_helios.size();
    for (int i = 0; i < n; ++i)
        if (ids[i] < 0 || ids[i] >= nh) return fail(QMC_E_RANGE, "qmc_cull_flags: heliostat out of range");
    qmc_status st = set_device(c);
    if (st) return st;
    std::vector<CullList> all((size_t)nh);
    std::vector<char> seen((size_t)nh, 0);
    for (size_t g = 0; g < c->grids.size(); ++g) {
        // each heliostat once per launch: k_cull writes its lists per heliostat id
        std::vector<int32_t> list;
        for (int i = 0; i < n; ++i)
            if (c->h_helios[ids[i]].grid == (int)g && !seen[ids[i]]) {
                seen[ids[i]] = 1;
                list.push_back(ids[i]);
            }
        if (list.empty()) continue;
        TraceBufs b;
        if ((st = ctx_trace_bufs(c, (int)list.size(), false, 0, 0, b))) return st;
        HIP_TRY(hipMemcpyAsync(b.list, list.data(), sizeof(int32_t) * list.size(), hipMemcpyHostToDevice, c->stream));
        if ((st = launch_cull_into(c, (int)g, b.list, (int)list.size(), b.cull, nullptr))) return st;
        HIP_TRY(hipMemcpyAsync(all.data(), b.cull, sizeof(CullList) * (size_t)nh, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));  // the list and candidate buffers are reused per grid
        for (int i = 0; i < n; ++i)
            if (c->h_helios[ids[i]].grid == (int)g)
                flags[i] = (uint8_t)((all[ids[i]].n[0] > 0 ? 1 : 0) | (all[ids[i]].n[1] > 0 ? 2 : 0));
    }
    return QMC_OK;
}

extern "C" qmc_status qmc_trace_records_many(qmc_ctx *c, const int32_t *ids, int32_t n, int32_t kernel,
                                             qmc_ray_record *rec, int64_t n_records) {
    if (!c || !rec || (n > 0 && !ids)) return fail(QMC_E_INVALID_ARG, "qmc_trace_records_many: null argument");
    if (!c->uploaded) return fail(QMC_E_STATE, "qmc_trace_records_many: upload a scene first");
    const int nh = (int)c->h_helios.size();
    std::vector<long long> base((size_t)std::max(n, 0));
    std::vector<char> seen((size_t)nh, 0);
    long long need = 0;
    for (int i = 0; i < n; ++i) {
        if (ids[i] < 0 || ids[i] >= nh) return fail(QMC_E_RANGE, "qmc_trace_records_many: heliostat out of range");
        if (seen[ids[i]]) return fail(QMC_E_INVALID_ARG, "qmc_trace_records_many: heliostat " +
                                                             std::to_string(ids[i]) + " listed twice");
        seen[ids[i]] = 1;
        base[i] = need;
        need += (long long)c->h_helios[ids[i]].F * c->L;
    }
    if (n_records != need)
        return fail(QMC_E_INVALID_ARG, "qmc_trace_records_many: need sum(F*L) = " + std::to_string(need) + " records");
    qmc_status st = set_device(c);
    if (st) return st;
    if ((size_t)need > c->rec_cap) {
        if ((st = dev_alloc(c->d_rec, (size_t)std::max<long long>(need, 1)))) {
            c->rec_cap = 0;
            return st;
        }
        c->rec_cap = (size_t)need;
    }
    if ((st = trace_common(c, ids, n, c->d_rec, base.data(), nullptr, kernel, nullptr))) return st;
    if (need > 0)
        HIP_TRY(hipMemcpyAsync(rec, c->d_rec, sizeof(qmc_ray_record) * (size_t)need, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return QMC_OK;
}

extern "C" qmc_status qmc_trace_records(qmc_ctx *c, int32_t h, qmc_ray_record *rec, int64_t n) {
    if (!c || !rec) return fail(QMC_E_INVALID_ARG, "qmc_trace_records: null argument");
    if (!c->uploaded) return fail(QMC_E_STATE, "qmc_trace_records: upload a scene first");
    if (h < 0 || h >= (int)c->h_helios.size()) return fail(QMC_E_RANGE, "qmc_trace_records: heliostat out of range");
    const int64_t need = (int64_t)c->h_helios[h].F * c->L;
    if (n != need) return fail(QMC_E_INVALID_ARG, "qmc_trace_records: need F*L = " + std::to_string(need) + " records");
    return qmc_trace_records_many(c, &h, 1, QMC_KERNEL_AUTO, rec, n);
}

extern "C" qmc_status qmc_trace_stats_run(qmc_ctx *c, const int32_t *ids, int32_t n, qmc_trace_stats *out) {
    if (!c || !out) return fail(QMC_E_INVALID_ARG, "qmc_trace_stats_run: null argument");
    qmc_status st = set_device(c);
    if (st) return st;
    HIP_TRY(hipMemsetAsync(c->d_stats, 0, sizeof(StatsAcc), c->stream));
    if ((st = trace_common(c, ids, n, nullptr, nullptr, c->d_stats, QMC_KERNEL_AUTO, nullptr))) return st;
    StatsAcc h;
    HIP_TRY(hipMemcpyAsync(&h, c->d_stats, sizeof(StatsAcc), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    out->rays = h.rays;
    out->unshadowed = h.unshadowed;
    out->unblocked = h.unblocked;
    out->deposits = h.deposits;
    out->cells = h.cells;
    out->candidates = h.candidates;
    out->tests = h.tests;
    out->facets = h.facets;
    out->edge_rays = h.edge;
    return QMC_OK;
}

// ---------------------------------------------------------------------------------------------
// Flux maps
// ---------------------------------------------------------------------------------------------
extern "C" qmc_status qmc_flux_clear(qmc_ctx *c, int32_t r) {
    if (!c) return fail(QMC_E_INVALID_ARG, "qmc_flux_clear: null ctx");
    if (r < 0 || r >= (int)c->recvs.size()) return fail(QMC_E_RANGE, "qmc_flux_clear: receiver out of range");
    qmc_status st = set_device(c);
    if (st) return st;
    const RecvBuf &rb = c->recvs[r];
    HIP_TRY(hipMemsetAsync(rb.flux, 0, sizeof(float) * (size_t)rb.w * rb.h, c->stream));
    return QMC_OK;
}

extern "C" qmc_status qmc_flux_read(qmc_ctx *c, int32_t r, float *out, int64_t n) {
    if (!c || !out) return fail(QMC_E_INVALID_ARG, "qmc_flux_read: null argument");
    if (r < 0 || r >= (int)c->recvs.size()) return fail(QMC_E_RANGE, "qmc_flux_read: receiver out of range");
    const RecvBuf &rb = c->recvs[r];
    if (n != (int64_t)rb.w * rb.h) return fail(QMC_E_INVALID_ARG, "qmc_flux_read: size mismatch");
    qmc_status st = set_device(c);
    if (st) return st;
    HIP_TRY(hipMemcpyAsync(out, rb.flux, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return QMC_OK;
}

extern "C" qmc_status qmc_flux_write(qmc_ctx *c, int32_t r, const float *in, int64_t n) {
    if (!c || !in) return fail(QMC_E_INVALID_ARG, "qmc_flux_write: null argument");
    if (r < 0 || r >= (int)c->recvs.size()) return fail(QMC_E_RANGE, "qmc_flux_write: receiver out of range");
    const RecvBuf &rb = c->recvs[r];
    if (n != (int64_t)rb.w * rb.h) return fail(QMC_E_INVALID_ARG, "qmc_flux_write: size mismatch");
    qmc_status st = set_device(c);
    if (st) return st;
    HIP_TRY(hipMemcpyAsync(rb.flux, in, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return QMC_OK;
}

extern "C" qmc_status qmc_flux_device_ptr(qmc_ctx *c, int32_t r, float **p, int32_t *w, int32_t *h) {
    if (!c || !p) return fail(QMC_E_INVALID_ARG, "qmc_flux_device_ptr: null argument");
    if (r < 0 || r >= (int)c->recvs.size()) return fail(QMC_E_RANGE, "qmc_flux_device_ptr: receiver out of range");
    *p = c->recvs[r].flux;
    if (w) *w = c->recvs[r].w;
    if (h) *h = c->recvs[r].h;
    return QMC_OK;
}

// ---------------------------------------------------------------------------------------------
// Timing
// ---------------------------------------------------------------------------------------------
extern "C" qmc_status qmc_event_record(qmc_ctx *c, int32_t slot) {
    if (!c || slot < 0 || slot >= 16) return fail(QMC_E_INVALID_ARG, "qmc_event_record: bad slot");
    HIP_TRY(hipEventRecord(c->ev[slot], c->stream));
    return QMC_OK;
}

extern "C" qmc_status qmc_kernel_events(qmc_ctx *c, int32_t before, int32_t after) {
    if (!c) return fail(QMC_E_INVALID_ARG, "qmc_kernel_events: null ctx");
    const bool off = before == -1 && after == -1;
    if (!off && (before < 0 || before >= 16 || after < 0 || after >= 16 || before == after))
        return fail(QMC_E_INVALID_ARG, "qmc_kernel_events: slots must be two distinct values in 0..15, or -1, -1");
    c->kev_before = before;
    c->kev_after = after;
    return QMC_OK;
}

extern "C" qmc_status qmc_event_elapsed_ms(qmc_ctx *c, int32_t a, int32_t b, float *ms) {
    if (!c || !ms || a < 0 || a >= 16 || b < 0 || b >= 16) return fail(QMC_E_INVALID_ARG, "qmc_event_elapsed_ms: bad slot");
    HIP_TRY(hipEventSynchronize(c->ev[b]));
    HIP_TRY(hipEventElapsedTime(ms, c->ev[a], c->ev[b]));
    return QMC_OK;
}
