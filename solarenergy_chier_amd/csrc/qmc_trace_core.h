// qmc_trace_core.h — per-ray device functions shared by the plain and the batched trace kernels, so
// both follow exactly the same arithmetic (and so stay bit-identical to the host oracle):
//   fr_apply         local2world_rotate with a precomputed frame (Util/global_function.cuh:19-36)
//   dda_walk         rectGridDDA::collision's cell walk (QuasiMonteCarloRayTracing/Grid/rectGridDDA.cu:31-119)
//   dda_collision    the walk with rectGridDDA::intersect on every candidate (rectGridDDA.cu:7-25)
//   facet_geometry   micro-facet origin / normal in registers (RectangleHelio.cu:54-80,
//                    FocusFlatRectangleHelio.cu:12-64)
//   receiver_hit     RectangleReceiver / CylinderReceiver GIntersect
#pragma once

#include <hip/hip_runtime.h>

#include "qmc_device.h"

namespace qmc {

__device__ __forceinline__ V3 fr_apply(const DevFrame &f, V3 d) {
    if (f.identity) return d;
    return V3{d.x * f.ux + d.y * f.nx + d.z * f.vx,
              d.x * f.uy + d.y * f.ny + d.z * f.vy,
              d.x * f.uz + d.y * f.nz + d.z * f.vz};
}

struct RayStats {
    unsigned cells, candidates, tests;
};

// A value kept in a vector register from here on (an empty asm with a "v" operand): for wave-uniform
// constants that a loop reads every iteration, when the loop's scalar registers would otherwise run
// out (the compiler then reloads kernel arguments inside the loop and spills to vector-register lanes).
template <class T>
__device__ __forceinline__ void in_vgpr(T &x) {
    asm("" : "+v"(x));
}
// A frame's nine components, likewise.
__device__ __forceinline__ void vgpr_frame(DevFrame &f) {
    in_vgpr(f.ux), in_vgpr(f.uy), in_vgpr(f.uz);
    in_vgpr(f.nx), in_vgpr(f.ny), in_vgpr(f.nz);
    in_vgpr(f.vx), in_vgpr(f.vy), in_vgpr(f.vz);
}

// Walk of the 3D-DDA over the uniform grid: cells in the reference's order (rectGridDDA.cu:31-119).
// cell(addr, ix, iy, iz, sx, sy, sz, hit_id, st) is called for every visited cell (address in x-major
// order, indices, step signs): it returns 1 at a hit (hit_id set), -1 when no later cell can hold one
// (the walk ends without a hit) and 0 to go on.
template <class Cell>
__device__ __forceinline__ bool dda_walk(V3 o, V3 d, const DevGrid &g, int &hit_id, RayStats &st, Cell cell) {
    const V3 gpos{g.pos[0], g.pos[1], g.pos[2]};
    const V3 gint{g.interval[0], g.interval[1], g.interval[2]};
    const V3 cellf = (o - gpos) / gint;
    int ix = (int)cellf.x, iy = (int)cellf.y, iz = (int)cellf.z;
    const int sx = (d.x >= 0) ? 1 : -1;
    const int sy = (d.y >= 0) ? 1 : -1;
    const int sz = (d.z >= 0) ? 1 : -1;
    float tx = abs_divide(cal_tmax(d.x, gint.x, ix, o.x - gpos.x), d.x);
    float ty = abs_divide(cal_tmax(d.y, gint.y, iy, o.y - gpos.y), d.y);
    float tz = abs_divide(cal_tmax(d.z, gint.z, iz, o.z - gpos.z), d.z);
    const float dx = abs_divide(gint.x, d.x);
    const float dy = abs_divide(gint.y, d.y);
    const float dz = abs_divide(gint.z, d.z);
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    // cell address kept incrementally (x-major, then y, then z); one unsigned compare per axis
    // covers both ends of the grid
    const int stride_x = sx * ny * nz, stride_y = sy * nz;
    int addr = ix * ny * nz + iy * nz + iz;
    while (((unsigned)ix < (unsigned)nx) & ((unsigned)iy < (unsigned)ny) & ((unsigned)iz < (unsigned)nz)) {
        const int r = cell(addr, ix, iy, iz, sx, sy, sz, hit_id, st);
        if (r) return r > 0;
        // Step rule of rectGridDDA (tie tolerance kEps, compared in double): if tx is clearly below
        // ty, step x and/or z by comparing tx with tz; else if ty is clearly below tx, step y and/or
        // z by comparing ty with tz; else step x and y together and/or z by comparing tx with tz.
        // Branch-free form of dda_less/dda_equal on the same double values: each bound t -/+ kEps
        // is formed once and every comparison is made once.
        const double dtx = tx, dty = ty, dtz = tz;
        const double xm = dtx - kEps, ym = dty - kEps, zm = dtz - kEps, zp = dtz + kEps;
        const bool ltxy = dtx < ym;               // dda_less(tx, ty)
        const bool use_y = !ltxy && dty < xm;     // dda_less(ty, tx) in the else branch
        const double ta = use_y ? dty : dtx, tam = use_y ? ym : xm;
        const bool eq_az = ta < zp && ta > zm;    // dda_equal(ta, tz)
        const bool step_a = ta < zm || eq_az;     // dda_less(ta, tz) || dda_equal(ta, tz)
        const bool fz = dtz < tam || eq_az;       // dda_less(tz, ta) || dda_equal(ta, tz)
        const bool fx = step_a && !use_y;
        const bool fy = step_a && !ltxy;
        ix += fx ? sx : 0;
        iy += fy ? sy : 0;
        iz += fz ? sz : 0;
        addr += (fx ? stride_x : 0) + (fy ? stride_y : 0) + (fz ? sz : 0);
        tx += fx ? dx : 0.0f;
        ty += fy ? dy : 0.0f;
        tz += fz ? dz : 0.0f;
    }
    return false;
}

// The candidates of one cell in list order (the first three from the packed record, the rest from
// the cell list), own sub-heliostats skipped; test(id) says whether the ray hits quad id.
template <bool STATS, class Test>
__device__ __forceinline__ int cell_candidates(const DevGrid &g, int addr, int own_lo, int own_hi, int &hit_id,
                                               RayStats &st, Test test) {
    const int4 rec = g.cells[addr];  // count, then the cell's first three candidates
    const int n = rec.x;
    if (STATS) {
        st.cells += 1;
        st.candidates += (unsigned)n;
    }
    for (int i = 0; i < n; ++i) {
        const int id = i == 0 ? rec.y : i == 1 ? rec.z : i == 2 ? rec.w : g.match[g.cell_start[addr] + i];
        if (id < own_lo || id >= own_hi) {
            if (STATS) st.tests += 1;
            if (test(id)) {
                hit_id = id;
                return 1;
            }
        }
    }
    return 0;
}

// 3D-DDA over the uniform grid; true when a non-own sub-heliostat quad is hit (first hit in
// traversal and cell-list order; its id goes to hit_id).
template <bool STATS>
__device__ __forceinline__ bool dda_collision(V3 o, V3 d, const DevGrid &g, const float4 *quads, int own_lo,
                                              int own_hi, int &hit_id, RayStats &st) {
    return dda_walk(o, d, g, hit_id, st, [&](int addr, int, int, int, int, int, int, int &hid, RayStats &s) {
        return cell_candidates<STATS>(g, addr, own_lo, own_hi, hid, s, [&](int id) {
            const float4 qa = quads[3 * id];
            const float4 qb = quads[3 * id + 1];
            const float4 qc = quads[3 * id + 2];
            float t, u, v;
            return ray_quad(o, d, V3{qa.x, qa.y, qa.z}, V3{qb.x, qb.y, qb.z}, V3{qc.x, qc.y, qc.z}, t, u, v);
        });
    });
}

// Origin of micro-facet m of a flat heliostat whose normal frame is fr (RectangleHelio.cu:64-79):
// local centre (col*pl + block_col*gap.x + pl/2 - sx/2, sy/2, row*pl + block_row*gap.y + pl/2 - sz/2)
// rotated into the frame and translated to the heliostat position.
__device__ __forceinline__ V3 flat_facet_origin(const DevHelio &H, const DevFrame &fr, int m) {
    const int cpr = H.cols * H.sub_cols;
    const int row = (int)fast_div((uint32_t)m, H.fd_cpr);
    const int col = m - row * cpr;
    const int brow = (int)fast_div((uint32_t)row, H.fd_sub_rows);
    const int bcol = (int)fast_div((uint32_t)col, H.fd_sub_cols);
    const float pl = H.pl;
    V3 local{col * pl + bcol * H.gap[0] + pl / 2 - H.size[0] / 2, H.size[1] / 2,
             row * pl + brow * H.gap[1] + pl / 2 - H.size[2] / 2};
    return fr_apply(fr, local) + V3{H.pos[0], H.pos[1], H.pos[2]};
}

// Wave-tiled ray order of a flat heliostat with R x C micro-facets (row-major, m = row*C + col) and
// L rays per facet, L in {1, 2, 4, 8}: slot t of a wave of 64 takes ray j of facet (row, col) inside
// a 2^ltw x 2^lth tile of facets (8x8, 8x4, 4x4 or 4x2; tile area * L = 64), tiles row-major over
// the padded facet grid. A wave's ray origins then span a compact patch instead of a long facet row,
// so its lanes walk similar grid cells. Slots past the grid's edge hold no ray. Other L keep facet
// order (with L >= 16 a wave holds at most four facets of one row already).
struct FacetTiling {
    int ltw, lth, lL, ntx, slots;
    bool tiled;
};

__host__ __device__ inline FacetTiling facet_tiling(int R, int C, int L) {
    FacetTiling t;
    t.lL = L == 1 ? 0 : L == 2 ? 1 : L == 4 ? 2 : L == 8 ? 3 : -1;
    t.tiled = t.lL >= 0;
    if (!t.tiled) {
        t.ltw = t.lth = t.ntx = 0;
        t.slots = R * C * L;
        return t;
    }
    const int lf = 6 - t.lL;  // log2 of facets per wave
    t.ltw = (lf + 1) / 2;
    t.lth = lf - t.ltw;
    t.ntx = (C + (1 << t.ltw) - 1) >> t.ltw;
    const int nty = (R + (1 << t.lth) - 1) >> t.lth;
    t.slots = t.ntx * nty * 64;
    return t;
}

// Ray id of slot t, or -1 for a slot outside the facet grid; fd_ntx divides by tl.ntx.
__host__ __device__ inline int facet_tiled_ray(const FacetTiling &tl, FastDiv fd_ntx, int R, int C, int t) {
    if (!tl.tiled) return t;
    const int tile = t >> 6, lane = t & 63;
    const int f = lane >> tl.lL, j = lane & ((1 << tl.lL) - 1);
    const int ty = (int)fast_div((uint32_t)tile, fd_ntx), tx = tile - ty * tl.ntx;
    const int col = (tx << tl.ltw) + (f & ((1 << tl.ltw) - 1));
    const int row = (ty << tl.lth) + (f >> tl.ltw);
    if (col >= C || row >= R) return -1;
    return ((row * C + col) << tl.lL) + j;
}

// The multiply-high divisor of a heliostat for tiles of 2^ltw facets across.
__device__ __forceinline__ FastDiv tile_divisor(const DevHelio &H, const FacetTiling &tl) {
    return tl.ltw == 3 ? H.fd_ntx8 : H.fd_ntx4;
}

// Ray slot -> ray id of heliostat H: flat heliostats take their rays in facet tiles (a wave's origins
// form a compact patch); -1 for slots past the heliostat's rays. HT: the heliostat type when the
// caller knows it at compile time (every heliostat of a grid has the grid's type), -1 to read it.
template <int HT = -1>
__device__ __forceinline__ long long slot_ray(const DevHelio &H, int L, long long slot) {
    if ((HT < 0 ? H.type : HT) == 0) {
        const int R = H.rows * H.sub_rows, C = H.cols * H.sub_cols;
        const FacetTiling tl = facet_tiling(R, C, L);
        return slot < tl.slots ? facet_tiled_ray(tl, tile_divisor(H, tl), R, C, (int)slot) : -1;
    }
    return slot < (long long)H.F * L ? slot : -1;
}

// Ray slots of a heliostat: its rays, or the tiled ray slots of a flat heliostat (slot_ray).
__host__ __device__ inline long long heliostat_slot_count(const DevHelio &H, int L) {
    return H.type == 0 ? (long long)facet_tiling(H.rows * H.sub_rows, H.cols * H.sub_cols, L).slots
                       : (long long)H.F * L;
}

// Culling shape of a flat heliostat with normal frame f: its micro-facet origins lie in the rectangle
// of half-extents size.x/2, size.z/2 at height size.y/2 above the position, and every facet normal is
// the heliostat normal.
__host__ __device__ inline DevCullShape flat_cull_shape(const DevHelio &H, const DevFrame &f) {
    const CD3 u = f.identity ? CD3{1, 0, 0} : CD3{f.ux, f.uy, f.uz};
    const CD3 n = f.identity ? CD3{0, 1, 0} : CD3{f.nx, f.ny, f.nz};
    const CD3 v = f.identity ? CD3{0, 0, 1} : CD3{f.vx, f.vy, f.vz};
    const CD3 c = cadd(CD3{H.pos[0], H.pos[1], H.pos[2]}, cmul(n, 0.5 * H.size[1]));
    const CD3 hu = cmul(u, 0.5 * H.size[0]), hv = cmul(v, 0.5 * H.size[2]);
    DevCullShape s;
    s.c[0] = c.x; s.c[1] = c.y; s.c[2] = c.z;
    s.hu[0] = hu.x; s.hu[1] = hu.y; s.hu[2] = hu.z;
    s.hn[0] = 0.0; s.hn[1] = 0.0; s.hn[2] = 0.0;
    s.hv[0] = hv.x; s.hv[1] = hv.y; s.hv[2] = hv.z;
    s.n[0] = n.x; s.n[1] = n.y; s.n[2] = n.z;
    s.cant = 0.0;
    return s;
}

// Lane src's value of x, for every lane (ds_bpermute; every lane of the wave must be active).
__device__ __forceinline__ int lane_read(int x, int src) { return __builtin_amdgcn_ds_bpermute(src << 2, x); }
__device__ __forceinline__ float lane_read(float x, int src) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(x)));
}

// Load of data that no kernel of the launch writes (per-heliostat and per-instance records) through
// the constant address space: with a wave-uniform address it becomes a scalar load into SGPRs.
template <class T>
__device__ __forceinline__ T ld_uniform(const T *p) {
    static_assert(sizeof(T) % 4 == 0, "records are whole dwords");
    typedef const __attribute__((address_space(4))) uint32_t cdword;
    cdword *src = reinterpret_cast<cdword *>(reinterpret_cast<uintptr_t>(p));
    T out;
    uint32_t *dst = reinterpret_cast<uint32_t *>(&out);
#pragma unroll
    for (size_t i = 0; i < sizeof(T) / 4; ++i) dst[i] = src[i];
    return out;
}

// ld_uniform when U (the address is wave-uniform), else an ordinary load.
template <bool U, class T>
__device__ __forceinline__ T ld_maybe(const T *p) {
    if constexpr (U) return ld_uniform(p);
    else return *p;
}

// Whether a traversal has to run for this ray, given its heliostat's candidate list for that
// traversal: not when no quad can be met (count 0) or when the ray misses every listed candidate —
// the traversal tests only quads, the others provably miss, and a listed one gives the same
// ray_quad result here as there. Otherwise (a candidate is hit, or too many to list) the traversal
// runs and decides, with the reference's first-hit order.
// The list is wave-uniform and written by an earlier launch: its fields are read with scalar loads.
__device__ __forceinline__ bool traversal_needed(const CullList *cl, int kind, V3 o, V3 d, const float4 *quads) {
    const int n = ld_uniform(&cl->n[kind]);
    if (n == 0) return false;
    if (n > kCullList) return true;
    for (int i = 0; i < n; ++i) {
        const int id = ld_uniform(&cl->ids[kind][i]);
        // the candidates are wave-uniform and no kernel writes the quads: scalar loads
        const float4 qa = ld_uniform(quads + 3 * id), qb = ld_uniform(quads + 3 * id + 1),
                     qc = ld_uniform(quads + 3 * id + 2);
        float t, u, v;
        if (ray_quad(o, d, V3{qa.x, qa.y, qa.z}, V3{qb.x, qb.y, qb.z}, V3{qc.x, qc.y, qc.z}, t, u, v)) return true;
    }
    return false;
}

// A ray's tests of its wave's masked candidates (traversal_plan): how many it hits, the first and (with
// two) the second in mask order, with the ray's parameter at each, and — with at most two hits — the
// union of the cell ranges whose lists hold them (x0..x1, y0..y1, z0..z1).
struct RayHits {
    int n;
    int id0, id1;
    float t0, t1;  // the ray's parameter at each hit
    int x0, y0, z0, x1, y1, z1;
};

// How the plain trace runs a ray's traversal of kind `kind`, given its heliostat's wide list and the
// mask of its wave (wm: the wave's two masks, from k_cull_waves): every quad outside the mask is out
// of reach of every ray of the wave, so ray_quad misses it. Returns
//   0 — no traversal: no candidate, or the ray misses every masked one, so no quad can be hit;
//   1 — the ray hits rh.n >= 1 masked candidates (rh): the traversal reports a hit. With at most two,
//       since every quad's ray_quad result is known (a hit for those, a miss for all others), the
//       traversal's cell walk only needs the cells listing them (dda_known), and its first hit is the
//       reference traversal's; with more, the first hit needs the traversal's own tests;
//   2 — the traversal runs with its own tests (too many candidates to list, or, unless ALL, more than
//       kRayTests in the wave's mask).
// ALL: every masked candidate is tested, however many. Records are read with scalar loads when U (the
// wave's lanes share the heliostat and the wave mask), with vector loads otherwise.
template <bool U = true, bool ALL = false>
__device__ __forceinline__ int traversal_plan(const WideList *wl, const uint64_t *wm, int kind, V3 o, V3 d,
                                              const DevGrid &g, RayHits &rh) {
    const int n = ld_maybe<U>(&wl->n[kind]);
    if (n == 0) return 0;
    if (n > kCullWide) return 2;
    uint64_t m = ld_maybe<U>(wm + kind);
    if (!ALL && __builtin_popcountll(m) > kRayTests) return 2;
    rh.n = 0;
    rh.id0 = rh.id1 = -1;
    auto test = [&](int i) {
        const int id = ld_maybe<U>(&wl->ids[kind][i]);
        const float4 qa = ld_maybe<U>(g.quads + 3 * id), qb = ld_maybe<U>(g.quads + 3 * id + 1),
                     qc = ld_maybe<U>(g.quads + 3 * id + 2);
        float t, u, v;
        if (ray_quad(o, d, V3{qa.x, qa.y, qa.z}, V3{qb.x, qb.y, qb.z}, V3{qc.x, qc.y, qc.z}, t, u, v)) {
            // id1 ends as the last hit, the second one when there are two (with more it is not read)
            if (rh.n == 0) rh.id0 = id, rh.t0 = t;
            else rh.id1 = id, rh.t1 = t;
            ++rh.n;
        }
    };
    if (ALL) {
        while (m) {
            const int i = __builtin_ctzll(m);
            m &= m - 1;
            test(i);
        }
    } else {
#pragma unroll
        for (int j = 0; j < kRayTests; ++j) {
            if (!m) break;
            const int i = __builtin_ctzll(m);
            m &= m - 1;
            test(i);
        }
    }
    if (rh.n == 0) return 0;
    // the cells that list the hit quads (grid_match's ranges; the union when there are two)
    const int4 lo0 = g.boxes[2 * rh.id0], hi0 = g.boxes[2 * rh.id0 + 1];
    rh.x0 = lo0.x, rh.y0 = lo0.y, rh.z0 = lo0.z, rh.x1 = hi0.x, rh.y1 = hi0.y, rh.z1 = hi0.z;
    if (rh.n == 2) {
        const int4 lo1 = g.boxes[2 * rh.id1], hi1 = g.boxes[2 * rh.id1 + 1];
        rh.x0 = min(rh.x0, lo1.x), rh.y0 = min(rh.y0, lo1.y), rh.z0 = min(rh.z0, lo1.z);
        rh.x1 = max(rh.x1, hi1.x), rh.y1 = max(rh.y1, hi1.y), rh.z1 = max(rh.z1, hi1.z);
    }
    return 1;
}

// The traversal of a ray with plan 1 and at most two hits (traversal_plan): the reference's cell walk
// in its order, where only cells inside the hit quads' range can list a quad the ray hits. There the
// cell's candidates are checked in list order against the hit ids; elsewhere the cell is stepped over
// without reading it; once the walk has passed the range on any axis (indices only grow along the ray)
// no later cell is in it and the walk ends without a hit. Its first hit is the reference traversal's.
template <bool STATS>
__device__ __forceinline__ bool dda_known(const RayHits &rh, V3 o, V3 d, const DevGrid &g, int own_lo, int own_hi,
                                          int &hit_id, RayStats &st) {
    if ((rh.x0 > rh.x1) | (rh.y0 > rh.y1) | (rh.z0 > rh.z1)) return false;  // listed in no cell
    return dda_walk(o, d, g, hit_id, st,
                    [&](int addr, int ix, int iy, int iz, int sx, int sy, int sz, int &hid, RayStats &s) {
                        const bool inside = (ix >= rh.x0) & (ix <= rh.x1) & (iy >= rh.y0) & (iy <= rh.y1) &
                                            (iz >= rh.z0) & (iz <= rh.z1);
                        if (inside)
                            return cell_candidates<STATS>(g, addr, own_lo, own_hi, hid, s,
                                                          [&](int id) { return (id == rh.id0) | (id == rh.id1); });
                        const bool past = (sx > 0 ? ix > rh.x1 : ix < rh.x0) | (sy > 0 ? iy > rh.y1 : iy < rh.y0) |
                                          (sz > 0 ? iz > rh.z1 : iz < rh.z0);
                        return past ? -1 : 0;
                    });
}

// The traversal's outcome for a ray with plan 1 without its walk, where that can be proven: the walk
// would visit a cell of the cell box of the first quad the ray's tests hit (dda_visit_proof at that
// hit; every cell of the box lists the quad), so the traversal reports a hit — that quad's or an
// earlier one. True with hit_id = that quad (not necessarily the traversal's first hit: for callers
// that need only whether a hit exists); false when the proof does not hold (the caller walks).
// Straight-line code: the frame is read with vector loads (its address moved to vector registers
// behind an empty barrier), so that a persistent caller keeps none of it in scalar registers across
// its loop.
__device__ __forceinline__ bool known_hit_proven(const RayHits &rh, V3 o, V3 d, const DevGrid &g, int &hit_id) {
    const ProofFrame *pf = g.frame;
    asm volatile("" : "+v"(pf));
    const ProofFrame f = *pf;
    const float4 m = g.pbox[2 * rh.id0], l = g.pbox[2 * rh.id0 + 1];
    const float mid[3] = {m.x, m.y, m.z}, lim[3] = {l.x, l.y, l.z};
    hit_id = rh.id0;
    return dda_visit_proof(o, d, rh.t0, f, mid, lim);
}

// Half-angles (radians) at or above which a cone is not culled: the bound needs theta < 90 deg.
constexpr double kCullMaxAngle = 80.0 * 3.14159265358979323846 / 180.0;

// Margin (metres) of the culling bound: far above the float rounding of traced origins, directions
// and the intersection test.
constexpr double kCullMargin = 0.01;

// Origin of micro-facet m of a focus-flat heliostat (FocusFlatRectangleHelio.cu:21-39,53-64): its
// centre in its canted facet's local frame, placed by the facet's frame and centre, then by the
// heliostat's focus-variant frame and position. fidx receives the canted facet's DevFacet index.
__device__ __forceinline__ V3 focus_facet_origin(const DevHelio &H, const DevFacet *facets, int m, int &fidx) {
    const int cpr = H.cols * H.sub_cols;
    int row = (int)fast_div((uint32_t)m, H.fd_cpr);
    int col = m - row * cpr;
    const int brow = (int)fast_div((uint32_t)row, H.fd_sub_rows), bcol = (int)fast_div((uint32_t)col, H.fd_sub_cols);
    fidx = H.facet_base + brow * H.cols + bcol;
    const DevFacet &Fc = facets[fidx];
    row -= brow * H.sub_rows;
    col -= bcol * H.sub_cols;
    const float pl = H.pl;
    V3 local{(col + 0.5f) * pl - H.len_col / 2, 0.0f, (row + 0.5f) * pl - H.len_row / 2};
    local = fr_apply(Fc.fr_local, local) + V3{Fc.center[0], Fc.center[1], Fc.center[2]};
    return fr_apply(H.fr_focus, local) + V3{H.pos[0], H.pos[1], H.pos[2]};
}

// Micro-facet m of heliostat H: world origin and the frame that rotates slope-error samples
// onto the facet normal (returned by value: H may be a copy in registers).
// Facet indices use the heliostat's multiply-high divisors (exact for m < 2^31).
// HT as for slot_ray.
template <int HT = -1>
__device__ __forceinline__ void facet_geometry(const DevHelio &H, const DevFacet *facets, int m, V3 &origin,
                                               DevFrame &pert_frame, V3 &normal) {
    if ((HT < 0 ? H.type : HT) == 0) {
        origin = flat_facet_origin(H, H.fr, m);
        pert_frame = H.fr;
        normal = V3{H.fr.nx, H.fr.ny, H.fr.nz};
    } else {
        int fidx;
        origin = focus_facet_origin(H, facets, m, fidx);
        const DevFacet &Fc = facets[fidx];
        pert_frame = Fc.fr_pert;
        normal = V3{Fc.normal[0], Fc.normal[1], Fc.normal[2]};
    }
}

// Receiver hit: true with (t, u, v) when the ray reaches the receiver surface. For a rectangle
// receiver, ot (nullable) holds the ray origin's terms of the test (quad_origin_terms with the
// receiver's A, E1, E2), computed once for the rays of a micro-facet. For a cylinder with column
// breakpoints and col non-null, *col receives the hit's map column, (int)(v * res_x) without the
// clamp, in place of v.
template <int RT>
__device__ __forceinline__ bool receiver_hit(const DevReceiver &R, V3 o, V3 d, float &t, float &u, float &v,
                                             const QuadOriginTerms *ot = nullptr, int *col = nullptr) {
    if (RT == 0) {
        // RectangleReceiver::GIntersect (RectangleReceiver.cuh:32-34)
        if (ot) return ray_quad_pre(d, V3{R.E1[0], R.E1[1], R.E1[2]}, V3{R.E2[0], R.E2[1], R.E2[2]}, *ot, t, u, v);
        return ray_quad(o, d, V3{R.A[0], R.A[1], R.A[2]}, V3{R.E1[0], R.E1[1], R.E1[2]},
                        V3{R.E2[0], R.E2[1], R.E2[2]}, t, u, v);
    } else {
        // CylinderReceiver::GIntersect (CylinderReceiver.cuh:39-89)
        const float R2 = R.radius * R.radius;
        V2 l{o.x - R.pos[0], o.z - R.pos[2]};
        if (dot2(l, l) <= R2) return false;
        V2 Ro{R.pos[0] - o.x, R.pos[2] - o.z};
        float tp = dot2(Ro, normalize2(V2{d.x, d.z}));
        float delta = dot2(Ro, Ro) - tp * tp;
        if (tp < -kEpsF || delta > R2) return false;  // tp < -Epsilon
        float t_plus = delta <= 0.0f ? R.radius : sqrtf(R2 - delta);
        t = (tp - t_plus) / length2(V2{d.x, d.z});
        V3 ip = t * d + o;
        u = (ip.y - R.pos[1]) / R.height + 0.5f;
        if (u < 0.0f || u > 1.0f) return false;
        V2 c = normalize2(V2{ip.x - R.pos[0], ip.z - R.pos[2]});
        float cx = fminf(fmaxf(c.x, -1.0f), 1.0f);
        if (col && R.colt) {
            // the map column directly, from the receiver's breakpoints (cylinder_column); v is not read
            *col = cylinder_column(R.colt, R.res_x, cx, c.y < 0);
            v = 0.0f;
            return true;
        }
        // float(double(a) / (2 pi)) computed as float(double(a) * (1 / (2 pi))): equal for every float a
        // in [0, float(pi)], the range of acos_r (checked exhaustively, tests/test_float_compares.py)
        v = (float)((double)acos_r(cx) * kInv2Pi);
        if (c.y < 0) v = 1 - v;
        return true;
    }
}

}  // namespace qmc
