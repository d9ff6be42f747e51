// qmc_cull.hip — traversal culling of the plain trace (no counterpart in the reference; it only
// skips work that cannot change a ray's result):
//   k_cull         per listed heliostat, the quads its shadow and blocking rays may meet (CullList, and
//                  up to kCullWide for the plain trace: WideList), scanning the grid's quad groups
//   k_cull_waves   per wave of 64 ray slots, which of those candidates that wave's rays can reach
//                  (a 64-bit mask per traversal), bounded from the wave's own micro-facet origins
//   upload_groups  the grid's quad groups (one bounding ball per heliostat's quads) for the scan
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "qmc_ctx.h"
#include "qmc_trace_core.h"

using namespace qmc;

namespace {

constexpr int kCullBS = 256;            // lanes of a k_cull workgroup
constexpr int kScanRound = 4 * kCullBS; // quad groups per round of the scan

// The traversal-culling scan of one heliostat (culling shape S, own quads [own_lo, own_hi)) over a
// grid's quad groups, by the whole workgroup: every quad that passes the bound (sphere_may_hit, then
// cone_may_hit) for the shadow cone (kind 0: tan_sun around the sun axis, the n axis of sunf) or the
// blocking cone (kind 1: tan_block around the sun axis reflected in the heliostat normal) is counted in
// n[kind] and, while the count is at most cap (<= kCullWide), listed in ids[kind] (in no particular
// order); a negative tangent skips that kind. Each round takes kScanRound groups: one lane per group
// tests the group's ball (sphere_may_hit is monotone in the ball: a ball inside one that fails fails
// too, so no quad that would pass is skipped); the passing groups' quads are then spread over all
// lanes, one (group, quad) pair each, so that a group near the heliostat does not leave one lane
// testing its quads while the others idle. A kind stops being scanned once its count exceeds cap (it
// is then traversed for every ray and its list is not used), so a final count <= cap is complete.
// n and ids are in LDS, n zeroed beforehand; every lane of the workgroup calls this.
__device__ void cull_scan(const DevCullShape &S, const DevFrame &sunf, const float4 *quads, const QuadGroup *groups,
                          int n_groups, int own_lo, int own_hi, double tan_sun, double tan_block, double margin,
                          int cap, int *n, int (*ids)[kCullWide]) {
    __shared__ int s_sel[kScanRound];  // passing groups of the round: index, kinds in bits 29 and 30
    __shared__ int s_end[kScanRound];  // inclusive prefix sums of their quad counts
    __shared__ int s_cnt;
    __shared__ int s_wsum[kCullBS / 64];
    constexpr int kIdx = (1 << 29) - 1, kS = 1 << 29, kB = 1 << 30;
    const CD3 c{S.c[0], S.c[1], S.c[2]}, hu{S.hu[0], S.hu[1], S.hu[2]}, hn{S.hn[0], S.hn[1], S.hn[2]},
        hv{S.hv[0], S.hv[1], S.hv[2]}, nrm{S.n[0], S.n[1], S.n[2]};
    const CD3 ash = sunf.identity ? CD3{0, 1, 0} : CD3{sunf.nx, sunf.ny, sunf.nz};
    CD3 r0 = cadd(cmul(ash, -1.0), cmul(nrm, 2.0 * cdot(ash, nrm)));  // reflect(-sun axis, n)
    r0 = cmul(r0, 1.0 / sqrt(cdot(r0, r0)));
    const double rs = sqrt(cdot(hu, hu) + cdot(hn, hn) + cdot(hv, hv));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int g0 = 0; g0 < n_groups; g0 += kScanRound) {
        // the counts were last written before a barrier: every lane reads the same values
        const bool want_s = tan_sun >= 0.0 && n[0] <= cap;
        const bool want_b = tan_block >= 0.0 && n[1] <= cap;
        if (!want_s && !want_b) break;
        if (threadIdx.x == 0) s_cnt = 0;
        __syncthreads();
        const int g1 = min(n_groups, g0 + kScanRound);
        for (int gi = g0 + (int)threadIdx.x; gi < g1; gi += kCullBS) {
            const QuadGroup G = groups[gi];
            if (G.begin >= own_lo && G.begin + G.count <= own_hi) continue;
            const CD3 Dg = csub(CD3{G.c[0], G.c[1], G.c[2]}, c);
            const bool gs = want_s && sphere_may_hit(Dg, rs + G.r, ash, tan_sun, margin);
            const bool gb = want_b && sphere_may_hit(Dg, rs + G.r, r0, tan_block, margin);
            if (gs || gb) s_sel[atomicAdd(&s_cnt, 1)] = gi | (gs ? kS : 0) | (gb ? kB : 0);
        }
        __syncthreads();
        const int cnt = s_cnt;
        // prefix sums of the selected groups' quad counts: four entries per lane, then across lanes
        int loc[4], run = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int e = 4 * (int)threadIdx.x + k;
            run += e < cnt ? groups[s_sel[e] & kIdx].count : 0;
            loc[k] = run;
        }
        int incl = run;
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        if (lane == 63) s_wsum[wave] = incl;
        __syncthreads();
        int base = incl - run;
        for (int w = 0; w < wave; ++w) base += s_wsum[w];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int e = 4 * (int)threadIdx.x + k;
            if (e < cnt) s_end[e] = base + loc[k];
        }
        __syncthreads();
        const int total = cnt ? s_end[cnt - 1] : 0;
        for (int t = (int)threadIdx.x; t < total; t += kCullBS) {
            // the selected group holding pair t: the first entry whose prefix sum exceeds t
            int lo = 0, hi = cnt - 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (s_end[mid] > t) hi = mid;
                else lo = mid + 1;
            }
            const int sel = s_sel[lo];
            const QuadGroup &G = groups[sel & kIdx];
            const int q = G.begin + t - (s_end[lo] - G.count);
            if (q >= own_lo && q < own_hi) continue;
            const float4 qa = quads[3 * q], qb = quads[3 * q + 1], qc = quads[3 * q + 2];
            const CD3 A{qa.x, qa.y, qa.z}, E1{qb.x, qb.y, qb.z}, E2{qc.x, qc.y, qc.z};
            const CD3 sum = cadd(E1, E2), diff = csub(E1, E2);
            const CD3 D = csub(cadd(A, cmul(sum, 0.5)), c);
            const double R = rs + 0.5 * sqrt(fmax(cdot(sum, sum), cdot(diff, diff)));
            if ((sel & kS) && sphere_may_hit(D, R, ash, tan_sun, margin) &&
                cone_may_hit(c, hu, hn, hv, ash, tan_sun, A, E1, E2, margin)) {
                const int k = atomicAdd(n, 1);
                if (k < cap) ids[0][k] = q;
            }
            if ((sel & kB) && sphere_may_hit(D, R, r0, tan_block, margin) &&
                cone_may_hit(c, hu, hn, hv, r0, tan_block, A, E1, E2, margin)) {
                const int k = atomicAdd(n + 1, 1);
                if (k < cap) ids[1][k] = q;
            }
        }
        __syncthreads();
    }
}

// Per heliostat of the launch: whether any ray of its shadow cone (directions within the sun-shape
// pool's half-angle of the sun axis) or of its blocking cone (within sun shape + twice the slope
// error of the ideal reflection) can meet another quad of the grid (cone_may_hit, conservative).
// When none can, every test of that traversal is a miss, so skipping it returns the same result for
// every ray. Candidates are listed up to kCullList (CullList) or, when a.wide is set, up to kCullWide
// (WideList, for the per-wave masks of the plain trace).
__global__ __launch_bounds__(kCullBS) void k_cull(CullArgs a) {
    __shared__ int s_n[2];
    __shared__ int s_ids[2][kCullWide];
    const int h = a.list[blockIdx.x];
    const DevHelio &H = a.helios[h];
    const DevCullShape S = a.shapes[h];
    if (a.zero && blockIdx.x == 0 && threadIdx.x == 0) *a.zero = 0u;
    // canted facets turn reflections further from the ideal one: twice the canting angle
    const double ang_block = a.ang_block < 0.0 ? -1.0 : a.ang_block + 2.0 * S.cant;
    const double tan_block = (ang_block >= 0.0 && ang_block < kCullMaxAngle) ? tan(ang_block) : -1.0;
    if (threadIdx.x < 2) s_n[threadIdx.x] = 0;
    __syncthreads();
    cull_scan(S, a.sun, a.quads, a.groups, a.n_groups, H.sub_begin, H.sub_begin + H.sub_count, a.tan_sun, tan_block,
              a.margin, a.wide ? kCullWide : kCullList, s_n, s_ids);
    __syncthreads();
    if (a.wide) {
        WideList *W = a.wide + h;
        const int nw0 = a.tan_sun < 0.0 ? kCullWide + 1 : s_n[0], nw1 = tan_block < 0.0 ? kCullWide + 1 : s_n[1];
        if (threadIdx.x < 2) W->n[threadIdx.x] = threadIdx.x ? nw1 : nw0;
        for (int t = threadIdx.x; t < 2 * kCullWide; t += blockDim.x) {
            const int k = t / kCullWide, i = t - k * kCullWide;
            W->ids[k][i] = i < (k ? nw1 : nw0) ? s_ids[k][i] : -1;
        }
    }
    if (threadIdx.x == 0) {
        CullList L;
        // a cone that cannot be bounded keeps its traversal for every ray
        L.n[0] = a.tan_sun < 0.0 ? kCullList + 1 : s_n[0];
        L.n[1] = tan_block < 0.0 ? kCullList + 1 : s_n[1];
        for (int k = 0; k < 2; ++k)
            for (int i = 0; i < kCullList; ++i) L.ids[k][i] = i < L.n[k] ? s_ids[k][i] : -1;
        a.lists[h] = L;
    }
}

// Axes of a frame in double (u, n, v: where local x, y, z go); the identity frame gives the world axes.
__device__ __forceinline__ void frame_axes(const DevFrame &f, CD3 &u, CD3 &n, CD3 &v) {
    u = f.identity ? CD3{1, 0, 0} : CD3{f.ux, f.uy, f.uz};
    n = f.identity ? CD3{0, 1, 0} : CD3{f.nx, f.ny, f.nz};
    v = f.identity ? CD3{0, 0, 1} : CD3{f.vx, f.vy, f.vz};
}

// A heliostat's listed candidates in double, staged once per workgroup of k_cull_waves: corner A,
// edges E1, E2, centre A + (E1 + E2)/2 and the radius of their bounding ball (half the longer
// diagonal), as k_cull computes them.
struct StagedQuad {
    double A[3], E1[3], E2[3], qc[3], r;
};

// Sets, in m[k], the bit of each wide-list candidate of traversal k (n[k] of them, 0 < n[k] <=
// kCullWide) that rays starting in the parallelogram c +- hu +- hv may meet: shadow rays within the
// sun-shape cone of the sun axis, blocking rays within the sun shape plus twice the slope error of the
// sun axis reflected about nrm. The same two tests as k_cull (sphere_may_hit, then cone_may_hit), with
// the wave's origins in place of the heliostat's.
__device__ void wave_shape_bits(const WaveArgs &a, const StagedQuad (*sq)[kCullWide], const int n[2], CD3 c, CD3 hu,
                                CD3 hv, CD3 nrm, uint64_t m[2]) {
    const CD3 zero{0, 0, 0};
    const CD3 ash = a.sun.identity ? CD3{0, 1, 0} : CD3{a.sun.nx, a.sun.ny, a.sun.nz};
    nrm = cmul(nrm, 1.0 / sqrt(cdot(nrm, nrm)));
    CD3 r0 = cadd(cmul(ash, -1.0), cmul(nrm, 2.0 * cdot(ash, nrm)));  // reflect(-sun axis, nrm)
    r0 = cmul(r0, 1.0 / sqrt(cdot(r0, r0)));
    const double rs = sqrt(cdot(hu, hu) + cdot(hv, hv));
    for (int k = 0; k < 2; ++k) {
        if (n[k] <= 0 || n[k] > kCullWide) continue;
        const CD3 ax = k == 0 ? ash : r0;
        const double tn = k == 0 ? a.tan_sun : a.tan_block;
        for (int i = 0; i < n[k]; ++i) {
            if ((m[k] >> i) & 1ull) continue;
            const StagedQuad &Q = sq[k][i];
            const CD3 D = csub(CD3{Q.qc[0], Q.qc[1], Q.qc[2]}, c);
            if (sphere_may_hit(D, rs + Q.r, ax, tn, a.margin) &&
                cone_may_hit_n<4>(c, hu, zero, hv, ax, tn, CD3{Q.A[0], Q.A[1], Q.A[2]}, CD3{Q.E1[0], Q.E1[1], Q.E1[2]},
                                  CD3{Q.E2[0], Q.E2[1], Q.E2[2]}, a.margin))
                m[k] |= 1ull << i;
        }
    }
}

// Waves of slots per lane of k_cull_waves: each lane bounds the rays of two adjacent waves together and
// gives both the same mask. That halves the lanes (the kernel's cost) for slightly looser masks, which
// leave the trace kernel's time unchanged on a flat field; with four per lane focus-flat runs cross too
// many facets, whose waves then keep every candidate.
constexpr int kWavesPerLane = 2;
static_assert(kWavesPerLane == 1 || kWavesPerLane == 2 || kWavesPerLane == 4, "one, two or four waves per lane");

// One lane per kWavesPerLane waves of slots (grid.y = list entry): the candidate masks of those waves
// (WaveArgs), one mask for all of them. Their rays start at the centres of micro-facets: one tile of the
// facet grid per wave (flat heliostats with tiled ray order, as slot_ray), or the lane's run of
// consecutive facets in row-major order (one row, or the end of one row and the start of the next).
// Every centre of a tile or row piece lies in the parallelogram spanned by its extreme rows and
// columns: on a flat heliostat the centre grows with column and row over the whole heliostat (affine
// within a facet, a gap between facets; RectangleHelio.cu:54-80), on a focus-flat one it is affine
// within each facet (FocusFlatRectangleHelio.cu:21-39), whose own perturbation frame then gives the
// reflection axis (no canting term). Origins are float-rounded; the 1 cm margin of the bound covers
// that. A lane whose run needs more than 8 parallelograms keeps every candidate of its heliostat.
__global__ __launch_bounds__(256) void k_cull_waves(WaveArgs a) {
    static_assert(kCullWide == 64, "one 64-bit mask per wave and traversal");
    __shared__ StagedQuad sq[2][kCullWide];
    const int li = blockIdx.y;
    const int w = (int)(blockIdx.x * 256u + threadIdx.x);
    const int h = ld_uniform(a.list + li);
    const WideList *wl = a.wide + h;
    const int n[2] = {ld_uniform(&wl->n[0]), ld_uniform(&wl->n[1])};
    // the heliostat's candidates, one lane each
    if (threadIdx.x < 2 * kCullWide) {
        const int k = threadIdx.x / kCullWide, i = threadIdx.x - k * kCullWide;
        if (n[k] <= kCullWide && i < n[k]) {
            const int q = wl->ids[k][i];
            const float4 qa = a.quads[3 * q], qb = a.quads[3 * q + 1], qc = a.quads[3 * q + 2];
            const CD3 A{qa.x, qa.y, qa.z}, E1{qb.x, qb.y, qb.z}, E2{qc.x, qc.y, qc.z};
            const CD3 sum = cadd(E1, E2), diff = csub(E1, E2);
            const CD3 ctr = cadd(A, cmul(sum, 0.5));
            StagedQuad &Q = sq[k][i];
            Q.A[0] = A.x; Q.A[1] = A.y; Q.A[2] = A.z;
            Q.E1[0] = E1.x; Q.E1[1] = E1.y; Q.E1[2] = E1.z;
            Q.E2[0] = E2.x; Q.E2[1] = E2.y; Q.E2[2] = E2.z;
            Q.qc[0] = ctr.x; Q.qc[1] = ctr.y; Q.qc[2] = ctr.z;
            Q.r = 0.5 * sqrt(fmax(cdot(sum, sum), cdot(diff, diff)));
        }
    }
    __syncthreads();
    if (w * kWavesPerLane >= a.wstride) return;
    uint64_t m[2] = {0ull, 0ull};
    const bool any = (n[0] > 0 && n[0] <= kCullWide) || (n[1] > 0 && n[1] <= kCullWide);
    const DevHelio &H = a.helios[h];
    if (any) {
        const int R = H.rows * H.sub_rows, C = H.cols * H.sub_cols;
        // slot order as slot_ray: flat heliostats in facet tiles, focus-flat ones in ray order
        const FacetTiling tl = facet_tiling(R, C, a.L);
        constexpr int kRects = kWavesPerLane > 2 ? kWavesPerLane : 2;
        int nr = 0, r0[kRects], r1[kRects], c0[kRects], c1[kRects];
        bool full = false;
        const long long s0 = 64LL * kWavesPerLane * w;
        if (H.type == 0 && tl.tiled) {
            for (int j = 0; j < kWavesPerLane; ++j) {
                const int wj = w * kWavesPerLane + j;
                if (64LL * wj >= tl.slots) break;
                const int ty = (int)fast_div((uint32_t)wj, tile_divisor(H, tl)), tx = wj - ty * tl.ntx;
                const int ra = ty << tl.lth, ca = tx << tl.ltw;
                if (ra < R && ca < C) {
                    r0[nr] = ra;
                    r1[nr] = min(R, ra + (1 << tl.lth)) - 1;
                    c0[nr] = ca;
                    c1[nr] = min(C, ca + (1 << tl.ltw)) - 1;
                    ++nr;
                }
            }
        } else {
            const long long nrays = (long long)H.F * a.L;
            if (s0 < nrays) {
                // slot counts stay below 2^31 (checked at launch): multiply-high quotients are exact
                const long long e = min(s0 + 64LL * kWavesPerLane - 1, nrays - 1);
                const int m0 = (int)fast_div((uint32_t)s0, a.fdL), m1 = (int)fast_div((uint32_t)e, a.fdL);
                const int ra = (int)fast_div((uint32_t)m0, H.fd_cpr), rb = (int)fast_div((uint32_t)m1, H.fd_cpr);
                if (ra == rb) {
                    r0[0] = r1[0] = ra;
                    c0[0] = m0 - ra * C;
                    c1[0] = m1 - rb * C;
                    nr = 1;
                } else if (rb == ra + 1) {
                    r0[0] = r1[0] = ra;
                    c0[0] = m0 - ra * C;
                    c1[0] = C - 1;
                    r0[1] = r1[1] = rb;
                    c0[1] = 0;
                    c1[1] = m1 - rb * C;
                    nr = 2;
                } else {
                    full = true;
                }
            }
        }
        // centres grow with column and row only for non-negative facet gaps; otherwise no bound
        if (H.type == 0 && !(H.gap[0] >= 0.0f && H.gap[1] >= 0.0f)) full = true;
        const CD3 pos{H.pos[0], H.pos[1], H.pos[2]};
        const double pl = H.pl;
        int shapes = 0;
        for (int g = 0; g < nr && !full; ++g) {
            if (H.type == 0) {
                CD3 u, nn, v;
                frame_axes(H.fr, u, nn, v);
                auto xcol = [&](int col) {
                    return col * pl + (int)fast_div((uint32_t)col, H.fd_sub_cols) * (double)H.gap[0] + pl / 2 -
                           (double)H.size[0] / 2;
                };
                auto zrow = [&](int row) {
                    return row * pl + (int)fast_div((uint32_t)row, H.fd_sub_rows) * (double)H.gap[1] + pl / 2 -
                           (double)H.size[2] / 2;
                };
                const double xa = xcol(c0[g]), xb = xcol(c1[g]), za = zrow(r0[g]), zb = zrow(r1[g]);
                const CD3 c = cadd(pos, cadd(cadd(cmul(u, 0.5 * (xa + xb)), cmul(nn, 0.5 * (double)H.size[1])),
                                             cmul(v, 0.5 * (za + zb))));
                wave_shape_bits(a, sq, n, c, cmul(u, 0.5 * (xb - xa)), cmul(v, 0.5 * (zb - za)), nn, m);
                ++shapes;
            } else {
                CD3 U, N, V;
                frame_axes(H.fr_focus, U, N, V);
                const int br0 = (int)fast_div((uint32_t)r0[g], H.fd_sub_rows),
                          br1 = (int)fast_div((uint32_t)r1[g], H.fd_sub_rows);
                const int bc0 = (int)fast_div((uint32_t)c0[g], H.fd_sub_cols),
                          bc1 = (int)fast_div((uint32_t)c1[g], H.fd_sub_cols);
                for (int br = br0; br <= br1 && !full; ++br) {
                    for (int bc = bc0; bc <= bc1 && !full; ++bc) {
                        if (++shapes > 8) {
                            full = true;
                            break;
                        }
                        const DevFacet &Fc = a.facets[H.facet_base + br * H.cols + bc];
                        const int ri0 = max(r0[g], br * H.sub_rows) - br * H.sub_rows;
                        const int ri1 = min(r1[g], br * H.sub_rows + H.sub_rows - 1) - br * H.sub_rows;
                        const int ci0 = max(c0[g], bc * H.sub_cols) - bc * H.sub_cols;
                        const int ci1 = min(c1[g], bc * H.sub_cols + H.sub_cols - 1) - bc * H.sub_cols;
                        const double xa = (ci0 + 0.5) * pl - (double)H.len_col / 2,
                                     xb = (ci1 + 0.5) * pl - (double)H.len_col / 2;
                        const double za = (ri0 + 0.5) * pl - (double)H.len_row / 2,
                                     zb = (ri1 + 0.5) * pl - (double)H.len_row / 2;
                        CD3 uf, nf, vf, up, np, vp;
                        frame_axes(Fc.fr_local, uf, nf, vf);
                        frame_axes(Fc.fr_pert, up, np, vp);
                        auto rot = [&](CD3 p) { return cadd(cadd(cmul(U, p.x), cmul(N, p.y)), cmul(V, p.z)); };
                        const CD3 ch = cadd(CD3{Fc.center[0], Fc.center[1], Fc.center[2]},
                                            cadd(cmul(uf, 0.5 * (xa + xb)), cmul(vf, 0.5 * (za + zb))));
                        wave_shape_bits(a, sq, n, cadd(pos, rot(ch)), cmul(rot(uf), 0.5 * (xb - xa)),
                                        cmul(rot(vf), 0.5 * (zb - za)), np, m);
                    }
                }
            }
        }
        if (full) {
            for (int k = 0; k < 2; ++k)
                m[k] = (n[k] > 0 && n[k] <= kCullWide) ? (n[k] >= 64 ? ~0ull : (1ull << n[k]) - 1ull) : 0ull;
        }
    }
    for (int j = 0; j < kWavesPerLane; ++j) {
        const int wj = w * kWavesPerLane + j;
        if (wj >= a.wstride) break;
        uint64_t *out = a.masks + 2 * ((size_t)li * a.wstride + wj);
        out[0] = m[0];
        out[1] = m[1];
    }
}

}  // namespace

// Culling scan groups of grid g's quads: each heliostat's sub-heliostats (runs of at most 64), every
// quad of no heliostat alone. The group ball holds each quad's bounding ball (centre A + (E1+E2)/2,
// radius half the longer diagonal, as the scan computes it from the same floats), widened by 1e-9
// relative for rounding.
static std::vector<QuadGroup> pack_groups(const std::vector<float4> &quads, int n_sub, const std::vector<DevHelio> &helios,
                                   int g) {
    std::vector<int> owner((size_t)n_sub, -1);
    for (size_t h = 0; h < helios.size(); ++h)
        if (helios[h].grid == g)
            for (int q = helios[h].sub_begin; q < helios[h].sub_begin + helios[h].sub_count; ++q)
                if (q >= 0 && q < n_sub) owner[q] = (int)h;
    std::vector<QuadGroup> out;
    for (int q = 0; q < n_sub;) {
        int e = q + 1;
        if (owner[q] >= 0)
            while (e < n_sub && owner[e] == owner[q] && e - q < 64) ++e;
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int k = q; k < e; ++k) {
            const float4 A = quads[3 * k], E1 = quads[3 * k + 1], E2 = quads[3 * k + 2];
            const double a[3] = {A.x, A.y, A.z}, e1[3] = {E1.x, E1.y, E1.z}, e2[3] = {E2.x, E2.y, E2.z};
            for (int i = 0; i < 4; ++i)
                for (int d = 0; d < 3; ++d) {
                    const double p = a[d] + ((i & 1) ? e1[d] : 0.0) + ((i & 2) ? e2[d] : 0.0);
                    lo[d] = std::min(lo[d], p);
                    hi[d] = std::max(hi[d], p);
                }
        }
        QuadGroup G;
        for (int d = 0; d < 3; ++d) G.c[d] = 0.5 * (lo[d] + hi[d]);
        double r = 0.0;
        for (int k = q; k < e; ++k) {
            const float4 A = quads[3 * k], E1 = quads[3 * k + 1], E2 = quads[3 * k + 2];
            const double s0 = (double)E1.x + E2.x, s1 = (double)E1.y + E2.y, s2 = (double)E1.z + E2.z;
            const double d0 = (double)E1.x - E2.x, d1 = (double)E1.y - E2.y, d2 = (double)E1.z - E2.z;
            const double cq[3] = {A.x + 0.5 * s0, A.y + 0.5 * s1, A.z + 0.5 * s2};
            const double rq = 0.5 * std::sqrt(std::max(s0 * s0 + s1 * s1 + s2 * s2, d0 * d0 + d1 * d1 + d2 * d2));
            const double dc = std::sqrt((cq[0] - G.c[0]) * (cq[0] - G.c[0]) + (cq[1] - G.c[1]) * (cq[1] - G.c[1]) +
                                        (cq[2] - G.c[2]) * (cq[2] - G.c[2]));
            r = std::max(r, dc + rq);
        }
        G.r = r * (1.0 + 1e-9) + 1e-9;
        G.begin = q;
        G.count = e - q;
        out.push_back(G);
        q = e;
    }
    return out;
}

// Uploads grid g's culling scan groups (after its quads and the heliostat records are known).
qmc_status qmc::upload_groups(qmc_ctx *c, int g, const std::vector<float4> &quads) {
    GridBuf &gb = c->grids[g];
    const std::vector<QuadGroup> groups = pack_groups(quads, gb.n_sub, c->h_helios, g);
    if ((size_t)gb.n_groups != groups.size() || !gb.groups) {
        if (qmc_status st = dev_alloc(gb.groups, groups.size())) return st;
        gb.n_groups = (int)groups.size();
    }
    if (!groups.empty())
        HIP_TRY(hipMemcpy(gb.groups, groups.data(), groups.size() * sizeof(QuadGroup), hipMemcpyHostToDevice));
    return QMC_OK;
}


// Traversal-culling candidate lists (k_cull) of the listed heliostats of a grid, on the ctx stream,
// into caller-provided per-heliostat arrays; `zero` (nullable) is set to 0 on the way.
qmc_status qmc::launch_cull_into(qmc_ctx *c, int grid, const int32_t *d_list, int n, CullList *lists, WideList *wide,
                                 unsigned *zero) {
    const GridBuf &gb = c->grids[grid];
    CullArgs ca;
    ca.helios = c->d_helios;
    ca.shapes = c->d_cull_shapes;
    ca.list = d_list;
    ca.quads = gb.quads;
    ca.groups = gb.groups;
    ca.n_groups = gb.n_groups;
    ca.sun = c->sun;
    ca.tan_sun = c->cull_tan_sun;
    ca.ang_block = c->cull_ang_block;
    ca.margin = kCullMargin;
    ca.lists = lists;
    ca.wide = wide;
    ca.zero = zero;
    if (gb.n_groups >= (1 << 29)) return fail(QMC_E_UNSUPPORTED, "trace: 2^29 or more culling groups in a grid");
    hipLaunchKernelGGL(k_cull, dim3((unsigned)n), dim3(kCullBS), 0, c->stream, ca);
    HIP_TRY(hipGetLastError());
    return QMC_OK;
}

// The same into the ctx-owned per-heliostat arrays (grown to the scene's heliostat count).
qmc_status qmc::launch_cull(qmc_ctx *c, int grid, const int32_t *d_list, int n, const CullList **lists,
                            const WideList **wide) {
    qmc_status st;
    if (c->cull_cap < c->h_helios.size()) {
        if ((st = dev_alloc(c->d_cull, c->h_helios.size()))) return st;
        c->cull_cap = c->h_helios.size();
    }
    if (wide && c->wide_cap < c->h_helios.size()) {
        if ((st = dev_alloc(c->d_wide, c->h_helios.size()))) {
            c->wide_cap = 0;
            return st;
        }
        c->wide_cap = c->h_helios.size();
    }
    if ((st = launch_cull_into(c, grid, d_list, n, c->d_cull, wide ? c->d_wide : nullptr))) return st;
    *lists = c->d_cull;
    if (wide) *wide = c->d_wide;
    return QMC_OK;
}

// Per-wave candidate masks (k_cull_waves) of the listed heliostats of a grid, after their wide lists,
// into `masks` (2 * n * wave_stride(max_slots) entries).
qmc_status qmc::launch_cull_waves_into(qmc_ctx *c, int grid, const int32_t *d_list, int n, long long max_slots,
                                       uint64_t *masks, TraceArgs &a) {
    const long long wstride = wave_stride(max_slots);
    if (wstride > 0x7fffffffLL) return fail(QMC_E_UNSUPPORTED, "trace: too many rays per heliostat");
    WaveArgs w;
    std::memset(&w, 0, sizeof(w));
    w.helios = c->d_helios;
    w.facets = c->d_facets;
    w.list = d_list;
    w.wide = a.wide;
    w.quads = c->grids[grid].quads;
    w.sun = c->sun;
    w.tan_sun = c->cull_tan_sun;
    w.tan_block = c->cull_tan_block;
    w.margin = kCullMargin;
    w.n_list = n;
    w.L = c->L;
    w.fdL = make_fastdiv((uint32_t)c->L);
    w.wstride = (int)wstride;
    w.masks = masks;
    hipLaunchKernelGGL(k_cull_waves, dim3((unsigned)((wstride + 256LL * kWavesPerLane - 1) / (256LL * kWavesPerLane)), (unsigned)n), dim3(256), 0, c->stream, w);
    HIP_TRY(hipGetLastError());
    a.wmask = masks;
    a.wstride = (int)wstride;
    return QMC_OK;
}
