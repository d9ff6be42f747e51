// qmc_adjoint.hip — differentiable trace: bilinear flux deposit, visibility masks, the flux-map loss
// and a hand-written adjoint giving d(loss)/d(per-heliostat aim-point offsets, sun direction).
// No reference counterpart (the reference only renders flux maps); the per-ray forward model is the
// reference's map_raytracing (QuasiMonteCarloRayTracing/RectangleReceiverRectangleGridRayTracing.cu:
// 6-41) with the nearest-pixel deposit replaced by a bilinear one.
//
// Differentiated model of one visible ray of heliostat h (flat heliostats, rectangle receiver):
//   n_h   = normalize(normalize(A_h - p_h) - s),  A_h = focus + o0*across + o1*up
//   (u,n,v) frame of n_h; O = p_h + lf.x*u + lf.y*n + lf.z*v (micro-facet origin, lf fixed)
//   d_in  = R(-s) q_in  (sun-shape sample), N = normalize(R(n_h) q_p) (slope-error sample)
//   d     = normalize(reflect(-d_in, N)); (t, uu, vv) = Moeller-Trumbore hit of the receiver plane
//   E     = |d.N| eta(t) factor;  splat E with bilinear weights at (vv*res_x - 0.5, uu*res_y - 0.5)
// Visibility (not shadowed, not blocked, inside the receiver) is decided by the float trace of the
// plain tracer and then frozen, so the gradient is the derivative of the masked sum. Rays are never
// stored: the backward kernel regenerates each ray from its pool indices (the QMC sample offsets).
//
// Precision: the splat's derivative jumps where a hit crosses a pixel centre, so a gradient is only
// reproducible if hit positions are: float geometry (about 1e-3 pixel of rounding at a 500 m slant
// range) flips enough rays across those kinks to move a field-wide gradient by 1e-2. The model
// (frames, reflection, plane hit, weights, adjoint) therefore runs in double — half the FP32 vector
// rate on CDNA, a small cost next to the float DDA that decides visibility. The forward deposits
// float(E*w) with float atomics; the masked forward accumulates in double (finite differences).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "qmc_ctx.h"
#include "qmc_trace_core.h"

using namespace qmc;

namespace {

struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 operator-(D3 a) { return D3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ D3 operator*(D3 a, double s) { return D3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 dcross(D3 a, D3 b) {
    return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ D3 ldd(const double *p) { return D3{p[0], p[1], p[2]}; }

struct AdjHelio {
    double u[3], n[3], v[3];
    double nw1, nw2;   // 1/|Y x n|, 1/|u x n|
    double p[3];       // position
    int64_t ray_base;  // first mask bit of this heliostat (multiple of 64); bit ray_base + slot per ray slot
};

struct AdjSun {
    double us[3], m[3], vs[3];  // frame of m = -s
    double nw1, nw2;            // 1/|Y x m|, 1/|us x m|
};

struct AdjArgs {
    DevGrid grid;
    DevReceiver recv;
    const DevHelio *helios;
    const AdjHelio *ah;
    AdjSun sun;
    DevFrame sunf;
    const int32_t *list;
    const float4 *pool;
    const float4 *pert;
    const int32_t *offsets;
    int32_t P, L;
    FastDiv fdP, fdL;
    int32_t n_list;
    float *flux;          // bilinear map (forward)
    double *flux64;       // bilinear map (masked forward)
    uint32_t *mask;       // visibility bits, one per ray slot (slot_ray order)
    const WideList *wide; // per heliostat id: candidate quads of each traversal
    const uint64_t *wmask;// per-wave candidate masks of the list's slots (k_cull_waves)
    int32_t wstride;      // waves per list entry in wmask
    const float *G;       // dL/dflux (backward)
    double *g_part;       // backward: per block, its rays' dL/dn then dL/dm (6 doubles), m = -s
};

// Forward values of the differentiated model for one ray (kept for the reverse sweep).
struct RayModel {
    D3 lf, N, ii, d, pv, T, q;
    double ilN, k, ilr, idet, uu, vv, t, c, eta, deta, E;  // ilN, ilr, idet: reciprocals
};

__device__ __forceinline__ void ray_model(const AdjArgs &a, const DevHelio &H, const AdjHelio &Ah, int m, float4 qi,
                                          float4 qp, RayModel &r) {
    const D3 u = ldd(Ah.u), n = ldd(Ah.n), v = ldd(Ah.v);
    const D3 us = ldd(a.sun.us), mm = ldd(a.sun.m), vs = ldd(a.sun.vs);
    // local facet coordinates (RectangleHelio.cu:64-79)
    const int cpr = H.cols * H.sub_cols;
    const int row = (int)fast_div((uint32_t)m, H.fd_cpr), col = m - row * cpr;
    const int bcol = (int)fast_div((uint32_t)col, H.fd_sub_cols), brow = (int)fast_div((uint32_t)row, H.fd_sub_rows);
    const double pl = H.pl;
    r.lf = D3{col * pl + bcol * (double)H.gap[0] + pl / 2.0 - H.size[0] / 2.0, H.size[1] / 2.0,
              row * pl + brow * (double)H.gap[1] + pl / 2.0 - H.size[2] / 2.0};
    const D3 O = ldd(Ah.p) + u * r.lf.x + n * r.lf.y + v * r.lf.z;
    const D3 din = us * (double)qi.x + mm * (double)qi.y + vs * (double)qi.z;
    const D3 Nraw = u * (double)qp.x + n * (double)qp.y + v * (double)qp.z;
    // divisions by a vector length or the determinant are products with one reciprocal
    r.ilN = 1.0 / sqrt(ddot(Nraw, Nraw));
    r.N = Nraw * r.ilN;
    r.ii = -din;
    r.k = ddot(r.N, r.ii);
    const D3 rr = r.ii - r.N * (r.k * 2.0);
    r.ilr = 1.0 / sqrt(ddot(rr, rr));
    r.d = rr * r.ilr;
    const DevReceiver &R = a.recv;
    const D3 A0{R.A[0], R.A[1], R.A[2]}, E1{R.E1[0], R.E1[1], R.E1[2]}, E2{R.E2[0], R.E2[1], R.E2[2]};
    r.pv = dcross(r.d, E2);
    r.idet = 1.0 / ddot(E1, r.pv);
    r.T = O - A0;
    r.uu = ddot(r.T, r.pv) * r.idet;
    r.q = dcross(r.T, E1);
    r.vv = ddot(r.d, r.q) * r.idet;
    r.t = ddot(E2, r.q) * r.idet;
    // atmospheric attenuation (global_function.cuh eta_aAlpha), with its derivative
    if (r.t <= 1000.0) {
        r.eta = 0.99331 - r.t * 0.0001176 + r.t * r.t * 1.97e-8;
        r.deta = -0.0001176 + 2.0 * 1.97e-8 * r.t;
    } else {
        r.eta = exp(r.t * -0.0001106);
        r.deta = -0.0001106 * r.eta;
    }
    r.c = ddot(r.d, r.N);
    r.E = fabs(r.c) * r.eta * (double)H.factor;
}

// Bilinear deposit of E at (vv*res_x - 0.5, uu*res_y - 0.5) into the double map (masked forward).
__device__ __forceinline__ void bilinear_splat_f64(const AdjArgs &a, double uu, double vv, double E) {
    const int resx = a.recv.res_x, resy = a.recv.res_y;
    const double x = vv * (double)resx - 0.5, y = uu * (double)resy - 0.5;
    const double fx0 = floor(x), fy0 = floor(y);
    const int i0 = (int)fx0, j0 = (int)fy0;
    const double fx = x - fx0, fy = y - fy0;
    const double wx[2] = {1.0 - fx, fx}, wy[2] = {1.0 - fy, fy};
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int i = i0 + c, j = j0 + b;
            if (i >= 0 && i < resx && j >= 0 && j < resy) unsafeAtomicAdd(a.flux64 + j * resx + i, E * wx[c] * wy[b]);
        }
}

// Visibility of ray r (slot `slot` of list entry li, heliostat h) by the plain trace's float rules
// (RectangleReceiverRectangleGridRayTracing.cu:14-40) with its culling: traversals proved hit-free
// are skipped, the others walk only the cells of the masked candidates the ray hits, unless the walk
// is proven to reach a hit quad's cells (the outcome, not which quad, matters here). True when the
// ray reaches the receiver, with its micro-facet m and pool samples s (sun shape) and q (slope error).
__device__ __forceinline__ bool adj_visible(const AdjArgs &a, const DevHelio &H, int h, int li, long long slot,
                                            long long r, int &m, float4 &s, float4 &q) {
    const int L = a.L, P = a.P;
    m = (int)fast_div((uint32_t)r, a.fdL);
    const int j = (int)r - m * L;
    const int32_t *grp = a.offsets + H.offsets_base;
    s = a.pool[fast_mod((uint32_t)(grp[m] + j), P, a.fdP)];
    q = a.pert[fast_mod((uint32_t)(grp[fast_mod((uint32_t)(m + 1), P, a.fdP)] + j), P, a.fdP)];
    V3 dir = fr_apply(a.sunf, V3{s.x, s.y, s.z});
    const V3 origin = flat_facet_origin(H, H.fr, m);
    const WideList *wl = a.wide + h;
    // the caller's slot blocks hold whole waves, so the wave's masks are uniform
    const uint64_t *wm = a.wmask + 2 * ((size_t)li * a.wstride + (size_t)(slot >> 6));
    RayStats st{0u, 0u, 0u};
    RayHits rh;
    int hit;
    const int lo = H.sub_begin, hi = H.sub_begin + H.sub_count;
    int plan = traversal_plan(wl, wm, 0, origin, dir, a.grid, rh);
    if (plan != 0 && ((plan == 1 && known_hit_proven(rh, origin, dir, a.grid, hit)) ||
                      (plan == 1 && rh.n <= 2 ? dda_known<false>(rh, origin, dir, a.grid, lo, hi, hit, st)
                                              : dda_collision<false>(origin, dir, a.grid, a.grid.quads, lo, hi, hit, st))))
        return false;
    const V3 normal = normalize(fr_apply(H.fr, V3{q.x, q.y, q.z}));
    dir = normalize(reflect(-dir, normal));
    plan = traversal_plan(wl, wm, 1, origin, dir, a.grid, rh);
    if (plan != 0 && ((plan == 1 && known_hit_proven(rh, origin, dir, a.grid, hit)) ||
                      (plan == 1 && rh.n <= 2 ? dda_known<false>(rh, origin, dir, a.grid, lo, hi, hit, st)
                                              : dda_collision<false>(origin, dir, a.grid, a.grid.quads, lo, hi, hit, st))))
        return false;
    float t, u, v;
    return receiver_hit<0>(a.recv, origin, dir, t, u, v);
}

// Bilinear deposit of E at (vv*res_x - 0.5, uu*res_y - 0.5): rows [row0, row0 + rows) into the
// workgroup's LDS band, the others as device-scope float atomics into the map.
__device__ __forceinline__ void bilinear_splat_band(const AdjArgs &a, float *band, int row0, int rows, double uu,
                                                    double vv, double E) {
    const int resx = a.recv.res_x, resy = a.recv.res_y;
    const double x = vv * (double)resx - 0.5, y = uu * (double)resy - 0.5;
    const double fx0 = floor(x), fy0 = floor(y);
    const int i0 = (int)fx0, j0 = (int)fy0;
    const double fx = x - fx0, fy = y - fy0;
    const double wx[2] = {1.0 - fx, fx}, wy[2] = {1.0 - fy, fy};
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int i = i0 + c, j = j0 + b;
            if (i >= 0 && i < resx && j >= 0 && j < resy) {
                const float w = (float)(E * wx[c] * wy[b]);
                const int br = j - row0;
                if ((unsigned)br < (unsigned)rows)
                    atomicAdd(band + br * resx + i, w);
                else
                    unsafeAtomicAdd(a.flux + j * resx + i, w);
            }
        }
}

// Forward with full visibility: visibility by adj_visible, double bilinear deposit, one visibility
// bit per ray slot. Persistent like the plain banded trace: each of a deposit's four pixels is an
// atomic, so the workgroup keeps map rows [row0, row0 + rows) in LDS (the whole map when it fits)
// and takes BS-slot blocks (list entry, block) round robin; at the end it stores its band to its own
// slice of `partials` (k_adj_band_fold adds them).
template <int BS>
__global__ __launch_bounds__(BS) void k_adj_forward(AdjArgs a, int row0, int rows, int blocks_per_helio,
                                                    float *partials) {
    extern __shared__ float band[];
    const int W = a.recv.res_x, nband = rows * W;
    for (int p = threadIdx.x; p < nband; p += BS) band[p] = 0.0f;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    // per-ray constants in vector registers (in_vgpr): the sun frames, the receiver's edges and (below)
    // the heliostat's frame and facet geometry; in scalar registers they overflow, and the compiler
    // spills and reloads inside the loop (differentiable forward 0.131 -> 0.125 ms on the PS10 sample)
    AdjArgs va = a;
    vgpr_frame(va.sunf);
    for (int k = 0; k < 3; ++k) {
        in_vgpr(va.recv.A[k]);
        in_vgpr(va.recv.E1[k]);
        in_vgpr(va.recv.E2[k]);
        in_vgpr(va.sun.us[k]);
        in_vgpr(va.sun.m[k]);
        in_vgpr(va.sun.vs[k]);
    }
    in_vgpr(va.sun.nw1);
    in_vgpr(va.sun.nw2);
    int li = (int)(blockIdx.x / (unsigned)blocks_per_helio);
    int bx = (int)blockIdx.x - li * blocks_per_helio;
#pragma nounroll
    while (li < a.n_list) {
        // the launch writes none of these records, so they are read with scalar loads
        const int h = ld_uniform(a.list + __builtin_amdgcn_readfirstlane(li));
        DevHelio H = ld_uniform(a.helios + h);
        vgpr_frame(H.fr);
        for (int k = 0; k < 3; ++k) {
            in_vgpr(H.pos[k]);
            in_vgpr(H.size[k]);
        }
        in_vgpr(H.gap[0]);
        in_vgpr(H.gap[1]);
        in_vgpr(H.pl);
        in_vgpr(H.factor);
        const long long slot = (long long)__builtin_amdgcn_readfirstlane(bx) * BS + threadIdx.x;
        const long long r = slot_ray(H, a.L, slot);
        bool visible = false;
        if (r >= 0) {
            int m;
            float4 s, q;
            visible = adj_visible(va, H, h, li, slot, r, m, s, q);
            if (visible) {
                RayModel rm;
                ray_model(va, H, a.ah[h], m, s, q, rm);
                bilinear_splat_band(va, band, row0, rows, rm.uu, rm.vv, rm.E);
            }
        }
        // ray_base is a multiple of 64 and blocks hold whole waves: each wave owns two mask words
        const unsigned long long bits = __ballot(visible);
        if ((lane & 31) == 0 && slot - lane < heliostat_slot_count(H, a.L)) {
            const long long bit = a.ah[h].ray_base + slot;
            a.mask[bit >> 5] = (uint32_t)(bits >> lane);
        }
        bx += (int)gridDim.x;
        while (bx >= blocks_per_helio) {
            bx -= blocks_per_helio;
            ++li;
        }
    }
    __syncthreads();
    float *mine = partials + (size_t)blockIdx.x * nband;
    for (int p = threadIdx.x; p < nband; p += BS) mine[p] = band[p];
}

// Map rows [row0, row0 + nband / W) of the forward: the sum (in double) of the workgroups' bands. A
// block takes 64 pixels; its wave w adds the bands w, w + 4, ... for them (coalesced rows of
// partials), and the four wave sums are added in a fixed order. Those rows take no device-scope
// deposits, so the sum is stored.
__global__ __launch_bounds__(256) void k_adj_band_fold(float *flux, const float *partials, int n_parts, int nband,
                                                       int offset) {
    __shared__ double red[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int p = blockIdx.x * 64 + lane;
    double s = 0.0;
    if (p < nband) {
#pragma unroll 8
        for (int k = w; k < n_parts; k += 4) s += (double)partials[(size_t)k * nband + p];
    }
    red[w][lane] = s;
    __syncthreads();
    if (w == 0 && p < nband) flux[offset + p] = (float)(((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane]);
}

// Forward of the frozen-visibility model in double (finite differences with common random numbers).
__global__ __launch_bounds__(256) void k_adj_forward_masked(AdjArgs a) {
    const int h = a.list[blockIdx.y];
    const DevHelio &H = a.helios[h];
    const int L = a.L, P = a.P;
    const long long slot = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long r = slot_ray(H, L, slot);
    if (r < 0) return;
    const long long bit = a.ah[h].ray_base + slot;
    if (!((a.mask[bit >> 5] >> (bit & 31)) & 1u)) return;
    const int m = (int)fast_div((uint32_t)r, a.fdL), j = (int)r - m * L;
    const int32_t *grp = a.offsets + H.offsets_base;
    RayModel rm;
    ray_model(a, H, a.ah[h], m, a.pool[fast_mod((uint32_t)(grp[m] + j), P, a.fdP)],
              a.pert[fast_mod((uint32_t)(grp[fast_mod((uint32_t)(m + 1), P, a.fdP)] + j), P, a.fdP)], rm);
    bilinear_splat_f64(a, rm.uu, rm.vv, rm.E);
}

__device__ __forceinline__ D3 norm_bwd(D3 y, double inv_len, D3 g) {  // y = x/|x|, inv_len = 1/|x|
    return (g - y * ddot(y, g)) * inv_len;
}

__global__ __launch_bounds__(256) void k_adj_backward(AdjArgs a) {
    __shared__ double red[4][6];
    const int h = a.list[blockIdx.y];
    const DevHelio &H = a.helios[h];
    const int L = a.L, P = a.P;
    const long long slot = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long r = slot_ray(H, L, slot);
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // dL/dn (3), dL/dm (3)
    bool vis = false;
    if (r >= 0) {
        const long long bit = a.ah[h].ray_base + slot;
        vis = (a.mask[bit >> 5] >> (bit & 31)) & 1u;
    }
    if (vis) {
        const AdjHelio &Ah = a.ah[h];
        const int m = (int)fast_div((uint32_t)r, a.fdL), j = (int)r - m * L;
        const int32_t *grp = a.offsets + H.offsets_base;
        const float4 qi4 = a.pool[fast_mod((uint32_t)(grp[m] + j), P, a.fdP)];
        const float4 qp4 = a.pert[fast_mod((uint32_t)(grp[fast_mod((uint32_t)(m + 1), P, a.fdP)] + j), P, a.fdP)];
        RayModel f;
        ray_model(a, H, Ah, m, qi4, qp4, f);
        // d(loss)/d(E), d/d(x), d/d(y) through the four bilinear weights
        const int resx = a.recv.res_x, resy = a.recv.res_y;
        const double x = f.vv * (double)resx - 0.5, y = f.uu * (double)resy - 0.5;
        const double fx0 = floor(x), fy0 = floor(y);
        const int i0 = (int)fx0, j0 = (int)fy0;
        const double fx = x - fx0, fy = y - fy0;
        const double wx[2] = {1.0 - fx, fx}, wy[2] = {1.0 - fy, fy}, dw[2] = {-1.0, 1.0};
        double gE = 0.0, gx = 0.0, gy = 0.0;
#pragma unroll
        for (int bb = 0; bb < 2; ++bb)
#pragma unroll
            for (int aa = 0; aa < 2; ++aa) {
                const int pi = i0 + aa, pj = j0 + bb;
                if (pi >= 0 && pi < resx && pj >= 0 && pj < resy) {
                    const double Gp = a.G[pj * resx + pi];
                    gE += Gp * wx[aa] * wy[bb];
                    gx += Gp * f.E * dw[aa] * wy[bb];
                    gy += Gp * f.E * wx[aa] * dw[bb];
                }
            }
        // reverse sweep
        const double fac = H.factor;
        const double g_vv = gx * (double)resx, g_uu = gy * (double)resy;
        const double g_c = gE * (f.c >= 0.0 ? 1.0 : -1.0) * f.eta * fac;
        const double g_t = gE * fabs(f.c) * fac * f.deta;
        D3 g_d = f.N * g_c;
        D3 g_N = f.d * g_c;
        const DevReceiver &R = a.recv;
        const D3 E1{R.E1[0], R.E1[1], R.E1[2]}, E2{R.E2[0], R.E2[1], R.E2[2]};
        const double g_a1 = g_uu * f.idet, g_a2 = g_vv * f.idet, g_a3 = g_t * f.idet;
        const double g_det = -(g_uu * f.uu + g_vv * f.vv + g_t * f.t) * f.idet;
        D3 g_T = f.pv * g_a1;
        const D3 g_pv = f.T * g_a1 + E1 * g_det;
        g_d = g_d + dcross(E2, g_pv) + f.q * g_a2;
        const D3 g_q = f.d * g_a2 + E2 * g_a3;
        g_T = g_T + dcross(E1, g_q);
        const D3 g_O = g_T;
        const D3 g_r = norm_bwd(f.d, f.ilr, g_d);
        const double grN = ddot(f.N, g_r);
        const D3 g_i = g_r - f.N * (2.0 * grN);
        g_N = g_N - (g_r * f.k + f.ii * grN) * 2.0;
        const D3 g_din = -g_i;
        const D3 g_Nraw = norm_bwd(f.N, f.ilN, g_N);
        const D3 u = ldd(Ah.u), n = ldd(Ah.n), v = ldd(Ah.v);
        D3 g_u = g_Nraw * (double)qp4.x + g_O * f.lf.x;
        D3 gn = g_Nraw * (double)qp4.y + g_O * f.lf.y;
        const D3 g_v = g_Nraw * (double)qp4.z + g_O * f.lf.z;
        // heliostat frame: w1 = Y x n, u = w1/|w1|, w2 = u x n, v = w2/|w2|
        const D3 Y{0.0, 1.0, 0.0};
        const D3 g_w2 = norm_bwd(v, Ah.nw2, g_v);
        g_u = g_u + dcross(n, g_w2);
        gn = gn + dcross(g_w2, u);
        const D3 g_w1 = norm_bwd(u, Ah.nw1, g_u);
        gn = gn + dcross(g_w1, Y);
        // sun frame (m = -s)
        const D3 us = ldd(a.sun.us), mm = ldd(a.sun.m), vs = ldd(a.sun.vs);
        const D3 g_us = g_din * (double)qi4.x, g_vs = g_din * (double)qi4.z;
        D3 gm = g_din * (double)qi4.y;
        const D3 g_ws2 = norm_bwd(vs, a.sun.nw2, g_vs);
        const D3 g_us2 = g_us + dcross(mm, g_ws2);
        gm = gm + dcross(g_ws2, us);
        const D3 g_ws1 = norm_bwd(us, a.sun.nw1, g_us2);
        gm = gm + dcross(g_ws1, Y);
        acc[0] = gn.x; acc[1] = gn.y; acc[2] = gn.z;
        acc[3] = gm.x; acc[4] = gm.y; acc[5] = gm.z;
    }
    // one block = a slice of one heliostat's rays: wave sums, then the block's sum in a fixed order
    // into its own partial (no atomics: every block shares dL/dm, and the result is reproducible)
#pragma unroll
    for (int k = 0; k < 6; ++k)
        for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 6; ++k) red[w][k] = acc[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        double s = 0.0;
        for (int ww = 0; ww < (int)(blockDim.x >> 6); ++ww) s += red[ww][k];
        a.g_part[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 6 + k] = s;
    }
}

// Per list entry li: the sum of its n_blocks backward partials (lane-strided, then a fixed shuffle
// tree), dL/dn into g_n[3 * list[li]] and the heliostat's share of dL/dm into g_m[3 * li].
__global__ __launch_bounds__(64) void k_adj_reduce(const double *part, int n_blocks, const int32_t *list, double *g_n,
                                                   double *g_m) {
    const int li = blockIdx.x, lane = threadIdx.x;
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int b = lane; b < n_blocks; b += 64)
#pragma unroll
        for (int k = 0; k < 6; ++k) acc[k] += part[((size_t)li * n_blocks + b) * 6 + k];
#pragma unroll
    for (int k = 0; k < 6; ++k)
        for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o);
    if (lane == 0) {
        const int h = list[li];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            g_n[3 * h + k] = acc[k];
            g_m[3 * li + k] = acc[3 + k];
        }
    }
}

// L = sum (F - target)^2 + lambda * sum max(0, F - peak)^2 and G = dL/dF; per-block partial sums.
__global__ __launch_bounds__(256) void k_flux_loss(const float *F, int n, float target, float peak, float lambda,
                                                  float *G, double *partial) {
    __shared__ double red[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double l = 0.0;
    if (i < n) {
        const float e = F[i] - target;
        const float over = fmaxf(F[i] - peak, 0.0f);
        G[i] = 2.0f * e + 2.0f * lambda * over;
        l = (double)e * e + (double)lambda * over * over;
    }
    for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = l;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += red[k];
        partial[blockIdx.x] = s;
    }
}

// ---------------------------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------------------------
struct AdjState {
    int grid = -1;
    std::vector<int32_t> list;
    std::vector<AdjHelio> h_ah;  // indexed by heliostat id
    AdjHelio *d_ah = nullptr;
    size_t ah_cap = 0;
    int32_t *d_list = nullptr;
    size_t list_cap = 0;
    // pinned staging of the uploads (read by the copies until up_done) and of the gradient results
    AdjHelio *p_ah = nullptr;
    size_t p_ah_cap = 0;
    int32_t *p_list = nullptr;
    size_t p_list_cap = 0;
    double *p_out = nullptr;
    size_t p_out_cap = 0;
    hipEvent_t up_done = nullptr;
    uint32_t *d_mask = nullptr;
    size_t mask_words = 0;
    double *d_gout = nullptr;    // per heliostat id dL/dn (3 doubles), then per list entry its share of dL/dm
    size_t gout_cap = 0;
    double *d_gpart = nullptr;   // backward partials, 6 doubles per block
    size_t gpart_cap = 0;
    float *d_G = nullptr;
    double *d_partial = nullptr;
    double *d_flux64 = nullptr;
    size_t flux64_cap = 0;
    size_t npix = 0;
    uint64_t *d_wmask = nullptr;
    size_t wmask_cap = 0;  // uint64 entries
    float *d_parts = nullptr;
    size_t parts_cap = 0;  // floats
    bool mask_ready = false;
    AdjSun sun{};
    long long max_slots = 0;
};

AdjState *state(qmc_ctx *c) {
    if (!c->adj) c->adj = new AdjState();
    return static_cast<AdjState *>(c->adj);
}

// Grows a device buffer to at least `need` elements, without keeping its contents; earlier launches
// may still use the old buffer, so the stream is drained first.
template <class T>
qmc_status grow(qmc_ctx *c, T *&p, size_t &cap, size_t need) {
    if (p && need <= cap) return QMC_OK;
    HIP_TRY(hipStreamSynchronize(c->stream));
    cap = 0;
    if (qmc_status st = dev_alloc(p, need)) return st;
    cap = std::max<size_t>(need, 1);
    return QMC_OK;
}

// The same for pinned host memory (the caller makes sure that no copy still reads or writes it).
template <class T>
qmc_status grow_pinned(T *&p, size_t &cap, size_t need) {
    if (p && need <= cap) return QMC_OK;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    need = std::max<size_t>(need, 1);
    HIP_TRY(hipHostMalloc((void **)&p, need * sizeof(T)));
    cap = need;
    return QMC_OK;
}

// Slots per block of the persistent forward (one workgroup per CU with the LDS band).
constexpr int kAdjBS = 1024;

// Grid of the masked forward and the backward: 256-slot blocks (x) of each list entry (y).
dim3 adj_grid(qmc_ctx *, AdjState *s) {
    return dim3((unsigned)((s->max_slots + 255) / 256), (unsigned)s->list.size());
}

struct H3 {
    double x, y, z;
};
H3 hsub(H3 a, H3 b) { return H3{a.x - b.x, a.y - b.y, a.z - b.z}; }
double hdot(H3 a, H3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
H3 hcross(H3 a, H3 b) { return H3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
H3 hnorm(H3 a, double &len) {
    len = std::sqrt(hdot(a, a));
    return H3{a.x / len, a.y / len, a.z / len};
}
void put(double *dst, H3 a) {
    dst[0] = a.x;
    dst[1] = a.y;
    dst[2] = a.z;
}

// (u, v) of the frame of n with u = normalize(Y x n), v = normalize(u x n), and the reciprocals of
// the two normalised lengths; false when n is (anti)parallel to Y, where the frame is not
// differentiable.
bool frame_f64(H3 n, H3 &u, H3 &v, double &inw1, double &inw2) {
    double nw1, nw2;
    u = hnorm(hcross(H3{0.0, 1.0, 0.0}, n), nw1);
    v = hnorm(hcross(u, n), nw2);
    inw1 = 1.0 / nw1;
    inw2 = 1.0 / nw2;
    return nw1 > 1e-9 && nw2 > 1e-9;
}

// Heliostat normal of the differentiable model: normalize(normalize(A - p) - s), in double.
H3 helio_normal_f64(const double *A, const float *pos, const float *sun, double &la, double &lb, H3 &rr) {
    const H3 a = hsub(H3{A[0], A[1], A[2]}, H3{pos[0], pos[1], pos[2]});
    rr = hnorm(a, la);
    return hnorm(hsub(rr, H3{sun[0], sun[1], sun[2]}), lb);
}

// Prepare the per-heliostat adjoint records of the current scene state for grid g.
qmc_status adj_prepare(qmc_ctx *c, int grid) {
    if (!c->uploaded || !c->pool_ready || !c->offsets_ready)
        return fail(QMC_E_STATE, "adjoint: upload a scene and set pools and offsets first");
    if (grid < 0 || grid >= (int)c->grids.size()) return fail(QMC_E_RANGE, "adjoint: grid out of range");
    const RecvBuf &rb = c->recvs[c->grids[grid].receiver];
    if (rb.r.type != 0) return fail(QMC_E_UNSUPPORTED, "adjoint: rectangle receivers only");
    AdjState *s = state(c);
    s->grid = grid;
    s->list.clear();
    s->h_ah.assign(c->h_helios.size(), AdjHelio{});
    int64_t base = 0;
    s->max_slots = 0;
    for (size_t h = 0; h < c->h_helios.size(); ++h) {
        const DevHelio &H = c->h_helios[h];
        if (H.grid != grid) continue;
        if (H.type != 0 || H.sub_count != 1) return fail(QMC_E_UNSUPPORTED, "adjoint: flat heliostats only");
        s->list.push_back((int32_t)h);
        AdjHelio &A = s->h_ah[h];
        double la, lb;
        H3 rr, u, v;
        const H3 n = helio_normal_f64(&c->h_aim64[3 * h], H.pos, c->cfg.sun_dir, la, lb, rr);
        if (!frame_f64(n, u, v, A.nw1, A.nw2))
            return fail(QMC_E_UNSUPPORTED, "adjoint: heliostat " + std::to_string(h) + " normal parallel to the y axis");
        put(A.u, u);
        put(A.n, n);
        put(A.v, v);
        for (int k = 0; k < 3; ++k) A.p[k] = H.pos[k];
        A.ray_base = base;
        const long long slots = heliostat_slot_count(H, c->L);
        base += ((slots + 63) / 64) * 64;
        s->max_slots = std::max(s->max_slots, slots);
    }
    // slots and ray ids are divided by multiply-high constants, exact below 2^31
    if (s->max_slots >= 0x7fffffffLL) return fail(QMC_E_UNSUPPORTED, "adjoint: 2^31 or more rays per heliostat");
    if (s->list.size() > 65535) return fail(QMC_E_UNSUPPORTED, "adjoint: more than 65535 heliostats in one grid");
    const H3 mvec{-(double)c->cfg.sun_dir[0], -(double)c->cfg.sun_dir[1], -(double)c->cfg.sun_dir[2]};
    H3 us, vs;
    if (!frame_f64(mvec, us, vs, s->sun.nw1, s->sun.nw2))
        return fail(QMC_E_UNSUPPORTED, "adjoint: sun direction parallel to the y axis");
    put(s->sun.us, us);
    put(s->sun.m, mvec);
    put(s->sun.vs, vs);
    qmc_status st;
    const size_t nh = s->h_ah.size(), nl = s->list.size();
    if ((st = grow(c, s->d_mask, s->mask_words, (size_t)(base / 32) + 2))) return st;
    if ((st = grow(c, s->d_ah, s->ah_cap, nh))) return st;
    if ((st = grow(c, s->d_list, s->list_cap, nl))) return st;
    if ((st = grow(c, s->d_gout, s->gout_cap, 3 * (nh + nl)))) return st;
    if ((st = grow(c, s->d_gpart, s->gpart_cap, 6 * (size_t)adj_grid(c, s).x * nl))) return st;
    const size_t npix = (size_t)rb.w * rb.h;
    if (npix != s->npix) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        if ((st = dev_alloc(s->d_G, npix))) return st;
        if ((st = dev_alloc(s->d_partial, (npix + 255) / 256))) return st;
        s->npix = npix;
    }
    // the records go up through pinned staging: the previous upload must have read it
    if (!s->up_done) HIP_TRY(hipEventCreateWithFlags(&s->up_done, hipEventDisableTiming));
    else HIP_TRY(hipEventSynchronize(s->up_done));
    if ((st = grow_pinned(s->p_ah, s->p_ah_cap, nh))) return st;
    if ((st = grow_pinned(s->p_list, s->p_list_cap, nl))) return st;
    std::memcpy(s->p_ah, s->h_ah.data(), sizeof(AdjHelio) * nh);
    std::memcpy(s->p_list, s->list.data(), sizeof(int32_t) * nl);
    HIP_TRY(hipMemcpyAsync(s->d_ah, s->p_ah, sizeof(AdjHelio) * nh, hipMemcpyHostToDevice, c->stream));
    if (nl) HIP_TRY(hipMemcpyAsync(s->d_list, s->p_list, sizeof(int32_t) * nl, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipEventRecord(s->up_done, c->stream));
    return QMC_OK;
}

AdjArgs adj_args(qmc_ctx *c, AdjState *s, float *flux) {
    AdjArgs a;
    std::memset(&a, 0, sizeof(a));
    const GridBuf &gb = c->grids[s->grid];
    a.grid = gb.g;
    a.recv = c->recvs[gb.receiver].r;
    a.helios = c->d_helios;
    a.ah = s->d_ah;
    a.sun = s->sun;
    a.sunf = c->sun;
    a.list = s->d_list;
    a.pool = c->d_pool;
    a.pert = c->d_pert;
    a.offsets = c->d_offsets;
    a.n_list = (int32_t)s->list.size();
    a.P = c->P;
    a.L = c->L;
    a.fdP = make_fastdiv((uint32_t)c->P);
    a.fdL = make_fastdiv((uint32_t)c->L);
    a.flux = flux;
    a.flux64 = s->d_flux64;
    a.mask = s->d_mask;
    a.G = s->d_G;
    a.g_part = s->d_gpart;
    return a;
}


}  // namespace

namespace qmc {
void qmc_adjoint_release(qmc_ctx *c) {
    if (!c->adj) return;
    AdjState *s = static_cast<AdjState *>(c->adj);
    dev_free(s->d_ah);
    dev_free(s->d_list);
    dev_free(s->d_mask);
    dev_free(s->d_gout);
    if (s->up_done) {
        (void)hipEventSynchronize(s->up_done);
        (void)hipEventDestroy(s->up_done);
    }
    if (s->p_ah) (void)hipHostFree(s->p_ah);
    if (s->p_list) (void)hipHostFree(s->p_list);
    if (s->p_out) (void)hipHostFree(s->p_out);
    dev_free(s->d_G);
    dev_free(s->d_partial);
    dev_free(s->d_flux64);
    dev_free(s->d_wmask);
    dev_free(s->d_parts);
    dev_free(s->d_gpart);
    delete s;
    c->adj = nullptr;
}
}  // namespace qmc

extern "C" qmc_status qmc_adjoint_forward(qmc_ctx *c, int32_t grid, float *d_flux) {
    if (!c || !d_flux) return fail(QMC_E_INVALID_ARG, "qmc_adjoint_forward: null argument");
    qmc_status st = set_device(c);
    if (st) return st;
    if ((st = adj_prepare(c, grid))) return st;
    AdjState *s = state(c);
    if (s->list.empty()) return fail(QMC_E_INVALID_ARG, "qmc_adjoint_forward: grid has no heliostats");
    const RecvBuf &rb = c->recvs[c->grids[grid].receiver];
    AdjArgs a = adj_args(c, s, d_flux);
    // the plain trace's culling: wide candidate lists, then per-wave masks over the same slots
    const int n = (int)s->list.size();
    const CullList *cull = nullptr;
    TraceArgs ta;
    std::memset(&ta, 0, sizeof(ta));
    if ((st = launch_cull(c, grid, s->d_list, n, &cull, &ta.wide))) return st;
    const size_t need = 2 * (size_t)n * (size_t)wave_stride(s->max_slots);
    if (need > s->wmask_cap) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        if ((st = dev_alloc(s->d_wmask, need))) {
            s->wmask_cap = 0;
            return st;
        }
        s->wmask_cap = need;
    }
    if ((st = launch_cull_waves_into(c, grid, s->d_list, n, s->max_slots, s->d_wmask, ta))) return st;
    a.wide = ta.wide;
    a.wmask = ta.wmask;
    a.wstride = ta.wstride;
    // the band: as many map rows as the LDS holds, centred (flux concentrates around the aim points)
    if ((st = device_limits(c))) return st;
    const int W = rb.w, Hh = rb.h;
    const int rows = (int)std::min<size_t>((size_t)Hh, std::min(c->max_lds, c->lds_per_cu) / (sizeof(float) * W));
    const size_t lds = sizeof(float) * (size_t)rows * W;
    int resident = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, k_adj_forward<kAdjBS>, kAdjBS, lds));
    if (resident < 1) return fail(QMC_E_UNSUPPORTED, "qmc_adjoint_forward: no workgroup is resident");
    const long long bph = (s->max_slots + kAdjBS - 1) / kAdjBS;
    const int nwg = (int)std::min<long long>(c->num_cus, bph * n);
    const long long nband = (long long)rows * W;
    if ((size_t)(nwg * nband) > s->parts_cap) {
        HIP_TRY(hipStreamSynchronize(c->stream));
        if ((st = dev_alloc(s->d_parts, (size_t)(nwg * nband)))) {
            s->parts_cap = 0;
            return st;
        }
        s->parts_cap = (size_t)(nwg * nband);
    }
    const int row0 = (Hh - rows) / 2;
    // band rows are stored by the fold; the others take atomics (every mask word of the list's slots
    // is written by the forward)
    if (rows < Hh) HIP_TRY(hipMemsetAsync(d_flux, 0, sizeof(float) * (size_t)W * Hh, c->stream));
    hipLaunchKernelGGL(k_adj_forward<kAdjBS>, dim3((unsigned)nwg), dim3(kAdjBS), lds, c->stream, a, row0, rows,
                       (int)bph, s->d_parts);
    HIP_TRY(hipGetLastError());
    if (nband > 0) {
        hipLaunchKernelGGL(k_adj_band_fold, dim3((unsigned)((nband + 63) / 64)), dim3(256), 0, c->stream, d_flux,
                           s->d_parts, nwg, (int)nband, row0 * W);
        HIP_TRY(hipGetLastError());
    }
    s->mask_ready = true;
    return QMC_OK;
}

extern "C" qmc_status qmc_adjoint_forward_masked(qmc_ctx *c, int32_t grid, double *flux_out, int64_t n) {
    if (!c || !flux_out) return fail(QMC_E_INVALID_ARG, "qmc_adjoint_forward_masked: null argument");
    AdjState *s = state(c);
    if (!s->mask_ready || s->grid != grid)
        return fail(QMC_E_STATE, "qmc_adjoint_forward_masked: run qmc_adjoint_forward on this grid first");
    qmc_status st = set_device(c);
    if (st) return st;
    // refresh frames and sun for the current (possibly perturbed) parameters, keeping the mask
    const std::vector<AdjHelio> keep = s->h_ah;
    if ((st = adj_prepare(c, grid))) return st;
    for (size_t h = 0; h < keep.size(); ++h)
        if (keep[h].ray_base != s->h_ah[h].ray_base)
            return fail(QMC_E_STATE, "qmc_adjoint_forward_masked: field layout changed since the forward pass");
    const RecvBuf &rb = c->recvs[c->grids[grid].receiver];
    const size_t npix = (size_t)rb.w * rb.h;
    if (n < (int64_t)npix)
        return fail(QMC_E_RANGE, "qmc_adjoint_forward_masked: output holds " + std::to_string(n) + " of " +
                                     std::to_string(npix) + " pixels");
    if (npix > s->flux64_cap) {
        if ((st = dev_alloc(s->d_flux64, npix))) return st;
        s->flux64_cap = npix;
    }
    HIP_TRY(hipMemsetAsync(s->d_flux64, 0, sizeof(double) * npix, c->stream));
    hipLaunchKernelGGL(k_adj_forward_masked, adj_grid(c, s), dim3(256), 0, c->stream, adj_args(c, s, nullptr));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(flux_out, s->d_flux64, sizeof(double) * npix, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return QMC_OK;
}

extern "C" qmc_status qmc_flux_loss(qmc_ctx *c, const float *d_flux, int32_t n, float target, float peak,
                                    float lambda, float *d_grad, double *loss) {
    if (!c || !d_flux || !loss || n <= 0) return fail(QMC_E_INVALID_ARG, "qmc_flux_loss: bad argument");
    qmc_status st = set_device(c);
    if (st) return st;
    AdjState *s = state(c);
    const int blocks = (n + 255) / 256;
    if ((size_t)n != s->npix) {
        if ((st = dev_alloc(s->d_G, (size_t)n))) return st;
        if ((st = dev_alloc(s->d_partial, (size_t)blocks))) return st;
        s->npix = (size_t)n;
    }
    float *G = d_grad ? d_grad : s->d_G;
    hipLaunchKernelGGL(k_flux_loss, dim3(blocks), dim3(256), 0, c->stream, d_flux, n, target, peak, lambda, G,
                       s->d_partial);
    HIP_TRY(hipGetLastError());
    if (d_grad && d_grad != s->d_G)
        HIP_TRY(hipMemcpyAsync(s->d_G, d_grad, sizeof(float) * n, hipMemcpyDeviceToDevice, c->stream));
    std::vector<double> part(blocks);
    HIP_TRY(hipMemcpyAsync(part.data(), s->d_partial, sizeof(double) * blocks, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    double l = 0.0;
    for (double p : part) l += p;
    *loss = l;
    return QMC_OK;
}

extern "C" qmc_status qmc_adjoint_backward(qmc_ctx *c, const qmc_scene *scene, int32_t grid,
                                           const float *d_dL_dflux, float *grad_aim, float *grad_sun) {
    if (!c || !scene || !grad_aim || !grad_sun) return fail(QMC_E_INVALID_ARG, "qmc_adjoint_backward: null argument");
    AdjState *s = state(c);
    if (!s->mask_ready || s->grid != grid)
        return fail(QMC_E_STATE, "qmc_adjoint_backward: run qmc_adjoint_forward on this grid first");
    if (!scene->processed || scene->heliostats.size() != c->h_helios.size())
        return fail(QMC_E_STATE, "qmc_adjoint_backward: scene does not match the uploaded one");
    qmc_status st = set_device(c);
    if (st) return st;
    if (d_dL_dflux && d_dL_dflux != s->d_G)
        HIP_TRY(hipMemcpyAsync(s->d_G, d_dL_dflux, sizeof(float) * s->npix, hipMemcpyDeviceToDevice, c->stream));
    const dim3 g = adj_grid(c, s);
    hipLaunchKernelGGL(k_adj_backward, g, dim3(256), 0, c->stream, adj_args(c, s, nullptr));
    HIP_TRY(hipGetLastError());
    const int nl = (int)s->list.size();
    const size_t nh3 = 3 * s->h_ah.size();
    // every list entry's dL/dn is written; other heliostats' entries are never read
    hipLaunchKernelGGL(k_adj_reduce, dim3((unsigned)nl), dim3(64), 0, c->stream, s->d_gpart, (int)g.x, s->d_list,
                       s->d_gout, s->d_gout + nh3);
    HIP_TRY(hipGetLastError());
    if ((st = grow_pinned(s->p_out, s->p_out_cap, nh3 + 3 * (size_t)nl))) return st;
    HIP_TRY(hipMemcpyAsync(s->p_out, s->d_gout, sizeof(double) * (nh3 + 3 * (size_t)nl), hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    const double *gn = s->p_out, *gml = s->p_out + nh3;
    double gm[3] = {0.0, 0.0, 0.0};
    for (int li = 0; li < nl; ++li)
        for (int k = 0; k < 3; ++k) gm[k] += gml[3 * (size_t)li + k];
    // chain n_h = normalize(normalize(A - p) - s), A = focus + o0*across + o1*up, in double
    double gs[3] = {-gm[0], -gm[1], -gm[2]};
    const int nh = (int)scene->heliostats.size();
    for (int h = 0; h < nh; ++h) {
        grad_aim[2 * h] = 0.0f;
        grad_aim[2 * h + 1] = 0.0f;
    }
    const DevReceiver &R = c->recvs[c->grids[grid].receiver].r;
    // unit edge directions (v3 - v0 across, v1 - v0 up), rounded as the aim points were built
    const V3 across = normalize(V3{R.E2[0], R.E2[1], R.E2[2]}), up = normalize(V3{R.E1[0], R.E1[1], R.E1[2]});
    for (int32_t h : s->list) {
        double la, lb;
        H3 rr;
        const H3 n = helio_normal_f64(&c->h_aim64[3 * h], c->h_helios[h].pos, c->cfg.sun_dir, la, lb, rr);
        const H3 g{gn[3 * h], gn[3 * h + 1], gn[3 * h + 2]};
        const double ng = hdot(n, g);
        const H3 gb{(g.x - n.x * ng) / lb, (g.y - n.y * ng) / lb, (g.z - n.z * ng) / lb};
        gs[0] -= gb.x;
        gs[1] -= gb.y;
        gs[2] -= gb.z;
        const double rg = hdot(rr, gb);
        const H3 ga{(gb.x - rr.x * rg) / la, (gb.y - rr.y * rg) / la, (gb.z - rr.z * rg) / la};
        grad_aim[2 * h] = (float)(ga.x * across.x + ga.y * across.y + ga.z * across.z);
        grad_aim[2 * h + 1] = (float)(ga.x * up.x + ga.y * up.y + ga.z * up.z);
    }
    for (int k = 0; k < 3; ++k) grad_sun[k] = (float)gs[k];
    return QMC_OK;
}
