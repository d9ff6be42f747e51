// qmc_adjoint.hip — differentiable trace: bilinear flux deposit, visibility masks, the flux-map loss
// and a hand-written adjoint giving d(loss)/d(per-heliostat aim-point offsets, sun direction).
// No reference counterpart (the reference only renders flux maps); the per-ray forward model is the
// reference's map_raytracing (QuasiMonteCarloRayTracing/RectangleReceiverRectangleGridRayTracing.cu:
// 6-41) with the nearest-pixel deposit replaced by a bilinear one.
//
// Differentiated model of one visible ray of heliostat h (flat heliostats, rectangle receiver):
//   n_h   = normalize(normalize(A_h - p_h) - s),  A_h = focus + o0*across + o1*up   (host, double)
//   (u,n,v) frame of n_h; O = p_h + lf.x*u + lf.y*n + lf.z*v (micro-facet origin, lf fixed)
//   d_in  = R(-s) q_in  (sun-shape sample), N = normalize(R(n_h) q_p) (slope-error sample)
//   d     = normalize(reflect(-d_in, N)); (t, uu, vv) = Moeller-Trumbore hit of the receiver plane
//   E     = |d.N| eta(t) factor;  splat E with bilinear weights at (vv*res_x - 0.5, uu*res_y - 0.5)
// Visibility (not shadowed, not blocked, inside the receiver) is decided by the forward pass and then
// frozen, so the gradient is the derivative of the masked sum. Rays are never stored: the backward
// kernel regenerates each ray from its pool indices (the QMC sample offsets) and recomputes it.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "qmc_ctx.h"
#include "qmc_trace_core.h"

using namespace qmc;

namespace {

struct AdjHelio {
    float u[3], n[3], v[3];
    float nw1, nw2;   // |Y x n|, |u x n|
    int64_t ray_base; // first mask bit of this heliostat (multiple of 64)
};

struct AdjSun {
    float us[3], m[3], vs[3];  // frame of m = -s
    float nw1, nw2;
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
    float *flux;          // bilinear map (forward)
    uint32_t *mask;       // visibility bits
    const float *G;       // dL/dflux (backward)
    float *g_n;           // per heliostat dL/dn (3 floats), indexed by heliostat id
    float *g_m;           // dL/dm (3 floats), m = -s
};

__device__ __forceinline__ V3 ld3(const float *p) { return V3{p[0], p[1], p[2]}; }

__device__ __forceinline__ void bilinear_splat(float *flux, int resx, int resy, float uu, float vv, float E) {
    const float x = vv * (float)resx - 0.5f, y = uu * (float)resy - 0.5f;
    const float fx0 = floorf(x), fy0 = floorf(y);
    const int i0 = (int)fx0, j0 = (int)fy0;
    const float fx = x - fx0, fy = y - fy0;
    const float wx[2] = {1.0f - fx, fx}, wy[2] = {1.0f - fy, fy};
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const int i = i0 + a, j = j0 + b;
            if (i >= 0 && i < resx && j >= 0 && j < resy) unsafeAtomicAdd(flux + j * resx + i, E * wx[a] * wy[b]);
        }
}

// Receiver-plane hit without the [0,1] acceptance test (visibility is frozen by the mask).
__device__ __forceinline__ void plane_hit(const DevReceiver &R, V3 O, V3 d, float &t, float &uu, float &vv) {
    const V3 A0{R.A[0], R.A[1], R.A[2]}, E1{R.E1[0], R.E1[1], R.E1[2]}, E2{R.E2[0], R.E2[1], R.E2[2]};
    const V3 pv = cross(d, E2);
    const float det = dot(E1, pv);
    const V3 T = O - A0;
    uu = dot(T, pv) / det;
    const V3 q = cross(T, E1);
    vv = dot(d, q) / det;
    t = dot(E2, q) / det;
}

// Forward with full visibility: bilinear deposit + visibility bit per ray.
__global__ __launch_bounds__(256) void k_adj_forward(AdjArgs a) {
    const int h = a.list[blockIdx.y];
    const DevHelio &H = a.helios[h];
    const int L = a.L, P = a.P;
    const long long nrays = (long long)H.F * L;
    const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    bool visible = false;
    if (r < nrays) {
        const int m = (int)(r / L), j = (int)(r % L);
        const int32_t *grp = a.offsets + H.offsets_base;
        const float4 s = a.pool[(grp[m] + j) % P];
        V3 dir = fr_apply(a.sunf, V3{s.x, s.y, s.z});
        const V3 origin = flat_facet_origin(H, H.fr, m);
        RayStats st{0u, 0u, 0u};
        int hit;
        const int lo = H.sub_begin, hi = H.sub_begin + H.sub_count;
        if (!dda_collision<false>(origin, dir, a.grid, a.grid.quads, lo, hi, hit, st)) {
            const float4 q = a.pert[(grp[(m + 1) % P] + j) % P];
            const V3 normal = normalize(fr_apply(H.fr, V3{q.x, q.y, q.z}));
            dir = normalize(reflect(-dir, normal));
            if (!dda_collision<false>(origin, dir, a.grid, a.grid.quads, lo, hi, hit, st)) {
                float t, u, v;
                if (receiver_hit<0>(a.recv, origin, dir, t, u, v)) {
                    visible = true;
                    bilinear_splat(a.flux, a.recv.res_x, a.recv.res_y, u, v, ray_energy(t, dir, normal, H.factor));
                }
            }
        }
    }
    // ray_base is a multiple of 64 and blocks hold whole waves: each wave owns two mask words
    const unsigned long long bits = __ballot(visible);
    const int lane = threadIdx.x & 63;
    if ((lane & 31) == 0 && r - lane < nrays) {
        const long long bit = a.ah[h].ray_base + r;
        a.mask[bit >> 5] = (uint32_t)(bits >> lane);
    }
}

// Forward of the frozen-visibility model (finite differences with common random numbers).
__global__ __launch_bounds__(256) void k_adj_forward_masked(AdjArgs a) {
    const int h = a.list[blockIdx.y];
    const DevHelio &H = a.helios[h];
    const int L = a.L, P = a.P;
    const long long nrays = (long long)H.F * L;
    const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrays) return;
    const long long bit = a.ah[h].ray_base + r;
    if (!((a.mask[bit >> 5] >> (bit & 31)) & 1u)) return;
    const int m = (int)(r / L), j = (int)(r % L);
    const int32_t *grp = a.offsets + H.offsets_base;
    const float4 s = a.pool[(grp[m] + j) % P];
    V3 dir = fr_apply(a.sunf, V3{s.x, s.y, s.z});
    const V3 origin = flat_facet_origin(H, H.fr, m);
    const float4 q = a.pert[(grp[(m + 1) % P] + j) % P];
    const V3 normal = normalize(fr_apply(H.fr, V3{q.x, q.y, q.z}));
    dir = normalize(reflect(-dir, normal));
    float t, uu, vv;
    plane_hit(a.recv, origin, dir, t, uu, vv);
    bilinear_splat(a.flux, a.recv.res_x, a.recv.res_y, uu, vv, ray_energy(t, dir, normal, H.factor));
}

__device__ __forceinline__ V3 norm_bwd(V3 y, float len, V3 g) {  // y = x/|x|
    return (g - y * dot(y, g)) * (1.0f / len);
}

__device__ __forceinline__ float block_sum(float v, float *red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    float s = 0.0f;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += red[k];
    return s;
}

__global__ __launch_bounds__(256) void k_adj_backward(AdjArgs a) {
    __shared__ float red[16];
    const int h = a.list[blockIdx.y];
    const DevHelio &H = a.helios[h];
    const int L = a.L, P = a.P;
    const long long nrays = (long long)H.F * L;
    const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    V3 gn{0, 0, 0}, gm{0, 0, 0};
    bool vis = false;
    if (r < nrays) {
        const long long bit = a.ah[h].ray_base + r;
        vis = (a.mask[bit >> 5] >> (bit & 31)) & 1u;
    }
    if (vis) {
        const AdjHelio &Ah = a.ah[h];
        const V3 u = ld3(Ah.u), n = ld3(Ah.n), v = ld3(Ah.v);
        const V3 us = ld3(a.sun.us), mm = ld3(a.sun.m), vs = ld3(a.sun.vs);
        const int m = (int)(r / L), j = (int)(r % L);
        const int32_t *grp = a.offsets + H.offsets_base;
        const float4 qi4 = a.pool[(grp[m] + j) % P];
        const float4 qp4 = a.pert[(grp[(m + 1) % P] + j) % P];
        const V3 qi{qi4.x, qi4.y, qi4.z}, qp{qp4.x, qp4.y, qp4.z};
        // local facet coordinates (RectangleHelio.cu:64-79)
        const int cpr = H.cols * H.sub_cols;
        const int row = m / cpr, col = m % cpr;
        const float pl = H.pl;
        const V3 lf{col * pl + (col / H.sub_cols) * H.gap[0] + pl / 2 - H.size[0] / 2, H.size[1] / 2,
                    row * pl + (row / H.sub_rows) * H.gap[1] + pl / 2 - H.size[2] / 2};
        // ---- forward ----
        const V3 O = ld3(H.pos) + u * lf.x + n * lf.y + v * lf.z;
        const V3 din = us * qi.x + mm * qi.y + vs * qi.z;
        const V3 Nraw = u * qp.x + n * qp.y + v * qp.z;
        const float lN = length(Nraw);
        const V3 N = Nraw * (1.0f / lN);
        const V3 ii = -din;
        const float k = dot(N, ii);
        const V3 rr = ii - N * (2.0f * k);
        const float lr = length(rr);
        const V3 d = rr * (1.0f / lr);
        const DevReceiver &R = a.recv;
        const V3 A0{R.A[0], R.A[1], R.A[2]}, E1{R.E1[0], R.E1[1], R.E1[2]}, E2{R.E2[0], R.E2[1], R.E2[2]};
        const V3 pv = cross(d, E2);
        const float det = dot(E1, pv);
        const V3 T = O - A0;
        const float uu = dot(T, pv) / det;
        const V3 q = cross(T, E1);
        const float vv = dot(d, q) / det;
        const float t = dot(E2, q) / det;
        const float c = dot(d, N);
        float eta, deta;
        if (t <= 1000.0f) {
            eta = 0.99331f - 0.0001176f * t + 1.97e-8f * t * t;
            deta = -0.0001176f + 2.0f * 1.97e-8f * t;
        } else {
            eta = expf(-0.0001106f * t);
            deta = -0.0001106f * eta;
        }
        const float f = H.factor;
        const float E = fabsf(c) * eta * f;
        const int resx = R.res_x, resy = R.res_y;
        const float x = vv * (float)resx - 0.5f, y = uu * (float)resy - 0.5f;
        const float fx0 = floorf(x), fy0 = floorf(y);
        const int i0 = (int)fx0, j0 = (int)fy0;
        const float fx = x - fx0, fy = y - fy0;
        const float wx[2] = {1.0f - fx, fx}, wy[2] = {1.0f - fy, fy}, dw[2] = {-1.0f, 1.0f};
        float gE = 0.0f, gx = 0.0f, gy = 0.0f;
#pragma unroll
        for (int bb = 0; bb < 2; ++bb)
#pragma unroll
            for (int aa = 0; aa < 2; ++aa) {
                const int pi = i0 + aa, pj = j0 + bb;
                if (pi >= 0 && pi < resx && pj >= 0 && pj < resy) {
                    const float Gp = a.G[pj * resx + pi];
                    gE += Gp * wx[aa] * wy[bb];
                    gx += Gp * E * dw[aa] * wy[bb];
                    gy += Gp * E * wx[aa] * dw[bb];
                }
            }
        // ---- backward ----
        const float g_vv = gx * (float)resx, g_uu = gy * (float)resy;
        const float g_c = gE * (c >= 0.0f ? 1.0f : -1.0f) * eta * f;
        const float g_t = gE * fabsf(c) * f * deta;
        V3 g_d = N * g_c;
        V3 g_N = d * g_c;
        const float inv_det = 1.0f / det;
        const float g_a1 = g_uu * inv_det, g_a2 = g_vv * inv_det, g_a3 = g_t * inv_det;
        const float g_det = -(g_uu * uu + g_vv * vv + g_t * t) * inv_det;
        V3 g_T = pv * g_a1;
        const V3 g_pv = T * g_a1 + E1 * g_det;
        g_d = g_d + cross(E2, g_pv) + q * g_a2;
        const V3 g_q = d * g_a2 + E2 * g_a3;
        g_T = g_T + cross(E1, g_q);
        const V3 g_O = g_T;
        const V3 g_r = norm_bwd(d, lr, g_d);
        const float grN = dot(N, g_r);
        const V3 g_i = g_r - N * (2.0f * grN);
        g_N = g_N - (g_r * k + ii * grN) * 2.0f;
        const V3 g_din = -g_i;
        const V3 g_Nraw = norm_bwd(N, lN, g_N);
        V3 g_u = g_Nraw * qp.x + g_O * lf.x;
        gn = g_Nraw * qp.y + g_O * lf.y;
        const V3 g_v = g_Nraw * qp.z + g_O * lf.z;
        // heliostat frame: w1 = Y x n, u = w1/|w1|, w2 = u x n, v = w2/|w2|
        const V3 Y{0.0f, 1.0f, 0.0f};
        const V3 g_w2 = norm_bwd(v, Ah.nw2, g_v);
        g_u = g_u + cross(n, g_w2);
        gn = gn + cross(g_w2, u);
        const V3 g_w1 = norm_bwd(u, Ah.nw1, g_u);
        gn = gn + cross(g_w1, Y);
        // sun frame (m = -s)
        const V3 g_us = g_din * qi.x, g_vs = g_din * qi.z;
        gm = g_din * qi.y;
        const V3 g_ws2 = norm_bwd(vs, a.sun.nw2, g_vs);
        const V3 g_us2 = g_us + cross(mm, g_ws2);
        gm = gm + cross(g_ws2, us);
        const V3 g_ws1 = norm_bwd(us, a.sun.nw1, g_us2);
        gm = gm + cross(g_ws1, Y);
    }
    // one block = one heliostat: reduce, one atomic per component
    const float s0 = block_sum(gn.x, red), s1 = block_sum(gn.y, red), s2 = block_sum(gn.z, red);
    const float s3 = block_sum(gm.x, red), s4 = block_sum(gm.y, red), s5 = block_sum(gm.z, red);
    if (threadIdx.x == 0) {
        unsafeAtomicAdd(a.g_n + 3 * h, s0);
        unsafeAtomicAdd(a.g_n + 3 * h + 1, s1);
        unsafeAtomicAdd(a.g_n + 3 * h + 2, s2);
        unsafeAtomicAdd(a.g_m, s3);
        unsafeAtomicAdd(a.g_m + 1, s4);
        unsafeAtomicAdd(a.g_m + 2, s5);
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
    int32_t *d_list = nullptr;
    uint32_t *d_mask = nullptr;
    size_t mask_words = 0;
    float *d_gn = nullptr;
    float *d_gm = nullptr;
    float *d_G = nullptr;
    double *d_partial = nullptr;
    size_t npix = 0;
    bool mask_ready = false;
    AdjSun sun{};
    int maxF = 0;
};

AdjState *state(qmc_ctx *c) {
    if (!c->adj) c->adj = new AdjState();
    return static_cast<AdjState *>(c->adj);
}

void frame_norms(V3 n, float &nw1, float &nw2, V3 &u, V3 &v) {
    const V3 w1 = cross(V3{0.0f, 1.0f, 0.0f}, n);
    nw1 = length(w1);
    u = normalize(w1);
    const V3 w2 = cross(u, n);
    nw2 = length(w2);
    v = normalize(w2);
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
    s->maxF = 0;
    for (size_t h = 0; h < c->h_helios.size(); ++h) {
        const DevHelio &H = c->h_helios[h];
        if (H.grid != grid) continue;
        if (H.type != 0 || H.sub_count != 1) return fail(QMC_E_UNSUPPORTED, "adjoint: flat heliostats only");
        s->list.push_back((int32_t)h);
        AdjHelio &A = s->h_ah[h];
        const V3 n{H.fr.nx, H.fr.ny, H.fr.nz};
        V3 u, v;
        frame_norms(n, A.nw1, A.nw2, u, v);
        A.u[0] = u.x; A.u[1] = u.y; A.u[2] = u.z;
        A.n[0] = n.x; A.n[1] = n.y; A.n[2] = n.z;
        A.v[0] = v.x; A.v[1] = v.y; A.v[2] = v.z;
        A.ray_base = base;
        base += (((int64_t)H.F * c->L + 63) / 64) * 64;
        s->maxF = std::max(s->maxF, H.F);
    }
    const V3 mvec{-c->cfg.sun_dir[0], -c->cfg.sun_dir[1], -c->cfg.sun_dir[2]};
    V3 us, vs;
    frame_norms(mvec, s->sun.nw1, s->sun.nw2, us, vs);
    s->sun.us[0] = us.x; s->sun.us[1] = us.y; s->sun.us[2] = us.z;
    s->sun.m[0] = mvec.x; s->sun.m[1] = mvec.y; s->sun.m[2] = mvec.z;
    s->sun.vs[0] = vs.x; s->sun.vs[1] = vs.y; s->sun.vs[2] = vs.z;
    qmc_status st;
    const size_t words = (size_t)(base / 32) + 2;
    if (words > s->mask_words) {
        if ((st = dev_alloc(s->d_mask, words))) return st;
        s->mask_words = words;
    }
    if ((st = dev_alloc(s->d_ah, s->h_ah.size()))) return st;
    if ((st = dev_alloc(s->d_list, std::max<size_t>(s->list.size(), 1)))) return st;
    if ((st = dev_alloc(s->d_gn, 3 * s->h_ah.size()))) return st;
    if (!s->d_gm && (st = dev_alloc(s->d_gm, 3))) return st;
    const size_t npix = (size_t)rb.w * rb.h;
    if (npix != s->npix) {
        if ((st = dev_alloc(s->d_G, npix))) return st;
        if ((st = dev_alloc(s->d_partial, (npix + 255) / 256))) return st;
        s->npix = npix;
    }
    HIP_TRY(hipMemcpyAsync(s->d_ah, s->h_ah.data(), sizeof(AdjHelio) * s->h_ah.size(), hipMemcpyHostToDevice,
                           c->stream));
    if (!s->list.empty())
        HIP_TRY(hipMemcpyAsync(s->d_list, s->list.data(), sizeof(int32_t) * s->list.size(), hipMemcpyHostToDevice,
                               c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
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
    a.P = c->P;
    a.L = c->L;
    a.flux = flux;
    a.mask = s->d_mask;
    a.G = s->d_G;
    a.g_n = s->d_gn;
    a.g_m = s->d_gm;
    return a;
}

dim3 adj_grid(qmc_ctx *c, AdjState *s) {
    const long long rays = (long long)s->maxF * c->L;
    return dim3((unsigned)((rays + 255) / 256), (unsigned)s->list.size());
}

}  // namespace

namespace qmc {
void qmc_adjoint_release(qmc_ctx *c) {
    if (!c->adj) return;
    AdjState *s = static_cast<AdjState *>(c->adj);
    dev_free(s->d_ah);
    dev_free(s->d_list);
    dev_free(s->d_mask);
    dev_free(s->d_gn);
    dev_free(s->d_gm);
    dev_free(s->d_G);
    dev_free(s->d_partial);
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
    HIP_TRY(hipMemsetAsync(d_flux, 0, sizeof(float) * (size_t)rb.w * rb.h, c->stream));
    HIP_TRY(hipMemsetAsync(s->d_mask, 0, sizeof(uint32_t) * s->mask_words, c->stream));
    hipLaunchKernelGGL(k_adj_forward, adj_grid(c, s), dim3(256), 0, c->stream, adj_args(c, s, d_flux));
    HIP_TRY(hipGetLastError());
    s->mask_ready = true;
    return QMC_OK;
}

extern "C" qmc_status qmc_adjoint_forward_masked(qmc_ctx *c, int32_t grid, float *d_flux) {
    if (!c || !d_flux) return fail(QMC_E_INVALID_ARG, "qmc_adjoint_forward_masked: null argument");
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
    HIP_TRY(hipMemsetAsync(d_flux, 0, sizeof(float) * (size_t)rb.w * rb.h, c->stream));
    hipLaunchKernelGGL(k_adj_forward_masked, adj_grid(c, s), dim3(256), 0, c->stream, adj_args(c, s, d_flux));
    HIP_TRY(hipGetLastError());
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
    HIP_TRY(hipMemsetAsync(s->d_gn, 0, sizeof(float) * 3 * s->h_ah.size(), c->stream));
    HIP_TRY(hipMemsetAsync(s->d_gm, 0, sizeof(float) * 3, c->stream));
    hipLaunchKernelGGL(k_adj_backward, adj_grid(c, s), dim3(256), 0, c->stream, adj_args(c, s, nullptr));
    HIP_TRY(hipGetLastError());
    std::vector<float> gn(3 * s->h_ah.size());
    float gm[3];
    HIP_TRY(hipMemcpyAsync(gn.data(), s->d_gn, sizeof(float) * gn.size(), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(gm, s->d_gm, sizeof(float) * 3, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    // chain n_h = normalize(normalize(A - p) - s), A = focus + o0*across + o1*up (double precision)
    const double sx = c->cfg.sun_dir[0], sy = c->cfg.sun_dir[1], sz = c->cfg.sun_dir[2];
    double gs[3] = {-(double)gm[0], -(double)gm[1], -(double)gm[2]};
    const int nh = (int)scene->heliostats.size();
    for (int h = 0; h < nh; ++h) {
        grad_aim[2 * h] = 0.0f;
        grad_aim[2 * h + 1] = 0.0f;
    }
    const HostReceiver &R = scene->receivers[c->grids[grid].receiver];
    const V3 across = normalize(R.vert[3] - R.vert[0]), up = normalize(R.vert[1] - R.vert[0]);
    for (int32_t h : s->list) {
        const HostHeliostat &hh = scene->heliostats[h];
        const double p[3] = {hh.d.pos[0], hh.d.pos[1], hh.d.pos[2]};
        const double A[3] = {hh.aim.x, hh.aim.y, hh.aim.z};
        double a[3] = {A[0] - p[0], A[1] - p[1], A[2] - p[2]};
        const double la = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
        const double rr[3] = {a[0] / la, a[1] / la, a[2] / la};
        const double b[3] = {rr[0] - sx, rr[1] - sy, rr[2] - sz};
        const double lb = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
        const double n[3] = {b[0] / lb, b[1] / lb, b[2] / lb};
        const double g[3] = {gn[3 * h], gn[3 * h + 1], gn[3 * h + 2]};
        const double ng = n[0] * g[0] + n[1] * g[1] + n[2] * g[2];
        double gb[3];
        for (int k = 0; k < 3; ++k) gb[k] = (g[k] - n[k] * ng) / lb;
        for (int k = 0; k < 3; ++k) gs[k] -= gb[k];
        const double rg = rr[0] * gb[0] + rr[1] * gb[1] + rr[2] * gb[2];
        double ga[3];
        for (int k = 0; k < 3; ++k) ga[k] = (gb[k] - rr[k] * rg) / la;
        grad_aim[2 * h] = (float)(ga[0] * across.x + ga[1] * across.y + ga[2] * across.z);
        grad_aim[2 * h + 1] = (float)(ga[0] * up.x + ga[1] * up.y + ga[2] * up.z);
    }
    for (int k = 0; k < 3; ++k) grad_sun[k] = (float)gs[k];
    return QMC_OK;
}
