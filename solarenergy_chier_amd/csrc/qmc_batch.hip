// qmc_batch.hip — batched traces: many sun positions x aim strategies of one heliostat field in
// persistent launches, each instance with its own flux map (no counterpart in the reference, which
// traces one configuration per process; its per-ray work is map_raytracing,
// QuasiMonteCarloRayTracing/RectangleReceiverRectangleGridRayTracing.cu:6-41).
//
// Two kernels per chunk of instances. k_batch_setup recomputes, for every (instance, heliostat), what
// depends on the sun and the aim point — heliostat normals and frames (Heliostat::CSetNormal,
// Heliostat.cu:15-20), blocking quads (CSetWorldVertex, Heliostat.cu:25-40), energy factors
// (QuasiMonteCarloRayTracer.cpp:58-60) — with the same float operations the host processor uses, into
// per-instance arrays in HBM. k_trace_batched is persistent: a workgroup owns whole instances, walks
// the instance's heliostats in order (their frames and factors are wave-uniform scalar loads) and
// traces every ray with the shared per-ray functions, accumulating the instance's map in LDS (one
// plain store per pixel at the end) or, when the map does not fit or the schedule says so, with
// workgroup-scope atomics into the instance's map in HBM (no other workgroup touches it, so the
// atomics complete in the XCD's L2). Grid cell lists, sample pools and micro-facet offsets are shared
// by all instances.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "qmc_ctx.h"
#include "qmc_trace_core.h"

using namespace qmc;

namespace {

// Instances whose per-heliostat data (frames, factors, quads: ~92 B per heliostat) is resident at
// once; the chunk is as large as this budget allows, so a sweep normally runs as one launch pair.
constexpr size_t kWorkBudget = size_t(2) << 30;

struct BatchArgs {
    DevGrid grid;
    DevReceiver recv;
    const DevHelio *__restrict__ helios;
    int32_t helio_begin, n_helio, n_total_helio, n_sub;
    const float *__restrict__ aims;  // [set][heliostat][3]
    const qmc_instance *__restrict__ inst;
    int32_t n_inst;
    const float4 *__restrict__ pool;
    const float4 *__restrict__ pert;
    const int32_t *__restrict__ offsets;
    int32_t P, L;
    FastDiv fdP, fdL;
    int32_t lLs;  // log2 L for the facet-shared order (0: it does not apply)
    double tan_sun, tan_block;  // traversal-culling cone tangents (negative: no culling)
    float rho;
    float *maps;
    int32_t npix;
    // per-instance data of the chunk (written by k_batch_setup, read by k_trace_batched)
    DevFrame *sun_fr;  // [instance]
    DevFrame *fr;      // [instance][heliostat]
    float *fac;        // [instance][heliostat]
    float4 *quads;     // [instance][sub-heliostat][3]
    CullList *cull;    // [instance][heliostat]: candidate quads of each traversal (null: no culling)
    int32_t *queue;    // next instance to take (reset by k_batch_setup)
};

__device__ __forceinline__ DevFrame dev_frame(const Frame &f) {
    DevFrame d;
    d.ux = f.u.x; d.uy = f.u.y; d.uz = f.u.z;
    d.nx = f.n.x; d.ny = f.n.y; d.nz = f.n.z;
    d.vx = f.v.x; d.vy = f.v.y; d.vz = f.v.z;
    d.identity = f.identity;
    return d;
}

__global__ __launch_bounds__(256) void k_batch_setup(BatchArgs a) {
    const int ii = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const qmc_instance I = a.inst[ii];
    const V3 sun{I.sun_dir[0], I.sun_dir[1], I.sun_dir[2]};
    if (k == 0) a.sun_fr[ii] = dev_frame(make_frame(-sun, false));
    if (k == 0 && ii == 0) *a.queue = 0;
    if (k >= a.n_helio) return;
    const DevHelio &H = a.helios[a.helio_begin + k];
    const float *ap = a.aims + ((size_t)I.aim_set * a.n_total_helio + a.helio_begin + k) * 3;
    const V3 aim{ap[0], ap[1], ap[2]};
    const V3 pos{H.pos[0], H.pos[1], H.pos[2]};
    const V3 normal = normalize(normalize(aim - pos) - sun);
    const Frame f = make_frame(normal, false);
    a.fr[(size_t)ii * a.n_helio + k] = dev_frame(f);
    // flat heliostat blocking quad: world vertices v0, v1, v2 -> (A, E1, E2) = (v1, v0-v1, v2-v1)
    const V3 v0l{-H.size[0] / 2, H.size[1] / 2, -H.size[2] / 2};
    const V3 v1l = V3{0, 0, H.size[2]} + v0l;
    const V3 v2l = V3{H.size[0], 0, H.size[2]} + v0l;
    const V3 w0 = frame_apply(f, v0l) + pos, w1 = frame_apply(f, v1l) + pos, w2 = frame_apply(f, v2l) + pos;
    const V3 E1 = w0 - w1, E2 = w2 - w1;
    float4 *q = a.quads + ((size_t)ii * a.n_sub + H.sub_begin) * 3;
    q[0] = make_float4(w1.x, w1.y, w1.z, 0.0f);
    q[1] = make_float4(E1.x, E1.y, E1.z, 0.0f);
    q[2] = make_float4(E2.x, E2.y, E2.z, 0.0f);
    a.fac[(size_t)ii * a.n_helio + k] = H.fac_rule == 0 ? I.dni * H.fac_x * H.fac_x * a.rho / float(a.L)
                                                          : I.dni * H.fac_x * a.rho / float(a.L);
}

// Traversal culling of every (instance, heliostat), one wavefront each, from the frames and quads
// k_batch_setup wrote: the per-instance counterpart of the plain trace's k_cull. The wave first scans
// the instance's quads 64 at a time with the bounding-ball test alone, collecting the quads that pass
// in LDS (ballot and lane ranks give each its slot), then runs the cone test once over the collected
// quads, one per lane, and lists the candidates of each traversal. More than 64 ball-passing quads
// for a traversal keep it for every ray (counted as more than kCullList).
constexpr int kBallCap = 64;
__global__ __launch_bounds__(256) void k_batch_cull(BatchArgs a) {
    __shared__ int s_q[4][2][kBallCap];
    const int ii = blockIdx.y;
    const int wv = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int k = blockIdx.x * (blockDim.x / 64) + wv;
    if (k >= a.n_helio) return;
    const DevHelio &H = a.helios[a.helio_begin + k];
    const DevCullShape S = flat_cull_shape(H, a.fr[(size_t)ii * a.n_helio + k]);
    const DevFrame sunf = a.sun_fr[ii];
    const float4 *quads = a.quads + (size_t)ii * a.n_sub * 3;
    CullList *out = a.cull + (size_t)ii * a.n_helio + k;
    const int own_lo = H.sub_begin, own_hi = H.sub_begin + H.sub_count;
    const CD3 c{S.c[0], S.c[1], S.c[2]}, hu{S.hu[0], S.hu[1], S.hu[2]}, hn{0, 0, 0}, hv{S.hv[0], S.hv[1], S.hv[2]},
        nrm{S.n[0], S.n[1], S.n[2]};
    const CD3 ash = sunf.identity ? CD3{0, 1, 0} : CD3{sunf.nx, sunf.ny, sunf.nz};
    CD3 r0 = cadd(cmul(ash, -1.0), cmul(nrm, 2.0 * cdot(ash, nrm)));  // reflect(-sun axis, n)
    r0 = cmul(r0, 1.0 / sqrt(cdot(r0, r0)));
    const double rs = sqrt(cdot(hu, hu) + cdot(hv, hv));
    const uint64_t below = (1ull << lane) - 1ull;  // lanes ranked before this one
    int cs = 0, cb = 0;                             // wave-uniform counts of ball-passing quads
    for (int base = 0; base < a.n_sub && (cs <= kBallCap || cb <= kBallCap); base += 64) {
        const int q = base + lane;
        bool ps = false, pb = false;
        if (q < a.n_sub && (q < own_lo || q >= own_hi)) {
            const float4 qa = quads[3 * q], qb = quads[3 * q + 1], qc = quads[3 * q + 2];
            const CD3 A{qa.x, qa.y, qa.z}, E1{qb.x, qb.y, qb.z}, E2{qc.x, qc.y, qc.z};
            const CD3 sum = cadd(E1, E2), diff = csub(E1, E2);
            const CD3 D = csub(cadd(A, cmul(sum, 0.5)), c);
            const double R = rs + 0.5 * sqrt(fmax(cdot(sum, sum), cdot(diff, diff)));
            ps = a.tan_sun >= 0.0 && sphere_may_hit(D, R, ash, a.tan_sun, kCullMargin);
            pb = a.tan_block >= 0.0 && sphere_may_hit(D, R, r0, a.tan_block, kCullMargin);
        }
        const uint64_t bs = __ballot(ps), bb = __ballot(pb);
        const int rs_ = cs + __popcll(bs & below), rb_ = cb + __popcll(bb & below);
        if (ps && rs_ < kBallCap) s_q[wv][0][rs_] = q;
        if (pb && rb_ < kBallCap) s_q[wv][1][rb_] = q;
        cs += __popcll(bs);
        cb += __popcll(bb);
    }
    // the wave's own LDS lists: visible to its lanes once the wave's stores are done
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int nk[2] = {0, 0};
    for (int kind = 0; kind < 2; ++kind) {
        const int cnt = kind == 0 ? cs : cb;
        const double tn = kind == 0 ? a.tan_sun : a.tan_block;
        if (cnt > kBallCap) {
            nk[kind] = kCullList + 1;
            continue;
        }
        bool pass = false;
        int q = -1;
        if (lane < cnt) {
            q = s_q[wv][kind][lane];
            const float4 qa = quads[3 * q], qb = quads[3 * q + 1], qc = quads[3 * q + 2];
            pass = cone_may_hit_n<4>(c, hu, hn, hv, kind == 0 ? ash : r0, tn, CD3{qa.x, qa.y, qa.z},
                                     CD3{qb.x, qb.y, qb.z}, CD3{qc.x, qc.y, qc.z}, kCullMargin);
        }
        const uint64_t bp = __ballot(pass);
        const int rank = __popcll(bp & below);
        if (pass && rank < kCullList) out->ids[kind][rank] = q;
        nk[kind] = __popcll(bp);
    }
    if (lane == 0) {
        out->n[0] = nk[0];
        out->n[1] = nk[1];
    }
}

// 1024-thread workgroups are built for two per CU (8 waves per SIMD, at most 64 VGPRs): with a 64 KB LDS
// map each, that is the only way an LDS-accumulated instance runs at full occupancy.
// ORDER: how a heliostat's rays are dealt to lanes. 0: ray order (slot = ray id, the workgroup strides
// over them); 1: facet tiles (slot_ray's tiled order); 2: facet-shared (L = 2^a.lLs rays per facet,
// 2 <= L <= 64): each wave takes blocks of 64 micro-facets, lane k computes facet k's origin and group
// offsets once, and the block's L steps of 64 rays read them from the computing lanes (ds_bpermute).
template <int RT, bool LDSMAP, int ORDER, int BS>
__global__ __launch_bounds__(BS, BS == 1024 ? 8 : 1) void k_trace_batched(BatchArgs a) {
    extern __shared__ float map_lds[];
    __shared__ int s_next;
    const int tid = threadIdx.x, bs = BS;
    const int L = a.L, P = a.P;
    RayStats st{0u, 0u, 0u};
    // Constants every ray reads — the receiver's geometry, the instance's sun frame, the heliostat's
    // slope-error frame, energy factor and (per-ray orders) facet geometry — are kept in vector registers
    // (in_vgpr): in scalar registers they overflow the 106 available, and the
    // compiler spills the rest to lanes of vector registers and reloads kernel arguments inside the ray
    // loop. 1024-lane blocks (at most 64 vector registers) keep them scalar. Batched sweeps: config-1
    // field 64.1k -> 71.9k scenes/s, PS10 sampled 40.9k -> 45.0k.
    DevReceiver recv = a.recv;
    if (BS < 1024) {
        for (int k = 0; k < 3; ++k) {
            if (RT == 0) {
                in_vgpr(recv.A[k]);
                in_vgpr(recv.E1[k]);
                in_vgpr(recv.E2[k]);
            } else {
                in_vgpr(recv.pos[k]);
            }
        }
        if (RT == 1) {
            in_vgpr(recv.radius);
            in_vgpr(recv.height);
        }
    }

    for (;;) {
        // instances are taken from a queue, so uneven instance costs do not leave workgroups idle;
        // the index is read from LDS, so every wave of the workgroup leaves the loop together
        __syncthreads();  // the previous instance (and its map store) is done
        if (tid == 0) s_next = atomicAdd(a.queue, 1);
        __syncthreads();
        const int ii = s_next;
        if (ii >= a.n_inst) break;
        if (LDSMAP) {
            for (int p = tid; p < a.npix; p += bs) map_lds[p] = 0.0f;
            __syncthreads();
        }
        DevFrame sunf = ld_uniform(a.sun_fr + ii);
        if (BS < 1024) vgpr_frame(sunf);
        const float4 *quads = a.quads + (size_t)ii * a.n_sub * 3;
        float *gmap = a.maps + (size_t)ii * a.npix;

#pragma nounroll
        for (int k = 0; k < a.n_helio; ++k) {
            DevHelio H = ld_uniform(a.helios + a.helio_begin + k);
            if (BS < 1024 && ORDER != 2) {  // the per-ray facet origin's geometry
                for (int c = 0; c < 3; ++c) {
                    in_vgpr(H.pos[c]);
                    in_vgpr(H.size[c]);
                }
                in_vgpr(H.gap[0]);
                in_vgpr(H.gap[1]);
                in_vgpr(H.pl);
            }
            DevFrame fr = ld_uniform(a.fr + (size_t)ii * a.n_helio + k);
            float factor = ld_uniform(a.fac + (size_t)ii * a.n_helio + k);
            if (BS < 1024) {
                vgpr_frame(fr);
                in_vgpr(factor);
            }
            const int32_t *grp = a.offsets + H.offsets_base;
            const int own_lo = H.sub_begin, own_hi = H.sub_begin + H.sub_count;
            const CullList *cl = a.cull ? a.cull + (size_t)ii * a.n_helio + k : nullptr;
            // whether the traversal of kind `kind` (0 shadow, 1 blocking) hits a quad: none when the ray
            // misses every listed candidate; a hit when it hits one whose cell box the reference's walk
            // is proven to reach (known_hit_proven; every cell of a heliostat's bounding-sphere box lists
            // its quad, whatever the instance's orientation); otherwise the walk decides
            auto hits_quad = [&](int kind, V3 o, V3 d) -> bool {
                int hit;
                if (cl) {
                    const int n = ld_uniform(&cl->n[kind]);
                    if (n == 0) return false;
                    if (n <= kCullList) {
                        RayHits rh;
                        rh.n = 0;
                        for (int i = 0; i < n; ++i) {
                            const int id = ld_uniform(&cl->ids[kind][i]);
                            const float4 qa = ld_uniform(quads + 3 * id), qb = ld_uniform(quads + 3 * id + 1),
                                         qc = ld_uniform(quads + 3 * id + 2);
                            float t, u, v;
                            if (ray_quad(o, d, V3{qa.x, qa.y, qa.z}, V3{qb.x, qb.y, qb.z}, V3{qc.x, qc.y, qc.z}, t,
                                         u, v)) {
                                rh.n = 1, rh.id0 = id, rh.t0 = t;
                                break;
                            }
                        }
                        if (rh.n == 0) return false;
                        if (known_hit_proven(rh, o, d, a.grid, hit)) return true;
                    }
                }
                return dda_collision<false>(o, d, a.grid, quads, own_lo, own_hi, hit, st);
            };
            // one ray from its facet's origin and its two pool samples (map_raytracing,
            // RectangleReceiverRectangleGridRayTracing.cu:14-41)
            auto trace_one = [&](V3 origin, float4 s, float4 q) {
                V3 dir = fr_apply(sunf, V3{s.x, s.y, s.z});
                if (hits_quad(0, origin, dir)) return;
                const V3 normal = normalize(fr_apply(fr, V3{q.x, q.y, q.z}));
                dir = normalize(reflect(-dir, normal));
                if (hits_quad(1, origin, dir)) return;
                float t, u, v;
                int col = -1;  // set by a cylinder's column lookup
                if (!receiver_hit<RT>(recv, origin, dir, t, u, v, nullptr, &col)) return;
                const float energy = ray_energy(t, dir, normal, factor);
                int row = (int)(u * (float)a.recv.res_y);
                if (col < 0) col = (int)(v * (float)a.recv.res_x);
                if (row >= a.recv.res_y) row = a.recv.res_y - 1;
                if (col >= a.recv.res_x) col = a.recv.res_x - 1;
                const int addr = row * a.recv.res_x + col;
                if (LDSMAP)
                    atomicAdd(map_lds + addr, energy);
                else
                    __hip_atomic_fetch_add(gmap + addr, energy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            };
            if constexpr (ORDER == 2) {
                const int lLs = a.lLs, fpi = 64 >> lLs;  // log2 of rays per facet, facets per step
                const int lane = tid & 63, lfac = lane >> lLs, lray = lane & (L - 1);
                const int nblk = (H.F + 63) >> 6;
#pragma nounroll
                for (int blk = __builtin_amdgcn_readfirstlane(tid >> 6); blk < nblk; blk += bs / 64) {
                    const int f0 = blk * 64, nf = min(64, H.F - f0);
                    V3 fo{0.0f, 0.0f, 0.0f};
                    int g0 = 0, g1 = 0;  // offsets in [0, P]; 0 past the last facet
                    if (lane < nf) {
                        fo = flat_facet_origin(H, fr, f0 + lane);
                        g0 = grp[f0 + lane];
                        g1 = grp[fast_mod((uint32_t)(f0 + lane + 1), P, a.fdP)];
                    }
#pragma nounroll
                    for (int i = 0; i * fpi < nf; ++i) {
                        const int src = i * fpi + lfac;
                        const V3 origin{lane_read(fo.x, src), lane_read(fo.y, src), lane_read(fo.z, src)};
                        const uint32_t u1 = (uint32_t)(lane_read(g0, src) + lray),
                                       u2 = (uint32_t)(lane_read(g1, src) + lray);
                        if (src < nf) {
                            // (g + j) mod P for g in [0, P] and j < L <= P: one conditional subtraction
                            const float4 s = a.pool[min(u1, u1 - (uint32_t)P)];
                            const float4 q = a.pert[min(u2, u2 - (uint32_t)P)];
                            trace_one(origin, s, q);
                        }
                    }
                }
            } else {
                const int R = H.rows * H.sub_rows, C = H.cols * H.sub_cols;
                const FacetTiling tl = facet_tiling(R, C, L);
                const int nslots = ORDER == 1 ? tl.slots : H.F * L;
#pragma nounroll
                for (int slot = tid; slot < nslots; slot += bs) {
                    const int r = ORDER == 1 ? facet_tiled_ray(tl, tile_divisor(H, tl), R, C, slot) : slot;
                    if (r < 0) continue;
                    const int m = (int)fast_div((uint32_t)r, a.fdL), j = r - m * L;
                    // both samples are read before the tests: two dependent load pairs, not a chain of four
                    const int a1 = (int)fast_mod((uint32_t)(grp[m] + j), P, a.fdP);
                    const int a2 =
                        (int)fast_mod((uint32_t)(grp[fast_mod((uint32_t)(m + 1), P, a.fdP)] + j), P, a.fdP);
                    const float4 s = a.pool[a1];
                    const float4 q = a.pert[a2];
                    trace_one(flat_facet_origin(H, fr, m), s, q);
                }
            }
        }
        if (LDSMAP) {
            __syncthreads();
            for (int p = tid; p < a.npix; p += bs) gmap[p] = map_lds[p];
        }
    }
}

// ---------------------------------------------------------------------------------------------
// schedule cache (process-wide; optionally persisted to the file named by QMC_TUNE_CACHE)
// ---------------------------------------------------------------------------------------------
std::mutex g_tune_mu;
std::map<std::string, qmc_schedule> g_tune;
bool g_tune_loaded = false;

void tune_load_file() {
    if (g_tune_loaded) return;
    g_tune_loaded = true;
    const char *path = std::getenv("QMC_TUNE_CACHE");
    if (!path) return;
    std::ifstream f(path);
    std::string key;
    qmc_schedule s;
    while (f >> key >> s.block_threads >> s.workgroups_per_cu >> s.lds_flux_tile >> s.facet_tiles) g_tune[key] = s;
}

void tune_store(const std::string &key, const qmc_schedule &s) {
    g_tune[key] = s;
    const char *path = std::getenv("QMC_TUNE_CACHE");
    if (!path) return;
    std::ofstream f(path, std::ios::app);
    f << key << " " << s.block_threads << " " << s.workgroups_per_cu << " " << s.lds_flux_tile << " " << s.facet_tiles
      << "\n";
}

std::string shape_key(qmc_ctx *c, int grid, int npix, int n_helio, int rays_per_instance) {
    hipDeviceProp_t prop;
    std::string arch = "gpu";
    if (hipGetDeviceProperties(&prop, c->device) == hipSuccess) arch = prop.gcnArchName;
    for (auto &ch : arch)
        if (ch == ' ' || ch == ':') ch = '_';
    std::ostringstream o;
    o << arch << "_cu" << c->num_cus << "_h" << n_helio << "_r" << rays_per_instance << "_px" << npix << "_rt"
      << c->recvs[c->grids[grid].receiver].r.type;
    return o.str();
}


struct BatchPlan {
    BatchArgs args;
    int rt;
};

size_t lds_bytes(const BatchArgs &a, bool lds_map) { return lds_map ? sizeof(float) * (size_t)a.npix : 0; }

// Per-instance data of one chunk, carved out of one ctx-owned allocation.
size_t work_bytes_per_instance(const BatchArgs &a) {
    return sizeof(DevFrame) * (size_t)(a.n_helio + 1) + (sizeof(float) + sizeof(CullList)) * a.n_helio +
           sizeof(float4) * 3 * a.n_sub;
}

qmc_status ensure_work(qmc_ctx *c, BatchArgs &a, int chunk) {
    const size_t fr = sizeof(DevFrame) * (size_t)chunk * a.n_helio;
    const size_t fac = sizeof(float) * (size_t)chunk * a.n_helio;
    const size_t quads = sizeof(float4) * 3 * (size_t)chunk * a.n_sub;
    const size_t sun = sizeof(DevFrame) * (size_t)chunk;
    const size_t cull = sizeof(CullList) * (size_t)chunk * a.n_helio;
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t total = up(quads) + up(fr) + up(fac) + up(sun) + up(cull) + 256;
    if (total > c->batch_work_cap) {
        qmc_status st = dev_alloc(c->d_batch_work, total);
        if (st) {
            c->batch_work_cap = 0;
            return st;
        }
        c->batch_work_cap = total;
    }
    char *p = c->d_batch_work;
    a.quads = reinterpret_cast<float4 *>(p);
    p += up(quads);
    a.fr = reinterpret_cast<DevFrame *>(p);
    p += up(fr);
    a.fac = reinterpret_cast<float *>(p);
    p += up(fac);
    a.sun_fr = reinterpret_cast<DevFrame *>(p);
    p += up(sun);
    // culling needs both cone tangents; without them every traversal runs
    a.cull = (a.tan_sun >= 0.0 && a.tan_block >= 0.0) ? reinterpret_cast<CullList *>(p) : nullptr;
    p += up(cull);
    a.queue = reinterpret_cast<int32_t *>(p);
    return QMC_OK;
}

template <int RT, bool LDSMAP, int ORDER>
void launch_trace_bs(int bs, dim3 g, size_t lds, hipStream_t st, const BatchArgs &b) {
    if (bs == 256) hipLaunchKernelGGL((k_trace_batched<RT, LDSMAP, ORDER, 256>), g, dim3(256), lds, st, b);
    else if (bs == 512) hipLaunchKernelGGL((k_trace_batched<RT, LDSMAP, ORDER, 512>), g, dim3(512), lds, st, b);
    else hipLaunchKernelGGL((k_trace_batched<RT, LDSMAP, ORDER, 1024>), g, dim3(1024), lds, st, b);
}

template <int RT, bool LDSMAP>
void launch_trace_order(int order, int bs, dim3 g, size_t lds, hipStream_t st, const BatchArgs &b) {
    if (order == 2) launch_trace_bs<RT, LDSMAP, 2>(bs, g, lds, st, b);
    else if (order == 1) launch_trace_bs<RT, LDSMAP, 1>(bs, g, lds, st, b);
    else launch_trace_bs<RT, LDSMAP, 0>(bs, g, lds, st, b);
}

void launch_trace(int rt, bool lds_map, int order, int bs, dim3 g, size_t lds, hipStream_t st, const BatchArgs &b) {
    if (rt == 0) {
        if (lds_map) launch_trace_order<0, true>(order, bs, g, lds, st, b);
        else launch_trace_order<0, false>(order, bs, g, lds, st, b);
    } else {
        if (lds_map) launch_trace_order<1, true>(order, bs, g, lds, st, b);
        else launch_trace_order<1, false>(order, bs, g, lds, st, b);
    }
}

template <int RT, bool LDSMAP, int ORDER>
hipError_t occupancy_bs(int bs, size_t lds, int &n) {
    if (bs == 256)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_trace_batched<RT, LDSMAP, ORDER, 256>, 256, lds);
    if (bs == 512)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_trace_batched<RT, LDSMAP, ORDER, 512>, 512, lds);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_trace_batched<RT, LDSMAP, ORDER, 1024>, 1024, lds);
}

template <int RT, bool LDSMAP>
hipError_t occupancy_order(int order, int bs, size_t lds, int &n) {
    if (order == 2) return occupancy_bs<RT, LDSMAP, 2>(bs, lds, n);
    if (order == 1) return occupancy_bs<RT, LDSMAP, 1>(bs, lds, n);
    return occupancy_bs<RT, LDSMAP, 0>(bs, lds, n);
}

qmc_status resident_per_cu(int rt, bool lds_map, int order, int bs, size_t lds, int &n) {
    hipError_t e;
    if (rt == 0) e = lds_map ? occupancy_order<0, true>(order, bs, lds, n) : occupancy_order<0, false>(order, bs, lds, n);
    else e = lds_map ? occupancy_order<1, true>(order, bs, lds, n) : occupancy_order<1, false>(order, bs, lds, n);
    if (e != hipSuccess) return fail(QMC_E_HIP, std::string("occupancy query: ") + hipGetErrorString(e));
    if (n < 1) return fail(QMC_E_UNSUPPORTED, "batched trace: the schedule does not fit on a CU");
    return QMC_OK;
}

// The facet-shared order's log2 of rays per facet, or 0 when it does not apply (L a power of two from 2
// to 64, at most the pool size).
int shared_order_lL(const BatchArgs &a) {
    for (int l = 1; l <= 6; ++l)
        if (a.L == (1 << l)) return a.P >= a.L ? l : 0;
    return 0;
}

// Trace instances [0, n_inst) of the plan into plan.args.maps, chunk by chunk.
qmc_status launch_batched(qmc_ctx *c, const BatchPlan &plan, const qmc_schedule &s, int n_inst) {
    BatchArgs a = plan.args;
    const bool lds_map = s.lds_flux_tile != 0;
    const size_t lds = lds_bytes(a, lds_map);
    a.lLs = shared_order_lL(a);
    if (s.facet_tiles == 2 && !a.lLs)
        return fail(QMC_E_UNSUPPORTED, "batched trace: the facet-shared order needs 2 to 64 rays per facet (a power of two)");
    if (lds > c->max_lds) return fail(QMC_E_UNSUPPORTED, "batched trace: the flux map does not fit in LDS");
    const size_t budget = c->batch_work_budget ? c->batch_work_budget : kWorkBudget;
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>(n_inst, budget / work_bytes_per_instance(a)));
    qmc_status st = ensure_work(c, a, chunk);
    if (st) return st;
    if (!lds_map) HIP_TRY(hipMemsetAsync(a.maps, 0, sizeof(float) * (size_t)n_inst * a.npix, c->stream));
    for (int base = 0; base < n_inst; base += chunk) {
        BatchArgs b = a;
        b.n_inst = std::min(chunk, n_inst - base);
        b.inst = a.inst + base;
        b.maps = a.maps + (size_t)base * a.npix;
        hipLaunchKernelGGL(k_batch_setup, dim3((unsigned)((a.n_helio + 255) / 256), (unsigned)b.n_inst), dim3(256), 0,
                           c->stream, b);
        if (b.cull)
            hipLaunchKernelGGL(k_batch_cull, dim3((unsigned)((a.n_helio + 3) / 4), (unsigned)b.n_inst), dim3(256), 0,
                               c->stream, b);
        // no more workgroups than can be resident: a queue-driven persistent grid
        int resident = 0;
        const int order = s.facet_tiles;
        if ((st = resident_per_cu(plan.rt, lds_map, order, s.block_threads, lds, resident))) return st;
        const int per_cu = std::max(1, std::min(s.workgroups_per_cu, resident));
        const int grid = std::max(1, std::min(b.n_inst, c->num_cus * per_cu));
        launch_trace(plan.rt, lds_map, order, s.block_threads, dim3(grid), lds, c->stream, b);
        HIP_TRY(hipGetLastError());
    }
    return QMC_OK;
}

qmc_status plan_batch(qmc_ctx *c, int grid, const qmc_instance *inst, int n, BatchPlan &plan) {
    if (!c->uploaded) return fail(QMC_E_STATE, "batched trace: upload a scene first");
    if (!c->pool_ready || !c->offsets_ready)
        return fail(QMC_E_STATE, "batched trace: set or generate the sample pools and micro-facet offsets first");
    if (!c->d_aims || c->n_aim_sets <= 0) return fail(QMC_E_STATE, "batched trace: call qmc_batch_set_aims first");
    if (grid < 0 || grid >= (int)c->grids.size()) return fail(QMC_E_RANGE, "batched trace: grid out of range");
    if (n <= 0 || !inst) return fail(QMC_E_INVALID_ARG, "batched trace: empty instance list");
    for (int i = 0; i < n; ++i)
        if (inst[i].aim_set < 0 || inst[i].aim_set >= c->n_aim_sets)
            return fail(QMC_E_RANGE, "batched trace: instance " + std::to_string(i) + " has aim set " +
                                         std::to_string(inst[i].aim_set) + " of " + std::to_string(c->n_aim_sets));
    qmc_status st = device_limits(c);
    if (st) return st;
    const GridBuf &gb = c->grids[grid];
    const RecvBuf &rb = c->recvs[gb.receiver];
    int begin = -1, count = 0;
    for (size_t h = 0; h < c->h_helios.size(); ++h)
        if (c->h_helios[h].grid == grid) {
            if (begin < 0) begin = (int)h;
            ++count;
            const DevHelio &H = c->h_helios[h];
            if (H.type != 0 || H.sub_count != 1)
                return fail(QMC_E_UNSUPPORTED, "batched trace: flat heliostats only (one blocking quad each)");
        }
    if (count == 0) return fail(QMC_E_INVALID_ARG, "batched trace: grid has no heliostats");
    if (gb.n_sub != count) return fail(QMC_E_UNSUPPORTED, "batched trace: one blocking quad per heliostat expected");
    BatchArgs &a = plan.args;
    std::memset(&a, 0, sizeof(a));
    a.grid = gb.g;
    a.recv = rb.r;
    a.helios = c->d_helios;
    a.helio_begin = begin;
    a.n_helio = count;
    a.n_total_helio = (int)c->h_helios.size();
    a.n_sub = gb.n_sub;
    a.aims = c->d_aims;
    a.pool = c->d_pool;
    a.pert = c->d_pert;
    a.offsets = c->d_offsets;
    a.P = c->P;
    a.L = c->L;
    a.fdP = make_fastdiv((uint32_t)c->P);
    a.fdL = make_fastdiv((uint32_t)c->L);
    a.tan_sun = c->cull_tan_sun;
    a.tan_block = c->cull_tan_block;
    a.rho = c->cfg.helio_reflected_rate;
    a.npix = rb.w * rb.h;
    plan.rt = rb.r.type;
    // instance table
    if ((size_t)n * sizeof(qmc_instance) > c->batch_inst_cap) {
        qmc_instance *p = nullptr;
        if ((st = dev_alloc(p, (size_t)n))) return st;
        if (c->d_batch_inst) (void)hipFree(c->d_batch_inst);
        c->d_batch_inst = p;
        c->batch_inst_cap = (size_t)n * sizeof(qmc_instance);
    }
    HIP_TRY(hipMemcpyAsync(c->d_batch_inst, inst, sizeof(qmc_instance) * n, hipMemcpyHostToDevice, c->stream));
    a.inst = reinterpret_cast<const qmc_instance *>(c->d_batch_inst);
    a.n_inst = n;
    return QMC_OK;
}

qmc_status ensure_maps(qmc_ctx *c, size_t floats) {
    if (floats <= c->batch_cap) return QMC_OK;
    qmc_status st = dev_alloc(c->d_batch_maps, floats);
    if (st) {
        c->batch_cap = 0;
        return st;
    }
    c->batch_cap = floats;
    return QMC_OK;
}

qmc_status autotune(qmc_ctx *c, const BatchPlan &plan, int n, qmc_schedule &best) {
    int rays = 0;
    for (int k = 0; k < plan.args.n_helio; ++k) rays += c->h_helios[plan.args.helio_begin + k].F * c->L;
    const std::string key = shape_key(c, plan.args.grid.nx >= 0 ? c->batch_grid : 0, plan.args.npix,
                                      plan.args.n_helio, rays);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        tune_load_file();
        auto it = g_tune.find(key);
        if (it != g_tune.end()) {
            best = it->second;
            return QMC_OK;
        }
    }
    // Candidates: block size x LDS map x facet tiles, each with as many workgroups per CU as are
    // resident (the occupancy query). A first round screens them on one instance per workgroup; the
    // two fastest are then timed on the same number of instances (three rounds of the larger grid),
    // which ranks them in the steady state the sweep runs in.
    const size_t lds_per_cu = 160 * 1024;
    struct Cand {
        qmc_schedule s;
        float rate;
    };
    std::vector<Cand> cands;
    int max_grid = 0;
    for (int bs : {256, 512, 1024})
        for (int lds : {1, 0})
            for (int tiles : {0, 1, 2}) {
                const size_t bytes = lds_bytes(plan.args, lds != 0);
                if (bytes > c->max_lds) continue;
                if (tiles == 2 && !shared_order_lL(plan.args)) continue;
                int resident = 0;
                qmc_status st = resident_per_cu(plan.rt, lds != 0, tiles, bs, bytes, resident);
                if (st) return st;
                resident = std::min(resident, 2048 / bs);  // 32 waves per CU
                if (bytes > 0) resident = std::min<int>(resident, (int)(lds_per_cu / bytes));
                if (resident < 1) continue;
                cands.push_back(Cand{qmc_schedule{bs, resident, lds, tiles}, 0.0f});
                max_grid = std::max(max_grid, c->num_cus * resident);
            }
    if (cands.empty()) return fail(QMC_E_UNSUPPORTED, "batched trace: no schedule fits this device");
    const int long_trial = std::min(n, 3 * max_grid);
    qmc_status st = ensure_maps(c, (size_t)std::max(long_trial, std::min(n, max_grid)) * plan.args.npix);
    if (st) return st;
    c->batch_n = 0;  // the trial traces overwrite the ctx-owned maps
    BatchPlan p = plan;
    p.args.maps = c->d_batch_maps;
    auto time_it = [&](const qmc_schedule &s, int trial, float &rate) -> qmc_status {
        HIP_TRY(hipEventRecord(c->ev[14], c->stream));
        qmc_status e = launch_batched(c, p, s, trial);
        if (e) return e;
        HIP_TRY(hipEventRecord(c->ev[15], c->stream));
        HIP_TRY(hipEventSynchronize(c->ev[15]));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, c->ev[14], c->ev[15]));
        rate = ms > 0.0f ? (float)trial / ms : 0.0f;
        return QMC_OK;
    };
    for (Cand &cd : cands)
        if ((st = time_it(cd.s, std::min(n, c->num_cus * cd.s.workgroups_per_cu), cd.rate))) return st;
    std::sort(cands.begin(), cands.end(), [](const Cand &x, const Cand &y) { return x.rate > y.rate; });
    const size_t finalists = std::min<size_t>(2, cands.size());
    for (size_t i = 0; i < finalists; ++i)
        if ((st = time_it(cands[i].s, long_trial, cands[i].rate))) return st;
    std::sort(cands.begin(), cands.begin() + finalists, [](const Cand &x, const Cand &y) { return x.rate > y.rate; });
    best = cands[0].s;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_store(key, best);
    return QMC_OK;
}

}  // namespace

extern "C" qmc_status qmc_batch_set_aims(qmc_ctx *c, const qmc_scene *s, const float *offsets, int32_t n_sets) {
    if (!c || !s || !offsets || n_sets <= 0) return fail(QMC_E_INVALID_ARG, "qmc_batch_set_aims: bad argument");
    if (!c->uploaded) return fail(QMC_E_STATE, "qmc_batch_set_aims: upload the scene first");
    if (!s->processed || s->heliostats.size() != c->h_helios.size())
        return fail(QMC_E_STATE, "qmc_batch_set_aims: scene does not match the uploaded one");
    qmc_status st = set_device(c);
    if (st) return st;
    const int nh = (int)s->heliostats.size();
    std::vector<float> aims((size_t)n_sets * nh * 3);
    for (int k = 0; k < n_sets; ++k)
        for (int h = 0; h < nh; ++h) {
            const float *off = offsets + ((size_t)k * nh + h) * 2;
            V3 a = heliostat_aim(*s, h, off);
            float *o = &aims[((size_t)k * nh + h) * 3];
            o[0] = a.x;
            o[1] = a.y;
            o[2] = a.z;
        }
    if ((st = dev_alloc(c->d_aims, aims.size()))) return st;
    HIP_TRY(hipMemcpyAsync(c->d_aims, aims.data(), sizeof(float) * aims.size(), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->n_aim_sets = n_sets;
    return QMC_OK;
}

extern "C" qmc_status qmc_batch_autotune(qmc_ctx *c, int32_t grid, const qmc_instance *inst, int32_t n,
                                         qmc_schedule *out) {
    if (!c || !out) return fail(QMC_E_INVALID_ARG, "qmc_batch_autotune: null argument");
    qmc_status st = set_device(c);
    if (st) return st;
    BatchPlan plan;
    if ((st = plan_batch(c, grid, inst, n, plan))) return st;
    c->batch_grid = grid;
    return autotune(c, plan, n, *out);
}

extern "C" qmc_status qmc_trace_batched(qmc_ctx *c, int32_t grid, const qmc_instance *inst, int32_t n,
                                        const qmc_schedule *sched, float *d_maps) {
    if (!c) return fail(QMC_E_INVALID_ARG, "qmc_trace_batched: null ctx");
    qmc_status st = set_device(c);
    if (st) return st;
    BatchPlan plan;
    if ((st = plan_batch(c, grid, inst, n, plan))) return st;
    c->batch_grid = grid;
    qmc_schedule s;
    if (sched) {
        s = *sched;
        if (s.block_threads != 256 && s.block_threads != 512 && s.block_threads != 1024)
            return fail(QMC_E_INVALID_ARG, "qmc_trace_batched: block_threads must be 256, 512 or 1024");
        if (s.workgroups_per_cu < 1 || s.workgroups_per_cu > 16)
            return fail(QMC_E_INVALID_ARG, "qmc_trace_batched: workgroups_per_cu must be 1..16");
        if (s.lds_flux_tile != 0 && s.lds_flux_tile != 1)
            return fail(QMC_E_INVALID_ARG, "qmc_trace_batched: lds_flux_tile must be 0 or 1");
        if (s.facet_tiles < 0 || s.facet_tiles > 2)
            return fail(QMC_E_INVALID_ARG, "qmc_trace_batched: facet_tiles must be 0, 1 or 2");
    } else if ((st = autotune(c, plan, n, s))) {
        return st;
    }
    c->batch_n = 0;
    if (d_maps) {
        plan.args.maps = d_maps;
    } else {
        if ((st = ensure_maps(c, (size_t)n * plan.args.npix))) return st;
        plan.args.maps = c->d_batch_maps;
    }
    if ((st = launch_batched(c, plan, s, n))) return st;
    c->batch_maps_caller = d_maps != nullptr;
    c->batch_n = n;
    c->batch_npix = plan.args.npix;
    return QMC_OK;
}

extern "C" qmc_status qmc_batch_read(qmc_ctx *c, int32_t first, int32_t count, float *out) {
    if (!c || !out) return fail(QMC_E_INVALID_ARG, "qmc_batch_read: null argument");
    if (c->batch_maps_caller)
        return fail(QMC_E_STATE, "qmc_batch_read: the last batched trace wrote into the caller's buffer");
    if (first < 0 || count < 0 || first + count > c->batch_n)
        return fail(QMC_E_RANGE, "qmc_batch_read: instances out of range of the last batched trace");
    qmc_status st = set_device(c);
    if (st) return st;
    HIP_TRY(hipMemcpyAsync(out, c->d_batch_maps + (size_t)first * c->batch_npix,
                           sizeof(float) * (size_t)count * c->batch_npix, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return QMC_OK;
}

extern "C" qmc_status qmc_batch_set_work_budget(qmc_ctx *c, uint64_t bytes) {
    if (!c) return fail(QMC_E_INVALID_ARG, "qmc_batch_set_work_budget: null ctx");
    c->batch_work_budget = (size_t)bytes;
    return QMC_OK;
}
