// qmc_batch.hip — batched traces: many sun positions x aim strategies of one heliostat field in one
// persistent launch, each instance with its own flux map (no counterpart in the reference, which
// traces one configuration per process; its per-ray work is map_raytracing,
// QuasiMonteCarloRayTracing/RectangleReceiverRectangleGridRayTracing.cu:6-41).
//
// A workgroup owns whole instances. For each instance it recomputes, in LDS, what depends on the sun
// and the aim point — heliostat normals and frames (Heliostat::CSetNormal, Heliostat.cu:15-20),
// blocking quads (CSetWorldVertex, Heliostat.cu:25-40), energy factors
// (QuasiMonteCarloRayTracer.cpp:58-60) — with the same float operations the host processor uses, then
// traces every ray of the field with the shared per-ray functions and accumulates the instance's map
// in LDS (one plain store per pixel at the end) or with global atomics when the map does not fit.
// The grid cell lists, sample pools and micro-facet offsets are shared by all instances.
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

struct BatchArgs {
    DevGrid grid;
    DevReceiver recv;
    const DevHelio *helios;
    int32_t helio_begin, n_helio, n_total_helio;
    const float *aims;  // [set][heliostat][3]
    const qmc_instance *inst;
    int32_t n_inst;
    const float4 *pool;
    const float4 *pert;
    const int32_t *offsets;
    int32_t P, L;
    float rho;
    float *maps;
    int32_t npix;
};

struct LdsLayout {
    size_t quads, frames, factors, sun, map, total;
};

inline LdsLayout lds_layout(int n_helio, int npix, bool lds_map) {
    LdsLayout l;
    l.quads = 0;
    l.frames = l.quads + sizeof(float4) * 3 * (size_t)n_helio;
    l.factors = l.frames + sizeof(DevFrame) * (size_t)n_helio;
    l.sun = l.factors + sizeof(float) * (size_t)n_helio;
    l.map = (l.sun + sizeof(DevFrame) + 15) & ~(size_t)15;
    l.total = l.map + (lds_map ? sizeof(float) * (size_t)npix : 0);
    return l;
}

template <int RT, bool LDSMAP>
__global__ void k_trace_batched(BatchArgs a, LdsLayout lay) {
    extern __shared__ float4 smem4[];
    char *smem = reinterpret_cast<char *>(smem4);
    float4 *q_lds = reinterpret_cast<float4 *>(smem + lay.quads);
    DevFrame *fr_lds = reinterpret_cast<DevFrame *>(smem + lay.frames);
    float *fac_lds = reinterpret_cast<float *>(smem + lay.factors);
    DevFrame *sun_lds = reinterpret_cast<DevFrame *>(smem + lay.sun);
    float *map_lds = reinterpret_cast<float *>(smem + lay.map);
    const int tid = threadIdx.x, bs = blockDim.x;
    const int L = a.L, P = a.P;
    RayStats st{0u, 0u, 0u};

    for (int ii = blockIdx.x; ii < a.n_inst; ii += gridDim.x) {
        const qmc_instance I = a.inst[ii];
        const V3 sun{I.sun_dir[0], I.sun_dir[1], I.sun_dir[2]};
        __syncthreads();  // the previous instance is done with LDS
        if (tid == 0) {
            const Frame f = make_frame(-sun, false);
            DevFrame d;
            d.ux = f.u.x; d.uy = f.u.y; d.uz = f.u.z;
            d.nx = f.n.x; d.ny = f.n.y; d.nz = f.n.z;
            d.vx = f.v.x; d.vy = f.v.y; d.vz = f.v.z;
            d.identity = f.identity;
            *sun_lds = d;
        }
        for (int k = tid; k < a.n_helio; k += bs) {
            const DevHelio &H = a.helios[a.helio_begin + k];
            const float *ap = a.aims + ((size_t)I.aim_set * a.n_total_helio + a.helio_begin + k) * 3;
            const V3 aim{ap[0], ap[1], ap[2]};
            const V3 pos{H.pos[0], H.pos[1], H.pos[2]};
            const V3 normal = normalize(normalize(aim - pos) - sun);
            const Frame f = make_frame(normal, false);
            DevFrame d;
            d.ux = f.u.x; d.uy = f.u.y; d.uz = f.u.z;
            d.nx = f.n.x; d.ny = f.n.y; d.nz = f.n.z;
            d.vx = f.v.x; d.vy = f.v.y; d.vz = f.v.z;
            d.identity = f.identity;
            fr_lds[k] = d;
            // flat heliostat blocking quad: world vertices v0, v1, v2 -> (A, E1, E2) = (v1, v0-v1, v2-v1)
            const V3 v0l{-H.size[0] / 2, H.size[1] / 2, -H.size[2] / 2};
            const V3 v1l = V3{0, 0, H.size[2]} + v0l;
            const V3 v2l = V3{H.size[0], 0, H.size[2]} + v0l;
            const V3 w0 = frame_apply(f, v0l) + pos, w1 = frame_apply(f, v1l) + pos, w2 = frame_apply(f, v2l) + pos;
            const V3 E1 = w0 - w1, E2 = w2 - w1;
            const int sid = H.sub_begin;
            q_lds[3 * sid] = make_float4(w1.x, w1.y, w1.z, 0.0f);
            q_lds[3 * sid + 1] = make_float4(E1.x, E1.y, E1.z, 0.0f);
            q_lds[3 * sid + 2] = make_float4(E2.x, E2.y, E2.z, 0.0f);
            fac_lds[k] = H.fac_rule == 0 ? I.dni * H.fac_x * H.fac_x * a.rho / float(L)
                                         : I.dni * H.fac_x * a.rho / float(L);
        }
        if (LDSMAP)
            for (int p = tid; p < a.npix; p += bs) map_lds[p] = 0.0f;
        __syncthreads();
        const DevFrame sunf = *sun_lds;
        float *gmap = a.maps + (size_t)ii * a.npix;

        for (int k = 0; k < a.n_helio; ++k) {
            const DevHelio &H = a.helios[a.helio_begin + k];
            const DevFrame fr = fr_lds[k];
            const float factor = fac_lds[k];
            const int32_t *grp = a.offsets + H.offsets_base;
            const int own_lo = H.sub_begin, own_hi = H.sub_begin + H.sub_count;
            const int nr = H.F * L;
            for (int r = tid; r < nr; r += bs) {
                const int m = r / L, j = r % L;
                const int a1 = (grp[m] + j) % P;
                const float4 s = a.pool[a1];
                V3 dir = fr_apply(sunf, V3{s.x, s.y, s.z});
                const V3 origin = flat_facet_origin(H, fr, m);
                int hit;
                if (dda_collision<false>(origin, dir, a.grid, q_lds, own_lo, own_hi, hit, st)) continue;
                const int a2 = (grp[(m + 1) % P] + j) % P;
                const float4 q = a.pert[a2];
                const V3 normal = normalize(fr_apply(fr, V3{q.x, q.y, q.z}));
                dir = normalize(reflect(-dir, normal));
                if (dda_collision<false>(origin, dir, a.grid, q_lds, own_lo, own_hi, hit, st)) continue;
                float t, u, v;
                if (!receiver_hit<RT>(a.recv, origin, dir, t, u, v)) continue;
                const float energy = ray_energy(t, dir, normal, factor);
                int row = (int)(u * (float)a.recv.res_y);
                int col = (int)(v * (float)a.recv.res_x);
                if (row >= a.recv.res_y) row = a.recv.res_y - 1;
                if (col >= a.recv.res_x) col = a.recv.res_x - 1;
                const int addr = row * a.recv.res_x + col;
                if (LDSMAP)
                    atomicAdd(map_lds + addr, energy);
                else
                    unsafeAtomicAdd(gmap + addr, energy);
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
    while (f >> key >> s.block_threads >> s.workgroups_per_cu >> s.lds_flux_tile >> s.reserved) g_tune[key] = s;
}

void tune_store(const std::string &key, const qmc_schedule &s) {
    g_tune[key] = s;
    const char *path = std::getenv("QMC_TUNE_CACHE");
    if (!path) return;
    std::ofstream f(path, std::ios::app);
    f << key << " " << s.block_threads << " " << s.workgroups_per_cu << " " << s.lds_flux_tile << " " << s.reserved
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

qmc_status ensure_device_limits(qmc_ctx *c) {
    if (c->num_cus > 0) return QMC_OK;
    int v = 0;
    HIP_TRY(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, c->device));
    c->num_cus = v;
    HIP_TRY(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
    c->max_lds = (size_t)v;
    return QMC_OK;
}

struct BatchPlan {
    BatchArgs args;
    int rt;
};

qmc_status launch_batched(qmc_ctx *c, const BatchPlan &plan, const qmc_schedule &s, int n_inst) {
    BatchArgs a = plan.args;
    a.n_inst = n_inst;
    const bool lds_map = s.lds_flux_tile != 0;
    const LdsLayout lay = lds_layout(a.n_helio, a.npix, lds_map);
    if (lay.total > c->max_lds) return fail(QMC_E_UNSUPPORTED, "batched trace: LDS layout exceeds the device limit");
    if (!lds_map) HIP_TRY(hipMemsetAsync(a.maps, 0, sizeof(float) * (size_t)n_inst * a.npix, c->stream));
    const int grid = std::max(1, std::min(n_inst, c->num_cus * std::max(1, s.workgroups_per_cu)));
    dim3 g(grid), b(s.block_threads);
    if (plan.rt == 0) {
        if (lds_map) hipLaunchKernelGGL((k_trace_batched<0, true>), g, b, lay.total, c->stream, a, lay);
        else hipLaunchKernelGGL((k_trace_batched<0, false>), g, b, lay.total, c->stream, a, lay);
    } else {
        if (lds_map) hipLaunchKernelGGL((k_trace_batched<1, true>), g, b, lay.total, c->stream, a, lay);
        else hipLaunchKernelGGL((k_trace_batched<1, false>), g, b, lay.total, c->stream, a, lay);
    }
    HIP_TRY(hipGetLastError());
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
    qmc_status st = ensure_device_limits(c);
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
    BatchArgs &a = plan.args;
    std::memset(&a, 0, sizeof(a));
    a.grid = gb.g;
    a.recv = rb.r;
    a.helios = c->d_helios;
    a.helio_begin = begin;
    a.n_helio = count;
    a.n_total_helio = (int)c->h_helios.size();
    a.aims = c->d_aims;
    a.pool = c->d_pool;
    a.pert = c->d_pert;
    a.offsets = c->d_offsets;
    a.P = c->P;
    a.L = c->L;
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
    const size_t lds_per_cu = 160 * 1024;
    const int trial = std::min(n, c->num_cus * 4);
    qmc_status st = ensure_maps(c, (size_t)trial * plan.args.npix);
    if (st) return st;
    BatchPlan p = plan;
    p.args.maps = c->d_batch_maps;
    float best_ms = 1e30f;
    bool found = false;
    for (int bs : {256, 512, 1024})
        for (int lds : {1, 0})
            for (int wpc : {1, 2, 4}) {
                const LdsLayout lay = lds_layout(p.args.n_helio, p.args.npix, lds != 0);
                if (lay.total > c->max_lds || lay.total * wpc > lds_per_cu) continue;
                if (bs * wpc > 2048) continue;
                qmc_schedule s{bs, wpc, lds, 0};
                HIP_TRY(hipEventRecord(c->ev[14], c->stream));
                if ((st = launch_batched(c, p, s, trial))) return st;
                HIP_TRY(hipEventRecord(c->ev[15], c->stream));
                HIP_TRY(hipEventSynchronize(c->ev[15]));
                float ms = 0;
                HIP_TRY(hipEventElapsedTime(&ms, c->ev[14], c->ev[15]));
                if (ms < best_ms) {
                    best_ms = ms;
                    best = s;
                    found = true;
                }
            }
    if (!found) return fail(QMC_E_UNSUPPORTED, "batched trace: no schedule fits this device");
    best.reserved = 0;
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
        if (s.block_threads != 64 && s.block_threads != 128 && s.block_threads != 256 && s.block_threads != 512 &&
            s.block_threads != 1024)
            return fail(QMC_E_INVALID_ARG, "qmc_trace_batched: block_threads must be 64..1024, a power of two");
        if (s.workgroups_per_cu < 1 || s.workgroups_per_cu > 16)
            return fail(QMC_E_INVALID_ARG, "qmc_trace_batched: workgroups_per_cu must be 1..16");
    } else if ((st = autotune(c, plan, n, s))) {
        return st;
    }
    if (d_maps) {
        plan.args.maps = d_maps;
    } else {
        if ((st = ensure_maps(c, (size_t)n * plan.args.npix))) return st;
        plan.args.maps = c->d_batch_maps;
    }
    if ((st = launch_batched(c, plan, s, n))) return st;
    c->batch_n = n;
    c->batch_npix = plan.args.npix;
    return QMC_OK;
}

extern "C" qmc_status qmc_batch_read(qmc_ctx *c, int32_t first, int32_t count, float *out) {
    if (!c || !out) return fail(QMC_E_INVALID_ARG, "qmc_batch_read: null argument");
    if (first < 0 || count < 0 || first + count > c->batch_n)
        return fail(QMC_E_RANGE, "qmc_batch_read: instances out of range of the last batched trace");
    qmc_status st = set_device(c);
    if (st) return st;
    HIP_TRY(hipMemcpyAsync(out, c->d_batch_maps + (size_t)first * c->batch_npix,
                           sizeof(float) * (size_t)count * c->batch_npix, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return QMC_OK;
}
