// qmc_probe.hip — device capability probe: the rates of the atomic operations the tracer deposits
// with (no counterpart in the reference). The trace's achieved deposit rates are reported against
// these ceilings: float and 64-bit device-scope atomic adds scattered over a map the size of the
// PS10-like field's (256 x 256), and float LDS atomic adds scattered over a 64 KB tile.
#include <hip/hip_runtime.h>

#include <string>

#include "qmc_ctx.h"

using namespace qmc;

namespace {

constexpr int kProbeThreads = 4096 * 1024;  // lanes per launch
constexpr int kProbePerLane = 64;           // atomics per lane
constexpr int kProbePixels = 65536;

__device__ __forceinline__ uint32_t probe_hash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

template <class T>
__global__ __launch_bounds__(256) void k_probe_global(T *map) {
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    for (int i = 0; i < kProbePerLane; ++i)
        unsafeAtomicAdd(map + (probe_hash(t * 131u + (uint32_t)i) & (kProbePixels - 1)), (T)1);
}

__global__ __launch_bounds__(256) void k_probe_lds(float *sink) {
    __shared__ float tile[16384];
    for (int p = threadIdx.x; p < 16384; p += 256) tile[p] = 0.0f;
    __syncthreads();
    const uint32_t t = blockIdx.x * 256u + threadIdx.x;
    for (int i = 0; i < kProbePerLane; ++i) atomicAdd(tile + (probe_hash(t * 131u + (uint32_t)i) & 16383u), 1.0f);
    __syncthreads();
    // one plain store per workgroup keeps the tile live; every value is a small positive count
    if (threadIdx.x == 0) sink[blockIdx.x] = tile[0];
}

}  // namespace

extern "C" qmc_status qmc_probe_atomic_rates(qmc_ctx *c, qmc_atomic_rates *out) {
    if (!c || !out) return fail(QMC_E_INVALID_ARG, "qmc_probe_atomic_rates: null argument");
    qmc_status st = set_device(c);
    if (st) return st;
    float *f = nullptr;
    double *d = nullptr;
    float *sink = nullptr;
    const int blocks = kProbeThreads / 256;
    if ((st = dev_alloc(f, kProbePixels))) return st;
    if ((st = dev_alloc(d, kProbePixels))) {
        dev_free(f);
        return st;
    }
    if ((st = dev_alloc(sink, blocks))) {
        dev_free(f);
        dev_free(d);
        return st;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double rates[3] = {0, 0, 0};
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    for (int k = 0; k < 3 && e == hipSuccess; ++k) {
        for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {  // the first launch warms up, the second is timed
            if (rep == 1) e = hipEventRecord(e0, c->stream);
            if (k == 0) hipLaunchKernelGGL(k_probe_global<float>, dim3(blocks), dim3(256), 0, c->stream, f);
            if (k == 1) hipLaunchKernelGGL(k_probe_global<double>, dim3(blocks), dim3(256), 0, c->stream, d);
            if (k == 2) hipLaunchKernelGGL(k_probe_lds, dim3(blocks), dim3(256), 0, c->stream, sink);
            if (e == hipSuccess) e = hipGetLastError();
            if (rep == 1 && e == hipSuccess) e = hipEventRecord(e1, c->stream);
        }
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e == hipSuccess && ms > 0.0f) rates[k] = (double)kProbeThreads * kProbePerLane / (ms * 1e-3) / 1e9;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dev_free(f);
    dev_free(d);
    dev_free(sink);
    if (e != hipSuccess) return fail(QMC_E_HIP, std::string("qmc_probe_atomic_rates: ") + hipGetErrorString(e));
    out->global_f32 = rates[0];
    out->global_f64 = rates[1];
    out->lds_f32 = rates[2];
    return QMC_OK;
}
