// qmc_probe.hip — device capability probes (no counterpart in the reference):
//  * the rates of the atomic operations the tracer deposits with. The trace's achieved deposit rates
//    are reported against these ceilings: float and 64-bit device-scope atomic adds scattered over a
//    map the size of the PS10-like field's (256 x 256), and float LDS atomic adds scattered over a
//    64 KB tile;
//  * the exhaustive check of the short square-root sequences of qmc_math.h against the general ones;
//  * the exhaustive host check of the cylinder map's column breakpoints (qmc_math.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <thread>
#include <vector>

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

// Every float x in [2^-8, 2^8] (bit patterns kFastSqrtLo + i): the short sequences against sqrtf and
// 1.0f / sqrtf, and those against the double computation rounded to float (double rounding of a float
// square root or quotient through double is exact: 53 >= 2 * 24 + 2).
__global__ __launch_bounds__(256) void k_check_fast_sqrt(unsigned long long *bad, uint32_t n) {
    unsigned long long b0 = 0, b1 = 0, b2 = 0;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
        const float x = __uint_as_float(kFastSqrtLo + i);
        const float s = sqrtf(x), r = 1.0f / s;
        const float fs = sqrt_in_range(x), fr = recip_in_range(fs);
        const float ds = (float)sqrt((double)x), dr = (float)(1.0 / (double)ds);
        b0 += __float_as_uint(fs) != __float_as_uint(s);
        b1 += __float_as_uint(fr) != __float_as_uint(r);
        b2 += (__float_as_uint(ds) != __float_as_uint(s)) | (__float_as_uint(dr) != __float_as_uint(r));
    }
    if (b0) atomicAdd(bad, b0);
    if (b1) atomicAdd(bad + 1, b1);
    if (b2) atomicAdd(bad + 2, b2);
}

}  // namespace

extern "C" qmc_status qmc_check_fast_sqrt(qmc_ctx *c, uint64_t *checked, uint64_t mismatches[3]) {
    if (!c || !checked || !mismatches) return fail(QMC_E_INVALID_ARG, "qmc_check_fast_sqrt: null argument");
    qmc_status st = set_device(c);
    if (st) return st;
    unsigned long long *bad = nullptr;
    if ((st = dev_alloc(bad, 3))) return st;
    const uint32_t n = kFastSqrtHi - kFastSqrtLo + 1u;
    unsigned long long h[3] = {0, 0, 0};
    hipError_t e = hipMemsetAsync(bad, 0, sizeof(h), c->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_check_fast_sqrt, dim3(4096), dim3(256), 0, c->stream, bad, n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(h, bad, sizeof(h), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dev_free(bad);
    if (e != hipSuccess) return fail(QMC_E_HIP, std::string("qmc_check_fast_sqrt: ") + hipGetErrorString(e));
    *checked = n;
    for (int k = 0; k < 3; ++k) mismatches[k] = h[k];
    return QMC_OK;
}

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

extern "C" qmc_status qmc_check_cylinder_columns(int32_t width, int32_t stride, uint64_t *checked,
                                                  uint64_t *mismatches, double *max_acos_error) {
    if (!checked || !mismatches || !max_acos_error)
        return fail(QMC_E_INVALID_ARG, "qmc_check_cylinder_columns: null argument");
    if (width < 1 || width > kColumnTableMaxW)
        return fail(QMC_E_RANGE, "qmc_check_cylinder_columns: width outside [1, " +
                                     std::to_string(kColumnTableMaxW) + "]");
    if (stride < 1) return fail(QMC_E_INVALID_ARG, "qmc_check_cylinder_columns: stride below 1");
    const int W = width;
    std::vector<float> T(2 * (size_t)W + 4);
    cylinder_column_table(W, T.data());
    // the floats of [-1, 1] (one zero) at the stride and, with a stride, the 513 floats centred on each
    // breakpoint (where the column changes); split over host threads
    const int64_t k0 = float_order(-1.0f), k1 = float_order(1.0f);
    std::vector<int64_t> near;
    if (stride > 1)
        for (float b : T)
            if (b >= -1.0f && b <= 1.0f)
                for (int64_t d = -256; d <= 256; ++d) {
                    const int64_t k = float_order(b) + d;
                    if (k >= k0 && k <= k1) near.push_back(k);
                }
    const int64_t n_grid = (k1 - k0) / stride + 1, n_all = n_grid + (int64_t)near.size();
    const int nt = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<uint64_t> bad(nt, 0);
    std::vector<double> err(nt, 0.0);
    std::vector<std::thread> th;
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&, t] {
            for (int64_t i = n_all * t / nt; i < n_all * (t + 1) / nt; ++i) {
                const float cx = order_float(i < n_grid ? k0 + i * stride : near[(size_t)(i - n_grid)]);
                for (int neg = 0; neg < 2; ++neg)
                    bad[t] += cylinder_column(T.data(), W, cx, neg != 0) != cylinder_column_exact(cx, neg != 0, W);
                // the estimate's own error (formed as on the device, with a correctly rounded sqrt here)
                const float ax = std::fabs(cx);
                float ac = std::sqrt(1.0f - ax) * (1.5707288f + ax * (-0.2121144f + ax * (0.0742610f - 0.0187293f * ax)));
                if (cx < 0.0f) ac = 3.14159265f - ac;
                err[t] = std::max(err[t], std::fabs((double)ac - std::acos((double)cx)));
            }
        });
    for (auto &x : th) x.join();
    *checked = (uint64_t)(2 * n_all);
    *mismatches = 0;
    *max_acos_error = 0.0;
    for (int t = 0; t < nt; ++t) {
        *mismatches += bad[t];
        *max_acos_error = std::max(*max_acos_error, err[t]);
    }
    return QMC_OK;
}
