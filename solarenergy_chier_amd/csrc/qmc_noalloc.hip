// qmc_noalloc.hip — qmc_trace_noalloc: the plain trace for callers that cannot allocate device memory
// (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). No counterpart in
// the reference.
//
// A plain trace needs device memory beyond its inputs (the uploaded scene, pools and offsets) and
// outputs (the receivers' flux maps and their accumulators): the call's heliostat list, per-heliostat
// candidate lists, the banded kernel's per-workgroup bands and the per-wave candidate masks. qmc_trace
// grows ctx-owned buffers for them; this entry point carves them from one file-scope array of
// kScratchBytes instead, of which every device holds its own copy until the process ends. A launch
// whose wave masks do not fit in what is left of it is split into launches over fewer heliostats.
// Nothing in the array has to start at zero: each kernel writes what it later reads, and the counts
// that must start at zero in every launch (the deferral queue's, set by k_cull, and the column-window
// trace's claim counters, set by k_zero_counters) are set by kernels of that launch, not by memsets.
// Calls that overlap on one device (two contexts on two streams) share the array and race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "qmc_ctx.h"

using namespace qmc;

namespace {
constexpr size_t kScratchBytes = 128ull << 20;  // 128 MiB per device
constexpr int kMaxList = 65536;                 // list entries per call
constexpr size_t kDeferEntries = 1u << 20;      // deferred rays queued per launch (8 MiB)
constexpr int kMaxDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_scratch_mu;
void *g_scratch_addr[kMaxDevices] = {};

// This device's address of `scratch`, taken once, at the first call on the device, with the device
// current. Host code uses only this address (kernel arguments, copies), never the symbol.
qmc_status scratch_address(int device, void **out) {
    if (device < 0 || device >= kMaxDevices)
        return fail(QMC_E_UNSUPPORTED, "qmc_trace_noalloc: device index " + std::to_string(device) + " above " +
                                           std::to_string(kMaxDevices - 1));
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(QMC_E_HIP, std::string("qmc_trace_noalloc: hipGetSymbolAddress(scratch): ") +
                                       hipGetErrorString(e));
        g_scratch_addr[device] = p;
    }
    *out = g_scratch_addr[device];
    return QMC_OK;
}

}  // namespace

extern "C" qmc_status qmc_trace_noalloc(qmc_ctx *c, const int32_t *ids, int32_t n) {
    if (!c) return fail(QMC_E_INVALID_ARG, "qmc_trace_noalloc: null ctx");
    if (!c->uploaded) return fail(QMC_E_STATE, "trace: upload a scene first");
    if (n > kMaxList)
        return fail(QMC_E_RANGE, "qmc_trace_noalloc: " + std::to_string(n) + " list entries exceed the maximum of " +
                                     std::to_string(kMaxList));
    qmc_status st = set_device(c);
    if (st) return st;
    void *base_addr = nullptr;
    if ((st = scratch_address(c->device, &base_addr))) return st;
    if ((st = device_limits(c))) return st;
    // layout: list | candidate lists (CullList, WideList per heliostat) | one band per CU | deferred-ray
    // queue (a full queue sends the rest through the full traversal in place) | wave masks
    char *base = static_cast<char *>(base_addr);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char *p = base + off;
        off += (bytes + 255) & ~(size_t)255;
        return p;
    };
    const size_t nh = c->h_helios.size();
    TraceBufs b;
    b.list = reinterpret_cast<int32_t *>(take(sizeof(int32_t) * (size_t)kMaxList));
    b.cull = reinterpret_cast<CullList *>(take(sizeof(CullList) * nh));
    b.wide = reinterpret_cast<WideList *>(take(sizeof(WideList) * nh));
    // bands, and for column windows their first columns and claim counters
    b.band_cap = (size_t)c->num_cus * (c->max_lds / sizeof(float) + 2);
    b.band = reinterpret_cast<float *>(take(sizeof(float) * b.band_cap));
    b.defer_n = reinterpret_cast<unsigned *>(take(sizeof(unsigned)));
    b.defer_cap = kDeferEntries;
    b.defer = reinterpret_cast<unsigned long long *>(take(sizeof(unsigned long long) * kDeferEntries));
    if (off + 256 > kScratchBytes)
        return fail(QMC_E_RANGE, "qmc_trace_noalloc: the candidate lists of " + std::to_string(nh) +
                                     " heliostats and the workgroup bands exceed the " +
                                     std::to_string(kScratchBytes >> 20) + " MiB scratch");
    b.wmask = reinterpret_cast<uint64_t *>(base + off);
    b.wmask_cap = (kScratchBytes - off) / sizeof(uint64_t);
    b.fixed = true;
    return trace_fixed(c, ids, n, b);
}
