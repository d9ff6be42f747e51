// qmc_smooth.hip — the trimmed-mean smoother (ImageSmoother::image_smooth,
// Tool/Smoother/ImageSmoother.cu:68-113; trimmed_mean :13-64; heaps Tool/Smoother/heap.cuh), one lane
// per pixel reading an LDS tile with its halo. Kept apart from the tracer kernels: it is built with
// the default loop-invariant code motion, which its heap loops need.
#include <hip/hip_runtime.h>

#include <string>

#include "qmc_ctx.h"

using namespace qmc;

namespace {

// ---------------------------------------------------------------------------------------------
// Trimmed-mean smoother (Tool/Smoother/ImageSmoother.cu:13-64, Tool/Smoother/heap.cuh)
// ---------------------------------------------------------------------------------------------
// Binary heap over e[1..cnt-1] (1-based, e[0] unused) with the reference's insert and
// delete-and-insert; MAXH selects the comparison (max-heap: a > b, min-heap: a < b).
template <bool MAXH>
__device__ __forceinline__ bool hcmp(float a, float b) {
    return MAXH ? (a > b) : (a < b);
}

template <bool MAXH>
__device__ __forceinline__ void heap_insert(float *e, int &num, float x) {
    int i = num++;
    for (; i != 1 && hcmp<MAXH>(x, e[i / 2]); i /= 2) e[i] = e[i / 2];
    e[i] = x;
}

template <bool MAXH>
__device__ __forceinline__ void heap_replace_top(float *e, int num, float x) {
    if (num == 1) return;  // empty heap: no change (returns a sentinel in the reference)
    int i = 1;
    for (; 2 * i < num;) {
        int c = (2 * i + 1 >= num || hcmp<MAXH>(e[2 * i], e[2 * i + 1])) ? 2 * i : 2 * i + 1;
        if (hcmp<MAXH>(e[c], x)) {
            e[i] = e[c];
            i = c;
        } else {
            break;
        }
    }
    e[i] = x;
}

__device__ __forceinline__ float heap_sum(const float *e, int num) {
    float s = 0.0f;
    for (int i = 1; i < num; ++i) s += e[i];
    return s;
}

__global__ void k_smooth(float *out, const float *src, int width, int height, int radius, float ratio, int kalloc,
                         int tile_w) {
    extern __shared__ float lds[];
    const int bx0 = blockIdx.x * blockDim.x, by0 = blockIdx.y * blockDim.y;
    // tile covers [bx0 - radius, bx0 + bdx + radius) x [by0 - radius, by0 + bdy + radius), clamped reads only
    float *tile = lds;
    const int tw = tile_w;
    const int th = blockDim.y + 2 * radius;
    const int tid = threadIdx.y * blockDim.x + threadIdx.x;
    const int nthreads = blockDim.x * blockDim.y;
    for (int idx = tid; idx < tw * th; idx += nthreads) {
        int ty = idx / tw, tx = idx % tw;
        int gx = bx0 - radius + tx, gy = by0 - radius + ty;
        float val = 0.0f;
        if (gx >= 0 && gx < width && gy >= 0 && gy < height) val = src[gy * width + gx];
        tile[idx] = val;
    }
    __syncthreads();
    const int x = bx0 + threadIdx.x, y = by0 + threadIdx.y;
    if (x > width - 1 || y > height - 1) return;
    float *max_e = lds + tw * th + (2 * tid) * kalloc;
    float *min_e = lds + tw * th + (2 * tid + 1) * kalloc;
    const int lx = (x - radius >= 0) ? (x - radius) : 0;
    const int ly = (y - radius >= 0) ? (y - radius) : 0;
    const int rx = (x + radius < width) ? (x + radius) : width - 1;
    const int ry = (y + radius < height) ? (y + radius) : height - 1;
    const int cnt = (rx - lx + 1) * (ry - ly + 1);
    const int k = (int)(cnt * ratio);
    const int max_size = k + 1;
    int nmax = 1, nmin = 1;
    float value = 0.0f;
    for (int rr = ly; rr <= ry; ++rr) {
        for (int cc = lx; cc <= rx; ++cc) {
            const float cur = tile[(rr - by0 + radius) * tw + (cc - bx0 + radius)];
            value += cur;
            if (nmax != max_size) {
                heap_insert<true>(max_e, nmax, cur);
                heap_insert<false>(min_e, nmin, cur);
            } else if (cur < (nmax > 1 ? max_e[1] : 0.0f)) {
                heap_replace_top<true>(max_e, nmax, cur);
            } else if (cur > (nmin > 1 ? min_e[1] : 0.0f)) {
                heap_replace_top<false>(min_e, nmin, cur);
            }
        }
    }
    out[y * width + x] = (value - heap_sum(max_e, nmax) - heap_sum(min_e, nmin)) / float(cnt - 2 * k);
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// Smoother
// ---------------------------------------------------------------------------------------------
extern "C" qmc_status qmc_smooth(qmc_ctx *c, float *img, int32_t w, int32_t h, int32_t radius, float ratio) {
    if (!c || !img) return fail(QMC_E_INVALID_ARG, "qmc_smooth: null argument");
    if (w <= 0 || h <= 0 || radius < 0 || !(ratio >= 0.0f) || ratio >= 0.5f)
        return fail(QMC_E_INVALID_ARG, "qmc_smooth: need w, h > 0, radius >= 0 and 0 <= ratio < 0.5");
    qmc_status st = set_device(c);
    if (st) return st;
    // heap slots per thread, as the reference sizes them (ImageSmoother.cu:103)
    const int kalloc = (int)((2 * radius + 1) * (2 * radius + 1) * ratio + 1);
    int bdx = 16, bdy = 16;
    auto lds_bytes = [&](int bx, int by) {
        return sizeof(float) * ((size_t)(bx + 2 * radius) * (by + 2 * radius) + (size_t)2 * kalloc * bx * by);
    };
    while (lds_bytes(bdx, bdy) > 64 * 1024 && bdx * bdy > 16) {
        if (bdx >= bdy) bdx /= 2;
        else bdy /= 2;
    }
    if (lds_bytes(bdx, bdy) > 64 * 1024)
        return fail(QMC_E_UNSUPPORTED, "qmc_smooth: radius/ratio need more than 64 KiB of LDS per block");
    const size_t n = (size_t)w * h;
    if (n > c->scratch_cap) {
        if ((st = dev_alloc(c->d_scratch, n))) return st;
        c->scratch_cap = n;
    }
    HIP_TRY(hipMemcpyAsync(c->d_scratch, img, sizeof(float) * n, hipMemcpyDeviceToDevice, c->stream));
    dim3 b(bdx, bdy), g((w + bdx - 1) / bdx, (h + bdy - 1) / bdy);
    hipLaunchKernelGGL(k_smooth, g, b, lds_bytes(bdx, bdy), c->stream, img, c->d_scratch, w, h, radius, ratio, kalloc,
                       bdx + 2 * radius);
    HIP_TRY(hipGetLastError());
    return QMC_OK;
}

extern "C" qmc_status qmc_smooth_receiver(qmc_ctx *c, int32_t r, int32_t radius, float ratio) {
    if (!c) return fail(QMC_E_INVALID_ARG, "qmc_smooth_receiver: null ctx");
    if (r < 0 || r >= (int)c->recvs.size()) return fail(QMC_E_RANGE, "qmc_smooth_receiver: receiver out of range");
    return qmc_smooth(c, c->recvs[r].flux, c->recvs[r].w, c->recvs[r].h, radius, ratio);
}

