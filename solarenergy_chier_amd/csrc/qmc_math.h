// qmc_math.h — float3 arithmetic with the exact operation order of the reference tracer,
// shared by the host scene processor and the gfx950 kernels so both round identically.
//
// Semantics restated (not copied) from the reference's vector helpers:
//   dot(a,b) = a.x*b.x + a.y*b.y + a.z*b.z (left to right), cross = the usual determinant
//   expansion, normalize(v) = v * (1.0f / sqrtf(dot(v,v))) with a correctly rounded sqrt and
//   divide (the reference's host rsqrtf, Util/vector_arithmetic.cuh:50-53,397-401),
//   reflect(i,n) = i - (2*n)*dot(n,i) (Util/vector_arithmetic.cuh:410-413),
//   float->int conversion truncates toward zero (Util/vector_arithmetic.cuh:571-574).
// Epsilon is the double 1e-5 (Util/global_constant.h:11): every comparison against it decides as
// in double, as C++ promotion does in the reference (in double, or as the equal float compare
// against kEpsF).
//
// All code including this header must be compiled with -ffp-contract=off and without
// fast-math, so that no multiply-add is fused and host and device produce the same bits.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define QMC_HD __host__ __device__ __forceinline__
#else
#define QMC_HD inline
#endif

namespace qmc {

constexpr double kEps = 1e-5;               // Util/global_constant.h:11 (a double)
// Epsilon as a float: the float just below it. Epsilon is not a float and no float lies in (kEpsF,
// Epsilon], so for a float x: x < Epsilon and x <= Epsilon exactly when x <= kEpsF, x < -Epsilon exactly
// when x < -kEpsF — float compares that decide the reference's double ones.
constexpr float kEpsF = (float)kEps;
static_assert((double)kEpsF < kEps && (double)kEpsF + 0x1p-40 > kEps, "kEpsF is the float below Epsilon");
constexpr float kMathPi = 3.141592654f;      // Util/global_constant.h:12
constexpr double kInv2Pi = 1.0 / (2 * M_PI);  // the double nearest 1 / (the double nearest 2 pi)
constexpr float kIntMaxF = 2147483648.0f;    // float(INT_MAX), rectGridDDA.cuh:17-20

struct V3 {
    float x, y, z;
};

QMC_HD V3 v3(float x, float y, float z) { return V3{x, y, z}; }
QMC_HD V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
QMC_HD V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
QMC_HD V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
QMC_HD V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
QMC_HD V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
QMC_HD V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
QMC_HD V3 operator/(V3 a, V3 b) { return V3{a.x / b.x, a.y / b.y, a.z / b.z}; }
QMC_HD V3 operator-(V3 a, float s) { return V3{a.x - s, a.y - s, a.z - s}; }
QMC_HD V3 operator+(V3 a, float s) { return V3{a.x + s, a.y + s, a.z + s}; }

QMC_HD float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
QMC_HD V3 cross(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// Correctly rounded sqrtf(x) and 1.0f / sqrtf(x) (each operation rounded once, as IEEE arithmetic and the
// host compute them). On the device, x in [2^-8, 2^8] takes a short sequence: the hardware square root
// (within 1 ulp), replaced by a neighbour when that neighbour's residual shows it is nearer, and the
// reciprocal of that root by a Newton-Raphson step on the hardware reciprocal followed by two residual
// corrections, the last one Markstein's (P. Markstein, "Computation of elementary functions on the IBM
// RISC System/6000 processor", IBM J. Res. Develop. 34(1), 1990). In that range no operand, quotient or
// residual is denormal, infinite or near overflow, so the general sequences' exponent scaling and
// special-value steps change nothing and are left out. Both results are compared with the general
// sequences for every float of the range on the GPU (qmc_check_fast_sqrt). Other x take the general
// sequences.
constexpr uint32_t kFastSqrtLo = 0x3B800000u;  // 2^-8
constexpr uint32_t kFastSqrtHi = 0x43800000u;  // 2^8
#if defined(__HIPCC__)
__device__ __forceinline__ bool fast_sqrt_range(float x) {
    return __float_as_uint(x) - kFastSqrtLo <= kFastSqrtHi - kFastSqrtLo;
}
__device__ __forceinline__ float sqrt_in_range(float x) {
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sm = __uint_as_float(__float_as_uint(s) - 1u), sp = __uint_as_float(__float_as_uint(s) + 1u);
    float r = __builtin_fmaf(-sm, s, x) <= 0.0f ? sm : s;
    r = __builtin_fmaf(-sp, s, x) > 0.0f ? sp : r;
    return r;
}
__device__ __forceinline__ float recip_in_range(float d) {
    const float y0 = __builtin_amdgcn_rcpf(d);
    const float y = __builtin_fmaf(__builtin_fmaf(-d, y0, 1.0f), y0, y0);
    const float q = __builtin_fmaf(__builtin_fmaf(-d, y, 1.0f), y, y);
    return __builtin_fmaf(__builtin_fmaf(-d, q, 1.0f), y, q);
}
#endif
QMC_HD float sqrt_cr(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    if (fast_sqrt_range(x)) return sqrt_in_range(x);
#endif
    return sqrtf(x);
}
QMC_HD float inv_sqrt_cr(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    if (fast_sqrt_range(x)) return recip_in_range(sqrt_in_range(x));
#endif
    return 1.0f / sqrtf(x);
}

QMC_HD float length(V3 v) { return sqrt_cr(dot(v, v)); }
QMC_HD V3 normalize(V3 v) {
    float inv = inv_sqrt_cr(dot(v, v));
    return v * inv;
}
QMC_HD V3 reflect(V3 i, V3 n) { return i - (2.0f * n) * dot(n, i); }

// 2-D helpers (cylinder receiver works in the xz plane).
struct V2 {
    float x, y;
};
QMC_HD float dot2(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
QMC_HD V2 normalize2(V2 v) {
    float inv = inv_sqrt_cr(dot2(v, v));
    return V2{v.x * inv, v.y * inv};
}
QMC_HD float length2(V2 v) { return sqrt_cr(dot2(v, v)); }

// Local frame of an "aligned normal": u = normalize((0,1,0) x n), v = normalize(u x n); a local
// vector d maps to d.x*u + d.y*n + d.z*v. When |n.x| and |n.z| are below Epsilon the map is the
// identity (Util/global_function.cuh:19-36). The focus-flat variant additionally negates u and v
// when n.z < 0 (Scene/DataStructure/Heliostat/FocusFlatRectangleHelio.cuh:15-46).
struct Frame {
    V3 u, n, v;
    int identity;
};

QMC_HD Frame make_frame(V3 n, bool focus_variant) {
    Frame f;
    f.n = n;
    if (fabsf(n.x) < kEps && fabsf(n.z) < kEps) {
        f.identity = 1;
        f.u = V3{1.0f, 0.0f, 0.0f};
        f.v = V3{0.0f, 0.0f, 1.0f};
        return f;
    }
    f.identity = 0;
    V3 u = normalize(cross(V3{0.0f, 1.0f, 0.0f}, n));
    V3 v = normalize(cross(u, n));
    if (focus_variant && n.z < 0) {
        u = -u;
        v = -v;
    }
    f.u = u;
    f.v = v;
    return f;
}

QMC_HD V3 frame_apply(const Frame &f, V3 d) {
    if (f.identity) return d;
    return V3{d.x * f.u.x + d.y * f.n.x + d.z * f.v.x,
              d.x * f.u.y + d.y * f.n.y + d.z * f.v.y,
              d.x * f.u.z + d.y * f.n.z + d.z * f.v.z};
}

QMC_HD V3 local2world_rotate(V3 d, V3 n) { return frame_apply(make_frame(n, false), d); }
QMC_HD V3 focus_rotate(V3 d, V3 n) { return frame_apply(make_frame(n, true), d); }

// Yaw-only rotation taking old_dir to new_dir (Util/global_function.cuh:51-59).
QMC_HD V3 rotate_y(V3 p, V3 old_dir, V3 new_dir) {
    int dir = (cross(old_dir, new_dir).y > 0) ? 1 : -1;
    float c = dot(old_dir, new_dir);
    float s = dir * sqrtf(1 - c * c);
    return V3{c * p.x + s * p.z, p.y, -s * p.x + c * p.z};
}

// Transcendentals evaluated in double and rounded once to float. Both glibc and the device
// math library return these double results to within an ulp of double, so the rounded float is
// the same on host and device (it is the correctly rounded float except in cases closer than
// ~1e-16 relative to a rounding boundary). Replaces the reference's single-precision expf,
// acosf, atanf, powf, sinf, cosf at the places where they feed traced geometry.
QMC_HD float exp_r(float x) { return (float)exp((double)x); }
QMC_HD float acos_r(float x) { return (float)acos((double)x); }
QMC_HD float atan_r(float x) { return (float)atan((double)x); }
QMC_HD float pow_r(float x, float y) { return (float)pow((double)x, (double)y); }
QMC_HD float sin_r(float x) { return (float)sin((double)x); }
QMC_HD float cos_r(float x) { return (float)cos((double)x); }

// Cylinder map columns. A hit whose direction from the axis has cosine cx (clamped to [-1, 1]) and a
// negative (neg) or non-negative z component lies in column (int)(v * W) of a W-column map, v =
// float(double(acos_r(cx)) / 2 pi), or 1 - v when neg (CylinderReceiver.cuh:82-86 and the pixel index).
QMC_HD int cylinder_column_exact(float cx, bool neg, int W) {
    float v = (float)((double)acos_r(cx) * kInv2Pi);  // equals the division by 2 pi (exhaustively)
    if (neg) v = 1 - v;
    return (int)(v * (float)W);
}

// For each sign that column is a monotone step function of the float cx (non-increasing without neg,
// non-decreasing with it), so a receiver keeps its breakpoints, T[0 .. 2W + 3] (cylinder_column_table,
// built on the host with cylinder_column_exact): T[k] (k <= W + 1) = the smallest float cx in [-1, 1]
// whose column is below k, T[W + 2 + k] = the smallest whose column is at least k with neg; 2 where
// there is none. A hit takes an estimate from Abramowitz and Stegun 4.4.45 (acos x = sqrt(1 - x)
// (a0 + a1 x + a2 x^2 + a3 x^3) on [0, 1], error at most 6.7e-5 rad: under 0.35 of a column for W <=
// kColumnTableMaxW, so within one of the exact column) and corrects it with two breakpoints. Equal to
// cylinder_column_exact for every float cx and both signs (checked exhaustively:
// qmc_check_cylinder_columns).
constexpr int kColumnTableMaxW = 32768;
QMC_HD int cylinder_column(const float *T, int W, float cx, bool neg) {
    const float ax = fabsf(cx);
#if defined(__HIP_DEVICE_COMPILE__)
    const float s = __builtin_amdgcn_sqrtf(1.0f - ax);
#else
    const float s = sqrtf(1.0f - ax);
#endif
    float ac = s * (1.5707288f + ax * (-0.2121144f + ax * (0.0742610f - 0.0187293f * ax)));
    if (cx < 0.0f) ac = 3.14159265f - ac;
    float va = ac * 0.159154943f;
    if (neg) va = 1.0f - va;
    int k = (int)(va * (float)W);
    k = k < 0 ? 0 : k > W ? W : k;
    const float *B = T + (neg ? W + 2 : 0);
    const float b0 = B[k], b1 = B[k + 1];
    return k - 1 + (neg ? (int)(cx >= b0) + (int)(cx >= b1) : (int)(cx < b0) + (int)(cx < b1));
}

// The breakpoint table of a W-column cylinder map (2 W + 4 floats), by binary search over the floats of
// [-1, 1] in order (host).
inline int64_t float_order(float x) {
    const uint32_t b = __builtin_bit_cast(uint32_t, x);
    return (b & 0x80000000u) ? -(int64_t)(b & 0x7fffffffu) : (int64_t)b;
}
inline float order_float(int64_t k) {
    return __builtin_bit_cast(float, k >= 0 ? (uint32_t)k : (0x80000000u | (uint32_t)(-k)));
}
inline void cylinder_column_table(int W, float *T) {
    for (int neg = 0; neg < 2; ++neg) {
        for (int k = 0; k <= W + 1; ++k) {
            // the first cx at which the predicate holds: column below k (no neg) or at least k (neg)
            auto holds = [&](float cx) {
                const int c = cylinder_column_exact(cx, neg != 0, W);
                return neg ? c >= k : c < k;
            };
            float b;
            if (!holds(1.0f)) b = 2.0f;
            else if (holds(-1.0f)) b = -1.0f;
            else {
                int64_t lo = float_order(-1.0f), hi = float_order(1.0f);  // holds at hi, not at lo
                while (hi - lo > 1) {
                    const int64_t mid = lo + (hi - lo) / 2;
                    if (holds(order_float(mid))) hi = mid;
                    else lo = mid;
                }
                b = order_float(hi);
            }
            T[neg * (W + 2) + k] = b;
        }
    }
}

// Atmospheric attenuation and per-ray energy (QuasiMonteCarloRayTracing/Receiver/
// receiverIntersectionUtil.cuh:17-21,31-34).
QMC_HD float eta_atm(float d) {
    if (d <= 1000.0f) return 0.99331f - 0.0001176f * d + 1.97f * (1e-8f) * d * d;
    return exp_r(-0.0001106f * d);
}
QMC_HD float ray_energy(float t, V3 dir, V3 normal, float factor) {
    return fabsf(dot(dir, normal)) * eta_atm(t) * factor;
}

// Ray vs parallelogram (A, A+E1, A+E2) — Moeller-Trumbore on a parallelogram, the reference's
// rayParallelogramIntersect (Util/global_function.cuh:101-136) with E1 = B-A and E2 = C-A
// supplied precomputed (the subtraction is exact-same on host and device).
QMC_HD bool ray_quad(V3 o, V3 d, V3 A, V3 E1, V3 E2, float &t, float &u, float &v) {
    V3 p = cross(d, E2);
    float det = dot(E1, p);
    if (fabsf(det) <= kEpsF) return false;  // |det| < Epsilon
    V3 T = o - A;
    u = dot(T, p) / det;
    if (u < 0 || u > 1) return false;
    V3 q = cross(T, E1);
    v = dot(d, q) / det;
    if (v < 0 || v > 1) return false;
    t = dot(E2, q) / det;
    if (t <= kEpsF) return false;  // t < Epsilon
    return true;
}

// The terms of ray_quad that depend on the ray's origin only: T = o - A, q = T x E1 and E2 . q. Rays
// from one micro-facet share them; ray_quad_pre takes them precomputed and performs the same operations
// in the same order as ray_quad.
struct QuadOriginTerms {
    V3 T, q;
    float tq;
};
QMC_HD QuadOriginTerms quad_origin_terms(V3 o, V3 A, V3 E1, V3 E2) {
    QuadOriginTerms r;
    r.T = o - A;
    r.q = cross(r.T, E1);
    r.tq = dot(E2, r.q);
    return r;
}
QMC_HD bool ray_quad_pre(V3 d, V3 E1, V3 E2, const QuadOriginTerms &ot, float &t, float &u, float &v) {
    V3 p = cross(d, E2);
    float det = dot(E1, p);
    if (fabsf(det) <= kEpsF) return false;  // |det| < Epsilon
    u = dot(ot.T, p) / det;
    if (u < 0 || u > 1) return false;
    v = dot(d, ot.q) / det;
    if (v < 0 || v > 1) return false;
    t = ot.tq / det;
    if (t <= kEpsF) return false;  // t < Epsilon
    return true;
}

// DDA helpers (QuasiMonteCarloRayTracing/Grid/rectGridDDA.cuh:15-32).
// Exact quotient of 0 <= n < 2^31 by a divisor d >= 1 fixed ahead of the kernel, as a multiply-high:
// l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1, q = (mulhi(n, m) + n) >> l (the round-up
// multiplier of Granlund and Montgomery, "Division by invariant integers using multiplication",
// PLDI 1994); mulhi(n, m) < n keeps the sum below 2^32.
struct FastDiv {
    uint32_t m, l;
};
inline FastDiv make_fastdiv(uint32_t d) {
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    return FastDiv{(uint32_t)(((1ull << 32) * ((1ull << l) - d)) / d + 1), l};
}
QMC_HD uint32_t fast_div(uint32_t n, FastDiv f) {
    return (uint32_t)((((uint64_t)n * f.m) >> 32) + n) >> f.l;
}
QMC_HD uint32_t fast_mod(uint32_t n, uint32_t d, FastDiv f) { return n - fast_div(n, f) * d; }

// ---------------------------------------------------------------------------------------------
// Conservative ray-cone culling (double precision). cone_may_hit returns false only when no ray
// p + t*d with p in the box c +- hu +- hn +- hv (hn = 0: a rectangle), t >= 0 and d within angle
// theta (tan_theta, theta below 90 deg) of the unit axis a can meet the parallelogram A + s*E1 + r*E2
// (s, r in [0, 1]) — then every ray/parallelogram test of such a ray against it is a miss.
// Bound: along the axis a ray point x = p + t*d has (x-p).a = t*cos(phi) <= Tmax (the largest axial
// gap between a quad vertex and a box vertex) and a sideways offset |(x-p).w| = t*|d.w| <=
// Tmax*tan(theta) for any unit w orthogonal to a. So if, along some such w, the projections of the
// box and the quad are further apart than Tmax*tan(theta) + margin, or the quad lies wholly behind the
// box along a, no ray reaches it. Candidate w: a x (each edge of both shapes) and the direction
// between their centres orthogonal to a. `margin` (metres) absorbs the float rounding of traced
// origins, directions and the intersection test.
struct CD3 {
    double x, y, z;
};
QMC_HD CD3 cd3(V3 v) { return CD3{(double)v.x, (double)v.y, (double)v.z}; }
QMC_HD CD3 cadd(CD3 a, CD3 b) { return CD3{a.x + b.x, a.y + b.y, a.z + b.z}; }
QMC_HD CD3 csub(CD3 a, CD3 b) { return CD3{a.x - b.x, a.y - b.y, a.z - b.z}; }
QMC_HD CD3 cmul(CD3 a, double s) { return CD3{a.x * s, a.y * s, a.z * s}; }
QMC_HD double cdot(CD3 a, CD3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
QMC_HD CD3 ccross(CD3 a, CD3 b) { return CD3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// NP = 8: a box; NP = 4: a rectangle (hn is then ignored). Projections use support functions: on a
// direction w the box spans c.w -+ (|hu.w| + |hv.w| + |hn.w|) and the parallelogram spans A.w +
// [min(0, E1.w) + min(0, E2.w), max(0, E1.w) + max(0, E2.w)]. Axes are not normalised: a gap g
// between the two spans on an axis w of length |w| is a distance g / |w|, so "separated by more than
// reach" is g > 0 and g^2 > reach^2 |w|^2.
template <int NP>
QMC_HD bool cone_may_hit_n(CD3 c, CD3 hu, CD3 hn, CD3 hv, CD3 a, double tan_theta, CD3 A, CD3 E1, CD3 E2,
                           double margin) {
    auto box_r = [&](CD3 w) {
        return fabs(cdot(hu, w)) + fabs(cdot(hv, w)) + (NP == 8 ? fabs(cdot(hn, w)) : 0.0);
    };
    const double e1a = cdot(E1, a), e2a = cdot(E2, a);
    const double pa_min = cdot(c, a) - box_r(a);
    const double qa_max = cdot(A, a) + fmax(0.0, e1a) + fmax(0.0, e2a);
    const double tmax = qa_max - pa_min;
    if (tmax < -margin) return false;  // the quad lies wholly behind every origin
    const double reach = fmax(tmax, 0.0) * tan_theta + margin;
    const CD3 qc = cadd(A, cmul(cadd(E1, E2), 0.5));
    const CD3 dc = csub(qc, c);
    CD3 axes[6] = {ccross(a, hu), ccross(a, hv), ccross(a, E1), ccross(a, E2), csub(dc, cmul(a, cdot(dc, a))),
                   ccross(a, hn)};
#pragma unroll
    for (int k = 0; k < (NP == 8 ? 6 : 5); ++k) {
        const CD3 w = axes[k];
        const double ww = cdot(w, w);
        if (!(ww > 1e-24)) continue;
        const double cw = cdot(c, w), r = box_r(w);
        const double aw = cdot(A, w), e1 = cdot(E1, w), e2 = cdot(E2, w);
        const double q_lo = aw + fmin(0.0, e1) + fmin(0.0, e2), q_hi = aw + fmax(0.0, e1) + fmax(0.0, e2);
        const double gap = fmax(q_lo - (cw + r), (cw - r) - q_hi);
        if (gap > 0.0 && gap * gap > reach * reach * ww) return false;
    }
    return true;
}

QMC_HD bool cone_may_hit(CD3 c, CD3 hu, CD3 hn, CD3 hv, CD3 a, double tan_theta, CD3 A, CD3 E1, CD3 E2,
                         double margin) {
    return cdot(hn, hn) > 0.0 ? cone_may_hit_n<8>(c, hu, hn, hv, a, tan_theta, A, E1, E2, margin)
                              : cone_may_hit_n<4>(c, hu, hn, hv, a, tan_theta, A, E1, E2, margin);
}

// Cheaper, weaker form of the same bound for a pre-test: origins within `r` of a centre, the quad
// within that distance plus its own radius (together R) of its centre, D = quad centre - source centre.
// A hit point y = p + t*d needs (y-p).a >= 0, and (y-p) differs from D by at most R, so
// D.a < -R - margin rules a hit out; and |(y-p) orthogonal to a| <= ((y-p).a)*tan(theta) <=
// (D.a + R)*tan(theta), so |D orthogonal to a| - R > (D.a + R)*tan(theta) + margin rules it out.
QMC_HD bool sphere_may_hit(CD3 D, double R, CD3 a, double tan_theta, double margin) {
    const double da = cdot(D, a);
    if (da < -R - margin) return false;
    const CD3 perp = csub(D, cmul(a, da));
    const double s = (da + R) * tan_theta + R + margin;  // > 0 here
    return cdot(perp, perp) <= s * s;
}

// Inverse CDF of a tabulated density, u in [0, 1] -> angle (table units). Bin i spans
// [theta[i], theta[i+1]] with cumulative probability cdf[i] .. cdf[i+1] and end densities lo[i],
// hi[i]; inside a bin the density is taken linear between them and scaled to the bin's mass, so the
// cumulative is a quadratic in the offset s: s = 2m / (lo + sqrt(lo^2 + 2 (hi - lo) m / h)), the
// cancellation-free root, with m the target mass in the bin's own density units.
QMC_HD double table_inverse_cdf(const double *theta, const double *cdf, const double *lo, const double *hi, int n,
                                double u) {
    int a = 0, b = n;  // cdf[a] <= u; b == n or cdf[b] > u
    while (b - a > 1) {
        const int mid = (a + b) >> 1;
        if (cdf[mid] <= u) a = mid;
        else b = mid;
    }
    const double h = theta[a + 1] - theta[a], mass = cdf[a + 1] - cdf[a];
    const double w = mass > 0.0 ? fmin(fmax((u - cdf[a]) / mass, 0.0), 1.0) : 0.0;
    const double p0 = lo[a], p1 = hi[a];
    const double m = w * (p0 + p1) * 0.5 * h;
    const double den = p0 + sqrt(fmax(p0 * p0 + 2.0 * (p1 - p0) / h * m, 0.0));
    const double s = den > 0.0 ? 2.0 * m / den : 0.0;
    return theta[a] + fmin(fmax(s, 0.0), h);
}

QMC_HD float abs_divide(float den, float num) {
    if (num <= kEpsF && num >= -kEpsF) return kIntMaxF;  // |num| <= Epsilon
    return fabsf(den / num);
}
QMC_HD float cal_tmax(float dir, float interval, int idx, float cur) {
    return fabsf(float(idx + (dir >= 0)) * interval - cur);
}
QMC_HD bool dda_equal(float a, float b) { return (a < b + kEps) && (a > b - kEps); }
QMC_HD bool dda_less(float a, float b) { return a < b - kEps; }

// Proof that the reference's cell walk of the ray o + t*d (rectGridDDA.cu:31-119: start cell
// (int)((o - pos) / interval), crossing times accumulated in float, steps chosen by the Epsilon rule)
// visits a cell of the cell box [lo, hi] (inclusive indices) of a grid at pos with cell sizes s and n
// cells per axis (extent E = n*s), given a point p = o + tp*d, tp > 0, inside the box. Then a quad listed
// in every cell of that box and hit by the ray at tp is hit by the walk, unless the walk stops at another
// hit first: either way the traversal reports a hit. False means "not proven" (the caller walks).
// Argument, per axis a, in position units along a (exact arithmetic): the walk's crossing of a cell face
// happens at a float time whose position o_a + t*d_a differs from the face by at most
// err_a = 2 ulp(E) (start cell and first crossing) + per step (distance + cell) * 2^-24 (accumulation
// and the rounded step) + 2*Epsilon*|d_a|. A step processes the earliest unprocessed crossing together
// with others at most 2*Epsilon later, and each axis's crossings in their order. Take the first step that
// processes a crossing later than tp: the cell the walk is in just before it has processed every crossing
// earlier than tp - 2*Epsilon and none later than tp, so on every axis its index is that of a point
// within err_a of p_a. The walk gets there: it starts inside the grid (o clear of the grid's faces) and
// cannot cross a grid face before p, which is inside. So if p is farther than err_a from the box's faces
// on every axis, the walk visits a cell of the box. An axis with |d_a| <= Epsilon is never stepped
// (abs_divide): the walk keeps o's cell there, and p_a is taken as o_a. Distances are at most tp
// (|d| = 1), so the per-step term is bounded once per ray with the largest 1/s and s. The bound is
// widened four times and by 1 mm. Evaluation is in float: every quantity is at most E in size and takes a
// few roundings, together below ulp(E) * 8 = E * 2^-20, which `slack` adds to every margin.
//
// The proof's per-grid constants (positions relative to the grid's): the largest 1/s and s over the
// axes, the grid's centre E/2 and its half extent less the origin's margin (1 mm + slack).
struct ProofFrame {
    float pos[3], inv_max;
    float gmid[3], cs_max;
    float glim[3], pad_;
};

// A cell box as the proof reads it: its centre relative to the grid's position, and its half size less
// the face margin before per-ray terms (1 mm + 4 * 2 ulp(E) + slack); negative for an empty box, which
// the proof then never claims.
struct ProofBox {
    float mid[3], lim[3];
};


// x rounded to a float not above it.
inline float round_down(double x) {
    const float f = (float)x;
    return (double)f > x ? nextafterf(f, -INFINITY) : f;
}

// x rounded to a float not below it.
inline float round_up(double x) {
    const float f = (float)x;
    return (double)f < x ? nextafterf(f, INFINITY) : f;
}

inline double proof_slack(double ext) { return ext * 0x1p-20; }

inline ProofFrame proof_frame(const float pos[3], const float s[3], const int n[3]) {
    ProofFrame f{};
    for (int a = 0; a < 3; ++a) {
        const double ext = (double)n[a] * (double)s[a];
        f.pos[a] = pos[a];
        f.inv_max = fmaxf(f.inv_max, round_up(1.0 / (double)s[a]));
        f.cs_max = fmaxf(f.cs_max, s[a]);
        f.gmid[a] = (float)(0.5 * ext);
        f.glim[a] = round_down(0.5 * ext - 1e-3 - proof_slack(ext));
    }
    return f;
}

inline ProofBox proof_box(const float s[3], const int n[3], const int lo[3], const int hi[3]) {
    ProofBox b{};
    const bool empty = lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2];
    for (int a = 0; a < 3; ++a) {
        const double ext = (double)n[a] * (double)s[a];
        const double margin = 1e-3 + 4.0 * ext * 0x1p-22 + proof_slack(ext);
        b.mid[a] = (float)(0.5 * (double)(lo[a] + hi[a] + 1) * (double)s[a]);
        b.lim[a] = empty ? -1.0f : round_down(0.5 * (double)(hi[a] - lo[a] + 1) * (double)s[a] - margin);
    }
    return b;
}

QMC_HD bool dda_visit_proof(V3 o, V3 d, float tp, const ProofFrame &f, const float mid[3], const float lim[3]) {
    // straight-line: every condition is evaluated and combined, no early exit
    const float oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
    // four times the per-step term at distance tp plus 2*Epsilon: the per-ray part of every axis's margin
    const float S = 4.0f * fmaf(fmaf(tp, f.inv_max, 2.0f) * (tp + f.cs_max), 0x1p-24f, 2.0e-5f);
    bool ok = tp > 0.0f;
    for (int a = 0; a < 3; ++a) {
        const float ro = oo[a] - f.pos[a];  // the origin relative to the grid
        const bool still = fabsf(dd[a]) <= kEpsF;
        const float rp = still ? ro : fmaf(tp, dd[a], ro);
        ok = ok & (fabsf(rp - mid[a]) < lim[a] - S) & (fabsf(ro - f.gmid[a]) < f.glim[a]);
    }
    return ok;
}

// Counter-based hashing (splitmix64 finaliser) for seeded offsets and Gaussian slope errors.
QMC_HD uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// Uniform float in (0, 1] from the top 24 bits of a 32-bit word (exactly representable).
QMC_HD float u01_open0(uint32_t x) { return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f); }

}  // namespace qmc
