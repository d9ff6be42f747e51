// oracle.cpp — CPU restatement of the reference tracer. TEST INFRASTRUCTURE ONLY: it is the
// checker of the GPU path (tests/, __graft_entry__.smoke) and the CPU baseline of bench.py.
//
// Built with g++ -O3 -ffp-contract=off (no fast-math), so every float expression rounds as the
// reference's does on a host: one IEEE operation at a time, left to right. Epsilon comparisons are
// done in double as C++ promotion does with the reference's `#define Epsilon 1e-5`.
//
// Function -> reference (JachyLikeCoding/SolarEnergy_Chier) -> what pins it (tests/):
//   vec math (dot, cross, normalize = v*(1/sqrtf), reflect)  Util/vector_arithmetic.cuh:50-53,379-413
//                                                   -> the Moeller-Trumbore and energy closed forms
//                                                   of test_oracle_invariants.py that use them
//   local2world / focus rotate        Util/global_function.cuh:19-36; FocusFlatRectangleHelio.cuh:15-46
//                                                   -> orthonormality + KATs below
//   rotateY                           Util/global_function.cuh:51-59 -> receiver vertex closed forms
//   rayParallelogramIntersect         Util/global_function.cuh:101-136 -> textbook Moeller-Trumbore
//                                                   closed forms and brute-force grids
//   receiver init (rect / cylinder)   RectangleReceiver.cu:9-101; CylinderReceiver.cu:4-31
//                                                   -> CylinderReceiverTest KATs (resolution 57x12,
//                                                   focus centres), rectangle closed forms; the tilt
//                                                   and resolution-override extensions -> vertex
//                                                   closed forms of the tilted 12 x 12 m receiver and
//                                                   a parallel-sun energy balance independent of the
//                                                   pixel area (test_oracle_pins.py)
//   cylinder GIntersect               CylinderReceiver.cuh:39-101 -> CylinderReceiverTest KATs
//   grid Cinit + CGridHelioMatch      RectangleGrid.cu:8-117 -> RectangleGridTest KATs
//   CSetNormal / CSetWorldVertex      Heliostat.cu:15-40 -> QuasiMonteCarloRayTracerTest vertex KATs
//   flat micro-facets                 RectangleHelio.cu:54-116 -> RectangleHelioTest KATs,
//                                                   QuasiMonteCarloRayTracerTest origin KATs
//   focus-flat facets / micro-facets  FocusFlatRectangleHelio.cu:12-227,333-339
//                                                   -> FocusFlatRectangleHelioTest KATs, read method
//                                                   -1 (analytic) and 1 (subhelio data: the same KAT
//                                                   from its centres and normals, test_oracle_pins.py)
//   setFlatRectangleHeliostatVertexs + subHeliostat_id  QuasiMonteCarloRayTracer.cpp:135-209
//                                                   -> subHeliostat_id == 7 KAT
//   factor                            QuasiMonteCarloRayTracer.cpp:58-60 -> energy conservation
//   Buie table + inverse CDF          Scene/SceneProcessor.cu:7-81,196-243 -> numeric integration of
//                                                   the Buie profile, KS tests of the sampled theta:
//                                                   core against the table, circumsolar tail (after
//                                                   undoing the remap) against e^k/(g+1) t^(g+1)
//   angle2xyz / map_permuation        Scene/SceneProcessor.cu:83-103; global_function.cuh:211-216
//                                                   -> unit norm, phi uniformity, slope moments
//   map_float2int                     Tool/RandomNumberGenerator/RandomGenerator.cu:20-25 -> range
//   3D-DDA collision / intersect      QuasiMonteCarloRayTracing/Grid/rectGridDDA.{cu,cuh}
//                                                   -> brute-force "any non-own quad" equivalence
//   map_raytracing (rect / cylinder)  RectangleReceiverRectangleGridRayTracing.cu:6-41;
//                                     CylinderReceiverRectangleGridRayTracing.cu:6-40
//                                                   -> parallel-sun analytic scenes, energy balance
//   receiver_drawing / calEnergy      Receiver/*.cu; receiverIntersectionUtil.cuh:17-34
//                                                   -> energy conservation closed form
//   trimmed_mean + heaps              Tool/Smoother/ImageSmoother.cu:13-64; heap.cuh:16-225
//                                                   -> sort-based textbook trimmed mean (k <= 1; any
//                                                   k on windows read in ascending or descending
//                                                   order) and a hand-derived 5x5 KAT on which the
//                                                   reference's `else if` departs from the textbook
//                                                   (test_oracle_pins.py)
// test_oracle_pins.py also plants a slip in a copy of this file for each of these pins (the `else
// if`, a heap child, the tilt sign, the pixel area, the tail exponent, read method 1) and checks
// that the pin rejects it.
//   physical Buie sampler (QMC_SUNSHAPE_BUIE, no reference counterpart) -> KS test of sampled angles
//                                     against quadrature of Buie's profile; the table itself against
//                                     independent quadrature
//   Sobol / counter hash (this build's generators, no reference counterpart; the reference used
//   cuRAND XORWOW seeded by time(NULL)) -> Sobol net property, Gaussian moments
//
// Decisions (pinned identically in the product): u or v exactly 1 clamps the pixel index to
// res-1 and flags the ray; the next-facet offset read at m = F-1 reads entry F (F+1 offsets);
// the cylinder's cosine is clamped to [-1, 1] before acos; expf/acosf/atanf/powf/sinf/cosf in
// traced geometry are evaluated in double and rounded to float once.
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

const double EPS = 1e-5;  // Util/global_constant.h:11

struct f3 {
    float x, y, z;
};
struct f2 {
    float x, y;
};
inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
inline f3 add(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline f3 sub(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline f3 neg(f3 a) { return mk(-a.x, -a.y, -a.z); }
inline f3 scale(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
inline f3 scale_l(float s, f3 a) { return mk(s * a.x, s * a.y, s * a.z); }
inline f3 addf(f3 a, float s) { return mk(a.x + s, a.y + s, a.z + s); }
inline f3 subf(f3 a, float s) { return mk(a.x - s, a.y - s, a.z - s); }
inline f3 divv(f3 a, f3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
inline f3 divf(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline f3 cross(f3 a, f3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline float rsqrt_host(float x) { return 1.0f / sqrtf(x); }  // vector_arithmetic.cuh:50-53
inline f3 normalize(f3 v) { return scale(v, rsqrt_host(dot(v, v))); }
inline float length(f3 v) { return sqrtf(dot(v, v)); }
inline f3 reflect(f3 i, f3 n) { return sub(i, scale(scale_l(2.0f, n), dot(n, i))); }
inline float dot2(f2 a, f2 b) { return a.x * b.x + a.y * b.y; }
inline f2 normalize2(f2 v) {
    float r = rsqrt_host(dot2(v, v));
    return f2{v.x * r, v.y * r};
}
inline float length2(f2 v) { return sqrtf(dot2(v, v)); }
inline f3 load3(const float *p) { return mk(p[0], p[1], p[2]); }
inline void store3(float *p, f3 v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}

// Transcendentals: double evaluation, one rounding to float.
inline float expf_d(float x) { return (float)std::exp((double)x); }
inline float acosf_d(float x) { return (float)std::acos((double)x); }
inline float atanf_d(float x) { return (float)std::atan((double)x); }
inline float powf_d(float x, float y) { return (float)std::pow((double)x, (double)y); }
inline float sinf_d(float x) { return (float)std::sin((double)x); }
inline float cosf_d(float x) { return (float)std::cos((double)x); }

// global_func::local2world_rotate (Util/global_function.cuh:19-36)
f3 local2world_rotate(f3 d, f3 n) {
    if (std::fabs(n.x) < EPS && std::fabs(n.z) < EPS) return d;
    f3 u = normalize(cross(mk(0.0f, 1.0f, 0.0f), n));
    f3 v = normalize(cross(u, n));
    return mk(d.x * u.x + d.y * n.x + d.z * v.x, d.x * u.y + d.y * n.y + d.z * v.y,
              d.x * u.z + d.y * n.z + d.z * v.z);
}

// focusFlatRectangleHeliostatLocal2WorldRotate (FocusFlatRectangleHelio.cuh:15-46)
f3 focus_rotate(f3 d, f3 n) {
    if (std::fabs(n.x) < EPS && std::fabs(n.z) < EPS) return d;
    f3 u = normalize(cross(mk(0.0f, 1.0f, 0.0f), n));
    f3 v = normalize(cross(u, n));
    if (n.z < 0) {
        u = neg(u);
        v = neg(v);
    }
    return mk(d.x * u.x + d.y * n.x + d.z * v.x, d.x * u.y + d.y * n.y + d.z * v.y,
              d.x * u.z + d.y * n.z + d.z * v.z);
}

// global_func::rotateY (Util/global_function.cuh:51-59)
f3 rotateY(f3 p, f3 old_dir, f3 new_dir) {
    int dir = (cross(old_dir, new_dir).y > 0) ? 1 : -1;
    float c = dot(old_dir, new_dir);
    float s = dir * sqrtf(1 - c * c);
    return mk(c * p.x + s * p.z, p.y, -s * p.x + c * p.z);
}

// global_func::rayParallelogramIntersect (Util/global_function.cuh:101-136)
bool rayParallelogramIntersect(f3 orig, f3 dir, f3 A, f3 B, f3 C, float &t, float &u, float &v) {
    f3 AB = sub(B, A);
    f3 AC = sub(C, A);
    f3 p = cross(dir, AC);
    float det = dot(AB, p);
    if (std::fabs(det) < EPS) return false;
    f3 T = sub(orig, A);
    u = dot(T, p) / det;
    if (u < 0 || u > 1) return false;
    f3 q = cross(T, AB);
    v = dot(dir, q) / det;
    if (v < 0 || v > 1) return false;
    t = dot(AC, q) / det;
    if (t < EPS) return false;
    return true;
}

// receiverIntersectionUtil.cuh:17-34
float eta_aAlpha(float d) {
    if (d <= 1000.0f) return 0.99331f - 0.0001176f * d + 1.97f * (1e-8f) * d * d;
    return expf_d(-0.0001106f * d);
}
float calEnergy(float distance, f3 dir, f3 normal, float factor) {
    return std::fabs(dot(dir, normal)) * eta_aAlpha(distance) * factor;
}


// ------------------------------------------------------------------------------------------
// Scene model (reference data structures)
// ------------------------------------------------------------------------------------------
struct Receiver {
    qmc_receiver_desc d{};
    int res_x = 0, res_y = 0;
    float pixel_length = 0.0f, area = 0.0f;
    bool override_res = false;
    f3 local_normal{0, 0, 0}, normal{0, 0, 0};
    f3 v[4]{};
    f3 focus{0, 0, 0};
};

struct Grid {
    qmc_grid_desc d{};
    int nx = 0, ny = 0, nz = 0;
    std::vector<int> index;  // d_grid_helio_index_ (cells + 1)
    std::vector<int> match;  // d_grid_helio_match_
    std::vector<f3> verts;   // sub-heliostat vertex array of the grid
};

struct Helio {
    qmc_heliostat_desc d{};
    std::vector<float> fc, fn;  // read method 1 local centres / normals
    int type = 0;
    f3 normal{0, 0, 0}, aim{0, 0, 0};
    f3 vertex[4]{};
    float focal = 0.0f, pl = 0.0f;
    int sub_size = 1, sub_id = 0;
    f2 sublen{0, 0};          // subhelio_row_col_length (x: along rows/z, y: along cols/x)
    int sub_rows = 0, sub_cols = 0, F = 0;
    std::vector<f3> lc, ln;   // focus-flat local centres / normals (CGetSubHeliostatVertexes)
};

}  // namespace

struct orc_scene {
    qmc_config cfg{};
    std::vector<Receiver> R;
    std::vector<Grid> G;
    std::vector<Helio> H;
};

namespace {

// RectangleReceiver::CInit (RectangleReceiver.cu:9-16,40-101)
void rect_receiver_init(Receiver &r, int ppm) {
    r.pixel_length = 1.0f / float(ppm);
    f3 size = load3(r.d.size);
    switch (r.d.face) {
        case 0: r.local_normal = mk(0.0f, 0.0f, 1.0f); break;
        case 1: r.local_normal = mk(1.0f, 0.0f, 0.0f); break;
        case 2: r.local_normal = mk(0.0f, 0.0f, -1.0f); break;
        case 3: r.local_normal = mk(-1.0f, 0.0f, 0.0f); break;
        default: break;
    }
    float hx = size.x / 2, hy = size.y / 2, hz = size.z / 2;
    switch (r.d.face) {
        case 0:
            r.v[0] = mk(-size.x / 2, -size.y / 2, hz); r.v[1] = mk(-size.x / 2, hy, hz);
            r.v[2] = mk(hx, hy, hz); r.v[3] = mk(hx, -size.y / 2, hz);
            break;
        case 1:
            r.v[0] = mk(hx, -size.y / 2, hz); r.v[1] = mk(hx, hy, hz);
            r.v[2] = mk(hx, hy, -size.z / 2); r.v[3] = mk(hx, -size.y / 2, -size.z / 2);
            break;
        case 2:
            r.v[0] = mk(hx, -size.y / 2, -size.z / 2); r.v[1] = mk(hx, hy, -size.z / 2);
            r.v[2] = mk(-size.x / 2, hy, -size.z / 2); r.v[3] = mk(-size.x / 2, -size.y / 2, -size.z / 2);
            break;
        case 3:
            r.v[0] = mk(-size.x / 2, -size.y / 2, -size.z / 2); r.v[1] = mk(-size.x / 2, hy, -size.z / 2);
            r.v[2] = mk(-size.x / 2, hy, hz); r.v[3] = mk(-size.x / 2, -size.y / 2, hz);
            break;
        default: break;
    }
    r.normal = normalize(load3(r.d.normal));
    if (r.d.tilt_deg != 0.0f) {
        double a = (double)r.d.tilt_deg * 3.14159265358979323846 / 180.0;
        float ca = (float)std::cos(a), sa = (float)std::sin(a);
        for (int k = 0; k < 4; ++k) {
            f3 p = r.v[k];
            r.v[k] = mk(p.x, p.y * ca - p.z * sa, p.y * sa + p.z * ca);
        }
    }
    for (int k = 0; k < 4; ++k) r.v[k] = rotateY(r.v[k], r.local_normal, r.normal);
    for (int k = 0; k < 4; ++k) r.v[k] = add(r.v[k], load3(r.d.pos));
    r.focus = divf(add(r.v[0], r.v[2]), 2);
    r.res_x = (int)(size.x * float(ppm));
    r.res_y = (int)(size.y * float(ppm));
    r.area = r.pixel_length * r.pixel_length;
    if (r.d.res_x > 0 && r.d.res_y > 0) {
        r.override_res = true;
        r.res_x = r.d.res_x;
        r.res_y = r.d.res_y;
        r.pixel_length = size.x / float(r.res_x);
        r.area = (size.x / float(r.res_x)) * (size.y / float(r.res_y));
    }
}

// CylinderReceiver::CInit / Cset_resolution (CylinderReceiver.cu:4-15)
void cyl_receiver_init(Receiver &r, int ppm) {
    r.pixel_length = 1.0f / float(ppm);
    r.res_x = int(std::ceil(2 * M_PI * r.d.size[0] * float(ppm)));
    r.res_y = int(r.d.size[1] * float(ppm));
    r.area = r.pixel_length * r.pixel_length;
    r.normal = load3(r.d.normal);
    if (r.d.res_x > 0 && r.d.res_y > 0) {
        r.override_res = true;
        r.res_x = r.d.res_x;
        r.res_y = r.d.res_y;
        float circ = (float)(2 * M_PI * r.d.size[0]);
        r.pixel_length = circ / float(r.res_x);
        r.area = (circ / float(r.res_x)) * (r.d.size[1] / float(r.res_y));
    }
}

bool cyl_inside(const float pos[3], const float size[3], f3 o) {
    f2 l{o.x - pos[0], o.z - pos[2]};
    return dot2(l, l) <= size[0] * size[0];
}

// CylinderReceiver::getFocusCenter (CylinderReceiver.cu:18-31)
f3 cyl_focus(const float pos[3], const float size[3], f3 hpos) {
    f3 p = load3(pos);
    if (cyl_inside(pos, size, hpos)) return p;
    f3 dir = normalize(sub(hpos, p));
    float radius = size[0] / length2(f2{dir.x, dir.z});
    float x = p.x + dir.x * radius;
    float z = p.z + dir.z * radius;
    return mk(x, p.y, z);
}

// CylinderReceiver::GIntersect (CylinderReceiver.cuh:39-89); the cosine is clamped before acos
bool cyl_intersect(const float pos[3], const float size[3], f3 orig, f3 dir, float &t, float &u, float &v) {
    if (cyl_inside(pos, size, orig)) return false;
    f2 Ro{pos[0] - orig.x, pos[2] - orig.z};
    float tp = dot2(Ro, normalize2(f2{dir.x, dir.z}));
    float delta = dot2(Ro, Ro) - tp * tp;
    float R2 = size[0] * size[0];
    if (tp < -EPS || delta > R2) return false;
    float t_plus = delta <= 0.0f ? size[0] : sqrtf(R2 - delta);
    t = (tp - t_plus) / length2(f2{dir.x, dir.z});
    f3 ip = add(scale_l(t, dir), orig);
    u = (ip.y - pos[1]) / size[1] + 0.5f;
    if (u < 0.0f || u > 1.0f) return false;
    f2 c = normalize2(f2{ip.x - pos[0], ip.z - pos[2]});
    float cx = c.x < -1.0f ? -1.0f : (c.x > 1.0f ? 1.0f : c.x);
    v = acosf_d(cx) / (2 * M_PI);
    if (c.y < 0) v = 1 - v;
    return true;
}

int helio_sub_size(const Helio &h) {
    if (h.type == 1) return h.d.rows * h.d.cols;  // FocusFlatRectangleHelio::getSubHelioSize
    return h.d.sub_count_override > 0 ? h.d.sub_count_override : 1;
}

// RectangleGrid::Cinit + boxIntersect + CGridHelioMatch (RectangleGrid.cu:8-117)
void grid_build(Grid &g, const std::vector<Helio> &hs) {
    f3 pos = load3(g.d.pos), size = load3(g.d.size), inter = load3(g.d.interval);
    g.nx = int(std::ceil(size.x / inter.x));
    g.ny = int(std::ceil(size.y / inter.y));
    g.nz = int(std::ceil(size.z / inter.z));
    std::vector<std::vector<int>> vec((size_t)g.nx * g.ny * g.nz);
    int start = 0;
    for (int i = g.d.start_heliostat; i < g.d.start_heliostat + g.d.n_heliostats; ++i) {
        f3 p = load3(hs[i].d.pos);
        float radius = length(load3(hs[i].d.size)) / 2;
        f3 minPos = subf(p, radius), maxPos = addf(p, radius);
        f3 a = divv(sub(minPos, pos), inter), b = divv(sub(maxPos, pos), inter);
        int mnx = int(a.x), mny = int(a.y), mnz = int(a.z);
        int mxx = int(b.x), mxy = int(b.y), mxz = int(b.z);
        int ss = helio_sub_size(hs[i]);
        for (int x = std::max(0, mnx); x <= std::min(g.nx - 1, mxx); ++x)
            for (int y = std::max(0, mny); y <= std::min(g.ny - 1, mxy); ++y)
                for (int z = std::max(0, mnz); z <= std::min(g.nz - 1, mxz); ++z) {
                    int id = x * g.ny * g.nz + y * g.nz + z;
                    for (int k = 0; k < ss; ++k) vec[id].push_back(start + k);
                }
        start += ss;
    }
    g.index.assign(vec.size() + 1, 0);
    g.match.clear();
    for (size_t i = 0; i < vec.size(); ++i) {
        g.index[i + 1] = int(g.index[i] + vec[i].size());
        for (int v : vec[i]) g.match.push_back(v);
    }
}

// Heliostat::CSetNormalAndRotate (+ focus-flat focal reset) and the micro-facet layout
void helio_process(Helio &h, f3 focus_center, f3 sun_dir, float pl) {
    h.aim = focus_center;
    h.pl = pl;
    f3 pos = load3(h.d.pos), size = load3(h.d.size);
    h.focal = h.d.surface_property[0];
    if (h.type == 1 && h.focal < 0.0f) h.focal = length(sub(focus_center, pos));
    f3 reflect_dir = normalize(sub(focus_center, pos));
    h.normal = normalize(sub(reflect_dir, sun_dir));
    h.vertex[0] = mk(-size.x / 2, size.y / 2, -size.z / 2);
    h.vertex[1] = add(mk(0, 0, size.z), h.vertex[0]);
    h.vertex[2] = add(mk(size.x, 0, size.z), h.vertex[0]);
    h.vertex[3] = add(mk(size.x, 0, 0), h.vertex[0]);
    for (int k = 0; k < 4; ++k) h.vertex[k] = local2world_rotate(h.vertex[k], h.normal);
    for (int k = 0; k < 4; ++k) h.vertex[k] = add(h.vertex[k], pos);
    h.sublen.x = (size.z - h.d.gap[1] * (h.d.rows - 1)) / float(h.d.rows);
    h.sublen.y = (size.x - h.d.gap[0] * (h.d.cols - 1)) / float(h.d.cols);
    h.sub_rows = int(h.sublen.x / pl);
    h.sub_cols = int(h.sublen.y / pl);
    h.F = h.d.rows * h.d.cols * h.sub_rows * h.sub_cols;
}

// FocusFlatRectangleHelio::CGetSubHeliostatVertexes (FocusFlatRectangleHelio.cu:123-227)
void focus_sub_vertices(Helio &h, std::vector<f3> &out) {
    f3 size = load3(h.d.size), pos = load3(h.d.pos);
    const f2 L = h.sublen;
    f3 local[3] = {mk(-L.y / 2, 0.0f, L.x / 2), mk(-L.y / 2, 0.0f, -L.x / 2), mk(L.y / 2, 0.0f, -L.x / 2)};
    int n = h.d.rows * h.d.cols;
    h.lc.assign(n, mk(0, 0, 0));
    h.ln.assign(n, mk(0, 0, 0));
    int method = int(h.d.surface_property[1]);
    for (int r = 0; r < h.d.rows; ++r)
        for (int c = 0; c < h.d.cols; ++c) {
            int id = r * h.d.cols + c;
            if (method == 1 && (int)h.fc.size() >= 3 * n && (int)h.fn.size() >= 3 * n) {
                h.lc[id] = load3(&h.fc[3 * id]);
                h.ln[id] = normalize(load3(&h.fn[3 * id]));
            } else {
                f3 cc = mk(c * (h.d.gap[0] + L.y) - size.x / 2 + L.y / 2, size.y / 2,
                           r * (h.d.gap[1] + L.x) - size.z / 2 + L.x / 2);
                cc.y += (cc.x * cc.x + cc.z * cc.z) / (4 * h.focal);
                h.lc[id] = cc;
                h.ln[id] = normalize(mk(-1 / (2 * h.focal) * cc.x, 1.0f, -1 / (2 * h.focal) * cc.z));
            }
            for (f3 sv : local) {
                sv = focus_rotate(sv, h.ln[id]);
                sv = add(sv, h.lc[id]);
                sv = focus_rotate(sv, h.normal);
                sv = add(sv, pos);
                out.push_back(sv);
            }
        }
}

// micro-facet origins and normals of a heliostat, materialised as the reference does
void micro_facets(const Helio &h, std::vector<f3> &origins, std::vector<f3> &normals) {
    origins.resize(h.F);
    normals.resize(h.F);
    f3 pos = load3(h.d.pos), size = load3(h.d.size);
    const int rc_y = h.d.cols;
    for (int myId = 0; myId < h.F; ++myId) {
        int row = myId / (rc_y * h.sub_cols);
        int col = myId % (rc_y * h.sub_cols);
        if (h.type == 0) {
            // rectangle_heliostat::map_microhelio_center_and_normal (RectangleHelio.cu:54-80)
            int block_row = row / h.sub_rows;
            int block_col = col / h.sub_cols;
            f3 c;
            c.x = col * h.pl + block_col * h.d.gap[0] + h.pl / 2 - size.x / 2;
            c.y = size.y / 2;
            c.z = row * h.pl + block_row * h.d.gap[1] + h.pl / 2 - size.z / 2;
            normals[myId] = h.normal;
            c = local2world_rotate(c, h.normal);
            origins[myId] = add(c, pos);
        } else {
            // focusFlatRectangle_heliostat kernels (FocusFlatRectangleHelio.cu:12-64)
            int block_id = (row / h.sub_rows) * rc_y + (col / h.sub_cols);
            f3 ln = h.ln[block_id];
            row %= h.sub_rows;
            col %= h.sub_cols;
            f3 subsize = mk(h.sublen.y, size.y, h.sublen.x);
            f3 lp = mk((col + 0.5f) * h.pl - subsize.x / 2, 0.0f, (row + 0.5f) * h.pl - subsize.z / 2);
            lp = focus_rotate(lp, ln);
            lp = add(lp, h.lc[block_id]);
            normals[myId] = focus_rotate(ln, h.normal);
            lp = focus_rotate(lp, h.normal);
            origins[myId] = add(lp, pos);
        }
    }
}

float factor_of(const orc_scene &s, int h) {
    const Helio &hh = s.H[h];
    const Receiver &r = s.R[s.G[hh.d.grid].d.receiver];
    if (!r.override_res) {
        float ratio = hh.pl / r.pixel_length;
        return s.cfg.dni * ratio * ratio * s.cfg.helio_reflected_rate / float(s.cfg.num_sunshape_lights_per_group);
    }
    return s.cfg.dni * (hh.pl * hh.pl / r.area) * s.cfg.helio_reflected_rate /
           float(s.cfg.num_sunshape_lights_per_group);
}

// ------------------------------------------------------------------------------------------
// 3D-DDA (QuasiMonteCarloRayTracing/Grid/rectGridDDA.{cu,cuh})
// ------------------------------------------------------------------------------------------
inline float absDivide(float den, float num) {
    if (num <= EPS && num >= -EPS) return float(2147483647);
    return std::fabs(den / num);
}
inline float calTMax(float dir, float interval, int idx, float cur) {
    return std::fabs(float(idx + (dir >= 0)) * interval - cur);
}
inline bool equal_(float a, float b) { return (a < b + EPS) && (a > b - EPS); }
inline bool less_(float a, float b) { return a < b - EPS; }

struct Counters {
    uint64_t cells = 0, candidates = 0, tests = 0;
};

bool intersect(f3 o, f3 d, const std::vector<f3> &V, const std::vector<int> &match, int b, int e, int own,
               int nown, int &hit, Counters *cnt) {
    float t, u, v;
    for (int i = b; i < e; ++i) {
        int sid = match[i];
        if (sid < own || sid >= own + nown) {
            if (cnt) cnt->tests++;
            if (rayParallelogramIntersect(o, d, V[3 * sid + 1], V[3 * sid], V[3 * sid + 2], t, u, v)) {
                hit = sid;
                return true;
            }
        }
    }
    return false;
}

bool collision(f3 o, f3 d, const Grid &g, int own, int nown, int &hit, Counters *cnt,
               std::vector<int> *visited = nullptr) {
    f3 gp = load3(g.d.pos), gi = load3(g.d.interval);
    f3 q = divv(sub(o, gp), gi);
    int px = int(q.x), py = int(q.y), pz = int(q.z);
    int sx = (d.x >= 0) ? 1 : -1, sy = (d.y >= 0) ? 1 : -1, sz = (d.z >= 0) ? 1 : -1;
    float tMx = absDivide(calTMax(d.x, gi.x, px, o.x - gp.x), d.x);
    float tMy = absDivide(calTMax(d.y, gi.y, py, o.y - gp.y), d.y);
    float tMz = absDivide(calTMax(d.z, gi.z, pz, o.z - gp.z), d.z);
    float tDx = absDivide(gi.x, d.x), tDy = absDivide(gi.y, d.y), tDz = absDivide(gi.z, d.z);
    while (px >= 0 && px < g.nx && py >= 0 && py < g.ny && pz >= 0 && pz < g.nz) {
        int addr = px * g.ny * g.nz + py * g.nz + pz;
        if (visited) visited->push_back(addr);
        if (cnt) {
            cnt->cells++;
            cnt->candidates += (uint64_t)(g.index[addr + 1] - g.index[addr]);
        }
        if (intersect(o, d, g.verts, g.match, g.index[addr], g.index[addr + 1], own, nown, hit, cnt)) return true;
        bool xf = false, yf = false, zf = false;
        if (less_(tMx, tMy)) {
            if (less_(tMx, tMz) || equal_(tMx, tMz)) { px += sx; xf = true; }
            if (less_(tMz, tMx) || equal_(tMx, tMz)) { pz += sz; zf = true; }
        } else if (less_(tMy, tMx)) {
            if (less_(tMy, tMz) || equal_(tMy, tMz)) { py += sy; yf = true; }
            if (less_(tMz, tMy) || equal_(tMy, tMz)) { pz += sz; zf = true; }
        } else {
            if (less_(tMx, tMz) || equal_(tMx, tMz)) { px += sx; py += sy; xf = yf = true; }
            if (less_(tMz, tMx) || equal_(tMx, tMz)) { pz += sz; zf = true; }
        }
        tMx += xf ? tDx : 0;
        tMy += yf ? tDy : 0;
        tMz += zf ? tDz : 0;
    }
    return false;
}

// ------------------------------------------------------------------------------------------
// Pools (Scene/SceneProcessor.cu) and generators
// ------------------------------------------------------------------------------------------
struct Buie {
    std::vector<float> cdf;  // (x, y) pairs
    std::vector<float> k;
    int n = 0;
    float A = 0, B = 0, C = 0, I = 0, gamma = 0;
};

Buie buie_table(const qmc_config &c) {
    Buie b;
    b.n = c.inverse_transform_sampling_groups;
    float csr = c.csr, upper = 4.65f;
    float kap = 0.9f * logf(13.5f * csr) * powf(csr, -0.3f);
    float gam = 2.2f * logf(0.52f * csr) * powf(csr, 0.43f) - 0.1f;
    auto intensity = [](float th) { return cosf(0.326f * th) / cosf(0.308f * th); };
    auto tail = [&](float th) { return expf(kap) / (gam + 1) * powf(th, gam + 1); };
    float between = tail(9.3f) - tail(upper);
    float len = upper / float(b.n);
    b.cdf.assign(2 * (b.n + 1), 0.0f);
    b.k.assign(b.n, 0.0f);
    float x = 0.0f, pre = intensity(x);
    for (int i = 1; i <= b.n; ++i) {
        x += len;
        float cur = intensity(x);
        float subarea = (cur + pre) / 2.0f * len;
        b.cdf[2 * i] = x;
        b.cdf[2 * i + 1] = subarea + b.cdf[2 * (i - 1) + 1];
        b.k[i - 1] = len / subarea;
        pre = cur;
    }
    b.I = b.cdf[2 * b.n + 1];
    b.A = b.I + between;
    b.B = (gam + 1) / expf(kap);
    b.C = powf(upper, gam + 1);
    b.gamma = gam;
    return b;
}

int max_less_index(const Buie &b, float value) {
    int left = 0, right = b.n, mid;
    while (left <= right) {
        mid = (left + right) >> 1;
        if (value > b.cdf[2 * mid + 1]) left = mid + 1;
        else if (value < b.cdf[2 * mid + 1]) right = mid - 1;
        else return mid;
    }
    return right;
}

float buie_theta(const Buie &b, float U) {
    float u = U * b.A;
    if (u < b.I) {
        int id = max_less_index(b, u);
        u = (u - b.cdf[2 * id + 1]) * b.k[id] + b.cdf[2 * id];
        u = atanf_d(sqrtf(u / 4.65f) * 0.5016916f);
    } else {
        u = powf_d((u - b.I) * b.B + b.C, 1 / (b.gamma + 1));
        float ri2 = 0.5016916f * 0.5016916f;
        float ro2 = 1.340874f * 1.340874f;
        u = atanf_d(sqrtf(ro2 + (u - 4.65f) / 4.65f * (ro2 - ri2)));
    }
    return u / 100.0f;
}

uint64_t splitmix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
float unit_open0(uint32_t x) { return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f); }

void sobol_dirs(uint32_t v0[32], uint32_t v1[32]) {
    for (int k = 0; k < 32; ++k) v0[k] = 1u << (31 - k);
    uint32_t m = 1;
    for (int k = 1; k <= 32; ++k) {
        if (k > 1) m = (m << 1) ^ m;
        v1[k - 1] = m << (32 - k);
    }
}

// Gray-code-free Sobol point: XOR of direction numbers of the set bits of i.
uint32_t sobol(const uint32_t v[32], uint32_t i) {
    uint32_t x = 0;
    for (int k = 0; i; ++k, i >>= 1)
        if (i & 1u) x ^= v[k];
    return x;
}

}  // namespace

// ------------------------------------------------------------------------------------------
// C API
// ------------------------------------------------------------------------------------------
extern "C" orc_scene *orc_scene_build(const qmc_config *cfg, const qmc_receiver_desc *R, int nr,
                                      const qmc_grid_desc *G, int ng, const qmc_heliostat_desc *H, int nh, char *err,
                                      int errlen) {
    auto set_err = [&](const std::string &m) {
        if (err && errlen > 0) std::snprintf(err, errlen, "%s", m.c_str());
    };
    auto *s = new orc_scene();
    s->cfg = *cfg;
    for (int i = 0; i < nr; ++i) {
        Receiver r;
        r.d = R[i];
        s->R.push_back(r);
    }
    for (int i = 0; i < ng; ++i) {
        Grid g;
        g.d = G[i];
        if (g.d.receiver < 0 || g.d.receiver >= nr || g.d.start_heliostat < 0 ||
            g.d.start_heliostat + g.d.n_heliostats > nh) {
            set_err("grid " + std::to_string(i) + " is inconsistent");
            delete s;
            return nullptr;
        }
        s->G.push_back(g);
    }
    for (int i = 0; i < nh; ++i) {
        Helio h;
        h.d = H[i];
        int n = h.d.rows * h.d.cols;
        if (h.d.local_centers) h.fc.assign(h.d.local_centers, h.d.local_centers + 3 * n);
        if (h.d.local_normals) h.fn.assign(h.d.local_normals, h.d.local_normals + 3 * n);
        h.d.local_centers = h.d.local_normals = nullptr;
        if (h.d.grid < 0 || h.d.grid >= ng) {
            set_err("heliostat " + std::to_string(i) + " has no grid");
            delete s;
            return nullptr;
        }
        h.type = G[h.d.grid].heliostat_type;
        s->H.push_back(std::move(h));
    }
    // SceneProcessor::set_grid_receiver_heliostat_content (SceneProcessor.cpp:13-39)
    int ppm = int(1.0f / cfg->receiver_pixel_length);
    f3 sun = load3(cfg->sun_dir);
    for (auto &g : s->G) {
        grid_build(g, s->H);
        Receiver &r = s->R[g.d.receiver];
        if (r.d.type == 0) rect_receiver_init(r, ppm);
        else cyl_receiver_init(r, ppm);
        for (int i = 0; i < g.d.n_heliostats; ++i) {
            Helio &h = s->H[g.d.start_heliostat + i];
            f3 fc = (r.d.type == 0) ? r.focus : cyl_focus(r.d.pos, r.d.size, load3(h.d.pos));
            if (h.d.aim_offset[0] != 0.0f || h.d.aim_offset[1] != 0.0f) {
                // aim-offset extension: focus + off0 * across + off1 * up in the receiver surface
                f3 across, up;
                if (r.d.type == 0) {
                    across = normalize(sub(r.v[3], r.v[0]));
                    up = normalize(sub(r.v[1], r.v[0]));
                } else {
                    f3 rad = sub(fc, load3(r.d.pos));
                    across = (rad.x == 0.0f && rad.z == 0.0f) ? mk(1.0f, 0.0f, 0.0f) : normalize(mk(-rad.z, 0.0f, rad.x));
                    up = mk(0.0f, 1.0f, 0.0f);
                }
                fc = add(add(fc, scale(across, h.d.aim_offset[0])), scale(up, h.d.aim_offset[1]));
            }
            helio_process(h, fc, sun, cfg->helio_pixel_length);
        }
        // QuasiMonteCarloRayTracer::setFlatRectangleHeliostatVertexs and the subHeliostat_id sums
        g.verts.clear();
        int sub = 0;
        for (int i = 0; i < g.d.n_heliostats; ++i) {
            Helio &h = s->H[g.d.start_heliostat + i];
            h.sub_size = helio_sub_size(h);
            h.sub_id = sub;
            sub += h.sub_size;
            if (h.type == 0) {
                g.verts.push_back(h.vertex[0]);
                g.verts.push_back(h.vertex[1]);
                g.verts.push_back(h.vertex[2]);
            } else {
                focus_sub_vertices(h, g.verts);
            }
        }
    }
    return s;
}

extern "C" void orc_scene_free(orc_scene *s) { delete s; }

extern "C" int orc_receiver_info(const orc_scene *s, int r, qmc_receiver_info *o) {
    if (r < 0 || r >= (int)s->R.size()) return -1;
    const Receiver &rr = s->R[r];
    std::memset(o, 0, sizeof(*o));
    o->type = rr.d.type;
    o->res_x = rr.res_x;
    o->res_y = rr.res_y;
    o->pixel_length = rr.pixel_length;
    o->factor_area = rr.area;
    if (rr.d.type == 0)
        for (int k = 0; k < 4; ++k) store3(o->vertices[k], rr.v[k]);
    store3(o->normal, rr.normal);
    store3(o->focus_center, rr.focus);
    return 0;
}

extern "C" int orc_heliostat_info(const orc_scene *s, int h, qmc_heliostat_info *o) {
    if (h < 0 || h >= (int)s->H.size()) return -1;
    const Helio &hh = s->H[h];
    std::memset(o, 0, sizeof(*o));
    store3(o->pos, load3(hh.d.pos));
    store3(o->size, load3(hh.d.size));
    store3(o->normal, hh.normal);
    store3(o->aim, hh.aim);
    for (int k = 0; k < 4; ++k) store3(o->vertices[k], hh.vertex[k]);
    o->focal_length = hh.focal;
    o->type = hh.type;
    o->grid = hh.d.grid;
    o->n_micro = hh.F;
    o->sub_rows = hh.sub_rows;
    o->sub_cols = hh.sub_cols;
    o->sub_begin = hh.sub_id;
    o->sub_count = hh.sub_size;
    return 0;
}

extern "C" int orc_grid_info(const orc_scene *s, int g, qmc_grid_info *o) {
    if (g < 0 || g >= (int)s->G.size()) return -1;
    const Grid &gg = s->G[g];
    o->nx = gg.nx;
    o->ny = gg.ny;
    o->nz = gg.nz;
    o->n_match = (int)gg.match.size();
    o->n_sub = (int)gg.verts.size() / 3;
    return 0;
}

extern "C" int orc_grid_cells(const orc_scene *s, int g, int32_t *cs, int32_t *m) {
    if (g < 0 || g >= (int)s->G.size()) return -1;
    const Grid &gg = s->G[g];
    if (cs) std::copy(gg.index.begin(), gg.index.end(), cs);
    if (m) std::copy(gg.match.begin(), gg.match.end(), m);
    return 0;
}

extern "C" int orc_subheliostat_vertices(const orc_scene *s, int g, float *out) {
    if (g < 0 || g >= (int)s->G.size()) return -1;
    const Grid &gg = s->G[g];
    if (out)
        for (size_t i = 0; i < gg.verts.size(); ++i) store3(out + 3 * i, gg.verts[i]);
    return (int)gg.verts.size();
}

extern "C" float orc_factor(const orc_scene *s, int h) { return factor_of(*s, h); }

extern "C" int orc_micro_count(const orc_scene *s, int h) { return s->H[h].F; }

extern "C" int orc_facets(const orc_scene *s, int h, float *origins, float *normals) {
    if (h < 0 || h >= (int)s->H.size()) return -1;
    std::vector<f3> o, n;
    micro_facets(s->H[h], o, n);
    for (size_t i = 0; i < o.size(); ++i) {
        if (origins) store3(origins + 3 * i, o[i]);
        if (normals) store3(normals + 3 * i, n[i]);
    }
    return (int)o.size();
}

extern "C" void orc_sunshape_table(const qmc_config *cfg, float *cdf_xy, float *slopes, float *abcig) {
    Buie b = buie_table(*cfg);
    if (cdf_xy) std::copy(b.cdf.begin(), b.cdf.end(), cdf_xy);
    if (slopes) std::copy(b.k.begin(), b.k.end(), slopes);
    if (abcig) {
        abcig[0] = b.A;
        abcig[1] = b.B;
        abcig[2] = b.C;
        abcig[3] = b.I;
        abcig[4] = b.gamma;
    }
}

// Physical Buie sampler (no reference counterpart): theta [mrad] with cumulative probability u, from
// a table of bins [theta_i, theta_i+1] with masses cdf_i+1 - cdf_i and the density linear between the
// bin's end values lo_i, hi_i. The bin is the last one whose cumulative start does not exceed u (a
// scan here); inside it the quadratic cumulative lo*s + (hi-lo)*s^2/(2h) = w*(lo+hi)*h/2 is solved as
// s = 2m / (lo + sqrt(lo^2 + 2(hi-lo)m/h)). Pinned by the KS test of sampled angles against the
// quadrature of Buie's profile (tests).
double table_theta(const double *theta, const double *cdf, const double *lo, const double *hi, int n, double u) {
    int a = 0;
    while (a + 1 < n && cdf[a + 1] <= u) ++a;
    double h = theta[a + 1] - theta[a], mass = cdf[a + 1] - cdf[a];
    double w = 0.0;
    if (mass > 0.0) w = std::min(std::max((u - cdf[a]) / mass, 0.0), 1.0);
    double m = w * (lo[a] + hi[a]) * 0.5 * h;
    double den = lo[a] + std::sqrt(std::max(lo[a] * lo[a] + 2.0 * (hi[a] - lo[a]) / h * m, 0.0));
    double s = den > 0.0 ? 2.0 * m / den : 0.0;
    return theta[a] + std::min(std::max(s, 0.0), h);
}

extern "C" void orc_table_theta(const double *theta, const double *cdf, const double *lo, const double *hi, int n,
                                const double *u, int count, double *out) {
    for (int i = 0; i < count; ++i) out[i] = table_theta(theta, cdf, lo, hi, n, u[i]);
}

struct PhysTable {
    const double *theta, *cdf, *lo, *hi;
    int n;
};

void pools_from_uniforms(const qmc_config *cfg, const PhysTable *phys, const float *ut, const float *up,
                         const float *gx, const float *gz, int P, float *dirs, float *perts) {
    Buie b = buie_table(*cfg);
    for (int i = 0; i < P; ++i) {
        // samplelights::linear_interpolation (or the physical table) + map_angle2xyz
        float theta = phys ? (float)(table_theta(phys->theta, phys->cdf, phys->lo, phys->hi, phys->n, (double)ut[i]) *
                                     1e-3)
                           : buie_theta(b, ut[i]);
        float phi = up[i] * 2 * 3.141592654f;
        float st = sinf_d(theta);
        store3(dirs + 3 * i, mk(st * cosf_d(phi), cosf_d(theta), st * sinf_d(phi)));
        // samplelights::map_permuation
        store3(perts + 3 * i, normalize(mk(gx[i], 1.0f, gz[i])));
    }
}

extern "C" void orc_pool_from_uniforms(const qmc_config *cfg, const float *ut, const float *up, const float *gx,
                                       const float *gz, int P, float *dirs, float *perts) {
    pools_from_uniforms(cfg, nullptr, ut, up, gx, gz, P, dirs, perts);
}

void seeded_uniforms(const qmc_config *cfg, uint64_t seed_sun, uint64_t seed_pert, std::vector<float> &ut,
                     std::vector<float> &up, std::vector<float> &gx, std::vector<float> &gz);

extern "C" void orc_pool_generate_buie(const qmc_config *cfg, uint64_t seed_sun, uint64_t seed_pert,
                                       const double *theta, const double *cdf, const double *lo, const double *hi,
                                       int n, float *dirs, float *perts) {
    std::vector<float> ut, up, gx, gz;
    seeded_uniforms(cfg, seed_sun, seed_pert, ut, up, gx, gz);
    PhysTable t{theta, cdf, lo, hi, n};
    pools_from_uniforms(cfg, &t, ut.data(), up.data(), gx.data(), gz.data(), (int)ut.size(), dirs, perts);
}

extern "C" void orc_pool_generate(const qmc_config *cfg, uint64_t seed_sun, uint64_t seed_pert, float *dirs,
                                  float *perts, float *ut_out, float *up_out, float *gx_out, float *gz_out) {
    std::vector<float> ut, up, gx, gz;
    seeded_uniforms(cfg, seed_sun, seed_pert, ut, up, gx, gz);
    const int P = (int)ut.size();
    orc_pool_from_uniforms(cfg, ut.data(), up.data(), gx.data(), gz.data(), P, dirs, perts);
    if (ut_out) std::copy(ut.begin(), ut.end(), ut_out);
    if (up_out) std::copy(up.begin(), up.end(), up_out);
    if (gx_out) std::copy(gx.begin(), gx.end(), gx_out);
    if (gz_out) std::copy(gz.begin(), gz.end(), gz_out);
}

// Scrambled Sobol (theta, phi) uniforms and counter-hash Box-Muller slope errors of a seeded pool.
void seeded_uniforms(const qmc_config *cfg, uint64_t seed_sun, uint64_t seed_pert, std::vector<float> &ut,
                     std::vector<float> &up, std::vector<float> &gx, std::vector<float> &gz) {
    const int P = cfg->num_sunshape_groups * cfg->num_sunshape_lights_per_group;
    ut.assign(P, 0.0f);
    up.assign(P, 0.0f);
    gx.assign(P, 0.0f);
    gz.assign(P, 0.0f);
    uint32_t v0[32], v1[32];
    sobol_dirs(v0, v1);
    uint32_t m0 = (uint32_t)(splitmix(seed_sun) >> 32);
    uint32_t m1 = (uint32_t)(splitmix(seed_sun ^ 0xA5A5A5A5A5A5A5A5ull) >> 32);
    uint64_t key = splitmix(seed_pert);
    for (int i = 0; i < P; ++i) {
        ut[i] = unit_open0(sobol(v0, (uint32_t)i) ^ m0);
        up[i] = unit_open0(sobol(v1, (uint32_t)i) ^ m1);
        uint64_t w1 = splitmix(key + 2ull * (uint32_t)i), w2 = splitmix(key + 2ull * (uint32_t)i + 1ull);
        double d1 = (double)((w1 >> 11) + 1ull) * (1.0 / 9007199254740992.0);
        double d2 = (double)(w2 >> 11) * (1.0 / 9007199254740992.0);
        double rad = std::sqrt(-2.0 * std::log(d1));
        double ang = 6.283185307179586 * d2;
        gx[i] = (float)((double)cfg->helio_disturb_std * rad * std::cos(ang));
        gz[i] = (float)((double)cfg->helio_disturb_std * rad * std::sin(ang));
    }
}

extern "C" void orc_offsets_generate(const orc_scene *s, uint64_t seed, int P, int h, int32_t *out) {
    const int F = s->H[h].F;
    for (int m = 0; m <= F; ++m) {
        uint64_t w = splitmix(splitmix(seed ^ ((uint64_t)(uint32_t)h << 32)) + (uint64_t)(uint32_t)m);
        float u = unit_open0((uint32_t)(w >> 32));
        out[m] = (int32_t)(u * float(P - 0) + 0);  // map_float2int
    }
}

namespace {

// map_raytracing for one heliostat over rays [r0, r1) with stride
void trace_range(const orc_scene &s, int h, const std::vector<f3> &origins, const std::vector<f3> &normals,
                 const float *dirs, const float *perts, int P, int L, const int32_t *grp, double *flux,
                 qmc_ray_record *rec, qmc_trace_stats *st, int64_t r0, int64_t r1, int64_t stride) {
    const Helio &H = s.H[h];
    const Grid &g = s.G[H.d.grid];
    const Receiver &R = s.R[g.d.receiver];
    const float factor = factor_of(s, h);
    const f3 sun = load3(s.cfg.sun_dir);
    Counters cnt;
    Counters *pc = st ? &cnt : nullptr;
    uint64_t unsh = 0, unbl = 0, dep = 0, edge = 0, rays = 0, fac = 0;
    for (int64_t myId = r0; myId < r1; myId += stride) {
        int m = (int)(myId / L), j = (int)(myId % L);
        ++rays;
        if (j == 0) ++fac;
        qmc_ray_record out{QMC_RAY_MISSED, -1, 0.0f, 0};
        int address = (int)((grp[m] + myId % L) % P);
        f3 dir = local2world_rotate(load3(dirs + 3 * address), neg(sun));
        f3 origin = origins[m];
        int hit = -1;
        if (collision(origin, dir, g, H.sub_id, H.sub_size, hit, pc)) {
            out.status = QMC_RAY_SHADOWED;
            out.index = hit;
        } else {
            ++unsh;
            f3 normal = normals[m];
            address = (int)((grp[(m + 1) % P] + myId % L) % P);
            normal = local2world_rotate(load3(perts + 3 * address), normal);
            normal = normalize(normal);
            dir = normalize(reflect(neg(dir), normal));
            if (collision(origin, dir, g, H.sub_id, H.sub_size, hit, pc)) {
                out.status = QMC_RAY_BLOCKED;
                out.index = hit;
            } else {
                ++unbl;
                float t, u, v;
                bool ok = (R.d.type == 0) ? rayParallelogramIntersect(origin, dir, R.v[0], R.v[1], R.v[3], t, u, v)
                                          : cyl_intersect(R.d.pos, R.d.size, origin, dir, t, u, v);
                if (ok) {
                    float energy = calEnergy(t, dir, normal, factor);
                    int row = int(u * R.res_y), col = int(v * R.res_x);
                    if (row >= R.res_y) { row = R.res_y - 1; out.flags |= QMC_RAYF_EDGE_CLAMPED; }
                    if (col >= R.res_x) { col = R.res_x - 1; out.flags |= QMC_RAYF_EDGE_CLAMPED; }
                    int addr = row * R.res_x + col;
                    flux[addr] += (double)energy;
                    out.status = QMC_RAY_DEPOSITED;
                    out.index = addr;
                    out.energy = energy;
                    ++dep;
                    if (out.flags) ++edge;
                }
            }
        }
        if (rec) rec[myId] = out;
    }
    if (st) {
        st->rays += rays;
        st->facets += fac;
        st->unshadowed += unsh;
        st->unblocked += unbl;
        st->deposits += dep;
        st->edge_rays += edge;
        st->cells += cnt.cells;
        st->candidates += cnt.candidates;
        st->tests += cnt.tests;
    }
}

}  // namespace

extern "C" int orc_trace_heliostat(const orc_scene *s, int h, const float *dirs, const float *perts, int P, int L,
                                   const int32_t *offsets, double *flux, qmc_ray_record *rec, qmc_trace_stats *st) {
    if (h < 0 || h >= (int)s->H.size()) return -1;
    std::vector<f3> o, n;
    micro_facets(s->H[h], o, n);
    trace_range(*s, h, o, n, dirs, perts, P, L, offsets, flux, rec, st, 0, (int64_t)s->H[h].F * L, 1);
    return 0;
}

extern "C" int orc_trace_many(const orc_scene *s, const int32_t *ids, int n, const float *dirs, const float *perts,
                              int P, int L, const int32_t *offsets_all, double *flux, int nthreads, int64_t stride,
                              qmc_trace_stats *st) {
    if (n <= 0) return 0;
    std::vector<int64_t> base(s->H.size() + 1, 0);
    for (size_t h = 0; h < s->H.size(); ++h) base[h + 1] = base[h] + s->H[h].F + 1;
    int rid = s->G[s->H[ids[0]].d.grid].d.receiver;
    for (int i = 0; i < n; ++i)
        if (s->G[s->H[ids[i]].d.grid].d.receiver != rid) return -2;
    const Receiver &R = s->R[rid];
    const size_t npix = (size_t)R.res_x * R.res_y;
    if (nthreads < 1) nthreads = 1;
    if (stride < 1) stride = 1;
    std::vector<std::vector<double>> maps(nthreads, std::vector<double>(npix, 0.0));
    std::vector<qmc_trace_stats> stats(nthreads);
    std::memset(stats.data(), 0, sizeof(qmc_trace_stats) * nthreads);
    std::atomic<int> next(0);
    auto worker = [&](int t) {
        std::vector<f3> o, nn;
        for (;;) {
            int i = next.fetch_add(1);
            if (i >= n) break;
            int h = ids[i];
            micro_facets(s->H[h], o, nn);
            trace_range(*s, h, o, nn, dirs, perts, P, L, offsets_all + base[h], maps[t].data(), nullptr,
                        st ? &stats[t] : nullptr, 0, (int64_t)s->H[h].F * L, stride);
        }
    };
    std::vector<std::thread> th;
    for (int t = 1; t < nthreads; ++t) th.emplace_back(worker, t);
    worker(0);
    for (auto &x : th) x.join();
    for (int t = 0; t < nthreads; ++t)
        for (size_t p = 0; p < npix; ++p) flux[p] += maps[t][p];
    if (st)
        for (int t = 0; t < nthreads; ++t) {
            st->rays += stats[t].rays;
            st->facets += stats[t].facets;
            st->unshadowed += stats[t].unshadowed;
            st->unblocked += stats[t].unblocked;
            st->deposits += stats[t].deposits;
            st->edge_rays += stats[t].edge_rays;
            st->cells += stats[t].cells;
            st->candidates += stats[t].candidates;
            st->tests += stats[t].tests;
        }
    return 0;
}

// ------------------------------------------------------------------------------------------
// Smoother (Tool/Smoother/ImageSmoother.cu:13-64, heap.cuh:16-225)
// ------------------------------------------------------------------------------------------
namespace {
class Heap {
public:
    Heap(int size, float *entry, bool maxheap) : num_(1), max_(size + 1), e_(entry), max_heap_(maxheap) {}
    bool full() const { return num_ == max_; }
    bool empty() const { return num_ == 1; }
    float top() const { return empty() ? 0.0f : e_[1]; }
    bool insert(float x) {
        if (full()) return false;
        int i = num_++;
        for (; i != 1 && cmp(x, e_[i / 2]); i /= 2) e_[i] = e_[i / 2];
        e_[i] = x;
        return true;
    }
    float delete_and_insert(float x) {
        if (empty()) return -100;
        float ret = e_[1];
        int i = 1, c;
        for (; 2 * i < num_;) {
            c = (2 * i + 1 >= num_ || cmp(e_[2 * i], e_[2 * i + 1])) ? 2 * i : 2 * i + 1;
            if (cmp(e_[c], x)) {
                e_[i] = e_[c];
                i = c;
            } else {
                break;
            }
        }
        e_[i] = x;
        return ret;
    }
    float sum() const {
        float s = 0.0f;
        for (int i = 1; i < num_; ++i) s += e_[i];
        return s;
    }

private:
    bool cmp(float a, float b) const { return max_heap_ ? a > b : a < b; }
    int num_, max_;
    float *e_;
    bool max_heap_;
};
}  // namespace

extern "C" void orc_smooth(float *img, int w, int h, int radius, float ratio) {
    std::vector<float> src(img, img + (size_t)w * h);
    int kalloc = (int)((2 * radius + 1) * (2 * radius + 1) * ratio + 1);
    std::vector<float> maxe(kalloc + 1), mine(kalloc + 1);
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            int lx = (x - radius >= 0) ? x - radius : 0, ly = (y - radius >= 0) ? y - radius : 0;
            int rx = (x + radius < w) ? x + radius : w - 1, ry = (y + radius < h) ? y + radius : h - 1;
            int cnt = (rx - lx + 1) * (ry - ly + 1);
            int k = (int)(cnt * ratio);
            Heap mx(k, maxe.data(), true), mn(k, mine.data(), false);
            float value = 0.0f;
            for (int r = ly; r <= ry; ++r)
                for (int c = lx; c <= rx; ++c) {
                    float cur = src[(size_t)r * w + c];
                    value += cur;
                    if (!mx.full()) {
                        mx.insert(cur);
                        mn.insert(cur);
                    } else if (cur < mx.top()) {
                        mx.delete_and_insert(cur);
                    } else if (cur > mn.top()) {
                        mn.delete_and_insert(cur);
                    }
                }
            img[(size_t)y * w + x] = (value - mx.sum() - mn.sum()) / float(cnt - 2 * k);
        }
}

// ------------------------------------------------------------------------------------------
// Primitives
// ------------------------------------------------------------------------------------------
extern "C" int orc_ray_parallelogram(const float o[3], const float d[3], const float A[3], const float B[3],
                                     const float C[3], float tuv[3]) {
    float t = 0, u = 0, v = 0;
    bool ok = rayParallelogramIntersect(load3(o), load3(d), load3(A), load3(B), load3(C), t, u, v);
    tuv[0] = t;
    tuv[1] = u;
    tuv[2] = v;
    return ok ? 1 : 0;
}

extern "C" int orc_cylinder_intersect(const float pos[3], const float size[3], const float o[3], const float d[3],
                                      float tuv[3]) {
    float t = 0, u = 0, v = 0;
    bool ok = cyl_intersect(pos, size, load3(o), load3(d), t, u, v);
    tuv[0] = t;
    tuv[1] = u;
    tuv[2] = v;
    return ok ? 1 : 0;
}

extern "C" void orc_cylinder_focus(const float pos[3], const float size[3], const float hpos[3], float out[3]) {
    store3(out, cyl_focus(pos, size, load3(hpos)));
}

extern "C" int orc_cylinder_resolution(const float size[3], int ppm, int res[2]) {
    Receiver r;
    std::memset(&r.d, 0, sizeof(r.d));
    r.d.type = 1;
    for (int k = 0; k < 3; ++k) r.d.size[k] = size[k];
    cyl_receiver_init(r, ppm);
    res[0] = r.res_x;
    res[1] = r.res_y;
    return 0;
}

extern "C" int orc_dda(const orc_scene *s, int g, const float o[3], const float d[3], int own_lo, int own_hi,
                       int *hit) {
    int hh = -1;
    bool r = collision(load3(o), load3(d), s->G[g], own_lo, own_hi - own_lo, hh, nullptr);
    if (hit) *hit = hh;
    return r ? 1 : 0;
}

extern "C" int orc_dda_cells(const orc_scene *s, int g, const float o[3], const float d[3], int32_t *cells, int cap) {
    const Grid &gg = s->G[g];
    std::vector<int> visited;
    int hh = -1;
    collision(load3(o), load3(d), gg, 0, (int)gg.verts.size() / 3, hh, nullptr, &visited);
    const int n = (int)visited.size();
    for (int i = 0; i < n && i < cap; ++i) cells[i] = visited[(size_t)i];
    return n;
}

extern "C" int orc_brute_any_hit(const orc_scene *s, int g, const float o[3], const float d[3], int own_lo,
                                 int own_hi) {
    const Grid &gg = s->G[g];
    int nsub = (int)gg.verts.size() / 3;
    for (int i = 0; i < nsub; ++i) {
        if (i >= own_lo && i < own_hi) continue;
        float t, u, v;
        if (rayParallelogramIntersect(load3(o), load3(d), gg.verts[3 * i + 1], gg.verts[3 * i], gg.verts[3 * i + 2], t,
                                      u, v))
            return 1;
    }
    return 0;
}

// ------------------------------------------------------------------------------------------
// FP64 host adjoint of the differentiable trace (extension; test infrastructure). Forward-mode
// dual numbers in double over the masked model: per visible ray the derivatives with respect to
// its heliostat's aim offset (across, up) and the sun direction (3 components). Visibility is
// decided by the float reference trace above (shadowing, blocking, receiver bounds) and frozen.
// ------------------------------------------------------------------------------------------
namespace {

struct D5 {
    double v;
    double d[5];
};
inline D5 dc(double v) {
    D5 r;
    r.v = v;
    for (double &x : r.d) x = 0.0;
    return r;
}
inline D5 operator+(const D5 &a, const D5 &b) {
    D5 r;
    r.v = a.v + b.v;
    for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
inline D5 operator-(const D5 &a, const D5 &b) {
    D5 r;
    r.v = a.v - b.v;
    for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
inline D5 operator-(const D5 &a) { return dc(0.0) - a; }
inline D5 operator*(const D5 &a, const D5 &b) {
    D5 r;
    r.v = a.v * b.v;
    for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
inline D5 operator*(const D5 &a, double s) {
    D5 r;
    r.v = a.v * s;
    for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] * s;
    return r;
}
inline D5 operator/(const D5 &a, const D5 &b) {
    D5 r;
    r.v = a.v / b.v;
    for (int k = 0; k < 5; ++k) r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) / (b.v * b.v);
    return r;
}
inline D5 dsqrt(const D5 &a) {
    D5 r;
    r.v = std::sqrt(a.v);
    for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] / (2.0 * r.v);
    return r;
}
inline D5 dexp(const D5 &a) {
    D5 r;
    r.v = std::exp(a.v);
    for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] * r.v;
    return r;
}
inline D5 dabs(const D5 &a) { return a.v >= 0 ? a : -a; }

struct V5 {
    D5 x, y, z;
};
inline V5 vc(double x, double y, double z) { return V5{dc(x), dc(y), dc(z)}; }
inline V5 operator+(const V5 &a, const V5 &b) { return V5{a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V5 operator-(const V5 &a, const V5 &b) { return V5{a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V5 operator*(const V5 &a, const D5 &s) { return V5{a.x * s, a.y * s, a.z * s}; }
inline V5 operator*(const V5 &a, double s) { return V5{a.x * s, a.y * s, a.z * s}; }
inline D5 vdot(const V5 &a, const V5 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V5 vcross(const V5 &a, const V5 &b) {
    return V5{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline V5 vnorm(const V5 &a) {
    D5 l = dsqrt(vdot(a, a));
    return V5{a.x / l, a.y / l, a.z / l};
}

}  // namespace

extern "C" int orc_adjoint_fp64(const orc_scene *s, int grid, const float *dirs, const float *perts, int P, int L,
                                const int32_t *offsets_all, float target, float peak, float lambda, int fd_helio,
                                const double *delta, double *flux_out, double *loss_out, double *grad_aim,
                                double *grad_sun, int64_t *n_visible) {
    if (grid < 0 || grid >= (int)s->G.size()) return -1;
    const Grid &g = s->G[grid];
    const Receiver &R = s->R[g.d.receiver];
    if (R.d.type != 0) return -2;
    std::vector<int64_t> base(s->H.size() + 1, 0);
    for (size_t h = 0; h < s->H.size(); ++h) base[h + 1] = base[h] + s->H[h].F + 1;
    const int resx = R.res_x, resy = R.res_y;
    const size_t npix = (size_t)resx * resy;
    std::vector<double> flux(npix, 0.0);
    // pass 1: visibility from the float reference trace
    std::vector<std::vector<qmc_ray_record>> recs(g.d.n_heliostats);
    for (int k = 0; k < g.d.n_heliostats; ++k) {
        int h = g.d.start_heliostat + k;
        if (s->H[h].type != 0) return -3;
        std::vector<double> tmp(npix, 0.0);
        recs[k].resize((size_t)s->H[h].F * L);
        orc_trace_heliostat(s, h, dirs, perts, P, L, offsets_all + base[h], tmp.data(), recs[k].data(), nullptr);
    }
    const f3 sunf = load3(s->cfg.sun_dir);
    const double Fx = R.focus.x, Fy = R.focus.y, Fz = R.focus.z;
    f3 acr = normalize(sub(R.v[3], R.v[0])), upv = normalize(sub(R.v[1], R.v[0]));
    const V5 A0 = vc(R.v[0].x, R.v[0].y, R.v[0].z);
    f3 e1 = sub(R.v[1], R.v[0]), e2 = sub(R.v[3], R.v[0]);
    const V5 E1 = vc(e1.x, e1.y, e1.z), E2 = vc(e2.x, e2.y, e2.z);
    const V5 Y = vc(0.0, 1.0, 0.0);
    // per-ray model evaluated on duals; mode 0 = forward flux, mode 1 = gradient with G
    std::vector<double> G(npix, 0.0);
    int64_t nvis = 0;
    for (int mode = 0; mode < 2; ++mode) {
        for (int k = 0; k < g.d.n_heliostats; ++k) {
            const int h = g.d.start_heliostat + k;
            const Helio &H = s->H[h];
            const float factor = factor_of(*s, h);
            const double dsh = (delta && h == fd_helio) ? 1.0 : 0.0;
            D5 o0 = dc(H.d.aim_offset[0] + dsh * (delta ? delta[0] : 0.0));
            D5 o1 = dc(H.d.aim_offset[1] + dsh * (delta ? delta[1] : 0.0));
            o0.d[0] = 1.0;
            o1.d[1] = 1.0;
            V5 sv = vc(sunf.x + (delta ? delta[2] : 0.0), sunf.y + (delta ? delta[3] : 0.0),
                       sunf.z + (delta ? delta[4] : 0.0));
            sv.x.d[2] = 1.0;
            sv.y.d[3] = 1.0;
            sv.z.d[4] = 1.0;
            const V5 p = vc(H.d.pos[0], H.d.pos[1], H.d.pos[2]);
            const V5 A = vc(Fx, Fy, Fz) + vc(acr.x, acr.y, acr.z) * o0 + vc(upv.x, upv.y, upv.z) * o1;
            const V5 n = vnorm(vnorm(A - p) - sv);
            const V5 u = vnorm(vcross(Y, n));
            const V5 v = vnorm(vcross(u, n));
            const V5 mm = vc(0, 0, 0) - sv;
            const V5 us = vnorm(vcross(Y, mm));
            const V5 vs = vnorm(vcross(us, mm));
            const int32_t *grp = offsets_all + base[h];
            const int rc_y = H.d.cols;
            double ga0 = 0.0, ga1 = 0.0;
            for (int64_t r = 0; r < (int64_t)H.F * L; ++r) {
                if (recs[k][r].status != QMC_RAY_DEPOSITED) continue;
                if (mode == 0) ++nvis;
                const int m = (int)(r / L), j = (int)(r % L);
                const int row = m / (rc_y * H.sub_cols), col = m % (rc_y * H.sub_cols);
                const double lx = col * (double)H.pl + (col / H.sub_cols) * (double)H.d.gap[0] + H.pl / 2.0 -
                                  H.d.size[0] / 2.0;
                const double ly = H.d.size[1] / 2.0;
                const double lz = row * (double)H.pl + (row / H.sub_rows) * (double)H.d.gap[1] + H.pl / 2.0 -
                                  H.d.size[2] / 2.0;
                const float *qi = dirs + 3 * ((grp[m] + j) % P);
                const float *qp = perts + 3 * ((grp[(m + 1) % P] + j) % P);
                const V5 O = p + u * lx + n * ly + v * lz;
                const V5 din = us * (double)qi[0] + mm * (double)qi[1] + vs * (double)qi[2];
                const V5 N = vnorm(u * (double)qp[0] + n * (double)qp[1] + v * (double)qp[2]);
                const V5 ii = vc(0, 0, 0) - din;
                const V5 rr = ii - N * (vdot(N, ii) * 2.0);
                const V5 d = vnorm(rr);
                const V5 pv = vcross(d, E2);
                const D5 det = vdot(E1, pv);
                const V5 T = O - A0;
                const D5 uu = vdot(T, pv) / det;
                const V5 q = vcross(T, E1);
                const D5 vv = vdot(d, q) / det;
                const D5 t = vdot(E2, q) / det;
                const D5 eta = t.v <= 1000.0 ? dc(0.99331) - t * 0.0001176 + t * t * 1.97e-8 : dexp(t * -0.0001106);
                const D5 E = dabs(vdot(d, N)) * eta * (double)factor;
                const D5 x = vv * (double)resx - dc(0.5), yv = uu * (double)resy - dc(0.5);
                const double fx0 = std::floor(x.v), fy0 = std::floor(yv.v);
                const D5 fx = x - dc(fx0), fy = yv - dc(fy0);
                const D5 wx[2] = {dc(1.0) - fx, fx}, wy[2] = {dc(1.0) - fy, fy};
                for (int b = 0; b < 2; ++b)
                    for (int a = 0; a < 2; ++a) {
                        const int pi = (int)fx0 + a, pj = (int)fy0 + b;
                        if (pi < 0 || pi >= resx || pj < 0 || pj >= resy) continue;
                        const D5 c = E * wx[a] * wy[b];
                        const size_t pix = (size_t)pj * resx + pi;
                        if (mode == 0) {
                            flux[pix] += c.v;
                        } else {
                            ga0 += G[pix] * c.d[0];
                            ga1 += G[pix] * c.d[1];
                            for (int kk = 0; kk < 3; ++kk) grad_sun[kk] += G[pix] * c.d[2 + kk];
                        }
                    }
            }
            if (mode == 1) {
                grad_aim[2 * h] = ga0;
                grad_aim[2 * h + 1] = ga1;
            }
        }
        if (mode == 0) {
            double loss = 0.0;
            for (size_t i = 0; i < npix; ++i) {
                const double e = flux[i] - target, over = std::max(flux[i] - (double)peak, 0.0);
                G[i] = 2.0 * e + 2.0 * lambda * over;
                loss += e * e + lambda * over * over;
            }
            if (loss_out) *loss_out = loss;
            for (int kk = 0; kk < 3; ++kk) grad_sun[kk] = 0.0;
            for (size_t hh = 0; hh < s->H.size(); ++hh) grad_aim[2 * hh] = grad_aim[2 * hh + 1] = 0.0;
        }
    }
    if (flux_out) std::copy(flux.begin(), flux.end(), flux_out);
    if (n_visible) *n_visible = nvis;
    return 0;
}
