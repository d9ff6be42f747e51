// Check of the walk-free blocking proof (dda_visit_proof in qmc_math.h) against the oracle's traversal.
// Built by tests/test_dda_proof.py as a shared library beside the oracle and called with the handle of an
// oracle scene: for many rays of one grid (from micro-facet origins and random points on heliostats, to
// points on nearby quads, many of them pushed onto cell faces, cell edges and near-axis directions), every
// quad the ray hits (the trace's own ray/parallelogram test, non-own quads) is offered to the proof with
// the quad's cell box (from the oracle's cell lists, checked to be a full box). Whenever the proof holds,
// the oracle's traversal of that ray must report a hit, and — the proof's own claim — visit a cell of the
// box. The claim is checked as well for the smallest box, the one cell holding the hit point (the
// origin's index on an axis the traversal never steps), where the traversal passing a cell by at an edge
// (stepping two axes at once) shows at once.
//
//   int dda_proof_check(const orc_scene *s, int g, const float pos[3], const float interval[3],
//                       int n_rays, uint64_t seed, long long out[8])
// out: rays, rays with a hit, proofs, oracle hits, violations (proof without an oracle hit), boxes not full,
// proofs (quad boxes and single cells) whose box the traversal does not visit, single-cell proofs.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

#include "oracle.h"
#include "qmc_math.h"

using namespace qmc;

namespace {

struct Quad {
    V3 A, E1, E2;
    int helio;
};

}  // namespace

extern "C" int dda_proof_check(const orc_scene *s, int g, const float pos[3], const float interval[3], int n_rays,
                               uint64_t seed, long long out[8]) {
    for (int k = 0; k < 8; ++k) out[k] = 0;
    qmc_grid_info gi{};
    if (orc_grid_info(s, g, &gi)) return 1;
    const int n[3] = {gi.nx, gi.ny, gi.nz};
    const size_t ncell = (size_t)gi.nx * gi.ny * gi.nz;
    std::vector<int32_t> start(ncell + 1), match((size_t)std::max(gi.n_match, 1));
    if (orc_grid_cells(s, g, start.data(), match.data())) return 2;
    // each quad's cell box from the cell lists, and whether every cell of the box lists it
    std::vector<int> lo(3 * (size_t)gi.n_sub, 1 << 30), hi(3 * (size_t)gi.n_sub, -1);
    std::vector<long long> listed((size_t)gi.n_sub, 0);
    for (size_t c = 0; c < ncell; ++c) {
        const int x = (int)(c / ((size_t)gi.ny * gi.nz)), y = (int)((c / gi.nz) % gi.ny), z = (int)(c % gi.nz);
        for (int i = start[c]; i < start[c + 1]; ++i) {
            const int q = match[(size_t)i];
            const int xyz[3] = {x, y, z};
            for (int a = 0; a < 3; ++a) {
                lo[3 * (size_t)q + a] = std::min(lo[3 * (size_t)q + a], xyz[a]);
                hi[3 * (size_t)q + a] = std::max(hi[3 * (size_t)q + a], xyz[a]);
            }
            ++listed[(size_t)q];
        }
    }
    for (int q = 0; q < gi.n_sub; ++q) {
        if (!listed[(size_t)q]) continue;
        long long vol = 1;
        for (int a = 0; a < 3; ++a) vol *= hi[3 * (size_t)q + a] - lo[3 * (size_t)q + a] + 1;
        if (vol != listed[(size_t)q]) ++out[5];
    }
    // quads as the trace packs them (A = v1, E1 = v0 - v1, E2 = v2 - v1) and their heliostats
    std::vector<float> verts(9 * (size_t)gi.n_sub);
    if (orc_subheliostat_vertices(s, g, verts.data()) != 3 * gi.n_sub) return 3;
    std::vector<Quad> quads((size_t)gi.n_sub);
    std::vector<int> helios;
    std::vector<qmc_heliostat_info> hinfo;
    for (int h = 0;; ++h) {
        qmc_heliostat_info hi_{};
        if (orc_heliostat_info(s, h, &hi_)) break;
        hinfo.push_back(hi_);
        if (hi_.grid != g) continue;
        helios.push_back(h);
        for (int k = 0; k < hi_.sub_count; ++k) quads[(size_t)(hi_.sub_begin + k)].helio = h;
    }
    for (int q = 0; q < gi.n_sub; ++q) {
        const float *v = &verts[9 * (size_t)q];
        const V3 v0{v[0], v[1], v[2]}, v1{v[3], v[4], v[5]}, v2{v[6], v[7], v[8]};
        quads[(size_t)q].A = v1;
        quads[(size_t)q].E1 = v0 - v1;
        quads[(size_t)q].E2 = v2 - v1;
    }
    if (helios.empty()) return 4;
    const double inv[3] = {1.0 / (double)interval[0], 1.0 / (double)interval[1], 1.0 / (double)interval[2]};
    const ProofFrame frame = proof_frame(pos, interval, n);
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    // rays start on up to 256 of the grid's heliostats (their micro-facet centres are kept)
    std::shuffle(helios.begin(), helios.end(), rng);
    if (helios.size() > 256) helios.resize(256);
    std::vector<std::vector<float>> facet_o(hinfo.size());
    std::vector<std::vector<int>> near(hinfo.size());  // quads within 60 m of each heliostat
    std::vector<float> fn;
    for (int it = 0; it < n_rays; ++it) {
        const int h = helios[(size_t)(rng() % helios.size())];
        const qmc_heliostat_info &H = hinfo[(size_t)h];
        // origin: a micro-facet centre, or a random point of the heliostat's quads lifted off them
        V3 o;
        if (it % 2 == 0 && H.n_micro > 0) {
            std::vector<float> &fo = facet_o[(size_t)h];
            if (fo.empty()) {
                fo.resize(3 * (size_t)H.n_micro);
                fn.resize(3 * (size_t)H.n_micro);
                if (orc_facets(s, h, fo.data(), fn.data()) != H.n_micro) return 5;
            }
            const size_t m = rng() % (size_t)H.n_micro;
            o = V3{fo[3 * m], fo[3 * m + 1], fo[3 * m + 2]};
        } else {
            const Quad &Q = quads[(size_t)(H.sub_begin + (int)(rng() % (uint64_t)H.sub_count))];
            const V3 nrm = normalize(cross(Q.E1, Q.E2));
            o = Q.A + Q.E1 * (float)U(rng) + Q.E2 * (float)U(rng) + nrm * (float)(0.05 * U(rng));
        }
        // target: a random point of a quad of a heliostat near this one
        std::vector<int> &nb = near[(size_t)h];
        if (nb.empty()) {
            const V3 hp{H.pos[0], H.pos[1], H.pos[2]};
            for (int q = 0; q < gi.n_sub; ++q) {
                const V3 c = quads[(size_t)q].A + (quads[(size_t)q].E1 + quads[(size_t)q].E2) * 0.5f;
                if (quads[(size_t)q].helio != h && length(c - hp) < 60.0f) nb.push_back(q);
            }
            if (nb.empty()) nb.push_back(-1);
        }
        const int tq = nb[(size_t)(rng() % nb.size())];
        if (tq < 0) continue;
        const Quad &T = quads[(size_t)tq];
        double p[3];
        {
            const V3 tp = T.A + T.E1 * (float)U(rng) + T.E2 * (float)U(rng);
            p[0] = tp.x, p[1] = tp.y, p[2] = tp.z;
        }
        const int mode = it % 5;
        if (mode == 1) {
            // a point of the quad next to a cell edge: two coordinates on cell faces (off by 0 .. 1 cm),
            // the third on the quad's plane, kept only if it lies on the quad — where the traversal
            // steps two axes at once and may pass one of the edge's cells by
            const V3 nq = cross(T.E1, T.E2);
            const double nn[3] = {nq.x, nq.y, nq.z}, A0[3] = {T.A.x, T.A.y, T.A.z};
            const int a = (int)(rng() % 3), b = (a + 1 + (int)(rng() % 2)) % 3, c = 3 - a - b;
            if (std::fabs(nn[c]) > 1e-3 * std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2])) {
                double q[3];
                for (int k : {a, b}) {
                    const double f = pos[k] + std::round((p[k] - pos[k]) * inv[k]) * interval[k];
                    q[k] = f + (U(rng) < 0.2 ? 0.0 : (U(rng) < 0.5 ? -1.0 : 1.0) * std::pow(10.0, -2.0 - 7.0 * U(rng)));
                }
                q[c] = A0[c] - (nn[a] * (q[a] - A0[a]) + nn[b] * (q[b] - A0[b])) / nn[c];
                // on the quad: (q - A) = s E1 + r E2 with s, r in [0, 1] (least squares on the plane)
                const double e1[3] = {T.E1.x, T.E1.y, T.E1.z}, e2[3] = {T.E2.x, T.E2.y, T.E2.z};
                double w[3], g11 = 0, g12 = 0, g22 = 0, b1 = 0, b2 = 0;
                for (int k = 0; k < 3; ++k) {
                    w[k] = q[k] - A0[k];
                    g11 += e1[k] * e1[k], g12 += e1[k] * e2[k], g22 += e2[k] * e2[k];
                    b1 += w[k] * e1[k], b2 += w[k] * e2[k];
                }
                const double det = g11 * g22 - g12 * g12;
                const double ss = (b1 * g22 - b2 * g12) / det, rr = (b2 * g11 - b1 * g12) / det;
                if (ss > 0.0 && ss < 1.0 && rr > 0.0 && rr < 1.0) p[0] = q[0], p[1] = q[1], p[2] = q[2];
            }
        }
        if (mode >= 2) {
            // push one or two coordinates onto a cell face (an edge for two), off by 0 .. 1 cm
            const int na = mode == 4 ? 2 : 1;
            const int a0 = (int)(rng() % 3);
            for (int k = 0; k < na; ++k) {
                const int a = (a0 + k) % 3;
                const double f = pos[a] + std::round((p[a] - pos[a]) * inv[a]) * interval[a];
                const double off = (U(rng) < 0.5 ? -1.0 : 1.0) * std::pow(10.0, -2.0 - 6.0 * U(rng));
                p[a] = f + (U(rng) < 0.2 ? 0.0 : off);
            }
        }
        V3 d = normalize(V3{(float)(p[0] - o.x), (float)(p[1] - o.y), (float)(p[2] - o.z)});
        if (it % 7 == 3) {
            // a direction component next to the walk's never-step threshold
            const int a = (int)(rng() % 3);
            const float c = (float)(kEps * (0.5 + U(rng)));
            if (a == 0) d.x = d.x < 0 ? -c : c;
            if (a == 1) d.y = d.y < 0 ? -c : c;
            if (a == 2) d.z = d.z < 0 ? -c : c;
            d = normalize(d);
        }
        ++out[0];
        bool any = false, proven = false;
        std::vector<int32_t> visited;
        const float of[3] = {o.x, o.y, o.z}, df[3] = {d.x, d.y, d.z};
        for (int q = 0; q < gi.n_sub; ++q) {
            const Quad &Q = quads[(size_t)q];
            if (Q.helio == h) continue;
            const V3 c = Q.A + (Q.E1 + Q.E2) * 0.5f;
            if (length(c - o) > 80.0f) continue;
            float t, u, v;
            if (!ray_quad(o, d, Q.A, Q.E1, Q.E2, t, u, v)) continue;
            any = true;
            // the traversal visits a cell of [bl, bh] (inclusive)
            auto visits = [&](const int *bl, const int *bh) {
                if (visited.empty()) {
                    visited.resize(4096);
                    const int nv = orc_dda_cells(s, g, of, df, visited.data(), (int)visited.size());
                    visited.resize((size_t)std::min(nv, 4096));
                }
                for (const int32_t c : visited) {
                    const int xyz[3] = {c / (n[1] * n[2]), (c / n[2]) % n[1], c % n[2]};
                    if (xyz[0] >= bl[0] && xyz[0] <= bh[0] && xyz[1] >= bl[1] && xyz[1] <= bh[1] &&
                        xyz[2] >= bl[2] && xyz[2] <= bh[2])
                        return true;
                }
                return false;
            };
            // the one cell holding the hit point (the origin's index on an axis that is never stepped)
            const double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z};
            int ci[3];
            for (int a = 0; a < 3; ++a) {
                const bool still = dd[a] <= kEps && dd[a] >= -kEps;
                const double r = oo[a] - pos[a] + (still ? 0.0 : (double)t * dd[a]);
                ci[a] = (int)std::floor(r * inv[a]);
            }
            const ProofBox cell = proof_box(interval, n, ci, ci);
            if (dda_visit_proof(o, d, t, frame, cell.mid, cell.lim)) {
                ++out[7];
                if (!visits(ci, ci)) ++out[6];
            }
            if (!listed[(size_t)q]) continue;
            const ProofBox box = proof_box(interval, n, &lo[3 * (size_t)q], &hi[3 * (size_t)q]);
            if (!dda_visit_proof(o, d, t, frame, box.mid, box.lim)) continue;
            proven = true;
            if (!visits(&lo[3 * (size_t)q], &hi[3 * (size_t)q])) ++out[6];
        }
        if (!any) continue;
        ++out[1];
        int hit = -1;
        const bool dda = orc_dda(s, g, of, df, H.sub_begin, H.sub_begin + H.sub_count, &hit) != 0;
        if (proven) ++out[2];
        if (dda) ++out[3];
        if (proven && !dda) ++out[4];
    }
    return 0;
}
