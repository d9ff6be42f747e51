// Brute-force check of the traversal-culling bounds (cone_may_hit and its bounding-sphere pre-test
// sphere_may_hit in qmc_math.h, each on half of the cases): whenever one says
// that no ray from a source region inside a cone can meet a quad, rays sampled from that region
// (corners, edges, faces, interior) in that cone (boundary and interior directions), traced with the
// float ray/parallelogram test the kernels use, must all miss. Source regions are flat rectangles
// (flat heliostats) and boxes (focus-flat heliostats, whose facets sag off the heliostat plane);
// quads are placed near the cone so that a good share of cases is culled.
// Prints "cases culled rays bad" and exits non-zero on any miss of the bound.
#include <cmath>
#include <cstdio>
#include <random>

#include "qmc_math.h"

using namespace qmc;

static CD3 unit(CD3 a) { return cmul(a, 1.0 / std::sqrt(cdot(a, a))); }

int main(int argc, char **argv) {
    const int n_cases = argc > 1 ? std::atoi(argv[1]) : 60000;
    std::mt19937_64 g(11);
    std::uniform_real_distribution<double> U(-1, 1);
    long long culled = 0, bad = 0, rays = 0;
    for (int it = 0; it < n_cases; ++it) {
        const CD3 a = unit(CD3{U(g), U(g) * 0.3 + 0.7, U(g)});
        const CD3 nrm = unit(CD3{U(g), 1.0, U(g)});
        const CD3 u = unit(ccross(CD3{0, 1, 0}, nrm)), v = ccross(u, nrm);
        const double sx = 1 + 6 * std::fabs(U(g)), sz = 1 + 5 * std::fabs(U(g));
        const double sy = (it % 2) ? 0.0 : 0.3 * std::fabs(U(g));  // every other case a box
        const CD3 c{U(g) * 5, 5 + U(g), U(g) * 5};
        const CD3 hu = cmul(u, sx / 2), hv = cmul(v, sz / 2), hn = cmul(nrm, sy / 2);
        const double dist = 2 + 40 * std::fabs(U(g));
        const CD3 side = cmul(ccross(a, CD3{U(g), U(g), U(g)}), 8 * U(g));
        const CD3 A = cadd(cadd(c, cmul(a, dist * U(g))), side);
        const CD3 E1{U(g) * 8, U(g) * 3, U(g) * 8}, E2{U(g) * 3, U(g) * 6, U(g) * 3};
        const double theta = (it % 3 == 0) ? 0.005 : (it % 3 == 1 ? 0.02 : 0.2);
        // the pre-test (bounding spheres) and the full test must each be conservative
        const CD3 sum = cadd(E1, E2), diff = csub(E1, E2);
        const CD3 D = csub(cadd(A, cmul(sum, 0.5)), c);
        const double R = std::sqrt(cdot(hu, hu) + cdot(hn, hn) + cdot(hv, hv)) +
                         0.5 * std::sqrt(std::fmax(cdot(sum, sum), cdot(diff, diff)));
        const bool use_sphere = it % 4 >= 2;
        if (use_sphere ? sphere_may_hit(D, R, a, std::tan(theta), 1e-3)
                       : cone_may_hit(c, hu, hn, hv, a, std::tan(theta), A, E1, E2, 1e-3))
            continue;
        ++culled;
        const CD3 e1 = unit(ccross(a, CD3{0.3, 0.1, 0.9})), e2 = ccross(a, e1);
        for (int k = 0; k < 400; ++k) {
            // source point: corners and edges (coordinates at +-1) as well as interior points
            const double s = (k % 4 < 2) ? (k % 2 ? 1.0 : -1.0) : U(g);
            const double r = (k % 8 < 4) ? U(g) : ((k / 2) % 2 ? 1.0 : -1.0);
            const double q = (k % 3 == 0) ? ((k / 3) % 2 ? 1.0 : -1.0) : U(g);
            const CD3 p = cadd(c, cadd(cadd(cmul(hu, s), cmul(hv, r)), cmul(hn, q)));
            // direction: on the cone's boundary four times in five, inside it otherwise
            const double ang = theta * ((k % 5) ? 1.0 : std::fabs(U(g)));
            const double golden = k * 0.6180339887;
            const double psi = 6.283185307179586 * (golden - std::floor(golden));
            const CD3 d = cadd(cmul(a, std::cos(ang)), cmul(cadd(cmul(e1, std::cos(psi)), cmul(e2, std::sin(psi))),
                                                             std::sin(ang)));
            const V3 o{(float)p.x, (float)p.y, (float)p.z}, df{(float)d.x, (float)d.y, (float)d.z};
            float t, uu, vv;
            ++rays;
            if (ray_quad(o, df, V3{(float)A.x, (float)A.y, (float)A.z}, V3{(float)E1.x, (float)E1.y, (float)E1.z},
                         V3{(float)E2.x, (float)E2.y, (float)E2.z}, t, uu, vv)) {
                if (++bad < 5) std::printf("culled but hit: case %d theta %g dist %g\n", it, theta, dist);
            }
        }
    }
    std::printf("%d %lld %lld %lld\n", n_cases, culled, rays, bad);
    return bad != 0;
}
