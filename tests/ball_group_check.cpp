// Check of the grouped culling scan's premise (cull_scan in qmc_cull.hip): sphere_may_hit is
// monotone in its ball, so a group ball that fails the test proves that every ball inside it fails
// too. Random cones (axis, half-angle, margin) and random groups of 1-28 balls, each group ball the
// tightest one about the group's box centre holding them (as the upload builds it) or a looser one;
// whenever the group test fails, every member's test must fail. Prints "groups rejected members bad"
// and exits non-zero on any violation.
#include <cmath>
#include <cstdio>
#include <random>

#include "qmc_math.h"

using namespace qmc;

int main(int argc, char **argv) {
    const int n = argc > 1 ? std::atoi(argv[1]) : 200000;
    std::mt19937_64 g(5);
    std::uniform_real_distribution<double> U(-1, 1);
    long rejected = 0, members = 0, bad = 0;
    for (int it = 0; it < n; ++it) {
        CD3 a{U(g), 0.3 + std::fabs(U(g)), U(g)};
        a = cmul(a, 1.0 / std::sqrt(cdot(a, a)));
        const double tan_t = std::tan(0.002 + 0.3 * std::fabs(U(g)));
        const double margin = (it % 2) ? 0.01 : 0.0;
        const double rs = 6.0 * std::fabs(U(g));  // source radius
        const CD3 base{U(g) * 80, U(g) * 20, U(g) * 80};
        const int k = 1 + (int)(28 * std::fabs(U(g)));
        CD3 cq[28];
        double rq[28];
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int i = 0; i < k; ++i) {
            cq[i] = cadd(base, CD3{U(g) * 6, U(g) * 3, U(g) * 6});
            rq[i] = 0.2 + 2.0 * std::fabs(U(g));
            const double c[3] = {cq[i].x, cq[i].y, cq[i].z};
            for (int d = 0; d < 3; ++d) {
                lo[d] = std::fmin(lo[d], c[d] - rq[i]);
                hi[d] = std::fmax(hi[d], c[d] + rq[i]);
            }
        }
        const CD3 cg{0.5 * (lo[0] + hi[0]), 0.5 * (lo[1] + hi[1]), 0.5 * (lo[2] + hi[2])};
        double rg = 0.0;
        for (int i = 0; i < k; ++i) rg = std::fmax(rg, std::sqrt(cdot(csub(cq[i], cg), csub(cq[i], cg))) + rq[i]);
        if (it % 3 == 0) rg *= 1.0 + std::fabs(U(g));  // a looser ball
        if (sphere_may_hit(cg, rs + rg, a, tan_t, margin)) continue;
        ++rejected;
        for (int i = 0; i < k; ++i) {
            ++members;
            if (sphere_may_hit(cq[i], rs + rq[i], a, tan_t, margin)) ++bad;
        }
    }
    std::printf("%d %ld %ld %ld\n", n, rejected, members, bad);
    return bad ? 1 : 0;
}
