// Host-side maintenance of the scaffold's air mesh between Newton steps.
//
// The reference builds a new air mesh with Triangle after every Newton step (Optimizer.cpp:251-257,
// Scaffold.cpp:27-176). Here the air mesh is kept: its vertices move with the steps (the step bound keeps its
// triangles positively oriented), its rest shape is reset to the current UV after each step, and when its
// triangles degrade it is re-meshed in place by Lawson edge flips: every air edge shared by two air triangles
// whose opposite vertex lies inside the other triangle's circumcircle is flipped, provided both new triangles are
// positively oriented, until no such edge is left. The result is the Delaunay triangulation of the current air
// vertices constrained by the air mesh's boundary edges (the map's boundary and the outer box, each in one air
// triangle, are never flipped); among the triangulations of the same points it maximises the smallest angle
// (Lawson 1977; de Berg et al., Computational Geometry, ch. 9). Vertex set and boundary stay, so only the air
// block of the Newton system's pattern changes.
#pragma once

#include <cmath>
#include <cstdint>
#include <unordered_map>
#include <vector>

namespace oc {

// Smallest interior angle (radians) of triangles F (n x 3 ids) at positions P (x, y per vertex).
inline double air_min_angle(const std::vector<double>& P, const std::vector<int>& F)
{
    double best = M_PI;
    const size_t nF = F.size() / 3;
    for (size_t t = 0; t < nF; ++t)
        for (int k = 0; k < 3; ++k) {
            const int a = F[3 * t + k], b = F[3 * t + (k + 1) % 3], c = F[3 * t + (k + 2) % 3];
            const double ux = P[2 * b] - P[2 * a], uy = P[2 * b + 1] - P[2 * a + 1];
            const double vx = P[2 * c] - P[2 * a], vy = P[2 * c + 1] - P[2 * a + 1];
            const double ang = std::atan2(std::fabs(ux * vy - uy * vx), ux * vx + uy * vy);
            if (ang < best) best = ang;
        }
    return best;
}

namespace detail {
inline long double orient2(const double* a, const double* b, const double* c)
{
    return ((long double)b[0] - a[0]) * ((long double)c[1] - a[1]) -
           ((long double)b[1] - a[1]) * ((long double)c[0] - a[0]);
}
// > 0 when d lies inside the circumcircle of the counter-clockwise triangle (a, b, c)
inline long double incircle(const double* a, const double* b, const double* c, const double* d)
{
    const long double ax = (long double)a[0] - d[0], ay = (long double)a[1] - d[1];
    const long double bx = (long double)b[0] - d[0], by = (long double)b[1] - d[1];
    const long double cx = (long double)c[0] - d[0], cy = (long double)c[1] - d[1];
    return (ax * ax + ay * ay) * (bx * cy - cx * by) - (bx * bx + by * by) * (ax * cy - cx * ay) +
           (cx * cx + cy * cy) * (ax * by - bx * ay);
}
}  // namespace detail

// Lawson flips on F (counter-clockwise triangles over positions P); returns the number of flips (at most
// maxFlips). Edges in one triangle only are the constraints.
inline long lawson_flips(const std::vector<double>& P, std::vector<int>& F, long maxFlips)
{
    const int nF = (int)(F.size() / 3);
    auto key = [](int u, int v) -> uint64_t {
        const uint32_t lo = (uint32_t)(u < v ? u : v), hi = (uint32_t)(u < v ? v : u);
        return ((uint64_t)lo << 32) | hi;
    };
    // per undirected edge: the (up to two) half-edges t*3+k (edge from corner k to corner k+1 of t)
    struct Slots {
        int h[2] = {-1, -1};
    };
    std::unordered_map<uint64_t, Slots> edges;
    edges.reserve((size_t)nF * 2);
    for (int t = 0; t < nF; ++t)
        for (int k = 0; k < 3; ++k) {
            Slots& s = edges[key(F[3 * t + k], F[3 * t + (k + 1) % 3])];
            (s.h[0] < 0 ? s.h[0] : s.h[1]) = 3 * t + k;
        }
    std::vector<uint64_t> stack;
    stack.reserve(edges.size());
    for (const auto& e : edges)
        if (e.second.h[1] >= 0) stack.push_back(e.first);
    // scale of the tolerances: the largest coordinate extent
    double lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
    for (size_t i = 0; i < P.size(); ++i) {
        lo[i & 1] = std::fmin(lo[i & 1], P[i]);
        hi[i & 1] = std::fmax(hi[i & 1], P[i]);
    }
    const long double L = std::fmax(hi[0] - lo[0], hi[1] - lo[1]);
    const long double tolIn = 1e-14L * L * L * L * L, tolOr = 1e-15L * L * L;
    auto replace = [&](uint64_t k, int oldH, int newH) {
        Slots& s = edges[k];
        if (s.h[0] == oldH) s.h[0] = newH;
        else if (s.h[1] == oldH) s.h[1] = newH;
    };
    long flips = 0;
    while (!stack.empty() && flips < maxFlips) {
        const uint64_t k = stack.back();
        stack.pop_back();
        auto it = edges.find(k);
        if (it == edges.end() || it->second.h[1] < 0) continue;
        const int h1 = it->second.h[0], h2 = it->second.h[1];
        const int t1 = h1 / 3, k1 = h1 % 3, t2 = h2 / 3, k2 = h2 % 3;
        const int a = F[3 * t1 + k1], b = F[3 * t1 + (k1 + 1) % 3], c = F[3 * t1 + (k1 + 2) % 3];
        const int d = F[3 * t2 + (k2 + 2) % 3];
        if (F[3 * t2 + k2] != b || F[3 * t2 + (k2 + 1) % 3] != a) continue;  // not a consistent pair
        const double *pa = &P[2 * a], *pb = &P[2 * b], *pc = &P[2 * c], *pd = &P[2 * d];
        if (!(detail::incircle(pa, pb, pc, pd) > tolIn)) continue;
        if (!(detail::orient2(pc, pa, pd) > tolOr) || !(detail::orient2(pc, pd, pb) > tolOr)) continue;
        if (edges.count(key(c, d))) continue;  // c-d already an edge elsewhere
        // t1 <- (c, a, d), t2 <- (c, d, b)
        const int o_ad = 3 * t2 + (k2 + 1) % 3, o_db = 3 * t2 + (k2 + 2) % 3;  // half-edges a->d, d->b in t2
        const int o_bc = 3 * t1 + (k1 + 1) % 3, o_ca = 3 * t1 + (k1 + 2) % 3;  // b->c, c->a in t1
        F[3 * t1] = c;
        F[3 * t1 + 1] = a;
        F[3 * t1 + 2] = d;
        F[3 * t2] = c;
        F[3 * t2 + 1] = d;
        F[3 * t2 + 2] = b;
        edges.erase(k);
        replace(key(c, a), o_ca, 3 * t1 + 0);
        replace(key(a, d), o_ad, 3 * t1 + 1);
        replace(key(d, b), o_db, 3 * t2 + 1);
        replace(key(b, c), o_bc, 3 * t2 + 2);
        Slots s;
        s.h[0] = 3 * t1 + 2;  // d->c in t1
        s.h[1] = 3 * t2 + 0;  // c->d in t2
        edges[key(c, d)] = s;
        stack.push_back(key(c, a));
        stack.push_back(key(a, d));
        stack.push_back(key(d, b));
        stack.push_back(key(b, c));
        ++flips;
    }
    return flips;
}

}  // namespace oc
