/*
 * jsx_delaunay.h -- host-side Delaunay triangulation of scattered (z, r)
 * simulation nodes and the bucket index the GPU locates query points with.
 * Included by jsx_host.hip (host code only).
 *
 * create_triangulation (src/SimulationInterpolater.cpp:105-117) hands the
 * nodes to CGAL's Delaunay_triangulation_2; SimulationInterpolater::
 * interpolated_value (:33-102) locates the face that holds a query point and
 * weights its three node values with Triangle_coordinates_2.  Here the
 * triangulation is built once on the host (Bowyer-Watson: each node
 * inserted into the current triangulation by removing the triangles whose
 * circumcircle holds it and fanning the cavity from the node) and shipped to
 * HBM with a uniform bucket index over its bounding box, so that a lane finds
 * its triangle by scanning one short list instead of walking the mesh.
 *
 * Robustness: the in-circle test runs in long double on coordinates taken
 * relative to the query node; every cavity is made star-shaped from the new
 * node (a cavity triangle whose boundary edge the node does not see strictly
 * on its left is returned to the mesh), so rounding can only cost the
 * Delaunay property on cocircular nodes, never the validity of the mesh.
 * Nodes repeated at the same (z, r) keep their first occurrence.
 */
#ifndef JSX_DELAUNAY_H
#define JSX_DELAUNAY_H

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <unordered_map>
#include <vector>

namespace jsx_del {

struct Tri {
    int v[3];   /* counter-clockwise in (z, r); a ghost has kInf as v[2] */
    int n[3];   /* n[i]: the triangle across the edge opposite v[i] */
    bool alive;
};

/* The vertex at infinity: a ghost triangle (a, b, kInf) lies across the hull
 * edge a -> b (the hull is clockwise seen from outside, so the mesh is on the
 * right of a -> b); its "circumcircle" is the open half-plane left of a -> b
 * together with the open segment (a, b).  With ghosts the mesh always covers
 * the plane and the triangulation of the nodes is exact up to its convex hull
 * (a large enclosing triangle instead leaves slivers along nearly straight
 * stretches of the hull uncovered). */
constexpr int kInf = -2;

struct Mesh {
    std::vector<double> x, y;
    std::vector<Tri> t;
    std::vector<int> free_slots;
    int last = 0;               /* walk start */
    std::vector<char> incav;    /* cavity membership per triangle (0 outside) */

    double orient(int a, int b, double px, double py) const {
        return (x[b] - x[a]) * (py - y[a]) - (y[b] - y[a]) * (px - x[a]);
    }
    static bool ghost(const Tri &T) { return T.v[2] == kInf; }
    /* true: (px, py) in conflict with T (strictly inside its circumcircle) */
    bool conflict(const Tri &T, double px, double py) const {
        if (ghost(T)) {
            const int a = T.v[0], b = T.v[1];
            const double o = orient(a, b, px, py);
            if (o > 0.0) return true;
            if (o < 0.0) return false;
            /* on the hull edge's line: in conflict inside the open segment */
            const double dx = x[b] - x[a], dy = y[b] - y[a];
            const double s = (px - x[a]) * dx + (py - y[a]) * dy;
            return s > 0.0 && s < dx * dx + dy * dy;
        }
        const long double adx = (long double)x[T.v[0]] - px, ady = (long double)y[T.v[0]] - py;
        const long double bdx = (long double)x[T.v[1]] - px, bdy = (long double)y[T.v[1]] - py;
        const long double cdx = (long double)x[T.v[2]] - px, cdy = (long double)y[T.v[2]] - py;
        return (adx * adx + ady * ady) * (bdx * cdy - cdx * bdy) + (bdx * bdx + bdy * bdy) * (cdx * ady - adx * cdy) +
                   (cdx * cdx + cdy * cdy) * (adx * bdy - bdx * ady) > 0.0L;
    }
    int add(const Tri &T) {
        if (!free_slots.empty()) {
            int k = free_slots.back();
            free_slots.pop_back();
            t[k] = T;
            return k;
        }
        t.push_back(T);
        incav.push_back(0);
        return (int)t.size() - 1;
    }
    /* a triangle holding (px, py): a real one that holds it on or inside its
       edges, or a ghost across whose hull edge it lies */
    int locate(double px, double py) const {
        int cur = last;
        if (cur < 0 || cur >= (int)t.size() || !t[cur].alive) {
            cur = -1;
            for (int k = 0; k < (int)t.size(); ++k)
                if (t[k].alive) { cur = k; break; }
        }
        const int cap = 4 * (int)t.size() + 64;
        for (int step = 0; step < cap; ++step) {
            const Tri &T = t[cur];
            if (ghost(T)) {
                if (orient(T.v[0], T.v[1], px, py) > 0.0) return cur;
                cur = T.n[2];
                continue;
            }
            int next = -1;
            for (int i = 0; i < 3; ++i) {
                const int a = T.v[(i + 1) % 3], b = T.v[(i + 2) % 3];
                if (orient(a, b, px, py) < 0.0) { next = T.n[i]; break; }
            }
            if (next < 0) return cur;
            cur = next;
        }
        /* the walk cycled (rounding): search every triangle */
        for (int k = 0; k < (int)t.size(); ++k) {
            if (!t[k].alive || ghost(t[k])) continue;
            const Tri &T = t[k];
            if (orient(T.v[0], T.v[1], px, py) >= 0.0 && orient(T.v[1], T.v[2], px, py) >= 0.0 &&
                orient(T.v[2], T.v[0], px, py) >= 0.0)
                return k;
        }
        for (int k = 0; k < (int)t.size(); ++k)
            if (t[k].alive && ghost(t[k]) && orient(t[k].v[0], t[k].v[1], px, py) > 0.0) return k;
        return cur;
    }
    /* may the cavity's boundary edge (a, b) be fanned from p */
    bool sees(int a, int b, double px, double py) const {
        if (a == kInf || b == kInf) return true;
        return orient(a, b, px, py) > 0.0;
    }
    /* insert node p; false if the cavity could not be formed (the node is
       then left out: it coincides with the mesh to rounding) */
    bool insert(int p) {
        const double px = x[p], py = y[p];
        const int t0 = locate(px, py);
        if (incav.size() < t.size()) incav.resize(t.size(), 0);
        /* the triangles in conflict with p, connected to t0 */
        std::vector<int> cav{t0}, stack{t0};
        incav[t0] = 1;
        while (!stack.empty()) {
            const int k = stack.back();
            stack.pop_back();
            for (int i = 0; i < 3; ++i) {
                const int nb = t[k].n[i];
                if (nb < 0 || incav[nb]) continue;
                if (conflict(t[nb], px, py)) {
                    incav[nb] = 1;
                    cav.push_back(nb);
                    stack.push_back(nb);
                }
            }
        }
        /* star-shaped from p: every boundary edge must see p strictly on its
           left; a triangle with an edge that does not leaves the cavity, and
           what stays connected to t0 is kept */
        for (;;) {
            int bad = -1;
            for (int k : cav) {
                for (int i = 0; i < 3 && bad < 0; ++i) {
                    const int nb = t[k].n[i];
                    if (nb >= 0 && incav[nb]) continue;
                    if (!sees(t[k].v[(i + 1) % 3], t[k].v[(i + 2) % 3], px, py)) bad = k;
                }
                if (bad >= 0) break;
            }
            if (bad < 0) break;
            if (bad == t0) {
                for (int k : cav) incav[k] = 0;
                return false;
            }
            incav[bad] = 0;
            std::vector<int> keep{t0};
            for (int k : cav) if (k != t0) incav[k] = 2;   /* 2: candidate */
            incav[t0] = 1;
            for (size_t c = 0; c < keep.size(); ++c)
                for (int i = 0; i < 3; ++i) {
                    const int nb = t[keep[c]].n[i];
                    if (nb >= 0 && incav[nb] == 2) {
                        incav[nb] = 1;
                        keep.push_back(nb);
                    }
                }
            for (int k : cav) if (incav[k] == 2) incav[k] = 0;
            cav.swap(keep);
        }
        /* fan the cavity boundary from p */
        struct Edge { int a, b, out, from; };
        std::vector<Edge> edges;
        for (int k : cav)
            for (int i = 0; i < 3; ++i) {
                const int nb = t[k].n[i];
                if (nb >= 0 && incav[nb]) continue;
                edges.push_back({t[k].v[(i + 1) % 3], t[k].v[(i + 2) % 3], nb, k});
            }
        for (int k : cav) {
            t[k].alive = false;
            incav[k] = 0;
            free_slots.push_back(k);
        }
        std::unordered_map<int, int> by_start, by_end;
        std::vector<int> made(edges.size());
        for (size_t e = 0; e < edges.size(); ++e) {
            Tri T;
            /* (a, b, p); an edge with the vertex at infinity gives a ghost,
               rotated so that kInf comes last */
            int v0 = edges[e].a, v1 = edges[e].b, v2 = p;
            int o2 = edges[e].out;
            T.alive = true;
            if (v0 == kInf) {          /* (inf, b, p) -> (b, p, inf) */
                T.v[0] = v1; T.v[1] = v2; T.v[2] = kInf;
                T.n[0] = -1; T.n[1] = o2; T.n[2] = -1;
            } else if (v1 == kInf) {   /* (a, inf, p) -> (p, a, inf) */
                T.v[0] = v2; T.v[1] = v0; T.v[2] = kInf;
                T.n[0] = o2; T.n[1] = -1; T.n[2] = -1;
            } else {
                T.v[0] = v0; T.v[1] = v1; T.v[2] = v2;
                T.n[0] = -1; T.n[1] = -1; T.n[2] = o2;
            }
            const int k = add(T);
            made[e] = k;
            if (o2 >= 0) {
                Tri &O = t[o2];
                for (int i = 0; i < 3; ++i)
                    if (O.n[i] == edges[e].from && O.v[(i + 1) % 3] == edges[e].b && O.v[(i + 2) % 3] == edges[e].a)
                        O.n[i] = k;
            }
            by_start[edges[e].a] = k;
            by_end[edges[e].b] = k;
        }
        /* the fan's inner edges: (b, p) of the triangle from edge (a, b) is
           (p, b) of the one from the edge starting at b; (p, a) is (a, p) of
           the one from the edge ending at a */
        for (size_t e = 0; e < edges.size(); ++e) {
            Tri &T = t[made[e]];
            auto s = by_start.find(edges[e].b);
            auto f = by_end.find(edges[e].a);
            if (s == by_start.end() || f == by_end.end()) return false;
            for (int i = 0; i < 3; ++i) {
                const int a = T.v[(i + 1) % 3], b = T.v[(i + 2) % 3];
                if (a == edges[e].b && b == p) T.n[i] = s->second;
                else if (a == p && b == edges[e].a) T.n[i] = f->second;
            }
        }
        last = made.empty() ? last : made[0];
        return true;
    }
};

/* 32-bit Morton code of (i, j) < 2^16 */
inline uint32_t morton(uint32_t i, uint32_t j) {
    auto spread = [](uint32_t v) {
        v &= 0xffffu;
        v = (v | (v << 8)) & 0x00ff00ffu;
        v = (v | (v << 4)) & 0x0f0f0f0fu;
        v = (v | (v << 2)) & 0x33333333u;
        v = (v | (v << 1)) & 0x55555555u;
        return v;
    };
    return spread(i) | (spread(j) << 1);
}

/* Delaunay triangles (counter-clockwise node indices into z, r) of the
 * distinct nodes; returns false when fewer than three nodes are not collinear */
inline bool triangulate(const std::vector<double> &z, const std::vector<double> &r, std::vector<int32_t> &tri) {
    const int n = (int)z.size();
    tri.clear();
    if (n < 3) return false;
    double zmin = z[0], zmax = z[0], rmin = r[0], rmax = r[0];
    for (int i = 1; i < n; ++i) {
        zmin = std::min(zmin, z[i]); zmax = std::max(zmax, z[i]);
        rmin = std::min(rmin, r[i]); rmax = std::max(rmax, r[i]);
    }
    const double span = std::max(zmax - zmin, rmax - rmin);
    if (!(span > 0.0)) return false;
    /* insertion order: Morton order of the nodes (short walks); a node equal
       to an earlier one in input order is skipped */
    std::vector<std::pair<uint64_t, int>> order(n);
    for (int i = 0; i < n; ++i) {
        const uint32_t qi = (uint32_t)std::min(65535.0, std::floor((z[i] - zmin) / span * 65535.0));
        const uint32_t qj = (uint32_t)std::min(65535.0, std::floor((r[i] - rmin) / span * 65535.0));
        order[i] = {((uint64_t)morton(qi, qj) << 32) | (uint32_t)i, i};
    }
    std::sort(order.begin(), order.end());
    Mesh M;
    M.x = z; M.y = r;
    std::vector<int> ins;
    ins.reserve(n);
    {
        struct PairHash {
            size_t operator()(const std::pair<double, double> &p) const {
                return std::hash<double>()(p.first) ^ (std::hash<double>()(p.second) * 1315423911u);
            }
        };
        std::unordered_map<std::pair<double, double>, int, PairHash> seen;
        seen.reserve((size_t)n * 2);
        for (int i = 0; i < n; ++i) seen.emplace(std::make_pair(z[i], r[i]), i);   /* first occurrence wins */
        for (auto &o : order)
            if (seen[std::make_pair(z[o.second], r[o.second])] == o.second) ins.push_back(o.second);
    }
    if (ins.size() < 3) return false;
    /* the first triangle: the first two nodes and the first node off their line */
    const int a0 = ins[0], b0 = ins[1];
    size_t c_at = 2;
    while (c_at < ins.size() && M.orient(a0, b0, z[ins[c_at]], r[ins[c_at]]) == 0.0) ++c_at;
    if (c_at == ins.size()) return false;   /* all nodes on one line */
    int A = a0, Bv = b0, Cv = ins[c_at];
    if (M.orient(A, Bv, z[Cv], r[Cv]) < 0.0) std::swap(A, Bv);
    /* real (A, B, C) = 0; ghosts across B->A, C->B, A->C: 1, 2, 3 */
    Tri T0{{A, Bv, Cv}, {2, 3, 1}, true};
    Tri G1{{Bv, A, kInf}, {-1, -1, 0}, true};
    Tri G2{{Cv, Bv, kInf}, {-1, -1, 0}, true};
    Tri G3{{A, Cv, kInf}, {-1, -1, 0}, true};
    /* ghost neighbours: across (A, inf) and (inf, B) etc. */
    G1.n[0] = 3; G1.n[1] = 2;   /* G1 = (B, A, inf): opp B is (A, inf) -> G3 = (A, C, inf) shares (inf, A); opp A is (inf, B) -> G2 */
    G2.n[0] = 1; G2.n[1] = 3;   /* G2 = (C, B, inf): opp C (B, inf) -> G1; opp B (inf, C) -> G3 */
    G3.n[0] = 2; G3.n[1] = 1;   /* G3 = (A, C, inf): opp A (C, inf) -> G2; opp C (inf, A) -> G1 */
    M.add(T0); M.add(G1); M.add(G2); M.add(G3);
    for (size_t k = 2; k < ins.size(); ++k) {
        if (k == c_at) continue;
        (void)M.insert(ins[k]);
    }
    for (const Tri &T : M.t) {
        if (!T.alive || Mesh::ghost(T)) continue;
        tri.push_back(T.v[0]); tri.push_back(T.v[1]); tri.push_back(T.v[2]);
    }
    /* a stable order: by the triangles' smallest node index, then the others */
    const size_t m = tri.size() / 3;
    std::vector<std::array<int32_t, 4>> key(m);
    for (size_t k = 0; k < m; ++k) {
        int32_t a = tri[3 * k], b = tri[3 * k + 1], c = tri[3 * k + 2];
        /* rotate so that the smallest index comes first (orientation kept) */
        while (a > b || a > c) { const int32_t s = a; a = b; b = c; c = s; }
        key[k] = {a, b, c, 0};
    }
    std::sort(key.begin(), key.end());
    for (size_t k = 0; k < m; ++k) { tri[3 * k] = key[k][0]; tri[3 * k + 1] = key[k][1]; tri[3 * k + 2] = key[k][2]; }
    return m > 0;
}

/* Uniform bucket index of a triangulation over its nodes' bounding box: each
 * bucket lists, in increasing order, every triangle whose bounding box --
 * widened by 1e-9 of its size, the slack of the containment test -- meets the
 * bucket.  About two triangles per bucket. */
struct Buckets {
    int nbz = 0, nbr = 0;
    double z0 = 0, dz = 1, r0 = 0, dr = 1, zlast = 0, rlast = 0;
    std::vector<int32_t> off, list;
};

inline void build_buckets(const std::vector<double> &z, const std::vector<double> &r, const std::vector<int32_t> &tri,
                          Buckets &B) {
    const size_t m = tri.size() / 3;
    double zmin = 1e300, zmax = -1e300, rmin = 1e300, rmax = -1e300;
    for (size_t k = 0; k < 3 * m; ++k) {
        zmin = std::min(zmin, z[tri[k]]); zmax = std::max(zmax, z[tri[k]]);
        rmin = std::min(rmin, r[tri[k]]); rmax = std::max(rmax, r[tri[k]]);
    }
    const double wz = std::max(zmax - zmin, 1e-300), wr = std::max(rmax - rmin, 1e-300);
    const double nb = std::max(1.0, (double)m / 2.0);
    double fz = std::sqrt(nb * wz / wr);
    fz = std::min(std::max(fz, 1.0), nb);
    B.nbz = std::max(1, std::min(1 << 15, (int)std::lround(fz)));
    B.nbr = std::max(1, std::min(1 << 15, (int)std::lround(nb / (double)B.nbz)));
    B.z0 = zmin; B.r0 = rmin;
    B.dz = wz / B.nbz; B.dr = wr / B.nbr;
    B.zlast = zmax; B.rlast = rmax;
    const size_t nbk = (size_t)B.nbz * (size_t)B.nbr;
    std::vector<int32_t> cnt(nbk + 1, 0);
    auto range = [&](size_t k, int &i0, int &i1, int &j0, int &j1) {
        double a = 1e300, b = -1e300, c = 1e300, d = -1e300;
        for (int q = 0; q < 3; ++q) {
            a = std::min(a, z[tri[3 * k + q]]); b = std::max(b, z[tri[3 * k + q]]);
            c = std::min(c, r[tri[3 * k + q]]); d = std::max(d, r[tri[3 * k + q]]);
        }
        const double mz = 1e-9 * (b - a) + 1e-12 * wz, mr = 1e-9 * (d - c) + 1e-12 * wr;
        i0 = (int)std::floor((a - mz - B.z0) / B.dz); i1 = (int)std::floor((b + mz - B.z0) / B.dz);
        j0 = (int)std::floor((c - mr - B.r0) / B.dr); j1 = (int)std::floor((d + mr - B.r0) / B.dr);
        i0 = std::max(i0, 0); j0 = std::max(j0, 0);
        i1 = std::min(i1, B.nbz - 1); j1 = std::min(j1, B.nbr - 1);
    };
    for (size_t k = 0; k < m; ++k) {
        int i0, i1, j0, j1;
        range(k, i0, i1, j0, j1);
        for (int i = i0; i <= i1; ++i)
            for (int j = j0; j <= j1; ++j) cnt[(size_t)i * B.nbr + j + 1]++;
    }
    for (size_t b = 0; b < nbk; ++b) cnt[b + 1] += cnt[b];
    B.off = cnt;
    B.list.assign((size_t)cnt[nbk], 0);
    std::vector<int32_t> fill(cnt.begin(), cnt.end() - 1);
    for (size_t k = 0; k < m; ++k) {
        int i0, i1, j0, j1;
        range(k, i0, i1, j0, j1);
        for (int i = i0; i <= i1; ++i)
            for (int j = j0; j <= j1; ++j) B.list[(size_t)fill[(size_t)i * B.nbr + j]++] = (int32_t)k;
    }
}

}  // namespace jsx_del

#endif /* JSX_DELAUNAY_H */
