// Geometry of the floating-random-walk capacitance walker (host code, no
// HIP): conductor and dielectric boxes in uniform grids, the largest
// conductor-free cube around a point within the limits of a query's
// structure description, that description, and the Gaussian surface around
// the master conductor.  Kept apart from the walk loop so that a CPU program
// can check every query against a scan over all boxes.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <string>
#include <utility>
#include <vector>

#include "query_limits.h"

namespace drw::walkgeo {

inline constexpr double kInf = std::numeric_limits<double>::infinity();

// L-infinity distance from p to an axis-aligned box (0 inside).
inline double box_dist(const double* b, const double* p)
{
    double d = 0;
    for (int a = 0; a < 3; a++) {
        const double lo = b[a] - p[a], hi = p[a] - b[a + 3];
        d = std::max(d, std::max(lo, hi));
    }
    return std::max(d, 0.0);
}

// Uniform grid over a set of boxes: each cell lists the boxes that overlap
// it.  A query around p visits the cells in rings of growing Chebyshev
// radius r around p's cell; after ring r every box not yet visited lies
// beyond outside_bound(p, c, r) in L-infinity distance (an axis-aligned box
// disjoint from the visited block of cells is separated from it along one
// axis).
class BoxGrid {
public:
    void build(const double* box, size_t n)
    {
        start_.assign(2, 0);
        item_.clear();
        for (int a = 0; a < 3; a++) { lo_[a] = 0; cs_[a] = 1; n_[a] = 1; }
        if (n == 0) return;
        double hi[3];
        for (int a = 0; a < 3; a++) { lo_[a] = kInf; hi[a] = -kInf; }
        for (size_t i = 0; i < n; i++)
            for (int a = 0; a < 3; a++) {
                lo_[a] = std::min(lo_[a], box[6 * i + a]);
                hi[a] = std::max(hi[a], box[6 * i + a + 3]);
            }
        double ext[3];
        for (int a = 0; a < 3; a++) ext[a] = std::max(hi[a] - lo_[a], 1e-9 * (1.0 + std::fabs(lo_[a])));
        // about four cells per box, cubic where the extents allow; coarser
        // while boxes spanning many cells would make the lists too long
        for (double target = std::min(4.0 * (double)n, 2.0e6);; target *= 0.25) {
            bool fixed[3] = {false, false, false};
            for (int pass = 0; pass < 3; pass++) {
                double vol = 1, cells = target;
                int k = 0;
                for (int a = 0; a < 3; a++)
                    if (fixed[a]) cells /= n_[a];
                    else { vol *= ext[a]; k++; }
                if (k == 0) break;
                const double s = std::pow(vol / std::max(cells, 1.0), 1.0 / k);
                for (int a = 0; a < 3; a++) {
                    if (fixed[a]) continue;
                    n_[a] = (int)std::min(1024.0, std::max(1.0, std::round(ext[a] / s)));
                    if (n_[a] == 1) fixed[a] = true;
                }
            }
            for (int a = 0; a < 3; a++) cs_[a] = ext[a] / n_[a];
            size_t entries = 0;
            for (size_t i = 0; i < n; i++) {
                size_t m = 1;
                for (int a = 0; a < 3; a++) m *= (size_t)(cell(a, box[6 * i + a + 3]) - cell(a, box[6 * i + a]) + 1);
                entries += m;
            }
            if (entries <= ((size_t)1 << 25) || target < 64) break;
        }
        const size_t ncell = (size_t)n_[0] * n_[1] * n_[2];
        start_.assign(ncell + 1, 0);
        for (int pass = 0; pass < 2; pass++) {
            for (size_t i = 0; i < n; i++) {
                int c0[3], c1[3];
                for (int a = 0; a < 3; a++) { c0[a] = cell(a, box[6 * i + a]); c1[a] = cell(a, box[6 * i + a + 3]); }
                for (int x = c0[0]; x <= c1[0]; x++)
                    for (int y = c0[1]; y <= c1[1]; y++)
                        for (int z = c0[2]; z <= c1[2]; z++) {
                            const size_t id = index(x, y, z);
                            if (pass == 0) start_[id + 1]++;
                            else item_[fill_[id]++] = (uint32_t)i;
                        }
            }
            if (pass == 0) {
                for (size_t c = 0; c < ncell; c++) start_[c + 1] += start_[c];
                item_.resize(start_[ncell]);
                fill_.assign(start_.begin(), start_.end() - 1);
            }
        }
        fill_.clear();
        fill_.shrink_to_fit();
    }

    void cell_of(const double* p, int* c) const
    {
        for (int a = 0; a < 3; a++) c[a] = cell(a, p[a]);
    }

    // true once the rings up to r cover the whole grid
    bool covers(const int* c, int r) const
    {
        for (int a = 0; a < 3; a++)
            if (c[a] - r > 0 || c[a] + r < n_[a] - 1) return false;
        return true;
    }

    // lower bound of the L-infinity distance from p to any box registered
    // only in cells outside the rings up to r (slightly low, so that the
    // rounding of cell boundaries never hides a box)
    double outside_bound(const double* p, const int* c, int r) const
    {
        double lb = kInf;
        for (int a = 0; a < 3; a++) {
            if (c[a] - r > 0) lb = std::min(lb, p[a] - (lo_[a] + (c[a] - r) * cs_[a]) - 1e-9 * cs_[a]);
            if (c[a] + r < n_[a] - 1) lb = std::min(lb, lo_[a] + (c[a] + r + 1) * cs_[a] - p[a] - 1e-9 * cs_[a]);
        }
        return lb;
    }

    // calls f(box index) for every box listed in a cell at Chebyshev ring r
    // around c (a box spanning several cells is visited once per cell)
    template <class F>
    void ring(const int* c, int r, F&& f) const
    {
        const int x0 = std::max(c[0] - r, 0), x1 = std::min(c[0] + r, n_[0] - 1);
        const int y0 = std::max(c[1] - r, 0), y1 = std::min(c[1] + r, n_[1] - 1);
        const int z0 = std::max(c[2] - r, 0), z1 = std::min(c[2] + r, n_[2] - 1);
        auto visit = [&](int x, int y, int z) {
            const size_t id = index(x, y, z);
            for (uint32_t k = start_[id]; k < start_[id + 1]; k++) f(item_[k]);
        };
        for (int x = x0; x <= x1; x++) {
            const bool ex = x == c[0] - r || x == c[0] + r;
            for (int y = y0; y <= y1; y++) {
                if (ex || y == c[1] - r || y == c[1] + r) {
                    for (int z = z0; z <= z1; z++) visit(x, y, z);
                } else {
                    if (c[2] - r >= 0) visit(x, y, c[2] - r);
                    if (r > 0 && c[2] + r < n_[2]) visit(x, y, c[2] + r);
                }
            }
        }
    }

private:
    int cell(int a, double x) const
    {
        double t = std::floor((x - lo_[a]) / cs_[a]);
        t = std::min(std::max(t, 0.0), (double)(n_[a] - 1));  // also maps NaN to cell 0
        return (int)t;
    }
    size_t index(int x, int y, int z) const { return ((size_t)x * n_[1] + y) * n_[2] + z; }

    double lo_[3], cs_[3];
    int n_[3];
    std::vector<uint32_t> start_, item_, fill_;
};

struct Geo {
    double win[6];
    std::vector<double> lz, leps;  // layer bottoms (ascending) and permittivities
    std::vector<double> box;       // [n][6] conductor boxes
    std::vector<int> id;           // conductor id per box
    std::vector<double> mbox;      // [m][6] dielectric boxes (override the layers, a later box an earlier one)
    std::vector<double> meps;      // [m]
    int ncond = 0;
    BoxGrid cgrid, mgrid;          // conductor and dielectric boxes
};

// Per-thread scratch of the geometry queries.
struct Near {
    uint32_t index;  // dielectric box
    double dist;     // L-infinity distance from the cube's centre
    double vis;      // the box shows in cubes of half side > vis
};
struct Scratch {
    std::vector<double> tmp;
    std::vector<Near> hit;        // dielectric boxes near a cube
    std::vector<uint32_t> stamp;  // last query that visited each dielectric box
    uint32_t tick = 0;
};

// Distance from p to the nearest conductor box and that box's conductor
// (the lowest box index among equally near ones).  The search stops once
// nothing closer than `enough` can remain; the result is then a lower bound
// >= enough, and the caller never uses more than `enough` of it.
inline double nearest_conductor(const Geo& g, const double* p, double enough, int* nearest)
{
    int c[3];
    g.cgrid.cell_of(p, c);
    double best = kInf, lb = kInf;
    uint32_t bi = UINT32_MAX;
    for (int r = 0;; r++) {
        g.cgrid.ring(c, r, [&](uint32_t b) {
            const double d = box_dist(&g.box[6 * b], p);
            if (d < best || (d == best && b < bi)) { best = d; bi = b; }
        });
        if (g.cgrid.covers(c, r)) { lb = kInf; break; }
        lb = g.cgrid.outside_bound(p, c, r);
        if (lb >= std::min(best, enough)) break;
    }
    *nearest = bi == UINT32_MAX ? -1 : g.id[bi];
    return std::min(best, lb);
}

// Half side up to which box B, listed before box A, is covered by A within
// the cube centred at c: B's part of the cube lies in A while the cube stays
// inside A on every side where B reaches past A (+inf: B lies in A).
inline double covered_until(const double* B, const double* A, const double* c)
{
    double t = kInf;
    for (int a = 0; a < 3; a++) {
        if (B[a] < A[a]) t = std::min(t, c[a] - A[a]);
        if (B[a + 3] > A[a + 3]) t = std::min(t, A[a + 3] - c[a]);
    }
    return t;
}

// Largest half side <= h for which the cube centred at c crosses at most
// kMaxLayers - 1 layer interfaces and shows at most kMaxCuboids dielectric
// boxes (the limits of a query's structure description,
// dnn_solver.h:28-29).  A later dielectric box overrides an earlier one
// where they overlap (in the reference's structures inner boxes are always
// listed after the boxes around them), so a box shows in the cube when it
// reaches into it and no single later box covers its part of the cube:
// above the half side vis = max(distance, covered_until by every later box).
// h drops to the (K+1)-th smallest vis, at which that box only begins to
// show.  On return s.hit holds every box within the search and its vis; the
// ones with vis < the result are the boxes that show.
inline double fit_cube(const Geo& g, const double* c, double h, Scratch& s)
{
    s.tmp.clear();
    for (size_t k = 1; k < g.lz.size(); k++) {
        const double d = std::fabs(g.lz[k] - c[2]);
        if (d < h) s.tmp.push_back(d);
    }
    if ((int)s.tmp.size() > drw::kMaxLayers - 1) {
        std::nth_element(s.tmp.begin(), s.tmp.begin() + (drw::kMaxLayers - 1), s.tmp.end());
        h = s.tmp[drw::kMaxLayers - 1];
    }
    s.hit.clear();
    if (g.meps.empty()) return h;
    if (++s.tick == 0) {
        std::fill(s.stamp.begin(), s.stamp.end(), 0u);
        s.tick = 1;
    }
    // rings until every box nearer than the current limit has been seen: a
    // box that covers another within a cube reaches into that cube, so the
    // boxes nearer than the limit settle every vis below it
    const size_t K = drw::kMaxCuboids;
    int cc[3];
    g.mgrid.cell_of(c, cc);
    double limit = h;
    for (int r = 0;; r++) {
        g.mgrid.ring(cc, r, [&](uint32_t b) {
            if (s.stamp[b] == s.tick) return;
            s.stamp[b] = s.tick;
            const double* X = &g.mbox[6 * b];
            const double d = box_dist(X, c);
            if (!(d < h)) return;
            Near x{b, d, d};
            for (Near& e : s.hit) {
                const double* E = &g.mbox[6 * (size_t)e.index];
                if (e.index > b) x.vis = std::max(x.vis, covered_until(X, E, c));
                else e.vis = std::max(e.vis, covered_until(E, X, c));
            }
            s.hit.push_back(x);
        });
        limit = h;
        if (s.hit.size() > K) {
            s.tmp.clear();
            for (const Near& e : s.hit) s.tmp.push_back(e.vis);
            std::nth_element(s.tmp.begin(), s.tmp.begin() + K, s.tmp.end());
            limit = std::min(limit, s.tmp[K]);
        }
        if (g.mgrid.covers(cc, r) || g.mgrid.outside_bound(c, cc, r) >= limit) break;
    }
    return limit;
}

// Structure description of the cube centred at c with half side h, in the
// cube's normalised coordinates (dnn_solver.h:27-30): the layers it crosses
// and the dielectric boxes that show in it (s.hit from fit_cube), the latest
// listed first, since the first box of a description that holds a voxel
// gives its value (cuda_kernels.cu:60-77).
inline void describe(const Geo& g, const double* c, double h, Scratch& s, float* layer, float* cub)
{
    for (int i = 0; i < 30; i++) layer[i] = drw::kPad;
    for (int i = 0; i < 105; i++) cub[i] = drw::kPad;
    // the bottom layer is the last whose interface lies at or below the
    // cube's bottom face in the cube's own coordinates (zn <= -1), the same
    // numbers that decide which interfaces are listed (-1 < zn < 1)
    size_t k0 = 0;
    for (size_t k = 1; k < g.lz.size(); k++)
        if ((g.lz[k] - c[2]) / h <= -1.0) k0 = k;
    layer[0] = -1.5f;
    layer[1] = (float)g.leps[k0];
    int n = 1;
    for (size_t k = k0 + 1; k < g.lz.size() && n < drw::kMaxLayers; k++) {
        const double zn = (g.lz[k] - c[2]) / h;
        if (zn > -1.0 && zn < 1.0) {
            layer[2 * n] = (float)zn;
            layer[2 * n + 1] = (float)g.leps[k];
            n++;
        }
    }
    s.hit.erase(std::remove_if(s.hit.begin(), s.hit.end(), [h](const Near& e) { return !(e.vis < h); }),
                s.hit.end());
    std::sort(s.hit.begin(), s.hit.end(), [](const Near& a, const Near& b) { return a.index > b.index; });
    int m = 0;
    for (size_t j = 0; j < s.hit.size() && m < drw::kMaxCuboids; j++) {
        const size_t i = s.hit[j].index;
        const double* b = &g.mbox[6 * i];
        for (int a = 0; a < 3; a++) {
            cub[7 * m + a] = (float)std::max((b[a] - c[a]) / h, -1.5);
            cub[7 * m + 3 + a] = (float)std::min((b[a + 3] - c[a]) / h, 1.5);
        }
        cub[7 * m + 6] = (float)g.meps[i];
        m++;
    }
}

// Gaussian surface: boundary of the union of the master's boxes, each grown
// by its own margin.  Draws pick face f with probability (area_f /
// margin_f) / Z and a uniform point on it; a draw weighs Z margin_f (the
// first hop's flux estimate scales as 1/h with h about the margin, so the
// weights come out even across boxes with tight and wide gaps), and draws
// off the union's boundary weigh zero.
struct Gauss {
    std::vector<double> box;     // [n][6] grown master boxes
    std::vector<double> margin;  // [n][6] margin of each face (face f = 6 i + 2 axis + side)
    std::vector<double> cum;     // cumulative area_f / margin_f over the faces
    double Z = 0;                // sum of area_f / margin_f
    double area = 0;             // total face area
    double min_margin = kInf;

    // Is p, a point of face (i, axis a, side s), on the boundary of the
    // union?  Not where it lies inside another grown box or on the face of
    // one that meets this face from the other side; where two faces lie in
    // one plane on the same side, the box listed first keeps the point.
    bool on_boundary(size_t i, int a, int s, const double* p) const
    {
        const double plane = box[6 * i + a + 3 * s];
        for (size_t j = 0; j < box.size() / 6; j++) {
            if (j == i) continue;
            const double* G = &box[6 * j];
            bool in = true;
            for (int b = 0; b < 3 && in; b++)
                if (b != a) in = p[b] >= G[b] && p[b] <= G[b + 3];
            if (!in) continue;
            if (plane == G[a + 3 * s]) {
                if (j < i) return false;
                continue;
            }
            if (plane >= G[a] && plane <= G[a + 3]) return false;
        }
        return true;
    }
};

// Gaussian surface around conductor `master`: each of its boxes grown by
// `frac` of the box's gap to the nearest box of another conductor and to the
// window (at most max_cube / frac), so every grown box stays clear of the
// other conductors.  Returns an empty string, or what makes the surface
// impossible.
inline std::string build_gauss(const Geo& g, int master, double frac, double max_cube, Gauss& gs)
{
    gs = Gauss();
    char msg[160];
    for (size_t i = 0; i < g.id.size(); i++) {
        if (g.id[i] != master) continue;
        const double* b = &g.box[6 * i];
        double gap = max_cube / frac;
        for (int a = 0; a < 3; a++) gap = std::min(gap, std::min(b[a] - g.win[a], g.win[a + 3] - b[a + 3]));
        size_t other = SIZE_MAX;
        for (size_t k = 0; k < g.id.size(); k++) {
            if (g.id[k] == master) continue;
            double d = -kInf;
            for (int a = 0; a < 3; a++)
                d = std::max(d, std::max(g.box[6 * k + a] - b[a + 3], b[a] - g.box[6 * k + a + 3]));
            if (d < gap) { gap = d; other = k; }
        }
        if (!(gap > 0)) {
            if (other == SIZE_MAX) std::snprintf(msg, sizeof msg, "master box %zu touches the window", i);
            else std::snprintf(msg, sizeof msg, "master box %zu touches conductor %d (box %zu)", i, g.id[other], other);
            return msg;
        }
        const double m = frac * gap;
        gs.min_margin = std::min(gs.min_margin, m);
        for (int a = 0; a < 3; a++) gs.box.push_back(b[a] - m);
        for (int a = 0; a < 3; a++) gs.box.push_back(b[a + 3] + m);
        for (int f = 0; f < 6; f++) gs.margin.push_back(m);
    }
    if (gs.box.empty()) return "master has no box";
    for (size_t i = 0; i < gs.box.size() / 6; i++) {
        const double* G = &gs.box[6 * i];
        const double ext[3] = {G[3] - G[0], G[4] - G[1], G[5] - G[2]};
        for (int a = 0; a < 3; a++) {
            const double fa = ext[(a + 1) % 3] * ext[(a + 2) % 3];
            for (int s = 0; s < 2; s++) {
                gs.area += fa;
                gs.Z += fa / gs.margin[6 * i + 2 * a + s];
                gs.cum.push_back(gs.Z);
            }
        }
    }
    return std::string();
}

// A draw on the Gaussian surface from a uniform u (the face) and three more
// (the point): p, the face's axis and side, and the draw's weight Z
// margin_f; returns whether p lies on the surface (else the draw weighs
// zero).
inline bool gauss_draw(const Gauss& gs, double u, const double* v, double* p, int* axis, int* side, double* weight)
{
    const size_t f = std::min((size_t)(std::upper_bound(gs.cum.begin(), gs.cum.end(), u * gs.Z) - gs.cum.begin()),
                              gs.cum.size() - 1);
    const size_t i = f / 6;
    const int a = (int)(f % 6) / 2, s = (int)(f % 2);
    const double* G = &gs.box[6 * i];
    for (int b = 0; b < 3; b++) p[b] = G[b] + v[b] * (G[b + 3] - G[b]);
    p[a] = G[a + 3 * s];
    *axis = a;
    *side = s;
    *weight = gs.Z * gs.margin[f];
    return gs.on_boundary(i, a, s, p);
}

}  // namespace drw::walkgeo
