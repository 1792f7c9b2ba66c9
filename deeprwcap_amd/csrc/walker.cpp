// Floating-random-walk capacitance extraction on the CPU, issuing its
// transition queries through the serving engine.
//
// The reference walker is a closed binary (executable/bin/deepRWCap); this is
// an FRW of the standard form its README and query interface imply
// (README.md:124-139, dnn_solver_gradient.cpp:149-162 for the gradient
// query's (x, y, z, W sign, eps_centre) result):
//  * a walk starts at a uniformly drawn point r of a Gaussian surface (the
//    master's bounding box grown by a margin) with outward normal n;
//  * the first hop is a gradient query along n's axis on the largest cube
//    centred at r that touches no conductor; it returns the next point and
//    W sign(g), so that d phi/dn(r) ~ (s/h) W sign(g) phi(next) (h = half
//    side, s = sign of n);
//  * further hops are Poisson queries on the largest conductor-free cube
//    around the current point; a walk ends within `stop_dist` of a conductor
//    (phi = 1 for that conductor's column) or of the window (phi = 0);
//  * C[master][k] = -E[ A eps(r) (s/h) W sign 1(end on k) ], A = area of the
//    Gaussian surface.
// Dielectrics are planar z-layers; each cube's description lists the layers
// it crosses in its own normalised coordinates (dnn_solver.h:27-30).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/dnnsolver_c.h"
#include "common.h"
#include "runtime.h"

namespace {

struct Geo {
    double win[6];
    std::vector<double> lz, leps;  // layer bottoms (ascending) and permittivities
    std::vector<double> box;       // [n][6] conductor boxes
    std::vector<int> id;           // conductor id per box
    std::vector<double> mbox;      // [m][6] dielectric boxes (override the layers, first listed wins)
    std::vector<double> meps;      // [m]
    int ncond = 0;
};

// L-infinity distance from p to an axis-aligned box (0 inside).
double box_dist(const double* b, const double* p)
{
    double d = 0;
    for (int a = 0; a < 3; a++) {
        const double lo = b[a] - p[a], hi = p[a] - b[a + 3];
        d = std::max(d, std::max(lo, hi));
    }
    return std::max(d, 0.0);
}

double eps_at(const Geo& g, double z)
{
    int i = 0;
    for (size_t k = 0; k < g.lz.size(); k++)
        if (z >= g.lz[k]) i = (int)k;
    return g.leps[i];
}

// Largest half side <= h for which the cube centred at c crosses at most
// kMaxLayers - 1 layer interfaces and overlaps at most kMaxCuboids
// dielectric boxes (the limits of a query's structure description,
// dnn_solver.h:28-29): h drops to the distance of the first element past the
// limit, which then only touches the cube.
double fit_cube(const Geo& g, const double* c, double h, std::vector<double>& tmp)
{
    tmp.clear();
    for (size_t k = 1; k < g.lz.size(); k++) {
        const double d = std::fabs(g.lz[k] - c[2]);
        if (d < h) tmp.push_back(d);
    }
    if ((int)tmp.size() > drw::kMaxLayers - 1) {
        std::nth_element(tmp.begin(), tmp.begin() + (drw::kMaxLayers - 1), tmp.end());
        h = tmp[drw::kMaxLayers - 1];
    }
    tmp.clear();
    for (size_t i = 0; i < g.meps.size(); i++) {
        const double d = box_dist(&g.mbox[6 * i], c);
        if (d < h) tmp.push_back(d);
    }
    if ((int)tmp.size() > drw::kMaxCuboids) {
        std::nth_element(tmp.begin(), tmp.begin() + drw::kMaxCuboids, tmp.end());
        h = tmp[drw::kMaxCuboids];
    }
    return h;
}

// Structure description of the cube centred at c with half side h, in the
// cube's normalised coordinates (dnn_solver.h:27-30): the layers it crosses
// and the dielectric boxes it overlaps (first listed wins, as in the
// voxelizer).
void describe(const Geo& g, const double* c, double h, float* layer, float* cub)
{
    for (int i = 0; i < 30; i++) layer[i] = drw::kPad;
    for (int i = 0; i < 105; i++) cub[i] = drw::kPad;
    int n = 0;
    layer[0] = -1.5f;
    layer[1] = (float)eps_at(g, c[2] - h);
    n = 1;
    for (size_t k = 0; k < g.lz.size() && n < drw::kMaxLayers; k++) {
        const double zn = (g.lz[k] - c[2]) / h;
        if (zn > -1.0 && zn < 1.0) {
            layer[2 * n] = (float)zn;
            layer[2 * n + 1] = (float)g.leps[k];
            n++;
        }
    }
    int m = 0;
    for (size_t i = 0; i < g.meps.size() && m < drw::kMaxCuboids; i++) {
        const double* b = &g.mbox[6 * i];
        if (!(box_dist(b, c) < h)) continue;
        for (int a = 0; a < 3; a++) {
            cub[7 * m + a] = (float)std::max((b[a] - c[a]) / h, -1.5);
            cub[7 * m + 3 + a] = (float)std::min((b[a + 3] - c[a]) / h, 1.5);
        }
        cub[7 * m + 6] = (float)g.meps[i];
        m++;
    }
}

struct Walk {
    double p[3];
    double w = 0;  // -A eps (s/h) W sign
    int hops = 0;
    bool first = true;
    int axis = 2;
    double sgn = 1, area_eps = 0, h = 0;
};

}  // namespace

extern "C" int drw_walk_capacitance(drw_server* sv, const drw_geometry* geo, int master, const drw_walk_config* wc,
                                    double* cap, double* stderr_out, double* stats)
{
    drw::clear_error();
    if (!sv || !geo || !wc || !cap || !stderr_out) return drw::set_error(DRW_EINVAL, "walk: null argument");
    Geo g;
    std::memcpy(g.win, geo->window, sizeof g.win);
    g.lz.assign(geo->layer_z, geo->layer_z + geo->n_layers);
    g.leps.assign(geo->layer_eps, geo->layer_eps + geo->n_layers);
    g.box.assign(geo->cond_box, geo->cond_box + 6 * geo->n_boxes);
    g.id.assign(geo->cond_id, geo->cond_id + geo->n_boxes);
    if (geo->n_media > 0) {
        if (!geo->media_box || !geo->media_eps) return drw::set_error(DRW_EINVAL, "walk: media arrays missing");
        g.mbox.assign(geo->media_box, geo->media_box + 6 * geo->n_media);
        g.meps.assign(geo->media_eps, geo->media_eps + geo->n_media);
    }
    for (int v : g.id) g.ncond = std::max(g.ncond, v + 1);
    if (geo->n_layers < 1 || geo->n_boxes < 1 || master < 0 || master >= g.ncond)
        return drw::set_error(DRW_EINVAL, "walk: bad geometry or master");
    if (wc->walks < 1 || wc->threads < 1 || wc->concurrent < 1 || !(wc->stop_dist > 0) || !(wc->max_cube > 0))
        return drw::set_error(DRW_EINVAL, "walk: bad configuration");
    // master bounding box and Gaussian surface margin
    double mb[6] = {1e300, 1e300, 1e300, -1e300, -1e300, -1e300};
    for (size_t i = 0; i < g.id.size(); i++)
        if (g.id[i] == master)
            for (int a = 0; a < 3; a++) {
                mb[a] = std::min(mb[a], g.box[6 * i + a]);
                mb[a + 3] = std::max(mb[a + 3], g.box[6 * i + a + 3]);
            }
    double gap = 1e300;
    for (size_t i = 0; i < g.id.size(); i++) {
        if (g.id[i] == master) continue;
        double d = 0;
        for (int a = 0; a < 3; a++) d = std::max(d, std::max(g.box[6 * i + a] - mb[a + 3], mb[a] - g.box[6 * i + a + 3]));
        gap = std::min(gap, d);
    }
    const double margin = wc->gauss_frac * (gap < 1e299 ? gap : wc->max_cube);
    double G[6];
    for (int a = 0; a < 3; a++) { G[a] = mb[a] - margin; G[a + 3] = mb[a + 3] + margin; }
    const double ext[3] = {G[3] - G[0], G[4] - G[1], G[5] - G[2]};
    const double face_area[3] = {ext[1] * ext[2], ext[0] * ext[2], ext[0] * ext[1]};
    const double area = 2 * (face_area[0] + face_area[1] + face_area[2]);

    std::vector<double> sum(g.ncond, 0.0), sum2(g.ncond, 0.0);
    std::atomic<long long> hops_total{0}, walks_done{0};
    std::atomic<int> failed{0};
    std::mutex mu;
    std::string first_bad;  // description of the first non-finite query result
    const int N = wc->N;

    auto dist_all = [&](const double* p, int* nearest) {
        double d = 1e300;
        int k = -1;
        for (size_t i = 0; i < g.id.size(); i++) {
            const double di = box_dist(&g.box[6 * i], p);
            if (di < d) { d = di; k = g.id[i]; }
        }
        *nearest = k;
        return d;
    };
    auto win_dist = [&](const double* p) {
        double d = 1e300;
        for (int a = 0; a < 3; a++) d = std::min(d, std::min(p[a] - g.win[a], g.win[a + 3] - p[a]));
        return d;
    };

    auto worker = [&](int tid) {
        uint64_t rng = (0x9E3779B97F4A7C15ull * (uint64_t)(tid + 1)) ^ wc->seed;
        auto U = [&rng]() {
            rng ^= rng >> 12; rng ^= rng << 25; rng ^= rng >> 27;
            return (double)((rng * 0x2545F4914F6CDD1Dull) >> 11) * (1.0 / 9007199254740992.0);
        };
        const long long quota = wc->walks / wc->threads + (tid < wc->walks % wc->threads ? 1 : 0);
        long long started = 0;
        std::vector<double> lsum(g.ncond, 0.0), lsum2(g.ncond, 0.0);
        long long lhops = 0, ldone = 0;
        auto start = [&](Walk& w) {
            // area-weighted face, uniform point
            double u = U() * area * 0.5;
            int ax = 0;
            while (ax < 2 && u > face_area[ax]) { u -= face_area[ax]; ax++; }
            w.sgn = U() < 0.5 ? -1.0 : 1.0;
            for (int a = 0; a < 3; a++) w.p[a] = G[a] + U() * ext[a];
            w.p[ax] = w.sgn < 0 ? G[ax] : G[ax + 3];
            w.axis = ax;
            w.first = true;
            w.hops = 0;
            w.w = 0;
            started++;
        };
        std::vector<Walk> ws(wc->concurrent);
        std::vector<char> active(wc->concurrent, 0);
        for (int i = 0; i < wc->concurrent && started < quota; i++) { start(ws[i]); active[i] = 1; }
        // one submission and one wait per round for all walks of this thread
        const int W = wc->concurrent;
        std::vector<int32_t> kk(W), nn(W), aa(W);
        std::vector<float> LL((size_t)W * 30), CC((size_t)W * 105), outs((size_t)W * 5);
        std::vector<uint64_t> tk(W);
        std::vector<int> idx;
        std::vector<double> tmp;
        for (;;) {
            idx.clear();
            for (int i = 0; i < W; i++) {
                if (!active[i]) continue;
                Walk& w = ws[i];
                int near = -1;
                double dc = dist_all(w.p, &near);
                double dw = win_dist(w.p);
                if (!w.first && (dc < wc->stop_dist || dw < wc->stop_dist || w.hops > 100000)) {
                    if (dc < wc->stop_dist && near >= 0) {
                        lsum[near] += w.w;
                        lsum2[near] += w.w * w.w;
                    }
                    lhops += w.hops;
                    ldone++;
                    if (started >= quota) { active[i] = 0; continue; }
                    // the slot starts a new walk: its first cube is sized at the new point
                    start(w);
                    dc = dist_all(w.p, &near);
                    dw = win_dist(w.p);
                }
                double h = std::min(std::min(dc, dw), wc->max_cube);
                if (w.first) h = std::min(h, margin);
                h = fit_cube(g, w.p, h, tmp);
                if (!(h > 0)) {  // a start point on a conductor or the window: geometry error
                    std::lock_guard<std::mutex> lk(mu);
                    if (first_bad.empty()) {
                        char msg[160];
                        std::snprintf(msg, sizeof msg, "walk at (%.6g, %.6g, %.6g) has no room for a cube", w.p[0],
                                      w.p[1], w.p[2]);
                        first_bad = msg;
                    }
                    failed++;
                    active[i] = 0;
                    continue;
                }
                w.h = h;
                const int j = (int)idx.size();
                describe(g, w.p, h, &LL[(size_t)j * 30], &CC[(size_t)j * 105]);
                kk[j] = w.first ? DRW_GRADIENT : DRW_POISSON;
                nn[j] = N;
                aa[j] = w.axis;
                idx.push_back(i);
            }
            const int n = (int)idx.size();
            if (n == 0) break;
            if (drw_server_submit_many(sv, n, kk.data(), nn.data(), aa.data(), LL.data(), CC.data(), nullptr,
                                       tk.data()) ||
                drw_server_wait_many(sv, n, tk.data(), outs.data(), -1)) {
                std::lock_guard<std::mutex> lk(mu);
                if (first_bad.empty()) first_bad = drw::last_error();
                failed += n;
                break;
            }
            for (int j = 0; j < n; j++) {
                Walk& w = ws[idx[j]];
                const float* out = &outs[(size_t)j * 5];
                const int nout = w.first ? 5 : 3;
                bool finite = true;
                for (int k = 0; k < nout; k++) finite = finite && std::isfinite(out[k]);
                if (!finite) {
                    std::lock_guard<std::mutex> lk(mu);
                    if (first_bad.empty()) {
                        char msg[256];
                        std::snprintf(msg, sizeof msg,
                                      "%s query at (%.6g, %.6g, %.6g) half side %.6g axis %d returned (%g, %g, %g%s",
                                      w.first ? "gradient" : "Poisson", w.p[0], w.p[1], w.p[2], w.h, w.axis,
                                      (double)out[0], (double)out[1], (double)out[2], w.first ? "" : ")");
                        first_bad = msg;
                        if (w.first) {
                            std::snprintf(msg, sizeof msg, ", %g, %g)", (double)out[3], (double)out[4]);
                            first_bad += msg;
                        }
                    }
                    failed++;
                    active[idx[j]] = 0;
                    continue;
                }
                if (w.first) {
                    // -A eps (s/h) W sign, applied when the walk ends
                    w.w = -area * (double)out[4] * w.sgn / w.h * (double)out[3];
                    w.first = false;
                }
                for (int a = 0; a < 3; a++) w.p[a] += w.h * (double)out[a];
                w.hops++;
            }
        }
        std::lock_guard<std::mutex> lk(mu);
        for (int k = 0; k < g.ncond; k++) { sum[k] += lsum[k]; sum2[k] += lsum2[k]; }
        hops_total += lhops;
        walks_done += ldone;
    };
    std::vector<std::thread> th;
    for (int t = 0; t < wc->threads; t++) th.emplace_back(worker, t);
    for (auto& t : th) t.join();
    if (failed)
        return drw::set_error(DRW_ESTATE, "walk: %d queries failed: %s", (int)failed,
                              first_bad.empty() ? drw::last_error() : first_bad.c_str());
    const double M = (double)walks_done.load();
    for (int k = 0; k < g.ncond; k++) {
        const double m = sum[k] / M, v = std::max(sum2[k] / M - m * m, 0.0);
        cap[k] = m;
        stderr_out[k] = std::sqrt(v / M);
    }
    if (stats) {
        stats[0] = M;
        stats[1] = (double)hops_total.load() / std::max(M, 1.0);
        stats[2] = area;
        stats[3] = margin;
    }
    return 0;
}
