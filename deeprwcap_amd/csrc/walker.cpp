// Floating-random-walk capacitance extraction on the CPU, issuing its
// transition queries through the serving engine.
//
// The reference walker is a closed binary (executable/bin/deepRWCap); this is
// an FRW of the standard form its README and query interface imply
// (README.md:124-139, dnn_solver_gradient.cpp:149-162 for the gradient
// query's (x, y, z, W sign, eps_centre) result):
//  * the Gaussian surface is the boundary of the union of the master's boxes,
//    each grown by its own margin (half its gap to the nearest other
//    conductor or the window), so it encloses the master and nothing else for
//    any box layout; a draw picks a face of the grown boxes with probability
//    proportional to area / margin and a uniform point on it, and is kept
//    when the point lies on the union's boundary (then the face's outward
//    normal n is the surface's); a kept draw weighs A = Z margin (Z = sum of
//    area / margin over the faces), a rejected one counts as a zero-weight
//    walk;
//  * the first hop is a gradient query along n's axis on the largest cube
//    centred at r that touches no conductor; it returns the next point and
//    W sign(g), so that d phi/dn(r) ~ (s/h) W sign(g) phi(next) (h = half
//    side, s = sign of n);
//  * further hops are Poisson queries on the largest conductor-free cube
//    around the current point; a walk ends within `stop_dist` of a conductor
//    (phi = 1 for that conductor's column) or of the window (phi = 0);
//  * with the walk weight w = -A eps(r) (s/h) W sign, C[master][k] = E[ w
//    1(end on k) ] over all draws for k != master, and the master's own
//    entry comes from charge conservation, C[m][m] = -E[ w 1(end on another
//    conductor or the window) ]: the gradient kernel sums to zero (the
//    gradient of a constant), so E[w] = 0 and the two forms agree, but the
//    walks that return to the master (most of them where the surface hugs the
//    master across a narrow gap) add nothing to the second, whose variance is
//    then several times smaller; it also does not multiply a kernel's
//    residual sum by the returning walks.  The direct form and E[w] are
//    reported beside it.
// Dielectrics are planar z-layers; each cube's description lists the layers
// it crosses in its own normalised coordinates (dnn_solver.h:27-30).
// Conductor and dielectric boxes are held in uniform grids, so a hop's
// geometry queries visit the cells near the walk instead of every box.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/dnnsolver_c.h"
#include "common.h"
#include "runtime.h"
#include "walker_geometry.h"

namespace {

using namespace drw::walkgeo;

struct Walk {
    double p[3];
    double w = 0;  // -A eps (s/h) W sign
    int hops = 0;
    bool first = true;
    bool stalled = false;  // no cube of at least the floor size fits
    int axis = 2;
    double sgn = 1, h = 0;
    double area = 0;  // the start draw's weight (Z margin of its face)
};

}  // namespace

extern "C" int drw_walk_capacitance(drw_server* sv, const drw_geometry* geo, int master, const drw_walk_config* wc,
                                    double* cap, double* stderr_out, double* stats)
{
    drw::clear_error();
    if (!sv || !geo || !wc || !cap || !stderr_out) return drw::set_error(DRW_EINVAL, "walk: null argument");
    if (geo->n_layers < 1 || geo->n_boxes < 1 || !geo->layer_z || !geo->layer_eps || !geo->cond_box || !geo->cond_id)
        return drw::set_error(DRW_EINVAL, "walk: bad geometry");
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
    for (int v : g.id) {
        if (v < 0) return drw::set_error(DRW_EINVAL, "walk: negative conductor id");
        g.ncond = std::max(g.ncond, v + 1);
    }
    if (master < 0 || master >= g.ncond) return drw::set_error(DRW_EINVAL, "walk: bad master %d", master);
    if (wc->walks < 1 || wc->threads < 1 || wc->concurrent < 1 || !(wc->stop_dist > 0) || !(wc->max_cube > 0) ||
        !(wc->gauss_frac > 0 && wc->gauss_frac < 1))
        return drw::set_error(DRW_EINVAL, "walk: bad configuration");
    g.cgrid.build(g.box.data(), g.id.size());
    g.mgrid.build(g.mbox.data(), g.meps.size());

    Gauss gs;
    const std::string gerr = build_gauss(g, master, wc->gauss_frac, wc->max_cube, gs);
    if (!gerr.empty()) return drw::set_error(DRW_EINVAL, "walk: %s", gerr.c_str());

    std::vector<double> sum(g.ncond, 0.0), sum2(g.ncond, 0.0);
    std::atomic<long long> hops_total{0}, walks_done{0}, draws_total{0}, window_total{0}, stalled_total{0};
    double area_total = 0, cons = 0, cons2 = 0, wsum = 0, wsum2 = 0;  // under mu
    // smallest cube worth a query: far below any feature, far above rounding
    const double h_floor = 1e-12 * std::max(g.win[3] - g.win[0], std::max(g.win[4] - g.win[1], g.win[5] - g.win[2]));
    constexpr int kMaxHops = 100000;
    std::atomic<int> failed{0};
    std::mutex mu;
    std::string first_bad;  // description of the first failure
    const int N = wc->N;

    auto win_dist = [&](const double* p) {
        double d = kInf;
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
        long long started = 0, draws = 0;
        double area_sum = 0;  // weights of the accepted draws
        std::vector<double> lsum(g.ncond, 0.0), lsum2(g.ncond, 0.0);
        long long lhops = 0, ldone = 0, lwindow = 0, lstalled = 0;
        double lcons = 0, lcons2 = 0, lw = 0, lw2 = 0;
        Scratch sc;
        sc.stamp.assign(g.meps.size(), 0u);
        // draws points until one lies on the Gaussian surface; every draw is
        // one sample of the estimator (the rejected ones weigh zero)
        auto start = [&](Walk& w) {
            for (long long tries = 0;; tries++) {
                if (tries > 10000000) return false;
                draws++;
                const double u = U(), v[3] = {U(), U(), U()};
                int a = 0, s = 0;
                if (!gauss_draw(gs, u, v, w.p, &a, &s, &w.area)) continue;
                area_sum += w.area;
                w.sgn = s ? 1.0 : -1.0;
                w.axis = a;
                w.first = true;
                w.stalled = false;
                w.hops = 0;
                w.w = 0;
                started++;
                return true;
            }
        };
        auto fail = [&](const char* msg) {
            std::lock_guard<std::mutex> lk(mu);
            if (first_bad.empty()) first_bad = msg;
            failed++;
        };
        std::vector<Walk> ws(wc->concurrent);
        std::vector<char> active(wc->concurrent, 0);
        for (int i = 0; i < wc->concurrent && started < quota; i++) {
            if (!start(ws[i])) { fail("walk: no point of the Gaussian surface found"); return; }
            active[i] = 1;
        }
        // one submission and one wait per round for all walks of this thread
        const int W = wc->concurrent;
        std::vector<int32_t> kk(W), nn(W), aa(W);
        std::vector<float> LL((size_t)W * 30), CC((size_t)W * 105), outs((size_t)W * 5);
        std::vector<uint64_t> tk(W);
        std::vector<int> idx;
        for (;;) {
            idx.clear();
            for (int i = 0; i < W; i++) {
                if (!active[i]) continue;
                Walk& w = ws[i];
                int near = -1;
                double dw = win_dist(w.p);
                double enough = std::max(std::min(dw, wc->max_cube), wc->stop_dist);
                double dc = nearest_conductor(g, w.p, enough, &near);
                if (w.stalled || (!w.first && (dc < wc->stop_dist || dw < wc->stop_dist || w.hops > kMaxHops))) {
                    if (w.stalled || w.hops > kMaxHops) lstalled++;
                    else {
                        const bool on = dc < wc->stop_dist && near >= 0;
                        if (on) {
                            lsum[near] += w.w;
                            lsum2[near] += w.w * w.w;
                        } else lwindow++;
                        if (!on || near != master) {  // charge conservation: -w off the master
                            lcons -= w.w;
                            lcons2 += w.w * w.w;
                        }
                        lw += w.w;
                        lw2 += w.w * w.w;
                    }
                    lhops += w.hops;
                    ldone++;
                    if (started >= quota) { active[i] = 0; continue; }
                    // the slot starts a new walk: its first cube is sized at the new point
                    if (!start(w)) {
                        fail("walk: no point of the Gaussian surface found");
                        active[i] = 0;
                        continue;
                    }
                    dw = win_dist(w.p);
                    enough = std::max(std::min(dw, wc->max_cube), wc->stop_dist);
                    dc = nearest_conductor(g, w.p, enough, &near);
                }
                double h = std::min(std::min(dc, dw), wc->max_cube);
                h = fit_cube(g, w.p, h, sc);
                if (!(h >= h_floor)) {
                    // more boxes than a description holds meet at this point
                    // (or a degenerate cube): the walk ends here, counted
                    w.stalled = true;
                    continue;
                }
                w.h = h;
                const int j = (int)idx.size();
                describe(g, w.p, h, sc, &LL[(size_t)j * 30], &CC[(size_t)j * 105]);
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
                    char msg[256];
                    int len = std::snprintf(msg, sizeof msg,
                                            "%s query at (%.6g, %.6g, %.6g) half side %.6g axis %d returned (%g, %g, %g",
                                            w.first ? "gradient" : "Poisson", w.p[0], w.p[1], w.p[2], w.h, w.axis,
                                            (double)out[0], (double)out[1], (double)out[2]);
                    if (len > 0 && len < (int)sizeof msg)
                        std::snprintf(msg + len, sizeof msg - len, w.first ? ", %g, %g)" : ")", (double)out[3],
                                      (double)out[4]);
                    fail(msg);
                    active[idx[j]] = 0;
                    continue;
                }
                if (w.first) {
                    // -A eps (s/h) W sign, applied when the walk ends (A: the draw's weight)
                    w.w = -w.area * (double)out[4] * w.sgn / w.h * (double)out[3];
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
        draws_total += draws;
        window_total += lwindow;
        stalled_total += lstalled;
        area_total += area_sum;
        cons += lcons;
        cons2 += lcons2;
        wsum += lw;
        wsum2 += lw2;
    };
    std::vector<std::thread> th;
    for (int t = 0; t < wc->threads; t++) th.emplace_back(worker, t);
    for (auto& t : th) t.join();
    if (failed)
        return drw::set_error(DRW_ESTATE, "walk: %d queries failed: %s", (int)failed,
                              first_bad.empty() ? drw::last_error() : first_bad.c_str());
    const double M = (double)draws_total.load();
    auto mean_err = [M](double s1, double s2, double* m, double* e) {
        *m = s1 / M;
        *e = std::sqrt(std::max(s2 / M - *m * *m, 0.0) / M);
    };
    for (int k = 0; k < g.ncond; k++) mean_err(sum[k], sum2[k], &cap[k], &stderr_out[k]);
    double direct, direct_err, wm, wm_err;
    direct = cap[master];
    direct_err = stderr_out[master];
    mean_err(cons, cons2, &cap[master], &stderr_out[master]);
    mean_err(wsum, wsum2, &wm, &wm_err);
    if (stats) {
        const double done = (double)walks_done.load();
        stats[0] = done;
        stats[1] = (double)hops_total.load() / std::max(done, 1.0);
        stats[2] = area_total / std::max(M, 1.0);  // the surface's area, estimated by the draws
        stats[3] = gs.min_margin;
        stats[4] = (double)window_total.load();
        stats[5] = (double)stalled_total.load();
        stats[6] = direct;
        stats[7] = direct_err;
        stats[8] = wm;
        stats[9] = wm_err;
    }
    return 0;
}
