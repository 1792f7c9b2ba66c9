// CPU check of the walker's geometry (deeprwcap_amd/csrc/walker_geometry.h)
// against scans over every box, on one structure:
//
//   walker_geometry_check <geometry.bin> <points> <seed>
//
// geometry.bin (little endian): int32 n_layers, n_boxes, n_media, master;
// f64 window[6], layer_z[n_layers], layer_eps[n_layers], box[n_boxes][6];
// int32 id[n_boxes]; f64 media_box[n_media][6], media_eps[n_media],
// gauss_frac, max_cube, stop_dist.
//
// Points: random points near random boxes, and the positions of geometric
// walks (each hop moves to a random point on the surface of its cube, as a
// walk does).  At each point:
//  * nearest_conductor with no bound equals the scan (distance, conductor of
//    the lowest-index nearest box); with a bound it is exact below the bound
//    and a lower bound of the distance otherwise;
//  * fit_cube equals the scan (the layer limit, then the (K+1)-th smallest
//    half side at which a dielectric box shows in the cube: it reaches into
//    the cube and no single later box covers its part), and its list of
//    boxes that show is the scan's;
//  * describe lists those boxes, the latest first, and on every fourth point
//    the description voxelizes (N = 23, as the query path does) to the
//    permittivity of the structure itself at each voxel centre: the last
//    listed dielectric box holding it, else its planar layer.
// Gaussian surface: no grown box overlaps a box of another conductor; draws
// on the grown boxes' faces: an accepted draw must lie on the union's
// boundary (the cell behind its face filled, the one in front empty) and a
// rejected one must not, except where a coplanar face of a box listed earlier
// claims it; the mean draw weight (zero for rejected draws) must equal the
// union's boundary area computed exactly on the grid of all box coordinates.
//
// Prints "key value" lines.
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../deeprwcap_amd/csrc/walker_geometry.h"

using namespace drw::walkgeo;

namespace {

template <class T>
bool rd(FILE* f, T* p, size_t n)
{
    return std::fread(p, sizeof(T), n, f) == n;
}

double scan_nearest(const Geo& g, const double* p, int* id)
{
    double best = kInf;
    size_t bi = SIZE_MAX;
    for (size_t i = 0; i < g.id.size(); i++) {
        const double d = box_dist(&g.box[6 * i], p);
        if (d < best) { best = d; bi = i; }
    }
    *id = bi == SIZE_MAX ? -1 : g.id[bi];
    return best;
}

// fit_cube by scans: the layer limit, then for every dielectric box nearer
// than it, vis = max(distance, half side up to which a later box covers it),
// and the (K+1)-th smallest vis; fills the indices of the boxes with vis
// below the result, ascending
double scan_fit(const Geo& g, const double* c, double h, std::vector<uint32_t>& hits)
{
    std::vector<double> t;
    for (size_t k = 1; k < g.lz.size(); k++) {
        const double d = std::fabs(g.lz[k] - c[2]);
        if (d < h) t.push_back(d);
    }
    if ((int)t.size() > drw::kMaxLayers - 1) {
        std::nth_element(t.begin(), t.begin() + (drw::kMaxLayers - 1), t.end());
        h = t[drw::kMaxLayers - 1];
    }
    std::vector<uint32_t> near;
    for (size_t i = 0; i < g.meps.size(); i++)
        if (box_dist(&g.mbox[6 * i], c) < h) near.push_back((uint32_t)i);
    std::vector<double> vis(near.size());
    for (size_t j = 0; j < near.size(); j++) {
        vis[j] = box_dist(&g.mbox[6 * (size_t)near[j]], c);
        for (size_t k = j + 1; k < near.size(); k++)
            vis[j] = std::max(vis[j], covered_until(&g.mbox[6 * (size_t)near[j]], &g.mbox[6 * (size_t)near[k]], c));
    }
    t = vis;
    if (t.size() > (size_t)drw::kMaxCuboids) {
        std::nth_element(t.begin(), t.begin() + drw::kMaxCuboids, t.end());
        h = std::min(h, t[drw::kMaxCuboids]);
    }
    hits.clear();
    for (size_t j = 0; j < near.size(); j++)
        if (vis[j] < h) hits.push_back(near[j]);
    return h;
}

// Permittivity at x of the structure itself: the last listed dielectric box
// holding x (boxes grown by tol), else the planar layer (interfaces moved by
// tol).  `boxes`: ascending indices of every dielectric box that can hold x.
double eps_true(const Geo& g, const std::vector<uint32_t>& boxes, const double* x, double tol)
{
    for (size_t j = boxes.size(); j-- > 0;) {
        const size_t i = boxes[j];
        const double* b = &g.mbox[6 * i];
        bool in = true;
        for (int a = 0; a < 3 && in; a++) in = x[a] >= b[a] - tol && x[a] <= b[a + 3] + tol;
        if (in) return g.meps[i];
    }
    int k = 0;
    for (size_t l = 0; l < g.lz.size(); l++)
        if (x[2] >= g.lz[l] - tol) k = (int)l;
    return g.leps[k];
}

// Permittivity of voxel (i, j, k) of an N^3 cube from its description, as
// the query path voxelizes it (cuda_kernels.cu:44-77: the last layer whose
// bottom lies at or below the voxel centre, then the first box holding it).
float eps_described(const float* layer, const float* cub, int N, int i, int j, int k)
{
    const float x = (2.0f * i + 1.0f) / N - 1.0f, y = (2.0f * j + 1.0f) / N - 1.0f, z = (2.0f * k + 1.0f) / N - 1.0f;
    int l = 0;
    for (int q = 0; q < drw::kMaxLayers; q++) {
        if (layer[2 * q] == drw::kPad) break;
        if (z >= layer[2 * q]) l = q;
        else break;
    }
    float v = layer[2 * l + 1];
    for (int q = 0; q < drw::kMaxCuboids; q++) {
        const float* b = &cub[7 * q];
        if (b[0] == drw::kPad) break;
        if (x >= b[0] && x <= b[3] && y >= b[1] && y <= b[4] && z >= b[2] && z <= b[5]) return b[6];
    }
    return v;
}

// Exact boundary area of the union of boxes on the grid of their coordinates,
// and a lookup of whether a cell is filled.
struct Union {
    std::vector<double> x[3];
    std::vector<uint8_t> fill;
    size_t n[3] = {0, 0, 0};
    size_t at(size_t i, size_t j, size_t k) const { return (i * n[1] + j) * n[2] + k; }
    bool filled(long i, long j, long k) const
    {
        if (i < 0 || j < 0 || k < 0 || i >= (long)n[0] || j >= (long)n[1] || k >= (long)n[2]) return false;
        return fill[at(i, j, k)] != 0;
    }
    void build(const std::vector<double>& box)
    {
        const size_t nb = box.size() / 6;
        for (int a = 0; a < 3; a++) {
            x[a].clear();
            for (size_t i = 0; i < nb; i++) { x[a].push_back(box[6 * i + a]); x[a].push_back(box[6 * i + a + 3]); }
            std::sort(x[a].begin(), x[a].end());
            x[a].erase(std::unique(x[a].begin(), x[a].end()), x[a].end());
            n[a] = x[a].size() - 1;
        }
        fill.assign(n[0] * n[1] * n[2], 0);
        for (size_t b = 0; b < nb; b++) {
            size_t lo[3], hi[3];
            for (int a = 0; a < 3; a++) {
                lo[a] = std::lower_bound(x[a].begin(), x[a].end(), box[6 * b + a]) - x[a].begin();
                hi[a] = std::lower_bound(x[a].begin(), x[a].end(), box[6 * b + a + 3]) - x[a].begin();
            }
            for (size_t i = lo[0]; i < hi[0]; i++)
                for (size_t j = lo[1]; j < hi[1]; j++)
                    for (size_t k = lo[2]; k < hi[2]; k++) fill[at(i, j, k)] = 1;
        }
    }
    double area() const
    {
        double A = 0;
        for (long i = 0; i < (long)n[0]; i++)
            for (long j = 0; j < (long)n[1]; j++)
                for (long k = 0; k < (long)n[2]; k++) {
                    if (!filled(i, j, k)) continue;
                    const double dx = x[0][i + 1] - x[0][i], dy = x[1][j + 1] - x[1][j], dz = x[2][k + 1] - x[2][k];
                    A += (!filled(i - 1, j, k) + !filled(i + 1, j, k)) * dy * dz;
                    A += (!filled(i, j - 1, k) + !filled(i, j + 1, k)) * dx * dz;
                    A += (!filled(i, j, k - 1) + !filled(i, j, k + 1)) * dx * dy;
                }
        return A;
    }
    // cell index along axis b containing coordinate v (v inside a cell)
    long cell(int b, double v) const
    {
        return (long)(std::upper_bound(x[b].begin(), x[b].end(), v) - x[b].begin()) - 1;
    }
};

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s geometry.bin points seed\n", argv[0]);
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int32_t hdr[4];
    Geo g;
    double tail[3];
    bool ok = rd(f, hdr, 4) && rd(f, g.win, 6);
    const int nl = hdr[0], nb = hdr[1], nm = hdr[2], master = hdr[3];
    g.lz.resize(nl);
    g.leps.resize(nl);
    g.box.resize(6 * (size_t)nb);
    g.id.resize(nb);
    g.mbox.resize(6 * (size_t)nm);
    g.meps.resize(nm);
    std::vector<int32_t> id32(nb);
    ok = ok && rd(f, g.lz.data(), nl) && rd(f, g.leps.data(), nl) && rd(f, g.box.data(), 6 * (size_t)nb) &&
         rd(f, id32.data(), nb) && rd(f, g.mbox.data(), 6 * (size_t)nm) && rd(f, g.meps.data(), nm) && rd(f, tail, 3);
    std::fclose(f);
    if (!ok) {
        std::fprintf(stderr, "short geometry file\n");
        return 2;
    }
    for (int i = 0; i < nb; i++) {
        g.id[i] = id32[i];
        g.ncond = std::max(g.ncond, id32[i] + 1);
    }
    const double frac = tail[0], max_cube = tail[1], stop = tail[2];
    const long npts = std::atol(argv[2]);
    std::mt19937_64 rng(std::strtoull(argv[3], nullptr, 0));
    std::uniform_real_distribution<double> U(0.0, 1.0);

    auto t0 = std::chrono::steady_clock::now();
    g.cgrid.build(g.box.data(), g.id.size());
    g.mgrid.build(g.mbox.data(), g.meps.size());
    const double t_build = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    // query points: near a random box (conductor or dielectric), then the
    // hops of a geometric walk from there
    std::vector<double> pts;
    while ((long)pts.size() / 3 < npts) {
        const bool cond = nm == 0 || U(rng) < 0.5;
        const double* b = cond ? &g.box[6 * (size_t)(U(rng) * nb)] : &g.mbox[6 * (size_t)(U(rng) * nm)];
        double p[3];
        for (int a = 0; a < 3; a++) {
            const double e = b[a + 3] - b[a], m = e * std::pow(U(rng), 3.0);
            p[a] = b[a] - m + U(rng) * (e + 2 * m);
        }
        if (U(rng) < 0.2) {  // on a face of the box
            const int a = (int)(U(rng) * 3);
            p[a] = b[a + 3 * (U(rng) < 0.5)];
        }
        Scratch sc;
        sc.stamp.assign(g.meps.size(), 0u);
        for (int hop = 0; hop < 40 && (long)pts.size() / 3 < npts; hop++) {
            bool inside = true;
            for (int a = 0; a < 3; a++) inside = inside && p[a] > g.win[a] && p[a] < g.win[a + 3];
            if (!inside) break;
            pts.insert(pts.end(), p, p + 3);
            int near;
            const double dc = scan_nearest(g, p, &near);
            double dw = kInf;
            for (int a = 0; a < 3; a++) dw = std::min(dw, std::min(p[a] - g.win[a], g.win[a + 3] - p[a]));
            if (dc < stop || dw < stop) break;
            const double h = fit_cube(g, p, std::min(std::min(dc, dw), max_cube), sc);
            // a random point on the cube's surface: one axis at +-h, the others uniform
            const int a = (int)(U(rng) * 3);
            for (int k = 0; k < 3; k++) p[k] += k == a ? (U(rng) < 0.5 ? -h : h) : h * (2 * U(rng) - 1);
        }
    }
    const long P = (long)pts.size() / 3;

    long bad_near = 0, bad_bound = 0, bad_fit = 0, bad_hits = 0, bad_desc = 0, voxels = 0, voxels_edge = 0,
         bad_vox = 0;
    Scratch sc;
    sc.stamp.assign(g.meps.size(), 0u);
    std::vector<uint32_t> hs, hg;
    double t_grid = 0, t_scan = 0;
    float layer[30], cub[105];
    for (long q = 0; q < P; q++) {
        const double* p = &pts[3 * q];
        auto a0 = std::chrono::steady_clock::now();
        int id_s;
        const double d_s = scan_nearest(g, p, &id_s);
        const double h0 = std::min(d_s, max_cube);
        const double h_s = h0 > 0 ? scan_fit(g, p, h0, hs) : 0;
        auto a1 = std::chrono::steady_clock::now();
        int id_g;
        const double d_g = nearest_conductor(g, p, kInf, &id_g);
        const double h_g = h0 > 0 ? fit_cube(g, p, h0, sc) : 0;
        auto a2 = std::chrono::steady_clock::now();
        t_scan += std::chrono::duration<double>(a1 - a0).count();
        t_grid += std::chrono::duration<double>(a2 - a1).count();
        if (d_g != d_s || id_g != id_s) {
            if (bad_near < 5)
                std::printf("near_mismatch %ld %.17g %d %.17g %d\n", q, d_s, id_s, d_g, id_g);
            bad_near++;
        }
        // with a bound: exact below it, a lower bound above it
        const double enough = d_s * (0.25 + 1.5 * U(rng));
        int id_b;
        const double d_b = nearest_conductor(g, p, enough, &id_b);
        if (d_s < enough ? (d_b != d_s || id_b != id_s) : !(d_b <= d_s && d_b >= enough)) bad_bound++;
        if (!(h0 > 0)) continue;
        if (h_g != h_s) {
            if (bad_fit < 5) std::printf("fit_mismatch %ld %.17g %.17g\n", q, h_s, h_g);
            bad_fit++;
        }
        hg.clear();
        for (auto& e : sc.hit)
            if (e.vis < h_g) hg.push_back(e.index);
        std::sort(hg.begin(), hg.end());
        if (hg != hs) bad_hits++;
        describe(g, p, h_g, sc, layer, cub);
        // the description lists exactly the boxes that show, latest first
        for (size_t m = 0; m < (size_t)drw::kMaxCuboids; m++) {
            const bool used = m < hs.size();
            if (used != (cub[7 * m + 6] != drw::kPad) ||
                (used && cub[7 * m + 6] != (float)g.meps[hs[hs.size() - 1 - m]])) {
                bad_desc++;
                break;
            }
        }
        // and voxelizes to the structure's own permittivities (voxel centres
        // within 1e-6 h of an interface, where rounding decides, aside)
        if (q % 4 == 0) {
            const int N = 23;
            std::vector<uint32_t> touching;  // every box within h (grown a little) of the centre
            for (size_t i = 0; i < g.meps.size(); i++)
                if (box_dist(&g.mbox[6 * i], p) <= h_g * (1 + 1e-6)) touching.push_back((uint32_t)i);
            for (int i = 0; i < N; i++)
                for (int j = 0; j < N; j++)
                    for (int k = 0; k < N; k++) {
                        const double x[3] = {p[0] + h_g * ((2.0 * i + 1) / N - 1), p[1] + h_g * ((2.0 * j + 1) / N - 1),
                                             p[2] + h_g * ((2.0 * k + 1) / N - 1)};
                        const double tol = 1e-6 * h_g;
                        const double lo = eps_true(g, touching, x, -tol), hi = eps_true(g, touching, x, tol);
                        if (lo != hi) { voxels_edge++; continue; }
                        voxels++;
                        const float got = eps_described(layer, cub, N, i, j, k);
                        if (got != (float)lo) {
                            if (bad_vox < 3)
                                std::printf("voxel_mismatch point %ld (%.9g %.9g %.9g) h %.9g voxel %d %d %d: %g, "
                                            "described %g\n", q, p[0], p[1], p[2], h_g, i, j, k, lo, (double)got);
                            bad_vox++;
                        }
                    }
        }
    }
    std::printf("points %ld\nbuild_s %.6g\nscan_us %.6g\ngrid_us %.6g\n", P, t_build, 1e6 * t_scan / P,
                1e6 * t_grid / P);
    std::printf("bad_nearest %ld\nbad_bounded %ld\nbad_fit %ld\nbad_hits %ld\nbad_describe %ld\n", bad_near,
                bad_bound, bad_fit, bad_hits, bad_desc);
    std::printf("voxels %ld\nvoxels_at_interfaces %ld\nbad_voxels %ld\n", voxels, voxels_edge, bad_vox);

    // Gaussian surface
    Gauss gs;
    const std::string err = build_gauss(g, master, frac, max_cube, gs);
    if (!err.empty()) {
        std::printf("gauss_error %s\n", err.c_str());
        return 0;
    }
    Union un;
    un.build(gs.box);
    const double exact = un.area();
    const long D = 4 * npts;
    long acc = 0, bad_acc = 0, bad_rej = 0, coplanar = 0;
    double wsum = 0, wsum2 = 0;  // weights of the draws (zero when rejected): mean = the surface's area
    for (long d = 0; d < D; d++) {
        const double u = U(rng), v[3] = {U(rng), U(rng), U(rng)};
        double p[3], weight;
        int a, s;
        const bool on = gauss_draw(gs, u, v, p, &a, &s, &weight);
        long c[3];
        for (int b = 0; b < 3; b++) c[b] = b == a ? 0 : un.cell(b, p[b]);
        const long k = (long)(std::lower_bound(un.x[a].begin(), un.x[a].end(), p[a]) - un.x[a].begin());
        long in[3] = {c[0], c[1], c[2]}, out[3] = {c[0], c[1], c[2]};
        in[a] = s ? k - 1 : k;
        out[a] = s ? k : k - 1;
        const bool behind = un.filled(in[0], in[1], in[2]), front = un.filled(out[0], out[1], out[2]);
        const bool boundary = behind && !front;
        if (on) {
            acc++;
            wsum += weight;
            wsum2 += weight * weight;
            if (!boundary) bad_acc++;
        } else if (boundary) {
            coplanar++;  // must be a face shared in one plane with a box listed earlier
            const size_t f = std::min((size_t)(std::upper_bound(gs.cum.begin(), gs.cum.end(), u * gs.Z) -
                                               gs.cum.begin()), gs.cum.size() - 1);
            const size_t i = f / 6;
            bool shared = false;
            for (size_t j = 0; j < i && !shared; j++) {
                const double* G = &gs.box[6 * j];
                bool in_j = G[a + 3 * s] == p[a];
                for (int b = 0; b < 3 && in_j; b++)
                    if (b != a) in_j = p[b] >= G[b] && p[b] <= G[b + 3];
                shared = in_j;
            }
            if (!shared) bad_rej++;
        }
    }
    const double mean = wsum / D, var = std::max(wsum2 / D - mean * mean, 0.0);
    // every grown box stays clear of every other conductor
    long bad_clear = 0;
    for (size_t i = 0; i < gs.box.size() / 6; i++)
        for (size_t k = 0; k < g.id.size(); k++) {
            if (g.id[k] == master) continue;
            bool overlap = true;
            for (int a = 0; a < 3 && overlap; a++)
                overlap = gs.box[6 * i + a] < g.box[6 * k + a + 3] && g.box[6 * k + a] < gs.box[6 * i + a + 3];
            bad_clear += overlap;
        }
    std::printf("gauss_boxes %zu\ngauss_face_area %.17g\ngauss_exact_area %.17g\ngauss_mc_area %.17g\n"
                "gauss_mc_stderr %.17g\ngauss_min_margin %.17g\ngauss_bad_accept %ld\ngauss_bad_reject %ld\n"
                "gauss_coplanar %ld\ngauss_accepted %ld\ngauss_bad_clearance %ld\n",
                gs.box.size() / 6, gs.area, exact, mean, std::sqrt(var / D), gs.min_margin, bad_acc, bad_rej, coplanar,
                acc, bad_clear);
    return 0;
}
