// Host side of the network queries: weight files, batch-norm folding, MFMA
// fragment packing and the per-batch launch sequence.
//
// Weight file (.drwnet, little endian): "DRWNET01", int32 tensor count, then
// per tensor: int32 name length, name bytes, int32 ndim, int32 dims[ndim],
// float32 data.  Names are those of the reference state dicts
// (training_pytorch/src/architecture.py), e.g. "solver.z_weight.0.weight".
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "nn_common.h"
#include "nn_face.h"
#include "nn_sel.h"
#include "query.h"

namespace drw {

namespace {

struct HostTensor {
    std::vector<int> shape;
    std::vector<float> data;
};
using TensorMap = std::map<std::string, HostTensor>;

constexpr float kBnEps = 1e-5f;  // nn.BatchNorm default

int read_drwnet(const std::string& path, TensorMap& out)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return set_error(-5, "cannot open weight file '%s'", path.c_str());
    char magic[8];
    f.read(magic, 8);
    if (!f || std::memcmp(magic, "DRWNET01", 8) != 0) return set_error(-5, "'%s' is not a DRWNET01 file", path.c_str());
    int32_t n = 0;
    f.read((char*)&n, 4);
    if (!f || n < 0 || n > 100000) return set_error(-5, "'%s': bad tensor count", path.c_str());
    for (int i = 0; i < n; i++) {
        int32_t len = 0, nd = 0;
        f.read((char*)&len, 4);
        if (!f || len <= 0 || len > 4096) return set_error(-5, "'%s': bad name length", path.c_str());
        std::string name(len, '\0');
        f.read(&name[0], len);
        f.read((char*)&nd, 4);
        if (!f || nd < 0 || nd > 8) return set_error(-5, "'%s': bad rank for %s", path.c_str(), name.c_str());
        HostTensor t;
        t.shape.resize(nd);
        size_t cnt = 1;
        for (int d = 0; d < nd; d++) {
            int32_t v;
            f.read((char*)&v, 4);
            if (!f || v < 0 || v > (1 << 24)) return set_error(-5, "'%s': bad dim for %s", path.c_str(), name.c_str());
            t.shape[d] = v;
            cnt *= (size_t)v;
        }
        t.data.resize(cnt);
        f.read((char*)t.data.data(), (std::streamsize)(cnt * 4));
        if (!f) return set_error(-5, "'%s': truncated at %s", path.c_str(), name.c_str());
        out[name] = std::move(t);
    }
    return 0;
}

int get(const TensorMap& m, const std::string& k, size_t numel, const HostTensor** t)
{
    auto it = m.find(k);
    if (it == m.end()) return set_error(-5, "weight tensor '%s' missing", k.c_str());
    if (it->second.data.size() != numel)
        return set_error(-5, "weight tensor '%s' has %zu values, expected %zu", k.c_str(), it->second.data.size(), numel);
    *t = &it->second;
    return 0;
}

// scale/shift of an eval-mode batch norm: y = scale * x + shift
int bn_fold(const TensorMap& m, const std::string& k, int C, std::vector<float>& scale, std::vector<float>& shift)
{
    const HostTensor *g, *b, *mu, *var;
    int rc;
    if ((rc = get(m, k + ".weight", C, &g)) || (rc = get(m, k + ".bias", C, &b)) ||
        (rc = get(m, k + ".running_mean", C, &mu)) || (rc = get(m, k + ".running_var", C, &var)))
        return rc;
    scale.resize(C);
    shift.resize(C);
    for (int c = 0; c < C; c++) {
        const double s = (double)g->data[c] / std::sqrt((double)var->data[c] + (double)kBnEps);
        scale[c] = (float)s;
        shift[c] = (float)((double)b->data[c] - (double)mu->data[c] * s);
    }
    return 0;
}

// Bump allocator: everything of one network in one device allocation.
struct Arena {
    std::vector<uint8_t> host;
    std::vector<std::pair<size_t, void**>> fix;  // (offset, where to store the device pointer)
    template <typename T>
    void add(const std::vector<T>& v, const T** slot)
    {
        const size_t off = (host.size() + 255) & ~size_t(255);
        host.resize(off + v.size() * sizeof(T));
        std::memcpy(host.data() + off, v.data(), v.size() * sizeof(T));
        fix.push_back({off, (void**)slot});
    }
    int upload(void** dev)
    {
        DRW_HIP(hipMalloc(dev, host.size() ? host.size() : 256));
        DRW_HIP(hipMemcpy(*dev, host.data(), host.size(), hipMemcpyHostToDevice));
        for (auto& f : fix) *f.second = (uint8_t*)*dev + f.first;
        return 0;
    }
};

// B fragments of an MFMA 16x16x16 GEMM with B[k][n] = W[n][k] (W row-major
// [N][K]); lane l of (n_tile, k_tile) holds column n*16 + l%16,
// k = k_tile*16 + 4(l/16) + j.
std::vector<half_t> pack_frag(const std::vector<float>& W, int N, int K)
{
    const int NT = (N + 15) / 16, KT = (K + 15) / 16;
    std::vector<half_t> out((size_t)NT * KT * 64 * 4);
    for (int n = 0; n < NT; n++)
        for (int kt = 0; kt < KT; kt++)
            for (int l = 0; l < 64; l++)
                for (int j = 0; j < 4; j++) {
                    const int col = n * 16 + (l & 15), k = kt * 16 + 4 * (l >> 4) + j;
                    const float v = (col < N && k < K) ? W[(size_t)col * K + k] : 0.f;
                    out[(((size_t)n * KT + kt) * 64 + l) * 4 + j] = (half_t)v;
                }
    return out;
}

// The same for the face predictors' input layer (W [N][25]): K slots in the
// kernel's operand order, input_slot_channel(lane group, 4 kt + j).
std::vector<half_t> pack_frag_input(const std::vector<float>& W, int N)
{
    constexpr int K = 25, KT = 2;
    const int NT = (N + 15) / 16;
    std::vector<half_t> out((size_t)NT * KT * 64 * 4);
    for (int n = 0; n < NT; n++)
        for (int kt = 0; kt < KT; kt++)
            for (int l = 0; l < 64; l++)
                for (int j = 0; j < 4; j++) {
                    const int col = n * 16 + (l & 15), ch = input_slot_channel(l >> 4, 4 * kt + j);
                    const float v = (col < N && ch >= 0 && ch < K) ? W[(size_t)col * K + ch] : 0.f;
                    out[(((size_t)n * KT + kt) * 64 + l) * 4 + j] = (half_t)v;
                }
    return out;
}

double phi_exact(double x) { return 0.5 * std::erfc(-x / std::sqrt(2.0)); }

// Table of gelu_tab: entry 0 = (0, 0) below the range, entry n + 1 = (1, 0)
// above it, entry i + 1 = the line a + b x of Phi on interval i with slope b
// of the chord and a shifted so that the largest deviations of Phi above and
// below the line on the interval (sampled at 257 points) are equal.
std::vector<float2> phi_line_table(int n = kGeluTab, double lo = kGeluLo, double inv_h = kGeluInvH)
{
    std::vector<float2> t(n + 2);
    t[0] = make_float2(0.f, 0.f);
    t[n + 1] = make_float2(1.f, 0.f);
    const double h = 1.0 / inv_h;
    for (int i = 0; i < n; i++) {
        const double x0 = lo + (double)i * h, x1 = x0 + h;
        const double b = (phi_exact(x1) - phi_exact(x0)) / h, a = phi_exact(x0) - b * x0;
        double dmax = -1e300, dmin = 1e300;
        for (int k = 0; k <= 256; k++) {
            const double x = x0 + h * k / 256.0, d = phi_exact(x) - (a + b * x);
            dmax = std::max(dmax, d);
            dmin = std::min(dmin, d);
        }
        t[i + 1] = make_float2((float)(a + 0.5 * (dmax + dmin)), (float)b);
    }
    return t;
}

// Table of gelu_tab_sel: (Phi(x_i), Phi(x_i+1) - Phi(x_i)) at the nodes
// x_i = lo + i h, Phi(x) = erfc(-x / sqrt 2) / 2 in double.
std::vector<float2> gelu_table(int n, double lo, double inv_h)
{
    std::vector<float2> t(n);
    auto phi = [](double x) { return 0.5 * std::erfc(-x / std::sqrt(2.0)); };
    for (int i = 0; i < n; i++) {
        const double x0 = lo + (double)i / inv_h, x1 = lo + (double)(i + 1) / inv_h;
        const double p0 = phi(x0);
        t[i] = make_float2((float)p0, (float)(phi(x1) - p0));
    }
    return t;
}

std::vector<float> pad(const std::vector<float>& v, size_t n)
{
    std::vector<float> o(std::max(n, v.size()), 0.f);
    std::copy(v.begin(), v.end(), o.begin());
    return o;
}

struct FaceArch {
    int c0;
    std::vector<int> cin, cout, dil;
    int clast;
};
const FaceArch kFaceSolver{16, {16, 16, 8, 4}, {16, 8, 4, 2}, {1, 1, 2, 3}, 2};
const FaceArch kFaceSolverPlus{32, {32, 32, 24, 24, 16, 16, 8}, {32, 24, 24, 16, 16, 8, 4}, {1, 1, 1, 1, 2, 2, 3}, 4};

// FaceSolver / FaceSolverPlus (architecture.py:44-105) -> FaceNetDev
int build_face(const TensorMap& m, const FaceArch& A, Arena& ar, FaceNetDev& d)
{
    std::memset(&d, 0, sizeof d);
    const std::string pre = "solver.";
    int rc;
    const HostTensor* t;
    std::vector<float> sc, sh;
    // z_weight: 25 -> c0
    if ((rc = get(m, pre + "z_weight.0.weight", (size_t)A.c0 * 25, &t)) || (rc = bn_fold(m, pre + "z_weight.1", A.c0, sc, sh)))
        return rc;
    {
        std::vector<float> W(t->data);
        for (int o = 0; o < A.c0; o++)
            for (int i = 0; i < 25; i++) W[o * 25 + i] *= sc[o];
        std::vector<float> Wlo(W.size());
        for (size_t i = 0; i < W.size(); i++) Wlo[i] = W[i] - (float)(half_t)W[i];
        ar.add(pack_frag_input(W, A.c0), (const half_t**)&d.frag[0]);
        ar.add(pack_frag_input(Wlo, A.c0), (const half_t**)&d.frag_lo[0]);
        ar.add(pad(sh, 32), &d.pw_bias[0]);
    }
    for (size_t i = 0; i < A.cin.size(); i++) {
        const std::string k = pre + "blocks." + std::to_string(i) + ".block.";
        const int ci = A.cin[i], co = A.cout[i];
        if ((rc = get(m, k + "0.weight", (size_t)ci * 9, &t)) || (rc = bn_fold(m, k + "1", ci, sc, sh))) return rc;
        std::vector<float> dw(t->data);
        for (int c = 0; c < ci; c++)
            for (int q = 0; q < 9; q++) dw[c * 9 + q] *= sc[c];
        ar.add(dw, &d.dw_w[i]);
        ar.add(pad(sh, 32), &d.dw_b[i]);
        if ((rc = get(m, k + "3.weight", (size_t)co * ci, &t)) || (rc = bn_fold(m, k + "4", co, sc, sh))) return rc;
        std::vector<float> W(t->data);
        for (int o = 0; o < co; o++)
            for (int c = 0; c < ci; c++) W[o * ci + c] *= sc[o];
        if (co >= 8) {
            std::vector<float> Wlo(W.size());
            for (size_t q = 0; q < W.size(); q++) Wlo[q] = W[q] - (float)(half_t)W[q];
            ar.add(pack_frag(W, co, ci), (const half_t**)&d.frag[i + 1]);
            ar.add(pack_frag(Wlo, co, ci), (const half_t**)&d.frag_lo[i + 1]);
        } else {
            ar.add(W, &d.pw_w[i + 1]);
        }
        ar.add(pad(sh, 32), &d.pw_bias[i + 1]);
    }
    if ((rc = get(m, pre + "head.weight", (size_t)A.clast, &t))) return rc;
    ar.add(pad(t->data, 4), &d.head_w);
    ar.add(phi_line_table(), &d.gtab);
    if ((rc = get(m, pre + "head.bias", 1, &t))) return rc;
    d.head_b = t->data[0];
    return 0;
}

// Weights of the selectors' direct first conv (conv1_direct): per output
// channel 14 fp16 pairs.  Pair r < 9 is tap row (tz, ty) = (r/3, r%3): x-taps
// (0, 1) when tz + ty is even, else (1, 2); pairs 9..13 hold the remaining
// taps of rows 0..8 two at a time (x-tap 2 or 0 of each row, in row order;
// the last pair's second weight is zero).  Tap q = 9 tz + 3 ty + tx of the
// [C][1][3][3][3] weight.
std::vector<half_t> conv1_pairs(const std::vector<float>& W, int C)
{
    std::vector<half_t> out((size_t)C * 28, (half_t)0.f);
    for (int c = 0; c < C; c++) {
        half_t* o = out.data() + (size_t)c * 28;
        int single[10];
        for (int r = 0; r < 9; r++) {
            const int tz = r / 3, ty = r % 3, par = (tz + ty) & 1, q0 = 9 * tz + 3 * ty;
            o[2 * r] = (half_t)W[(size_t)c * 27 + q0 + par];
            o[2 * r + 1] = (half_t)W[(size_t)c * 27 + q0 + par + 1];
            single[r] = q0 + (par ? 0 : 2);
        }
        for (int k = 0; k < 9; k++) o[18 + k] = (half_t)W[(size_t)c * 27 + single[k]];
    }
    return out;
}

// FaceSelector / FaceSelectorWeight (architecture.py:176-240) -> SelNetDev
int build_sel(const TensorMap& m, bool weighted, Arena& ar, SelNetDev& d)
{
    std::memset(&d, 0, sizeof d);
    const int ch[4] = {1, weighted ? 8 : 4, weighted ? 16 : 8, weighted ? 64 : 16};
    int rc;
    const HostTensor *t, *bb;
    std::vector<float> sc, sh;
    for (int l = 0; l < 3; l++) {
        const int ci = ch[l], co = ch[l + 1], K = ci * 27;
        const std::string k = "backbone." + std::to_string(3 * l);
        if ((rc = get(m, k + ".weight", (size_t)co * K, &t)) || (rc = get(m, k + ".bias", co, &bb)) ||
            (rc = bn_fold(m, "backbone." + std::to_string(3 * l + 1), co, sc, sh)))
            return rc;
        std::vector<float> W(t->data), b(co);
        for (int o = 0; o < co; o++) {
            for (int q = 0; q < K; q++) W[(size_t)o * K + q] *= sc[o];
            b[o] = sc[o] * bb->data[o] + sh[o];
        }
        if (l == 0) {
            ar.add(conv1_pairs(W, co), &d.conv1_w);
        } else {
            // conv3d_mfma's K order: k = tap * ci + c (channel innermost)
            std::vector<float> Wt(W.size());
            for (int o = 0; o < co; o++)
                for (int c = 0; c < ci; c++)
                    for (int q = 0; q < 27; q++) Wt[(size_t)o * K + q * ci + c] = W[(size_t)o * K + c * 27 + q];
            ar.add(pack_frag(Wt, co, K), (const half_t**)&d.frag[l]);
        }
        ar.add(pad(b, 64), &d.bias[l]);
    }
    ar.add(gelu_table(kSelGeluTab, kSelGeluLo, kSelGeluInvH), &d.gtab);
    const std::vector<int> dims = weighted ? std::vector<int>{512, 256, 128, 7} : std::vector<int>{128, 64, 6};
    for (size_t j = 0; j + 1 < dims.size(); j++) {
        const std::string k = "head." + std::to_string(2 * j + 1);
        const int in = dims[j], out = dims[j + 1];
        if ((rc = get(m, k + ".weight", (size_t)out * in, &t)) || (rc = get(m, k + ".bias", out, &bb))) return rc;
        ar.add(pack_frag(t->data, out, in), (const half_t**)&d.frag[3 + j]);
        ar.add(pad(bb->data, 256), &d.bias[3 + j]);
    }
    return 0;
}

}  // namespace

struct NetWeights {
    int kind = 0;
    SelNetDev sel{};
    FaceNetDev pred{}, g1{}, g2{};
    void* dev[3] = {nullptr, nullptr, nullptr};
};

// models.txt: skip blank and '#' lines, strip trailing '# comment', trim
// (inference_cpp/src/dnn_solver.cpp:65-87).
static std::vector<std::string> model_paths(const std::string& file)
{
    std::vector<std::string> paths;
    std::ifstream f(file);
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::string p = line.substr(0, line.find('#'));
        const size_t a = p.find_first_not_of(" \t\r"), b = p.find_last_not_of(" \t\r");
        if (a != std::string::npos) paths.push_back(p.substr(a, b - a + 1));
    }
    return paths;
}

static std::string resolve(const std::string& base_file, const std::string& p)
{
    if (!p.empty() && p[0] == '/') return p;
    std::ifstream probe(p);
    if (probe) return p;
    const size_t s = base_file.find_last_of('/');
    return s == std::string::npos ? p : base_file.substr(0, s + 1) + p;
}

int net_weights_load(const char* models_txt, int kind, NetWeights** out)
{
    *out = nullptr;
    const std::string mt(models_txt);
    {
        std::ifstream probe(mt);
        if (!probe) return set_error(-5, "cannot open model list '%s'", models_txt);
    }
    const std::vector<std::string> paths = model_paths(mt);
    if (paths.size() < 5) return set_error(-5, "model list '%s' has %zu entries, expected 5", models_txt, paths.size());
    NetWeights* w = new NetWeights();
    w->kind = kind;
    int rc = 0;
    // [0] PoissonSelector, [1] PoissonPredictor, [2] GradientSelectorWeight,
    // [3] Gradient2Predictor, [4] Gradient1Predictor (inference_cpp/models.txt:4-8)
    const int idx[2][3] = {{0, 1, -1}, {2, 3, 4}};
    for (int j = 0; j < 3 && !rc; j++) {
        const int i = idx[kind][j];
        if (i < 0) break;
        TensorMap m;
        if ((rc = read_drwnet(resolve(mt, paths[i]), m))) break;
        Arena ar;
        if (j == 0) rc = build_sel(m, kind == 1, ar, w->sel);
        else if (kind == 0) rc = build_face(m, kFaceSolver, ar, w->pred);
        else if (j == 1) rc = build_face(m, kFaceSolverPlus, ar, w->g2);
        else rc = build_face(m, kFaceSolver, ar, w->g1);
        if (!rc) rc = ar.upload(&w->dev[j]);
    }
    if (rc) {
        net_weights_free(w);
        return rc;
    }
    *out = w;
    return 0;
}

const SelNetDev& net_selector(const NetWeights* w) { return w->sel; }

const FaceNetDev& net_face(const NetWeights* w, int which) { return which == 0 ? w->pred : which == 1 ? w->g1 : w->g2; }

void net_weights_free(NetWeights* w)
{
    if (!w) return;
    for (void* p : w->dev)
        if (p) (void)hipFree(p);
    delete w;
}

static int grow(uint8_t** buf, size_t* cap, size_t want)
{
    if (*cap >= want) return 0;
    if (*buf) (void)hipFree(*buf);
    *buf = nullptr;
    *cap = 0;
    const size_t n = want + want / 4 + 4096;
    DRW_HIP(hipMalloc(buf, n));
    *cap = n;
    return 0;
}

int query_nn_launch(const NetWeights* w, const QueryArgs& q, uint8_t** scratch, size_t* cap, hipStream_t stream)
{
    if (!w) return set_error(-4, "models not loaded");
    const int B = q.B;
    if (B <= 0) return 0;
    const bool grad = w->kind == 1;
    const int F = grad ? 512 : 128;
    // scratch: feat [B*F] f32 | probs [B*6] f32 | weight [B] f32 | face [B] i32
    const size_t off_probs = (size_t)B * F * 4, off_w = off_probs + (size_t)B * 24, off_face = off_w + (size_t)B * 4;
    const size_t need = off_face + (size_t)B * 4;
    int rc = grow(scratch, cap, need);
    if (rc) return rc;
    uint8_t* s = *scratch;
    float* feat = (float*)s;
    float* probs = q.probs_out ? q.probs_out : (float*)(s + off_probs);
    float* wgt = (float*)(s + off_w);
    int32_t* face = q.face_out ? q.face_out : (int32_t*)(s + off_face);
    SelBackboneArgs ba{B, q.layers, q.cuboids, grad ? q.axis : nullptr, feat};
    SelHeadArgs ha{B, feat, probs, grad ? wgt : nullptr, q.full ? nullptr : face, q.seed, q.counter, q.qid, q.ctr_dev};
    StageTimer* tm = q.timer;
    auto timed = [&](int stage, auto&& fn) -> int {
        int r;
        if (tm && (r = tm->begin(stage, stream))) return r;
        if ((r = fn())) return r;
        return tm ? tm->end(stream) : 0;
    };
    // Poisson sample mode: the fused query kernel unless the launch-shape
    // rules ask for the separate kernels (threads 1)
    const int fused = !grad && !q.full ? launch_shape(kShapePoissonFused, B) : 1;
    if (fused != 1) {
        FaceQueryArgs fa{};
        fa.B = B; fa.layers = q.layers; fa.cuboids = q.cuboids; fa.out = q.out; fa.kern_out = q.kern_out;
        fa.seed = q.seed; fa.counter = q.counter; fa.qid = q.qid; fa.ctr_dev = q.ctr_dev;
        fa.fb_list = q.fb_list; fa.fb_count = q.fb_count; fa.fb_cap = q.fb_cap; fa.fb_stats = q.fb_stats;
        fa.probs_out = q.probs_out; fa.face_out = q.face_out;
        return timed(7, [&] { return poisson_query_launch(fa, w->sel, w->pred, fused, stream); });
    }
    if ((rc = timed(0, [&] { return selector_backbone_launch(grad, ba, w->sel, stream); }))) return rc;
    if (!grad && !q.full) {
        // Poisson sample mode with the backbone as its own kernel: head, face
        // draw and predictor in one launch (the predictor's launch shape)
        FaceQueryArgs fa{};
        fa.B = B; fa.layers = q.layers; fa.cuboids = q.cuboids; fa.out = q.out; fa.kern_out = q.kern_out;
        fa.seed = q.seed; fa.counter = q.counter; fa.qid = q.qid; fa.ctr_dev = q.ctr_dev;
        fa.fb_list = q.fb_list; fa.fb_count = q.fb_count; fa.fb_cap = q.fb_cap; fa.fb_stats = q.fb_stats;
        fa.probs_out = q.probs_out; fa.face_out = q.face_out; fa.feat_in = feat;
        return timed(2, [&] { return poisson_query_launch(fa, w->sel, w->pred, launch_shape(kShapeFacePoisson, B), stream); });
    }
    if ((rc = timed(1, [&] { return selector_head_launch(grad, ha, w->sel, stream); }))) return rc;
    FaceQueryArgs fa{};
    fa.B = B; fa.full = q.full; fa.layers = q.layers; fa.cuboids = q.cuboids; fa.face = face;
    fa.axis = grad ? q.axis : nullptr; fa.weight = wgt; fa.out = q.out; fa.kern_out = q.kern_out;
    fa.seed = q.seed; fa.counter = q.counter; fa.qid = q.qid; fa.ctr_dev = q.ctr_dev;
    fa.fb_list = q.fb_list; fa.fb_count = q.fb_count; fa.fb_cap = q.fb_cap; fa.fb_stats = q.fb_stats;
    if (!grad) return timed(2, [&] { return face_query_launch(0, fa, w->pred, stream); });
    rc = timed(2, [&] {
        int r = face_query_launch(1, fa, w->g1, stream);
        return r ? r : face_query_launch(2, fa, w->g2, stream);
    });
    if (rc) return rc;
    if (q.weight_out) DRW_HIP(hipMemcpyAsync(q.weight_out, wgt, sizeof(float) * B, hipMemcpyDeviceToDevice, stream));
    return 0;
}

}  // namespace drw
