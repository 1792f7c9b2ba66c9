// extern "C" boundary of libdnnsolver (declared in include/dnnsolver_c.h).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dnnsolver_c.h"
#include "common.h"
#include "fdm.h"
#include "kernels.h"
#include "query.h"
#include "runtime.h"
#include "solver.h"

using namespace drw;

extern "C" {

const char* drw_last_error(void) { return drw::last_error(); }
int drw_abi_version(void) { return 1; }

int drw_fdm_solve(int B, const int32_t* N, const int32_t* code, const float* eps, const int64_t* eps_off, double* G,
                  const int64_t* G_off, const int32_t* order, double tol, int32_t* iters, float* relres, void* stream)
{
    clear_error();
    if (B < 0) return set_error(DRW_EINVAL, "drw_fdm_solve: B < 0");
    if (B == 0) return 0;
    if (!N || !code || !eps || !eps_off || !G || !G_off) return set_error(DRW_EINVAL, "drw_fdm_solve: null pointer");
    if (!(tol > 0)) return set_error(DRW_EINVAL, "drw_fdm_solve: tol must be > 0");
    Workspace* ws = nullptr;
    int rc = get_workspace((hipStream_t)stream, &ws);
    if (rc) return rc;
    FdmBatch a{};
    a.B = B;
    a.N = N;
    a.code = code;
    a.eps = eps;
    a.eps_off = eps_off;
    a.G = G;
    a.G_off = G_off;
    a.order = order;
    a.tol = tol;
    a.inner_floor = kFdmInnerFloor;
    a.max_inner = 2000;
    a.max_rounds = 12;
    a.iters = iters;
    a.relres = relres;
    a.y_ws = ws->y_ws;
    a.s_ws = ws->s_ws;
    a.counters = ws->counters;
    a.n_hint = 0;  // per-cube N from the device array
    return fdm_launch(a, ws->fdm_grid, (hipStream_t)stream);
}

int drw_voxelize(int B, const float* layers, const float* cuboids, const int32_t* N, int fixed_N, int normalise,
                 float* out, const int64_t* out_off, float* maxv, void* stream)
{
    clear_error();
    if (B < 0) return set_error(DRW_EINVAL, "drw_voxelize: B < 0");
    if (!layers || !cuboids || !out) return set_error(DRW_EINVAL, "drw_voxelize: null pointer");
    if (!N && (fixed_N < 1 || fixed_N > 64)) return set_error(DRW_EINVAL, "drw_voxelize: fixed_N out of range");
    VoxelArgs a{B, layers, cuboids, N, fixed_N, normalise, out, out_off, maxv};
    return voxelize_launch(a, (hipStream_t)stream);
}

int drw_sample_flat(int B, const int32_t* N, const int32_t* axis, const double* G, const int64_t* G_off,
                    const float* center_eps, uint64_t seed, uint64_t counter, float* out, void* stream)
{
    clear_error();
    if (B < 0) return set_error(DRW_EINVAL, "drw_sample_flat: B < 0");
    if (!N || !G || !G_off || !out) return set_error(DRW_EINVAL, "drw_sample_flat: null pointer");
    FlatSampleArgs a{B, N, axis, G, G_off, center_eps, seed, counter, out, nullptr, nullptr, nullptr};
    return sample_flat_launch(a, (hipStream_t)stream);
}

// ------------------------------------------------------------------ solvers
int drw_solver_create(int kind, int device_idx, drw_solver** out)
{
    clear_error();
    if (!out) return set_error(DRW_EINVAL, "drw_solver_create: out is null");
    *out = nullptr;
    if (kind != DRW_POISSON && kind != DRW_GRADIENT) return set_error(DRW_EINVAL, "drw_solver_create: bad kind %d", kind);
    int ndev = 0;
    DRW_HIP(hipGetDeviceCount(&ndev));
    if (device_idx < 0 || device_idx >= ndev)
        return set_error(DRW_EINVAL, "drw_solver_create: device %d out of range (%d devices)", device_idx, ndev);
    DRW_HIP(hipSetDevice(device_idx));
    int rc = noalloc_scratch_init(device_idx);
    if (rc) return rc;
    drw_solver* s = new drw_solver();
    s->kind = kind;
    s->device = device_idx;
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&s->stats, 4 * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMemset(s->stats, 0, 4 * sizeof(uint64_t));
    if (e != hipSuccess) {
        if (s->stats) (void)hipFree(s->stats);
        if (s->stream) (void)hipStreamDestroy(s->stream);
        delete s;
        return set_error(DRW_EHIP, "drw_solver_create: %s", hipGetErrorString(e));
    }
    *out = s;
    return 0;
}

void drw_solver_destroy(drw_solver* s)
{
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    s->layers.release(); s->cuboids.release(); s->out.release(); s->eps.release(); s->cen.release();
    s->axis.release(); s->Nq.release(); s->code.release(); s->eps_off.release(); s->G_off.release();
    s->G.release(); s->scratch.release();
    s->status.release(); s->list.release(); s->count.release(); s->bad.release();
    s->fb_eps.release(); s->fb_G.release(); s->fb_y.release(); s->fb_list.release(); s->fb_count.release();
    s->fb_ctr.release();
    if (s->stats) (void)hipFree(s->stats);
    if (s->nets) net_weights_free(s->nets);
    s->timer.release();
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

int drw_solver_load_models(drw_solver* s, const char* models_txt)
{
    clear_error();
    if (!s) return set_error(DRW_EINVAL, "drw_solver_load_models: null solver");
    DRW_HIP(hipSetDevice(s->device));
    NetWeights* w = nullptr;
    int rc = net_weights_load(models_txt ? models_txt : "models.txt", s->kind, &w);
    if (rc) return rc;
    if (s->nets) net_weights_free(s->nets);
    s->nets = w;
    s->backend = DRW_BACKEND_NN;
    // FDM fallback slots of the network path (fixed size: no allocation per call)
    if ((rc = s->fb_eps.ensure((size_t)kFallbackSlots * kMaxCells)) ||
        (rc = s->fb_G.ensure((size_t)kFallbackSlots * 6 * kNetN * kNetN)) ||
        (rc = s->fb_y.ensure((size_t)kFallbackSlots * kMaxCells)) || (rc = s->fb_list.ensure(kFallbackCap)) ||
        (rc = s->fb_count.ensure(1)) || (rc = s->fb_ctr.ensure(2)))
        return rc;
    DRW_HIP(hipMemset(s->fb_ctr.p, 0, sizeof(uint32_t) * s->fb_ctr.n));
    DRW_HIP(hipMemset(s->fb_count.p, 0, sizeof(int32_t)));  // the face kernels append from zero
    DRW_HIP(hipDeviceSynchronize());
    return 0;
}

int drw_solver_use_fdm(drw_solver* s, int N, double tol)
{
    clear_error();
    if (!s) return set_error(DRW_EINVAL, "drw_solver_use_fdm: null solver");
    if (!fdm_supported_n(N)) return set_error(DRW_EINVAL, "drw_solver_use_fdm: N=%d not supported", N);
    if (!(tol > 0)) return set_error(DRW_EINVAL, "drw_solver_use_fdm: tol must be > 0");
    s->fdm_N = N;
    s->fdm_tol = tol;
    s->backend = DRW_BACKEND_FDM;
    return 0;
}

int drw_solver_backend(const drw_solver* s) { return s ? s->backend : -1; }

int drw_solver_set_seed(drw_solver* s, uint64_t seed, uint64_t counter)
{
    if (!s) return set_error(DRW_EINVAL, "drw_solver_set_seed: null solver");
    s->seed = seed;
    s->counter = counter;
    return 0;
}

int drw_solver_get_counter(const drw_solver* s, uint64_t* counter)
{
    if (!s || !counter) return set_error(DRW_EINVAL, "drw_solver_get_counter: null argument");
    *counter = s->counter;
    return 0;
}

int drw_solver_output_width(const drw_solver* s) { return s ? (s->kind == DRW_GRADIENT ? 5 : 3) : -1; }

int drw_set_launch_shape(int which, int min_items, int threads)
{
    clear_error();
    return drw::set_launch_shape(which, min_items, threads);
}

int drw_get_launch_shape(int which, int items)
{
    clear_error();
    const int t = drw::launch_shape(which, items);
    return t > 0 ? t : drw::set_error(DRW_EINVAL, "launch shape %d: no such kernel", which);
}

int drw_solver_set_timing(drw_solver* s, int enable)
{
    if (!s) return set_error(DRW_EINVAL, "set_timing: null solver");
    s->timer.enabled = enable != 0;
    return 0;
}

int drw_solver_stage_times(drw_solver* s, double* total_ms, int64_t* counts, int reset)
{
    clear_error();
    if (!s || !total_ms || !counts) return set_error(DRW_EINVAL, "stage_times: null argument");
    DRW_HIP(hipSetDevice(s->device));
    int rc = s->timer.collect();
    if (rc) return rc;
    for (int i = 0; i < StageTimer::kStages; i++) {
        total_ms[i] = s->timer.total_ms[i];
        counts[i] = s->timer.count[i];
        if (reset) { s->timer.total_ms[i] = 0; s->timer.count[i] = 0; }
    }
    return 0;
}

struct Diag {
    const uint64_t* qid = nullptr;      // per-query Philox counters (device)
    const uint64_t* ctr_dev = nullptr;  // device-resident counter base
    int32_t* face = nullptr;
    float* probs = nullptr;
    float* weight = nullptr;
    float* kern = nullptr;
};

static int solver_run(drw_solver* s, int B, const float* L, const float* C, const int32_t* ax, float* out,
                      hipStream_t st, const Diag& dg = Diag())
{
    const bool grad = s->kind == DRW_GRADIENT;
    if (s->backend == DRW_BACKEND_FDM) {
        const int N = s->fdm_N;
        int rc;
        if ((rc = s->eps.ensure((size_t)B * N * N * N)) || (rc = s->G.ensure((size_t)B * 6 * N * N)) ||
            (rc = s->Nq.ensure(B)) || (rc = s->code.ensure(B)) || (rc = s->eps_off.ensure(B)) ||
            (rc = s->G_off.ensure(B)) || (rc = s->cen.ensure(B)))
            return rc;
        StageTimer* tm = &s->timer;
        VoxelArgs va{B, L, C, nullptr, N, 0, s->eps.p, nullptr, nullptr};
        if ((rc = tm->begin(3, st)) || (rc = voxelize_launch(va, st)) || (rc = tm->end(st))) return rc;
        if ((rc = s->status.ensure(B)) || (rc = s->list.ensure(B)) || (rc = s->count.ensure(1)) ||
            (rc = s->bad.ensure(B)))
            return rc;
        if ((rc = fill_meta_launch(B, N, grad ? ax : nullptr, s->Nq.p, s->code.p, s->eps_off.p, s->G_off.p, s->bad.p,
                                   s->count.p, st)))
            return rc;
        Workspace* ws = nullptr;
        if ((rc = get_workspace(st, &ws))) return rc;
        FdmBatch a{};
        a.B = B; a.N = s->Nq.p; a.code = s->code.p; a.eps = s->eps.p; a.eps_off = s->eps_off.p;
        a.G = s->G.p; a.G_off = s->G_off.p; a.order = nullptr; a.tol = s->fdm_tol; a.inner_floor = kFdmInnerFloor;
        a.max_inner = s->fdm_max_inner; a.max_rounds = s->fdm_max_rounds; a.iters = nullptr; a.relres = nullptr;
        a.status = s->status.p;
        a.y_ws = ws->y_ws; a.s_ws = ws->s_ws; a.counters = ws->counters; a.n_hint = N;
        // Acceptance as in ggft_solver::solve (ggft_solver.cpp:565-626): cubes that did
        // not converge, or whose kernel is not finite (or, for GFT, does not sum to 1
        // within 0.01 or has a negative panel) are solved again in FP64 mode; those
        // that still fail get NaN hops.  The re-solve launch takes its count from the
        // device and exits at once when the list is empty.
        FdmBatch r = a;
        r.order = s->list.p; r.B_dev = s->count.p; r.tol = a.tol < 1e-12 ? a.tol : 1e-12;
        r.max_rounds = 40; r.max_inner = 4000;
        FdmAcceptArgs acc1{B, s->Nq.p, s->code.p, s->G.p, s->G_off.p, s->status.p, nullptr, nullptr, s->list.p,
                           s->count.p, nullptr, nullptr};
        FdmAcceptArgs acc2{B, s->Nq.p, s->code.p, s->G.p, s->G_off.p, s->status.p, s->list.p, s->count.p, nullptr,
                           nullptr, s->bad.p, s->stats};
        if ((rc = tm->begin(4, st)) || (rc = fdm_launch(a, ws->fdm_grid, st)) || (rc = fdm_accept_launch(acc1, st)) ||
            (rc = fdm_launch(r, ws->fdm_grid, st)) || (rc = fdm_accept_launch(acc2, st)) || (rc = tm->end(st)))
            return rc;
        if (grad && (rc = center_eps_launch(B, s->Nq.p, ax, s->eps.p, s->eps_off.p, s->cen.p, st))) return rc;
        FlatSampleArgs fa{B, s->Nq.p, grad ? ax : nullptr, s->G.p, s->G_off.p, grad ? s->cen.p : nullptr,
                          s->seed, s->counter, out, dg.qid, dg.ctr_dev, s->bad.p};
        if ((rc = tm->begin(5, st)) || (rc = sample_flat_launch(fa, st)) || (rc = tm->end(st))) return rc;
        s->counter += (uint64_t)B;
        return 0;
    }
    if (s->backend == DRW_BACKEND_NN) {
        QueryArgs q{};
        q.B = B; q.layers = L; q.cuboids = C; q.axis = grad ? ax : nullptr; q.out = out;
        q.face_out = dg.face; q.probs_out = dg.probs; q.weight_out = dg.weight; q.kern_out = dg.kern;
        q.qid = dg.qid;
        q.ctr_dev = dg.ctr_dev;
        q.seed = s->seed; q.counter = s->counter; q.timer = &s->timer;
        if (!s->fb_ctr.p) return set_error(DRW_ESTATE, "network fallback buffers missing: load the models first");
        nn_fallback_attach(s, q);
        int rc = query_nn_launch(s->nets, q, &s->scratch.p, &s->scratch.n, st);
        // the FDM fallback of this batch, always enqueued: it reads the count of
        // failed rows on the device
        if (!rc) rc = nn_fallback_enqueue(s, q, st, &s->timer);
        if (rc) return rc;
        s->counter += (uint64_t)B;
        return 0;
    }
    return set_error(DRW_ESTATE, "solver has no backend: call drw_solver_load_models or drw_solver_use_fdm");
}

int drw_solver_set_fdm_limits(drw_solver* s, int max_rounds, int max_inner)
{
    clear_error();
    if (!s) return set_error(DRW_EINVAL, "set_fdm_limits: null solver");
    if (max_rounds < 1 || max_inner < 1) return set_error(DRW_EINVAL, "set_fdm_limits: limits must be >= 1");
    s->fdm_max_rounds = max_rounds;
    s->fdm_max_inner = max_inner;
    return 0;
}

int drw_solver_failures(drw_solver* s, uint64_t* fdm_resolved, uint64_t* failed, uint64_t* nn_fallbacks)
{
    clear_error();
    if (!s) return set_error(DRW_EINVAL, "failures: null solver");
    DRW_HIP(hipSetDevice(s->device));
    DRW_HIP(hipStreamSynchronize(s->stream));
    uint64_t h[4] = {};
    DRW_HIP(hipMemcpy(h, s->stats, sizeof h, hipMemcpyDeviceToHost));
    if (fdm_resolved) *fdm_resolved = h[0];
    if (failed) *failed = h[1];
    if (nn_fallbacks) *nn_fallbacks = h[2];
    return 0;
}

static void warn_axis_ignored()
{
    static std::once_flag once;
    std::call_once(once, [] {
        fprintf(stderr,
                "Warning: Axis parameter provided to Green's function solver but will be ignored. "
                "(This warning will only be shown once)\n");
    });
}

int drw_solver_sample_device(drw_solver* s, int B, const float* layers, const float* cuboids, const int32_t* axis,
                             float* out, void* stream)
{
    clear_error();
    if (!s) return set_error(DRW_EINVAL, "sample: null solver");
    if (B < 0) return set_error(DRW_EINVAL, "sample: B < 0");
    if (B == 0) return 0;
    if (!layers || !cuboids || !out) return set_error(DRW_EINVAL, "sample: null pointer");
    if (s->kind == DRW_GRADIENT && !axis) return set_error(DRW_EINVAL, "Gradient solver requires axis parameter");
    if (s->kind == DRW_POISSON && axis) warn_axis_ignored();
    DRW_HIP(hipSetDevice(s->device));
    return solver_run(s, B, layers, cuboids, axis, out, stream ? (hipStream_t)stream : s->stream);
}

int drw_solver_sample_ex(drw_solver* s, int B, const float* layers, const float* cuboids, const int32_t* axis,
                         float* out, int32_t* face_out, float* probs_out, float* weight_out, float* kern_out,
                         void* stream)
{
    clear_error();
    if (!s) return set_error(DRW_EINVAL, "sample_ex: null solver");
    if (B < 0) return set_error(DRW_EINVAL, "sample_ex: B < 0");
    if (B == 0) return 0;
    if (!layers || !cuboids || !out) return set_error(DRW_EINVAL, "sample_ex: null pointer");
    if (s->kind == DRW_GRADIENT && !axis) return set_error(DRW_EINVAL, "Gradient solver requires axis parameter");
    if (s->backend != DRW_BACKEND_NN) return set_error(DRW_ESTATE, "sample_ex: network backend not loaded");
    DRW_HIP(hipSetDevice(s->device));
    Diag dg;
    dg.face = face_out; dg.probs = probs_out; dg.weight = weight_out; dg.kern = kern_out;
    return solver_run(s, B, layers, cuboids, s->kind == DRW_GRADIENT ? axis : nullptr, out,
                      stream ? (hipStream_t)stream : s->stream, dg);
}

int drw_solver_sample_ids(drw_solver* s, int B, const float* layers, const float* cuboids, const int32_t* axis,
                          const uint64_t* qid, float* out, void* stream)
{
    clear_error();
    if (!s) return set_error(DRW_EINVAL, "sample_ids: null solver");
    if (B < 0) return set_error(DRW_EINVAL, "sample_ids: B < 0");
    if (B == 0) return 0;
    if (!layers || !cuboids || !out || !qid) return set_error(DRW_EINVAL, "sample_ids: null pointer");
    if (s->kind == DRW_GRADIENT && !axis) return set_error(DRW_EINVAL, "Gradient solver requires axis parameter");
    DRW_HIP(hipSetDevice(s->device));
    Diag dg;
    dg.qid = qid;
    const uint64_t keep = s->counter;
    const int rc = solver_run(s, B, layers, cuboids, s->kind == DRW_GRADIENT ? axis : nullptr, out,
                              stream ? (hipStream_t)stream : s->stream, dg);
    s->counter = keep;  // per-query counters given: the solver's own sequence is untouched
    return rc;
}

struct drw_graph {
    drw_solver* s = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint64_t* d_counter = nullptr;
    int B = 0;
    hipEvent_t before = nullptr, after = nullptr;  // orders a caller's stream with the solver's
};

int drw_solver_graph_create(drw_solver* s, int B, const float* layers, const float* cuboids, const int32_t* axis,
                            float* out, drw_graph** g_out)
{
    clear_error();
    if (!s || !g_out || !layers || !cuboids || !out || B <= 0) return set_error(DRW_EINVAL, "graph_create: bad argument");
    if (s->kind == DRW_GRADIENT && !axis) return set_error(DRW_EINVAL, "Gradient solver requires axis parameter");
    if (s->backend < 0) return set_error(DRW_ESTATE, "graph_create: solver has no backend");
    DRW_HIP(hipSetDevice(s->device));
    drw_graph* g = new drw_graph();
    g->s = s;
    g->B = B;
    hipError_t e = hipMalloc(&g->d_counter, sizeof(uint64_t));
    if (e == hipSuccess) e = hipMemcpy(g->d_counter, &s->counter, sizeof(uint64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&g->before, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&g->after, hipEventDisableTiming);
    if (e != hipSuccess) {
        delete g;
        return set_error(DRW_EHIP, "graph_create: %s", hipGetErrorString(e));
    }
    const int32_t* ax = s->kind == DRW_GRADIENT ? axis : nullptr;
    // One eager step sizes every buffer (no allocation may happen under
    // capture).  Replay k then uses query counters c0 + k B, c0 = the
    // solver's counter now; the solver's own counter is left unchanged.
    const uint64_t keep = s->counter;
    Diag dg;
    dg.ctr_dev = g->d_counter;
    int rc = solver_run(s, B, layers, cuboids, ax, out, s->stream, dg);
    s->counter = keep;
    if (!rc && hipStreamSynchronize(s->stream) != hipSuccess) rc = set_error(DRW_EHIP, "graph_create: warm-up failed");
    if (!rc) {
        if (hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            rc = set_error(DRW_EHIP, "graph_create: begin capture failed");
        } else {
            rc = solver_run(s, B, layers, cuboids, ax, out, s->stream, dg);
            if (!rc) rc = advance_counter_launch(g->d_counter, (uint64_t)B, s->stream);
            const hipError_t ce = hipStreamEndCapture(s->stream, &g->graph);
            if (!rc && ce != hipSuccess) rc = set_error(DRW_EHIP, "graph_create: end capture failed: %s", hipGetErrorString(ce));
        }
        s->counter = keep;
        if (!rc && hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
            rc = set_error(DRW_EHIP, "graph_create: instantiate failed");
    }
    if (rc) {
        drw_graph_destroy(g);
        return rc;
    }
    *g_out = g;
    return 0;
}

int drw_graph_launch(drw_graph* g, void* stream)
{
    clear_error();
    if (!g || !g->exec) return set_error(DRW_EINVAL, "graph_launch: null graph");
    // The captured step uses the solver's buffers and the workspace and ticket
    // counters of the solver's stream, so it always replays on that stream.
    // A caller's stream is ordered around it: the replay starts after the work
    // queued on `stream` so far, and later work on `stream` waits for it.
    hipStream_t own = g->s->stream;
    hipStream_t st = stream ? (hipStream_t)stream : own;
    if (st != own) {
        DRW_HIP(hipEventRecord(g->before, st));
        DRW_HIP(hipStreamWaitEvent(own, g->before, 0));
    }
    DRW_HIP(hipGraphLaunch(g->exec, own));
    if (st != own) {
        DRW_HIP(hipEventRecord(g->after, own));
        DRW_HIP(hipStreamWaitEvent(st, g->after, 0));
    }
    return 0;
}

int drw_graph_counter(drw_graph* g, uint64_t* counter)
{
    clear_error();
    if (!g || !counter) return set_error(DRW_EINVAL, "graph_counter: null argument");
    DRW_HIP(hipStreamSynchronize(g->s->stream));
    DRW_HIP(hipMemcpy(counter, g->d_counter, sizeof(uint64_t), hipMemcpyDeviceToHost));
    return 0;
}

void drw_graph_destroy(drw_graph* g)
{
    if (!g) return;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->d_counter) (void)hipFree(g->d_counter);
    if (g->before) (void)hipEventDestroy(g->before);
    if (g->after) (void)hipEventDestroy(g->after);
    delete g;
}

int drw_solver_eval_kernels(drw_solver* s, int B, const float* layers, const float* cuboids, const int32_t* axis,
                            float* probs_out, float* weight_out, float* kern_out, void* stream)
{
    clear_error();
    if (!s) return set_error(DRW_EINVAL, "eval_kernels: null solver");
    if (B < 0) return set_error(DRW_EINVAL, "eval_kernels: B < 0");
    if (B == 0) return 0;
    if (!layers || !cuboids || !probs_out || !kern_out) return set_error(DRW_EINVAL, "eval_kernels: null pointer");
    if (s->kind == DRW_GRADIENT && (!axis || !weight_out))
        return set_error(DRW_EINVAL, "eval_kernels: gradient solvers need axis and weight_out");
    if (s->backend != DRW_BACKEND_NN || !s->nets) return set_error(DRW_ESTATE, "eval_kernels: network backend not loaded");
    DRW_HIP(hipSetDevice(s->device));
    QueryArgs q{};
    q.B = B; q.full = 1; q.layers = layers; q.cuboids = cuboids;
    q.axis = s->kind == DRW_GRADIENT ? axis : nullptr;
    q.probs_out = probs_out; q.weight_out = weight_out; q.kern_out = kern_out;
    q.seed = s->seed; q.counter = s->counter;
    return query_nn_launch(s->nets, q, &s->scratch.p, &s->scratch.n, stream ? (hipStream_t)stream : s->stream);
}

int drw_solver_sample(drw_solver* s, int B, const float* layers, const float* cuboids, const int32_t* axis,
                      float* out)
{
    clear_error();
    if (!s) return set_error(DRW_EINVAL, "sample: null solver");
    if (B < 0) return set_error(DRW_EINVAL, "sample: B < 0");
    if (B == 0) return 0;
    if (!layers || !cuboids || !out) return set_error(DRW_EINVAL, "sample: null pointer");
    if (s->kind == DRW_GRADIENT && !axis) return set_error(DRW_EINVAL, "Gradient solver requires axis parameter");
    if (s->kind == DRW_POISSON && axis) warn_axis_ignored();
    if (s->backend < 0)
        return set_error(DRW_ESTATE, "solver has no backend: call drw_solver_load_models or drw_solver_use_fdm");
    DRW_HIP(hipSetDevice(s->device));
    const int width = drw_solver_output_width(s);
    int rc;
    if ((rc = s->layers.ensure((size_t)B * kMaxLayers * 2)) || (rc = s->cuboids.ensure((size_t)B * kMaxCuboids * 7)) ||
        (rc = s->out.ensure((size_t)B * width)) || (rc = s->axis.ensure(B)))
        return rc;
    hipStream_t st = s->stream;
    DRW_HIP(hipMemcpyAsync(s->layers.p, layers, sizeof(float) * B * kMaxLayers * 2, hipMemcpyHostToDevice, st));
    DRW_HIP(hipMemcpyAsync(s->cuboids.p, cuboids, sizeof(float) * B * kMaxCuboids * 7, hipMemcpyHostToDevice, st));
    const bool grad = s->kind == DRW_GRADIENT;
    if (grad) DRW_HIP(hipMemcpyAsync(s->axis.p, axis, sizeof(int32_t) * B, hipMemcpyHostToDevice, st));
    rc = solver_run(s, B, s->layers.p, s->cuboids.p, grad ? s->axis.p : nullptr, s->out.p, st);
    if (rc) return rc;
    DRW_HIP(hipMemcpyAsync(out, s->out.p, sizeof(float) * B * width, hipMemcpyDeviceToHost, st));
    DRW_HIP(hipStreamSynchronize(st));
    return 0;
}

}  // extern "C"
