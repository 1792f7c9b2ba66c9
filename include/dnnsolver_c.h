/*
 * libdnnsolver (MI355X / gfx950) - C ABI of the deepRWCap transition-query path.
 *
 * The reference library exposes only C++ classes whose members are LibTorch
 * types (inference_cpp/src/dnn_solver.h:43-201).  This header is the plain C
 * boundary those classes are rebuilt on: opaque handles, plain pointers and
 * sizes, int return codes (0 = success, < 0 = error; the message is in
 * drw_last_error() of the calling thread) and no exceptions across the ABI.
 * include/dnn_solver.h restores the reference's class interface on top.
 *
 * Data layouts (reference conventions):
 *   layers  float[B][15][2]  (min_z, eps), padded with -9999   dnn_solver.h:27-30
 *   cuboids float[B][15][7]  (minx,miny,minz,maxx,maxy,maxz,eps) dnn_solver.cpp:264-272
 *   axis    int32[B] in {0:X, 1:Y, 2:Z}                         dnn_solver.h:32-36
 *   Poisson output  float[B][3]  next-hop point                  dnn_solver_poisson.cpp:111
 *   gradient output float[B][5]  (x, y, z, W*sign, eps_centre)   dnn_solver_gradient.cpp:268-273
 *   cubes  float, [z][y][x] (x fastest); FDM kernels G double[6 N^2] in GGFT
 *   panel order (faces z-, z+, y-, y+, x-, x+; ggft/src/ggft_solver.h:65-77).
 *
 * Functions taking a `stream` argument take device pointers and enqueue on
 * that hipStream_t (NULL = the null stream) without synchronising.
 */
#ifndef DNNSOLVER_C_H
#define DNNSOLVER_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DRW_OK 0
#define DRW_EINVAL (-2)   /* bad argument */
#define DRW_EHIP (-3)     /* HIP runtime error */
#define DRW_ESTATE (-4)   /* object not ready (e.g. models not loaded) */
#define DRW_EIO (-5)      /* file missing or malformed */
#define DRW_ETIMEOUT (-6) /* a wait ran out of time; the ticket stays valid */

enum drw_kind { DRW_POISSON = 0, DRW_GRADIENT = 1 };
enum drw_backend { DRW_BACKEND_NN = 0, DRW_BACKEND_FDM = 1 };
enum drw_code { DRW_GFT = 0, DRW_WVTZ = 1, DRW_WVTX = 2, DRW_WVTY = 3 };

/* Message of the last failed call on this thread ("" if none). */
const char* drw_last_error(void);
int drw_abi_version(void);

/* ------------------------------------------------------------------------
 * Batched FDM surface Green's function (the GGFT accuracy reference / fallback).
 * Replaces ggft_solver::solve (ggft/src/ggft_solver.cpp:521-636) for a batch:
 * per cube b of resolution N[b] in {8, 9, 16, 21, 23}, right-hand side
 * code[b] (DRW_GFT/WVTZ/WVTX/WVTY), writes G = 2 eps y on the 6 N^2 panels
 * (the reference's gft = A13 x).  FP32 Jacobi-PCG with FP64 residual
 * refinement until ||e - L y|| <= tol ||e||; tol <= 1e-12 gives FP64-accurate
 * kernels.  order (optional) is the processing order of the cubes; iters /
 * relres (optional) receive the FP32 iteration count (-1: unsupported N) and
 * the final FP64 relative residual of each cube.
 * ------------------------------------------------------------------------ */
int drw_fdm_solve(int B, const int32_t* N, const int32_t* code, const float* eps, const int64_t* eps_off,
                  double* G, const int64_t* G_off, const int32_t* order, double tol, int32_t* iters, float* relres,
                  void* stream);

/* Voxelize the structure description (fill_structure_kernel,
 * inference_cpp/src/cuda_kernels.cu:21-84, with the per-query eps/max
 * normalisation of dnn_solver.cpp:274-285 when normalise != 0).  N (optional)
 * gives each query's resolution, else fixed_N; out_off (optional) the float
 * offset of each cube, else b*N^3; maxv (optional) receives the per-query max. */
int drw_voxelize(int B, const float* layers, const float* cuboids, const int32_t* N, int fixed_N, int normalise,
                 float* out, const int64_t* out_off, float* maxv, void* stream);

/* Flat inverse-CDF next hop over each query's FDM kernel (ggft_solver.cpp:
 * 638-685 semantics) with Philox4x32-10 uniforms (key seed, counter
 * counter + b).  axis == NULL: Poisson output [B][3]; else gradient output
 * [B][5] with W = sum|G| and eps_centre from center_eps (optional). */
int drw_sample_flat(int B, const int32_t* N, const int32_t* axis, const double* G, const int64_t* G_off,
                    const float* center_eps, uint64_t seed, uint64_t counter, float* out, void* stream);

/* ------------------------------------------------------------------------
 * Solver objects: DNNSolverPoisson / DNNSolverGrad (dnn_solver.h:98-149).
 * ------------------------------------------------------------------------ */
typedef struct drw_solver drw_solver;

/* DNNSolver(int device_idx) ctor, dnn_solver.cpp:13-38.  The solver owns a
 * HIP stream and its device buffers.  Not thread-safe (like the reference). */
int drw_solver_create(int kind, int device_idx, drw_solver** out);
void drw_solver_destroy(drw_solver* s);

/* loadModels(), dnn_solver_poisson.cpp:11-33 / dnn_solver_gradient.cpp:11-38:
 * reads the index-addressed model list (models.txt semantics,
 * dnn_solver.cpp:65-87; NULL = "models.txt" in the working directory) and
 * selects the network backend. */
int drw_solver_load_models(drw_solver* s, const char* models_txt);

/* Select the FDM backend (the plain batched-FDM fallback and accuracy reference): queries
 * are voxelized at resolution N and solved to relative residual tol. */
int drw_solver_use_fdm(drw_solver* s, int N, double tol);
int drw_solver_backend(const drw_solver* s);

/* Philox key and the next query counter (each sampled query consumes one). */
int drw_solver_set_seed(drw_solver* s, uint64_t seed, uint64_t counter);
int drw_solver_get_counter(const drw_solver* s, uint64_t* counter);

/* Failure handling (ggft_solver::solve, ggft/src/ggft_solver.cpp:565-626).
 * FDM backend: a cube that did not converge within the first solve's limits,
 * or whose kernel is not finite, or (GFT) does not sum to 1 within 0.01 or
 * has a negative panel, is solved again in FP64 mode (tol 1e-12, where the
 * reference falls back to SparseLU); if it still fails, its hop is NaN.
 * Network backend: a query whose result holds a NaN (a failed draw: a
 * non-finite or all-zero kernel or face distribution) is solved by the FDM at
 * N = 23 with the same acceptance and drawn with its own Philox counter (no
 * counterpart in the reference, which has no fallback for its networks).
 * set_fdm_limits bounds the first FDM solve (refinement rounds, FP32
 * iterations per round; defaults 12, 2000).  failures waits for the solver's
 * stream and returns the running counts: FDM re-solves, queries left failed
 * (NaN hops), network queries sent to the FDM fallback. */
int drw_solver_set_fdm_limits(drw_solver* s, int max_rounds, int max_inner);
int drw_solver_failures(drw_solver* s, uint64_t* fdm_resolved, uint64_t* failed, uint64_t* nn_fallbacks);

/* Output floats per query: 3 (Poisson) or 5 (gradient). */
int drw_solver_output_width(const drw_solver* s);

/* Per-stage kernel timing with HIP events on the launch stream (the
 * counterpart of the walker's "GPU Sampler Statistics",
 * executable/run_script.py:108-162).  Stages: 0 selector backbone,
 * 1 selector head, 2 face predictor, 3 voxelize, 4 FDM solve, 5 flat
 * sampler, 6 network fallback, 7 fused Poisson query.  stage_times waits for the recorded events and returns the
 * accumulated milliseconds and launch counts of 8 stages (reset != 0 clears). */
int drw_solver_set_timing(drw_solver* s, int enable);
int drw_solver_stage_times(drw_solver* s, double* total_ms, int64_t* counts, int reset);

/* Launch shapes (no counterpart in the reference, whose TensorRT engines fix
 * theirs at build time): threads per workgroup of the tunable kernels,
 * process-wide and read at every launch, so they apply to every solver and
 * server.  Per kernel, rules (min_items, threads): a launch of `items`
 * workgroups uses the rule with the largest min_items <= items, else 256
 * threads.  set with min_items >= 0 adds or replaces that rule (threads 0
 * removes it; at most 8 rules); with min_items < 0 it makes `threads` the
 * shape of every size, or with threads 0 restores the built-in rules (the
 * autotuner's choice on MI355X).  get returns the threads a launch of
 * `items` workgroups uses.  Chosen per machine by the autotuner
 * (deeprwcap_amd.autotune).  DRW_SHAPE_FACE_* are the face predictors'
 * kernels (one workgroup per query and face, 256 or 512 threads).
 * DRW_SHAPE_POISSON_FUSED is the Poisson query in one kernel (sample mode:
 * selector, head, face draw and predictor in one workgroup per query, 256 or
 * 512 threads), or 1: the selector backbone as its own kernel, then head,
 * face draw and predictor in one kernel (DRW_SHAPE_FACE_POISSON's threads);
 * both give bit-identical results. */
enum drw_shape {
    DRW_SHAPE_FACE_POISSON = 0,
    DRW_SHAPE_FACE_GRAD1 = 1,
    DRW_SHAPE_FACE_GRAD2 = 2,
    DRW_SHAPE_POISSON_FUSED = 3
};
int drw_set_launch_shape(int which, int min_items, int threads);
int drw_get_launch_shape(int which, int items);

/* sample(layers, cuboids[, axis]) with host pointers; blocks until `out`
 * holds B * width floats.  B = size(layers) / 30 in the reference
 * (dnn_solver_poisson.cpp:40).  Gradient solvers require axis (the reference
 * throws, dnn_solver_gradient.cpp:40-47); Poisson solvers ignore it. */
int drw_solver_sample(drw_solver* s, int B, const float* layers, const float* cuboids, const int32_t* axis,
                      float* out);

/* Same with device pointers, enqueued on `stream` (NULL: the solver's own). */
int drw_solver_sample_device(drw_solver* s, int B, const float* layers, const float* cuboids, const int32_t* axis,
                             float* out, void* stream);

/* Same with a caller-given Philox counter per query (device array qid[B]):
 * a query's hop depends only on its own input and counter, not on the batch
 * it travels in (used by the serving engine). */
int drw_solver_sample_ids(drw_solver* s, int B, const float* layers, const float* cuboids, const int32_t* axis,
                          const uint64_t* qid, float* out, void* stream);

/* hipGraph-captured step: one call records the whole query batch (voxelize,
 * networks or FDM, draws, locate) on fixed device buffers plus a
 * device-resident counter advance, so each drw_graph_launch is a single
 * submission that draws fresh uniforms: launch k uses query counters
 * c0 + k B with c0 the solver's counter at creation (the solver's own counter
 * is not advanced; drw_graph_counter reads the graph's).  The graph must be
 * destroyed before its solver.  A replay always runs on the solver's own
 * stream (the captured step uses that stream's workspace and ticket
 * counters); a `stream` argument other than that is ordered around it with
 * events (the replay waits for the work queued on `stream`, later work on
 * `stream` waits for the replay).  drw_graph_counter waits for the solver's
 * stream.  A solver, and its graphs, serve one stream of work at a time. */
typedef struct drw_graph drw_graph;
int drw_solver_graph_create(drw_solver* s, int B, const float* layers, const float* cuboids, const int32_t* axis,
                            float* out, drw_graph** graph);
int drw_graph_launch(drw_graph* graph, void* stream);
int drw_graph_counter(drw_graph* graph, uint64_t* counter);
void drw_graph_destroy(drw_graph* graph);

/* Network backend with diagnostics (device pointers, each optional):
 * face_out [B] drawn face, probs_out [B][6] selector probabilities,
 * weight_out [B] predicted W (gradient), kern_out [B][529] normalised kernel
 * of the drawn face (post-processed for gradients) from which the tile was
 * drawn.  Replaces the intermediate tensors of samplePoissonFunction /
 * sampleGradient (dnn_solver_poisson.cpp:86-111, dnn_solver_gradient.cpp:
 * 111-164). */
int drw_solver_sample_ex(drw_solver* s, int B, const float* layers, const float* cuboids, const int32_t* axis,
                         float* out, int32_t* face_out, float* probs_out, float* weight_out, float* kern_out,
                         void* stream);

/* Allocation-free drw_solver_sample_device.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels, results and counter use as drw_solver_sample_device; the per-query
 * intermediates live in a fixed 16.25 MiB device array that the library holds
 * on each GPU it runs on until the process ends, so no call allocates or frees
 * anything.  Network backend only (DRW_ESTATE otherwise); B <= 8192
 * (DRW_EINVAL beyond).  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (other streams, threads or solvers) race. */
int drw_solver_sample_device_noalloc(drw_solver* s, int B, const float* layers, const float* cuboids,
                                     const int32_t* axis, float* out, void* stream);

/* Full transition kernels from the networks (no draws), for accuracy checks
 * against the FDM reference: probs_out [B][6], kern_out [B][6][529] (face f of
 * the query's frame, tile t = 23 b + a in GGFT panel order), weight_out [B]
 * (gradient; required there).  Poisson: G[f][t] = probs[f] * kern[f][t];
 * gradient: G[remap(f)][t] = W * probs[f] * kern[f][t]. */
int drw_solver_eval_kernels(drw_solver* s, int B, const float* layers, const float* cuboids, const int32_t* axis,
                            float* probs_out, float* weight_out, float* kern_out, void* stream);

/* ------------------------------------------------------------------------
 * Serving engine (replaces ParallelSampler, parallel_sampler.cpp:42-211, for
 * callers that mix query kinds and resolutions): thread-safe submit/wait,
 * continuous batching with up to max_inflight batches on the device (one HIP
 * stream each), admission bounded by an HBM budget, per-request Philox
 * counters, latency percentiles.  N = 23 queries go to the networks when a
 * model list is given; other N (16, 21) and model-less servers use the FDM
 * fallback at the query's N.
 * ------------------------------------------------------------------------ */
typedef struct drw_server drw_server;
typedef struct {
    int max_batch;         /* queries per launch (default 2048) */
    int max_inflight;      /* concurrent batches / streams (default 4) */
    double hbm_budget_mb;  /* device memory for in-flight queries (default 1024) */
    double fdm_tol;        /* FDM relative residual (default 1e-6) */
    uint64_t seed;         /* Philox key */
    /* Resident loop (no counterpart in the reference): with resident = 1 and
     * a model list, N = 23 queries of both kinds skip the batches; a
     * persistent kernel takes them from a pinned request ring as they are
     * submitted, serves each in one workgroup and writes the result back,
     * and waiters read it from the ring.  FDM-route queries (N = 16, 21)
     * still run as batches: the dispatcher stops the resident kernel while
     * they run and restarts it afterwards.  Results are bit-identical to the
     * batched route's for the same counter. */
    int resident;              /* 0: batches only (default); 1: resident loop for N = 23 */
    int resident_threads;      /* threads per resident workgroup, 256 or 512 (0: 512) */
    double resident_idle_ms;   /* the resident kernel ends after this long without queries (0: 20 ms) */
} drw_server_config;
typedef struct {
    uint64_t completed, batches, inflight_peak;
    double mean_batch, p50_ms, p99_ms, max_ms, elapsed_s, qps;
} drw_server_stats;

int drw_server_create(int device, const char* models_txt, const drw_server_config* cfg, drw_server** out);
/* kind DRW_POISSON/DRW_GRADIENT, N in {16, 21, 23}, axis (gradient),
 * layer[30], cuboid[105]; qid = Philox counter of this query (0: the ticket). */
int drw_server_submit(drw_server* sv, int kind, int N, int axis, const float* layer, const float* cuboid, uint64_t qid,
                      uint64_t* ticket);
/* Blocks until the query is done (timeout_ms < 0: no limit; DRW_ETIMEOUT when
 * the limit passes); out receives 3 or 5 floats.  Each ticket can be waited
 * for once; at most 65536 tickets can be outstanding (submitted and not yet
 * waited for). */
int drw_server_wait(drw_server* sv, uint64_t ticket, float* out, int timeout_ms);
/* Batch forms for callers with many queries in flight (a walker thread with
 * many walks): one call submits n queries (kinds/Ns/axes[n], layers[n][30],
 * cuboids[n][105], qids[n] or NULL) and returns their tickets; one call waits
 * for n tickets in order, outs[n][5] (Poisson results use the first 3).  On
 * DRW_ETIMEOUT the tickets before the one that timed out are consumed and the
 * rest stay valid.  No counterpart in the reference, whose walker submits one
 * task at a time (parallel_sampler.cpp:42-70). */
int drw_server_submit_many(drw_server* sv, int n, const int32_t* kinds, const int32_t* Ns, const int32_t* axes,
                           const float* layers, const float* cuboids, const uint64_t* qids, uint64_t* tickets);
int drw_server_wait_many(drw_server* sv, int n, const uint64_t* tickets, float* outs, int timeout_ms);
int drw_server_stats_get(drw_server* sv, drw_server_stats* stats, int reset);
/* Resident-loop counters (zeros for a batches-only server), out[10]: kernel
 * launches, Poisson and gradient queries served, scout polls of the request
 * ring, workgroups that left idle and on the stop flag, claimed and
 * published sequences, workgroups per launch, device wall clock (kHz).  No
 * counterpart in the reference. */
int drw_server_resident_stats(drw_server* sv, uint64_t* out);
/* Drains pending queries, then frees everything. */
void drw_server_destroy(drw_server* sv);

/* Open-loop load generator standing in for the walker threads: `threads` host
 * threads submit queries from a pool (layers[pool][30], cuboids[pool][105],
 * kinds/Ns/axes[pool]) with Poisson arrivals at `rate_qps` in total (<= 0:
 * back-to-back), each keeping at most `max_outstanding` tickets, for
 * `duration_s`; returns the server statistics of the run (p50/p99 latency is
 * submit-to-completion). */
typedef struct {
    int threads;
    int max_outstanding;
    double rate_qps;
    double duration_s;
    uint64_t seed;
} drw_load_config;
int drw_server_run_load(drw_server* sv, const drw_load_config* load, const float* layers, const float* cuboids,
                        const int32_t* kinds, const int32_t* Ns, const int32_t* axes, int pool, drw_server_stats* stats);

/* ------------------------------------------------------------------------
 * Floating-random-walk capacitance extraction (the consumer of the query
 * path; stands in for the closed walker executable/bin/deepRWCap).  Planar
 * z-layers of dielectric, optional dielectric boxes over them, conductors
 * as unions of boxes (cond_id: the conductor, i.e. net, of each box),
 * Dirichlet 0 on the window.  Walks start on a Gaussian surface around
 * `master` (the boundary of the union of its boxes, each grown by gauss_frac
 * of its gap to the nearest other conductor or the window; faces drawn in
 * proportion to area / margin), take one gradient query then Poisson queries
 * through `sv` until they come within stop_dist of a conductor; cap[k]
 * receives C[master][k] (the units of the coordinates times eps; multiply by
 * eps0 for farads), the master's own entry from charge conservation (minus
 * the flux of the walks that end on another conductor or the window),
 * stderr_out[k] its standard error; stats (optional, 10 doubles): walks,
 * mean hops per walk, Gaussian-surface area, smallest Gaussian margin, walks
 * that ended at the window, walks that ended elsewhere (where more
 * dielectric boxes meet than a description holds, no cube of 1e-12 of the
 * window fits; or after 100,000 hops), the master's entry estimated
 * directly (the flux of the walks that return to it) and its standard
 * error, the mean walk weight E[w] (zero for kernels that sum to zero: the
 * difference of the two estimates) and its standard error.  Fails
 * (DRW_EINVAL) when a master box touches another conductor or the window.
 * ------------------------------------------------------------------------ */
typedef struct {
    double window[6];          /* xmin ymin zmin xmax ymax zmax */
    int n_layers;
    const double* layer_z;     /* [n_layers] bottom z of each layer, ascending (first = -inf side) */
    const double* layer_eps;   /* [n_layers] */
    int n_boxes;
    const double* cond_box;    /* [n_boxes][6] */
    const int32_t* cond_id;    /* [n_boxes] conductor index of each box */
    int n_media;               /* dielectric boxes overriding the layers (0: none) */
    const double* media_box;   /* [n_media][6] */
    const double* media_eps;   /* [n_media] */
} drw_geometry;
typedef struct {
    int walks;                 /* total walks */
    int threads;               /* host walker threads */
    int concurrent;            /* walks in flight per thread */
    int N;                     /* cube resolution (23: networks when the server has models) */
    double stop_dist;          /* termination distance to a conductor / the window */
    double gauss_frac;         /* Gaussian margin of each master box as a fraction of its gap (0..1) */
    double max_cube;           /* largest cube half side */
    uint64_t seed;
} drw_walk_config;
int drw_walk_capacitance(drw_server* sv, const drw_geometry* geo, int master, const drw_walk_config* cfg, double* cap,
                         double* stderr_out, double* stats);

#ifdef __cplusplus
}
#endif

#endif /* DNNSOLVER_C_H */
