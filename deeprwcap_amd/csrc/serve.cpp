// Serving engine: continuous batching of transition queries of mixed kind and
// resolution in front of the GPU, built on the public C ABI.
//
// The reference serves walker threads through ParallelSampler
// (inference_cpp/src/parallel_sampler.cpp:42-211): one worker takes up to
// 2048 tasks, runs one synchronous batch, and only then looks at the queue
// again.  Here a dispatcher thread keeps up to `max_inflight` batches on the
// device at once, each on its own HIP stream with pinned staging buffers:
// requests that arrive while batches are running join the next batch as soon
// as a slot and the HBM budget allow, so host staging, H2D, kernels and D2H
// of different batches overlap.  Each request carries its own Philox counter,
// so its hop is identical whether it runs alone or in a mix.
//
// Classes: (kind, route) with route = network at N = 23 when models are
// loaded, else the FDM fallback at the request's N.  The oldest pending
// request decides which class is launched next.
//
// Host path: requests live in a fixed ring indexed by ticket (no per-request
// allocation); an entry's state is an atomic (free -> claimed -> queued ->
// done -> free), the class queues, the completion wake-ups and the statistics
// have separate locks, and callers that issue many queries at once (a walker
// thread with many walks in flight) submit and wait for them in one call.
//
// Resident mode (cfg.resident): network queries (N = 23, both kinds) bypass
// the batches.  submit() writes them into the pinned request ring of the
// resident kernel (resident.h) and publishes them; the kernel, once
// started by the dispatcher, serves them with no host step per query, and
// waiters read the results from the response ring.  While FDM-route batches
// are pending or running, the dispatcher stops the resident kernel (its
// workgroups hold every CU) and starts it again afterwards.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/dnnsolver_c.h"
#include "common.h"
#include "query.h"
#include "resident.h"
#include "runtime.h"

using Clock = std::chrono::steady_clock;

namespace {

constexpr int kIn = drw::kMaxLayers * 2 + drw::kMaxCuboids * 7;  // 135 floats
constexpr int kRoutes = 4;                                        // NN23, FDM16, FDM21, FDM23
constexpr uint64_t kRing = 1u << 16;  // tickets outstanding (submitted, not yet waited for) at most
constexpr int64_t kFdmHoldNs = 5000000;  // resident mode: network queries batch for 5 ms after an FDM-route query

enum : int { kFree = 0, kClaimed = 1, kQueued = 2, kDone = 3 };

struct Req {
    std::atomic<uint64_t> ticket{0};  // ticket holding this entry (0: none)
    std::atomic<int> state{kFree};
    uint64_t qid = 0;
    int kind = 0, N = 23, axis = 2, cls = 0, rc = 0;
    float in[kIn];
    float out[5];
    Clock::time_point t_submit;
    uint64_t rseq = 0;  // resident-ring sequence
    bool res = false;   // waiting for the resident kernel
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    float *h_L = nullptr, *h_C = nullptr, *h_out = nullptr;
    int32_t* h_ax = nullptr;
    uint64_t* h_q = nullptr;
    float *d_L = nullptr, *d_C = nullptr, *d_out = nullptr;
    int32_t* d_ax = nullptr;
    uint64_t* d_q = nullptr;
    drw_solver* solvers[2 * kRoutes] = {};
    std::vector<Req*> batch;
    size_t bytes = 0;
    int cls = -1, launch_rc = 0;
    bool busy = false;
};

int route_of(int N, bool nn)
{
    if (N == 23 && nn) return 0;
    return N == 16 ? 1 : N == 21 ? 2 : 3;
}

int route_N(int r) { return r == 1 ? 16 : r == 2 ? 21 : 23; }

// Device bytes one query of a route occupies while its batch is in flight.
size_t route_bytes(int r)
{
    const size_t io = sizeof(float) * (kIn + 5) + 16;
    if (r == 0) return io + sizeof(float) * (512 + 8);  // selector features + probabilities
    const size_t n = (size_t)route_N(r);
    return io + n * n * n * sizeof(float) + 6 * n * n * sizeof(double) + 64;
}

}  // namespace

struct drw_server {
    int device = 0;
    drw_server_config cfg{};
    std::string models;
    bool nn = false;
    std::vector<Slot> slots;
    std::unique_ptr<Req[]> ring{new Req[kRing]};
    std::atomic<uint64_t> next_ticket{1};
    // class queues, admission state, shutdown
    std::mutex mu;
    std::condition_variable cv_work;
    std::deque<Req*> queue[2 * kRoutes];
    size_t bytes_inflight = 0;
    bool stop = false;
    // completions
    std::mutex mu_done;
    std::condition_variable cv_done;
    // statistics
    std::mutex mu_stats;
    std::vector<float> lat_ms;
    uint64_t completed = 0, batches = 0, batched = 0, inflight_peak = 0;
    Clock::time_point t0 = Clock::now();
    std::thread worker;
    // resident loop (cfg.resident)
    bool resident = false;
    std::atomic<bool> resident_ok{true};  // false after a failed launch: network queries take the batches
    drw::NetWeights* rw[2] = {nullptr, nullptr};
    void* rnets = nullptr;          // device views of both model sets
    drw::ResReq* rreq = nullptr;    // pinned request ring (host view)
    drw::ResResp* rresp = nullptr;  // pinned response ring (host view)
    uint32_t* rstop = nullptr;      // pinned stop flag
    drw::ResCtl* rctl = nullptr;    // device control block
    drw::ResidentArgs rargs{};
    hipStream_t rstream = nullptr;
    hipEvent_t rdone = nullptr;
    int rgrid = 0;
    std::mutex rmu;                       // sequence allocation and publication
    uint64_t rnext = 0;                   // next sequence (under rmu)
    std::atomic<uint64_t> rpub{0};        // sequences published
    std::unique_ptr<std::atomic<uint64_t>[]> rfree;  // per slot: the sequence that may take it next
    std::unique_ptr<std::atomic<Req*>[]> rowner;  // per slot: the request of its current sequence
    uint64_t rscan = 0;                   // dispatcher: lowest sequence not yet reaped
    std::vector<Req*> rdone_buf;          // dispatcher: requests completed in one reaper pass
    std::atomic<bool> rrunning{false};    // resident kernel launched and not seen to end (dispatcher-owned)
    std::atomic<uint64_t> rlaunches{0};
    std::atomic<int64_t> rfdm_last{INT64_MIN / 2};  // steady-clock ns of the last FDM-route submit
    int rate_khz = 0;                     // device wall clock
    Clock::duration ridle{};              // the dispatcher stops the kernel after this long idle
    Clock::time_point rlast{};            // last resident submit or completion

    int slot_solver(Slot& s, int cls, drw_solver** out);
    int res_create();
    void res_destroy();
    int res_launch(bool drain);
    void res_abandon();
    void res_manage(int busy, int rout);
    int res_reap();
    int launch(Slot& s, int cls, std::vector<Req*>& batch);
    void reap(Slot& s);
    void loop();
    int submit(int n, const int32_t* kinds, const int32_t* Ns, const int32_t* axes, const float* layers,
               const float* cuboids, const uint64_t* qids, uint64_t* tickets, const Clock::time_point* t_arrival);
    Req* lookup(uint64_t ticket);
    int take(uint64_t ticket, float* out, int timeout_ms);
    int try_take(uint64_t ticket, float* out);
};

int drw_server::slot_solver(Slot& s, int cls, drw_solver** out)
{
    if (!s.solvers[cls]) {
        const int kind = cls / kRoutes, r = cls % kRoutes;
        drw_solver* h = nullptr;
        int rc = drw_solver_create(kind, device, &h);
        if (rc) return rc;
        rc = r == 0 ? drw_solver_load_models(h, models.c_str()) : drw_solver_use_fdm(h, route_N(r), cfg.fdm_tol);
        if (!rc) rc = drw_solver_set_seed(h, cfg.seed, 0);
        if (rc) {
            drw_solver_destroy(h);
            return rc;
        }
        s.solvers[cls] = h;
    }
    *out = s.solvers[cls];
    return 0;
}

int drw_server::launch(Slot& s, int cls, std::vector<Req*>& batch)
{
    const int B = (int)batch.size();
    const int kind = cls / kRoutes;
    for (int i = 0; i < B; i++) {
        const Req* q = batch[i];
        std::memcpy(s.h_L + (size_t)i * 30, q->in, sizeof(float) * 30);
        std::memcpy(s.h_C + (size_t)i * 105, q->in + 30, sizeof(float) * 105);
        s.h_ax[i] = q->axis;
        s.h_q[i] = q->qid;
    }
    drw_solver* sol = nullptr;
    int rc = slot_solver(s, cls, &sol);
    if (rc) return rc;
    DRW_HIP(hipMemcpyAsync(s.d_L, s.h_L, sizeof(float) * B * 30, hipMemcpyHostToDevice, s.stream));
    DRW_HIP(hipMemcpyAsync(s.d_C, s.h_C, sizeof(float) * B * 105, hipMemcpyHostToDevice, s.stream));
    DRW_HIP(hipMemcpyAsync(s.d_ax, s.h_ax, sizeof(int32_t) * B, hipMemcpyHostToDevice, s.stream));
    DRW_HIP(hipMemcpyAsync(s.d_q, s.h_q, sizeof(uint64_t) * B, hipMemcpyHostToDevice, s.stream));
    rc = drw_solver_sample_ids(sol, B, s.d_L, s.d_C, kind == DRW_GRADIENT ? s.d_ax : nullptr, s.d_q, s.d_out, s.stream);
    if (rc) return rc;
    const int w = kind == DRW_GRADIENT ? 5 : 3;
    DRW_HIP(hipMemcpyAsync(s.h_out, s.d_out, sizeof(float) * B * w, hipMemcpyDeviceToHost, s.stream));
    DRW_HIP(hipEventRecord(s.done, s.stream));
    return 0;
}

void drw_server::reap(Slot& s)
{
    const int w = s.cls / kRoutes == DRW_GRADIENT ? 5 : 3;
    const auto now = Clock::now();
    const size_t n = s.batch.size();
    for (size_t i = 0; i < n; i++) {
        Req* q = s.batch[i];
        if (s.launch_rc == 0) std::memcpy(q->out, s.h_out + i * w, sizeof(float) * w);
        q->rc = s.launch_rc;
    }
    {
        std::lock_guard<std::mutex> lk(mu_stats);
        for (size_t i = 0; i < n && lat_ms.size() < (1u << 22); i++)
            lat_ms.push_back(std::chrono::duration<float, std::milli>(now - s.batch[i]->t_submit).count());
        completed += n;
        batches += 1;
        batched += n;
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        bytes_inflight -= s.bytes;
    }
    {
        // the waiter frees an entry once it sees kDone: nothing here touches
        // a request after its state store
        std::lock_guard<std::mutex> lk(mu_done);
        for (size_t i = 0; i < n; i++) s.batch[i]->state.store(kDone, std::memory_order_release);
    }
    cv_done.notify_all();
    s.batch.clear();
    s.busy = false;
}

// Resident loop: pinned rings, device control block, weights of both kinds.
int drw_server::res_create()
{
    int rc = drw::net_weights_load(models.c_str(), 0, &rw[0]);
    if (!rc) rc = drw::net_weights_load(models.c_str(), 1, &rw[1]);
    if (rc) return rc;
    const int threads = cfg.resident_threads ? cfg.resident_threads : 512;
    if ((rc = drw::resident_grid(device, threads, &rgrid))) return rc;
    if ((rc = drw::resident_nets_create(rw[0], rw[1], &rnets))) return rc;
    DRW_HIP(hipHostMalloc((void**)&rreq, sizeof(drw::ResReq) * drw::kResRing, hipHostMallocMapped | hipHostMallocCoherent));
    DRW_HIP(hipHostMalloc((void**)&rresp, sizeof(drw::ResResp) * drw::kResRing,
                          hipHostMallocMapped | hipHostMallocCoherent));
    DRW_HIP(hipHostMalloc((void**)&rstop, 2 * sizeof(uint32_t), hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(rreq, 0, sizeof(drw::ResReq) * drw::kResRing);
    std::memset(rresp, 0, sizeof(drw::ResResp) * drw::kResRing);
    rstop[0] = rstop[1] = 0;
    rfree.reset(new std::atomic<uint64_t>[drw::kResRing]);
    for (uint32_t i = 0; i < drw::kResRing; i++) rfree[i].store(i, std::memory_order_relaxed);
    rowner.reset(new std::atomic<Req*>[drw::kResRing]);
    for (uint32_t i = 0; i < drw::kResRing; i++) rowner[i].store(nullptr, std::memory_order_relaxed);
    DRW_HIP(hipStreamCreateWithFlags(&rstream, hipStreamNonBlocking));
    DRW_HIP(hipEventCreateWithFlags(&rdone, hipEventDisableTiming));
    // fine-grained: the kernel's plain atomic loads of it are coherent across XCDs
    DRW_HIP(hipExtMallocWithFlags((void**)&rctl, sizeof(drw::ResCtl), hipDeviceMallocFinegrained));
    const drw::ResCtl zero{};
    DRW_HIP(hipMemcpyAsync(rctl, &zero, sizeof zero, hipMemcpyHostToDevice, rstream));
    DRW_HIP(hipStreamSynchronize(rstream));
    DRW_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, device));
    if (rate_khz <= 0) return drw::set_error(DRW_EHIP, "server_create: no device wall clock");
    rargs.ctl = rctl;
    rargs.seed = cfg.seed;
    // the dispatcher stops an idle kernel after resident_idle_ms; the kernel's
    // own limit is only a safety net (a host that stopped polling)
    const double idle_ms = cfg.resident_idle_ms > 0 ? cfg.resident_idle_ms : 20.0;
    ridle = std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double, std::milli>(idle_ms));
    rargs.idle_ticks = (uint64_t)(std::max(1000.0, 100.0 * idle_ms) * rate_khz);
    rargs.scout_gap = (uint64_t)std::max(1, rate_khz / 1000);  // 1 us
    DRW_HIP(hipHostGetDevicePointer((void**)&rargs.req, rreq, 0));
    DRW_HIP(hipHostGetDevicePointer((void**)&rargs.resp, rresp, 0));
    DRW_HIP(hipHostGetDevicePointer((void**)&rargs.host_stop, rstop, 0));
    return 0;
}

void drw_server::res_destroy()
{
    if (rstop) __atomic_store_n(rstop, 1u, __ATOMIC_RELEASE);
    if (rstream) (void)hipStreamSynchronize(rstream);
    if (rdone) (void)hipEventDestroy(rdone);
    if (rstream) (void)hipStreamDestroy(rstream);
    if (rctl) (void)hipFree(rctl);
    if (rnets) (void)hipFree(rnets);
    if (rreq) (void)hipHostFree(rreq);
    if (rresp) (void)hipHostFree(rresp);
    if (rstop) (void)hipHostFree(rstop);
    for (auto*& w : rw) {
        drw::net_weights_free(w);
        w = nullptr;
    }
}

// (Re)starts the resident kernel: clears the stop flags and rewinds the
// ticket head to the oldest unanswered sequence (tickets held by
// workgroups of the last launch when it stopped are served again), then
// launches.
int drw_server::res_launch(bool drain)
{
    // a drain launch starts with the stop flag set: its workgroups serve
    // what is published and leave
    __atomic_store_n(rstop, drain ? 1u : 0u, __ATOMIC_RELEASE);
    const unsigned long long hs[2] = {rscan, drain ? 1ull : 0ull};  // head, stop
    DRW_HIP(hipMemcpyAsync(&rctl->head, &hs[0], sizeof hs[0], hipMemcpyHostToDevice, rstream));
    DRW_HIP(hipMemcpyAsync(&rctl->stop, &hs[1], sizeof hs[1], hipMemcpyHostToDevice, rstream));
    DRW_HIP(hipStreamSynchronize(rstream));  // the copies have read hs
    const int threads = cfg.resident_threads ? cfg.resident_threads : 512;
    const int rc = drw::resident_launch(rargs, rnets, threads, rgrid, rstream);
    if (rc) return rc;
    DRW_HIP(hipEventRecord(rdone, rstream));
    rrunning.store(true, std::memory_order_release);
    rlaunches++;
    return 0;
}

// After a failed launch (no kernel runs): the ring's unanswered queries go
// to the batched network route (same results: the counters are per query),
// and so do later ones.
void drw_server::res_abandon()
{
    resident_ok.store(false, std::memory_order_release);
    const uint64_t end = rpub.load(std::memory_order_acquire);
    std::lock_guard<std::mutex> lk(mu);
    for (uint64_t sq = rscan; sq < end; sq++) {
        const uint32_t slot = (uint32_t)(sq & (drw::kResRing - 1));
        if (rfree[slot].load(std::memory_order_relaxed) != sq) continue;
        Req* q = rowner[slot].load(std::memory_order_acquire);
        rfree[slot].store(sq + drw::kResRing, std::memory_order_release);
        q->res = false;
        queue[q->cls].push_back(q);
    }
    rscan = end;
}

// Dispatcher step for the resident kernel: notice its end, stop it while
// batches are pending or in flight (its workgroups hold every CU), and
// start it when published queries are waiting: as a resident loop when no
// batch needs the device, else as a drain launch that serves what is
// published and ends.
void drw_server::res_manage(int busy, int rout)
{
    const auto now = Clock::now();
    if (rout > 0) rlast = now;
    if (rrunning.load(std::memory_order_relaxed) && hipEventQuery(rdone) == hipSuccess)
        rrunning.store(false, std::memory_order_release);
    bool fdm = busy > 0, closing;
    {
        std::lock_guard<std::mutex> lk(mu);
        for (auto& qd : queue) fdm |= !qd.empty();
        closing = stop;
    }
    if (rrunning.load(std::memory_order_relaxed)) {
        // stop it for FDM batches, or when nothing has come for ridle (at
        // once when the server is closing)
        if (fdm || (rout == 0 && (closing || now - rlast > ridle))) __atomic_store_n(rstop, 1u, __ATOMIC_RELEASE);
        return;
    }
    if (rpub.load(std::memory_order_acquire) > rscan) {
        rlast = now;
        if (res_launch(fdm)) res_abandon();
    }
}

void drw_server::loop()
{
    (void)hipSetDevice(device);
    for (;;) {
        // 1) completions
        bool progressed = false;
        int busy = 0;
        for (auto& s : slots) {
            if (!s.busy) continue;
            const hipError_t e = s.launch_rc ? hipSuccess : hipEventQuery(s.done);
            if (e == hipSuccess) {
                reap(s);
                progressed = true;
            } else if (e != hipErrorNotReady) {
                s.launch_rc = DRW_EHIP;
                reap(s);
                progressed = true;
            } else {
                busy++;
            }
        }
        int rout = 0;  // resident queries published and not yet reaped
        if (resident) {
            rout = res_reap();
            res_manage(busy, rout);
        }
        // 2) admissions: the class of the oldest pending request, limited by
        //    free slots, the batch cap and the HBM budget
        for (auto& s : slots) {
            if (s.busy) continue;
            int cls = -1;
            {
                std::unique_lock<std::mutex> lk(mu);
                Clock::time_point oldest = Clock::time_point::max();
                for (int c = 0; c < 2 * kRoutes; c++)
                    if (!queue[c].empty() && queue[c].front()->t_submit < oldest) {
                        oldest = queue[c].front()->t_submit;
                        cls = c;
                    }
                if (cls < 0) break;
                const size_t per = route_bytes(cls % kRoutes);
                const size_t budget = (size_t)(cfg.hbm_budget_mb * 1048576.0);
                size_t room = budget > bytes_inflight ? (budget - bytes_inflight) / per : 0;
                if (bytes_inflight == 0 && room == 0) room = 1;  // always make progress
                const size_t n = std::min<size_t>({queue[cls].size(), (size_t)cfg.max_batch, room});
                if (n == 0) break;  // wait for a batch to finish and free memory
                s.batch.assign(queue[cls].begin(), queue[cls].begin() + (ptrdiff_t)n);
                queue[cls].erase(queue[cls].begin(), queue[cls].begin() + (ptrdiff_t)n);
                s.bytes = n * per;
                bytes_inflight += s.bytes;
            }
            s.cls = cls;
            s.busy = true;
            s.launch_rc = launch(s, cls, s.batch);
            busy++;
            progressed = true;
        }
        {
            std::lock_guard<std::mutex> lk(mu_stats);
            inflight_peak = std::max<uint64_t>(inflight_peak, (uint64_t)busy);
        }
        if (progressed) continue;
        if (busy > 0 || rout > 0) {
            // batches or resident queries in flight: keep polling their
            // events and the response ring (the dispatcher owns a core while
            // work is on the device; completion latency is the serving
            // latency)
            std::this_thread::yield();
            continue;
        }
        std::unique_lock<std::mutex> lk(mu);
        bool pending = false;
        for (auto& qd : queue) pending |= !qd.empty();
        if (resident) {
            // published resident queries not yet claimed when the kernel ended
            const bool rwait = !rrunning.load(std::memory_order_relaxed) && rpub.load(std::memory_order_acquire) > rscan;
            if (stop && !pending && !rwait && !rrunning.load(std::memory_order_relaxed)) return;
            if (pending || rwait) continue;
            // the kernel runs (or is idle-ended, seen at the next wake-up):
            // check on it every millisecond
            cv_work.wait_for(lk, std::chrono::milliseconds(1), [&] {
                for (auto& qd : queue)
                    if (!qd.empty()) return true;
                return rpub.load(std::memory_order_acquire) > rscan;
            });
            continue;
        }
        if (stop && !pending) return;
        cv_work.wait(lk, [&] {
            if (stop) return true;
            for (auto& qd : queue)
                if (!qd.empty()) return true;
            return false;
        });
    }
}

int drw_server::submit(int n, const int32_t* kinds, const int32_t* Ns, const int32_t* axes, const float* layers,
                       const float* cuboids, const uint64_t* qids, uint64_t* tickets,
                       const Clock::time_point* t_arrival)
{
    for (int i = 0; i < n; i++) {
        if (kinds[i] != DRW_POISSON && kinds[i] != DRW_GRADIENT)
            return drw::set_error(DRW_EINVAL, "server_submit: bad kind (query %d)", i);
        if (Ns[i] != 16 && Ns[i] != 21 && Ns[i] != 23)
            return drw::set_error(DRW_EINVAL, "server_submit: N must be 16, 21 or 23 (query %d)", i);
        if (kinds[i] == DRW_GRADIENT && (!axes || axes[i] < 0 || axes[i] > 2))
            return drw::set_error(DRW_EINVAL, "Gradient solver requires axis parameter");
    }
    std::vector<Req*> mine((size_t)n);
    const uint64_t t0 = next_ticket.fetch_add((uint64_t)n);
    const auto now = Clock::now();
    for (int i = 0; i < n; i++) {
        const uint64_t t = t0 + (uint64_t)i;
        Req& q = ring[t & (kRing - 1)];
        int expect = kFree;
        if (!q.state.compare_exchange_strong(expect, kClaimed, std::memory_order_acquire)) {
            for (int j = 0; j < i; j++) {
                mine[j]->ticket.store(0, std::memory_order_relaxed);
                mine[j]->state.store(kFree, std::memory_order_release);
            }
            return drw::set_error(DRW_ESTATE, "server_submit: more than %llu queries outstanding (wait for tickets)",
                                  (unsigned long long)kRing);
        }
        q.ticket.store(t, std::memory_order_relaxed);
        q.kind = kinds[i];
        q.N = Ns[i];
        q.axis = kinds[i] == DRW_GRADIENT ? axes[i] : 2;
        q.cls = kinds[i] * kRoutes + route_of(Ns[i], nn);
        q.qid = qids && qids[i] ? qids[i] : t;
        q.rc = 0;
        std::memcpy(q.in, layers + (size_t)i * 30, sizeof(float) * 30);
        std::memcpy(q.in + 30, cuboids + (size_t)i * 105, sizeof(float) * 105);
        q.t_submit = t_arrival ? t_arrival[i] : now;
        mine[i] = &q;
    }
    auto release = [&mine] {
        for (auto* q : mine) {
            q->ticket.store(0, std::memory_order_relaxed);
            q->state.store(kFree, std::memory_order_release);
        }
    };
    // While FDM-route queries are about (submitted within the last
    // kFdmHold), network queries take the batches too: the resident kernel
    // holds every CU, so it is stopped while batches run.  Results are the
    // same either way (per-query counters).
    bool batch_nn = resident && !resident_ok.load(std::memory_order_acquire);
    if (resident && !batch_nn) {
        const int64_t t_now = std::chrono::duration_cast<std::chrono::nanoseconds>(now.time_since_epoch()).count();
        bool has_fdm = false;
        for (auto* q : mine) has_fdm |= q->cls % kRoutes != 0;
        if (has_fdm) rfdm_last.store(t_now, std::memory_order_relaxed);
        batch_nn = has_fdm || t_now - rfdm_last.load(std::memory_order_relaxed) < kFdmHoldNs;
    }
    int nres = 0;
    for (auto* q : mine) nres += resident && !batch_nn && q->cls % kRoutes == 0;
    {
        std::lock_guard<std::mutex> lk(mu);
        if (stop) {
            release();
            return drw::set_error(DRW_ESTATE, "server is shutting down");
        }
    }
    if (nres) {
        // resident queries: sequences and publication, all or none (the
        // kernel claims sequences in order, so none may stay unpublished);
        // each request is kQueued before its seq word makes it visible
        std::lock_guard<std::mutex> lr(rmu);
        for (int j = 0; j < nres; j++)
            if (rfree[(rnext + (uint64_t)j) & (drw::kResRing - 1)].load(std::memory_order_acquire) !=
                rnext + (uint64_t)j) {
                release();
                return drw::set_error(DRW_ESTATE, "server_submit: resident ring full (wait for tickets)");
            }
        for (auto* q : mine) {
            if (q->cls % kRoutes != 0) continue;  // nres > 0: every network query goes to the ring
            const uint64_t sq = rnext++;
            const uint32_t slot = (uint32_t)(sq & (drw::kResRing - 1));
            q->rseq = sq;
            q->res = true;
            q->state.store(kQueued, std::memory_order_relaxed);
            rowner[slot].store(q, std::memory_order_release);
            drw::ResReq& r = rreq[slot];
            std::memcpy(r.layers, q->in, sizeof(float) * 30);
            std::memcpy(r.cuboids, q->in + 30, sizeof(float) * 105);
            r.kind = q->kind;
            r.axis = q->axis;
            r.qid = q->qid;
            __atomic_store_n(&r.seq, sq + 1, __ATOMIC_RELEASE);
        }
        rpub.store(rnext, std::memory_order_release);
    }
    if (nres < n) {
        std::lock_guard<std::mutex> lk(mu);
        for (auto* q : mine) {
            if (q->res) continue;
            q->state.store(kQueued, std::memory_order_relaxed);
            queue[q->cls].push_back(q);
        }
    }
    cv_work.notify_one();
    for (int i = 0; i < n; i++) tickets[i] = t0 + (uint64_t)i;
    return 0;
}

Req* drw_server::lookup(uint64_t ticket)
{
    if (ticket == 0) return nullptr;
    Req* q = &ring[ticket & (kRing - 1)];
    if (q->ticket.load(std::memory_order_acquire) != ticket) return nullptr;
    const int st = q->state.load(std::memory_order_acquire);
    return st == kQueued || st == kDone ? q : nullptr;
}

// Result of a finished request; frees its entry.
static int consume(Req* q, float* out)
{
    std::memcpy(out, q->out, sizeof(float) * (q->kind == DRW_GRADIENT ? 5 : 3));
    const int rc = q->rc;
    q->ticket.store(0, std::memory_order_relaxed);
    q->state.store(kFree, std::memory_order_release);
    return rc ? drw::set_error(rc, "server: batch failed") : 0;
}

// Dispatcher: collects the resident kernel's responses.  Sequences in
// [rscan, published) are checked in order of age; a response whose six words
// carry its tag completes its request as a batch completion does (result,
// latency, kDone, waiters woken) and frees its ring slot; a NaN result
// (status 1) is queued on the FDM route at N = 23 with the same counter,
// which is what the batched network route's fallback computes for it.
// Returns the number of requests still waiting for the kernel.
int drw_server::res_reap()
{
    const uint64_t end = rpub.load(std::memory_order_acquire);
    uint64_t first_open = end;
    std::vector<Req*>& done = rdone_buf;
    done.clear();
    bool rerouted = false;
    for (uint64_t sq = rscan; sq < end; sq++) {
        const uint32_t slot = (uint32_t)(sq & (drw::kResRing - 1));
        if (rfree[slot].load(std::memory_order_relaxed) != sq) continue;  // reaped already
        const drw::ResResp& r = rresp[slot];
        const uint32_t tag = (uint32_t)(sq + 1);
        uint64_t w[drw::kResWords];
        bool ready = true;
        for (int j = 0; j < drw::kResWords && ready; j++) {
            w[j] = __atomic_load_n(&r.w[j], __ATOMIC_ACQUIRE);
            ready = (uint32_t)(w[j] >> 32) == tag;
        }
        if (!ready) {
            first_open = std::min(first_open, sq);
            continue;
        }
        Req* q = rowner[slot].load(std::memory_order_acquire);
        for (int j = 0; j < 5; j++) {
            const uint32_t v = (uint32_t)w[j];
            std::memcpy(&q->out[j], &v, sizeof v);
        }
        q->res = false;
        rfree[slot].store(sq + drw::kResRing, std::memory_order_release);
        if ((uint32_t)w[5] != 0) {
            std::lock_guard<std::mutex> lk(mu);
            q->cls = q->kind * kRoutes + 3;
            queue[q->cls].push_back(q);
            rerouted = true;
            continue;
        }
        q->rc = 0;
        done.push_back(q);
    }
    rscan = first_open;
    if (!done.empty()) {
        const auto now = Clock::now();
        {
            std::lock_guard<std::mutex> lk(mu_stats);
            for (Req* q : done)
                if (lat_ms.size() < (1u << 22))
                    lat_ms.push_back(std::chrono::duration<float, std::milli>(now - q->t_submit).count());
            completed += done.size();
        }
        {
            // the waiter frees an entry once it sees kDone: nothing here
            // touches a request after its state store
            std::lock_guard<std::mutex> lk(mu_done);
            for (Req* q : done) q->state.store(kDone, std::memory_order_release);
        }
        cv_done.notify_all();
    }
    (void)rerouted;
    return (int)std::min<uint64_t>(end - rscan, 1u << 30);
}

int drw_server::take(uint64_t ticket, float* out, int timeout_ms)
{
    Req* q = lookup(ticket);
    if (!q) return drw::set_error(DRW_EINVAL, "server_wait: unknown ticket %llu", (unsigned long long)ticket);
    auto done = [q] { return q->state.load(std::memory_order_acquire) == kDone; };
    for (int i = 0; i < 64 && !done(); i++) std::this_thread::yield();
    if (!done()) {
        std::unique_lock<std::mutex> lk(mu_done);
        if (timeout_ms < 0) cv_done.wait(lk, done);
        else if (!cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done))
            return drw::set_error(DRW_ETIMEOUT, "server_wait: timed out");
    }
    return consume(q, out);
}

int drw_server::try_take(uint64_t ticket, float* out)
{
    Req* q = lookup(ticket);
    if (!q) return drw::set_error(DRW_EINVAL, "server_wait: unknown ticket %llu", (unsigned long long)ticket);
    if (q->state.load(std::memory_order_acquire) != kDone) return 1;
    return consume(q, out) ? -1 : 0;
}

extern "C" {

int drw_server_create(int device, const char* models_txt, const drw_server_config* cfg, drw_server** out)
{
    drw::clear_error();
    if (!out) return drw::set_error(DRW_EINVAL, "server_create: out is null");
    *out = nullptr;
    drw_server_config c{};
    c.max_batch = 2048;
    c.max_inflight = 4;
    c.hbm_budget_mb = 1024.0;
    c.fdm_tol = 1e-6;
    c.seed = 0x5eed5eedULL;
    if (cfg) c = *cfg;
    if (c.max_batch < 1 || c.max_batch > 65536 || c.max_inflight < 1 || c.max_inflight > 16 || !(c.hbm_budget_mb > 0) ||
        !(c.fdm_tol > 0) || (c.resident != 0 && c.resident != 1) ||
        (c.resident_threads != 0 && c.resident_threads != 256 && c.resident_threads != 512) ||
        !(c.resident_idle_ms >= 0))
        return drw::set_error(DRW_EINVAL, "server_create: bad configuration");
    if (c.resident && !(models_txt && models_txt[0]))
        return drw::set_error(DRW_EINVAL, "server_create: the resident loop serves the networks: give a model list");
    int ndev = 0;
    DRW_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return drw::set_error(DRW_EINVAL, "server_create: bad device %d", device);
    DRW_HIP(hipSetDevice(device));
    auto* sv = new drw_server();
    sv->device = device;
    sv->cfg = c;
    if (models_txt && models_txt[0]) {
        sv->models = models_txt;
        sv->nn = true;
    }
    sv->slots.resize(c.max_inflight);
    const size_t B = (size_t)c.max_batch;
    for (auto& s : sv->slots) {
        hipError_t e = hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
        if (e == hipSuccess) e = hipHostMalloc((void**)&s.h_L, B * 30 * sizeof(float), hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void**)&s.h_C, B * 105 * sizeof(float), hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void**)&s.h_out, B * 5 * sizeof(float), hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void**)&s.h_ax, B * sizeof(int32_t), hipHostMallocDefault);
        if (e == hipSuccess) e = hipHostMalloc((void**)&s.h_q, B * sizeof(uint64_t), hipHostMallocDefault);
        if (e == hipSuccess) e = hipMalloc((void**)&s.d_L, B * 30 * sizeof(float));
        if (e == hipSuccess) e = hipMalloc((void**)&s.d_C, B * 105 * sizeof(float));
        if (e == hipSuccess) e = hipMalloc((void**)&s.d_out, B * 5 * sizeof(float));
        if (e == hipSuccess) e = hipMalloc((void**)&s.d_ax, B * sizeof(int32_t));
        if (e == hipSuccess) e = hipMalloc((void**)&s.d_q, B * sizeof(uint64_t));
        if (e != hipSuccess) {
            drw_server_destroy(sv);
            return drw::set_error(DRW_EHIP, "server_create: %s", hipGetErrorString(e));
        }
    }
    // Create the per-slot solvers of the network route up front so that a
    // bad model list fails here rather than on the first request.
    if (sv->nn) {
        for (auto& s : sv->slots)
            for (int k = 0; k < 2; k++) {
                drw_solver* h = nullptr;
                const int rc = sv->slot_solver(s, k * kRoutes, &h);
                if (rc) {
                    const std::string msg = drw::last_error();
                    drw_server_destroy(sv);
                    return drw::set_error(rc, "%s", msg.c_str());
                }
            }
    }
    if (c.resident) {
        sv->resident = true;
        const int rc = sv->res_create();
        if (rc) {
            const std::string msg = drw::last_error();
            drw_server_destroy(sv);
            return drw::set_error(rc, "%s", msg.c_str());
        }
    }
    sv->t0 = Clock::now();
    sv->worker = std::thread([sv] { sv->loop(); });
    *out = sv;
    return 0;
}

int drw_server_submit(drw_server* sv, int kind, int N, int axis, const float* layer, const float* cuboid, uint64_t qid,
                      uint64_t* ticket)
{
    drw::clear_error();
    if (!sv || !layer || !cuboid || !ticket) return drw::set_error(DRW_EINVAL, "server_submit: null argument");
    const int32_t k = kind, n = N, a = axis;
    return sv->submit(1, &k, &n, &a, layer, cuboid, &qid, ticket, nullptr);
}

int drw_server_submit_many(drw_server* sv, int n, const int32_t* kinds, const int32_t* Ns, const int32_t* axes,
                           const float* layers, const float* cuboids, const uint64_t* qids, uint64_t* tickets)
{
    drw::clear_error();
    if (!sv || n < 0 || (n > 0 && (!kinds || !Ns || !layers || !cuboids || !tickets)))
        return drw::set_error(DRW_EINVAL, "server_submit_many: bad argument");
    if ((uint64_t)n > kRing) return drw::set_error(DRW_EINVAL, "server_submit_many: n above %llu", (unsigned long long)kRing);
    if (n == 0) return 0;
    return sv->submit(n, kinds, Ns, axes, layers, cuboids, qids, tickets, nullptr);
}

int drw_server_wait(drw_server* sv, uint64_t ticket, float* out, int timeout_ms)
{
    drw::clear_error();
    if (!sv || !out) return drw::set_error(DRW_EINVAL, "server_wait: null argument");
    return sv->take(ticket, out, timeout_ms);
}

int drw_server_wait_many(drw_server* sv, int n, const uint64_t* tickets, float* outs, int timeout_ms)
{
    drw::clear_error();
    if (!sv || n < 0 || (n > 0 && (!tickets || !outs))) return drw::set_error(DRW_EINVAL, "server_wait_many: bad argument");
    const auto deadline = Clock::now() + std::chrono::milliseconds(timeout_ms < 0 ? 0 : timeout_ms);
    int first = 0;
    for (int i = 0; i < n; i++) {
        int left = -1;
        if (timeout_ms >= 0) {
            const auto ms = std::chrono::duration_cast<std::chrono::milliseconds>(deadline - Clock::now()).count();
            left = (int)std::max<long long>(0, ms);
        }
        const int rc = sv->take(tickets[i], outs + (size_t)i * 5, left);
        if (rc == DRW_ETIMEOUT) return rc;  // this ticket and the later ones stay valid
        if (rc && !first) first = rc;
    }
    return first ? drw::set_error(first, "server_wait_many: %s", drw::last_error()) : 0;
}

int drw_server_resident_stats(drw_server* sv, uint64_t* out)
{
    drw::clear_error();
    if (!sv || !out) return drw::set_error(DRW_EINVAL, "server_resident_stats: null argument");
    std::memset(out, 0, 10 * sizeof(uint64_t));
    if (!sv->resident) return 0;
    (void)hipSetDevice(sv->device);
    drw::ResCtl c{};
    hipStream_t st = nullptr;
    DRW_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    hipError_t e = hipMemcpyAsync(&c, sv->rctl, sizeof c, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    if (e != hipSuccess) return drw::set_error(DRW_EHIP, "server_resident_stats: %s", hipGetErrorString(e));
    const uint64_t v[10] = {sv->rlaunches.load(), c.served[0], c.served[1], c.scouts, c.exits_idle, c.exits_stop,
                            c.head, c.published, (uint64_t)sv->rgrid, (uint64_t)sv->rate_khz};
    std::memcpy(out, v, sizeof v);
    return 0;
}

int drw_server_stats_get(drw_server* sv, drw_server_stats* st, int reset)
{
    if (!sv || !st) return drw::set_error(DRW_EINVAL, "server_stats: null argument");
    std::lock_guard<std::mutex> lk(sv->mu_stats);
    std::memset(st, 0, sizeof *st);
    st->completed = sv->completed;
    st->batches = sv->batches;
    st->mean_batch = sv->batches ? (double)sv->batched / (double)sv->batches : 0.0;
    st->inflight_peak = sv->inflight_peak;
    st->elapsed_s = std::chrono::duration<double>(Clock::now() - sv->t0).count();
    st->qps = st->elapsed_s > 0 ? (double)sv->completed / st->elapsed_s : 0.0;
    if (!sv->lat_ms.empty()) {
        std::vector<float> v(sv->lat_ms);
        auto pct = [&](double p) {
            const size_t k = std::min(v.size() - 1, (size_t)(p * (double)(v.size() - 1) + 0.5));
            std::nth_element(v.begin(), v.begin() + k, v.end());
            return (double)v[k];
        };
        st->p50_ms = pct(0.50);
        st->p99_ms = pct(0.99);
        st->max_ms = *std::max_element(sv->lat_ms.begin(), sv->lat_ms.end());
    }
    if (reset) {
        sv->lat_ms.clear();
        sv->completed = sv->batches = sv->batched = sv->inflight_peak = 0;
        sv->t0 = Clock::now();
    }
    return 0;
}

// Open-loop load: `threads` host threads (the walker threads) each draw
// exponential inter-arrival times (aggregate rate `rate_qps`; <= 0 means
// back-to-back), submit every query that is due in one call, with its
// arrival time as the start of its latency, collect finished tickets oldest
// first, and block only when `max_outstanding` tickets are in flight.
int drw_server_run_load(drw_server* sv, const drw_load_config* lc, const float* layers, const float* cuboids,
                        const int32_t* kinds, const int32_t* Ns, const int32_t* axes, int pool, drw_server_stats* st)
{
    drw::clear_error();
    if (!sv || !lc || !layers || !cuboids || !kinds || !Ns || !axes || pool <= 0 || !st)
        return drw::set_error(DRW_EINVAL, "server_run_load: bad argument");
    if (lc->threads < 1 || lc->threads > 1024 || !(lc->duration_s > 0) || lc->max_outstanding < 1 ||
        (uint64_t)lc->threads * (uint64_t)lc->max_outstanding > kRing / 2)
        return drw::set_error(DRW_EINVAL, "server_run_load: bad load configuration");
    drw_server_stats dummy;
    drw_server_stats_get(sv, &dummy, 1);
    std::atomic<int> failed{0};
    std::string first_error;
    std::mutex emu;
    const auto t_end = Clock::now() + std::chrono::duration_cast<Clock::duration>(
                                          std::chrono::duration<double>(lc->duration_s));
    auto worker = [&](int tid) {
        uint64_t rng = 0x9E3779B97F4A7C15ull * (uint64_t)(tid + 1) ^ lc->seed;
        auto next_u = [&rng]() {  // xorshift64*
            rng ^= rng >> 12; rng ^= rng << 25; rng ^= rng >> 27;
            return (double)((rng * 0x2545F4914F6CDD1Dull) >> 11) * (1.0 / 9007199254740992.0);
        };
        auto fail = [&]() {
            failed++;
            std::lock_guard<std::mutex> lk(emu);
            if (first_error.empty()) first_error = drw::last_error();
        };
        const double rate = lc->rate_qps > 0 ? lc->rate_qps / lc->threads : 0.0;
        const size_t maxo = (size_t)lc->max_outstanding;
        std::deque<uint64_t> mine;
        std::vector<int32_t> kk, nn, aa;
        std::vector<float> LL, CC;
        std::vector<Clock::time_point> ta;
        std::vector<uint64_t> tk;
        float o[5];
        auto t_next = Clock::now();
        int i = tid;
        for (;;) {
            const auto now = Clock::now();
            if (now >= t_end) break;
            // 1) every arrival that is due, in one submission
            kk.clear(); nn.clear(); aa.clear(); LL.clear(); CC.clear(); ta.clear();
            while (mine.size() + kk.size() < maxo && kk.size() < 512 && (rate <= 0 || t_next <= now)) {
                const int k = i % pool;
                i += lc->threads;
                kk.push_back(kinds[k]);
                nn.push_back(Ns[k]);
                aa.push_back(axes[k]);
                LL.insert(LL.end(), layers + (size_t)k * 30, layers + (size_t)k * 30 + 30);
                CC.insert(CC.end(), cuboids + (size_t)k * 105, cuboids + (size_t)k * 105 + 105);
                ta.push_back(rate > 0 ? t_next : now);
                if (rate > 0)
                    t_next += std::chrono::duration_cast<Clock::duration>(
                        std::chrono::duration<double>(-std::log(1.0 - next_u()) / rate));
            }
            if (!kk.empty()) {
                tk.resize(kk.size());
                if (sv->submit((int)kk.size(), kk.data(), nn.data(), aa.data(), LL.data(), CC.data(), nullptr,
                               tk.data(), ta.data())) {
                    fail();
                    break;
                }
                mine.insert(mine.end(), tk.begin(), tk.end());
            }
            // 2) finished tickets, oldest first
            while (!mine.empty()) {
                const int r = sv->try_take(mine.front(), o);
                if (r == 1) break;
                if (r < 0) fail();
                mine.pop_front();
            }
            // 3) block at the outstanding limit, else pace to the next arrival
            if (mine.size() >= maxo) {
                if (sv->take(mine.front(), o, -1)) fail();
                mine.pop_front();
            } else if (rate > 0) {
                const auto wake = std::min(t_next, Clock::now() + std::chrono::microseconds(100));
                if (wake - Clock::now() > std::chrono::microseconds(60)) std::this_thread::sleep_until(wake);
                else std::this_thread::yield();
            }
        }
        for (auto t : mine)
            if (sv->take(t, o, -1)) fail();
    };
    std::vector<std::thread> th;
    for (int t = 0; t < lc->threads; t++) th.emplace_back(worker, t);
    for (auto& t : th) t.join();
    drw_server_stats_get(sv, st, 0);
    if (failed)
        return drw::set_error(DRW_ESTATE, "server_run_load: %d submissions or waits failed: %s", (int)failed,
                              first_error.c_str());
    return 0;
}

void drw_server_destroy(drw_server* sv)
{
    if (!sv) return;
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stop = true;
    }
    sv->cv_work.notify_all();
    if (sv->worker.joinable()) sv->worker.join();  // drains the queues first
    (void)hipSetDevice(sv->device);
    sv->res_destroy();
    for (auto& s : sv->slots) {
        if (s.stream) (void)hipStreamSynchronize(s.stream);
        for (auto* h : s.solvers)
            if (h) drw_solver_destroy(h);
        if (s.h_L) (void)hipHostFree(s.h_L);
        if (s.h_C) (void)hipHostFree(s.h_C);
        if (s.h_out) (void)hipHostFree(s.h_out);
        if (s.h_ax) (void)hipHostFree(s.h_ax);
        if (s.h_q) (void)hipHostFree(s.h_q);
        if (s.d_L) (void)hipFree(s.d_L);
        if (s.d_C) (void)hipFree(s.d_C);
        if (s.d_out) (void)hipFree(s.d_out);
        if (s.d_ax) (void)hipFree(s.d_ax);
        if (s.d_q) (void)hipFree(s.d_q);
        if (s.done) (void)hipEventDestroy(s.done);
        if (s.stream) (void)hipStreamDestroy(s.stream);
    }
    delete sv;
}

}  // extern "C"
