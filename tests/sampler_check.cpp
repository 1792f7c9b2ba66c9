// Behaviour of ParallelSampler (inference_cpp/src/parallel_sampler.cpp:42-211)
// against a CPU test double of the solver: every task gets its own rows of
// the batch result, batches never exceed min(batch_size, 2048), a failing
// batch fails every task in it and only those, tasks without an axis run
// with axis 0 when their batch has one, and the destructor drains the queue.
// Built and run by tests/test_parallel_sampler.py; prints one line per case
// and exits non-zero on the first failure.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <future>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "dnn_solver.h"

#define CHECK(c)                                                              \
    do {                                                                      \
        if (!(c)) {                                                           \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
            std::exit(1);                                                     \
        }                                                                     \
    } while (0)

// Row of query q: (layer[0], cuboid[0], batch size, axis or -1).  layer[0] = -1
// in any query makes the batch throw; -2 makes it return a wrong-size result.
class FakeSolver : public DNNSolver {
public:
    FakeSolver() : DNNSolver(NoDevice{}, DRW_POISSON) {}
    void loadModels(bool) override {}
    std::vector<float> sample(const std::vector<float>& l, const std::vector<float>& c) override
    {
        return rows(l, c, nullptr);
    }
    std::vector<float> sample(const std::vector<float>& l, const std::vector<float>& c,
                              const std::vector<int>& axis) override
    {
        return rows(l, c, &axis);
    }
    void hold() { gate_ = std::promise<void>(); held_ = true; }
    void release() { gate_.set_value(); }
    std::atomic<int> calls{0}, max_batch{0}, entered{0};
    int delay_us = 0;

private:
    std::vector<float> rows(const std::vector<float>& l, const std::vector<float>& c, const std::vector<int>* axis)
    {
        entered++;
        if (held_) {
            gate_.get_future().wait();
            held_ = false;
        }
        if (delay_us) std::this_thread::sleep_for(std::chrono::microseconds(delay_us));
        const size_t B = l.size() / (MAX_LAYERS * 2);
        CHECK(l.size() == B * MAX_LAYERS * 2 && c.size() == B * MAX_CUBOIDS * 7);
        CHECK(!axis || axis->size() == B);
        calls++;
        int mb = max_batch.load();
        while ((int)B > mb && !max_batch.compare_exchange_weak(mb, (int)B)) {}
        std::vector<float> out;
        for (size_t q = 0; q < B; q++) {
            const float tag = l[q * MAX_LAYERS * 2];
            if (tag == -1.f) throw std::runtime_error("fake solver failure");
            if (tag == -2.f) return std::vector<float>(B * 4 + 1, 0.f);
            out.insert(out.end(), {tag, c[q * MAX_CUBOIDS * 7], (float)B, axis ? (float)(*axis)[q] : -1.f});
        }
        return out;
    }
    std::promise<void> gate_;
    bool held_ = false;
};

static std::vector<float> layer_of(float tag)
{
    std::vector<float> v(MAX_LAYERS * 2, (float)PADDING_VALUE);
    v[0] = tag;
    return v;
}

static std::vector<float> cuboid_of(float tag)
{
    std::vector<float> v(MAX_CUBOIDS * 7, (float)PADDING_VALUE);
    v[0] = 2.f * tag;
    return v;
}

// Submits one task and waits until the worker sits inside the solver with it,
// so that everything submitted afterwards forms the next batch.
static std::future<std::vector<float>> block_worker(ParallelSampler& ps, FakeSolver& fs, float tag)
{
    fs.hold();
    const int before = fs.entered.load();
    auto f = ps.submit(layer_of(tag), cuboid_of(tag));
    while (fs.entered.load() == before) std::this_thread::sleep_for(std::chrono::microseconds(50));
    return f;
}

static void case_split_concurrent(size_t batch_size, size_t cap)
{
    FakeSolver fs;
    ParallelSampler ps(&fs, 1, batch_size);
    constexpr int T = 8, K = 600;
    std::vector<std::vector<std::future<std::vector<float>>>> futs(T);
    std::vector<std::thread> th;
    for (int t = 0; t < T; t++)
        th.emplace_back([&, t] {
            for (int i = 0; i < K; i++) futs[t].push_back(ps.submit(layer_of(t * 1000 + i), cuboid_of(t * 1000 + i)));
        });
    for (auto& x : th) x.join();
    for (int t = 0; t < T; t++)
        for (int i = 0; i < K; i++) {
            const std::vector<float> r = futs[t][i].get();
            CHECK(r.size() == 4);
            CHECK(r[0] == (float)(t * 1000 + i) && r[1] == 2.f * (float)(t * 1000 + i));
            CHECK(r[2] >= 1.f && r[2] <= (float)cap && r[3] == -1.f);
        }
    const auto st = ps.stats();
    CHECK(st.tasks == (uint64_t)T * K && st.batches == (uint64_t)fs.calls.load());
    CHECK(fs.max_batch.load() <= (int)cap);
    std::printf("split: %d tasks from %d threads in %llu batches (max %d, cap %zu)\n", T * K, T,
                (unsigned long long)st.batches, fs.max_batch.load(), cap);
}

static void case_cap_2048()
{
    FakeSolver fs;
    ParallelSampler ps(&fs, 1, 100000);
    auto first = block_worker(ps, fs, 0.f);
    std::vector<std::future<std::vector<float>>> f;
    for (int i = 1; i <= 5000; i++) f.push_back(ps.submit(layer_of(i), cuboid_of(i)));
    fs.release();
    CHECK(first.get()[2] == 1.f);
    for (int i = 0; i < 5000; i++) {
        const auto r = f[i].get();
        CHECK(r[0] == (float)(i + 1));
        // the queued 5000 run as batches of 2048, 2048 and 904, in order
        CHECK(r[2] == (i < 4096 ? 2048.f : 904.f));
    }
    std::printf("cap: 5000 queued tasks ran as 2048 + 2048 + 904\n");
}

static void case_failure_fan_out()
{
    FakeSolver fs;
    ParallelSampler ps(&fs, 1, 64);
    auto first = block_worker(ps, fs, 100.f);
    std::vector<std::future<std::vector<float>>> bad;
    for (int i = 0; i < 9; i++) bad.push_back(ps.submit(layer_of(i == 4 ? -1.f : (float)i), cuboid_of(i)));
    fs.release();
    CHECK(first.get()[0] == 100.f);
    for (auto& f : bad) {
        bool threw = false;
        try {
            f.get();
        } catch (const std::runtime_error& e) {
            threw = std::string(e.what()) == "fake solver failure";
        }
        CHECK(threw);
    }
    auto size_first = block_worker(ps, fs, 101.f);
    std::vector<std::future<std::vector<float>>> mis;
    for (int i = 0; i < 3; i++) mis.push_back(ps.submit(layer_of(i == 1 ? -2.f : (float)i), cuboid_of(i)));
    fs.release();
    CHECK(size_first.get()[0] == 101.f);
    for (auto& f : mis) {
        bool threw = false;
        try {
            f.get();
        } catch (const std::runtime_error& e) {
            threw = std::string(e.what()).find("Result size") != std::string::npos;
        }
        CHECK(threw);
    }
    auto after = ps.submit(layer_of(7.f), cuboid_of(7.f));
    CHECK(after.get()[0] == 7.f);
    std::printf("failure: a throwing batch of 9 and a wrong-size batch of 3 failed every task, the next task ran\n");
}

static void case_axis_default()
{
    FakeSolver fs;
    ParallelSampler ps(&fs, 1, 64);
    auto first = block_worker(ps, fs, 0.f);
    auto a = ps.submit(layer_of(1.f), cuboid_of(1.f));        // no axis
    auto b = ps.submit(layer_of(2.f), cuboid_of(2.f), 2);     // axis Z
    auto c = ps.submit(layer_of(3.f), cuboid_of(3.f), 1);     // axis Y
    fs.release();
    CHECK(first.get()[3] == -1.f);  // a batch without axes takes the two-argument sample()
    const auto ra = a.get(), rb = b.get(), rc = c.get();
    CHECK(ra[2] == 3.f && ra[3] == 0.f && rb[3] == 2.f && rc[3] == 1.f);
    std::printf("axis: tasks without an axis ran with axis 0 beside tasks with one\n");
}

static void case_drain_on_shutdown()
{
    FakeSolver fs;
    fs.delay_us = 200;
    std::vector<std::future<std::vector<float>>> f;
    {
        ParallelSampler ps(&fs, 1, 7);
        for (int i = 0; i < 700; i++) f.push_back(ps.submit(layer_of(i), cuboid_of(i)));
    }  // destructor: signal, join, after the queue is empty
    for (int i = 0; i < 700; i++) {
        CHECK(f[i].wait_for(std::chrono::seconds(0)) == std::future_status::ready);
        CHECK(f[i].get()[0] == (float)i);
    }
    std::printf("drain: 700 tasks submitted just before destruction all completed\n");
}

int main()
{
    case_split_concurrent(2048, 2048);
    case_split_concurrent(7, 7);
    case_cap_2048();
    case_failure_fan_out();
    case_axis_default();
    case_drain_on_shutdown();
    {
        bool threw = false;
        try {
            ParallelSampler ps(nullptr, 1, 16);
        } catch (const std::invalid_argument&) {
            threw = true;
        }
        CHECK(threw);
        FakeSolver fs;
        threw = false;
        try {
            fs.setSeed(1, 0);  // no device solver behind a NoDevice subclass
        } catch (const std::runtime_error&) {
            threw = true;
        }
        CHECK(threw);
    }
    std::printf("all cases passed\n");
    return 0;
}
