// The drop-in path as a walker uses it: include/dnn_solver.h + libdnnsolver.so
// and nothing else (the INTEGRATION.md section-1 program, run on real data).
//
// Reads a batch of structure descriptions, runs it through
//   DNNSolverPoisson / DNNSolverGrad::sample (dnn_solver_poisson.cpp:35-63,
//   dnn_solver_gradient.cpp:40-73) on the network backend (loadModels(),
//   ./models.txt) and on the FDM backend (useFDM(23, tol)), one fixed batch
//   each after setSeed(seed, c0);
//   ParallelSampler (parallel_sampler.cpp:42-211) in front of the same four
//   solvers, re-seeded to (seed, c1), with 8 threads submitting one task at a
//   time (thread t submits queries t, t + 8, ...),
// and writes every result and each solver's counter after the ParallelSampler
// run.  The checks against the oracle are in tests/test_gpu_dropin.py.
//
// usage: dropin_check IN OUT FDM_TOL   (run in a directory holding models.txt)
// IN : int32 B, uint64 seed, c0, c1, float layers[B*30], cuboids[B*105], int32 axis[B]
// OUT: float fixed[4][B*w], float parallel[4][B*w], uint64 counters[4]
//      (w = 3 for the Poisson solvers 0 and 2, 5 for the gradient solvers 1 and 3)
#include <cstdio>
#include <cstdlib>
#include <future>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "dnn_solver.h"

static void die(const std::string& m)
{
    std::fprintf(stderr, "dropin_check: %s\n", m.c_str());
    std::exit(1);
}

template <typename T>
static void read_into(FILE* f, T* p, size_t n)
{
    if (std::fread(p, sizeof(T), n, f) != n) die("short input file");
}

int main(int argc, char** argv)
{
    if (argc != 4) die("usage: dropin_check IN OUT FDM_TOL");
    FILE* in = std::fopen(argv[1], "rb");
    if (!in) die("cannot open input");
    int32_t B = 0;
    uint64_t seed = 0, c0 = 0, c1 = 0;
    read_into(in, &B, 1);
    read_into(in, &seed, 1);
    read_into(in, &c0, 1);
    read_into(in, &c1, 1);
    std::vector<float> layers((size_t)B * MAX_LAYERS * 2), cuboids((size_t)B * MAX_CUBOIDS * 7);
    std::vector<int> axis(B);
    read_into(in, layers.data(), layers.size());
    read_into(in, cuboids.data(), cuboids.size());
    read_into(in, axis.data(), axis.size());
    std::fclose(in);
    const double tol = std::atof(argv[3]);

    try {
        // solvers 0, 1: networks; 2, 3: FDM backend at N = 23
        std::vector<std::unique_ptr<DNNSolver>> solvers;
        solvers.emplace_back(new DNNSolverPoisson(0));
        solvers.emplace_back(new DNNSolverGrad(0));
        solvers.emplace_back(new DNNSolverPoisson(0));
        solvers.emplace_back(new DNNSolverGrad(0));
        solvers[0]->loadModels();
        solvers[1]->loadModels();
        solvers[2]->useFDM(23, tol);
        solvers[3]->useFDM(23, tol);

        std::vector<std::vector<float>> fixed(4), par(4);
        std::vector<uint64_t> counters(4);
        for (int k = 0; k < 4; k++) {
            DNNSolver* s = solvers[k].get();
            s->setSeed(seed, c0);
            fixed[k] = s->kind() == DRW_GRADIENT ? s->sample(layers, cuboids, axis) : s->sample(layers, cuboids);
            if (fixed[k].size() != (size_t)B * s->outputWidth()) die("fixed batch: wrong result size");
        }
        for (int k = 0; k < 4; k++) {
            DNNSolver* s = solvers[k].get();
            s->setSeed(seed, c1);
            const int w = s->outputWidth();
            par[k].assign((size_t)B * w, 0.f);
            {
                ParallelSampler ps(s, 1, 2048);
                std::vector<std::thread> th;
                for (int t = 0; t < 8; t++) {
                    th.emplace_back([&, t] {
                        std::vector<std::pair<int, std::future<std::vector<float>>>> fut;
                        for (int q = t; q < B; q += 8) {
                            std::vector<float> l(layers.begin() + (size_t)q * 30, layers.begin() + (size_t)q * 30 + 30);
                            std::vector<float> c(cuboids.begin() + (size_t)q * 105,
                                                 cuboids.begin() + (size_t)q * 105 + 105);
                            fut.emplace_back(q, s->kind() == DRW_GRADIENT ? ps.submit(l, c, axis[q])
                                                                          : ps.submit(l, c));
                        }
                        for (auto& f : fut) {
                            std::vector<float> r = f.second.get();
                            if (r.size() != (size_t)w) die("parallel: wrong result size");
                            for (int j = 0; j < w; j++) par[k][(size_t)f.first * w + j] = r[j];
                        }
                    });
                }
                for (auto& x : th) x.join();
                const auto st = ps.stats();
                std::printf("solver %d: %llu tasks in %llu batches (mean %.1f)\n", k, (unsigned long long)st.tasks,
                            (unsigned long long)st.batches, st.mean_batch);
            }
            if (drw_solver_get_counter(s->handle(), &counters[k]) != 0) die(drw_last_error());
        }
        FILE* out = std::fopen(argv[2], "wb");
        if (!out) die("cannot open output");
        for (auto& v : fixed) std::fwrite(v.data(), sizeof(float), v.size(), out);
        for (auto& v : par) std::fwrite(v.data(), sizeof(float), v.size(), out);
        std::fwrite(counters.data(), sizeof(uint64_t), 4, out);
        std::fclose(out);
    } catch (const std::exception& e) {
        die(std::string("exception: ") + e.what());
    }
    std::printf("dropin_check done\n");
    return 0;
}
