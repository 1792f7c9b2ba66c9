// Source-compatible C++ interface of the reference dnnsolver library
// (inference_cpp/src/dnn_solver.h:21-201), rebuilt on the C ABI of
// dnnsolver_c.h.  Class names, method signatures, argument meaning, output
// layouts and error behaviour follow the reference; members are opaque
// handles instead of LibTorch types, so code written against the reference
// header recompiles unchanged (binary layout differs).
//
// Additions (no counterpart in the reference): handle(), useFDM(),
// setSeed(), ParallelSampler::stats(), the NoDevice constructor.
#pragma once

#include <atomic>
#include <condition_variable>
#include <cstddef>
#include <cstdint>
#include <deque>
#include <future>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "dnnsolver_c.h"

#define MODELS_PATH "models.txt"

// Structure related parameters (dnn_solver.h:26-30)
const int MAX_SAMPLES = 2048;
const int MAX_CUBOIDS = 15;
const int MAX_LAYERS = 15;
const int PADDING_VALUE = -9999;

enum class Axis : int { X = 0, Y = 1, Z = 2 };
enum class DataType { FLOAT, DOUBLE };

class DNNSolver {
public:
    static constexpr int N = 23;
    static constexpr int NN = N * N;
    static constexpr int NNN = N * N * N;

    explicit DNNSolver(int device_idx = 0) : DNNSolver(DRW_POISSON, device_idx) {}
    virtual ~DNNSolver();
    DNNSolver(const DNNSolver&) = delete;
    DNNSolver& operator=(const DNNSolver&) = delete;

    // Offline readers (dnn_solver.cpp:89-250).  Errors throw
    // std::runtime_error: a file that cannot be opened, and (readBinaryFile)
    // one that ends before numSamples records.
    static void readBinaryFile(const std::string& filePath, std::vector<float>& dielectricConfigs,
                               std::vector<float>& targetFunctions, int numSamples, int structureSize,
                               DataType inputType, bool verbose);
    int readStructureFile(const std::string& filePath, std::vector<float>& layerStructures,
                          std::vector<float>& cuboidStructures, int numSamples, bool verbose);

    virtual void loadModels(bool verbose = false) = 0;
    virtual std::vector<float> sample(const std::vector<float>& layerStructures,
                                      const std::vector<float>& cuboidStructures) = 0;
    virtual std::vector<float> sample(const std::vector<float>& layerStructures,
                                      const std::vector<float>& cuboidStructures, const std::vector<int>& axis) = 0;

    drw_solver* handle() const { return handle_; }
    int kind() const { return kind_; }
    int outputWidth() const { return kind_ == DRW_GRADIENT ? 5 : 3; }
    void useFDM(int n, double tol);
    void setSeed(uint64_t seed, uint64_t counter = 0);

protected:
    DNNSolver(int kind, int device_idx);
    // For subclasses that implement both sample() overloads themselves and
    // need no device solver (a recorded or remote sampler, a test double in
    // front of ParallelSampler): no handle is created, and the handle-based
    // calls (useFDM, setSeed, loadModelsFrom, run) throw.
    struct NoDevice {};
    DNNSolver(NoDevice, int kind) : kind_(kind) {}
    void loadModelsFrom(const std::string& models_txt, bool verbose);
    std::vector<float> run(const std::vector<float>& layers, const std::vector<float>& cuboids, const int32_t* axis);

    drw_solver* handle_ = nullptr;
    int kind_ = DRW_POISSON;
    int device_ = 0;
};

class DNNSolverPoisson : public DNNSolver {
public:
    explicit DNNSolverPoisson(int device_idx = 0) : DNNSolver(DRW_POISSON, device_idx) {}
    void loadModels(bool verbose = false) override;
    std::vector<float> sample(const std::vector<float>& layerStructures,
                              const std::vector<float>& cuboidStructures) override;
    // The axis is ignored for Green's-function sampling (warns once).
    std::vector<float> sample(const std::vector<float>& layerStructures, const std::vector<float>& cuboidStructures,
                              const std::vector<int>& axis) override;
};

class DNNSolverGrad : public DNNSolver {
public:
    explicit DNNSolverGrad(int device_idx = 0) : DNNSolver(DRW_GRADIENT, device_idx) {}
    void loadModels(bool verbose = false) override;
    // Throws std::runtime_error: the gradient solver requires an axis.
    std::vector<float> sample(const std::vector<float>& layerStructures,
                              const std::vector<float>& cuboidStructures) override;
    std::vector<float> sample(const std::vector<float>& layerStructures, const std::vector<float>& cuboidStructures,
                              const std::vector<int>& axis) override;
};

// Host request batching in front of one solver (parallel_sampler.cpp):
// submit() is thread-safe and returns a future; a worker thread takes up to
// min(batch_size, MAX_SAMPLES) pending tasks, runs them as one batch and
// splits the result evenly; a failed batch fails every task in it; the
// destructor drains the queue.
class ParallelSampler {
public:
    ParallelSampler(DNNSolver* solver, size_t sample_size, size_t batch_size);
    ~ParallelSampler();
    std::future<std::vector<float>> submit(const std::vector<float>& layer, const std::vector<float>& cuboid);
    std::future<std::vector<float>> submit(const std::vector<float>& layer, const std::vector<float>& cuboid,
                                           const int axis);

    struct Stats {
        uint64_t tasks = 0, batches = 0;
        double mean_batch = 0.0;
    };
    Stats stats() const;

private:
    struct Task {
        std::vector<float> input_a;
        std::vector<float> input_b;
        int axis = -1;
        std::promise<std::vector<float>> promise;
    };
    void gpuWorkerLoop();
    void processBatch(std::vector<std::unique_ptr<Task>>& batch);

    DNNSolver* solver_;
    const size_t sample_size_;
    const size_t batch_size_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::unique_ptr<Task>> queue_;
    std::atomic<bool> shutdown_{false};
    std::thread gpu_thread_;
    std::atomic<uint64_t> tasks_{0}, batches_{0};
};
