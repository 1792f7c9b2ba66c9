// C++ facade of include/dnn_solver.h over the C ABI.
#include "../../include/dnn_solver.h"

#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>

namespace {
void throw_on(int rc, const char* what)
{
    if (rc != 0) throw std::runtime_error(std::string(what) + ": " + drw_last_error());
}
}  // namespace

DNNSolver::DNNSolver(int kind, int device_idx) : kind_(kind), device_(device_idx)
{
    throw_on(drw_solver_create(kind, device_idx, &handle_), "DNNSolver");
}

DNNSolver::~DNNSolver()
{
    if (handle_) drw_solver_destroy(handle_);
}

void DNNSolver::useFDM(int n, double tol) { throw_on(drw_solver_use_fdm(handle_, n, tol), "useFDM"); }

void DNNSolver::setSeed(uint64_t seed, uint64_t counter)
{
    throw_on(drw_solver_set_seed(handle_, seed, counter), "setSeed");
}

void DNNSolver::loadModelsFrom(const std::string& models_txt, bool verbose)
{
    throw_on(drw_solver_load_models(handle_, models_txt.c_str()), "loadModels");
    if (verbose) std::cout << "\tModels loaded from " << models_txt << "\n";
}

std::vector<float> DNNSolver::run(const std::vector<float>& layers, const std::vector<float>& cuboids,
                                  const int32_t* axis)
{
    const int B = (int)(layers.size() / (MAX_LAYERS * 2));  // dnn_solver_poisson.cpp:40
    if (cuboids.size() != (size_t)B * MAX_CUBOIDS * 7)
        throw std::runtime_error("sample: cuboid array size does not match the layer array");
    std::vector<float> out((size_t)B * outputWidth());
    throw_on(drw_solver_sample(handle_, B, layers.data(), cuboids.data(), axis, out.data()), "sample");
    return out;
}

void DNNSolverPoisson::loadModels(bool verbose) { loadModelsFrom(MODELS_PATH, verbose); }

std::vector<float> DNNSolverPoisson::sample(const std::vector<float>& layerStructures,
                                            const std::vector<float>& cuboidStructures)
{
    return run(layerStructures, cuboidStructures, nullptr);
}

std::vector<float> DNNSolverPoisson::sample(const std::vector<float>& layerStructures,
                                            const std::vector<float>& cuboidStructures, const std::vector<int>& axis)
{
    static std::once_flag once;
    std::call_once(once, [] {
        std::cerr << "Warning: Axis parameter provided to Green's function solver but will be ignored. "
                  << "Green's function sampling does not use axis information. "
                  << "(This warning will only be shown once)" << std::endl;
    });
    (void)axis;
    return run(layerStructures, cuboidStructures, nullptr);
}

void DNNSolverGrad::loadModels(bool verbose) { loadModelsFrom(MODELS_PATH, verbose); }

std::vector<float> DNNSolverGrad::sample(const std::vector<float>&, const std::vector<float>&)
{
    throw std::runtime_error("Gradient solver requires axis parameter");  // dnn_solver_gradient.cpp:40-47
}

std::vector<float> DNNSolverGrad::sample(const std::vector<float>& layerStructures,
                                         const std::vector<float>& cuboidStructures, const std::vector<int>& axis)
{
    const size_t B = layerStructures.size() / (MAX_LAYERS * 2);
    if (axis.size() != B) throw std::runtime_error("sample: axis array size does not match the batch");
    std::vector<int32_t> ax(axis.begin(), axis.end());
    return run(layerStructures, cuboidStructures, ax.data());
}

// ----------------------------------------------------------- offline readers
// readBinaryFile: GGFT dataset records (dnn_solver.cpp:89-163): a 2-value
// header, then per sample N^3 permittivities, structureSize ignored values and
// 6 N^2 kernel values, as float64 or float32.
void DNNSolver::readBinaryFile(const std::string& filePath, std::vector<float>& dielectricConfigs,
                               std::vector<float>& targetFunctions, int numSamples, int structureSize,
                               DataType inputType, bool verbose)
{
    std::ifstream file(filePath, std::ios::binary);
    if (!file) throw std::runtime_error("Error opening file: " + filePath);
    const auto t0 = std::chrono::high_resolution_clock::now();
    const size_t ds = NNN, ts = (size_t)NN * 6, w = inputType == DataType::DOUBLE ? 8 : 4;
    dielectricConfigs.resize((size_t)numSamples * ds);
    targetFunctions.resize((size_t)numSamples * ts);
    file.ignore(2 * w);
    std::vector<double> tmp(std::max(ds, ts));
    for (int i = 0; i < numSamples; ++i) {
        auto rd = [&](float* dst, size_t n) {
            if (w == 8) {
                file.read(reinterpret_cast<char*>(tmp.data()), n * 8);
                for (size_t j = 0; j < n; j++) dst[j] = (float)tmp[j];
            } else {
                file.read(reinterpret_cast<char*>(dst), n * 4);
            }
        };
        rd(&dielectricConfigs[(size_t)i * ds], ds);
        file.ignore((std::streamsize)(structureSize * w));
        rd(&targetFunctions[(size_t)i * ts], ts);
        if (!file) throw std::runtime_error("readBinaryFile: file ends before sample " + std::to_string(i));
    }
    if (verbose) {
        const double dt = std::chrono::duration<double>(std::chrono::high_resolution_clock::now() - t0).count();
        std::cout << "Read " << numSamples << " dielectric configurations and target functions in " << 1000 * dt
                  << " ms" << std::endl;
    }
}

// readStructureFile: one sample per line, ';'-separated groups of
// comma/space-separated numbers; 2 values = layer, 7 = cuboid; padded with
// PADDING_VALUE to 15 each (dnn_solver.cpp:165-250).
int DNNSolver::readStructureFile(const std::string& filePath, std::vector<float>& layerStructures,
                                 std::vector<float>& cuboidStructures, int numSamples, bool verbose)
{
    std::ifstream file(filePath);
    if (!file) throw std::runtime_error("Error opening file");
    layerStructures.clear();
    cuboidStructures.clear();
    std::string line;
    int count = 0;
    while (count < numSamples && std::getline(file, line)) {
        std::vector<float> ly, cb;
        std::stringstream ss(line);
        std::string group;
        while (std::getline(ss, group, ';')) {
            std::stringstream gs(group);
            std::vector<float> v;
            float x;
            while (gs >> x) {
                v.push_back(x);
                if (gs.peek() == ',') gs.ignore();
            }
            if (v.size() == 2) ly.insert(ly.end(), v.begin(), v.end());
            else if (v.size() == 7) cb.insert(cb.end(), v.begin(), v.end());
        }
        if (ly.size() < (size_t)MAX_LAYERS * 2) ly.resize(MAX_LAYERS * 2, (float)PADDING_VALUE);
        if (cb.size() < (size_t)MAX_CUBOIDS * 7) cb.resize(MAX_CUBOIDS * 7, (float)PADDING_VALUE);
        layerStructures.insert(layerStructures.end(), ly.begin(), ly.end());
        cuboidStructures.insert(cuboidStructures.end(), cb.begin(), cb.end());
        count++;
    }
    if (verbose) std::cout << "Read " << count << " structures from " << filePath << std::endl;
    return count;
}

// ------------------------------------------------------------ ParallelSampler
ParallelSampler::ParallelSampler(DNNSolver* solver, size_t sample_size, size_t batch_size)
    : solver_(solver), sample_size_(sample_size), batch_size_(batch_size)
{
    if (!solver_) throw std::invalid_argument("Solver pointer cannot be null");
    gpu_thread_ = std::thread([this] { gpuWorkerLoop(); });
}

ParallelSampler::~ParallelSampler()
{
    {
        std::lock_guard<std::mutex> lk(mu_);
        shutdown_ = true;
    }
    cv_.notify_all();
    if (gpu_thread_.joinable()) gpu_thread_.join();
}

std::future<std::vector<float>> ParallelSampler::submit(const std::vector<float>& layer,
                                                        const std::vector<float>& cuboid)
{
    return submit(layer, cuboid, -1);
}

std::future<std::vector<float>> ParallelSampler::submit(const std::vector<float>& layer,
                                                        const std::vector<float>& cuboid, const int axis)
{
    auto t = std::make_unique<Task>();
    t->input_a = layer;
    t->input_b = cuboid;
    t->axis = axis;
    auto fut = t->promise.get_future();
    {
        std::lock_guard<std::mutex> lk(mu_);
        queue_.push_back(std::move(t));
    }
    cv_.notify_one();
    return fut;
}

void ParallelSampler::gpuWorkerLoop()
{
    const size_t cap = std::max<size_t>(1, std::min<size_t>(batch_size_, MAX_SAMPLES));
    for (;;) {
        std::vector<std::unique_ptr<Task>> batch;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [this] { return shutdown_ || !queue_.empty(); });
            if (queue_.empty() && shutdown_) return;  // drained
            while (!queue_.empty() && batch.size() < cap) {
                batch.push_back(std::move(queue_.front()));
                queue_.pop_front();
            }
        }
        processBatch(batch);
    }
}

void ParallelSampler::processBatch(std::vector<std::unique_ptr<Task>>& batch)
{
    if (batch.empty()) return;
    std::vector<float> a, b;
    std::vector<int> ax;
    bool has_axis = false;
    for (auto& t : batch) has_axis |= t->axis >= 0;
    for (auto& t : batch) {
        a.insert(a.end(), t->input_a.begin(), t->input_a.end());
        b.insert(b.end(), t->input_b.begin(), t->input_b.end());
        if (has_axis) ax.push_back(t->axis >= 0 ? t->axis : 0);  // default axis 0 (parallel_sampler.cpp:165-171)
    }
    std::vector<float> result;
    try {
        result = has_axis ? solver_->sample(a, b, ax) : solver_->sample(a, b);
    } catch (...) {
        auto e = std::current_exception();
        for (auto& t : batch) t->promise.set_exception(e);
        return;
    }
    if (result.size() % batch.size() != 0) {
        auto e = std::make_exception_ptr(std::runtime_error("Result size doesn't match expected batch output size"));
        for (auto& t : batch) t->promise.set_exception(e);
        return;
    }
    const size_t per = result.size() / batch.size();
    // counted before the results are handed out, so that a caller holding
    // every result also sees them in stats()
    tasks_ += batch.size();
    batches_ += 1;
    for (size_t i = 0; i < batch.size(); ++i)
        batch[i]->promise.set_value(std::vector<float>(result.begin() + i * per, result.begin() + (i + 1) * per));
}

ParallelSampler::Stats ParallelSampler::stats() const
{
    Stats s;
    s.tasks = tasks_.load();
    s.batches = batches_.load();
    s.mean_batch = s.batches ? (double)s.tasks / (double)s.batches : 0.0;
    return s;
}
