// The offline readers of the C++ facade (inference_cpp/src/dnn_solver.cpp:
// 89-250) on files written here: readStructureFile (one sample per line,
// ';'-separated groups, 2 values = layer, 7 = cuboid, anything else ignored,
// padded with PADDING_VALUE to 15 of each, stops after numSamples lines) and
// readBinaryFile (GGFT dataset records: a 2-value header, then per sample N^3
// permittivities, structureSize skipped values and 6 N^2 kernel values, in
// float64 or float32).  A file that ends early is an error here (the
// reference reads past the end silently).  Built and run by
// tests/test_offline_readers.py; exits non-zero on the first failure.
// With --read-bin it only reads a given float64 dataset file and writes what
// it read (tests/test_ggft_bin.py).
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "dnn_solver.h"

#define CHECK(c)                                                              \
    do {                                                                      \
        if (!(c)) {                                                           \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
            std::exit(1);                                                     \
        }                                                                     \
    } while (0)

// readStructureFile is a member: a solver without a device behind it
class FileOnlySolver : public DNNSolver {
public:
    FileOnlySolver() : DNNSolver(NoDevice{}, DRW_POISSON) {}
    void loadModels(bool) override {}
    std::vector<float> sample(const std::vector<float>&, const std::vector<float>&) override { return {}; }
    std::vector<float> sample(const std::vector<float>&, const std::vector<float>&, const std::vector<int>&) override
    {
        return {};
    }
};

template <typename T>
static void write_records(const std::string& path, int samples, int structure_size, bool truncate)
{
    std::ofstream f(path, std::ios::binary);
    const T hdr[2] = {(T)23, (T)1};
    f.write((const char*)hdr, sizeof hdr);
    for (int s = 0; s < samples; s++) {
        std::vector<T> d(DNNSolver::NNN), st(structure_size), g(6 * DNNSolver::NN);
        for (size_t i = 0; i < d.size(); i++) d[i] = (T)(1.0 + 0.001 * (double)i + s);
        for (size_t i = 0; i < st.size(); i++) st[i] = (T)(-7.0 - (double)i);
        for (size_t i = 0; i < g.size(); i++) g[i] = (T)(0.5 * (double)i - 3.0 * s);
        f.write((const char*)d.data(), d.size() * sizeof(T));
        f.write((const char*)st.data(), st.size() * sizeof(T));
        const size_t gn = truncate && s == samples - 1 ? g.size() / 2 : g.size();
        f.write((const char*)g.data(), gn * sizeof(T));
    }
}

static void check_records(const std::vector<float>& d, const std::vector<float>& g, int samples)
{
    CHECK(d.size() == (size_t)samples * DNNSolver::NNN && g.size() == (size_t)samples * 6 * DNNSolver::NN);
    for (int s = 0; s < samples; s++) {
        for (int i = 0; i < DNNSolver::NNN; i += 997)
            CHECK(d[(size_t)s * DNNSolver::NNN + i] == (float)(1.0 + 0.001 * (double)i + s));
        for (int i = 0; i < 6 * DNNSolver::NN; i += 101)
            CHECK(g[(size_t)s * 6 * DNNSolver::NN + i] == (float)(0.5 * (double)i - 3.0 * s));
    }
}

int main(int argc, char** argv)
{
    // readers_check --read-bin FILE SAMPLES STRUCTURE_SIZE OUT: read a float64
    // dataset file and write its dielectric and target floats to OUT
    if (argc == 6 && std::string(argv[1]) == "--read-bin") {
        std::vector<float> d, g;
        DNNSolver::readBinaryFile(argv[2], d, g, std::atoi(argv[3]), std::atoi(argv[4]), DataType::DOUBLE, false);
        std::ofstream o(argv[5], std::ios::binary);
        o.write(reinterpret_cast<const char*>(d.data()), d.size() * sizeof(float));
        o.write(reinterpret_cast<const char*>(g.data()), g.size() * sizeof(float));
        return o ? 0 : 1;
    }
    const std::string dir = argc > 1 ? argv[1] : ".";
    // ---- structure text file
    {
        const std::string p = dir + "/structures.txt";
        std::ofstream f(p);
        f << "-1.5, 2.0; 0.2 4.5; -0.5,-0.5,-1.5, 0.5,0.5,1.5, 7.25\n";  // 2 layers, 1 cuboid
        f << "0.1,0.2,0.3,0.4,0.5,0.6,3;1,2,3;-1,-1,-1,1,1,1,2\n";       // 2 cuboids; a 3-value group is ignored
        f << "-1.5,9\n";                                                 // not read (numSamples = 2)
        f.close();
        FileOnlySolver s;
        std::vector<float> L, C;
        const int n = s.readStructureFile(p, L, C, 2, false);
        CHECK(n == 2);
        CHECK(L.size() == 2 * MAX_LAYERS * 2 && C.size() == 2 * MAX_CUBOIDS * 7);
        CHECK(L[0] == -1.5f && L[1] == 2.0f && L[2] == 0.2f && L[3] == 4.5f && L[4] == (float)PADDING_VALUE);
        CHECK(C[0] == -0.5f && C[5] == 1.5f && C[6] == 7.25f && C[7] == (float)PADDING_VALUE);
        const float* C1 = C.data() + MAX_CUBOIDS * 7;
        const float* L1 = L.data() + MAX_LAYERS * 2;
        CHECK(L1[0] == (float)PADDING_VALUE && L1[MAX_LAYERS * 2 - 1] == (float)PADDING_VALUE);
        CHECK(C1[0] == 0.1f && C1[6] == 3.0f && C1[7] == -1.0f && C1[13] == 2.0f && C1[14] == (float)PADDING_VALUE);
        std::vector<float> L3, C3;
        CHECK(s.readStructureFile(p, L3, C3, 10, false) == 3);  // stops at the end of the file
        CHECK(L3[2 * MAX_LAYERS * 2] == -1.5f && L3[2 * MAX_LAYERS * 2 + 1] == 9.0f);
        bool threw = false;
        try {
            s.readStructureFile(dir + "/missing.txt", L3, C3, 1, false);
        } catch (const std::runtime_error&) {
            threw = true;
        }
        CHECK(threw);
        std::printf("readStructureFile: layers, cuboids, ignored groups, padding, sample limit\n");
    }
    // ---- GGFT dataset records
    {
        const int S = 3, SS = 7 * 16;  // ggft_solver.h STRUCTURE: 7 x 16 values per record
        std::vector<float> d, g;
        write_records<double>(dir + "/ds64.bin", S, SS, false);
        DNNSolver::readBinaryFile(dir + "/ds64.bin", d, g, S, SS, DataType::DOUBLE, false);
        check_records(d, g, S);
        write_records<float>(dir + "/ds32.bin", S, SS, false);
        DNNSolver::readBinaryFile(dir + "/ds32.bin", d, g, S, SS, DataType::FLOAT, false);
        check_records(d, g, S);
        write_records<double>(dir + "/short.bin", S, SS, true);
        bool threw = false;
        try {
            DNNSolver::readBinaryFile(dir + "/short.bin", d, g, S, SS, DataType::DOUBLE, false);
        } catch (const std::runtime_error&) {
            threw = true;
        }
        CHECK(threw);
        std::printf("readBinaryFile: float64 and float32 records, structure block skipped, short file rejected\n");
    }
    std::printf("all cases passed\n");
    return 0;
}
