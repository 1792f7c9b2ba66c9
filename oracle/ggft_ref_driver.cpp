// Driver for the reference GGFT solver, compiled from the reference sources
// where they lie (ggft/src/ggft_solver.cpp + vendored Eigen 3.4.0), never
// copied.  TEST INFRASTRUCTURE ONLY: it writes golden vectors and times the
// reference CPU path; no product code uses it.
//
// The reference fixes N at compile time (ggft/src/ggft_solver.h:33): the
// driver takes N from the header it is compiled against (oracle/Makefile
// builds it at N = 23 from the sources as they are and at N = 16, 21 from
// copies in which only that constant differs).
//
// usage : ggft_ref [dataset.bin nstruct]
// stdin : int32 count, int32 code (0 GFT, 1 WVTZ, 2 WVTX, 3 WVTY),
//         float64 tol, then count cubes of N^3 float64 permittivities
//         (x fastest, ggft_solver.h:28), each followed by nstruct * 7
//         float64 structure values when a dataset file is named.
// stdout: per cube 6 N^2 float64 G (ggft_solver.cpp:595) then float64 seconds.
// With a dataset file, the reference solver itself writes it: the [N,
// block_width] header (ggft_solver.cpp:35-48) and after each solve the
// record DIEL, STRUCTURE, G (ggft_solver.h:180-191, ggft_solver.cpp:631).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <cstdint>
#include <iostream>
#include <vector>

#include "ggft_solver.h"

int main(int argc, char** argv)
{
    const std::string write_to = argc > 1 ? argv[1] : "/dev/null";
    const int nstruct = argc > 2 ? std::atoi(argv[2]) : 0;
    int32_t count = 0, code = 0;
    double tol = 1e-13;
    if (fread(&count, 4, 1, stdin) != 1 || fread(&code, 4, 1, stdin) != 1 || fread(&tol, 8, 1, stdin) != 1) return 2;
    // The solver reports progress on std::cout; keep stdout for the vectors.
    std::cout.rdbuf(std::cerr.rdbuf());
    deeprwcap::ggft_solver solver(write_to);
    solver.set_tolerance(tol);
    const int n = deeprwcap::index::NNN;
    std::vector<double> eps(n);
    for (int c = 0; c < count; c++) {
        if (fread(eps.data(), 8, n, stdin) != (size_t)n) return 3;
        for (int i = 0; i < n; i++) solver.DIEL[i] = eps[i];
        solver.STRUCTURE.assign((size_t)nstruct * 7, 0.0);
        if (nstruct > 0 && fread(solver.STRUCTURE.data(), 8, solver.STRUCTURE.size(), stdin) != solver.STRUCTURE.size())
            return 3;
        auto t0 = std::chrono::high_resolution_clock::now();
        bool ok = solver.solve(static_cast<deeprwcap::ggft_solver::solver_code>(code));
        auto t1 = std::chrono::high_resolution_clock::now();
        if (!ok) return 4;
        fwrite(solver.gft.data(), 8, solver.gft.size(), stdout);
        double s = std::chrono::duration<double>(t1 - t0).count();
        fwrite(&s, 8, 1, stdout);
    }
    return 0;
}
