// Captured FP64 solves (HIP graphs) shared by the C-ABI and the serving engine.
#pragma once

#include <hip/hip_runtime.h>

namespace moao {

struct SolveGraph {
    long n = 0, m = 0, ngal = 0;
    const double* Cmm = nullptr;  // buffers the recorded launches read and write
    const double* Ctm = nullptr;
    const double* invL = nullptr;
    const int* info = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void* own_ws = nullptr;       // workspace owned by this graph (freed with it), or nullptr
    SolveGraph() = default;
    SolveGraph(const SolveGraph&) = delete;
    SolveGraph& operator=(const SolveGraph&) = delete;
    ~SolveGraph();
    bool matches(long n_, long m_, long g_, const double* C, const double* B, const double* w,
                 const int* i) const {
        return exec && n == n_ && m == m_ && ngal == g_ && Cmm == C && Ctm == B && invL == w &&
               info == i;
    }
};

// Records reconstructor_f64_dev(n, m, Cmm, ldCmm, Ctm, ldCtm, ngal, invL, info) on
// capture_stream (idle, not shared) into out->graph / out->exec.
int capture_solve(long n, long m, double* Cmm, long ldCmm, double* Ctm, long ldCtm, long ngal,
                  double* invL, int* info, hipStream_t capture_stream, SolveGraph* out);

}  // namespace moao
