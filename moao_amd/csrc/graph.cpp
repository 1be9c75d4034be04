// Captured solves: the FP64 reconstructor's launch sequence recorded once as a HIP graph and
// replayed with one hipGraphLaunch.
//
// A solve of order n is 18 (n = 456) to ~1,300 (n = 40k) kernel launches; below a few
// thousand the host-side launch work and the gaps between kernels are a sizeable part of the
// solve.  The recursion depends only on (n, m, ngal) and the buffer addresses, so a caller that
// solves the same shape repeatedly on the same buffers -- the refine loop of one system, a
// serving slot whose pool hands back the same block -- records the sequence once and replays
// it.  The capture holds no allocation, no memset node and no host synchronisation: the status
// word is cleared by a kernel, and the leaf inverses live in workspace the graph object owns.
#include <hip/hip_runtime.h>

#include <string>

#include "common.hpp"
#include "graph.hpp"
#include "recon.hpp"

namespace moao {

SolveGraph::~SolveGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (own_ws) (void)hipFree(own_ws);
}

int capture_solve(long n, long m, double* Cmm, long ldCmm, double* Ctm, long ldCtm, long ngal,
                  double* invL, int* info, hipStream_t capture_stream, SolveGraph* out) {
    out->n = n;
    out->m = m;
    out->ngal = ngal;
    out->Cmm = Cmm;
    out->Ctm = Ctm;
    out->invL = invL;
    out->info = info;
    MOAO_HIP_TRY(hipStreamBeginCapture(capture_stream, hipStreamCaptureModeThreadLocal));
    const int rc = reconstructor_f64_dev(n, m, Cmm, ldCmm, Ctm, ldCtm, ngal, invL, info,
                                         capture_stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(capture_stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess || !g) {
        set_error(std::string("solve graph: capture failed: ") + hipGetErrorString(e));
        return MOAO_ERR_HIP;
    }
    out->graph = g;
    MOAO_HIP_TRY(hipGraphInstantiate(&out->exec, g, nullptr, nullptr, 0));
    return MOAO_OK;
}

}  // namespace moao

struct moao_solve_graph {
    moao::SolveGraph g;
};

extern "C" {

int moao_hip_solve_graph_create(long nmeas, long nmeasts, double* dCmm, long ldCmm, double* dCtm,
                                long ldCtm, long ngal, int* d_info, moao_solve_graph** out) {
    if (int rc = moao::ensure_init()) return rc;
    if (!out || !dCmm || !d_info || nmeas < 1 || nmeasts < 0 || (nmeasts > 0 && !dCtm)) {
        moao::set_error("solve_graph_create: invalid arguments");
        return MOAO_ERR_ARG;
    }
    *out = nullptr;
    auto* h = new moao_solve_graph;
    const size_t bytes = sizeof(double) * moao::recon_invl_doubles(nmeas);
    if (hipMalloc(&h->g.own_ws, bytes) != hipSuccess) {
        (void)hipGetLastError();
        delete h;
        moao::set_error("solve_graph_create: workspace allocation failed");
        return MOAO_ERR_NOMEM;
    }
    hipStream_t cs = nullptr;
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        moao::set_error("solve_graph_create: capture stream");
        return MOAO_ERR_HIP;
    }
    const int rc = moao::capture_solve(nmeas, nmeasts, dCmm, ldCmm, dCtm, ldCtm, ngal,
                                       static_cast<double*>(h->g.own_ws), d_info, cs, &h->g);
    (void)hipStreamDestroy(cs);
    if (rc) {
        delete h;
        return rc;
    }
    *out = h;
    return MOAO_OK;
}

int moao_hip_solve_graph_launch(moao_solve_graph* g, void* stream) {
    if (!g || !g->g.exec) {
        moao::set_error("solve_graph_launch: no graph");
        return MOAO_ERR_ARG;
    }
    if (int rc = moao::ensure_init()) return rc;
    MOAO_HIP_TRY(hipGraphLaunch(g->g.exec, moao::pick_stream(stream)));
    return MOAO_OK;
}

void moao_hip_solve_graph_destroy(moao_solve_graph* g) { delete g; }

}  // extern "C"
