#include "profile.hpp"

#include <mutex>
#include <vector>

#include "common.hpp"

namespace moao {
namespace {

struct Rec {
    int cls;
    double work;
    hipEvent_t a, b;
};

std::mutex g_mu;
bool g_on = false;
std::vector<Rec> g_recs;
std::vector<hipEvent_t> g_pool;
size_t g_used = 0;

hipEvent_t take() {
    if (g_used < g_pool.size()) return g_pool[g_used++];
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    g_pool.push_back(e);
    ++g_used;
    return e;
}

}  // namespace

bool profiling() { return g_on; }

int prof_start(int cls, double work, hipStream_t s) {
    if (!g_on) return -1;
    // launches recorded into a graph are not timed here (no event records in a capture)
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) return -1;
    std::lock_guard<std::mutex> g(g_mu);
    hipEvent_t a = take(), b = take();
    if (!a || !b) return -1;
    if (hipEventRecord(a, s) != hipSuccess) return -1;
    g_recs.push_back({cls, work, a, b});
    return (int)g_recs.size() - 1;
}

void prof_stop(int idx, hipStream_t s) {
    if (idx < 0) return;
    std::lock_guard<std::mutex> g(g_mu);
    (void)hipEventRecord(g_recs[idx].b, s);
}

}  // namespace moao

extern "C" {

int moao_hip_profile_begin(void) {
    if (int rc = moao::ensure_init()) return rc;
    std::lock_guard<std::mutex> g(moao::g_mu);
    moao::g_recs.clear();
    moao::g_used = 0;
    moao::g_on = true;
    return MOAO_OK;
}

// out_ms[c], out_work[c], out_count[c] for c < kProfClasses (7); returns the class count.
int moao_hip_profile_end(double* out_ms, double* out_work, long* out_count) {
    std::lock_guard<std::mutex> g(moao::g_mu);
    moao::g_on = false;
    for (int c = 0; c < moao::kProfClasses; ++c) {
        if (out_ms) out_ms[c] = 0;
        if (out_work) out_work[c] = 0;
        if (out_count) out_count[c] = 0;
    }
    for (auto& r : moao::g_recs) {
        if (hipEventSynchronize(r.b) != hipSuccess) {
            moao::set_error("profile: event synchronisation failed");
            return MOAO_ERR_HIP;
        }
        float ms = 0;
        if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) continue;
        if (out_ms) out_ms[r.cls] += ms;
        if (out_work) out_work[r.cls] += r.work;
        if (out_count) out_count[r.cls] += 1;
    }
    moao::g_recs.clear();
    moao::g_used = 0;
    return moao::kProfClasses;
}

}  // extern "C"
