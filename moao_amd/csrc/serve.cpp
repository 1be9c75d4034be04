// Serving engine: continuous batching of reconstructor requests on one GPU.
//
// The reference runs one system at a time (lapack/moao_lapack.cpp:95-188: matcov -> solve per
// atmospheric profile).  A serving deployment instead receives solves for new profiles and for
// systems of uneven size while earlier ones are still on the device.  Here:
//   * submit() builds the request's Tomo on the caller's thread (host work, no GPU), which fixes
//     n and m and so the request's HBM footprint, and queues it;
//   * `slots` worker threads, each with its own HIP stream and scratch, take requests from the
//     queue as soon as a slot is free AND the footprint fits in the byte budget (first fit in
//     arrival order; a request passed over max_bypass times blocks later ones until it fits, so
//     large requests are not starved);
//   * a slot allocates the request's device block stream-ordered (hipMallocAsync from the
//     device pool, which keeps freed memory cached), generates Cmm and Ctm with matcov, solves
//     FP64 or mixed precision, copies R out if asked, and frees the block on its stream;
//   * an FP64 solve of order up to kGraphMaxN that repeats the slot's previous solve (same
//     shape, and the pool handed back the same block -- the common case for a stream of one
//     system's refine steps) runs as a captured graph: one hipGraphLaunch instead of its launch
//     sequence.  The slot records the graph the second time it sees that solve and replays it
//     after; a shape seen once runs eagerly, so one-off requests pay no capture.
// Work of different slots overlaps on the device; the kernels, their order and their inputs for
// one request do not depend on what else runs, so a result is bit-identical alone or in a mix.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"
#include "graph.hpp"
#include "recon.hpp"
#include "stages.hpp"
#include "tomo.hpp"

namespace moao {
namespace {

using Clock = std::chrono::steady_clock;

moao_tomo_view view_of(const Tomo& t) {
    moao_tomo_view v;
    v.Nw = t.sys.nW;
    v.Nlayer = t.atm.nLayer;
    v.nlgs = t.sys.nLgs;
    v.Nx = t.Nx;
    v.diam = t.sys.diam;
    v.lgs_cst = t.sys.lgsCst;
    v.Nsubap = t.Nsubap.data();
    v.Nssp = t.Nssp.data();
    v.sspSizeL = t.sspSizeL.data();
    v.u = t.u.data();
    v.v = t.v.data();
    v.X = t.X.data();
    v.Y = t.Y.data();
    v.L0 = t.L0diff.data();
    v.cn2 = t.cn2.data();
    v.noiseNGS = t.noiseNGS.data();
    v.noiseLGSxx = t.noiseLGSxx.data();
    v.noiseLGSyy = t.noiseLGSyy.data();
    v.noiseLGSxy = t.noiseLGSxy.data();
    return v;
}

size_t al256(size_t b) { return (b + 255) & ~size_t(255); }

struct Request {
    long ticket = 0;
    std::unique_ptr<Tomo> tomo;
    moao_tomo_view view{};
    int mode = MOAO_FP64;
    double tol = 1e-11;
    int max_iter = 5;
    double* R = nullptr;
    long ldR = 0;
    long n = 0, m = 0;
    size_t solve_bytes = 0;  // scratch of the solve (inverse leaves or mixed workspace)
    size_t footprint = 0;
    int bypassed = 0;
    bool done = false;
    moao_serve_result res{};
    std::string error;
};

constexpr long kGraphMaxN = 16384;  // FP64 solves up to this order run as captured graphs

struct Slot {
    hipStream_t stream = nullptr;
    hipStream_t capture = nullptr;       // idle stream the slot records its graphs on
    DevBuf tables;                       // matcov upload tables (small, kept)
    std::unique_ptr<SolveGraph> graph;   // the graph of this slot's last captured solve
    struct {
        long n = 0, m = 0;
        const void *Cmm = nullptr, *Ctm = nullptr, *invL = nullptr, *info = nullptr;
    } last;                              // shape and buffers of the slot's previous FP64 solve
};

}  // namespace
}  // namespace moao

struct moao_server {
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<moao::Request*> queue;
    std::map<long, std::unique_ptr<moao::Request>> reqs;
    std::vector<moao::Slot> slots;
    std::vector<std::thread> workers;
    size_t budget = 0, in_use = 0, peak_bytes = 0;
    long running = 0, peak_running = 0, next_ticket = 1, completed = 0;
    long graph_captures = 0, graph_replays = 0;
    int max_bypass = 8;
    bool stop = false;
    moao::Clock::time_point t0 = moao::Clock::now();

    double now() const {
        return std::chrono::duration<double>(moao::Clock::now() - t0).count();
    }
    // First request (arrival order) whose footprint fits; a head passed over max_bypass times
    // blocks everything behind it.  Called with mu held.
    moao::Request* pick() {
        for (size_t i = 0; i < queue.size(); ++i) {
            moao::Request* r = queue[i];
            if (r->footprint <= budget - in_use) {
                for (size_t j = 0; j < i; ++j) ++queue[j]->bypassed;
                queue.erase(queue.begin() + i);
                return r;
            }
            if (r->bypassed >= max_bypass) return nullptr;
        }
        return nullptr;
    }
    void worker(int k);
    int run(moao::Request& r, moao::Slot& s);
    int run_body(moao::Request& r, moao::Slot& s);
};

int moao_server::run(moao::Request& r, moao::Slot& sl) {
    const int rc = run_body(r, sl);
    if (rc && r.error.empty()) r.error = moao_last_error();
    return rc;
}

int moao_server::run_body(moao::Request& r, moao::Slot& sl) {
    using namespace moao;
    hipStream_t s = sl.stream;
    const long n = r.n, m = r.m;
    char* block = nullptr;
    if (hipMallocAsync(reinterpret_cast<void**>(&block), r.footprint, s) != hipSuccess) {
        (void)hipGetLastError();
        set_error("serve: device allocation of " + std::to_string(r.footprint) + " B failed");
        return MOAO_ERR_NOMEM;
    }
    double* Cmm = reinterpret_cast<double*>(block);
    double* Ctm = reinterpret_cast<double*>(block + al256(sizeof(double) * n * n));
    char* scratch = block + al256(sizeof(double) * n * n) + al256(sizeof(double) * m * n);
    int rc = MOAO_OK;
    int hinfo = 0;
    do {
        rc = matcov_tile_launch(Cmm, (int)n, (int)n, 0, 0, (int)n, &r.view, 1, s, &sl.tables);
        if (rc) break;
        rc = matcov_tile_launch(Ctm, (int)m, (int)n, 0, 0, (int)m, &r.view, 3, s, &sl.tables);
        if (rc) break;
        if (r.mode == MOAO_FP64) {
            double* invL = reinterpret_cast<double*>(scratch);
            int* info = reinterpret_cast<int*>(scratch + recon_workspace_bytes(n) - 256);
            const bool again = sl.last.n == n && sl.last.m == m && sl.last.Cmm == Cmm &&
                               sl.last.Ctm == Ctm && sl.last.invL == invL && sl.last.info == info;
            sl.last.n = n;
            sl.last.m = m;
            sl.last.Cmm = Cmm;
            sl.last.Ctm = Ctm;
            sl.last.invL = invL;
            sl.last.info = info;
            const bool replay = sl.graph && sl.graph->matches(n, m, 1, Cmm, Ctm, invL, info);
            if (n <= kGraphMaxN && (replay || again)) {
                if (replay) {
                    std::lock_guard<std::mutex> g(mu);
                    ++graph_replays;
                } else {
                    sl.graph.reset(new SolveGraph);
                    rc = capture_solve(n, m, Cmm, n, Ctm, m, 1, invL, info, sl.capture,
                                       sl.graph.get());
                    if (rc) {
                        sl.graph.reset();
                        break;
                    }
                    std::lock_guard<std::mutex> g(mu);
                    ++graph_captures;
                }
                if (hipGraphLaunch(sl.graph->exec, s) != hipSuccess) {
                    set_error("serve: graph launch failed");
                    rc = MOAO_ERR_HIP;
                    break;
                }
            } else {
                rc = reconstructor_f64_dev(n, m, Cmm, n, Ctm, m, 1, invL, info, s);
                if (rc) break;
            }
            if (hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s) !=
                hipSuccess) {
                set_error("serve: status copy failed");
                rc = MOAO_ERR_HIP;
                break;
            }
        } else {
            double resid = 0;
            int iters = 0;
            rc = reconstructor_mixed_dev(n, m, Cmm, n, Ctm, m, &r.view, r.mode, r.tol, r.max_iter,
                                         &resid,
                                         &iters, scratch, r.solve_bytes, s);
            r.res.resid = resid;
            r.res.iters = iters;
        }
        if (rc) break;
        if (r.R) {
            hipError_t e = hipMemcpy2DAsync(r.R, r.ldR * sizeof(double), Ctm, m * sizeof(double),
                                            m * sizeof(double), n, hipMemcpyDeviceToHost, s);
            if (e != hipSuccess) {
                set_error(std::string("serve: copy of R: ") + hipGetErrorString(e));
                rc = MOAO_ERR_HIP;
                break;
            }
        }
    } while (false);
    if (rc) r.error = moao_last_error();
    hipError_t e1 = hipFreeAsync(block, s);
    hipError_t e2 = hipStreamSynchronize(s);
    if (!rc && (e1 != hipSuccess || e2 != hipSuccess)) {
        set_error(std::string("serve: ") + hipGetErrorString(e1 != hipSuccess ? e1 : e2));
        rc = MOAO_ERR_HIP;
    }
    if (!rc && hinfo) {
        set_error("serve: Cmm is not positive definite (dpotrf info " + std::to_string(hinfo) +
                  ")");
        r.error = moao_last_error();
        rc = hinfo;
    }
    return rc;
}

void moao_server::worker(int k) {
    if (hipSetDevice(moao::ctx().device) != hipSuccess) return;
    moao::Slot& sl = slots[k];
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        moao::Request* r = nullptr;
        cv_work.wait(lk, [&] { return stop || (r = pick()) != nullptr; });
        if (!r) return;  // stop with nothing admissible
        in_use += r->footprint;
        peak_bytes = std::max(peak_bytes, in_use);
        ++running;
        peak_running = std::max(peak_running, running);
        r->res.t_start = now();
        r->res.slot = k;
        lk.unlock();
        const int rc = run(*r, sl);
        lk.lock();
        r->res.t_end = now();
        r->res.status = rc;
        in_use -= r->footprint;
        --running;
        ++completed;
        r->done = true;
        cv_done.notify_all();
        cv_work.notify_all();  // memory freed: others may fit now
    }
}

extern "C" {

int moao_serve_create(const moao_serve_config* cfg, moao_server** out) {
    if (int rc = moao::ensure_init()) return rc;
    if (!out) return MOAO_ERR_ARG;
    *out = nullptr;
    const int nslots = cfg && cfg->slots > 0 ? cfg->slots : 4;
    if (nslots > 16) {
        moao::set_error("serve: at most 16 slots");
        return MOAO_ERR_ARG;
    }
    size_t free_b = 0, total_b = 0;
    MOAO_HIP_TRY(hipMemGetInfo(&free_b, &total_b));
    size_t budget = cfg && cfg->hbm_budget ? cfg->hbm_budget : (size_t)(0.85 * (double)free_b);
    if (budget > free_b) {
        moao::set_error("serve: HBM budget exceeds free device memory");
        return MOAO_ERR_NOMEM;
    }
    // Keep freed blocks cached in the device pool: a steady request stream then reuses memory
    // instead of returning it to the driver after every request.
    hipMemPool_t pool;
    MOAO_HIP_TRY(hipDeviceGetDefaultMemPool(&pool, moao::ctx().device));
    uint64_t thr = UINT64_MAX;
    MOAO_HIP_TRY(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr));
    auto* s = new moao_server;
    s->budget = budget;
    if (cfg && cfg->max_bypass > 0) s->max_bypass = cfg->max_bypass;
    s->slots.resize(nslots);
    for (auto& sl : s->slots) {
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&sl.capture, hipStreamNonBlocking) != hipSuccess) {
            for (auto& o : s->slots) {
                if (o.stream) (void)hipStreamDestroy(o.stream);
                if (o.capture) (void)hipStreamDestroy(o.capture);
            }
            delete s;
            moao::set_error("serve: stream creation failed");
            return MOAO_ERR_HIP;
        }
    }
    for (int k = 0; k < nslots; ++k) s->workers.emplace_back([s, k] { s->worker(k); });
    *out = s;
    return MOAO_OK;
}

int moao_serve_submit(moao_server* s, const moao_serve_request* q, long* ticket) {
    if (!s || !q || !q->sys_file || !q->atm_file || !ticket) {
        moao::set_error("serve_submit: null argument");
        return MOAO_ERR_ARG;
    }
    if (q->mode != MOAO_FP64 && q->mode != MOAO_MIXED_FP32 && q->mode != MOAO_MIXED_BF16) {
        moao::set_error("serve_submit: unknown mode");
        return MOAO_ERR_ARG;
    }
    auto r = std::make_unique<moao::Request>();
    const double t_submit = s->now();
    try {
        r->tomo.reset(new moao::Tomo(q->sys_file, q->atm_file));
        r->tomo->set_target(q->target);
    } catch (const std::exception& e) {
        moao::set_error(std::string("serve_submit: ") + e.what());
        return MOAO_ERR_IO;
    }
    r->view = moao::view_of(*r->tomo);
    r->n = r->tomo->nmeas();
    r->m = r->tomo->nmeasts();
    if (r->n < 1 || r->m < 1 || r->n >= (1L << 31)) {
        moao::set_error("serve_submit: invalid system");
        return MOAO_ERR_ARG;
    }
    // R is m x n: the caller states the extent it allocated, and it must be exactly that
    if (q->R && (q->ldR < r->m || q->R_cols != r->n)) {
        moao::set_error("serve_submit: R is " + std::to_string(r->m) + " x " +
                        std::to_string(r->n) + " but the output holds ldR " +
                        std::to_string(q->ldR) + " x " + std::to_string(q->R_cols) + " columns");
        return MOAO_ERR_ARG;
    }
    r->mode = q->mode;
    r->tol = q->tol > 0 ? q->tol : 1e-11;
    r->max_iter = q->max_iter > 0 ? q->max_iter : 5;
    r->R = q->R;
    r->ldR = q->ldR;
    r->solve_bytes = r->mode == MOAO_FP64 ? moao::recon_workspace_bytes(r->n)
                                          : moao::mixed_workspace_bytes(r->n, r->m, &r->view, r->mode);
    r->footprint = moao::al256(sizeof(double) * r->n * r->n) +
                   moao::al256(sizeof(double) * r->m * r->n) + moao::al256(r->solve_bytes);
    r->res.n = r->n;
    r->res.m = r->m;
    r->res.footprint = r->footprint;
    r->res.t_submit = t_submit;
    std::lock_guard<std::mutex> g(s->mu);
    if (r->footprint > s->budget) {
        moao::set_error("serve_submit: request footprint " + std::to_string(r->footprint) +
                        " B exceeds the HBM budget " + std::to_string(s->budget) + " B");
        return MOAO_ERR_NOMEM;
    }
    r->ticket = s->next_ticket++;
    *ticket = r->ticket;
    s->queue.push_back(r.get());
    s->reqs[r->ticket] = std::move(r);
    s->cv_work.notify_all();
    return MOAO_OK;
}

int moao_serve_wait(moao_server* s, long ticket, moao_serve_result* out) {
    if (!s) return MOAO_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) {
        moao::set_error("serve_wait: unknown ticket");
        return MOAO_ERR_ARG;
    }
    moao::Request* r = it->second.get();
    s->cv_done.wait(lk, [&] { return r->done; });
    if (out) *out = r->res;
    const int rc = r->res.status;
    if (rc) moao::set_error(r->error);
    s->reqs.erase(it);
    return rc;
}

int moao_serve_stats(moao_server* s, moao_serve_counters* st) {
    if (!s || !st) return MOAO_ERR_ARG;
    std::lock_guard<std::mutex> g(s->mu);
    st->queued = (long)s->queue.size();
    st->running = s->running;
    st->completed = s->completed;
    st->bytes_in_use = s->in_use;
    st->peak_bytes = s->peak_bytes;
    st->peak_running = s->peak_running;
    st->budget = s->budget;
    st->graph_captures = s->graph_captures;
    st->graph_replays = s->graph_replays;
    return MOAO_OK;
}

void moao_serve_destroy(moao_server* s) {
    if (!s) return;
    {
        // finish everything already submitted, then stop the workers
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_done.wait(lk, [&] { return s->queue.empty() && s->running == 0; });
        s->stop = true;
        s->cv_work.notify_all();
    }
    for (auto& t : s->workers) t.join();
    for (auto& sl : s->slots) {
        sl.tables.release();
        sl.graph.reset();
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.capture) (void)hipStreamDestroy(sl.capture);
    }
    delete s;
}

}  // extern "C"
