// Device context, memory helpers and the system-model part of the C-ABI.
#include <hip/hip_runtime.h>

#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>

#include "common.hpp"
#include "recon.hpp"
#include "tomo.hpp"

namespace moao {

namespace {
thread_local std::string g_last_error;
Context g_ctx;
}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }

int DevBuf::ensure(size_t need) {
    if (need <= bytes) return MOAO_OK;
    if (ptr) {
        hipError_t e = hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
        if (e != hipSuccess) {
            set_error(std::string("hipFree: ") + hipGetErrorString(e));
            return MOAO_ERR_HIP;
        }
    }
    hipError_t e = hipMalloc(&ptr, need);
    if (e != hipSuccess) {
        ptr = nullptr;
        set_error("hipMalloc(" + std::to_string(need) + "): " + hipGetErrorString(e));
        return MOAO_ERR_NOMEM;
    }
    bytes = need;
    return MOAO_OK;
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

Context& ctx() { return g_ctx; }

static int init_device(int device) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= device || device < 0) {
        set_error("no HIP device " + std::to_string(device) + " (" + hipGetErrorString(e) + ")");
        return MOAO_ERR_NODEV;
    }
    hipDeviceProp_t prop;
    MOAO_HIP_TRY(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_error(std::string("device is ") + prop.gcnArchName + ", this library targets gfx950");
        return MOAO_ERR_NODEV;
    }
    MOAO_HIP_TRY(hipSetDevice(device));
    // Default-priority streams only: mixing stream priorities can stall work indefinitely.
    MOAO_HIP_TRY(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
    g_ctx.device = device;
    // the allocation-free reconstructor's scratch address on this device
    return noalloc_scratch_setup();
}

int ensure_init() {
    if (g_ctx.device >= 0) {
        hipError_t e = hipSetDevice(g_ctx.device);
        if (e != hipSuccess) {
            set_error(std::string("hipSetDevice: ") + hipGetErrorString(e));
            return MOAO_ERR_HIP;
        }
        return MOAO_OK;
    }
    return init_device(0);
}

}  // namespace moao

using moao::ctx;
using moao::set_error;

struct moao_tomo {
    std::unique_ptr<moao::Tomo> t;
    std::vector<double> L0;  // kept so the view can point at doubles
};

extern "C" {

const char* moao_last_error(void) { return moao::g_last_error.c_str(); }
const char* moao_version(void) { return "moao_hip 0.1 (gfx950)"; }

int moao_tomo_create(const char* sys_file, const char* atm_file, moao_tomo** out) {
    if (!sys_file || !atm_file || !out) {
        set_error("moao_tomo_create: null argument");
        return MOAO_ERR_ARG;
    }
    try {
        auto* h = new moao_tomo;
        h->t.reset(new moao::Tomo(sys_file, atm_file));
        *out = h;
    } catch (const std::exception& e) {
        set_error(std::string("moao_tomo_create: ") + e.what());
        *out = nullptr;
        return MOAO_ERR_IO;
    }
    return MOAO_OK;
}

int moao_tomo_update_atm(moao_tomo* t, const char* atm_file) {
    if (!t || !atm_file) return MOAO_ERR_ARG;
    try {
        t->t->update_atm(std::string(atm_file));
    } catch (const std::exception& e) {
        set_error(std::string("moao_tomo_update_atm: ") + e.what());
        return MOAO_ERR_IO;
    }
    return MOAO_OK;
}

int moao_tomo_set_target(moao_tomo* t, int target) {
    if (!t) return MOAO_ERR_ARG;
    try {
        t->t->set_target(target);
    } catch (const std::exception& e) {
        set_error(std::string("moao_tomo_set_target: ") + e.what());
        return MOAO_ERR_ARG;
    }
    return MOAO_OK;
}

int moao_tomo_set_target_coord(moao_tomo* t, double ax, double ay, int to_rad) {
    if (!t) return MOAO_ERR_ARG;
    t->t->set_target_coord(ax, ay, to_rad != 0);
    return MOAO_OK;
}

int moao_tomo_get_view(const moao_tomo* h, moao_tomo_view* v) {
    if (!h || !v) return MOAO_ERR_ARG;
    const moao::Tomo& t = *h->t;
    v->Nw = t.sys.nW;
    v->Nlayer = t.atm.nLayer;
    v->nlgs = t.sys.nLgs;
    v->Nx = t.Nx;
    v->diam = t.sys.diam;
    v->lgs_cst = t.sys.lgsCst;
    v->Nsubap = t.Nsubap.data();
    v->Nssp = t.Nssp.data();
    v->sspSizeL = t.sspSizeL.data();
    v->u = t.u.data();
    v->v = t.v.data();
    v->X = t.X.data();
    v->Y = t.Y.data();
    v->L0 = t.L0diff.data();
    v->cn2 = t.cn2.data();
    v->noiseNGS = t.noiseNGS.data();
    v->noiseLGSxx = t.noiseLGSxx.data();
    v->noiseLGSyy = t.noiseLGSyy.data();
    v->noiseLGSxy = t.noiseLGSxy.data();
    return MOAO_OK;
}

long moao_tomo_nmeas(const moao_tomo* t) { return t ? t->t->nmeas() : MOAO_ERR_ARG; }
long moao_tomo_nmeasts(const moao_tomo* t) { return t ? t->t->nmeasts() : MOAO_ERR_ARG; }

int moao_tomo_dims(const moao_tomo* h, long* out) {
    if (!h || !out) return MOAO_ERR_ARG;
    const moao::Tomo& t = *h->t;
    out[0] = t.sys.nW;
    out[1] = t.sys.nLgs;
    out[2] = t.sys.nNgs;
    out[3] = t.sys.nTarget;
    out[4] = t.atm.nLayer;
    out[5] = t.sys.nssp;
    return MOAO_OK;
}

long moao_tomo_field(const moao_tomo* h, const char* name, double* out, long cap) {
    if (!h || !name) return MOAO_ERR_ARG;
    const moao::Tomo& t = *h->t;
    const std::string f(name);
    auto put = [&](const auto& vec) -> long {
        const long n = (long)vec.size();
        if (out)
            for (long i = 0; i < n && i < cap; ++i) out[i] = (double)vec[i];
        return n;
    };
    if (f == "X") return put(t.X);
    if (f == "Y") return put(t.Y);
    if (f == "u") return put(t.u);
    if (f == "v") return put(t.v);
    if (f == "sspSizeL") return put(t.sspSizeL);
    if (f == "L0") return put(t.L0diff);
    if (f == "cn2") return put(t.cn2);
    if (f == "noiseNGS") return put(t.noiseNGS);
    if (f == "noiseLGSxx") return put(t.noiseLGSxx);
    if (f == "noiseLGSyy") return put(t.noiseLGSyy);
    if (f == "noiseLGSxy") return put(t.noiseLGSxy);
    if (f == "lgsExt") return put(t.lgsExt);
    if (f == "lgsTheta") return put(t.lgsTheta);
    if (f == "Nsubap") return put(t.Nsubap);
    if (f == "Nssp") return put(t.Nssp);
    if (f == "alphaX") return put(t.alphaX);
    if (f == "alphaY") return put(t.alphaY);
    set_error("moao_tomo_field: unknown field " + f);
    return MOAO_ERR_ARG;
}

void moao_tomo_destroy(moao_tomo* t) { delete t; }

int moao_hip_init(int device) {
    if (ctx().device >= 0) {
        if (ctx().device == device) return MOAO_OK;
        set_error("moao_hip_init: already initialised on another device");
        return MOAO_ERR_ARG;
    }
    return moao::init_device(device);
}

int moao_hip_finalize(void) {
    if (ctx().device < 0) return MOAO_OK;
    moao::Context& c = ctx();
    (void)hipStreamSynchronize(c.stream);
    c.matcov_tables.release();
    c.host_stage.release();
    c.host_stage2.release();
    c.recon_ws.release();
    return MOAO_OK;
}

int moao_hip_malloc(void** dptr, size_t bytes) {
    if (int rc = moao::ensure_init()) return rc;
    if (!dptr) return MOAO_ERR_ARG;
    hipError_t e = hipMalloc(dptr, bytes ? bytes : 1);
    if (e != hipSuccess) {
        *dptr = nullptr;
        set_error("hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
        return MOAO_ERR_NOMEM;
    }
    return MOAO_OK;
}

int moao_hip_free(void* dptr) {
    if (!dptr) return MOAO_OK;
    MOAO_HIP_TRY(hipFree(dptr));
    return MOAO_OK;
}

int moao_hip_memcpy_h2d(void* dst, const void* src, size_t bytes) {
    if (int rc = moao::ensure_init()) return rc;
    MOAO_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx().stream));
    MOAO_HIP_TRY(hipStreamSynchronize(ctx().stream));
    return MOAO_OK;
}

int moao_hip_memcpy_d2h(void* dst, const void* src, size_t bytes) {
    if (int rc = moao::ensure_init()) return rc;
    MOAO_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx().stream));
    MOAO_HIP_TRY(hipStreamSynchronize(ctx().stream));
    return MOAO_OK;
}

int moao_hip_memcpy_d2d(void* dst, const void* src, size_t bytes) {
    if (int rc = moao::ensure_init()) return rc;
    MOAO_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx().stream));
    return MOAO_OK;
}

int moao_hip_synchronize(void) {
    if (int rc = moao::ensure_init()) return rc;
    MOAO_HIP_TRY(hipStreamSynchronize(ctx().stream));
    return MOAO_OK;
}

int moao_hip_mem_info(size_t* free_bytes, size_t* total_bytes) {
    if (int rc = moao::ensure_init()) return rc;
    MOAO_HIP_TRY(hipMemGetInfo(free_bytes, total_bytes));
    return MOAO_OK;
}

int moao_hip_stream_create(void** out) {
    if (int rc = moao::ensure_init()) return rc;
    if (!out) return MOAO_ERR_ARG;
    hipStream_t s = nullptr;
    // default priority, like every stream of the library
    MOAO_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *out = s;
    return MOAO_OK;
}

int moao_hip_stream_synchronize(void* stream) {
    if (int rc = moao::ensure_init()) return rc;
    MOAO_HIP_TRY(hipStreamSynchronize(moao::pick_stream(stream)));
    return MOAO_OK;
}

int moao_hip_stream_destroy(void* stream) {
    if (!stream) return MOAO_OK;
    MOAO_HIP_TRY(hipStreamDestroy(static_cast<hipStream_t>(stream)));
    return MOAO_OK;
}

}  // extern "C"
