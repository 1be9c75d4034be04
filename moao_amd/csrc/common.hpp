// Shared runtime pieces of libmoao_hip: status codes, HIP error checks, per-device context.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdio>
#include <mutex>
#include <string>

#include "../../include/moao_hip.h"

namespace moao {

// Last error message of the calling thread, readable through moao_last_error().
void set_error(const std::string& msg);

#define MOAO_HIP_TRY(expr)                                                            \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) {                                                       \
            ::moao::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));     \
            return MOAO_ERR_HIP;                                                      \
        }                                                                             \
    } while (0)

// Device scratch that grows on demand and is kept for the life of the context.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t need);
    void release();
    template <typename T>
    T* as() const { return static_cast<T*>(ptr); }
};

// One per process (single GPU).  Owns the library stream and reusable scratch.
struct Context {
    int device = -1;
    hipStream_t stream = nullptr;
    DevBuf matcov_tables;   // per-call tomo upload for matcov
    DevBuf host_stage;      // device copy of a host matrix for the host-pointer entry points
    DevBuf host_stage2;
    DevBuf recon_ws;        // reconstructor workspace (inverse leaves + status word)
    std::mutex mu;
};

Context& ctx();
// Initialises the context on first use (device 0 unless moao_hip_init chose another).
int ensure_init();

inline hipStream_t pick_stream(void* s) {
    return s ? static_cast<hipStream_t>(s) : ctx().stream;
}

}  // namespace moao
