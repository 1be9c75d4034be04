// Tile descriptors and sequences: the call surface of MOAO_CHAMELEON::reconstructor
// (chameleon/compute/reconstructor.cpp:31-65), which takes its matrices as Chameleon tile
// descriptors (MORSE_desc_t, tile size ts = mb = nb) and runs asynchronously on a
// MORSE_sequence_t, waiting after each step when `sync` is set.
//
// Here a descriptor is a device matrix in tile layout -- tile (i, j) of mb x nb elements,
// column-major with leading dimension mb, at element offset (i + j*mt) * mb*nb, edge tiles padded
// to mb x nb -- and a sequence is a stream plus a device status word.  The solve itself is the
// recursive FP64 (or FP32-storage) reconstructor on column-major data: the async entry converts
// the tiles into column-major scratch, solves, and converts R back, all ordered on the
// sequence's stream, and returns at once; with sync it waits after the solve (the
// factorization and the first triangular solve are one fused pass here, so their two waits in
// the reference become one).  `check` is accepted and ignored, as in the reference (a TODO
// there).
#include <hip/hip_runtime.h>

#include <string>

#include "common.hpp"
#include "recon.hpp"

struct moao_hip_sequence {
    hipStream_t stream = nullptr;
    int* d_status = nullptr;  // first non-zero solve status of the sequence (device)
    int h_status = 0;         // first argument / launch error (host)
};

namespace moao {
namespace {

// column-major (ld) <-> tile layout, one thread per element
template <typename T, bool ToTile>
__global__ void tile_convert(T* lap, long ld, T* tiles, long lm, long ln, int mb, int nb,
                             long mt) {
    const long i = (long)blockIdx.x * 64 + (threadIdx.x & 63);
    const long j = (long)blockIdx.y * 4 + (threadIdx.x >> 6);
    if (i >= lm || j >= ln) return;
    const long ti = i / mb, tj = j / nb, ii = i - ti * mb, jj = j - tj * nb;
    T* t = tiles + (ti + tj * mt) * (long)mb * nb + ii + jj * (long)mb;
    if (ToTile) *t = lap[i + j * ld];
    else lap[i + j * ld] = *t;
}

__global__ void keep_first_status(int* seq_status, const int* info) {
    if (*seq_status == 0) *seq_status = *info;
}

int dsize(int dtype) { return dtype == 0 ? 8 : 4; }

int check_desc(const moao_hip_desc* d, const char* what) {
    if (!d || !d->mat || d->mb < 1 || d->nb < 1 || d->lm < 0 || d->ln < 0 ||
        (d->dtype != 0 && d->dtype != 1)) {
        set_error(std::string(what) + ": invalid descriptor");
        return MOAO_ERR_ARG;
    }
    return MOAO_OK;
}

template <typename T>
void convert(T* lap, long ld, const moao_hip_desc* d, bool to_tile, hipStream_t s) {
    const long mt = (d->lm + d->mb - 1) / d->mb;
    const dim3 g((unsigned)((d->lm + 63) / 64), (unsigned)((d->ln + 3) / 4));
    if (to_tile)
        hipLaunchKernelGGL((tile_convert<T, true>), g, dim3(256), 0, s, lap, ld, (T*)d->mat,
                           d->lm, d->ln, d->mb, d->nb, mt);
    else
        hipLaunchKernelGGL((tile_convert<T, false>), g, dim3(256), 0, s, lap, ld, (T*)d->mat,
                           d->lm, d->ln, d->mb, d->nb, mt);
}

template <typename T>
int solve_tiles(const moao_hip_desc* Cmm, moao_hip_desc* Ctm, moao_hip_sequence* q, int sync) {
    hipStream_t s = q->stream;
    const long n = Cmm->lm, m = Ctm->lm;
    const size_t a = ((size_t)n * n * sizeof(T) + 255) & ~size_t(255);
    const size_t b = ((size_t)m * n * sizeof(T) + 255) & ~size_t(255);
    const size_t w = recon_workspace_bytes(n);
    char* ws = nullptr;
    if (hipMallocAsync(reinterpret_cast<void**>(&ws), a + b + w, s) != hipSuccess) {
        (void)hipGetLastError();
        set_error("reconstructor_tile: workspace allocation failed");
        return MOAO_ERR_NOMEM;
    }
    T* A = reinterpret_cast<T*>(ws);
    T* B = reinterpret_cast<T*>(ws + a);
    char* rw = ws + a + b;
    int* info = reinterpret_cast<int*>(rw + w - 256);
    convert<T>(A, n, Cmm, false, s);
    convert<T>(B, m, Ctm, false, s);
    int rc;
    if constexpr (sizeof(T) == 8)
        rc = reconstructor_f64_dev(n, m, A, n, B, m, 1, reinterpret_cast<double*>(rw), info, s);
    else
        rc = reconstructor_f32_dev(n, m, A, n, B, m, 1, reinterpret_cast<float*>(rw), info, s);
    if (rc == MOAO_OK) {
        // Cmm receives the factor, Ctm receives R (partial results after a failing pivot, as
        // the asynchronous reference leaves them)
        convert<T>(A, n, Cmm, true, s);
        convert<T>(B, m, Ctm, true, s);
        hipLaunchKernelGGL(keep_first_status, dim3(1), dim3(1), 0, s, q->d_status, info);
    }
    (void)hipFreeAsync(ws, s);
    MOAO_HIP_TRY(hipGetLastError());
    if (rc) return rc;
    if (sync) MOAO_HIP_TRY(hipStreamSynchronize(s));
    return MOAO_OK;
}

}  // namespace
}  // namespace moao

extern "C" {

int moao_hip_sequence_create(moao_hip_sequence** out) {
    if (int rc = moao::ensure_init()) return rc;
    if (!out) return MOAO_ERR_ARG;
    auto* q = new moao_hip_sequence;
    if (hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(reinterpret_cast<void**>(&q->d_status), sizeof(int)) != hipSuccess ||
        hipMemset(q->d_status, 0, sizeof(int)) != hipSuccess) {
        if (q->stream) (void)hipStreamDestroy(q->stream);
        if (q->d_status) (void)hipFree(q->d_status);
        delete q;
        moao::set_error("sequence_create failed");
        return MOAO_ERR_HIP;
    }
    *out = q;
    return MOAO_OK;
}

int moao_hip_sequence_wait(moao_hip_sequence* q) {
    if (!q) return MOAO_ERR_ARG;
    MOAO_HIP_TRY(hipStreamSynchronize(q->stream));
    if (q->h_status) return q->h_status;
    int st = 0;
    MOAO_HIP_TRY(hipMemcpy(&st, q->d_status, sizeof(int), hipMemcpyDeviceToHost));
    return st;
}

void moao_hip_sequence_destroy(moao_hip_sequence* q) {
    if (!q) return;
    (void)hipStreamSynchronize(q->stream);
    (void)hipFree(q->d_status);
    (void)hipStreamDestroy(q->stream);
    delete q;
}

int moao_hip_lapack_to_tile(const void* dA, long lda, moao_hip_desc* d, moao_hip_sequence* q) {
    if (int rc = moao::check_desc(d, "lapack_to_tile")) return rc;
    if (!dA || !q || lda < d->lm) return MOAO_ERR_ARG;
    if (d->dtype == 0) moao::convert<double>((double*)dA, lda, d, true, q->stream);
    else moao::convert<float>((float*)dA, lda, d, true, q->stream);
    MOAO_HIP_TRY(hipGetLastError());
    return MOAO_OK;
}

int moao_hip_tile_to_lapack(const moao_hip_desc* d, void* dA, long lda, moao_hip_sequence* q) {
    if (int rc = moao::check_desc(d, "tile_to_lapack")) return rc;
    if (!dA || !q || lda < d->lm) return MOAO_ERR_ARG;
    if (d->dtype == 0) moao::convert<double>((double*)dA, lda, d, false, q->stream);
    else moao::convert<float>((float*)dA, lda, d, false, q->stream);
    MOAO_HIP_TRY(hipGetLastError());
    return MOAO_OK;
}

int moao_hip_reconstructor_tile_async(const moao_hip_desc* Cmm, moao_hip_desc* Ctm,
                                      moao_hip_sequence* seq, int check, int sync) {
    (void)check;
    if (int rc = moao::ensure_init()) return rc;
    if (!seq) return MOAO_ERR_ARG;
    int rc = moao::check_desc(Cmm, "reconstructor_tile");
    if (!rc) rc = moao::check_desc(Ctm, "reconstructor_tile");
    if (!rc && (Cmm->lm != Cmm->ln || Ctm->ln != Cmm->lm || Cmm->dtype != Ctm->dtype ||
                Cmm->lm < 1)) {
        moao::set_error("reconstructor_tile: Cmm must be n x n and Ctm m x n, same type");
        rc = MOAO_ERR_ARG;
    }
    if (!rc)
        rc = Cmm->dtype == 0 ? moao::solve_tiles<double>(Cmm, Ctm, seq, sync)
                             : moao::solve_tiles<float>(Cmm, Ctm, seq, sync);
    if (rc && !seq->h_status) seq->h_status = rc;
    return rc;
}

}  // extern "C"
