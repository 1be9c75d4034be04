// The reference's intersample boundary (common/tomo_struct/intersample.hpp:22-63,
// intersample.cpp:175-333) over the device PSF stage: prepare fills the public fields and the
// three file names, init reads idx.fits / abs2fi.fits / otftel.fits and sets up the device stage,
// process<T> leaves the PSF in dphi (host, N x N doubles), free releases everything.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

namespace moao {
int fits_read(const std::string& path, long* nx, long* ny, std::vector<double>* out);
}

extern "C" {

void moao_hip_intersample_prepare(moao_isample_struct* s, long nidx, long np, float Dtel,
                                  const char* path) {
    (void)Dtel;  // intersample_prepare ignores it too (N comes from np alone)
    if (!s) return;
    std::memset(s, 0, sizeof *s);
    s->np = np;
    s->N = np >= 2 ? moao_intersample_default_n(np) : 0;
    s->nidx = nidx;
    s->lambda = 1.65e-06;
    // intersample.cpp:199: N/(3*(np-1))+1 in long arithmetic
    s->dactupix = np >= 2 ? (double)(s->N / (3 * (np - 1)) + 1) : 0.0;
    s->precision = 1;  // FP64 transforms: the reference's FFTW arithmetic
    const char* p = path ? path : "";
    std::snprintf(s->filename_idx, sizeof s->filename_idx, "%s/idx.fits", p);
    std::snprintf(s->filename_abs2fi, sizeof s->filename_abs2fi, "%s/abs2fi.fits", p);
    std::snprintf(s->filename_otftel, sizeof s->filename_otftel, "%s/otftel.fits", p);
}

int moao_hip_intersample_init(moao_isample_struct* s) {
    if (!s || s->nidx < 1 || s->np < 2 || s->N < 1) {
        moao::set_error("intersample_init: prepare was not called with a valid nidx / np");
        return MOAO_ERR_ARG;
    }
    if (s->dev) {
        moao::set_error("intersample_init: already initialised");
        return MOAO_ERR_ARG;
    }
    long nx = 0, ny = 0;
    std::vector<double> idxd, a2f, otf;
    // the reference checks the first axis only (naxes[0] == nidx, == N)
    if (int rc = moao::fits_read(s->filename_idx, &nx, &ny, &idxd)) return rc;
    if (nx != s->nidx || (long)idxd.size() < s->nidx) {
        moao::set_error(std::string("intersample_init: ") + s->filename_idx + " has " +
                        std::to_string(nx) + " entries, expected nidx = " +
                        std::to_string(s->nidx));
        return MOAO_ERR_IO;
    }
    const long nn = s->N * s->N;
    if (int rc = moao::fits_read(s->filename_abs2fi, &nx, &ny, &a2f)) return rc;
    if (nx != s->N || (long)a2f.size() < nn) {
        moao::set_error(std::string("intersample_init: ") + s->filename_abs2fi +
                        " is not N x N with N = " + std::to_string(s->N));
        return MOAO_ERR_IO;
    }
    if (int rc = moao::fits_read(s->filename_otftel, &nx, &ny, &otf)) return rc;
    if (nx != s->N || (long)otf.size() < nn) {
        moao::set_error(std::string("intersample_init: ") + s->filename_otftel +
                        " is not N x N with N = " + std::to_string(s->N));
        return MOAO_ERR_IO;
    }
    s->idx = static_cast<long*>(std::malloc(sizeof(long) * s->nidx));
    s->abs2fi = static_cast<double*>(std::malloc(sizeof(double) * nn));
    s->otf_tel = static_cast<double*>(std::malloc(sizeof(double) * nn));
    s->dphi = static_cast<double*>(std::calloc(nn, sizeof(double)));
    if (!s->idx || !s->abs2fi || !s->otf_tel || !s->dphi) {
        moao_hip_intersample_free(s);
        moao::set_error("intersample_init: host allocation failed");
        return MOAO_ERR_NOMEM;
    }
    for (long i = 0; i < s->nidx; ++i) s->idx[i] = (long)idxd[i];
    std::memcpy(s->abs2fi, a2f.data(), sizeof(double) * nn);
    std::memcpy(s->otf_tel, otf.data(), sizeof(double) * nn);
    const int rc = moao_hip_intersample_create(s->nidx, s->np, s->N, s->idx, s->abs2fi,
                                               s->otf_tel, s->precision, &s->dev);
    if (rc) {
        std::string msg = moao_last_error();
        moao_hip_intersample_free(s);
        moao::set_error(msg);
    }
    return rc;
}

int moao_hip_intersample_process_d(moao_isample_struct* s, const double* data) {
    if (!s || !s->dev || !data) {
        moao::set_error("intersample_process: not initialised or null data");
        return MOAO_ERR_ARG;
    }
    return moao_hip_intersample_process(s->dev, data, s->dphi);
}

int moao_hip_intersample_process_f(moao_isample_struct* s, const float* data) {
    if (!s || !s->dev || !data) {
        moao::set_error("intersample_process: not initialised or null data");
        return MOAO_ERR_ARG;
    }
    // the reference accumulates the map in double whatever T is (intersample.cpp:90)
    std::vector<double> d(data, data + s->nidx);
    return moao_hip_intersample_process(s->dev, d.data(), s->dphi);
}

void moao_hip_intersample_free(moao_isample_struct* s) {
    if (!s) return;
    if (s->dev) moao_hip_intersample_destroy(s->dev);
    std::free(s->idx);
    std::free(s->abs2fi);
    std::free(s->otf_tel);
    std::free(s->dphi);
    s->dev = nullptr;
    s->idx = nullptr;
    s->abs2fi = s->otf_tel = s->dphi = nullptr;
}

}  // extern "C"
