// moao_hip.hpp -- header-only C++ drop-in layer over the C-ABI (moao_hip.h).
//
// A caller of the reference swaps namespaces and keeps its call sites:
//   MOAO_COMMON::matcov_tile(...)  ->  MOAO_HIP::matcov_tile(...)   (common/compute/matcov.hpp:14-18)
//   MOAO_LAPACK::reconstructor(...) -> MOAO_HIP::reconstructor(...) (lapack/compute/reconstructor.hpp:19-21)
//   MOAO_LAPACK::Cee_Cvv(...)       -> MOAO_HIP::Cee_Cvv(...)       (lapack/compute/cee_cvv.hpp:21)
//   intersample_prepare/init/process/free and struct isample_struct
//                                   -> MOAO_HIP:: of the same names (intersample.hpp:22-63)
// Arguments, in-place semantics and void returns are the reference's; buffers are host
// memory.  The functions take any Tomo_struct-like object exposing the reference's fields
// (sys.nW, sys.nLgs, sys.diam, sys.lgsCst, atm.nLayer, atm.cn2, Nx, Nsubap, Nssp, sspSizeL,
// u, v, X, Y, L0diff, noiseNGS, noiseLGSxx/yy/xy), so the reference's own Tomo_struct<double>
// works unchanged.  Errors are reported the reference's way (a message on stdout/stderr) and
// the status of the last call is kept in MOAO_HIP::last_status().
#ifndef MOAO_HIP_HPP
#define MOAO_HIP_HPP

#include <cstdio>
#include <cstdlib>
#include <type_traits>
#include <vector>

#include "moao_hip.h"

namespace MOAO_HIP {

inline int& last_status() {
    static thread_local int s = 0;
    return s;
}

// Fill the POD view the kernels read from a Tomo_struct-like object (host pointers into it).
template <typename Tomo>
moao_tomo_view make_view(Tomo* tomo) {
    static_assert(std::is_same<typename std::remove_reference<decltype(tomo->u[0])>::type,
                               double>::value,
                  "MOAO_HIP runs the FP64 path: use a Tomo_struct<double>");
    moao_tomo_view v;
    v.Nw = tomo->sys.nW;
    v.Nlayer = tomo->atm.nLayer;
    v.nlgs = tomo->sys.nLgs;
    v.Nx = tomo->Nx;
    v.diam = tomo->sys.diam;
    v.lgs_cst = tomo->sys.lgsCst;
    v.Nsubap = tomo->Nsubap.data();
    v.Nssp = tomo->Nssp.data();
    v.sspSizeL = tomo->sspSizeL.data();
    v.u = tomo->u.data();
    v.v = tomo->v.data();
    v.X = tomo->X.data();
    v.Y = tomo->Y.data();
    v.L0 = tomo->L0diff.data();
    v.cn2 = tomo->atm.cn2.data();
    v.noiseNGS = tomo->noiseNGS.data();
    v.noiseLGSxx = tomo->noiseLGSxx.data();
    v.noiseLGSyy = tomo->noiseLGSyy.data();
    v.noiseLGSxy = tomo->noiseLGSxy.data();
    return v;
}

// MOAO_COMMON::matcov_tile (common/compute/matcov.cpp:17-98).  Unknown type_mat: the
// reference prints and exits; here it prints and returns (status MOAO_ERR_ARG).
template <typename T, typename Tomo>
void matcov_tile(T* data, int nrows, int ncols, int xoffset, int yoffset, int lda, Tomo* tomo,
                 int type_mat) {
    static_assert(std::is_same<T, double>::value, "MOAO_HIP::matcov_tile is FP64");
    const moao_tomo_view v = make_view(tomo);
    const int rc = moao_hip_matcov_tile(data, nrows, ncols, xoffset, yoffset, lda, &v, type_mat);
    last_status() = rc;
    if (rc) std::fprintf(stderr, "ERROR: matcov_tile: %s\n", moao_last_error());
}

// MOAO_LAPACK::reconstructor (lapack/compute/reconstructor.cpp:32-58): Cmm -> L, Ctm -> R.
// T = double: the lapack path's FP64 arithmetic; T = float: the Chameleon production
// precision (chameleon/moao_chameleon.cpp:41), spotrf + 2 strsm semantics.
template <typename T>
void reconstructor(long nmeas, long nmeasts, T* Cmm, long ldCmm, T* Ctm, long ldCtm, long ngal) {
    static_assert(std::is_same<T, double>::value || std::is_same<T, float>::value,
                  "MOAO_HIP::reconstructor runs in FP64 or FP32");
    int rc;
    if (std::is_same<T, double>::value)
        rc = moao_hip_reconstructor(nmeas, nmeasts, reinterpret_cast<double*>(Cmm), ldCmm,
                                    reinterpret_cast<double*>(Ctm), ldCtm, ngal);
    else
        rc = moao_hip_sreconstructor(nmeas, nmeasts, reinterpret_cast<float*>(Cmm), ldCmm,
                                     reinterpret_cast<float*>(Ctm), ldCtm, ngal);
    last_status() = rc;
    if (rc > 0) std::printf("An error occured in dpotrf: %d\n", rc);
    if (rc < 0) std::fprintf(stderr, "ERROR: reconstructor: %s\n", moao_last_error());
}

// MOAO_LAPACK::Cee_Cvv (lapack/compute/cee_cvv.cpp:50-124).
template <typename T>
void Cee_Cvv(long nmeas, long nmeasts, long nact, T* Cmm, long ldCmm, T* Cpp, long ldCpp, T* Cpm,
             long ldCpm, T* R, long ldR, T* Dx, long ldDx, T* Cee, long ldCee, T* Cvv, long ldCvv,
             T* Tmp, long ldTmp) {
    static_assert(std::is_same<T, double>::value, "MOAO_HIP::Cee_Cvv is FP64");
    const int rc = moao_hip_cee_cvv(nmeas, nmeasts, nact, Cmm, ldCmm, Cpp, ldCpp, Cpm, ldCpm, R,
                                    ldR, Dx, ldDx, Cee, ldCee, Cvv, ldCvv, Tmp, ldTmp);
    last_status() = rc;
    if (rc) std::fprintf(stderr, "ERROR: Cee_Cvv: %s\n", moao_last_error());
}

// The intersample PSF (common/tomo_struct/intersample.hpp:22-63, intersample.cpp:175-333).
// A caller keeps `struct isample_struct isample; intersample_prepare(&isample, nact*nact, np,
// Dtel, path); intersample_init(&isample); intersample_process(&isample, Cvv);` and reads the
// PSF from isample.dphi (N x N).  init / process return 1 on success and 0 on failure, as the
// reference's do; the message goes to stderr and the status to last_status().
struct isample_struct : moao_isample_struct {};

inline void intersample_prepare(isample_struct* s, long nidx, long np, float Dtel,
                                const char* path) {
    moao_hip_intersample_prepare(s, nidx, np, Dtel, path);
    last_status() = 0;
}

inline int intersample_init(isample_struct* s) {
    const int rc = moao_hip_intersample_init(s);
    last_status() = rc;
    if (rc) std::fprintf(stderr, "ERROR: intersample_init: %s\n", moao_last_error());
    return rc == 0 ? 1 : 0;
}

template <typename T>
int intersample_process(isample_struct* s, T* data) {
    static_assert(std::is_same<T, double>::value || std::is_same<T, float>::value,
                  "intersample_process takes double or float Cvv");
    const int rc = std::is_same<T, double>::value
                       ? moao_hip_intersample_process_d(s, reinterpret_cast<const double*>(data))
                       : moao_hip_intersample_process_f(s, reinterpret_cast<const float*>(data));
    last_status() = rc;
    if (rc) std::fprintf(stderr, "ERROR: intersample_process: %s\n", moao_last_error());
    return rc == 0 ? 1 : 0;
}

inline void intersample_free(isample_struct* s) { moao_hip_intersample_free(s); }

}  // namespace MOAO_HIP

#endif  // MOAO_HIP_HPP
