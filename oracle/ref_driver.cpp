// Checker-side driver around the reference's own CPU code (test infrastructure only).
//
// This file is compiled together with the unmodified reference sources that lie under the
// reference tree (common/kernels/matcov_kernels.cpp, common/compute/matcov.cpp,
// common/compute/check_utils.cpp, common/tomo_struct/*.cpp, common/io/IO.cpp,
// common/utils/*.cpp) by oracle/Makefile.ref into oracle/_ref/libmoao_ref.so.  It exposes a
// plain C interface so the Python tests and the golden-fixture script can ask the reference
// for Tomo_struct fields and matcov tiles.  Nothing in the product links or loads it.
#include <cstring>
#include <string>
#include <vector>
#include <iostream>
#include <sstream>

#include "tomo_struct/tomo_struct.hpp"
#include "compute/matcov.hpp"
#include "compute/check_utils.hpp"

namespace {
Tomo_struct<double>* g_tomo = nullptr;

// Silence the reference's progress printing to stdout while it runs.
struct Mute {
    std::streambuf* old;
    std::ostringstream sink;
    Mute() : old(std::cout.rdbuf(sink.rdbuf())) {}
    ~Mute() { std::cout.rdbuf(old); }
};

template <typename T>
int copy_vec(const std::vector<T>& v, double* out, long cap) {
    long n = (long)v.size();
    if (out) {
        for (long i = 0; i < n && i < cap; ++i) out[i] = (double)v[i];
    }
    return (int)n;
}
}  // namespace

extern "C" {

// Build Tomo_struct<double>(sys_file, atm_file)  (common/tomo_struct/tomo_struct.cpp:11-17).
int ref_tomo_load(const char* sys_file, const char* atm_file) {
    Mute m;
    try {
        delete g_tomo;
        g_tomo = new Tomo_struct<double>(sys_file, atm_file);
    } catch (const std::exception& e) {
        std::cerr << "ref_tomo_load: " << e.what() << std::endl;
        g_tomo = nullptr;
        return -1;
    }
    return 0;
}

int ref_tomo_update_atm(const char* atm_file) {
    if (!g_tomo) return -1;
    Mute m;
    try {
        g_tomo->updateAtm(atm_file);
    } catch (const std::exception& e) {
        std::cerr << "ref_tomo_update_atm: " << e.what() << std::endl;
        return -1;
    }
    return 0;
}

int ref_tomo_set_target(int t) {
    if (!g_tomo) return -1;
    Mute m;
    g_tomo->setTarget(t);
    return 0;
}

// dims: [nW, nLgs, nNgs, nLayer, Nx, nmeas, nmeasts, nTarget]
int ref_tomo_dims(long* dims) {
    if (!g_tomo) return -1;
    dims[0] = g_tomo->sys.nW;
    dims[1] = g_tomo->sys.nLgs;
    dims[2] = g_tomo->sys.nNgs;
    dims[3] = g_tomo->atm.nLayer;
    dims[4] = g_tomo->Nx;
    dims[5] = g_tomo->getNMeas();
    dims[6] = g_tomo->getNMeasTS();
    dims[7] = g_tomo->sys.nTarget;
    return 0;
}

// Copy a named Tomo_struct vector into out (as double); returns its length, -1 if unknown.
int ref_tomo_field(const char* name, double* out, long cap) {
    if (!g_tomo) return -1;
    Tomo_struct<double>& t = *g_tomo;
    std::string f(name);
    if (f == "X") return copy_vec(t.X, out, cap);
    if (f == "Y") return copy_vec(t.Y, out, cap);
    if (f == "Nssp") return copy_vec(t.Nssp, out, cap);
    if (f == "Nsubap") return copy_vec(t.Nsubap, out, cap);
    if (f == "u") return copy_vec(t.u, out, cap);
    if (f == "v") return copy_vec(t.v, out, cap);
    if (f == "sspSizeL") return copy_vec(t.sspSizeL, out, cap);
    if (f == "L0diff") return copy_vec(t.L0diff, out, cap);
    if (f == "cn2") return copy_vec(t.atm.cn2, out, cap);
    if (f == "h") return copy_vec(t.atm.h, out, cap);
    if (f == "noiseNGS") return copy_vec(t.noiseNGS, out, cap);
    if (f == "noiseLGSxx") return copy_vec(t.noiseLGSxx, out, cap);
    if (f == "noiseLGSyy") return copy_vec(t.noiseLGSyy, out, cap);
    if (f == "noiseLGSxy") return copy_vec(t.noiseLGSxy, out, cap);
    if (f == "lgsExt") return copy_vec(t.lgsExt, out, cap);
    if (f == "lgsTheta") return copy_vec(t.lgsTheta, out, cap);
    if (f == "alphaX") return copy_vec(t.alphaX, out, cap);
    if (f == "alphaY") return copy_vec(t.alphaY, out, cap);
    if (f == "mr") return copy_vec(t.mr, out, cap);
    if (f == "sspSize") return copy_vec(t.sspSize, out, cap);
    if (f == "scalars") {
        std::vector<double> s = {t.sys.diam, t.sys.obs, t.sys.lgsCst, t.sNGS2, t.sLGS2,
                                 t.qr, t.bdw};
        return copy_vec(s, out, cap);
    }
    return -1;
}

// MOAO_COMMON::matcov_tile on the loaded Tomo_struct (common/compute/matcov.cpp:17-98).
int ref_matcov_tile(double* data, int nrows, int ncols, int xoffset, int yoffset, int lda,
                    int type_mat) {
    if (!g_tomo) return -1;
    if (type_mat < 1 || type_mat > 4) return -2;  // the reference would exit(1)
    MOAO_COMMON::matcov_tile<double>(data, nrows, ncols, xoffset, yoffset, lda, g_tomo,
                                     type_mat);
    return 0;
}

// compareMatrices (common/compute/check_utils.cpp:20-57), double/double instance.
void ref_compare_matrices(double* A, double* B, long nrows, long ncols, char uplo, char trans,
                          double* maxErr, double* frob) {
    double me = 0, fn = 0;
    compareMatrices<double, double>(A, B, nrows, ncols, uplo, trans, me, fn);
    *maxErr = me;
    *frob = fn;
}

}  // extern "C"
