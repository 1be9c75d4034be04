// A reference-style caller of the MOAO_HIP:: drop-in layer (include/moao_hip.hpp).
//
// TomoLike has the member names and types the reference's Tomo_struct<double> exposes
// (common/tomo_struct/tomo_struct.hpp:9-94: sys.nW, sys.nLgs, sys.diam, sys.lgsCst, atm.nLayer,
// atm.cn2, Nx, Nsubap, Nssp, sspSizeL, u, v, X, Y, L0diff, noise*), filled here from the
// library's own system model.  The calls are the reference driver's refine step
// (lapack/moao_lapack.cpp:108-121): matcov Cmm (type 1), Ctm (type 3), reconstructor.
//
//   dropin <sys-params> <atm-params> <R.bin>      writes R (m x n, column-major doubles)
//   dropin --psf <dir> <nact> <np> <Cvv.bin> <psf.bin> [f32]
//                                                 the reference's intersample sequence
//                                                 (lapack/moao_lapack.cpp:57-66, 177-181) on
//                                                 dir/idx.fits, abs2fi.fits, otftel.fits; Cvv
//                                                 read as doubles (passed as float with f32)
//   dropin --link-only                            checks the program links, no GPU use
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "moao_hip.hpp"

struct SysLike {
    int nW = 0, nLgs = 0;
    double diam = 0, lgsCst = 0;
};
struct AtmLike {
    int nLayer = 0;
    std::vector<double> cn2;
};
struct TomoLike {
    SysLike sys;
    AtmLike atm;
    int Nx = 0;
    std::vector<long> Nsubap, Nssp;
    std::vector<double> sspSizeL, u, v, X, Y, L0diff, noiseNGS, noiseLGSxx, noiseLGSyy, noiseLGSxy;
};

static std::vector<double> field(moao_tomo* t, const char* name) {
    const long n = moao_tomo_field(t, name, nullptr, 0);
    std::vector<double> out(n > 0 ? n : 0);
    if (n > 0) moao_tomo_field(t, name, out.data(), n);
    return out;
}

int main(int argc, char** argv) {
    if (argc == 2 && std::strcmp(argv[1], "--link-only") == 0) {
        std::printf("linked %s\n", moao_version());
        return 0;
    }
    if (argc >= 7 && std::strcmp(argv[1], "--psf") == 0) {
        const long nact = std::atol(argv[3]), np = std::atol(argv[4]);
        std::vector<double> cvv((size_t)nact * nact);
        FILE* f = std::fopen(argv[5], "rb");
        if (!f || std::fread(cvv.data(), sizeof(double), cvv.size(), f) != cvv.size()) return 1;
        std::fclose(f);
        struct MOAO_HIP::isample_struct isample;
        MOAO_HIP::intersample_prepare(&isample, nact * nact, np, 39.0f, argv[2]);
        if (!MOAO_HIP::intersample_init(&isample)) return 1;
        int ok;
        if (argc == 8 && std::strcmp(argv[7], "f32") == 0) {
            std::vector<float> c32(cvv.begin(), cvv.end());
            ok = MOAO_HIP::intersample_process(&isample, c32.data());
        } else {
            ok = MOAO_HIP::intersample_process(&isample, cvv.data());
        }
        if (!ok) return 1;
        FILE* g = std::fopen(argv[6], "wb");
        if (!g) return 1;
        std::fwrite(isample.dphi, sizeof(double), (size_t)isample.N * isample.N, g);
        std::fclose(g);
        std::printf("N=%ld dactupix=%g\n", isample.N, isample.dactupix);
        MOAO_HIP::intersample_free(&isample);
        return 0;
    }
    if (argc != 4) {
        std::fprintf(stderr, "usage: dropin <sys-params> <atm-params> <R.bin>\n");
        return 2;
    }
    moao_tomo* h = nullptr;
    if (moao_tomo_create(argv[1], argv[2], &h) || moao_tomo_set_target(h, 0)) {
        std::fprintf(stderr, "tomo: %s\n", moao_last_error());
        return 1;
    }
    moao_tomo_view view;
    moao_tomo_get_view(h, &view);
    TomoLike tomo;
    tomo.sys.nW = view.Nw;
    tomo.sys.nLgs = view.nlgs;
    tomo.sys.diam = view.diam;
    tomo.sys.lgsCst = view.lgs_cst;
    tomo.atm.nLayer = view.Nlayer;
    tomo.atm.cn2 = field(h, "cn2");
    tomo.Nx = (int)view.Nx;
    for (double d : field(h, "Nsubap")) tomo.Nsubap.push_back((long)d);
    for (double d : field(h, "Nssp")) tomo.Nssp.push_back((long)d);
    tomo.sspSizeL = field(h, "sspSizeL");
    tomo.u = field(h, "u");
    tomo.v = field(h, "v");
    tomo.X = field(h, "X");
    tomo.Y = field(h, "Y");
    tomo.L0diff = field(h, "L0");
    tomo.noiseNGS = field(h, "noiseNGS");
    tomo.noiseLGSxx = field(h, "noiseLGSxx");
    tomo.noiseLGSyy = field(h, "noiseLGSyy");
    tomo.noiseLGSxy = field(h, "noiseLGSxy");
    const long nmeas = moao_tomo_nmeas(h), nmeasts = moao_tomo_nmeasts(h);
    moao_tomo_destroy(h);

    std::vector<double> Cmm((size_t)nmeas * nmeas, 0.0), R((size_t)nmeasts * nmeas, 0.0);
    MOAO_HIP::matcov_tile(Cmm.data(), (int)nmeas, (int)nmeas, 0, 0, (int)nmeas, &tomo, 1);
    MOAO_HIP::matcov_tile(R.data(), (int)nmeasts, (int)nmeas, 0, 0, (int)nmeasts, &tomo, 3);
    MOAO_HIP::reconstructor(nmeas, nmeasts, Cmm.data(), nmeas, R.data(), nmeasts, 1L);
    if (MOAO_HIP::last_status() != 0) return 1;
    FILE* f = std::fopen(argv[3], "wb");
    if (!f) return 1;
    std::fwrite(R.data(), sizeof(double), R.size(), f);
    std::fclose(f);
    std::printf("n=%ld m=%ld status=%d\n", nmeas, nmeasts, MOAO_HIP::last_status());
    return 0;
}
