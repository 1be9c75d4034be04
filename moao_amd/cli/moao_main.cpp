// moao: the reference's refine / observation loop (lapack/moao_lapack.cpp:30-188) on the
// MI355X path, through the C-ABI only.
//
//   moao --sys_path DIR --atm_path DIR [--maxrefine 1] [--maxobs 1] [--snapPerNight 1]
//        [--nact N] [--verbose 0] [--mode fp64|mixed] [--tol 1e-11] [--psf 0|1] [--out DIR]
//
// Files and loop are the reference's: DIR/sys-params.txt, profiles
// prof{snap*snapPerNight+obs}-atmos-night0.txt (obs = -1 for the refine step), Dx.fits (nact x
// nmeasts) when present, idx.fits / abs2fi.fits / otftel.fits for the PSF.  Refine step:
// matcov Cmm (type 1), Ctm of every target into R (type 3), reconstructor.  Observation step:
// Cmm, then per target Ctm, Ctt (type 4), Cee/Cvv and the PSF (psf_oloop*_iloop*_gal*.fits in
// --out).  Matrices stay in device memory between stages; --out also receives R.fits,
// Cee.fits and Cvv.fits of the last step (FITS images of doubles, memory order, as writeFits).
// --mode mixed solves each target with the mixed-precision reconstructor instead.
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "moao_hip.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Timer {
    double total = 0;
    Clock::time_point t0;
    void start() { t0 = Clock::now(); }
    void stop() { total += std::chrono::duration<double>(Clock::now() - t0).count(); }
};

[[noreturn]] void die(const std::string& what) {
    std::fprintf(stderr, "moao: %s: %s\n", what.c_str(), moao_last_error());
    std::exit(1);
}

void check(int rc, const char* what) {
    if (rc < 0) die(what);
    if (rc > 0) {
        std::printf("An error occured in dpotrf: %d\n", rc);
        std::exit(1);
    }
}

// ---- FITS primary images (the layout of common/io/fits.cpp writeFits / readFits)
void fits_write(const std::string& path, long nx, long ny, const double* data) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) {
        std::fprintf(stderr, "moao: cannot write %s\n", path.c_str());
        std::exit(1);
    }
    std::string hdr;
    auto card = [&](const std::string& k, const std::string& v) {
        char c[81];
        std::snprintf(c, sizeof c, "%-8s= %20s", k.c_str(), v.c_str());
        std::string s(c);
        s.resize(80, ' ');
        hdr += s;
    };
    card("SIMPLE", "T");
    card("BITPIX", "-64");
    card("NAXIS", "2");
    card("NAXIS1", std::to_string(nx));
    card("NAXIS2", std::to_string(ny));
    card("EXTEND", "T");
    std::string end = "END";
    end.resize(80, ' ');
    hdr += end;
    hdr.resize((hdr.size() + 2879) / 2880 * 2880, ' ');
    std::fwrite(hdr.data(), 1, hdr.size(), f);
    const size_t n = (size_t)nx * ny;
    std::vector<uint64_t> be(n);
    for (size_t i = 0; i < n; ++i) {
        uint64_t u;
        std::memcpy(&u, data + i, 8);
        be[i] = __builtin_bswap64(u);
    }
    std::fwrite(be.data(), 8, n, f);
    const size_t pad = (2880 - (n * 8) % 2880) % 2880;
    std::vector<char> z(pad, 0);
    std::fwrite(z.data(), 1, pad, f);
    std::fclose(f);
}

// Reads a 2-D image (BITPIX -64, -32, 64, 32, 16 or 8) into doubles; false if absent.
bool fits_read(const std::string& path, long* nx, long* ny, std::vector<double>* out) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::map<std::string, std::string> keys;
    char card[81] = {0};
    size_t nread = 0;
    for (;;) {
        if (std::fread(card, 1, 80, f) != 80) die("truncated FITS header in " + path);
        nread += 80;
        std::string c(card, 80), k = c.substr(0, 8);
        k.erase(k.find_last_not_of(' ') + 1);
        if (k == "END") break;
        if (c.compare(8, 2, "= ") == 0) {
            std::string v = c.substr(10);
            v = v.substr(0, v.find('/'));
            v.erase(0, v.find_first_not_of(' '));
            v.erase(v.find_last_not_of(' ') + 1);
            keys[k] = v;
        }
    }
    std::fseek(f, (long)((nread + 2879) / 2880 * 2880), SEEK_SET);
    const int bitpix = std::atoi(keys["BITPIX"].c_str());
    *nx = std::atol(keys["NAXIS1"].c_str());
    *ny = keys.count("NAXIS2") ? std::atol(keys["NAXIS2"].c_str()) : 1;
    const size_t n = (size_t)(*nx) * (*ny), w = std::abs(bitpix) / 8;
    std::vector<unsigned char> raw(n * w);
    if (std::fread(raw.data(), 1, raw.size(), f) != raw.size()) die("truncated FITS data in " + path);
    std::fclose(f);
    out->resize(n);
    for (size_t i = 0; i < n; ++i) {
        uint64_t u = 0;
        for (size_t b = 0; b < w; ++b) u = (u << 8) | raw[i * w + b];  // big-endian
        double v;
        if (bitpix == -64) {
            std::memcpy(&v, &u, 8);
        } else if (bitpix == -32) {
            uint32_t u32 = (uint32_t)u;
            float fv;
            std::memcpy(&fv, &u32, 4);
            v = fv;
        } else if (bitpix == 8) {
            v = (double)u;
        } else {  // signed integers
            const int sh = 64 - 8 * (int)w;
            v = (double)((int64_t)(u << sh) >> sh);
        }
        (*out)[i] = v;
    }
    return true;
}

struct Args {
    std::map<std::string, std::string> kv{{"maxrefine", "1"}, {"maxobs", "1"},
                                          {"snapPerNight", "1"}, {"nact", "0"},
                                          {"verbose", "0"}, {"mode", "fp64"},
                                          {"tol", "1e-11"}, {"psf", "0"}, {"out", ""}};
    std::string s(const char* k) { return kv[k]; }
    long i(const char* k) { return std::atol(kv[k].c_str()); }
};

struct Dev {
    double* p = nullptr;
    explicit Dev(size_t n) {
        if (moao_hip_malloc(reinterpret_cast<void**>(&p), n * sizeof(double))) die("malloc");
    }
    ~Dev() { moao_hip_free(p); }
};

std::string prof(const std::string& atm, long idx) {
    return atm + "/prof" + std::to_string(idx) + "-atmos-night0.txt";
}

}  // namespace

int main(int argc, char** argv) {
    Args a;
    for (int k = 1; k + 1 < argc; k += 2) {
        if (std::strncmp(argv[k], "--", 2) != 0) {
            std::fprintf(stderr, "moao: unexpected argument %s\n", argv[k]);
            return 2;
        }
        a.kv[argv[k] + 2] = argv[k + 1];
    }
    if (!a.kv.count("sys_path") || !a.kv.count("atm_path")) {
        std::fprintf(stderr,
                     "usage: moao --sys_path DIR --atm_path DIR [--maxrefine 1] [--maxobs 1] "
                     "[--snapPerNight 1] [--nact N] [--verbose 0] [--mode fp64|mixed] "
                     "[--tol 1e-11] [--psf 0|1] [--out DIR]\n");
        return 2;
    }
    const std::string sys = a.s("sys_path"), atm = a.s("atm_path"), out = a.s("out");
    const long maxrefine = a.i("maxrefine"), maxobs = a.i("maxobs"), spn = a.i("snapPerNight");
    const bool verbose = a.i("verbose") != 0, mixed = a.s("mode") == "mixed";
    const double tol = std::atof(a.s("tol").c_str());
    Timer t_matcov, t_R, t_CeeCvv, t_psf, t_pre, t_glob;
    t_glob.start();
    t_pre.start();
    check(moao_hip_init(0), "init");
    moao_tomo* tomo = nullptr;
    if (moao_tomo_create((sys + "/sys-params.txt").c_str(), prof(atm, -1).c_str(), &tomo))
        die("tomo");
    long dims[6];
    moao_tomo_dims(tomo, dims);
    const long nTarget = dims[3];
    const long n = moao_tomo_nmeas(tomo), m = moao_tomo_nmeasts(tomo);
    // Dx (nact x nmeasts) from Dx.fits when present (its first axis gives nact)
    std::vector<double> Dx;
    long nact = a.i("nact"), dxn = 0, dxm = 0;
    if (fits_read(sys + "/Dx.fits", &dxn, &dxm, &Dx)) {
        nact = dxn;
        if (dxm != m) die("Dx.fits does not have nmeasts columns");
    } else {
        Dx.assign((size_t)nact * m, 0.0);
    }
    std::printf("using %ld meas and %ld ts meas\nUsing %ld actuators\n", n, m, nact);
    moao_isample* isample = nullptr;
    long N = 0;
    if (a.i("psf")) {
        long nidx, one, nx, ny, ox, oy;
        std::vector<double> idxd, a2f, otf;
        if (!fits_read(sys + "/idx.fits", &nidx, &one, &idxd) ||
            !fits_read(sys + "/abs2fi.fits", &nx, &ny, &a2f) ||
            !fits_read(sys + "/otftel.fits", &ox, &oy, &otf)) {
            std::fprintf(stderr, "moao: --psf 1 needs idx.fits, abs2fi.fits, otftel.fits\n");
            return 1;
        }
        std::vector<double> ns(dims[0]);
        moao_tomo_field(tomo, "Nssp", ns.data(), dims[0]);
        std::vector<long> idx(idxd.begin(), idxd.end());
        check(moao_hip_intersample_create(nidx, (long)ns[dims[0] - 1] + 1, nx, idx.data(),
                                          a2f.data(), otf.data(), 0, &isample),
              "intersample");
        N = moao_hip_intersample_n(isample);
    }
    t_pre.stop();

    Dev dCmm((size_t)n * n), dR((size_t)m * n * nTarget), dCtm((size_t)m * n),
        dCtt((size_t)m * m), dCee((size_t)m * m), dCvv((size_t)std::max(nact, 1L) * std::max(nact, 1L)),
        dDx((size_t)std::max(nact, 1L) * m), dTmp((size_t)std::max(nact, 1L) * m),
        dPsf((size_t)std::max(N, 1L) * std::max(N, 1L));
    if (nact) check(moao_hip_memcpy_h2d(dDx.p, Dx.data(), Dx.size() * 8), "Dx upload");
    moao_tomo_view view;

    for (long r = 0; r < maxrefine; ++r) {
        if (verbose) std::printf("\nnbrefine%ld\n", r + 1);
        check(moao_tomo_update_atm(tomo, prof(atm, r * spn - 1).c_str()), "updateAtm");
        moao_tomo_get_view(tomo, &view);
        t_matcov.start();
        check(moao_hip_matcov_tile_dev(dCmm.p, n, n, 0, 0, n, &view, 1, nullptr), "matcov Cmm");
        for (long g = 0; g < nTarget; ++g) {
            moao_tomo_set_target(tomo, (int)g);
            moao_tomo_get_view(tomo, &view);
            check(moao_hip_matcov_tile_dev(dR.p + g * n * m, m, n, 0, 0, m, &view, 3, nullptr),
                  "matcov Ctm");
        }
        check(moao_hip_synchronize(), "sync");
        t_matcov.stop();
        t_R.start();
        if (!mixed) {
            check(moao_hip_reconstructor_dev(n, m, dCmm.p, n, dR.p, m, nTarget, nullptr),
                  "reconstructor");
        } else {
            for (long g = 0; g < nTarget; ++g) {
                double resid;
                int it;
                check(moao_hip_reconstructor_mixed_dev(n, m, dCmm.p, n, dR.p + g * n * m, m,
                                                       &view, MOAO_MIXED_FP32, tol, 8, &resid,
                                                       &it, nullptr),
                      "reconstructor (mixed)");
                if (verbose) std::printf("target %ld: %d refinement steps, residual %.3e\n", g,
                                         it, resid);
            }
        }
        check(moao_hip_synchronize(), "sync");
        t_R.stop();

        for (long o = 0; o < maxobs; ++o) {
            if (verbose) std::printf(" nbobs %ld\n", o + 1);
            check(moao_tomo_update_atm(tomo, prof(atm, r * spn + o).c_str()), "updateAtm");
            moao_tomo_get_view(tomo, &view);
            t_matcov.start();
            check(moao_hip_matcov_tile_dev(dCmm.p, n, n, 0, 0, n, &view, 1, nullptr), "Cmm");
            check(moao_hip_synchronize(), "sync");
            t_matcov.stop();
            for (long g = 0; g < nTarget; ++g) {
                moao_tomo_set_target(tomo, (int)g);
                moao_tomo_get_view(tomo, &view);
                t_matcov.start();
                check(moao_hip_matcov_tile_dev(dCtm.p, m, n, 0, 0, m, &view, 3, nullptr), "Ctm");
                check(moao_hip_matcov_tile_dev(dCtt.p, m, m, 0, 0, m, &view, 4, nullptr), "Ctt");
                check(moao_hip_synchronize(), "sync");
                t_matcov.stop();
                t_CeeCvv.start();
                check(moao_hip_cee_cvv_dev(n, m, nact, dCmm.p, n, dCtt.p, m, dCtm.p, m,
                                           dR.p + g * n * m, m, dDx.p, std::max(nact, 1L),
                                           dCee.p, m, dCvv.p, std::max(nact, 1L), dTmp.p,
                                           std::max(nact, 1L), nullptr),
                      "Cee_Cvv");
                check(moao_hip_synchronize(), "sync");
                t_CeeCvv.stop();
                if (isample) {
                    t_psf.start();
                    check(moao_hip_intersample_process_dev(isample, dCvv.p, 1, dPsf.p, nullptr),
                          "intersample");
                    std::vector<double> psf((size_t)N * N);
                    check(moao_hip_memcpy_d2h(psf.data(), dPsf.p, psf.size() * 8), "psf");
                    t_psf.stop();
                    if (!out.empty())
                        fits_write(out + "/psf_oloop" + std::to_string(r) + "_iloop" +
                                       std::to_string(o) + "_gal" + std::to_string(g) + ".fits",
                                   N, N, psf.data());
                }
            }
        }
    }
    t_glob.stop();
    if (!out.empty()) {
        std::vector<double> h((size_t)m * n * nTarget);
        check(moao_hip_memcpy_d2h(h.data(), dR.p, h.size() * 8), "R");
        fits_write(out + "/R.fits", m, n * nTarget, h.data());
        if (maxobs > 0) {
            h.resize((size_t)m * m);
            check(moao_hip_memcpy_d2h(h.data(), dCee.p, h.size() * 8), "Cee");
            fits_write(out + "/Cee.fits", m, m, h.data());
            if (nact) {
                h.resize((size_t)nact * nact);
                check(moao_hip_memcpy_d2h(h.data(), dCvv.p, h.size() * 8), "Cvv");
                fits_write(out + "/Cvv.fits", nact, nact, h.data());
            }
        }
    }
    if (isample) moao_hip_intersample_destroy(isample);
    moao_tomo_destroy(tomo);
    std::printf("========================================\nsimu characteristics\n"
                "ngpus    : 1 (%s)\nnmeas    : %ld\nnmeasts  : %ld\nnact     : %ld\n"
                "nLayers  : %ld\nmaxrefine: %ld\nmaxobs   : %ld\nmode     : %s\n"
                "========================================\ntimes (s)\n"
                "preprocess: %.3f\nmatcov    : %.3f\nR         : %.3f\nCee_Cvv   : %.3f\n"
                "psf       : %.3f\ntotal     : %.3f\n",
                moao_version(), n, m, nact, dims[4], maxrefine, maxobs, mixed ? "mixed" : "fp64",
                t_pre.total, t_matcov.total, t_R.total, t_CeeCvv.total, t_psf.total,
                t_glob.total);
    return 0;
}
