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
// Cmm, then per target Ctm, Ctt (type 4), Cee/Cvv (all targets in one batched call) and the
// PSF (psf_oloop*_iloop*_gal*.fits in --out).  Matrices stay in device memory between stages;
// --out also receives R.fits (every target) and target 0's Cee.fits and Cvv.fits of the last
// step (FITS images of doubles, memory order, as writeFits).
// --mode mixed (FP32 updates) or bf16 (split-BF16 updates) solves each target with the
// mixed-precision reconstructor instead.  With --psf 1 the targets' PSFs of one observation
// step are computed in one batched call, and the long-exposure PSF of each target (the mean of
// its short-exposure PSFs over every refine and observation step, which the reference leaves
// as a TODO at chameleon/moao_chameleon.cpp:324-345) is written as
// psf_nbrefine{maxrefine}_maxobs{maxobs}_gal{g}.fits (--write_short 0 skips the short-exposure
// files).  The DM has Nssp(truth) + 1 actuators across, as in the reference, unless --dm_np N
// says otherwise (a DM finer than the truth sensor's grid).
//
// Check mode (lapack/check_moao_lapack.cpp, also selected by running the binary as
// check_moao): --sys_check DIR [--compare_all 0|1] [--write_all 0|1] [--accuracy 1e-7]
// [--out_path DIR] runs one pass Cmm -> Ctm -> R -> Cmm, Ctm, Ctt -> Cee/Cvv -> PSF on
// prof-1-atmos-night0.txt, compares each stage with DIR/{Cmm,Ctm,R,Ctt,Cee,Cvv,psf}.fits by
// compareMatrices (all of them with --compare_all 1, the PSF only otherwise), prints the
// reference's table and exits 1 when a maximum error exceeds the accuracy.
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
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
                                          {"tol", "1e-11"}, {"psf", "0"}, {"out", ""},
                                          {"compare_all", "0"}, {"write_all", "0"},
                                          {"accuracy", "1.e-7"}, {"out_path", ""},
                                          {"write_short", "1"}, {"dm_np", "0"}};
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

int mode_of(const std::string& m) {
    if (m == "fp64") return MOAO_FP64;
    if (m == "mixed") return MOAO_MIXED_FP32;
    if (m == "bf16") return MOAO_MIXED_BF16;
    std::fprintf(stderr, "moao: unknown --mode %s (fp64, mixed or bf16)\n", m.c_str());
    std::exit(2);
}

// One check pass (lapack/check_moao_lapack.cpp:28-273) over the C-ABI.
int run_check(Args& a) {
    const std::string sys = a.s("sys_path"), atm = a.s("atm_path"), chk = a.s("sys_check");
    const std::string outp = a.s("out_path");
    const bool compare_all = a.i("compare_all") != 0, write_all = a.i("write_all") != 0;
    const double accuracy = std::atof(a.s("accuracy").c_str());
    Timer t_Cmm, t_Ctm, t_Ctt, t_R, t_CeeCvv, t_psf, t_glob;
    t_glob.start();
    check(moao_hip_init(0), "init");
    moao_tomo* tomo = nullptr;
    if (moao_tomo_create((sys + "/sys-params.txt").c_str(), (atm + "/prof-1-atmos-night0.txt").c_str(),
                         &tomo))
        die("tomo");
    const long n = moao_tomo_nmeas(tomo), m = moao_tomo_nmeasts(tomo);
    std::printf("using %ld meas and %ld ts meas\n", n, m);
    long dims[6];
    moao_tomo_dims(tomo, dims);
    std::vector<double> Dx, idxd, a2f, otf;
    long nact = 0, dxm = 0, nidx = 0, one = 0, N = 0, ny = 0, ox = 0, oy = 0;
    const bool isample = fits_read(sys + "/Dx.fits", &nact, &dxm, &Dx) &&
                         fits_read(sys + "/idx.fits", &nidx, &one, &idxd) &&
                         fits_read(sys + "/abs2fi.fits", &N, &ny, &a2f) &&
                         fits_read(sys + "/otftel.fits", &ox, &oy, &otf);
    if (isample) std::printf("Using %ld actuators\n", nact);
    // matrices start zeroed (the reference callocs them: Cmm's strict upper part stays 0)
    auto zeroed = [&](size_t cnt) {
        auto d = std::make_unique<Dev>(std::max<size_t>(cnt, 1));
        std::vector<double> z(std::max<size_t>(cnt, 1), 0.0);
        check(moao_hip_memcpy_h2d(d->p, z.data(), z.size() * 8), "zero");
        return d;
    };
    auto dCmm = zeroed((size_t)n * n), dR = zeroed((size_t)m * n), dCtm = zeroed((size_t)m * n),
         dCtt = zeroed((size_t)m * m), dCee = zeroed((size_t)m * m),
         dCvv = zeroed((size_t)nact * nact), dDx = zeroed((size_t)nact * m),
         dTmp = zeroed((size_t)nact * m),
         dRef = zeroed((size_t)std::max(std::max(n * n, m * n), std::max(nact * nact, N * N)));
    std::map<std::string, std::pair<double, double>> err;  // stage -> (max, frobenius)
    auto compare = [&](const char* name, const double* d, long nr, long nc, bool always) {
        if (!outp.empty() && write_all) {
            std::vector<double> h((size_t)nr * nc);
            check(moao_hip_memcpy_d2h(h.data(), d, h.size() * 8), "download");
            fits_write(outp + "/" + name + ".fits", nr, nc, h.data());
        }
        if (!compare_all && !always) return;
        long rx = 0, ry = 0;
        std::vector<double> ref;
        if (!fits_read(chk + "/" + name + ".fits", &rx, &ry, &ref) || rx * ry != nr * nc)
            die(std::string("reference file ") + chk + "/" + name + ".fits missing or misshaped");
        check(moao_hip_memcpy_h2d(dRef->p, ref.data(), ref.size() * 8), "upload");
        double mx = 0, fr = 0;
        check(moao_hip_compare_matrices_dev(d, dRef->p, nr, nc, 'N', 'T', &mx, &fr, nullptr),
              "compareMatrices");
        err[name] = {mx, fr};
    };
    moao_tomo_view view;
    moao_tomo_get_view(tomo, &view);
    auto stage = [&](Timer& t, const char* label, auto&& body) {
        std::printf("%-8s", label);
        std::fflush(stdout);
        t.start();
        body();
        check(moao_hip_synchronize(), "sync");
        t.stop();
        std::printf("done\n");
    };
    stage(t_Cmm, "Cmm", [&] { check(moao_hip_matcov_tile_dev(dCmm->p, n, n, 0, 0, n, &view, 1, nullptr), "Cmm"); });
    compare("Cmm", dCmm->p, n, n, false);
    moao_tomo_set_target(tomo, 0);
    moao_tomo_get_view(tomo, &view);
    stage(t_Ctm, "Ctm", [&] { check(moao_hip_matcov_tile_dev(dR->p, m, n, 0, 0, m, &view, 3, nullptr), "Ctm"); });
    compare("Ctm", dR->p, m, n, false);
    stage(t_R, "R", [&] { check(moao_hip_reconstructor_dev(n, m, dCmm->p, n, dR->p, m, 1, nullptr), "R"); });
    compare("R", dR->p, m, n, false);
    // Cmm and Ctm again (destroyed by the reconstructor), then Ctt
    stage(t_Cmm, "Cmm", [&] { check(moao_hip_matcov_tile_dev(dCmm->p, n, n, 0, 0, n, &view, 1, nullptr), "Cmm"); });
    stage(t_Ctm, "Ctm", [&] { check(moao_hip_matcov_tile_dev(dCtm->p, m, n, 0, 0, m, &view, 3, nullptr), "Ctm"); });
    stage(t_Ctt, "Ctt", [&] { check(moao_hip_matcov_tile_dev(dCtt->p, m, m, 0, 0, m, &view, 4, nullptr), "Ctt"); });
    compare("Ctt", dCtt->p, m, m, false);
    if (isample) {
        check(moao_hip_memcpy_h2d(dDx->p, Dx.data(), Dx.size() * 8), "Dx");
        stage(t_CeeCvv, "Cee,Cvv", [&] {
            check(moao_hip_cee_cvv_dev(n, m, nact, dCmm->p, n, dCtt->p, m, dCtm->p, m, dR->p, m,
                                       dDx->p, nact, dCee->p, m, dCvv->p, nact, dTmp->p, nact,
                                       nullptr), "Cee_Cvv");
        });
        compare("Cee", dCee->p, m, m, false);
        compare("Cvv", dCvv->p, nact, nact, false);
        std::vector<double> ns(dims[0]);
        moao_tomo_field(tomo, "Nssp", ns.data(), dims[0]);
        std::vector<long> idx(idxd.begin(), idxd.end());
        moao_isample* is = nullptr;
        // FP64 transforms: the reference's FFTW arithmetic
        check(moao_hip_intersample_create(nidx, (long)ns[dims[0] - 1] + 1, N, idx.data(),
                                          a2f.data(), otf.data(), 1, &is), "intersample");
        N = moao_hip_intersample_n(is);
        auto dPsf = zeroed((size_t)N * N);
        stage(t_psf, "psf", [&] {
            check(moao_hip_intersample_process_dev(is, dCvv->p, 1, dPsf->p, nullptr), "psf");
        });
        if (!outp.empty()) {
            std::vector<double> h((size_t)N * N);
            check(moao_hip_memcpy_d2h(h.data(), dPsf->p, h.size() * 8), "psf");
            fits_write(outp + "/psf.fits", N, N, h.data());
        }
        compare("psf", dPsf->p, N, N, true);
        moao_hip_intersample_destroy(is);
    }
    t_glob.stop();
    moao_tomo_destroy(tomo);
    if (!compare_all) std::printf("WARNING: error only for psf!\n");
    auto e = [&](const char* k, int w) {
        auto it = err.find(k);
        return it == err.end() ? 0.0 : (w ? it->second.first : it->second.second);
    };
    std::printf("==============================\n");
    std::printf("operation: Cmm           Ctm           Ctt           R             Cee           "
                "Cvv           psf          total\n");
    std::printf("time (s) : %e  %e  %e  %e  %e                %e %e\n", t_Cmm.total,
                t_Ctm.total, t_Ctt.total, t_R.total, t_CeeCvv.total, t_psf.total, t_glob.total);
    std::printf("F norm   : %e  %e  %e  %e  %e  %e  %e    \n", e("Cmm", 0), e("Ctm", 0),
                e("Ctt", 0), e("R", 0), e("Cee", 0), e("Cvv", 0), e("psf", 0));
    std::printf("max err  : %e  %e  %e  %e  %e  %e  %e    \n", e("Cmm", 1), e("Ctm", 1),
                e("Ctt", 1), e("R", 1), e("Cee", 1), e("Cvv", 1), e("psf", 1));
    std::printf("==============================\n");
    for (const auto& kv : err)
        if (kv.second.first > accuracy) {
            std::fprintf(stderr, "test failed: bad accuracy\n\t expected %g\n", accuracy);
            return 1;
        }
    return 0;
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
    const std::string prog = argv[0];
    const bool as_check = prog.size() >= 10 && prog.compare(prog.size() - 10, 10, "check_moao") == 0;
    if (a.kv.count("sys_check") || as_check) {
        if (!a.kv.count("sys_path") || !a.kv.count("atm_path") || !a.kv.count("sys_check")) {
            std::fprintf(stderr, "usage: check_moao --sys_path DIR --atm_path DIR --sys_check DIR "
                                 "[--compare_all 0|1] [--write_all 0|1] [--accuracy 1e-7] "
                                 "[--out_path DIR]\n");
            return 2;
        }
        return run_check(a);
    }
    if (!a.kv.count("sys_path") || !a.kv.count("atm_path")) {
        std::fprintf(stderr,
                     "usage: moao --sys_path DIR --atm_path DIR [--maxrefine 1] [--maxobs 1] "
                     "[--snapPerNight 1] [--nact N] [--verbose 0] [--mode fp64|mixed|bf16] "
                     "[--tol 1e-11] [--psf 0|1] [--out DIR]\n");
        return 2;
    }
    const std::string sys = a.s("sys_path"), atm = a.s("atm_path"), out = a.s("out");
    const long maxrefine = a.i("maxrefine"), maxobs = a.i("maxobs"), spn = a.i("snapPerNight");
    const int mode = mode_of(a.s("mode"));
    const bool verbose = a.i("verbose") != 0, mixed = mode != MOAO_FP64;
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
        // actuators across the DM: the reference's rule Nssp(truth) + 1 (moao_lapack.cpp:64)
        // unless --dm_np gives the DM explicitly
        const long np = a.i("dm_np") > 0 ? a.i("dm_np") : (long)ns[dims[0] - 1] + 1;
        check(moao_hip_intersample_create(nidx, np, nx, idx.data(), a2f.data(), otf.data(), 0,
                                          &isample),
              "intersample");
        N = moao_hip_intersample_n(isample);
    }
    t_pre.stop();

    // every target's Ctm, Ctt, Cee, Cvv and Tmp side by side (block g at g block strides), so
    // one batched Cee/Cvv call serves all targets of an observation step
    Dev dCmm((size_t)n * n), dR((size_t)m * n * nTarget), dCtm((size_t)m * n * nTarget),
        dCtt((size_t)m * m * nTarget), dCee((size_t)m * m * nTarget),
        dCvv((size_t)std::max(nact, 1L) * std::max(nact, 1L) * nTarget),
        dDx((size_t)std::max(nact, 1L) * m), dTmp((size_t)std::max(nact, 1L) * m * nTarget),
        dPsf((size_t)std::max(N, 1L) * std::max(N, 1L) * nTarget);
    // long-exposure PSF of each target: the sum of its short-exposure PSFs, on the device
    std::unique_ptr<Dev> dCumul;
    if (isample) {
        dCumul.reset(new Dev((size_t)N * N * nTarget));
        std::vector<double> z((size_t)N * N * nTarget, 0.0);
        check(moao_hip_memcpy_h2d(dCumul->p, z.data(), z.size() * 8), "zero");
    }
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
                                                       &view, mode, tol, 8, &resid,
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
            t_matcov.start();
            for (long g = 0; g < nTarget; ++g) {
                moao_tomo_set_target(tomo, (int)g);
                moao_tomo_get_view(tomo, &view);
                check(moao_hip_matcov_tile_dev(dCtm.p + g * m * n, m, n, 0, 0, m, &view, 3,
                                               nullptr),
                      "Ctm");
                check(moao_hip_matcov_tile_dev(dCtt.p + g * m * m, m, m, 0, 0, m, &view, 4,
                                               nullptr),
                      "Ctt");
            }
            check(moao_hip_synchronize(), "sync");
            t_matcov.stop();
            // Cee/Cvv of every target (lapack/moao_lapack.cpp:165-175 calls it per target)
            t_CeeCvv.start();
            const long na = std::max(nact, 1L);
            check(moao_hip_cee_cvv_batched_dev(n, m, nact, nTarget, dCmm.p, n, dCtt.p, m, dCtm.p,
                                               m, dR.p, m, dDx.p, na, dCee.p, m, dCvv.p, na,
                                               dTmp.p, na, nullptr),
                  "Cee_Cvv");
            check(moao_hip_synchronize(), "sync");
            t_CeeCvv.stop();
            if (isample) {
                // every target's PSF of this step in one batched call
                t_psf.start();
                check(moao_hip_intersample_process_dev(isample, dCvv.p, (int)nTarget, dPsf.p,
                                                       nullptr),
                      "intersample");
                check(moao_hip_psf_accumulate_dev(dPsf.p, N * N * nTarget, dCumul->p, nullptr),
                      "psf accumulate");
                // short-exposure PSFs reach the host only when they are written
                std::vector<double> psf;
                if (!out.empty() && a.i("write_short")) {
                    psf.resize((size_t)N * N * nTarget);
                    check(moao_hip_memcpy_d2h(psf.data(), dPsf.p, psf.size() * 8), "psf");
                } else {
                    check(moao_hip_synchronize(), "sync");
                }
                t_psf.stop();
                if (!psf.empty())
                    for (long g = 0; g < nTarget; ++g)
                        fits_write(out + "/psf_oloop" + std::to_string(r) + "_iloop" +
                                       std::to_string(o) + "_gal" + std::to_string(g) + ".fits",
                                   N, N, psf.data() + (size_t)g * N * N);
            }
        }
    }
    t_glob.stop();
    if (isample && !out.empty() && maxrefine * maxobs > 0) {
        std::vector<double> cumul((size_t)N * N * nTarget);
        check(moao_hip_memcpy_d2h(cumul.data(), dCumul->p, cumul.size() * 8), "long exposure");
        const double w = 1.0 / (double)(maxrefine * maxobs);
        for (auto& v : cumul) v *= w;
        for (long g = 0; g < nTarget; ++g)
            fits_write(out + "/psf_nbrefine" + std::to_string(maxrefine) + "_maxobs" +
                           std::to_string(maxobs) + "_gal" + std::to_string(g) + ".fits",
                       N, N, cumul.data() + (size_t)g * N * N);
    }
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
                moao_version(), n, m, nact, dims[4], maxrefine, maxobs, a.s("mode").c_str(),
                t_pre.total, t_matcov.total, t_R.total, t_CeeCvv.total, t_psf.total,
                t_glob.total);
    return 0;
}
