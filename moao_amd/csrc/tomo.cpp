// Host-side system model.  Each member follows the reference function named in its comment
// (file:line in common/tomo_struct/tomo_struct.cpp unless stated otherwise); the arithmetic
// is kept in the reference's order so the generated u, v, noise vectors agree to the last bit.
#include "tomo.hpp"

#include <cmath>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace moao {

namespace {

constexpr double kRasc = 206265.0;

// One "comment line + value line" record (common/io/named.cpp:19-31, IO.cpp.in:35-44).
class ParamReader {
  public:
    explicit ParamReader(const std::string& path) : path_(path) {
        std::ifstream in(path);
        if (!in.is_open()) throw std::runtime_error("Could not open file '" + path + "'!");
        std::string line;
        while (std::getline(in, line)) lines_.push_back(line);
    }
    template <typename T>
    std::vector<T> vec(const char* name) {
        std::istringstream ss(next(name));
        std::vector<T> out;
        T x;
        while (ss >> x) out.push_back(x);
        return out;
    }
    template <typename T>
    T scalar(const char* name) {
        std::vector<T> v = vec<T>(name);
        if (v.empty()) throw std::runtime_error(path_ + ": no value for " + name);
        return v[0];
    }

  private:
    std::string next(const char* name) {
        pos_ += 1;  // comment line
        if (pos_ >= lines_.size()) throw std::runtime_error(path_ + ": truncated before " + name);
        return lines_[pos_++];
    }
    std::string path_;
    std::vector<std::string> lines_;
    size_t pos_ = 0;
};

template <typename T>
void need(const std::vector<T>& v, size_t n, const char* name) {
    if (v.size() < n)
        throw std::runtime_error(std::string("parameter vector too short: ") + name);
}

// noise.cpp:16-31
double flux(double qr, double mr, double throughput, double d, double t, double bdw) {
    return std::pow(10.0, qr - 0.4 * mr) * throughput * d * d * t * bdw * 1.e4;
}
// noise.cpp:36-52
double magnitude_from_flux(double qr, double F, double throughput, double d, double t, double bdw) {
    return -(std::log10(F / (throughput * d * d * bdw * 1.e4 * t)) - qr) / 0.4;
}

}  // namespace

// Field order of SysParams::getParams (common/tomo_struct/sysParams.cpp:165-207).
SysParams read_sys_params(const std::string& path) {
    ParamReader r(path);
    SysParams s;
    s.diam = r.scalar<double>("diam");
    s.obs = r.scalar<double>("obs");
    s.tFrame = r.scalar<double>("tFrame");
    s.nW = r.scalar<int>("nW");
    s.nLgs = r.scalar<int>("nLgs");
    s.nTarget = r.scalar<int>("nTarget");
    s.nssp = r.scalar<int>("nssp");
    s.gsAlt = r.vec<double>("gsAlt");
    s.gsType = r.vec<int>("gsType");
    s.alphaX_as = r.vec<double>("alphaX_as");
    s.alphaY_as = r.vec<double>("alphaY_as");
    s.XPup = r.vec<double>("XPup");
    s.YPup = r.vec<double>("YPup");
    s.thetaML = r.vec<double>("thetaML");
    s.thetaCam = r.vec<double>("thetaCam");
    s.sensibilite = r.vec<double>("sensibilite");
    s.tracking = r.vec<double>("tracking");
    s.pasDphi = r.scalar<double>("pasDphi");
    s.ncpu = r.scalar<int>("ncpu");
    s.mrNGS = r.vec<double>("mrNGS");
    s.lgsFlux = r.scalar<double>("lgsFlux");
    s.ngsPixSize = r.scalar<double>("ngsPixSize");
    s.lgsPixSize = r.scalar<double>("lgsPixSize");
    s.lambdaNGS = r.scalar<double>("lambdaNGS");
    s.lambdaLGS = r.scalar<double>("lambdaLGS");
    s.bdw_m = r.scalar<double>("bdw_m");
    s.throughNGS = r.scalar<double>("throughNGS");
    s.throughLGS = r.scalar<double>("throughLGS");
    s.throughAtm = r.scalar<double>("throughAtm");
    s.RON = r.scalar<int>("RON");
    s.lgsCst = r.scalar<double>("lgsCst");
    s.spotWidth = r.scalar<double>("spotWidth");
    s.lgsAlt = r.scalar<double>("lgsAlt");
    s.lgsDepth = r.scalar<double>("lgsDepth");
    s.targetX_as = r.vec<double>("targetX_as");
    s.targetY_as = r.vec<double>("targetY_as");
    s.nNgs = s.nW - s.nLgs - 1;  // sysParams.cpp:146
    if (s.nW < 2 || s.nLgs < 0 || s.nNgs < 0 || s.nssp < 1 || s.nTarget < 1)
        throw std::runtime_error(path + ": inconsistent WFS counts");
    // SysParams::vectorSize (sysParams.cpp:113-132)
    need(s.gsAlt, s.nW, "gsAlt");
    need(s.gsType, s.nW, "type");
    need(s.alphaX_as, s.nW, "alphaX_as");
    need(s.alphaY_as, s.nW, "alphaY_as");
    need(s.XPup, s.nW, "XPup");
    need(s.YPup, s.nW, "YPup");
    need(s.thetaML, s.nW, "thetaML");
    need(s.thetaCam, s.nW, "thetaCam");
    need(s.sensibilite, s.nW, "sensibilite");
    need(s.tracking, 3, "tracking");
    need(s.mrNGS, s.nW - s.nLgs, "mrNGS");
    need(s.targetX_as, s.nTarget, "targetX_as");
    need(s.targetY_as, s.nTarget, "targetY_as");
    return s;
}

// Field order of AtmParams::getParams (common/tomo_struct/atmParams.cpp:93-101).
AtmParams read_atm_params(const std::string& path) {
    ParamReader r(path);
    AtmParams a;
    a.nLayer = r.scalar<int>("nLayer");
    a.r0 = r.scalar<double>("r0");
    a.cn2 = r.vec<double>("cn2");
    a.h = r.vec<double>("h");
    a.L0 = r.vec<double>("L0");
    if (a.nLayer < 1) throw std::runtime_error(path + ": nLayer < 1");
    need(a.cn2, a.nLayer, "cn2");
    need(a.h, a.nLayer, "h");
    need(a.L0, a.nLayer, "L0");
    return a;
}

Tomo::Tomo(const std::string& sys_file, const std::string& atm_file)
    : Tomo(read_sys_params(sys_file), read_atm_params(atm_file)) {}

Tomo::Tomo(const SysParams& s, const AtmParams& a) : sys(s) {
    init_sys();
    update_atm(a);
}

// getSysParams (90-133)
void Tomo::init_sys() {
    const int nW = sys.nW;
    qr = 2.9;
    bdw = sys.bdw_m * 1.e10;
    Nssp.assign(nW, sys.nssp);
    pixSize.assign(nW, sys.ngsPixSize);
    throughput.assign(nW, sys.throughNGS);
    for (int w = 0; w < sys.nLgs; ++w) {
        pixSize[w] = sys.lgsPixSize;
        throughput[w] = sys.throughLGS;
    }
    diamPup.assign(Nssp.begin(), Nssp.end());
    alphaX.resize(nW);
    alphaY.resize(nW);
    for (int w = 0; w < nW; ++w) {
        alphaX[w] = (1 / kRasc) * sys.alphaX_as[w];
        alphaY[w] = (1 / kRasc) * sys.alphaY_as[w];
    }
    targetX.resize(sys.targetX_as.size());
    targetY.resize(sys.targetY_as.size());
    for (size_t t = 0; t < targetX.size(); ++t) targetX[t] = (1 / kRasc) * sys.targetX_as[t];
    for (size_t t = 0; t < targetY.size(); ++t) targetY[t] = (1 / kRasc) * sys.targetY_as[t];
    sspSize.resize(nW);
    for (int w = 0; w < nW; ++w) sspSize[w] = sys.diam / (int)Nssp[w];
    mr.assign(nW, 0.0);
    for (int i = 0; i < nW - sys.nLgs; ++i) mr[sys.nLgs + i] = sys.mrNGS[i];
    const double fpspf = sys.lgsFlux * sys.tFrame * std::pow(sys.diam / Nssp[0], 2.) *
                         throughput[0] / sys.throughAtm;
    for (int w = 0; w < sys.nLgs; ++w)
        mr[w] = magnitude_from_flux(qr, fpspf, throughput[w], sspSize[w], sys.tFrame, bdw);
    valid_subap();
    generate_elongation();
}

// validSubap (379-471): centres of an nssp x nssp grid kept when inside the annulus.
void Tomo::valid_subap() {
    Nsubap.assign(sys.nW, 0);
    X.clear();
    Y.clear();
    const double bornemin = -sys.diam / 2.;
    const double rtel2 = (sys.diam * sys.diam) / 4.;
    for (int n = 0; n < sys.nW; ++n) {
        const long nssp = Nssp[n];
        const double pas = sys.diam / (1. * nssp);
        double robs2;
        if (nssp != 7 || (sys.obs <= 0.285 || sys.obs >= 0.29))
            robs2 = sys.diam * sys.obs / 2. * sys.diam * sys.obs / 2.;
        else
            robs2 = sys.diam * 0.285 / 2. * sys.diam * 0.285 / 2.;
        if (nssp == 1) {
            X.push_back(0.);
            Y.push_back(0.);
            Nsubap[n] = 1;
            continue;
        }
        long nsap = 0;
        for (long i = 0; i < nssp; ++i) {
            const double tp = bornemin + pas / 2. * (2. * i + 1.);
            for (long j = 0; j < nssp; ++j) {
                const double x = bornemin + pas / 2. * (2. * j + 1.);
                const double r2 = x * x + tp * tp;
                if (!(r2 < robs2 || r2 >= rtel2)) {
                    X.push_back(x);
                    Y.push_back(tp);
                    ++nsap;
                }
            }
        }
        Nsubap[n] = nsap;
    }
    Nx = (long)X.size();
}

// generateElongation (477-495)
void Tomo::generate_elongation() {
    lgsExt.assign(Nx, 0.0);
    lgsTheta.assign(Nx, 0.0);
    const double to_fwhm = kRasc * sys.lgsDepth / (sys.lgsAlt * sys.lgsAlt);
    long k = 0;
    for (int w = 0; w < sys.nLgs; ++w) {
        const double den = std::sqrt(alphaX[w] * alphaX[w] + alphaY[w] * alphaY[w]);
        const double lltx = alphaX[w] * (sys.diam / 2. / den);
        const double llty = alphaY[w] * (sys.diam / 2. / den);
        for (long i = 0; i < Nsubap[w]; ++i, ++k) {
            lgsExt[k] = std::sqrt(std::pow(X[k] - lltx, 2) + std::pow(Y[k] - llty, 2)) * to_fwhm;
            lgsTheta[k] = std::atan2((Y[k] - llty) * to_fwhm, (X[k] - lltx) * to_fwhm);
        }
    }
}

void Tomo::update_atm(const std::string& atm_file) { update_atm(read_atm_params(atm_file)); }

// updateAtm (212-262)
void Tomo::update_atm(const AtmParams& a) {
    atm = a;
    const int L = atm.nLayer;
    sspSizeL.assign((size_t)sys.nW * L, 0.0);
    for (int cc = 0; cc < sys.nW * L; ++cc) {
        const int n = cc / L, l = cc - n * L;
        sspSizeL[cc] = sspSize[n] * (1. - sys.gsAlt[n] * atm.h[l]);
    }
    L0diff.assign(atm.L0.begin(), atm.L0.begin() + L);
    cn2.assign(atm.cn2.begin(), atm.cn2.begin() + L);
    projected_subap();
    update_noise();
}

// setTarget / setTargetCoord (275-303)
void Tomo::set_target(int t) {
    if (t < 0 || t >= (int)targetX.size() || t >= (int)targetY.size())
        throw std::runtime_error("target index out of range");
    set_target_coord(targetX[t], targetY[t], false);
}

void Tomo::set_target_coord(double ax, double ay, bool to_rad) {
    if (to_rad) {
        ax /= kRasc;
        ay /= kRasc;
    }
    alphaX[sys.nW - 1] = ax;
    alphaY[sys.nW - 1] = ay;
    projected_subap();
    update_noise();
}

// updateSeeing (313-332)
void Tomo::update_seeing() {
    const double lambda = 500. * 1.e-9;
    double s = 0.;
    for (double c : atm.cn2) s += c;  // std::accumulate over the whole vector
    const double r0 = std::pow(s, -3. / 5.);
    sNGS2 = lambda / r0 * std::pow(lambda / sys.lambdaNGS, 0.2) * kRasc;
    sNGS2 *= sNGS2;
    sLGS2 = lambda / r0 * std::pow(lambda / sys.lambdaLGS, 0.2) * kRasc;
    sLGS2 *= sLGS2;
}

// updateNoise (338-374)
void Tomo::update_noise() {
    noiseNGS.assign(sys.nNgs + 1, 0.0);
    noiseLGSxx.assign(Nx, 0.0);
    noiseLGSyy.assign(Nx, 0.0);
    noiseLGSxy.assign(Nx, 0.0);
    update_seeing();
    long k = 0;
    for (int i = 0; i < sys.nW; ++i) {
        const double d = sys.diam / Nssp[i];
        const double F = flux(qr, mr[i], throughput[i], d, sys.tFrame, bdw);
        const double fact1 = 0.32 / F;
        const double fact2 = 0.82 * std::pow(sys.RON / (F * pixSize[i]), 2);
        if (i < sys.nLgs) {
            const double pe = sLGS2 + (sys.spotWidth * sys.spotWidth);
            const double shrt = (fact1 * pe + fact2 * std::pow(pe, 1.5) * std::pow(pe, 0.5));
            for (long j = 0; j < Nsubap[i]; ++j, ++k) {
                const double e2 = lgsExt[k] * lgsExt[k];
                const double lng = (fact1 * (pe + e2) +
                                    fact2 * std::pow(pe + e2, 1.5) * std::pow(pe, 0.5));
                const double c = std::cos(lgsTheta[k]), sn = std::sin(lgsTheta[k]);
                noiseLGSxx[k] = lng * std::pow(c, 2) + shrt * std::pow(sn, 2);
                noiseLGSyy[k] = lng * std::pow(sn, 2) + shrt * std::pow(c, 2);
                noiseLGSxy[k] = (lng - shrt) * sn * c;
            }
        } else {
            noiseNGS[i - sys.nLgs] = (fact1 * sNGS2 + fact2 * sNGS2 * sNGS2);
        }
    }
}

// projectedSubap (501-565): u,v of every valid subaperture on every layer.
void Tomo::projected_subap() {
    const int L = atm.nLayer;
    const double rad = 3.14159265358979323846 / 180.;
    std::vector<long> ioff(sys.nW, 0);
    for (int i = 1; i < sys.nW; ++i) ioff[i] = ioff[i - 1] + Nsubap[i - 1];
    u.assign((size_t)L * Nx, 0.0);
    v.assign((size_t)L * Nx, 0.0);
    for (int n = 0; n < sys.nW; ++n) {
        const double G = diamPup[n] / (double)Nssp[n];
        const double th = sys.thetaML[n] * rad;
        const double ct = std::cos(th), st = std::sin(th);
        for (int l = 0; l < L; ++l) {
            const double h = atm.h[l];
            const double dX = alphaX[n] * h, dY = alphaY[n] * h;
            const double rr = 1. - h * sys.gsAlt[n];
            for (long i = 0; i < Nsubap[n]; ++i) {
                const double xtp = X[ioff[n] + i] * G, ytp = Y[ioff[n] + i] * G;
                double uu = xtp * ct - ytp * st;
                double vv = xtp * st + ytp * ct;
                uu += sys.XPup[n];
                vv += sys.YPup[n];
                u[(size_t)l * Nx + ioff[n] + i] = uu * rr + dX;
                v[(size_t)l * Nx + ioff[n] + i] = vv * rr + dY;
            }
        }
    }
}

}  // namespace moao
