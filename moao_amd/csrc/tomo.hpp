// Host-side system model: what the matcov generator reads.
//
// Mirrors the reference's Tomo_struct<double> (common/tomo_struct/tomo_struct.hpp:9-94) for the
// fields on the reconstructor path, built from the same sys-params / atm-params text files.
// Everything here is O(nLayer * Nx) host work and never touches the GPU.
#pragma once

#include <string>
#include <vector>

namespace moao {

struct SysParams {
    double diam = 0, obs = 0, tFrame = 0;
    int nW = 0, nLgs = 0, nNgs = 0, nTarget = 1, nssp = 0;
    std::vector<double> gsAlt;
    std::vector<int> gsType;
    std::vector<double> alphaX_as, alphaY_as, XPup, YPup, thetaML, thetaCam, sensibilite, tracking;
    double pasDphi = 0;
    int ncpu = 1;
    std::vector<double> mrNGS;
    double lgsFlux = 0, ngsPixSize = 0, lgsPixSize = 0, lambdaNGS = 0, lambdaLGS = 0, bdw_m = 0;
    double throughNGS = 0, throughLGS = 0, throughAtm = 0;
    int RON = 0;
    double lgsCst = 0, spotWidth = 0, lgsAlt = 0, lgsDepth = 0;
    std::vector<double> targetX_as, targetY_as;
};

struct AtmParams {
    int nLayer = 0;
    double r0 = 0;
    std::vector<double> cn2, h, L0;
};

// Throws std::runtime_error on a missing file, a truncated file or a vector that is too short.
SysParams read_sys_params(const std::string& path);
AtmParams read_atm_params(const std::string& path);

class Tomo {
  public:
    SysParams sys;
    AtmParams atm;

    std::vector<long> Nssp, Nsubap;
    std::vector<double> X, Y;              // valid subaperture centres, all WFS
    std::vector<double> alphaX, alphaY;    // radians; last entry = truth sensor / target
    std::vector<double> targetX, targetY;  // radians
    std::vector<double> sspSize, pixSize, throughput, diamPup, mr;
    std::vector<double> lgsExt, lgsTheta;
    std::vector<double> sspSizeL;          // [w * nLayer + l]
    std::vector<double> L0diff, cn2;       // [nLayer]
    std::vector<double> u, v;              // [l * Nx + pos]
    std::vector<double> noiseNGS;          // [nNgs + 1]
    std::vector<double> noiseLGSxx, noiseLGSyy, noiseLGSxy;  // [Nx]
    double qr = 2.9, bdw = 0, sNGS2 = 0, sLGS2 = 0;
    long Nx = 0;

    Tomo(const std::string& sys_file, const std::string& atm_file);
    Tomo(const SysParams& s, const AtmParams& a);

    void update_atm(const AtmParams& a);
    void update_atm(const std::string& atm_file);
    void set_target(int t);
    void set_target_coord(double ax, double ay, bool to_rad);

    long nmeas() const { return 2 * (Nx - Nsubap[sys.nW - 1]); }
    long nmeasts() const { return 2 * Nsubap[sys.nLgs + sys.nNgs]; }

  private:
    void init_sys();
    void valid_subap();
    void generate_elongation();
    void projected_subap();
    void update_seeing();
    void update_noise();
};

}  // namespace moao
