// A caller of the MOAO_HIP:: drop-in layer written against the reference's own Tomo_struct
// (common/tomo_struct/tomo_struct.hpp), making the calls of the reference's lapack driver
// (lapack/moao_lapack.cpp:57-181) with MOAO_HIP:: in place of MOAO_COMMON:: / MOAO_LAPACK:: and
// the global intersample functions.  Built (compiled and linked) by the CPU test suite where
// the reference tree is present; it is not run.
#include <string>
#include <vector>

#include "moao_hip.hpp"
#include "tomo_struct/tomo_struct.hpp"

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    Tomo_struct<double> tomo(std::string(argv[1]) + "/sys-params.txt", argv[2]);
    const long nmeas = tomo.getNMeas(), nmeasts = tomo.getNMeasTS(), nact = 64;
    std::vector<double> Cmm(nmeas * nmeas), R(nmeasts * nmeas), Ctt(nmeasts * nmeasts);
    std::vector<double> Cee(nmeasts * nmeasts), Cvv(nact * nact), Tmp(nact * nmeasts);
    std::vector<double> Dx(nact * nmeasts, 1e-3);
    MOAO_HIP::matcov_tile(Cmm.data(), nmeas, nmeas, 0, 0, nmeas, &tomo, 1);
    tomo.setTarget(0);
    MOAO_HIP::matcov_tile(R.data(), nmeasts, nmeas, 0, 0, nmeasts, &tomo, 3);
    MOAO_HIP::reconstructor(nmeas, nmeasts, Cmm.data(), nmeas, R.data(), nmeasts, 1L);
    MOAO_HIP::matcov_tile(Ctt.data(), nmeasts, nmeasts, 0, 0, nmeasts, &tomo, 4);
    std::vector<double> Ctm(nmeasts * nmeas);
    MOAO_HIP::Cee_Cvv(nmeas, nmeasts, nact, Cmm.data(), nmeas, Ctt.data(), nmeasts, Ctm.data(),
                      nmeasts, R.data(), nmeasts, Dx.data(), nact, Cee.data(), nmeasts,
                      Cvv.data(), nact, Tmp.data(), nact);
    struct MOAO_HIP::isample_struct isample;
    MOAO_HIP::intersample_prepare(&isample, nact * nact, tomo.Nssp[tomo.sys.nW - 1] + 1,
                                  tomo.sys.diam, argv[1]);
    MOAO_HIP::intersample_init(&isample);
    MOAO_HIP::intersample_process(&isample, Cvv.data());
    MOAO_HIP::intersample_free(&isample);
    // the FP32 production precision of the Chameleon driver
    std::vector<float> Cmm32(Cmm.begin(), Cmm.end()), R32(R.begin(), R.end());
    MOAO_HIP::reconstructor(nmeas, nmeasts, Cmm32.data(), nmeas, R32.data(), nmeasts, 1L);
    return MOAO_HIP::last_status() == 0 ? 0 : 1;
}
