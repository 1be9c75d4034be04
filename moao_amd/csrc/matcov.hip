// FP64 covariance-tile generator for gfx950: Cmm (type 1), Cmaa (2), Ctm (3), Ctt (4).
//
// Arithmetic follows the reference element by element (common/kernels/matcov_kernels.cpp):
//   vk_series      macdo_x56      37-62     x^(5/6) K_{5/6}(x) by its 11-term series
//   vk_asymptote   asymp_macdo    64-89     3-term Hankel asymptote for x > 3*pi/2
//   vk_dphi        rodconan/DPHI_gb 91-140  von Karman phase structure function
//   slope_cov      cov_XX/YY/XY/YX_gb 146-193  4-point second differences
//   meas_kernel    compute_element_tiled_4 + matcov_kernel_4   205-400
//   truth_kernel   compute_element_ts_tile + matcov_ts_kernel_tile 422-515
// with the tile/storage rules of MOAO_COMMON::matcov_tile (common/compute/matcov.cpp:17-98).
//
// GPU design (not a translation of matcov_kernels_gpu.cu):
//  * one lane per element, a wave runs down 64 consecutive rows of one column, so the column's
//    u/v/noise reads are wave-uniform and the 64 FP64 stores are one contiguous 512-byte run
//    of the column-major tile;
//  * the per-element WFS/subaperture/axis decode (a linear scan over WFS in the reference) is
//    a host-built int32 table per row and per column;
//  * per-layer constants 2*pi/L0 and k1*L0^(5/3) are computed once on the host with the
//    reference's own expressions (identical bits), not per call of the structure function;
//  * type 1 skips workgroups whose tile lies wholly above the diagonal;
//  * a whole Cmm or Ctm (the lapack path's call) runs pair_kernel instead: one lane per pair
//    of subapertures computes the pair's xx, xy, yx and yy elements, sharing the structure
//    functions those four have in common (9 evaluations per layer instead of 14 for sensors
//    of equal subaperture size), bit-identical to the element kernel.
// Contraction into FMA is disabled here so each product/sum rounds as in the reference; the
// only differences left are the last-ulp differences of the device pow/exp/sqrt.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "profile.hpp"
#include "stages.hpp"

namespace moao {
namespace {

constexpr double kPi = 3.1415926535897932384626433;
constexpr double kXlim = 0.75 * 2 * kPi;
constexpr double kLambda2 = 0.00026942094446267851;
constexpr double kK1 = 0.1716613621245709486;
constexpr double kK2 = 1.00563491799858928388289314170833;
constexpr double kK3 = 1.25331413731550012081;
constexpr double kA1 = 0.22222222222222222222;
constexpr double kA2 = -0.08641975308641974829;
constexpr double kA3 = 0.08001828989483310284;

// Series coefficients of the reference (common/moao_defs.h:18-27): Gamma-function values.
__constant__ double cGa[11] = {0, 12.067619015983075, 5.17183672113560444,
                               0.795667187867016068, 0.0628158306210802181,
                               0.00301515986981185091, 9.72632216068338833e-05,
                               2.25320204494595251e-06, 3.93000356676612095e-08,
                               5.34694362825451923e-10, 5.83302941264329804e-12};
__constant__ double cGma[11] = {-3.74878707653729304, -2.04479295083852408,
                                -0.360845814853857083, -0.0313778969438136685,
                                -0.001622994669507603, -5.56455315259749673e-05,
                                -1.35720808599938951e-06, -2.47515152461894642e-08,
                                -3.50257291219662472e-10, -3.95770950530691961e-12,
                                -3.65327031259100284e-14};

constexpr int kRows = 64;       // rows per workgroup (one wave tall)
constexpr int kColsPerWave = 4; // columns per wave
constexpr int kWaves = 4;       // waves per workgroup
constexpr int kCols = kColsPerWave * kWaves;

struct MatcovArgs {
    double* out;
    long lda;
    int nrows, ncols, xoffset, yoffset;
    int Nlayer, nlgs, Nx, type_mat;
    double lgs_cst;
    const int* rowtab;      // packed (pos << 9) | (wfs << 1) | xy
    const int* coltab;
    const double* u;        // [Nlayer*Nx]
    const double* v;
    const double* sspL;     // [Nw*Nlayer]
    const double* lk;       // [Nlayer] 2*pi/L0
    const double* amp;      // [Nlayer] k1*L0^(5/3)
    const double* cn2;      // [Nlayer]
    const double* sspsz;    // [Nw] diam/Nssp
    const double* nNGS;     // [Nw-nlgs]
    const double* nxx;      // [Nx]
    const double* nyy;
    const double* nxy;
    const double* X;        // [Nx]
    const double* Y;
    int truth_base;         // Nsubap_t*(Nw-1), index of the truth sensor's first subap
    int truth_nsub;         // Nsubap of the truth sensor
    // pair kernel (whole matrices of types 1 and 3): one entry per subaperture of the row and
    // column sensors, {(pos << 9) | (wfs << 1), matrix index of its x slope}; the y slope sits
    // nsub[wfs] further
    const int2* rsub;
    const int2* csub;
    const int* nsub;        // [Nw]
    int nrsub, ncsub;
};

// cx = cbrt(x), shared with vk_asymptote: lanes of one wave often take both branches.
__device__ __forceinline__ double vk_series(double x, double cx) {
    // x^(5/3) as x * cbrt(x)^2: a cube root costs a fraction of a general pow
    const double x2a = x * cx * cx;
    const double x22 = x * x / 4.;
    double xn = 0.5;
    double s = cGma[0] * x2a;
    s *= xn;
    xn *= x22;
#pragma unroll
    for (int n = 1; n <= 10; n++) {
        s += (cGma[n] * x2a + cGa[n]) * xn;
        xn *= x22;
    }
    return s;
}

__device__ __forceinline__ double vk_asymptote(double x, double cx) {
    const double ix = 1. / x;
    return kK2 - kK3 * exp(-x) * cx * (1.0 + ix * (kA1 + ix * (kA2 + ix * kA3)));
}

__device__ __forceinline__ double vk_dphi(double x, double y, double lk, double amp) {
    const double r = sqrt(x * x + y * y);
    const double q = lk * r;
    const double cq = cbrt(q);
    double res = (q > kXlim) ? vk_asymptote(q, cq) : -vk_series(q, cq);
    res *= amp;
    return res;
}

__device__ __forceinline__ double slope_cov(int kind, double du, double dv, double s1, double s2,
                                            double lk, double amp) {
    const double ac = s1 - s2, ad = s1 + s2, bc = -ad, bd = -ac;
    // Equal sizes (s1 == s2, most WFS pairs): ac = +0 and bd = -0, so the first and last
    // terms share one argument exactly; evaluating it once keeps every operation and its order.
    switch (kind) {
        case 0:
            if (ac == 0.0) {
                const double d0 = vk_dphi(du, dv, lk, amp);
                return -d0 + vk_dphi(du + ad, dv, lk, amp) + vk_dphi(du + bc, dv, lk, amp) - d0;
            }
            return -vk_dphi(du + ac, dv, lk, amp) + vk_dphi(du + ad, dv, lk, amp) +
                   vk_dphi(du + bc, dv, lk, amp) - vk_dphi(du + bd, dv, lk, amp);
        case 3:
            if (ac == 0.0) {
                const double d0 = vk_dphi(du, dv, lk, amp);
                return -d0 + vk_dphi(du, dv + ad, lk, amp) + vk_dphi(du, dv + bc, lk, amp) - d0;
            }
            return -vk_dphi(du, dv + ac, lk, amp) + vk_dphi(du, dv + ad, lk, amp) +
                   vk_dphi(du, dv + bc, lk, amp) - vk_dphi(du, dv + bd, lk, amp);
        case 1:
            return -vk_dphi(du + s1, dv - s2, lk, amp) + vk_dphi(du + s1, dv + s2, lk, amp) +
                   vk_dphi(du - s1, dv - s2, lk, amp) - vk_dphi(du - s1, dv + s2, lk, amp);
        default:
            return -vk_dphi(du - s2, dv + s1, lk, amp) + vk_dphi(du + s2, dv + s1, lk, amp) +
                   vk_dphi(du - s2, dv - s1, lk, amp) - vk_dphi(du + s2, dv - s1, lk, amp);
    }
}

__global__ void __launch_bounds__(kWaves * 64) meas_kernel(MatcovArgs a) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int lx = blockIdx.x * kRows + lane;
    const int ly0 = blockIdx.y * kCols + wave * kColsPerWave;
    if (a.type_mat == 1) {
        // whole workgroup above the diagonal: nothing to write
        const int last_row = a.xoffset + min(a.nrows - 1, (int)blockIdx.x * kRows + kRows - 1);
        if (last_row < a.yoffset + (int)blockIdx.y * kCols) return;
    }
    if (lx >= a.nrows) return;
    const int ri = a.rowtab[lx];
    const int m = (ri >> 1) & 0xff, xi = ri & 1, p1 = ri >> 9;
    const int gx = lx + a.xoffset;
    for (int c = 0; c < kColsPerWave; ++c) {
        const int ly = ly0 + c;
        if (ly >= a.ncols) break;
        if (a.type_mat == 1 && ly + a.yoffset > gx) continue;
        const int ci = a.coltab[ly];
        const int n = (ci >> 1) & 0xff, xj = ci & 1, p2 = ci >> 9;
        const int kind = xi * 2 + xj;
        const double kk = kLambda2 / (a.sspsz[m] * a.sspsz[n]);
        double covar = 0.0;
        for (int l = 0; l < a.Nlayer; ++l) {
            const double sml = a.sspL[m * a.Nlayer + l];
            const double snl = a.sspL[n * a.Nlayer + l];
            if ((sml > 0) && (snl > 0)) {
                const long o = (long)l * a.Nx;
                const double du = a.u[o + p1] - a.u[o + p2];
                const double dv = a.v[o + p1] - a.v[o + p2];
                covar += 0.5 * slope_cov(kind, du, dv, sml * 0.5, snl * 0.5, a.lk[l], a.amp[l]) *
                         kk * a.cn2[l];
            }
        }
        if (m == n) {
            if (m < a.nlgs) {
                if (p1 == p2) {
                    if (kind == 0) covar += a.nxx[p1];
                    else if (kind == 3) covar += a.nyy[p1];
                    else covar += a.nxy[p1];
                }
                if (kind == 0 || kind == 3) covar += a.lgs_cst;
            } else if (p1 == p2 && (kind == 0 || kind == 3)) {
                covar += a.nNGS[m - a.nlgs];
            }
        }
        a.out[(long)ly * a.lda + lx] = covar;
    }
}

// Whole Cmm (type 1) or Ctm (type 3): one lane per pair of subapertures computes the 2 x 2
// block xx, xy, yx, yy of that pair, so the four slope covariances share their structure
// functions.  Per layer, for equal subaperture sizes (every pair of sensors at one altitude),
// xx and yy share D(du, dv), and xy and yx use the same four corners (du +- s, dv +- s): 9
// evaluations instead of the 14 the four elements take one by one.  Every element keeps the
// element kernel's expression and order (same arguments, same sums, same layer order), so the
// result is bit-identical to it.  Lanes run down consecutive row subapertures: the x and y
// rows of a pair are two contiguous runs of each column.
__global__ void __launch_bounds__(kWaves * 64) pair_kernel(MatcovArgs a) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int rs = blockIdx.x * kRows + lane;
    const int cs0 = blockIdx.y * kCols + wave * kColsPerWave;
    if (rs >= a.nrsub) return;
    const int2 re = a.rsub[rs];
    const int m = (re.x >> 1) & 0xff, p1 = re.x >> 9;
    const int xr = re.y, yr = re.y + a.nsub[m];
    for (int c = 0; c < kColsPerWave; ++c) {
        const int cs = cs0 + c;
        if (cs >= a.ncsub) break;
        const int2 ce = a.csub[cs];
        const int n = (ce.x >> 1) & 0xff, p2 = ce.x >> 9;
        const int xc = ce.y, yc = ce.y + a.nsub[n];
        // type 1 writes only row >= column: nothing of this pair when its largest row is above
        // its smallest column; in a sensor's diagonal block a pair may need only some of its
        // four elements, and only their structure functions are evaluated
        const bool lo = a.type_mat == 1;
        const bool nxx = !lo || xr >= xc, nyy = !lo || yr >= yc;
        const bool nxy = !lo || xr >= yc, nyx = !lo || yr >= xc;
        if (!nyx) continue;  // yx has the largest row and the smallest column
        const double kk = kLambda2 / (a.sspsz[m] * a.sspsz[n]);
        double cxx = 0.0, cxy = 0.0, cyx = 0.0, cyy = 0.0;
        for (int l = 0; l < a.Nlayer; ++l) {
            const double sml = a.sspL[m * a.Nlayer + l];
            const double snl = a.sspL[n * a.Nlayer + l];
            if ((sml > 0) && (snl > 0)) {
                const long o = (long)l * a.Nx;
                const double du = a.u[o + p1] - a.u[o + p2];
                const double dv = a.v[o + p1] - a.v[o + p2];
                const double s1 = sml * 0.5, s2 = snl * 0.5;
                const double lk = a.lk[l], amp = a.amp[l];
                const double ac = s1 - s2, ad = s1 + s2, bc = -ad, bd = -ac;
                double sxx, syy, sxy, syx;
                sxx = syy = sxy = 0.0;
                if (ac == 0.0) {
                    if (nxx || nyy) {
                        const double d0 = vk_dphi(du, dv, lk, amp);
                        if (nxx)
                            sxx = -d0 + vk_dphi(du + ad, dv, lk, amp) +
                                  vk_dphi(du + bc, dv, lk, amp) - d0;
                        if (nyy)
                            syy = -d0 + vk_dphi(du, dv + ad, lk, amp) +
                                  vk_dphi(du, dv + bc, lk, amp) - d0;
                    }
                    // s1 == s2: the corners of xy and yx are the same four arguments
                    const double q1 = vk_dphi(du + s1, dv - s2, lk, amp);
                    const double q2 = vk_dphi(du + s1, dv + s2, lk, amp);
                    const double q3 = vk_dphi(du - s1, dv - s2, lk, amp);
                    const double q4 = vk_dphi(du - s1, dv + s2, lk, amp);
                    if (nxy) sxy = -q1 + q2 + q3 - q4;
                    syx = -q4 + q2 + q3 - q1;
                } else {
                    if (nxx)
                        sxx = -vk_dphi(du + ac, dv, lk, amp) + vk_dphi(du + ad, dv, lk, amp) +
                              vk_dphi(du + bc, dv, lk, amp) - vk_dphi(du + bd, dv, lk, amp);
                    if (nyy)
                        syy = -vk_dphi(du, dv + ac, lk, amp) + vk_dphi(du, dv + ad, lk, amp) +
                              vk_dphi(du, dv + bc, lk, amp) - vk_dphi(du, dv + bd, lk, amp);
                    if (nxy)
                        sxy = -vk_dphi(du + s1, dv - s2, lk, amp) +
                              vk_dphi(du + s1, dv + s2, lk, amp) +
                              vk_dphi(du - s1, dv - s2, lk, amp) - vk_dphi(du - s1, dv + s2, lk, amp);
                    syx = -vk_dphi(du - s2, dv + s1, lk, amp) + vk_dphi(du + s2, dv + s1, lk, amp) +
                          vk_dphi(du - s2, dv - s1, lk, amp) - vk_dphi(du + s2, dv - s1, lk, amp);
                }
                cxx += 0.5 * sxx * kk * a.cn2[l];
                cxy += 0.5 * sxy * kk * a.cn2[l];
                cyx += 0.5 * syx * kk * a.cn2[l];
                cyy += 0.5 * syy * kk * a.cn2[l];
            }
        }
        if (m == n) {
            if (m < a.nlgs) {
                if (p1 == p2) {
                    cxx += a.nxx[p1];
                    cyy += a.nyy[p1];
                    cxy += a.nxy[p1];
                    cyx += a.nxy[p1];
                }
                cxx += a.lgs_cst;
                cyy += a.lgs_cst;
            } else if (p1 == p2) {
                cxx += a.nNGS[m - a.nlgs];
                cyy += a.nNGS[m - a.nlgs];
            }
        }
        double* ox = a.out + (long)xc * a.lda;
        double* oy = a.out + (long)yc * a.lda;
        if (nxx) ox[xr] = cxx;
        ox[yr] = cyx;
        if (nxy) oy[xr] = cxy;
        if (nyy) oy[yr] = cyy;
    }
}

__global__ void __launch_bounds__(kWaves * 64) truth_kernel(MatcovArgs a) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int lx = blockIdx.x * kRows + lane;
    const int ly0 = blockIdx.y * kCols + wave * kColsPerWave;
    if (lx >= a.nrows) return;
    const int gx = lx + a.xoffset;
    const int ns = a.truth_nsub;
    const int i = gx < ns ? gx : gx - ns;
    const int xi = gx < ns ? 0 : 1;
    const double ss = a.sspsz[0];  // diam / Nssp of the truth sensor (set by the host)
    const double kk = kLambda2 / (ss * ss);
    const double s = ss * 0.5;
    for (int c = 0; c < kColsPerWave; ++c) {
        const int ly = ly0 + c;
        if (ly >= a.ncols) break;
        const int gy = ly + a.yoffset;
        const int j = gy < ns ? gy : gy - ns;
        const int xj = gy < ns ? 0 : 1;
        const int kind = xi * 2 + xj;
        const double du = a.X[a.truth_base + i] - a.X[a.truth_base + j];
        const double dv = a.Y[a.truth_base + i] - a.Y[a.truth_base + j];
        double covar = 0.0;
        for (int l = 0; l < a.Nlayer; ++l) {
            if (ss > 0) covar += 0.5 * slope_cov(kind, du, dv, s, s, a.lk[l], a.amp[l]) * kk * a.cn2[l];
        }
        a.out[(long)ly * a.lda + lx] = covar;
    }
}

// Packed decode of measurement index g for the non-truth sensors (type 1/2/3 columns, type 1/2
// rows): WFS found by the cumulative 2*Nsubap scan of compute_element_tiled_4, x slopes first.
inline int pack_meas(long g, const moao_tomo_view* t, const std::vector<long>& off) {
    int w = 0;
    while (w < t->Nw - 1 && g >= 2 * (off[w] + t->Nsubap[w])) ++w;
    long i = g - 2 * off[w];
    int xy = 0;
    if (i >= t->Nsubap[w]) { i -= t->Nsubap[w]; xy = 1; }
    return (int)(((off[w] + i) << 9) | (w << 1) | xy);
}

// Row decode of type 3 (and column decode of type 2): the truth sensor.
inline int pack_truth(long g, const moao_tomo_view* t, const std::vector<long>& off) {
    const int w = t->Nw - 1;
    long i = g;
    int xy = 0;
    if (i >= t->Nsubap[w]) { i -= t->Nsubap[w]; xy = 1; }
    return (int)(((off[w] + i) << 9) | (w << 1) | xy);
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

int matcov_tile_launch(double* d_data, int nrows, int ncols, int xoffset, int yoffset, int lda,
                       const moao_tomo_view* t, int type_mat, hipStream_t stream, DevBuf* tables) {
    if (type_mat < 1 || type_mat > 4) {
        set_error("matcov_tile: unrecognized type_mat " + std::to_string(type_mat));
        return MOAO_ERR_ARG;
    }
    if (!t || nrows < 0 || ncols < 0 || lda < std::max(1, nrows) || xoffset < 0 || yoffset < 0 ||
        t->Nw < 2 || t->Nw > 255 || t->Nlayer < 1 || t->Nx < 1 || t->Nx >= (1L << 22)) {
        set_error("matcov_tile: invalid arguments");
        return MOAO_ERR_ARG;
    }
    if (nrows == 0 || ncols == 0) return MOAO_OK;
    const int Nw = t->Nw, L = t->Nlayer;
    const long Nx = t->Nx;
    std::vector<long> off(Nw, 0);
    for (int w = 1; w < Nw; ++w) off[w] = off[w - 1] + t->Nsubap[w - 1];
    const long nmeas = 2 * (Nx - t->Nsubap[Nw - 1]);
    const long nts = 2 * t->Nsubap[Nw - 1];
    // Row/column ranges must stay inside the matrix the type describes.
    long row_lim = (type_mat == 3 || type_mat == 4) ? nts : nmeas;
    long col_lim = (type_mat == 2 || type_mat == 4) ? nts : nmeas;
    if ((long)xoffset + nrows > row_lim || (long)yoffset + ncols > col_lim) {
        set_error("matcov_tile: tile exceeds the matrix of this type");
        return MOAO_ERR_ARG;
    }
    if (type_mat == 2 || type_mat == 4) {
        for (int w = 0; w < Nw; ++w)
            if (t->Nsubap[w] != t->Nsubap[0]) {
                set_error("matcov_tile: types 2 and 4 assume equal Nsubap on every WFS");
                return MOAO_ERR_ARG;
            }
    }

    std::vector<int> rowtab(nrows), coltab(ncols);
    if (type_mat != 4) {
        for (int r = 0; r < nrows; ++r)
            rowtab[r] = (type_mat == 3) ? pack_truth(r + xoffset, t, off)
                                        : pack_meas(r + xoffset, t, off);
        for (int c = 0; c < ncols; ++c)
            coltab[c] = (type_mat == 2) ? pack_truth(c + yoffset, t, off)
                                        : pack_meas(c + yoffset, t, off);
    }
    // Whole Cmm / Ctm: the pair kernel (one lane per subaperture pair).
    const bool pairs = (type_mat == 1 || type_mat == 3) && xoffset == 0 && yoffset == 0 &&
                       nrows == row_lim && ncols == col_lim;
    std::vector<int2> rsub, csub;
    std::vector<int> nsub(Nw);
    for (int w = 0; w < Nw; ++w) nsub[w] = (int)t->Nsubap[w];
    if (pairs) {
        for (int w = 0; w < Nw - 1; ++w)
            for (long i = 0; i < t->Nsubap[w]; ++i)
                csub.push_back(make_int2((int)(((off[w] + i) << 9) | (w << 1)),
                                         (int)(2 * off[w] + i)));
        if (type_mat == 1) {
            rsub = csub;
        } else {
            const int w = Nw - 1;
            for (long i = 0; i < t->Nsubap[w]; ++i)
                rsub.push_back(make_int2((int)(((off[w] + i) << 9) | (w << 1)), (int)i));
        }
    }
    std::vector<double> lk(L), amp(L), sspsz(Nw);
    for (int l = 0; l < L; ++l) {
        const double L0 = t->L0[l];
        lk[l] = (2 * kPi / L0);
        amp[l] = kK1 * std::pow(L0, 5. / 3.0);
    }
    for (int w = 0; w < Nw; ++w) sspsz[w] = t->diam / t->Nssp[w];
    if (type_mat == 4) sspsz[0] = t->diam / t->Nssp[Nw - 1];
    const int nngs = Nw - t->nlgs;

    // One contiguous upload of every vector the kernel reads.
    struct Piece { const void* src; size_t bytes; size_t off; };
    std::vector<Piece> pieces;
    size_t total = 0;
    auto add = [&](const void* src, size_t bytes) {
        pieces.push_back({src, bytes, total});
        total += align256(bytes);
        return pieces.size() - 1;
    };
    const size_t iRow = add(rowtab.data(), sizeof(int) * nrows);
    const size_t iCol = add(coltab.data(), sizeof(int) * ncols);
    const size_t iU = add(t->u, sizeof(double) * L * Nx);
    const size_t iV = add(t->v, sizeof(double) * L * Nx);
    const size_t iS = add(t->sspSizeL, sizeof(double) * Nw * L);
    const size_t iLk = add(lk.data(), sizeof(double) * L);
    const size_t iAmp = add(amp.data(), sizeof(double) * L);
    const size_t iCn2 = add(t->cn2, sizeof(double) * L);
    const size_t iSz = add(sspsz.data(), sizeof(double) * Nw);
    const size_t iNg = add(t->noiseNGS, sizeof(double) * nngs);
    const size_t iXX = add(t->noiseLGSxx, sizeof(double) * Nx);
    const size_t iYY = add(t->noiseLGSyy, sizeof(double) * Nx);
    const size_t iXY = add(t->noiseLGSxy, sizeof(double) * Nx);
    const size_t iX = add(t->X, sizeof(double) * Nx);
    const size_t iY = add(t->Y, sizeof(double) * Nx);
    const size_t iRs = add(rsub.data(), sizeof(int2) * rsub.size());
    const size_t iCs = add(csub.data(), sizeof(int2) * csub.size());
    const size_t iNs = add(nsub.data(), sizeof(int) * Nw);

    std::vector<unsigned char> host(total);
    for (auto& p : pieces)
        if (p.src && p.bytes) std::memcpy(host.data() + p.off, p.src, p.bytes);
    // Tables of the library context are shared by every caller: hold the lock until the kernel
    // that reads them has run.  A caller-owned buffer (one per serving slot) needs no lock.
    std::unique_lock<std::mutex> g;
    if (!tables) {
        g = std::unique_lock<std::mutex>(ctx().mu);
        tables = &ctx().matcov_tables;
    }
    if (int rc = tables->ensure(total)) return rc;
    unsigned char* d = tables->as<unsigned char>();
    // The upload buffer is reused by the next call: order it on the stream and make the host
    // copy complete before returning (hipMemcpyAsync from pageable memory is staged).
    MOAO_HIP_TRY(hipMemcpyAsync(d, host.data(), total, hipMemcpyHostToDevice, stream));

    MatcovArgs a;
    a.out = d_data;
    a.lda = lda;
    a.nrows = nrows;
    a.ncols = ncols;
    a.xoffset = xoffset;
    a.yoffset = yoffset;
    a.Nlayer = L;
    a.nlgs = t->nlgs;
    a.Nx = (int)Nx;
    a.type_mat = type_mat;
    a.lgs_cst = t->lgs_cst;
    auto P = [&](size_t i) { return (const void*)(d + pieces[i].off); };
    a.rowtab = (const int*)P(iRow);
    a.coltab = (const int*)P(iCol);
    a.u = (const double*)P(iU);
    a.v = (const double*)P(iV);
    a.sspL = (const double*)P(iS);
    a.lk = (const double*)P(iLk);
    a.amp = (const double*)P(iAmp);
    a.cn2 = (const double*)P(iCn2);
    a.sspsz = (const double*)P(iSz);
    a.nNGS = (const double*)P(iNg);
    a.nxx = (const double*)P(iXX);
    a.nyy = (const double*)P(iYY);
    a.nxy = (const double*)P(iXY);
    a.X = (const double*)P(iX);
    a.Y = (const double*)P(iY);
    a.truth_nsub = (int)t->Nsubap[Nw - 1];
    a.truth_base = (int)(t->Nsubap[Nw - 1] * (Nw - 1));
    a.rsub = (const int2*)P(iRs);
    a.csub = (const int2*)P(iCs);
    a.nsub = (const int*)P(iNs);
    a.nrsub = (int)rsub.size();
    a.ncsub = (int)csub.size();

    dim3 grid((nrows + kRows - 1) / kRows, (ncols + kCols - 1) / kCols);
    if (grid.y > 65535) {
        set_error("matcov_tile: too many columns for one launch");
        return MOAO_ERR_ARG;
    }
    {
        // work = elements written (bytes = 8 x elements)
        double elems = (double)nrows * ncols;
        if (type_mat == 1) {
            elems = 0;
            for (int c = 0; c < ncols; ++c)
                elems += std::max(0, std::min(nrows, xoffset + nrows - (yoffset + c)));
        }
        ProfScope ps(kProfMatcov, elems, stream);
        if (pairs) {
            const dim3 pg((a.nrsub + kRows - 1) / kRows, (a.ncsub + kCols - 1) / kCols);
            hipLaunchKernelGGL(pair_kernel, pg, dim3(kWaves * 64), 0, stream, a);
        } else if (type_mat == 4)
            hipLaunchKernelGGL(truth_kernel, grid, dim3(kWaves * 64), 0, stream, a);
        else
            hipLaunchKernelGGL(meas_kernel, grid, dim3(kWaves * 64), 0, stream, a);
    }
    MOAO_HIP_TRY(hipGetLastError());
    // The staging buffer must not be rewritten by another call before this kernel ran.
    MOAO_HIP_TRY(hipStreamSynchronize(stream));
    return MOAO_OK;
}

}  // namespace moao

extern "C" {

int moao_hip_matcov_tile_dev(double* d_data, int nrows, int ncols, int xoffset, int yoffset,
                             int lda, const moao_tomo_view* tomo, int type_mat, void* stream) {
    if (int rc = moao::ensure_init()) return rc;
    return moao::matcov_tile_launch(d_data, nrows, ncols, xoffset, yoffset, lda, tomo, type_mat,
                                    moao::pick_stream(stream), nullptr);
}

int moao_hip_matcov_tile(double* data, int nrows, int ncols, int xoffset, int yoffset, int lda,
                         const moao_tomo_view* tomo, int type_mat) {
    if (int rc = moao::ensure_init()) return rc;
    if (!data || nrows < 0 || ncols < 0 || lda < std::max(1, nrows)) {
        moao::set_error("matcov_tile: invalid host buffer");
        return MOAO_ERR_ARG;
    }
    if (nrows == 0 || ncols == 0) return MOAO_OK;
    moao::Context& c = moao::ctx();
    hipStream_t s = c.stream;
    const size_t span = ((size_t)(ncols - 1) * lda + nrows) * sizeof(double);
    std::lock_guard<std::mutex> g(c.mu);
    if (int rc = c.host_stage.ensure(span)) return rc;
    double* d = c.host_stage.as<double>();
    // Type 1 leaves the strictly-upper part untouched and lda may leave gaps between columns:
    // start from the caller's contents so the copy back preserves them.
    if (type_mat == 1 || lda != nrows)
        MOAO_HIP_TRY(hipMemcpyAsync(d, data, span, hipMemcpyHostToDevice, s));
    // c.mu is held: hand the context's table buffer over explicitly
    int rc = moao::matcov_tile_launch(d, nrows, ncols, xoffset, yoffset, lda, tomo, type_mat, s,
                                      &c.matcov_tables);
    if (rc) return rc;
    MOAO_HIP_TRY(hipMemcpyAsync(data, d, span, hipMemcpyDeviceToHost, s));
    MOAO_HIP_TRY(hipStreamSynchronize(s));
    return MOAO_OK;
}

}  // extern "C"
