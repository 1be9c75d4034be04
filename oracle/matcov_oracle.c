/* matcov restated in plain C -- ORACLE, test infrastructure only (never linked by the product).
 *
 * Follows common/kernels/matcov_kernels.cpp and common/compute/matcov.cpp of the reference:
 *   vk_series      <- macdo_x56            matcov_kernels.cpp:37-62   (11-term K_{5/6} series)
 *   vk_asymptote   <- asymp_macdo          matcov_kernels.cpp:64-89   (3-term Hankel asymptote)
 *   vk_dphi        <- rodconan + DPHI_gb   matcov_kernels.cpp:91-140
 *   slope_cov      <- cov_XX/YY/XY/YX_gb   matcov_kernels.cpp:146-193
 *   element_meas   <- compute_element_tiled_4   matcov_kernels.cpp:205-333
 *   element_truth  <- compute_element_ts_tile   matcov_kernels.cpp:422-483
 *   oracle_matcov_tile <- MOAO_COMMON::matcov_tile + matcov_kernel_4 / matcov_ts_kernel_tile
 *                                          matcov.cpp:17-98, matcov_kernels.cpp:354-400, 495-515
 * Series and asymptote coefficients: common/moao_defs.h:15-38.  They are values of the Gamma
 * function (Ga[n] = (-1)^n G(-n+5/6) 2^(5/6)/n!, Gma[n] = (-1)^n G(-n-5/6) 2^(-5/6)/n!) and
 * are pinned against scipy.special.gamma in tests/test_oracle_matcov.py.
 *
 * Compiled with gcc -O2 -ffp-contract=off so every operation rounds exactly as the reference's
 * own x86-64 build does; with the same libm this reproduces the reference's values bit for bit.
 */
#include <math.h>
#include <stdio.h>

#define VK_PI 3.1415926535897932384626433
#define VK_XLIM (0.75 * 2 * VK_PI)

static const double GA[11] = {0, 12.067619015983075, 5.17183672113560444,
                              0.795667187867016068, 0.0628158306210802181,
                              0.00301515986981185091, 9.72632216068338833e-05,
                              2.25320204494595251e-06, 3.93000356676612095e-08,
                              5.34694362825451923e-10, 5.83302941264329804e-12};
static const double GMA[11] = {-3.74878707653729304, -2.04479295083852408,
                               -0.360845814853857083, -0.0313778969438136685,
                               -0.001622994669507603, -5.56455315259749673e-05,
                               -1.35720808599938951e-06, -2.47515152461894642e-08,
                               -3.50257291219662472e-10, -3.95770950530691961e-12,
                               -3.65327031259100284e-14};
static const double K1 = 0.1716613621245709486;
static const double K2 = 1.00563491799858928388289314170833;
static const double K3 = 1.25331413731550012081;
static const double A1 = 0.22222222222222222222;
static const double A2 = -0.08641975308641974829;
static const double A3 = 0.08001828989483310284;
/* (lambda/(2 pi) * RASC)^2 with lambda = 0.5e-6 m (matcov_kernels.cpp:213) */
static const double LAMBDA2 = 0.00026942094446267851;

typedef struct {
    int Nw, Nlayer, nlgs;
    long Nx;
    double diam, lgs_cst;
    const long *Nsubap, *Nssp;
    const double *sspSizeL, *u, *v, *X, *Y, *L0, *cn2;
    const double *noiseNGS, *noiseLGSxx, *noiseLGSyy, *noiseLGSxy;
} oracle_tomo;

/* x^(5/6) K_{5/6}(x) without its constant term, by the series of K_a (a = 5/6). */
double vk_series(double x) {
    const double a = 5. / 6.;
    const double x2a = pow(x, 2. * a);
    const double x22 = x * x / 4.;
    double xn = 0.5;
    double s = GMA[0] * x2a;
    s *= xn;
    xn *= x22;
    for (int n = 1; n <= 10; n++) {
        s += (GMA[n] * x2a + GA[n]) * xn;
        xn *= x22;
    }
    return s;
}

double vk_asymptote(double x) {
    const double ix = 1. / x;
    return K2 - K3 * exp(-x) * pow(x, 1 / 3.) * (1.0 + ix * (A1 + ix * (A2 + ix * A3)));
}

/* von Karman phase structure function at separation (x, y) for outer scale L0, in units of
 * r0^(-5/3). */
double vk_dphi(double x, double y, double L0) {
    const double r = sqrt(x * x + y * y);
    const double q = (2 * VK_PI / L0) * r;
    double res = (q > VK_XLIM) ? vk_asymptote(q) : -vk_series(q);
    res *= K1 * pow(L0, 5. / 3.0);
    return res;
}

/* 4-point second difference of D for slope covariances; kind 0=XX 3=YY 1=XY 2=YX. */
static double slope_cov(int kind, double du, double dv, double s1, double s2, double L0) {
    const double ac = s1 - s2, ad = s1 + s2, bc = -ad, bd = -ac;
    switch (kind) {
        case 0:
            return -vk_dphi(du + ac, dv, L0) + vk_dphi(du + ad, dv, L0) +
                   vk_dphi(du + bc, dv, L0) - vk_dphi(du + bd, dv, L0);
        case 3:
            return -vk_dphi(du, dv + ac, L0) + vk_dphi(du, dv + ad, L0) +
                   vk_dphi(du, dv + bc, L0) - vk_dphi(du, dv + bd, L0);
        case 1:
            return -vk_dphi(du + s1, dv - s2, L0) + vk_dphi(du + s1, dv + s2, L0) +
                   vk_dphi(du - s1, dv - s2, L0) - vk_dphi(du - s1, dv + s2, L0);
        default:
            return -vk_dphi(du - s2, dv + s1, L0) + vk_dphi(du + s2, dv + s1, L0) +
                   vk_dphi(du - s2, dv - s1, L0) - vk_dphi(du + s2, dv - s1, L0);
    }
}

/* Element (ipos, jpos) of Cmm (type 1), Cmaa (type 2) or Ctm (type 3). */
static double element_meas(int ipos, int jpos, const oracle_tomo* t, int type_mat) {
    const long* Ns = t->Nsubap;
    long offx = Ns[0];
    int m = 0;
    if (type_mat == 3) {
        m = t->Nw - 1;
        offx = t->Nx;
        ipos += 2 * (t->Nx - Ns[m]);
    } else {
        while (ipos / (2 * offx) >= 1) {
            m++;
            offx += Ns[m];
        }
    }
    offx -= Ns[m];
    long offy = Ns[0];
    int n = 0;
    if (type_mat == 2) {
        n = t->Nw - 1;
        offy = t->Nx;
        jpos += 2 * (t->Nx - Ns[m]); /* reference uses the row WFS here (matcov_kernels.cpp:238) */
    } else {
        while (jpos / (2 * offy) >= 1) {
            n++;
            offy += Ns[n];
        }
    }
    offy -= Ns[n];
    int i = ipos - 2 * (int)offx;
    int j = jpos - 2 * (int)offy;
    int xi = 0, xj = 0;
    if (i >= Ns[m]) { i -= Ns[m]; xi = 1; }
    if (j >= Ns[n]) { j -= Ns[n]; xj = 1; }
    const double sm = t->diam / t->Nssp[m];
    const double sn = t->diam / t->Nssp[n];
    const double kk = LAMBDA2 / (sm * sn);
    const int kind = xi * 2 + xj;
    double covar = 0.0;
    for (int l = 0; l < t->Nlayer; l++) {
        const double sml = t->sspSizeL[m * t->Nlayer + l];
        const double snl = t->sspSizeL[n * t->Nlayer + l];
        if (sml > 0 && snl > 0) {
            const int p1 = i + (int)offx + l * (int)t->Nx;
            const int p2 = j + (int)offy + l * (int)t->Nx;
            const double du = t->u[p1] - t->u[p2];
            const double dv = t->v[p1] - t->v[p2];
            covar += 0.5 * slope_cov(kind, du, dv, sml * 0.5, snl * 0.5, t->L0[l]) * kk * t->cn2[l];
        }
    }
    if (m == n) {
        if (m < t->nlgs) {
            if (i == j) {
                const int p = i + (int)offx;
                if (kind == 0) covar += t->noiseLGSxx[p];
                else if (kind == 3) covar += t->noiseLGSyy[p];
                else covar += t->noiseLGSxy[p];
            }
            if (kind == 0 || kind == 3) covar += t->lgs_cst;
        } else if (i == j && (kind == 0 || kind == 3)) {
            covar += t->noiseNGS[m - t->nlgs];
        }
    }
    return covar;
}

/* Element (ipos, jpos) of Ctt (type 4): unprojected truth-sensor positions, no noise. */
static double element_truth(int ipos, int jpos, const oracle_tomo* t) {
    const int Nsub = (int)t->Nsubap[t->Nw - 1];
    const int i = ipos < Nsub ? ipos : ipos - Nsub;
    const int j = jpos < Nsub ? jpos : jpos - Nsub;
    const int kind = (ipos < Nsub ? 0 : 2) + (jpos < Nsub ? 0 : 1);
    const double ss = t->diam / t->Nssp[t->Nw - 1];
    const double kk = LAMBDA2 / (ss * ss);
    const double s = ss * 0.5;
    const long base = (long)Nsub * (t->Nw - 1);
    const double du = t->X[base + i] - t->X[base + j];
    const double dv = t->Y[base + i] - t->Y[base + j];
    double covar = 0.0;
    for (int l = 0; l < t->Nlayer; l++) {
        if (ss > 0) covar += 0.5 * slope_cov(kind, du, dv, s, s, t->L0[l]) * kk * t->cn2[l];
    }
    return covar;
}

/* Tile generator with the reference's storage rules: column-major data[ly*lda+lx]; type 1
 * writes only global row >= global col; types 2, 3, 4 write the full tile.  Returns 0, or -1
 * for an unknown type (where the reference exits). */
int oracle_matcov_tile(double* data, int nrows, int ncols, int xoffset, int yoffset, int lda,
                       const oracle_tomo* t, int type_mat) {
    if (type_mat < 1 || type_mat > 4) return -1;
    for (int lx = 0; lx < nrows; lx++) {
        for (int ly = 0; ly < ncols; ly++) {
            const int gx = lx + xoffset, gy = ly + yoffset;
            if (type_mat == 4) {
                data[(long)ly * lda + lx] = element_truth(gx, gy, t);
            } else if (type_mat == 1) {
                if (gy <= gx) data[(long)ly * lda + lx] = element_meas(gx, gy, t, 1);
            } else {
                data[(long)ly * lda + lx] = element_meas(gx, gy, t, type_mat);
            }
        }
    }
    return 0;
}
