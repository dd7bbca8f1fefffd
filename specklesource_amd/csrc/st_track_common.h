// Device pieces shared by the strack stages: sample access, the
// correlation-peak search with the reference scan order, the FP64
// sub-pixel refinement and the 1-D Gaussian Levenberg-Marquardt fit.
#pragma once

#include "st_common.h"

namespace sttrack {

constexpr int NOVER = 12;  // Strack/speckleTrack.c:15
constexpr int NFAST = 8;   // speckleTrack.c:16
constexpr int LA = 2;      // speckleTrack.c:18
constexpr int NGAUSS = 18; // speckleTrack.c:29
constexpr int kFine = 2 * NOVER * 2;  // 48: side of the searched oversampled box
constexpr int NT = 256;

struct Images {
    const void *img1, *img2;
    int nr1, na1, nr2, na2, stype;
};

__device__ inline float2 load_sample(const void *img, int stype, int nr, int a, int r)
{
    const size_t o = (size_t)a * nr + r;
    if (stype == ST_SAMPLE_C64)
        return ((const float2 *)img)[o];
    const short2 s = ((const short2 *)img)[o];
    return make_float2((float)s.x, (float)s.y);
}

// Per-attempt result of the peak stage.
struct PeakRes {
    int I, J;         // coarse peak in the fftshifted surface, -1000 = rejected
    float raw;        // |C| at the coarse peak (findCPeak's cMax), -1 when rejected
    float gap;        // (best - runner-up) / best on the coarse surface
    float snr;        // coarse peak / mean |C| over the surface
    int fi, fj;       // oversampled-box argmax (0..47) or -1
    double fineMax;   // |Q| at the refined peak
    float fineGap;
    int lmOk;         // 1 when the 18-sample Gaussian fits are defined
    float lmX0[2];    // first abscissa of the range / azimuth samples
    float lmY[2][NGAUSS];
    int gI, gJ;       // -gauss mode: fitted peak in 1/NOVER units (-LARGEINT: failed)
    double gC;        // -gauss mode: fitted amplitude, capped at 1
};

// Dirichlet kernel of the 8 -> 96 spectral zero pad (speckleTrack.c:852-863,
// 2014-2038): D(t) = sum_{k=-4}^{3} exp(2 pi i k t / 96), t = -96 .. 96.
extern __constant__ double2 c_dir96[193];

#pragma clang fp contract(off)
// 1-D Gaussian LM fit (gaussPeak1D / mrqminMod / mrqcofMod / myGauss1D,
// speckleTrack.c:2137-2202, 2309-2465), all in float like the reference.
__device__ inline void lm_gauss(float x, const float a[3], float *y, float dyda[3])
{
    float sig = a[2];
    float t = x - a[1];
    float argx = (float)(0.5 * (double)t * (double)t / (double)(sig * sig));
    float ex = (float)exp(-(double)argx);
    *y = a[0] * ex;
    dyda[0] = ex;
    dyda[1] = *y * t / (sig * sig);
    dyda[2] = *y * t * t / (sig * sig * sig);
}

__device__ inline void lm_coef(const float *x, const float *y, const float a[3], float alpha[3][3], float beta[3],
                               float *chisq)
{
    for (int j = 0; j < 3; j++) {
        for (int k = 0; k <= j; k++)
            alpha[j][k] = 0.0f;
        beta[j] = 0.0f;
    }
    *chisq = 0.0f;
    const float sg = 0.1f;
    const float sig2i = (float)(1.0 / (double)(sg * sg));
    for (int i = 0; i < NGAUSS; i++) {
        float ymod, dyda[3];
        lm_gauss(x[i], a, &ymod, dyda);
        float dy = y[i] - ymod;
        for (int l = 0; l < 3; l++) {
            float wt = dyda[l] * sig2i;
            for (int m = 0; m <= l; m++)
                alpha[l][m] += wt * dyda[m];
            beta[l] += dy * wt;
        }
        *chisq += dy * dy * sig2i;
    }
    for (int j = 1; j < 3; j++)
        for (int k = 0; k < j; k++)
            alpha[k][j] = alpha[j][k];
}

// Gauss-Jordan with full pivoting on a 3x3 system; 0 when singular.
__device__ inline int solve3(float m[3][3], float b[3])
{
    int used[3] = {0, 0, 0};
    for (int it = 0; it < 3; it++) {
        float big = 0;
        int r = -1, c = -1;
        for (int j = 0; j < 3; j++) {
            if (used[j])
                continue;
            for (int k = 0; k < 3; k++)
                if (!used[k] && fabsf(m[j][k]) >= big) {
                    big = fabsf(m[j][k]);
                    r = j;
                    c = k;
                }
        }
        if (r < 0)
            return 0;
        used[c] = 1;
        if (r != c) {
            for (int l = 0; l < 3; l++) {
                float t = m[r][l];
                m[r][l] = m[c][l];
                m[c][l] = t;
            }
            float t = b[r];
            b[r] = b[c];
            b[c] = t;
        }
        if (m[c][c] == 0.0f)
            return 0;
        float pinv = (float)(1.0 / (double)m[c][c]);
        m[c][c] = 1.0f;
        for (int l = 0; l < 3; l++)
            m[c][l] *= pinv;
        b[c] *= pinv;
        for (int ll = 0; ll < 3; ll++)
            if (ll != c) {
                float d = m[ll][c];
                m[ll][c] = 0.0f;
                for (int l = 0; l < 3; l++)
                    m[ll][l] -= m[c][l] * d;
                b[ll] -= b[c] * d;
            }
    }
    return 1;
}

// gaussPeak1D: returns sigma = width / 2 of the fit through y[0..17] at
// abscissae x0, x0+1, ...
__device__ inline float gauss_peak_1d(const float *yin, float x0)
{
    float x[NGAUSS], y[NGAUSS];
    float maxC = 0.0f;
    for (int i = 0; i < NGAUSS; i++) {
        x[i] = x0 + (float)i;
        y[i] = yin[i];
        maxC = y[i] > maxC ? y[i] : maxC;
    }
    for (int i = 0; i < NGAUSS; i++)
        y[i] /= maxC;
    float a[3] = {1.0f, x[NGAUSS / 2], 2.0f};
    float alpha[3][3], beta[3], ochisq, chisq;
    float lambda = 0.001f;
    lm_coef(x, y, a, alpha, beta, &chisq);
    ochisq = chisq;
    float lasti = 2e9f;
    for (int it = 0; it < 5; it++) {
        float covar[3][3], da[3];
        for (int j = 0; j < 3; j++) {
            for (int k = 0; k < 3; k++)
                covar[j][k] = alpha[j][k];
            covar[j][j] = (float)((double)alpha[j][j] * (1.0 + (double)lambda));
            da[j] = beta[j];
        }
        if (!solve3(covar, da))
            break;
        float atry[3];
        for (int l = 0; l < 3; l++)
            atry[l] = a[l] + da[l];
        float nalpha[3][3], nbeta[3];
        lm_coef(x, y, atry, nalpha, nbeta, &chisq);
        if (chisq < ochisq) {
            lambda = (float)((double)lambda * 0.1);
            ochisq = chisq;
            for (int j = 0; j < 3; j++) {
                for (int k = 0; k < 3; k++)
                    alpha[j][k] = nalpha[j][k];
                beta[j] = nbeta[j];
            }
            for (int l = 0; l < 3; l++)
                a[l] = atry[l];
        } else {
            lambda = (float)((double)lambda * 10.0);
        }
        if (fabs((double)(a[1] - lasti)) < 0.001)
            break;
        lasti = a[1];
    }
    return (float)(a[2] * 0.5f);
}
// 2-D Gaussian LM fit of the -gauss peak mode (cmpTrackGauss / mrqminMod /
// mrqcofMod / myGauss, speckleTrack.c:2205-2303, 2309-2443), in float like
// the reference, one thread per fit.  Parameters a = {amplitude, mu_i,
// sigma_i, mu_j, sigma_j}; data: the 18 x 18 magnitudes around the coarse
// peak times `scale`, weights 1.
__device__ inline void lm_gauss2d(float xx, float yy, const float a[5], float *y, float dyda[5])
{
    const float dx = xx - a[1], dy = yy - a[3];
    const float argx = 0.5f * (dx * dx) / (a[2] * a[2]);
    const float argy = 0.5f * (dy * dy) / (a[4] * a[4]);
    const float ex = (float)exp((double)(-argx - argy));
    *y = a[0] * ex;
    dyda[0] = ex;
    dyda[1] = (*y) * (xx - a[1]) / (a[2] * a[2]);
    dyda[2] = *y * (xx - a[1]) * (xx - a[1]) / (a[2] * a[2] * a[2]);
    dyda[3] = *y * (yy - a[3]) / (a[4] * a[4]);
    dyda[4] = *y * (yy - a[3]) * (yy - a[3]) / (a[4] * a[4] * a[4]);
}

__device__ inline void lm2d_coef(const float *M, int W, int i0, int j0, double scale, const float a[5],
                                 float alpha[5][5], float beta[5], float *chisq)
{
    for (int j = 0; j < 5; j++) {
        for (int k = 0; k <= j; k++)
            alpha[j][k] = 0.0f;
        beta[j] = 0.0f;
    }
    *chisq = 0.0f;
    for (int i = 0; i < NGAUSS; i++)
        for (int j = 0; j < NGAUSS; j++) {
            const float g = (float)((double)M[(i0 + i) * W + j0 + j] * scale);
            float ymod, dyda[5];
            lm_gauss2d((float)(i0 + i), (float)(j0 + j), a, &ymod, dyda);
            const float dy = g - ymod;
            for (int l = 0; l < 5; l++) {
                const float wt = dyda[l];
                for (int m = 0; m <= l; m++)
                    alpha[l][m] += wt * dyda[m];
                beta[l] += dy * wt;
            }
            *chisq += dy * dy;
        }
    for (int j = 1; j < 5; j++)
        for (int k = 0; k < j; k++)
            alpha[k][j] = alpha[j][k];
}

// Gauss-Jordan with full pivoting on a 5x5 system; 0 when singular.
__device__ inline int solve5(float m[5][5], float b[5])
{
    int used[5] = {0, 0, 0, 0, 0};
    for (int it = 0; it < 5; it++) {
        float big = 0;
        int r = -1, c = -1;
        for (int j = 0; j < 5; j++) {
            if (used[j])
                continue;
            for (int k = 0; k < 5; k++)
                if (!used[k] && fabsf(m[j][k]) >= big) {
                    big = fabsf(m[j][k]);
                    r = j;
                    c = k;
                }
        }
        if (r < 0)
            return 0;
        used[c] = 1;
        if (r != c) {
            for (int l = 0; l < 5; l++) {
                float t = m[r][l];
                m[r][l] = m[c][l];
                m[c][l] = t;
            }
            float t = b[r];
            b[r] = b[c];
            b[c] = t;
        }
        if (m[c][c] == 0.0f)
            return 0;
        float pinv = (float)(1.0 / (double)m[c][c]);
        m[c][c] = 1.0f;
        for (int l = 0; l < 5; l++)
            m[c][l] *= pinv;
        b[c] *= pinv;
        for (int ll = 0; ll < 5; ll++)
            if (ll != c) {
                float d = m[ll][c];
                m[ll][c] = 0.0f;
                for (int l = 0; l < 5; l++)
                    m[ll][l] -= m[c][l] * d;
                b[ll] -= b[c] * d;
            }
    }
    return 1;
}

// cmpTrackGauss on the fftshifted magnitude M (row stride W) around the
// coarse peak (I, J) with findCPeak's raw maximum cMax; wR / wA: the window
// of the surface; scale = hanningCorrection / (sqrt(p1 p2) (wR wA)^2).
// Returns the peak in 1/NOVER units and the fitted amplitude capped at 1, or
// -LARGEINT / 0 near the edge or when a width exceeds 2000.
__device__ inline void gauss_fit_2d(const float *M, int W, int I, int J, double cMax, int wR, int wA, double scale,
                                    int *oI, int *oJ, double *oC)
{
    constexpr int ng = NGAUSS;
    if ((I - ng / 2) < 0 || (I + ng / 2) >= (wA * 2 - 1) || (J - ng / 2) < 0 || (J + ng / 2) >= (wR * 2 - 1)) {
        *oI = *oJ = -2000000000;
        *oC = 0;
        return;
    }
    const int i0 = I - ng / 2, j0 = J - ng / 2;
    float a[5] = {(float)(cMax * scale), (float)I, 3.0f, (float)J, 3.0f};
    float alpha[5][5], beta[5], ochisq, chisq;
    float lambda = 0.001f;
    lm2d_coef(M, W, i0, j0, scale, a, alpha, beta, &chisq);
    ochisq = chisq;
    float lasti = 2e9f, lastj = 2.0e9f;
    for (int it = 0; it < 8; it++) {
        float covar[5][5], da[5];
        for (int j = 0; j < 5; j++) {
            for (int k = 0; k < 5; k++)
                covar[j][k] = alpha[j][k];
            covar[j][j] = (float)((double)alpha[j][j] * (1.0 + (double)lambda));
            da[j] = beta[j];
        }
        if (!solve5(covar, da))
            break;
        float atry[5];
        for (int l = 0; l < 5; l++)
            atry[l] = a[l] + da[l];
        float nalpha[5][5], nbeta[5];
        lm2d_coef(M, W, i0, j0, scale, atry, nalpha, nbeta, &chisq);
        if (chisq < ochisq) {
            lambda = (float)((double)lambda * 0.1);
            ochisq = chisq;
            for (int j = 0; j < 5; j++) {
                for (int k = 0; k < 5; k++)
                    alpha[j][k] = nalpha[j][k];
                beta[j] = nbeta[j];
            }
            for (int l = 0; l < 5; l++)
                a[l] = atry[l];
        } else {
            lambda = (float)((double)lambda * 10.0);
        }
        if (fabs((double)(a[1] - lasti)) < 0.001 && fabs((double)(a[3] - lastj)) < 0.001)
            break;
        lasti = a[1];
        lastj = a[3];
    }
    *oI = (int)round((double)(a[1] * (float)NOVER));
    *oJ = (int)round((double)(a[3] * (float)NOVER));
    *oC = (double)a[0] < 1.0 ? (double)a[0] : 1.0;
    if ((0.5 * (double)a[2]) > 1000.0 || (0.5 * (double)a[4]) > 1000.0) {
        *oI = *oJ = -2000000000;
        *oC = 0;
    }
}
#pragma clang fp contract(on)

// Doppler estimate of a w x w patch pair (estDopCarrier1,
// Strack/dopplerEstimation.c:25-88): lag-1 azimuth autocorrelation after
// complex mean removal, in double.
template <class Get>
__device__ inline int doppler_shift(Get get, int wA, int wR, int nAz, int wAnorm, double *sh)
{
    const int N = wA * wR;
    double m[4] = {0, 0, 0, 0};
    for (int e = threadIdx.x; e < N; e += NT) {
        float2 a, b;
        get(e, a, b);
        m[0] += a.x;
        m[1] += a.y;
        m[2] += b.x;
        m[3] += b.y;
    }
    for (int k = 0; k < 4; k++)
        m[k] = block_sum_d<NT>(m[k], sh) / ((double)wA * (double)wR);
    double z[4] = {0, 0, 0, 0};
    for (int e = threadIdx.x; e < N - wR; e += NT) {
        float2 a0, b0, a1, b1;
        get(e, a0, b0);
        get(e + wR, a1, b1);
        double ar = a1.x - m[0], ai = a1.y - m[1], br = a0.x - m[0], bi = a0.y - m[1];
        z[0] += ar * br + ai * bi;
        z[1] += ai * br - ar * bi;
        ar = b1.x - m[2], ai = b1.y - m[3], br = b0.x - m[2], bi = b0.y - m[3];
        z[2] += ar * br + ai * bi;
        z[3] += ai * br - ar * bi;
    }
    for (int k = 0; k < 4; k++)
        z[k] = block_sum_d<NT>(z[k], sh);
    double ang1 = (z[0] == 0.0 && z[1] == 0.0) ? NAN : atan2(z[1], z[0]);
    double ang2 = (z[2] == 0.0 && z[3] == 0.0) ? NAN : atan2(z[3], z[2]);
    double v = (double)(nAz / 2) * (ang1 + ang2) * .5 / 3.14159265358979323846;
    int shift = isnan(v) ? 0 : ((v >= 0.0) ? (int)(v + 0.5) : (int)(v - 0.5));
    if (shift < 0)
        shift += nAz;
    return (int)((float)shift * (float)wAnorm / (float)nAz);
}

}  // namespace sttrack
