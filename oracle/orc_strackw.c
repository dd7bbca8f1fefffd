/*
 * ORACLE -- test infrastructure (see orc_common.h).  Not the product.
 *
 * CPU restatement of strackw's edge-padded normalised cross-correlation
 * tracker, Strackw/corrTrackFast.c of fastice/speckleSource, on whole
 * in-memory images (the reference streams 2000-row strips,
 * corrTrackFast.c:554-602; the strip buffer does not change results).
 *
 * Kept quirks: edgePad += NFAST/4 internally (:116-117), the corner-only
 * bounds test (:170-171), boxcar smoothing only when navg > 1 with the
 * azimuth clamp written in terms of edgePadR (:790-799), the N-1 chip
 * variance minus mean^2 (:428-430), the float-stored running maximum of
 * the oversampled search (:532-543), type = AMPMATCH even on failure
 * (:402).  Deviations pinned here: samples outside image 1 / 2 read as 0
 * (the reference reads past the strip buffer there); the initial-offset
 * grid is interpolated with a plain bilinear rule because bilinearInterp
 * (mosaicSource) is not vendored ("parity unpinned").
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc_common.h"
#include "../include/speckle_amd.h"

#define NOVER 16 /* corrTrackFast.c:11 */
#define NFAST 20 /* corrTrackFast.c:12 */
#define OS 2     /* corrTrackFast.c:16 */
#define AMPMATCH 2

typedef struct {
    const st_track_params *par;
    const void *img1, *img2;
    int nr1, na1, nr2, na2, stype;
    int wRa, wAa, W_R, W_A, epR, epA, wR2, wA2, nLagR, nLagA;
    int azShift;
    ocplx *im1in, *im2in, *f1a, *f2a, *img1c, *img2c, *cbuf;
    oreal *dataR, *dataS, *mag;
    double *tmpS1, *tmpS2, *meanS, *sigmaS, *corrResult;
    ocplx cFast[(NFAST + 1) * (NFAST + 1)];
    ocplx *cFastOver;
    float gapCoarse, gapFine;
} orc_stw;

static inline void sample0(const void *img, int stype, int nr, int na, int a, int r, oreal *re, oreal *im)
{
    if (a < 0 || a >= na || r < 0 || r >= nr) {
        *re = 0;
        *im = 0;
        return;
    }
    if (stype == ST_SAMPLE_C64) {
        const float *p = (const float *)img + 2 * ((size_t)a * nr + r);
        *re = (oreal)p[0];
        *im = (oreal)p[1];
    } else {
        const int16_t *p = (const int16_t *)img + 2 * ((size_t)a * nr + r);
        *re = (oreal)p[0];
        *im = (oreal)p[1];
    }
}

/* dopplerEstimation.c:42-88 (same routine strack uses) */
static double single_dop(const ocplx *img, int nAz, int nRg)
{
    double mean_re = 0., mean_im = 0., sum_re = 0.0, sum_im = 0.0;
    long double denom = (long double)nRg * (long double)nAz;
    for (int j = 0; j < nRg; ++j)
        for (int i = 0; i < nAz; ++i) {
            mean_re += (long double)img[i * nRg + j].re;
            mean_im += (long double)img[i * nRg + j].im;
        }
    mean_re /= denom;
    mean_im /= denom;
    for (int j = 0; j < nRg; ++j)
        for (int i = 0; i < nAz - 1; ++i) {
            const ocplx *p0 = &img[i * nRg + j], *p1 = &img[(i + 1) * nRg + j];
            double a_re = (long double)p1->re - mean_re, a_im = (long double)p1->im - mean_im;
            double b_re = (long double)p0->re - mean_re, b_im = (long double)p0->im - mean_im;
            sum_re += a_re * b_re + a_im * b_im;
            sum_im += a_im * b_re - a_re * b_im;
        }
    if (sum_re == 0.0 && sum_im == 0.0)
        return NAN;
    return atan2(sum_im, sum_re);
}

static int iround_d(double x) { return (x >= 0.0) ? (int)(x + 0.5) : (int)(x - 0.5); }

/* Plain bilinear interpolation standing in for mosaicSource's
 * bilinearInterp (corrTrackFast.c:851-852). */
static float bilinear(const float *g, double x, double y, int nx, int ny, float minvalue, float nodata)
{
    if (!(x >= 0.0 && y >= 0.0 && x <= nx - 1 && y <= ny - 1))
        return nodata;
    int ix = (int)x, iy = (int)y;
    if (ix > nx - 2)
        ix = nx - 2;
    if (iy > ny - 2)
        iy = ny - 2;
    if (ix < 0 || iy < 0)
        return nodata;
    double t = x - ix, u = y - iy;
    float p1 = g[(size_t)iy * nx + ix], p2 = g[(size_t)iy * nx + ix + 1];
    float p3 = g[(size_t)(iy + 1) * nx + ix + 1], p4 = g[(size_t)(iy + 1) * nx + ix];
    if (p1 <= minvalue || p2 <= minvalue || p3 <= minvalue || p4 <= minvalue)
        return nodata;
    return (float)((1.0 - t) * (1.0 - u) * p1 + t * (1.0 - u) * p2 + t * u * p3 + (1.0 - t) * u * p4);
}

/* findImage2Pos + getShifts (strackw), corrTrackFast.c:257-292, 843-858 */
static void find_image2_pos_w(const st_track_params *par, int r1, int a1, int *r2, int *a2)
{
    double aShift, rShift;
    double sc = (double)par->scaleFactor;
    if (par->polyShift || par->initDr == NULL) {
        rShift = par->rShiftPoly[0] + (double)(r1 * sc) * par->rShiftPoly[1] + (double)(a1 * sc) * par->rShiftPoly[2];
        aShift = par->aShiftPoly[0] + (double)(r1 * sc) * par->aShiftPoly[1] + (double)(a1 * sc) * par->aShiftPoly[2];
    } else {
        float minvalue = (float)-ORC_LARGEINT * 0.9f;
        double range = (r1 - par->initR0) / par->initDeltaR;
        double azimuth = (a1 - par->initA0) / par->initDeltaA;
        rShift = bilinear(par->initDr, range, azimuth, par->initNr, par->initNa, minvalue, (float)-ORC_LARGEINT);
        aShift = bilinear(par->initDa, range, azimuth, par->initNr, par->initNa, minvalue, (float)-ORC_LARGEINT);
        if (rShift <= minvalue || aShift <= minvalue) {
            rShift = minvalue;
            aShift = minvalue;
        }
        if (rShift < (-ORC_LARGEINT + 1)) {
            rShift = par->rShiftPoly[0] + (double)(r1 * sc) * par->rShiftPoly[1] + (double)(a1 * sc) * par->rShiftPoly[2];
            aShift = par->aShiftPoly[0] + (double)(r1 * sc) * par->aShiftPoly[1] + (double)(a1 * sc) * par->aShiftPoly[2];
        }
    }
    rShift /= sc;
    aShift /= sc;
    *r2 = (rShift >= 0) ? r1 + (int)(rShift + 0.5) : r1 + (int)(rShift - 0.5);
    *a2 = (aShift >= 0) ? a1 + (int)(aShift + 0.5) : a1 + (int)(aShift - 0.5);
}

/* maskValue (strackw), corrTrackFast.c:1141-1160 */
static int mask_value_w(const st_track_params *p, int wRa, int wAa, int r1, int a1)
{
    if (p->mask == NULL)
        return 1;
    int dr = (wRa / 2 - p->maskR0) * p->scaleFactor;
    int da = (wAa / 2 - p->maskA0) * p->scaleFactor;
    int ia1 = (a1 + da) / p->maskNal, ir1 = (r1 + dr) / p->maskNrl;
    if (ia1 < 0 || ia1 >= p->maskNa || ir1 < 0 || ir1 >= p->maskNr)
        return 0;
    return (int)p->mask[(size_t)ia1 * p->maskNr + ir1];
}

static void zero_pad(ocplx *pad, int pSA, int pSR, const ocplx *orig, int oSA, int oSR, int dA)
{
    memset(pad, 0, sizeof(ocplx) * (size_t)pSA * pSR);
    for (int i = 0; i < oSA / 2; i++) {
        int i1 = pSA - oSA / 2 + i, i2 = (oSA / 2 + i + dA) % oSA, ia = (i + dA) % oSA;
        int j1 = pSR - oSR / 2, j2 = oSR / 2;
        for (int j = 0; j < oSR / 2; j++) {
            int ja = j % oSR, j2a = j2 % oSR;
            pad[i * pSR + j] = orig[ia * oSR + ja];
            pad[i1 * pSR + j] = orig[i2 * oSR + ja];
            pad[i1 * pSR + j1] = orig[i2 * oSR + j2a];
            pad[i * pSR + j1] = orig[ia * oSR + j2a];
            j1++;
            j2++;
        }
    }
}

/* detectPatch, corrTrackFast.c:765-839 */
static void detect_patch(orc_stw *s, oreal *data, ocplx *im, int wR, int wA, int epR, int epA, float scale, int large)
{
    const st_track_params *p = s->par;
    int extraR = large ? 0 : p->navgR, extraA = large ? 0 : p->navgA;
    int W_R = s->W_R, W_A = s->W_A;
    for (int i = 0; i < wA; i++) {
        int i1 = epA * OS + i;
        for (int j = 0; j < wR; j++) {
            int j1 = epR * OS + j;
            oreal v;
            if (p->navgA > 1 || p->navgR > 1) {
                v = 0;
                for (int m = -p->navgA; m <= p->navgA; m++) {
                    int k = i1 + m < 0 ? 0 : i1 + m;
                    int kmax = epR * OS + wA + extraA - 1;
                    k = k < kmax ? k : kmax;
                    for (int n = -p->navgR; n <= p->navgR; n++) {
                        int l = j1 + n < 0 ? 0 : j1 + n;
                        int lmax = epR * OS + wR + extraR - 1;
                        l = l < lmax ? l : lmax;
                        ocplx z = im[k * W_R + l];
                        v += (z.re * z.re + z.im * z.im) * scale;
                    }
                }
            } else {
                ocplx z = im[i1 * W_R + j1];
                v = (z.re * z.re + z.im * z.im) * scale;
            }
            data[i * wR + j] = v;
        }
    }
    for (int i = 0; i < wA; i++) {
        int i1 = epA * OS + i;
        for (int j = 0; j < wR; j++) {
            int j1 = epR * OS + j;
            data[i * wR + j] = (oreal)sqrt((double)data[i * wR + j]);
            im[i1 * W_R + j1].re = data[i * wR + j];
            im[i1 * W_R + j1].im = 0.0;
        }
    }
    for (int i = 0; i < W_A; i++)
        for (int j = 0; j < epR * OS; j++) {
            im[i * W_R + j].re = im[i * W_R + j].im = 0;
            im[i * W_R + W_R - j - 1].re = im[i * W_R + W_R - j - 1].im = 0;
        }
    for (int i = 0; i < epA * OS; i++)
        for (int j = 0; j < W_R; j++) {
            im[i * W_R + j].re = im[i * W_R + j].im = 0;
            im[(W_A - i - 1) * W_R + j].re = im[(W_A - i - 1) * W_R + j].im = 0;
        }
}

/* computeMeanSigma, corrTrackFast.c:407-477 (double accumulators) */
static void mean_sigma(orc_stw *s, double *meanR, double *sigmaR)
{
    int wR2 = s->wR2, wA2 = s->wA2, W_R = s->W_R, W_A = s->W_A;
    *sigmaR = 0;
    *meanR = 0;
    for (int la = 0; la < wA2; la++)
        for (int lr = 0; lr < wR2; lr++) {
            oreal d = s->dataR[la * wR2 + lr];
            *sigmaR += d * d;
            *meanR += d;
        }
    *meanR /= (double)(wR2 * wA2);
    *sigmaR = *sigmaR / (double)(wR2 * wA2 - 1.);
    *sigmaR -= *meanR * (*meanR);
    for (int l = 0; l < W_A; l++) {
        double sum1 = 0.0, sum2 = 0.0;
        int s1 = 0;
        for (int c = 0; c < W_R; c++) {
            oreal d = s->dataS[l * W_R + c];
            sum1 += d;
            sum2 += d * d;
            if (c >= (wR2 - 1)) {
                int s2 = c - (wR2 - 1);
                s->tmpS1[l * W_R + s1] = sum1;
                s->tmpS2[l * W_R + s1] = sum2;
                oreal e = s->dataS[l * W_R + s2];
                sum1 -= e;
                sum2 -= e * e;
                s1++;
            }
        }
    }
    for (int c = 0; c < s->nLagR; c++) {
        double sum1 = 0.0, sum2 = 0.0;
        int l1 = 0;
        for (int l = 0; l < W_A; l++) {
            sum1 += s->tmpS1[l * W_R + c];
            sum2 += s->tmpS2[l * W_R + c];
            if (l >= (wA2 - 1)) {
                int l2 = l - (wA2 - 1);
                double m = sum1 / (double)(wR2 * wA2);
                s->meanS[l1 * s->nLagR + c] = m;
                s->sigmaS[l1 * s->nLagR + c] = sum2 / (double)(wR2 * wA2 - 1) - m * m;
                sum1 -= s->tmpS1[l2 * W_R + c];
                sum2 -= s->tmpS2[l2 * W_R + c];
                l1++;
            }
        }
    }
}

/* overSampleC (strackw), corrTrackFast.c:951-990: 21x21 -> 336x336 */
static void over_sample_w(orc_stw *s)
{
    const int n = NFAST + 1, N = (NFAST + 1) * NOVER, half = NFAST / 2;
    ocplx ps[(NFAST + 1) * (NFAST + 1)];
    memcpy(ps, s->cFast, sizeof(ps));
    orc_fft2d(ps, n, n, ORC_FORWARD);
    memset(s->cFastOver, 0, sizeof(ocplx) * (size_t)N * N);
    for (int i = 0; i <= half; i++) {
        int i1 = N - half - 1 + i, i2 = half + i;
        int j1 = N - half - 1, j2 = half;
        for (int j = 0; j <= half; j++) {
            s->cFastOver[i * N + j] = ps[i * n + j];
            if (j != 0)
                s->cFastOver[i * N + j1] = ps[i * n + j2];
            if (i != 0) {
                s->cFastOver[i1 * N + j] = ps[i2 * n + j];
                if (j != 0)
                    s->cFastOver[i1 * N + j1] = ps[i2 * n + j2];
            }
            j1++;
            j2++;
        }
    }
    orc_fft2d(s->cFastOver, N, N, ORC_BACKWARD);
}

/* getCorrPatchesFast + correlateFast + corrMatch:
 * corrTrackFast.c:299-404, 482-551, 862-947.  Returns cMax. */
static float track_point(orc_stw *s, int r1, int a1, int r2, int a2, float *rShift1, float *aShift1, int *pI, int *pJ,
                         int *fI, int *fJ)
{
    const st_track_params *p = s->par;
    int wRa = s->wRa, wAa = s->wAa, W_R = s->W_R, W_A = s->W_A;
    float scale = (float)(1. / ((float)wRa * (float)wAa * (float)wRa * (float)wAa * (float)(OS * OS * OS * OS) *
                                (OS * OS * (p->navgA + 1) * (p->navgR + 1))));
    for (int i = 0; i < wAa; i++)
        for (int j = 0; j < wRa; j++) {
            sample0(s->img1, s->stype, s->nr1, s->na1, a1 + i, r1 + j, &s->im1in[i * wRa + j].re, &s->im1in[i * wRa + j].im);
            sample0(s->img2, s->stype, s->nr2, s->na2, a2 + i, r2 + j, &s->im2in[i * wRa + j].re, &s->im2in[i * wRa + j].im);
        }
    /* estDopCarrier1 with wA forced to wAa (corrTrackFast.c:322-323) */
    {
        double ang1 = single_dop(s->im1in, wAa, wRa), ang2 = single_dop(s->im2in, wAa, wRa);
        double v = (double)(wAa / 2) * (ang1 + ang2) * .5 / ORC_PI;
        int shift = isnan(v) ? 0 : iround_d(v);
        if (shift < 0)
            shift += wAa;
        s->azShift = (int)((float)shift * (float)wAa / (float)wAa);
    }
    orc_fft2d(s->im1in, wAa, wRa, ORC_FORWARD);
    orc_fft2d(s->im2in, wAa, wRa, ORC_FORWARD);
    int da = s->azShift;
    if (da < 0 || da >= wAa)
        da = 0;
    zero_pad(s->img1c, W_A, W_R, s->im1in, wAa, wRa, da);
    zero_pad(s->img2c, W_A, W_R, s->im2in, wAa, wRa, da);
    orc_fft2d(s->img1c, W_A, W_R, ORC_BACKWARD);
    orc_fft2d(s->img2c, W_A, W_R, ORC_BACKWARD);
    detect_patch(s, s->dataR, s->img2c, s->wR2, s->wA2, s->epR, s->epA, scale, 0);
    detect_patch(s, s->dataS, s->img1c, W_R, W_A, 0, 0, scale, 1);
    /* correlateFast */
    double meanR, sigmaR;
    mean_sigma(s, &meanR, &sigmaR);
    /* ampMatchEdge */
    orc_fft2d(s->img1c, W_A, W_R, ORC_FORWARD);
    orc_fft2d(s->img2c, W_A, W_R, ORC_FORWARD);
    for (int k = 0; k < W_A * W_R; k++) {
        ocplx a = s->img1c[k], b = s->img2c[k];
        s->cbuf[k].re = a.re * b.re + a.im * b.im;
        s->cbuf[k].im = a.im * b.re - a.re * b.im;
    }
    orc_fft2d(s->cbuf, W_A, W_R, ORC_BACKWARD);
    for (int i = 0; i < W_A / 2; i++) {
        int i1 = W_A / 2 + i, j1 = W_R / 2;
        for (int j = 0; j < W_R / 2; j++) {
            ocplx z;
            z = s->cbuf[i1 * W_R + j1];
            s->mag[i * W_R + j] = (oreal)sqrt((double)(z.re * z.re + z.im * z.im));
            z = s->cbuf[i * W_R + j];
            s->mag[i1 * W_R + j1] = (oreal)sqrt((double)(z.re * z.re + z.im * z.im));
            z = s->cbuf[i * W_R + j1];
            s->mag[i1 * W_R + j] = (oreal)sqrt((double)(z.re * z.re + z.im * z.im));
            z = s->cbuf[i1 * W_R + j];
            s->mag[i * W_R + j1] = (oreal)sqrt((double)(z.re * z.re + z.im * z.im));
            j1++;
        }
    }
    /* getPeakCorr */
    double wR2d = wRa - 2 * s->epR, wA2d = wAa - 2 * s->epA;
    double scaleW = 1.0 / ((OS * OS * OS * OS) * wAa * wA2d * wRa * wR2d);
    double maxCorr = -1e30, second = -1e30;
    int iMax = 0, jMax = 0;
    for (int l = -OS * s->epA; l <= OS * s->epA; l++) {
        int i = l + OS * s->epA, m = l + wAa * OS / 2;
        for (int c = -OS * s->epR; c <= OS * s->epR; c++) {
            int j = c + OS * s->epR, n = c + wRa * OS / 2;
            double cr = s->mag[m * W_R + n] * scaleW;
            cr = cr - (meanR * s->meanS[i * s->nLagR + j]);
            cr = cr / sqrt(sigmaR * s->sigmaS[i * s->nLagR + j]);
            s->corrResult[i * s->nLagR + j] = cr;
            if (l >= -OS * (s->epA - NFAST / 4) && c >= -OS * (s->epR - NFAST / 4) && l <= OS * (s->epA - NFAST / 4) &&
                c <= OS * (s->epR - NFAST / 4)) {
                if (cr > maxCorr) {
                    second = maxCorr;
                    maxCorr = cr;
                    iMax = i;
                    jMax = j;
                } else if (cr > second)
                    second = cr;
            }
        }
    }
    s->gapCoarse = maxCorr > 0 ? (float)((maxCorr - second) / maxCorr) : 0.f;
    *pI = iMax;
    *pJ = jMax;
    *fI = *fJ = -2000;
    if (maxCorr < 0)
        return 0.0f;
    for (int i = 0; i <= NFAST; i++) {
        int i1 = iMax - NFAST / 2 + i, j1 = jMax - NFAST / 2;
        for (int j = 0; j <= NFAST; j++) {
            s->cFast[i * (NFAST + 1) + j].re = (oreal)s->corrResult[i1 * s->nLagR + j1];
            s->cFast[i * (NFAST + 1) + j].im = 0.0;
            j1++;
        }
    }
    over_sample_w(s);
    const int N = (NFAST + 1) * NOVER;
    float cMax = -1.0e30f;
    double cMaxD = -1e30, secondD = -1e30;
    int iMax1 = 0, jMax1 = 0;
    for (int i = ((NFAST + 1) * NOVER) / 2 - NOVER * 2; i < ((NFAST + 1) * NOVER) / 2 + NOVER * 2; i++)
        for (int j = ((NFAST + 1) * NOVER) / 2 - NOVER * 2; j < ((NFAST + 1) * NOVER) / 2 + NOVER * 2; j++) {
            ocplx z = s->cFastOver[i * N + j];
            double corr = (z.re * z.re + z.im * z.im);
            if (corr > cMax) {
                cMax = (float)corr;
                iMax1 = i;
                jMax1 = j;
            }
            if (corr > cMaxD) {
                secondD = cMaxD;
                cMaxD = corr;
            } else if (corr > secondD)
                secondD = corr;
        }
    s->gapFine = cMaxD > 0 ? (float)((cMaxD - secondD) / cMaxD) : 0.f;
    cMax = (float)(sqrt(cMax) / ((NFAST + 1) * (NFAST + 1)));
    iMax = (iMax * NOVER) + iMax1 - (NOVER * (NFAST)) / 2;
    jMax = (jMax * NOVER) + jMax1 - (NOVER * (NFAST)) / 2;
    *fI = iMax;
    *fJ = jMax;
    *rShift1 = ((float)jMax - s->epR * OS * NOVER) / (NOVER * OS);
    *aShift1 = ((float)iMax - s->epA * OS * NOVER) / (NOVER * OS);
    return cMax;
}

static int stw_setup(orc_stw *s, const st_track_params *par, const void *img1, int nr1, int na1, const void *img2,
                     int nr2, int na2, int stype)
{
    memset(s, 0, sizeof(*s));
    s->par = par;
    s->img1 = img1;
    s->img2 = img2;
    s->nr1 = nr1;
    s->na1 = na1;
    s->nr2 = nr2;
    s->na2 = na2;
    s->stype = stype;
    s->wRa = par->wRa;
    s->wAa = par->wAa;
    s->W_R = 2 * par->wRa;
    s->W_A = 2 * par->wAa;
    s->epR = par->edgePadR + NFAST / 4;
    s->epA = par->edgePadA + NFAST / 4;
    s->wR2 = (s->wRa - 2 * s->epR) * OS;
    s->wA2 = (s->wAa - 2 * s->epA) * OS;
    s->nLagR = 2 * s->epR * OS + 1;
    s->nLagA = 2 * s->epA * OS + 1;
    if (s->wR2 <= 0 || s->wA2 <= 0 || (s->wRa & 1) || (s->wAa & 1))
        return ST_ERR_ARG;
    size_t nbig = (size_t)s->W_A * s->W_R, nsm = (size_t)s->wAa * s->wRa, nlag = (size_t)s->nLagA * s->nLagR;
    s->im1in = (ocplx *)calloc(nsm, sizeof(ocplx));
    s->im2in = (ocplx *)calloc(nsm, sizeof(ocplx));
    s->img1c = (ocplx *)calloc(nbig, sizeof(ocplx));
    s->img2c = (ocplx *)calloc(nbig, sizeof(ocplx));
    s->cbuf = (ocplx *)calloc(nbig, sizeof(ocplx));
    s->dataR = (oreal *)calloc((size_t)s->wA2 * s->wR2, sizeof(oreal));
    s->dataS = (oreal *)calloc(nbig, sizeof(oreal));
    s->mag = (oreal *)calloc(nbig, sizeof(oreal));
    s->tmpS1 = (double *)calloc(nbig, sizeof(double));
    s->tmpS2 = (double *)calloc(nbig, sizeof(double));
    s->meanS = (double *)calloc(nlag, sizeof(double));
    s->sigmaS = (double *)calloc(nlag, sizeof(double));
    s->corrResult = (double *)calloc(nlag, sizeof(double));
    s->cFastOver = (ocplx *)calloc((size_t)(NFAST + 1) * NOVER * (NFAST + 1) * NOVER, sizeof(ocplx));
    return ST_OK;
}

static void stw_free(orc_stw *s)
{
    free(s->im1in);
    free(s->im2in);
    free(s->img1c);
    free(s->img2c);
    free(s->cbuf);
    free(s->dataR);
    free(s->dataS);
    free(s->mag);
    free(s->tmpS1);
    free(s->tmpS2);
    free(s->meanS);
    free(s->sigmaS);
    free(s->corrResult);
    free(s->cFastOver);
}

/* Test hook: one patch, returning the coarse NCC surface (nLagA x nLagR),
 * the detected search (2wAa x 2wRa) and chip (wA2 x wR2) images and
 * res = {cMax, rShift1, aShift1, azShift}. */
int orc_stw_point(const st_track_params *par, const void *img1, int nr1, int na1, const void *img2, int nr2, int na2,
                  int stype, int r1, int a1, int r2, int a2, double *corrResult, float *dataS, float *dataR,
                  float *res)
{
    orc_stw S;
    int rc = stw_setup(&S, par, img1, nr1, na1, img2, nr2, na2, stype);
    if (rc != ST_OK)
        return rc;
    float rs = 0, as = 0;
    int pI, pJ, fI, fJ;
    float c = track_point(&S, r1, a1, r2, a2, &rs, &as, &pI, &pJ, &fI, &fJ);
    size_t nlag = (size_t)S.nLagA * S.nLagR;
    for (size_t k = 0; corrResult && k < nlag; k++)
        corrResult[k] = S.corrResult[k];
    for (size_t k = 0; dataS && k < (size_t)S.W_A * S.W_R; k++)
        dataS[k] = (float)S.dataS[k];
    for (size_t k = 0; dataR && k < (size_t)S.wA2 * S.wR2; k++)
        dataR[k] = (float)S.dataR[k];
    res[0] = c;
    res[1] = rs;
    res[2] = as;
    res[3] = (float)S.azShift;
    stw_free(&S);
    return ST_OK;
}

/* Test hook: overSampleC of strackw on a 21 x 21 real patch; pw receives
 * |z|^2 of the 336 x 336 result. */
void orc_stw_oversample(const float *c, float *pw)
{
    orc_stw S;
    memset(&S, 0, sizeof(S));
    const int n = NFAST + 1, N = n * NOVER;
    S.cFastOver = (ocplx *)calloc((size_t)N * N, sizeof(ocplx));
    for (int k = 0; k < n * n; k++) {
        S.cFast[k].re = (oreal)c[k];
        S.cFast[k].im = 0;
    }
    over_sample_w(&S);
    for (int k = 0; k < N * N; k++)
        pw[k] = (float)(S.cFastOver[k].re * S.cFastOver[k].re + S.cFastOver[k].im * S.cFastOver[k].im);
    free(S.cFastOver);
}

int orc_corr_track_fast(const st_track_params *par, const void *img1, int nr1, int na1, const void *img2, int nr2,
                        int na2, int stype, st_track_out *out, st_track_diag *dg)
{
    orc_stw S;
    orc_stw *s = &S;
    int rc = stw_setup(s, par, img1, nr1, na1, img2, nr2, na2, stype);
    if (rc != ST_OK)
        return rc;
    int nAmp = 0, nFail = 0;
    for (int i = 0; i < par->nA; i++) {
        int a1 = par->aStart + i * par->deltaA - par->wAa / 2;
        for (int j = 0; j < par->nR; j++) {
            size_t o = (size_t)i * par->nR + j;
            int r1 = par->rStart + j * par->deltaR - par->wRa / 2;
            int r2, a2;
            find_image2_pos_w(par, r1, a1, &r2, &a2);
            float rShift = (float)(r2 - r1), aShift = (float)(a2 - a1);
            int maskVal = mask_value_w(par, s->wRa, s->wAa, r1 * par->scaleFactor, a1 * par->scaleFactor);
            out->offR[o] = (float)-ORC_LARGEINT;
            out->offA[o] = (float)-ORC_LARGEINT;
            out->corr[o] = 0;
            out->type[o] = 0;
            if (dg) {
                if (dg->r2) dg->r2[o] = r2;
                if (dg->a2) dg->a2[o] = a2;
                if (dg->azShift) dg->azShift[o] = -1;
                if (dg->peakI[0]) dg->peakI[0][o] = -2000;
                if (dg->peakJ[0]) dg->peakJ[0][o] = -2000;
                if (dg->fineI[0]) dg->fineI[0][o] = -2000;
                if (dg->fineJ[0]) dg->fineJ[0][o] = -2000;
                if (dg->cmax[0]) dg->cmax[0][o] = 0;
                if (dg->peakGap[0]) dg->peakGap[0][o] = 0;
                if (dg->fineGap[0]) dg->fineGap[0][o] = 0;
            }
            if (maskVal == 1) {
                if (r1 > 0 && r1 < nr1 && a1 > 0 && a1 < na1 && r2 > 0 && r2 < nr2 && a2 > 0 && a2 < na2) {
                    float rShift1 = 0, aShift1 = 0;
                    int pI, pJ, fI, fJ;
                    s->gapFine = 0;
                    float cMax = track_point(s, r1, a1, r2, a2, &rShift1, &aShift1, &pI, &pJ, &fI, &fJ);
                    if (cMax > .00) {
                        rShift -= rShift1;
                        aShift -= aShift1;
                        nAmp++;
                    } else {
                        rShift = (float)-ORC_LARGEINT;
                        aShift = (float)-ORC_LARGEINT;
                        nFail++;
                    }
                    out->offR[o] = rShift;
                    out->offA[o] = aShift;
                    out->corr[o] = cMax;
                    out->type[o] = AMPMATCH;
                    if (dg) {
                        if (dg->azShift) dg->azShift[o] = s->azShift;
                        if (dg->peakI[0]) dg->peakI[0][o] = pI;
                        if (dg->peakJ[0]) dg->peakJ[0][o] = pJ;
                        if (dg->fineI[0]) dg->fineI[0][o] = fI;
                        if (dg->fineJ[0]) dg->fineJ[0][o] = fJ;
                        if (dg->cmax[0]) dg->cmax[0][o] = cMax;
                        if (dg->peakGap[0]) dg->peakGap[0][o] = s->gapCoarse;
                        if (dg->fineGap[0]) dg->fineGap[0][o] = s->gapFine;
                    }
                }
            }
        }
    }
    out->nComplex = 0;
    out->nAmp = nAmp;
    out->nAmpL = 0;
    out->nFail = nFail;
    stw_free(s);
    return ST_OK;
}
