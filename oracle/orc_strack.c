/*
 * ORACLE -- test infrastructure (see orc_common.h).  Not the product.
 *
 * CPU restatement of strack's tracking driver and per-patch matcher,
 * Strack/speckleTrack.c and Strack/dopplerEstimation.c of
 * fastice/speckleSource, operating on whole in-memory images (the
 * reference streams 1000-row strips, speckleTrack.c:1369-1399; the strip
 * buffer is transparent to the result).
 *
 * Quirks kept (they change results): findCPeak's interleaved quadrant scan
 * with strict '>', Nyquist bins on the negative side of the spectral pads,
 * Hanning on patch 2 only, image-1-only mean masking in ampMatch, the
 * persistent azShift that the amplitude path inherits from the last
 * complex estimate, stale r2/a2 when the initial-offset grid has no value,
 * finalizeCmpxMatch only reached when the 18-sample Gaussian fit fits.
 * Quirks pinned to their benign value: rangeShift = 0 (never assigned in
 * the reference, speckleTrack.c:626-630), the uninitialised aZDefocus[i][j]
 * write in ampTrackFast (speckleTrack.c:537) is skipped, an all-zero
 * correlation surface rejects the patch, NaN Doppler angles give shift 0.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc_common.h"
#include "../include/speckle_amd.h"

#define NOVER 12 /* speckleTrack.c:15 */
#define NFAST 8  /* speckleTrack.c:16 */
#define LA 2     /* speckleTrack.c:18 */
#define OS 2     /* speckleTrack.c:25 */
#define OSA 2    /* speckleTrack.c:26 */
#define NGAUSS 18
#define AZ 1
#define RG 2
#define BAD 0
#define CMATCH 1
#define AMPMATCH 2
#define AMPMATCHLARGE 3

typedef struct {
    const st_track_params *par;
    const void *img1, *img2;
    int nr1, na1, nr2, na2, stype;
    int wR, wA, wRa, wAa;
    /* file-scope state of the reference */
    int azShift;   /* trackPar->azShift */
    double p1, p2; /* trackPar->p1 / p2 */
    int nComplex, nAmp, nAmpL, nFail;
    /* work */
    oreal *hanning;
    ocplx *patch1, *patch2;
    ocplx *ps, *cbuf;        /* padded spectrum / correlation surface */
    oreal *mag;              /* fftshifted |C| (cNoPadMag / caNoPadMag) */
    ocplx *im1in, *im2in, *f1a, *f2a, *im1, *im2;
    ocplx cFast[NFAST * NFAST];
    ocplx *cFastOver;        /* 96 x 96 */
    float *azd;              /* output aZDefocus grid */
    float gapCoarse, gapFine;
    double snr;              /* coarse peak over the mean surface magnitude */
    float sigRgLast, sigAzLast;
} orc_st;

/* ------------------------------------------------------------------ */
/* sample access                                                        */

static inline void sample(const void *img, int stype, int nr, int a, int r, oreal *re, oreal *im)
{
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

static inline oreal abs2(ocplx x) { return x.re * x.re + x.im * x.im; }

/* absCmpx(x, FALSE): float magnitude through a double sqrt
 * (speckleTrack.c:1810-1817). */
static inline oreal absc(ocplx x)
{
    oreal x2 = x.re * x.re + x.im * x.im;
    return (oreal)sqrt((double)x2);
}

/* ------------------------------------------------------------------ */
/* computeHanning, speckleTrack.c:1830-1852.  The reference stores the
 * window as one contiguous wA x wR float block with rows of wR
 * (mallocFloatMat(wA, wR), speckleTrack.c:1775) and indexes its columns
 * with wA / 2 (:1843-1844).  For square windows that is the separable
 * taper; otherwise hanning[i][j] lands at flat offset i * wR + j, which can
 * fall in a neighbouring row or outside the block.  Restated on the flat
 * block in the reference's loop order: writes inside [0, wA wR) land where
 * the reference's land (a later write to the same cell wins), writes
 * outside it are dropped (the reference corrupts its heap there) and cells
 * never written hold 0 (the reference leaves them indeterminate). */
static void hann_put(oreal *h, int wA, int wR, int i, int j, oreal v)
{
    const long f = (long)i * wR + j;
    if (f >= 0 && f < (long)wA * wR)
        h[f] = v;
}

static void compute_hanning(orc_st *s)
{
    int wA = s->wA, wR = s->wR;
    for (int i = 0; i < wA * wR; i++)
        s->hanning[i] = 0;
    for (int i = 0; i < wA / 2; i++) {
        int i1 = wA / 2 + i, i2 = wA / 2 - i - 1;
        double wi = 0.5 * (1.0 + cos((2.0 * ORC_PI) * (double)i / (double)(wA - 1)));
        for (int j = 0; j < wR / 2; j++) {
            int j1 = wA / 2 + j, j2 = wA / 2 - j - 1;
            double wj = 0.5 * (1.0 + cos((2.0 * ORC_PI) * (double)j / (double)(wR - 1)));
            /* stored as float (hanning is float**) */
            oreal h = (oreal)(float)(wi * wj);
            hann_put(s->hanning, wA, wR, i1, j1, h);
            hann_put(s->hanning, wA, wR, i2, j2, h);
            hann_put(s->hanning, wA, wR, i2, j1, h);
            hann_put(s->hanning, wA, wR, i1, j2, h);
        }
    }
}

/* ------------------------------------------------------------------ */
/* estDopCarrier1 / computeSingleDop, dopplerEstimation.c:25-88 */
static double single_dop(const ocplx *img, int nAz, int nRg)
{
    double mean_re = 0., mean_im = 0.;
    double sum_re = 0.0, sum_im = 0.0;
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
            double a_re = (long double)p1->re - mean_re;
            double a_im = (long double)p1->im - mean_im;
            double b_re = (long double)p0->re - mean_re;
            double b_im = (long double)p0->im - mean_im;
            double z_re = a_re * b_re + a_im * b_im;
            double z_im = a_im * b_re - a_re * b_im;
            sum_re += z_re;
            sum_im += z_im;
        }
    if (sum_re == 0.0 && sum_im == 0.0)
        return NAN;
    return atan2(sum_im, sum_re);
}

static int iround_d(double x) { return (x >= 0.0) ? (int)(x + 0.5) : (int)(x - 0.5); }

static void est_dop_carrier1(orc_st *s, const ocplx *im1, const ocplx *im2, int nAz, int nRg)
{
    double angle1 = single_dop(im1, nAz, nRg);
    double angle2 = single_dop(im2, nAz, nRg);
    double v = (double)(nAz / 2) * (angle1 + angle2) * .5 / ORC_PI;
    int shift = isnan(v) ? 0 : iround_d(v);
    if (shift < 0)
        shift += nAz;
    s->azShift = shift;
    s->azShift = (int)((float)s->azShift * (float)s->wA / (float)nAz);
}

/* ------------------------------------------------------------------ */
/* findImage2Pos + getShifts, speckleTrack.c:810-847, 1857-1940 */
static int get_shifts(const st_track_params *par, double range, double azimuth, double *dr, double *da)
{
    float minvalue = (float)-ORC_LARGEINT * 0.9f;
    range = (range - par->initR0) / par->initDeltaR;
    azimuth = (azimuth - par->initA0) / par->initDeltaA;
    if (range < 0.0 || azimuth < 0.0 || (int)range >= par->initNr || (int)azimuth >= par->initNa) {
        *dr = -ORC_LARGEINT;
        *da = -ORC_LARGEINT;
        return 0;
    }
    const float *rim = par->initDr, *aim = par->initDa;
    int nr = par->initNr;
    int j = (int)range, i = (int)azimuth;
    if (j == par->initNr - 1 || i == par->initNa - 1) {
        *dr = rim[i * nr + j];
        *da = aim[i * nr + j];
        if (*dr < minvalue || *da < minvalue) {
            *dr = -ORC_LARGEINT;
            *da = -ORC_LARGEINT;
        }
        return 0;
    }
    float t = (float)(range - (double)j), u = (float)(azimuth - (double)i);
    float p1 = rim[i * nr + j], p2 = rim[i * nr + j + 1], p3 = rim[(i + 1) * nr + j + 1], p4 = rim[(i + 1) * nr + j];
    if (p1 <= minvalue || p2 <= minvalue || p3 <= minvalue || p4 <= minvalue) {
        float m = fmaxf(fmaxf(fmaxf(p1, p2), p3), p4);
        *dr = m;
        if (*dr < minvalue) {
            *dr = -ORC_LARGEINT;
            *da = -ORC_LARGEINT;
            return 0;
        }
    } else
        *dr = (float)((1.0 - t) * (1.0 - u) * p1 + t * (1.0 - u) * p2 + t * u * p3 + (1.0 - t) * u * p4);
    p1 = aim[i * nr + j];
    p2 = aim[i * nr + j + 1];
    p3 = aim[(i + 1) * nr + j + 1];
    p4 = aim[(i + 1) * nr + j];
    if (p1 <= minvalue || p2 <= minvalue || p3 <= minvalue || p4 <= minvalue) {
        float m = fmaxf(fmaxf(fmaxf(p1, p2), p3), p4);
        *da = m;
        if (*da < minvalue) {
            *dr = -ORC_LARGEINT;
            *da = -ORC_LARGEINT;
            return 0;
        }
    } else
        *da = (float)((1.0 - t) * (1.0 - u) * p1 + t * (1.0 - u) * p2 + t * u * p3 + (1.0 - t) * u * p4);
    return 0;
}

static void find_image2_pos(const st_track_params *par, int r1, int a1, int *r2, int *a2)
{
    double aShift = 0, rShift = 0;
    double sc = (double)par->scaleFactor;
    int haveShift = 0;
    if (par->polyShift == 0) {
        get_shifts(par, r1, a1, &rShift, &aShift);
        if (rShift > (-ORC_LARGEINT + 1))
            haveShift = 1;
        else
            return; /* r2, a2 keep their previous values (speckleTrack.c:823-826) */
    }
    if (par->polyShift || !haveShift) {
        rShift = par->rShiftPoly[0] + (double)(r1 * sc) * par->rShiftPoly[1] + (double)(a1 * sc) * par->rShiftPoly[2];
        aShift = par->aShiftPoly[0] + (double)(r1 * sc) * par->aShiftPoly[1] + (double)(a1 * sc) * par->aShiftPoly[2];
    }
    rShift /= sc;
    aShift /= sc;
    *r2 = (rShift >= 0) ? r1 + (int)(rShift + 0.5) : r1 + (int)(rShift - 0.5);
    *a2 = (aShift >= 0) ? a1 + (int)(aShift + 0.5) : a1 + (int)(aShift - 0.5);
}

/* inBounds / boundsCheck / maskValue, speckleTrack.c:1822-1828,1943-1965 */
static void zero_pad_fft(ocplx *pad, int pSA, int pSR, const ocplx *orig, int oSA, int oSR, int dA, int dR);

static int in_bounds(int r1, int a1, int wR, int wA, int sR, int sA)
{
    return (r1 > 0 && (r1 + wR) < sR && a1 > 0 && (a1 + wA) < sA);
}

static int bounds_check(orc_st *s, int r1, int a1, int r2, int a2, int wR, int wA)
{
    return in_bounds(r1, a1, wR, wA, s->nr1, s->na1) && in_bounds(r2, a2, wR, wA, s->nr2, s->na2);
}

static int mask_value(orc_st *s, int r1, int a1)
{
    const st_track_params *p = s->par;
    if (p->mask == NULL)
        return 1;
    int ia1 = (a1 + s->wA / 2 - p->maskA0) / p->maskNal;
    int ir1 = (r1 + s->wR / 2 - p->maskR0) / p->maskNrl;
    if (ia1 < 0 || ia1 >= p->maskNa || ir1 < 0 || ir1 >= p->maskNr)
        return 0;
    return (int)p->mask[(size_t)ia1 * p->maskNr + ir1];
}

/* ------------------------------------------------------------------ */
/* phaseCorrectBaseline, speckleTrack.c:868-907 */
static void phase_correct_baseline(orc_st *s, int r1, int a1)
{
    const st_track_params *p = s->par;
    if (!p->baseFlag)
        return;
    double x = (a1 - s->wA / 2 - (p->nSlpA1 / 2)) / (double)(p->nSlpA1);
    double deltaX = 1. / (double)(p->nSlpA1);
    double bn = p->Bn + x * p->dBn + x * x * p->dBnQ;
    double ptod = -(4.0 * ORC_PI) / p->lambda;
    for (int i = 0; i < s->wA; i++) {
        double bp = p->Bp + x * p->dBp + x * x * p->dBpQ;
        double bsq = bn * bn + bp * bp;
        for (int j = 0; j < s->wR; j++) {
            int j1 = (r1 - s->wR / 2 > 0 ? r1 - s->wR / 2 : 0) + j;
            double ph = -bn * p->sinThetaD[j1] - bp * p->cosThetaD[j1] + (0.5 * bsq * p->overRange[j1]);
            ph *= ptod;
            double pr = cos(ph), pi = sin(ph);
            double cr = s->patch1[i * s->wR + j].re, ci = s->patch1[i * s->wR + j].im;
            s->patch1[i * s->wR + j].re = (oreal)(cr * pr - ci * pi);
            s->patch1[i * s->wR + j].im = (oreal)(cr * pi + ci * pr);
        }
        x += deltaX;
    }
}

/* ------------------------------------------------------------------ */
/* phaseCorrectInt, speckleTrack.c:909-1000, with the patch size of
 * speckleTrack.c:1576-1580 (osF = 1, speckleTrack.c:176).  The spectrum
 * of the ps x ps interferogram patch keeps the real part of the bins whose
 * power is at least 0.99 of the maximum; the imaginary part of every bin
 * is cleared (the loop body at :960-962 sets .im unconditionally).  The
 * zero-padded inverse (unnormalised) gives y; patch 1 is multiplied by
 * conj(y) / |y|, or left alone where |Re y| and |Im y| are both below
 * 1e-6. */
int orc_int_patch_size(const st_track_params *p)
{
    int ps = p->wA / p->intNal;
    if (p->wR / p->intNrl > ps)
        ps = p->wR / p->intNrl;
    return ps > 8 ? ps : 8;
}

static void phase_correct_int(orc_st *s, int r1, int a1)
{
    const st_track_params *p = s->par;
    if (!p->intf)
        return;
    const int ps = orc_int_patch_size(p);
    const int ia1 = (a1 + s->wA / 2) / p->intNal - ps / 2;
    const int ir1 = (r1 + s->wR / 2) / p->intNrl - ps / 2;
    if (!in_bounds(ir1, ia1, ps, ps, p->intNr, p->intNa))
        return;
    const int Pa = ps * p->intNal, Pr = ps * p->intNrl;
    ocplx *X = (ocplx *)malloc(sizeof(ocplx) * (size_t)ps * ps);
    ocplx *Y = (ocplx *)malloc(sizeof(ocplx) * (size_t)Pa * Pr);
    for (int i = 0; i < ps; i++)
        for (int j = 0; j < ps; j++) {
            const float *v = p->intf + 2 * ((size_t)(ia1 + i) * p->intNr + (ir1 + j));
            X[i * ps + j].re = (oreal)v[0];
            X[i * ps + j].im = (oreal)v[1];
        }
    orc_fft2d(X, ps, ps, ORC_FORWARD);
    double maxf = 0.;
    for (int k = 0; k < ps * ps; k++) {
        const double pw = (double)(X[k].re * X[k].re + X[k].im * X[k].im);
        maxf = pw > maxf ? pw : maxf;
    }
    for (int k = 0; k < ps * ps; k++) {
        const double pw = (double)(X[k].re * X[k].re + X[k].im * X[k].im);
        if (pw < 0.99 * maxf)
            X[k].re = 0;
        X[k].im = 0;
    }
    zero_pad_fft(Y, Pa, Pr, X, ps, ps, 0, 0);
    orc_fft2d(Y, Pa, Pr, ORC_BACKWARD);
    int i1 = Pa / 2 - s->wA / 2;
    for (int i = 0; i < s->wA; i++, i1++) {
        int j1 = Pr / 2 - s->wR / 2;
        for (int j = 0; j < s->wR; j++, j1++) {
            double pr = Y[i1 * Pr + j1].re, pi = -Y[i1 * Pr + j1].im, scale;
            if (fabs(pr) < 1.0e-6 && fabs(pi) < 1.0e-6) {
                pr = 1.0;
                pi = 0.0;
                scale = 1.0;
            } else
                scale = 1.0 / sqrt(pr * pr + pi * pi);
            ocplx *c = &s->patch1[i * s->wR + j];
            const double cr = c->re, ci = c->im;
            c->re = (oreal)((cr * pr - ci * pi) * scale);
            c->im = (oreal)((cr * pi + ci * pr) * scale);
        }
    }
    free(X);
    free(Y);
}

/* ------------------------------------------------------------------ */
/* getPatches, speckleTrack.c:1194-1330 */
static int get_patches(orc_st *s, int r1, int a1, int r2, int a2)
{
    int wA = s->wA, wR = s->wR;
    int isf = (s->stype == ST_SAMPLE_C64);
    float threshold = isf ? 1e-15f : 1e-9f;
    float thresholdAll = threshold * wA * wR;
    double p1 = 0., p2 = 0.;
    int nZero = 0;
    for (int i = 0; i < wA; i++)
        for (int j = 0; j < wR; j++) {
            ocplx a, b;
            sample(s->img1, s->stype, s->nr1, a1 + i, r1 + j, &a.re, &a.im);
            sample(s->img2, s->stype, s->nr2, a2 + i, r2 + j, &b.re, &b.im);
            if (isf) {
                a.re *= 1000;
                a.im *= 1000;
                b.re *= 1000;
                b.im *= 1000;
            }
            if (s->par->hanningFlag) {
                b.re *= s->hanning[i * wR + j];
                b.im *= s->hanning[i * wR + j];
            }
            s->patch1[i * wR + j] = a;
            s->patch2[i * wR + j] = b;
            oreal power1 = abs2(a), power2 = abs2(b);
            if (power1 < threshold || power2 < threshold)
                nZero++;
            if (nZero > 20)
                return 0;
            p1 += power1;
            p2 += power2;
        }
    if (p1 < thresholdAll || p2 < thresholdAll)
        return 0;
    est_dop_carrier1(s, s->patch1, s->patch2, wA, wR);
    s->p1 = p1 / (double)(wA * wR);
    s->p2 = p2 / (double)(wA * wR);
    return 1;
}

/* ------------------------------------------------------------------ */
/* findCPeak, speckleTrack.c:744-805: fftshift |C| into mag with the
 * interleaved four-quadrant scan, strict '>' keeps the first maximum. */
static void find_cpeak(orc_st *s, oreal *mag, const ocplx *C, int W_R, int W_A, int *iMax, int *jMax, double *cMax)
{
    double best = 0.0;
    int bi = -1000, bj = -1000;
    for (int i = 0; i < W_A / 2; i++) {
        int i1 = W_A / 2 + i;
        int j1 = W_R / 2;
        for (int j = 0; j < W_R / 2; j++) {
            oreal v;
            v = mag[i * W_R + j] = absc(C[i1 * W_R + j1]);
            if (v > best) { best = v; bi = i; bj = j; }
            v = mag[i1 * W_R + j1] = absc(C[i * W_R + j]);
            if (v > best) { best = v; bi = i1; bj = j1; }
            v = mag[i1 * W_R + j] = absc(C[i * W_R + j1]);
            if (v > best) { best = v; bi = i1; bj = j; }
            v = mag[i * W_R + j1] = absc(C[i1 * W_R + j]);
            if (v > best) { best = v; bi = i; bj = j1; }
            j1++;
        }
    }
    /* runner-up for the near-tie diagnostic; SNR = peak / mean |C| */
    double second = 0.0, sum = 0.0;
    for (int k = 0; k < W_R * W_A; k++) {
        sum += (double)mag[k];
        if (k != bi * W_R + bj && mag[k] > second)
            second = mag[k];
    }
    s->gapCoarse = best > 0 ? (float)((best - second) / best) : 0.f;
    s->snr = sum > 0 ? best * (double)(W_R * W_A) / sum : 0.0;
    *cMax = best;
    *iMax = bi;
    *jMax = bj;
    if (*iMax < (NFAST / 2) || *iMax > (W_A - NFAST / 2) || *jMax < (NFAST / 2) || *jMax > (W_R - NFAST / 2)) {
        *cMax = -1.0;
        *jMax = -1000;
        *iMax = -1000;
    }
}

/* zeroPadFFT, speckleTrack.c:2014-2038 */
static void zero_pad_fft(ocplx *pad, int pSA, int pSR, const ocplx *orig, int oSA, int oSR, int dA, int dR)
{
    memset(pad, 0, sizeof(ocplx) * (size_t)pSA * pSR);
    for (int i = 0; i < oSA / 2; i++) {
        int i1 = pSA - oSA / 2 + i;
        int i2 = (oSA / 2 + i + dA) % oSA;
        int ia = (i + dA) % oSA;
        int j1 = pSR - oSR / 2;
        int j2 = oSR / 2 + dR;
        for (int j = 0; j < oSR / 2; j++) {
            int ja = (j + dR) % oSR;
            int j2a = j2 % oSR;
            pad[i * pSR + j] = orig[ia * oSR + ja];
            pad[i1 * pSR + j] = orig[i2 * oSR + ja];
            pad[i1 * pSR + j1] = orig[i2 * oSR + j2a];
            pad[i * pSR + j1] = orig[ia * oSR + j2a];
            j1++;
            j2++;
        }
    }
}

/* overSampleC, speckleTrack.c:852-863: 8x8 -> 96x96 spectral zero pad */
static void over_sample_c(orc_st *s)
{
    ocplx ps[NFAST * NFAST];
    memcpy(ps, s->cFast, sizeof(ps));
    orc_fft2d(ps, NFAST, NFAST, ORC_FORWARD);
    zero_pad_fft(s->cFastOver, NFAST * NOVER, NFAST * NOVER, ps, NFAST, NFAST, 0, 0);
    orc_fft2d(s->cFastOver, NFAST * NOVER, NFAST * NOVER, ORC_BACKWARD);
}

/* Load the 8x8 real neighbourhood, oversample it and return the argmax of
 * the magnitude over the central [24,72)^2 box (speckleTrack.c:673-700). */
static double oversample_peak(orc_st *s, const oreal *mag, int W_R, int iMax, int jMax, int *iMax1, int *jMax1)
{
    for (int i = 0; i < NFAST; i++) {
        int i1 = iMax - NFAST / 2 + i;
        int j1 = jMax - NFAST / 2;
        for (int j = 0; j < NFAST; j++) {
            s->cFast[i * NFAST + j].re = mag[i1 * W_R + j1];
            s->cFast[i * NFAST + j].im = 0.0;
            j1++;
        }
    }
    over_sample_c(s);
    int bi = 0, bj = 0;
    double cMax = 0.0;
    const int lo = (NFAST * NOVER) / 2 - NOVER * 2, hi = (NFAST * NOVER) / 2 + NOVER * 2;
    for (int i = lo; i < hi; i++)
        for (int j = lo; j < hi; j++) {
            double corr = absc(s->cFastOver[i * NFAST * NOVER + j]);
            if (corr > cMax) {
                cMax = corr;
                bi = i;
                bj = j;
            }
        }
    double second = 0.0;
    for (int i = lo; i < hi; i++)
        for (int j = lo; j < hi; j++) {
            double corr = absc(s->cFastOver[i * NFAST * NOVER + j]);
            if ((i != bi || j != bj) && corr > second)
                second = corr;
        }
    s->gapFine = cMax > 0 ? (float)((cMax - second) / cMax) : 0.f;
    *iMax1 = bi;
    *jMax1 = bj;
    return cMax;
}

/* ------------------------------------------------------------------ */
/* 1-D Gaussian Levenberg-Marquardt fit (gaussPeak1D, mrqminMod,
 * mrqcofMod, myGauss1D: speckleTrack.c:2137-2202, 2309-2416, 2445-2465).
 * Parameters a = {amplitude, centre, width}; all three are fitted. */
static void gauss1d(oreal x, const oreal a[3], oreal *y, oreal dyda[3])
{
    oreal sig = a[2];
    oreal t = x - a[1];
    oreal argx = (oreal)(0.5 * t * t / (sig * sig));
    oreal ex = (oreal)exp(-argx);
    *y = a[0] * ex;
    dyda[0] = ex;
    dyda[1] = *y * t / (sig * sig);
    dyda[2] = *y * t * t / (sig * sig * sig);
}

static void lm_coef(const oreal *x, const oreal *y, const oreal *sg, int n, const oreal a[3], oreal alpha[3][3],
                    oreal beta[3], oreal *chisq)
{
    for (int j = 0; j < 3; j++) {
        for (int k = 0; k <= j; k++)
            alpha[j][k] = 0.0;
        beta[j] = 0.0;
    }
    *chisq = 0.0;
    for (int i = 0; i < n; i++) {
        oreal ymod, dyda[3];
        gauss1d(x[i], a, &ymod, dyda);
        oreal sig2i = (oreal)(1.0 / (sg[i] * sg[i]));
        oreal dy = y[i] - ymod;
        for (int l = 0; l < 3; l++) {
            oreal wt = dyda[l] * sig2i;
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

/* Gauss-Jordan elimination with full pivoting on a 3x3 system; returns 0
 * if the matrix is singular (the reference's gaussj aborts there). */
static int solve3(oreal m[3][3], oreal b[3])
{
    int used[3] = {0, 0, 0}, prow[3], pcol[3];
    for (int it = 0; it < 3; it++) {
        oreal big = 0;
        int r = -1, c = -1;
        for (int j = 0; j < 3; j++) {
            if (used[j])
                continue;
            for (int k = 0; k < 3; k++)
                if (!used[k] && fabs((double)m[j][k]) >= big) {
                    big = (oreal)fabs((double)m[j][k]);
                    r = j;
                    c = k;
                }
        }
        if (r < 0)
            return 0;
        used[c] = 1;
        if (r != c) {
            for (int l = 0; l < 3; l++) {
                oreal t = m[r][l];
                m[r][l] = m[c][l];
                m[c][l] = t;
            }
            oreal t = b[r];
            b[r] = b[c];
            b[c] = t;
        }
        prow[it] = r;
        pcol[it] = c;
        if (m[c][c] == 0)
            return 0;
        oreal pinv = (oreal)(1.0 / m[c][c]);
        m[c][c] = 1.0;
        for (int l = 0; l < 3; l++)
            m[c][l] *= pinv;
        b[c] *= pinv;
        for (int ll = 0; ll < 3; ll++)
            if (ll != c) {
                oreal d = m[ll][c];
                m[ll][c] = 0.0;
                for (int l = 0; l < 3; l++)
                    m[ll][l] -= m[c][l] * d;
                b[ll] -= b[c] * d;
            }
    }
    (void)prow;
    (void)pcol;
    return 1;
}

typedef struct {
    oreal alpha[3][3], beta[3], ochisq, lambda;
} lm_state;

/* One Levenberg-Marquardt step (the mrqmin contract: the first call, with
 * lambda < 0, initialises; later calls try a step and keep it only if chi^2
 * falls). */
static int lm_step(lm_state *st, const oreal *x, const oreal *y, const oreal *sg, int n, oreal a[3])
{
    oreal chisq;
    if (st->lambda < 0.0) {
        st->lambda = (oreal)0.001;
        lm_coef(x, y, sg, n, a, st->alpha, st->beta, &chisq);
        st->ochisq = chisq;
    }
    oreal covar[3][3], da[3];
    for (int j = 0; j < 3; j++) {
        for (int k = 0; k < 3; k++)
            covar[j][k] = st->alpha[j][k];
        covar[j][j] = (oreal)(st->alpha[j][j] * (1.0 + st->lambda));
        da[j] = st->beta[j];
    }
    if (!solve3(covar, da))
        return 0;
    oreal atry[3];
    for (int l = 0; l < 3; l++)
        atry[l] = a[l] + da[l];
    oreal nalpha[3][3], nbeta[3];
    lm_coef(x, y, sg, n, atry, nalpha, nbeta, &chisq);
    if (chisq < st->ochisq) {
        st->lambda = (oreal)(st->lambda * 0.1);
        st->ochisq = chisq;
        memcpy(st->alpha, nalpha, sizeof(nalpha));
        memcpy(st->beta, nbeta, sizeof(nbeta));
        for (int l = 0; l < 3; l++)
            a[l] = atry[l];
    } else {
        st->lambda = (oreal)(st->lambda * 10.0);
    }
    return 1;
}

/* ------------------------------------------------------------------ */
/* 2-D Gaussian Levenberg-Marquardt fit of the -gauss peak mode
 * (cmpTrackGauss, mrqminMod, mrqcofMod, myGauss: speckleTrack.c:2205-2303,
 * 2309-2416, 2421-2443).  Parameters a = {amplitude, mu_i, sigma_i, mu_j,
 * sigma_j}, all five fitted; data weights 1.  Gauss-Jordan with full
 * pivoting as NR's gaussj; a singular system ends the iterations (the
 * reference's gaussj aborts the program there). */
#define NG2 (NGAUSS * NGAUSS)

static void gauss2d(oreal xx, oreal yy, const oreal a[5], oreal *y, oreal dyda[5])
{
    oreal dx = xx - a[1], dy = yy - a[3];
    oreal argx = (oreal)0.5 * (dx * dx) / (a[2] * a[2]);
    oreal argy = (oreal)0.5 * (dy * dy) / (a[4] * a[4]);
    oreal ex = (oreal)exp((double)(-argx - argy));
    *y = a[0] * ex;
    dyda[0] = ex;
    dyda[1] = (*y) * (xx - a[1]) / (a[2] * a[2]);
    dyda[2] = *y * (xx - a[1]) * (xx - a[1]) / (a[2] * a[2] * a[2]);
    dyda[3] = *y * (yy - a[3]) / (a[4] * a[4]);
    dyda[4] = *y * (yy - a[3]) * (yy - a[3]) / (a[4] * a[4] * a[4]);
}

static void lm5_coef(const oreal *x, const oreal *yv, const oreal *g, const oreal a[5], oreal alpha[5][5],
                     oreal beta[5], oreal *chisq)
{
    for (int j = 0; j < 5; j++) {
        for (int k = 0; k <= j; k++)
            alpha[j][k] = 0.0;
        beta[j] = 0.0;
    }
    *chisq = 0.0;
    const oreal sig2i = (oreal)(1.0 / (1.0f * 1.0f));
    for (int i = 0; i < NG2; i++) {
        oreal ymod, dyda[5];
        gauss2d(x[i], yv[i], a, &ymod, dyda);
        oreal dy = g[i] - ymod;
        for (int l = 0; l < 5; l++) {
            oreal wt = dyda[l] * sig2i;
            for (int m = 0; m <= l; m++)
                alpha[l][m] += wt * dyda[m];
            beta[l] += dy * wt;
        }
        *chisq += dy * dy * sig2i;
    }
    for (int j = 1; j < 5; j++)
        for (int k = 0; k < j; k++)
            alpha[k][j] = alpha[j][k];
}

static int solve5(oreal m[5][5], oreal b[5])
{
    int used[5] = {0, 0, 0, 0, 0};
    for (int it = 0; it < 5; it++) {
        oreal big = 0;
        int r = -1, c = -1;
        for (int j = 0; j < 5; j++) {
            if (used[j])
                continue;
            for (int k = 0; k < 5; k++)
                if (!used[k] && fabs((double)m[j][k]) >= big) {
                    big = (oreal)fabs((double)m[j][k]);
                    r = j;
                    c = k;
                }
        }
        if (r < 0)
            return 0;
        used[c] = 1;
        if (r != c) {
            for (int l = 0; l < 5; l++) {
                oreal t = m[r][l];
                m[r][l] = m[c][l];
                m[c][l] = t;
            }
            oreal t = b[r];
            b[r] = b[c];
            b[c] = t;
        }
        if (m[c][c] == 0)
            return 0;
        oreal pinv = (oreal)(1.0 / m[c][c]);
        m[c][c] = 1.0;
        for (int l = 0; l < 5; l++)
            m[c][l] *= pinv;
        b[c] *= pinv;
        for (int ll = 0; ll < 5; ll++)
            if (ll != c) {
                oreal d = m[ll][c];
                m[ll][c] = 0.0;
                for (int l = 0; l < 5; l++)
                    m[ll][l] -= m[c][l] * d;
                b[ll] -= b[c] * d;
            }
    }
    return 1;
}

/* cmpTrackGauss: fit the 18 x 18 samples of the fftshifted magnitude mag
 * (row stride W_R) around the coarse peak (iMax, jMax); cMax is findCPeak's
 * raw maximum, wR / wA the window the surface was made from.  Returns the
 * peak in 1/NOVER units and the fitted amplitude (capped at 1), or
 * -LARGEINT / 0 when the peak is too close to the edge or a width exceeds
 * 2000. */
static void gauss_fit_2d(const oreal *mag, int W_R, int iMax, int jMax, double cMax, int wR, int wA, double p1,
                         double p2, double hanningCorrection, int *oI, int *oJ, double *oC)
{
    const int ng = NGAUSS;
    if ((iMax - ng / 2) < 0 || (iMax + ng / 2) >= (wA * OS - 1) || (jMax - ng / 2) < 0 ||
        (jMax + ng / 2) >= (wR * OS - 1)) {
        *oI = *oJ = (int)-ORC_LARGEINT;
        *oC = 0;
        return;
    }
    const double scale = hanningCorrection / (sqrt(p1 * p2) * pow(wR * wA, 2));
    oreal x[NG2], yv[NG2], g[NG2];
    int k = 0;
    for (int i = 0; i < ng; i++) {
        int i1 = iMax - ng / 2 + i, j1 = jMax - ng / 2;
        for (int j = 0; j < ng; j++, j1++, k++) {
            x[k] = (oreal)(float)i1;
            yv[k] = (oreal)(float)j1;
            g[k] = (oreal)(mag[i1 * W_R + j1] * scale);
        }
    }
    oreal a[5] = {(oreal)(cMax * scale), (oreal)iMax, 3, (oreal)jMax, 3};
    oreal alpha[5][5], beta[5], ochisq, chisq, lambda = (oreal)0.001;
    lm5_coef(x, yv, g, a, alpha, beta, &chisq);
    ochisq = chisq;
    oreal lasti = (oreal)2e9, lastj = (oreal)2.0e9;
    for (int it = 0; it < 8; it++) {
        oreal covar[5][5], da[5];
        for (int j = 0; j < 5; j++) {
            for (int kk = 0; kk < 5; kk++)
                covar[j][kk] = alpha[j][kk];
            covar[j][j] = (oreal)(alpha[j][j] * (1.0 + lambda));
            da[j] = beta[j];
        }
        if (!solve5(covar, da))
            break;
        oreal atry[5];
        for (int l = 0; l < 5; l++)
            atry[l] = a[l] + da[l];
        oreal nalpha[5][5], nbeta[5];
        lm5_coef(x, yv, g, atry, nalpha, nbeta, &chisq);
        if (chisq < ochisq) {
            lambda = (oreal)(lambda * 0.1);
            ochisq = chisq;
            memcpy(alpha, nalpha, sizeof(nalpha));
            memcpy(beta, nbeta, sizeof(nbeta));
            for (int l = 0; l < 5; l++)
                a[l] = atry[l];
        } else {
            lambda = (oreal)(lambda * 10.0);
        }
        if (fabs((double)(a[1] - lasti)) < 0.001 && fabs((double)(a[3] - lastj)) < 0.001)
            break;
        lasti = a[1];
        lastj = a[3];
    }
    *oI = (int)round((double)(a[1] * NOVER));
    *oJ = (int)round((double)(a[3] * NOVER));
    *oC = a[0] < 1.0 ? (double)a[0] : 1.0;
    const float sigThresh = 1000;
    if ((0.5 * a[2]) > sigThresh || (0.5 * a[4]) > sigThresh) {
        *oI = *oJ = (int)-ORC_LARGEINT;
        *oC = 0;
    }
}

/* gaussPeak1D: fit 18 samples through (iMax, jMax) along azimuth or range;
 * returns sigma = width / 2. */
static float gauss_peak_1d(const oreal *cMag, int W_R, int iMax, int jMax, int azOrRg)
{
    oreal x[NGAUSS], y[NGAUSS], sg[NGAUSS];
    oreal maxC = 0;
    for (int i = 0; i < NGAUSS; i++) {
        int i1;
        if (azOrRg == AZ) {
            i1 = iMax - NGAUSS / 2 + i;
            y[i] = cMag[i1 * W_R + jMax];
        } else {
            i1 = jMax - NGAUSS / 2 + i;
            y[i] = cMag[iMax * W_R + i1];
        }
        sg[i] = (oreal)0.1;
        x[i] = (oreal)i1;
        maxC = (y[i] > maxC) ? y[i] : maxC;
    }
    for (int i = 0; i < NGAUSS; i++)
        y[i] /= maxC;
    oreal a[3];
    a[0] = 1;
    a[1] = x[NGAUSS / 2];
    a[2] = 2;
    lm_state st;
    st.lambda = -1.;
    oreal lasti = (oreal)2e9;
    for (int it = 0; it < 5; it++) {
        if (!lm_step(&st, x, y, sg, NGAUSS, a))
            break;
        if (fabs((double)(a[1] - lasti)) < 0.001)
            break;
        lasti = a[1];
    }
    return (float)(a[2] * (oreal).5);
}

/* ------------------------------------------------------------------ */
/* complexMatch, cmpPSWithPad, cmpTrackFast, finalizeCmpxMatch:
 * speckleTrack.c:349-394, 604-739 */
static int complex_match(orc_st *s, int r1, int a1, int iOut, int jOut, int *pI, int *pJ, int *fI, int *fJ,
                         double *cMaxOut, float *rShift1, float *aShift1)
{
    const st_track_params *par = s->par;
    int wA = s->wA, wR = s->wR;
    int WA = wA * OSA, WR = wR * OSA;
    phase_correct_baseline(s, r1, a1);
    phase_correct_int(s, r1, a1);
    /* cmpPSWithPad */
    orc_fft2d(s->patch1, wA, wR, ORC_FORWARD);
    orc_fft2d(s->patch2, wA, wR, ORC_FORWARD);
    memset(s->ps, 0, sizeof(ocplx) * (size_t)WA * WR);
    const ocplx *F1 = s->patch1, *F2 = s->patch2;
#define CROSS(dst, src)                                                              \
    do {                                                                             \
        (dst).re = F1[src].re * F2[src].re + F1[src].im * F2[src].im;                \
        (dst).im = F1[src].im * F2[src].re - F1[src].re * F2[src].im;                \
    } while (0)
    for (int i = 0; i < wA / 2; i++) {
        int i1 = OSA * wA - wA / OSA + i;
        int i2 = (wA / OSA + i + s->azShift) % wA;
        int ia = (i + s->azShift) % wA;
        int j1 = wR * OSA - wR / OSA;
        int j2 = wR / OSA;
        for (int j = 0; j < wR / 2; j++) {
            int ja = j % wR;
            int j2a = j2 % wR;
            CROSS(s->ps[i * WR + j], ia * wR + ja);
            CROSS(s->ps[i1 * WR + j], i2 * wR + ja);
            CROSS(s->ps[i1 * WR + j1], i2 * wR + j2a);
            CROSS(s->ps[i * WR + j1], ia * wR + j2a);
            j1++;
            j2++;
        }
    }
#undef CROSS
    /* cmpTrackFast */
    double hanningCorrection = par->hanningFlag ? 1.5 : 1.0;
    memcpy(s->cbuf, s->ps, sizeof(ocplx) * (size_t)WA * WR);
    orc_fft2d(s->cbuf, WA, WR, ORC_BACKWARD);
    int iMax, jMax;
    double cMax;
    find_cpeak(s, s->mag, s->cbuf, WR, WA, &iMax, &jMax, &cMax);
    *pI = iMax;
    *pJ = jMax;
    s->gapFine = 0;
    if (par->gaussFlag) {
        /* complexMatch's -gauss branch (speckleTrack.c:367-372) */
        gauss_fit_2d(s->mag, WR, iMax, jMax, cMax, wR, wA, s->p1, s->p2, hanningCorrection, &iMax, &jMax, &cMax);
    } else if (iMax >= -999) {
        int iMax1, jMax1;
        cMax = oversample_peak(s, s->mag, WR, iMax, jMax, &iMax1, &jMax1);
        cMax = cMax / (NFAST * NFAST);
        cMax = hanningCorrection * cMax / (sqrt(s->p1 * s->p2) * pow(wR * wA, 2));
        iMax = (iMax * NOVER) + iMax1 - (NOVER * NFAST) / 2;
        jMax = (jMax * NOVER) + jMax1 - (NOVER * NFAST) / 2;
    }
    *fI = iMax;
    *fJ = jMax;
    *cMaxOut = cMax;
    /* complexMatch tail */
    int good = 0;
    int iM = iMax / NOVER, jM = jMax / NOVER;
    s->azd[(size_t)iOut * par->nR + jOut] = (float)-ORC_LARGEINT;
    s->sigRgLast = s->sigAzLast = 0;
    if ((iM - NGAUSS / 2) >= 0 && (iM + NGAUSS / 2) < WA && (jM - NGAUSS / 2) >= 0 && (jM + NGAUSS / 2) < WR) {
        float sigmaRg = gauss_peak_1d(s->mag, WR, iM, jM, RG);
        float sigmaAz = gauss_peak_1d(s->mag, WR, iM, jM, AZ);
        s->sigRgLast = sigmaRg;
        s->sigAzLast = sigmaAz;
        float azd = sigmaAz / sigmaRg;
        s->azd[(size_t)iOut * par->nR + jOut] = azd;
        if (fabsf(azd) < par->azDefocusThresh || par->checkAzFocus == 0) {
            /* finalizeCmpxMatch */
            double invNOVER = 1.0 / (float)NOVER;
            double wRover2exp = wR / 2. * NOVER * OSA, wAover2exp = wA / 2. * NOVER * OSA;
            if (cMax > .185 && par->noComplex == 0) {
                *rShift1 = (float)(((float)jMax - wRover2exp) * invNOVER / OS);
                *aShift1 = (float)(((float)iMax - wAover2exp) * invNOVER / OS);
                if (fabs((double)*rShift1) < 0.125 * wR * OS && fabs((double)*aShift1) < 0.125 * wA * OS) {
                    good = 1;
                    s->nComplex++;
                }
            }
        }
    }
    return good;
}

/* ------------------------------------------------------------------ */
/* getAmpPatches, speckleTrack.c:1006-1189 */
static int get_amp_patches(orc_st *s, int r1, int a1, int r2, int a2, int large)
{
    int isf = (s->stype == ST_SAMPLE_C64);
    float threshold = isf ? 1.0e-15f : 1.0e-6f;
    int wAa = s->wAa * (large ? LA : 1), wRa = s->wRa * (large ? LA : 1);
    size_t swRa = (size_t)wRa, swAa = (size_t)wAa;
    float scale = (float)(1. / (double)(swRa * swAa * swRa * swAa * OS * OS * OS));
    for (int i = 0; i < wAa; i++)
        for (int j = 0; j < wRa; j++) {
            sample(s->img1, s->stype, s->nr1, a1 + i, r1 + j, &s->im1in[i * wRa + j].re, &s->im1in[i * wRa + j].im);
            sample(s->img2, s->stype, s->nr2, a2 + i, r2 + j, &s->im2in[i * wRa + j].re, &s->im2in[i * wRa + j].im);
        }
    /* The reference transforms out of place and estimates the carrier on
     * the untransformed small patches in amplitude-only mode
     * (speckleTrack.c:1128-1144). */
    if (s->par->noComplex && !large)
        est_dop_carrier1(s, s->im1in, s->im2in, s->wAa, s->wRa);
    orc_fft2d(s->im1in, wAa, wRa, ORC_FORWARD);
    orc_fft2d(s->im2in, wAa, wRa, ORC_FORWARD);
    int azShift = (int)((float)s->azShift * (float)wAa / (float)s->wA);
    int W_A = wAa * OS, W_R = wRa * OS;
    zero_pad_fft(s->f1a, W_A, W_R, s->im1in, wAa, wRa, azShift, 0);
    zero_pad_fft(s->f2a, W_A, W_R, s->im2in, wAa, wRa, azShift, 0);
    memcpy(s->im1, s->f1a, sizeof(ocplx) * (size_t)W_A * W_R);
    memcpy(s->im2, s->f2a, sizeof(ocplx) * (size_t)W_A * W_R);
    orc_fft2d(s->im1, W_A, W_R, ORC_BACKWARD);
    orc_fft2d(s->im2, W_A, W_R, ORC_BACKWARD);
    double i1Sum = 0.0, i2Sum = 0.0;
    int nZero = 0;
    for (int k = 0; k < W_A * W_R; k++) {
        s->im1[k].re = abs2(s->im1[k]) * scale;
        s->im1[k].im = 0.0;
        s->im2[k].re = abs2(s->im2[k]) * scale;
        s->im2[k].im = 0.0;
        if (s->im1[k].re < threshold || s->im2[k].re < threshold)
            nZero++;
        i1Sum += s->im1[k].re;
        i2Sum += s->im2[k].re;
    }
    if (nZero > 20)
        return 0;
    if (i1Sum < 1.0e-6 || i2Sum < 1.0e-6)
        return 0;
    return 1;
}

/* ampMatch + ampTrackFast, speckleTrack.c:398-597 */
static void amp_match(orc_st *s, int large, int iOut, int jOut, int *pI, int *pJ, int *iMaxO, int *jMaxO,
                      double *cMaxO)
{
    const st_track_params *par = s->par;
    int wA = s->wAa * (large ? LA : 1), wR = s->wRa * (large ? LA : 1);
    int W_A = wA * OS, W_R = wR * OS, N = W_A * W_R;
    double avgAmp1 = 0.0, avgAmp2 = 0.0;
    for (int k = 0; k < N; k++) {
        s->im1[k].re = (oreal)sqrt((double)s->im1[k].re);
        s->im2[k].re = (oreal)sqrt((double)s->im2[k].re);
        avgAmp1 += s->im1[k].re;
        avgAmp2 += s->im2[k].re;
    }
    avgAmp1 = avgAmp1 / (double)(N);
    avgAmp2 = avgAmp2 / (double)(N);
    double p1 = 0.0, p2 = 0.0;
    for (int k = 0; k < N; k++) {
        if (s->im1[k].re > 1.0e-20)
            s->im1[k].re = (oreal)(s->im1[k].re - avgAmp1);
        s->im2[k].re = (oreal)(s->im2[k].re - avgAmp2);
        p1 += s->im1[k].re * s->im1[k].re;
        p2 += s->im2[k].re * s->im2[k].re;
    }
    p1 = p1 / (double)N;
    p2 = p2 / N;
    orc_fft2d(s->im1, W_A, W_R, ORC_FORWARD);
    orc_fft2d(s->im2, W_A, W_R, ORC_FORWARD);
    for (int k = 0; k < N; k++) {
        ocplx a = s->im1[k], b = s->im2[k];
        s->cbuf[k].re = a.re * b.re + a.im * b.im;
        s->cbuf[k].im = a.im * b.re - a.re * b.im;
    }
    /* ampTrackFast */
    orc_fft2d(s->cbuf, W_A, W_R, ORC_BACKWARD);
    int iMax, jMax;
    double cMax;
    find_cpeak(s, s->mag, s->cbuf, W_R, W_A, &iMax, &jMax, &cMax);
    *pI = iMax;
    *pJ = jMax;
    s->gapFine = 0;
    s->sigRgLast = s->sigAzLast = 0;
    float *azdp = &s->azd[(size_t)iOut * par->nR + jOut];
    if (iMax > NGAUSS / 2 && iMax < (W_A - NGAUSS / 2) && jMax > NGAUSS / 2 && jMax < (W_R - NGAUSS / 2)) {
        float sigmaRg = gauss_peak_1d(s->mag, W_R, iMax, jMax, RG);
        float sigmaAz = gauss_peak_1d(s->mag, W_R, iMax, jMax, AZ);
        s->sigRgLast = sigmaRg;
        s->sigAzLast = sigmaAz;
        *azdp = sigmaAz / sigmaRg;
    }
    if (iMax < -999 || (fabsf(*azdp) >= par->azDefocusThresh && par->checkAzFocus)) {
        *iMaxO = iMax;
        *jMaxO = jMax;
        *cMaxO = cMax;
        return;
    }
    if (par->gaussFlag) {
        /* ampTrackFast's -gauss branch (speckleTrack.c:550-557): powers
         * scaled by pScale = 16, no Hanning correction */
        const double pScale = 16.;
        gauss_fit_2d(s->mag, W_R, iMax, jMax, cMax, wR, wA, p1 * pScale, p2 * pScale, 1.0, iMaxO, jMaxO, cMaxO);
        return;
    }
    int iMax1, jMax1;
    cMax = oversample_peak(s, s->mag, W_R, iMax, jMax, &iMax1, &jMax1);
    cMax = cMax / (pow((double)(W_A), 2) * pow((double)(W_R), 2) * pow((double)NFAST, 2));
    cMax = cMax / sqrt(p1 * p2);
    *iMaxO = (iMax * NOVER) + iMax1 - (NOVER * NFAST) / 2;
    *jMaxO = (jMax * NOVER) + jMax1 - (NOVER * NFAST) / 2;
    *cMaxO = cMax;
}

/* ------------------------------------------------------------------ */
/* speckleTrack driver, speckleTrack.c:159-346 */
int orc_speckle_track(const st_track_params *par, const void *img1, int nr1, int na1, const void *img2, int nr2,
                      int na2, int stype, st_track_out *out, st_track_diag *dg)
{
    /* windows of any even size, range and azimuth independent
     * (parseTrack.c:167-171) */
    if ((par->wR & 1) || (par->wA & 1) || (par->wRa & 1) || (par->wAa & 1) || par->wR < 8 || par->wA < 8 ||
        par->wRa < 8 || par->wAa < 8)
        return ST_ERR_UNSUPPORTED;
    if (par->intf && !par->noComplex) {
        /* the corrected window must lie inside the oversampled patch (the
         * reference indexes outside it otherwise) */
        if (par->intNr < 1 || par->intNa < 1 || par->intNrl < 1 || par->intNal < 1)
            return ST_ERR_ARG;
        const int ps = orc_int_patch_size(par);
        if ((ps * par->intNal) / 2 < par->wA / 2 || (ps * par->intNrl) / 2 < par->wR / 2)
            return ST_ERR_ARG;
    }
    orc_st S;
    memset(&S, 0, sizeof(S));
    orc_st *s = &S;
    s->par = par;
    s->img1 = img1;
    s->img2 = img2;
    s->nr1 = nr1;
    s->na1 = na1;
    s->nr2 = nr2;
    s->na2 = na2;
    s->stype = stype;
    s->wR = par->wR;
    s->wA = par->wA;
    s->wRa = par->wRa;
    s->wAa = par->wAa;
    int wmax = s->wA > s->wAa * LA ? s->wA : s->wAa * LA;
    if (s->wR > wmax)
        wmax = s->wR;
    if (s->wRa * LA > wmax)
        wmax = s->wRa * LA;
    size_t big = (size_t)(2 * wmax) * (2 * wmax);
    s->hanning = (oreal *)calloc((size_t)s->wA * s->wR, sizeof(oreal));
    s->patch1 = (ocplx *)calloc((size_t)s->wA * s->wR, sizeof(ocplx));
    s->patch2 = (ocplx *)calloc((size_t)s->wA * s->wR, sizeof(ocplx));
    s->ps = (ocplx *)calloc(big, sizeof(ocplx));
    s->cbuf = (ocplx *)calloc(big, sizeof(ocplx));
    s->mag = (oreal *)calloc(big, sizeof(oreal));
    s->im1in = (ocplx *)calloc(big / 4, sizeof(ocplx));
    s->im2in = (ocplx *)calloc(big / 4, sizeof(ocplx));
    s->f1a = (ocplx *)calloc(big, sizeof(ocplx));
    s->f2a = (ocplx *)calloc(big, sizeof(ocplx));
    s->im1 = (ocplx *)calloc(big, sizeof(ocplx));
    s->im2 = (ocplx *)calloc(big, sizeof(ocplx));
    s->cFastOver = (ocplx *)calloc((size_t)NFAST * NOVER * NFAST * NOVER, sizeof(ocplx));
    size_t npt = (size_t)par->nA * par->nR;
    float *azdLocal = NULL;
    s->azd = out->azd;
    if (!s->azd)
        s->azd = azdLocal = (float *)calloc(npt, sizeof(float));
    else
        memset(s->azd, 0, npt * sizeof(float));
    if (par->hanningFlag && !par->noComplex)
        compute_hanning(s);
    s->azShift = 0;
    const double invNOVER = 1.0 / (float)NOVER;
    int r2 = 0, a2 = 0, iMax = 0, jMax = 0;
    float rShift1 = 0, aShift1 = 0;
    for (int i = 0; i < par->nA; i++) {
        int a1 = par->aStart + i * par->deltaA - par->wA / 2;
        for (int j = 0; j < par->nR; j++) {
            size_t o = (size_t)i * par->nR + j;
            int haveData = 1, type = BAD, good = 0;
            double cMax = 0.;
            int r1 = par->rStart + j * par->deltaR - par->wR / 2;
            find_image2_pos(par, r1, a1, &r2, &a2);
            float rShiftAmp = (float)(r2 - r1), aShiftAmp = (float)(a2 - a1);
            int maskVal = mask_value(s, r1 * par->scaleFactor, a1 * par->scaleFactor);
            maskVal = maskVal < 1 ? maskVal : 1;
            if (dg) {
                if (dg->r2) dg->r2[o] = r2;
                if (dg->a2) dg->a2[o] = a2;
                if (dg->azShift) dg->azShift[o] = -1;
                for (int t = 0; t < 3; t++) {
                    if (dg->peakI[t]) dg->peakI[t][o] = -2000;
                    if (dg->peakJ[t]) dg->peakJ[t][o] = -2000;
                    if (dg->fineI[t]) dg->fineI[t][o] = -2000;
                    if (dg->fineJ[t]) dg->fineJ[t][o] = -2000;
                    if (dg->cmax[t]) dg->cmax[t][o] = 0;
                    if (dg->sigRg[t]) dg->sigRg[t][o] = 0;
                    if (dg->sigAz[t]) dg->sigAz[t][o] = 0;
                    if (dg->peakGap[t]) dg->peakGap[t][o] = 0;
                    if (dg->fineGap[t]) dg->fineGap[t][o] = 0;
                    if (dg->snr[t]) dg->snr[t][o] = 0;
                }
            }
            if (bounds_check(s, r1, a1, r2, a2, par->wR, par->wA) && maskVal > 0 && par->noComplex == 0) {
                haveData = get_patches(s, r1, a1, r2, a2);
                if (haveData) {
                    int pI, pJ;
                    if (dg && dg->azShift) dg->azShift[o] = s->azShift;
                    good = complex_match(s, r1, a1, i, j, &pI, &pJ, &iMax, &jMax, &cMax, &rShift1, &aShift1);
                    if (good)
                        type = CMATCH;
                    if (dg) {
                        if (dg->peakI[0]) dg->peakI[0][o] = pI;
                        if (dg->peakJ[0]) dg->peakJ[0][o] = pJ;
                        if (dg->fineI[0]) dg->fineI[0][o] = iMax;
                        if (dg->fineJ[0]) dg->fineJ[0][o] = jMax;
                        if (dg->cmax[0]) dg->cmax[0][o] = cMax;
                        if (dg->sigRg[0]) dg->sigRg[0][o] = s->sigRgLast;
                        if (dg->sigAz[0]) dg->sigAz[0][o] = s->sigAzLast;
                        if (dg->peakGap[0]) dg->peakGap[0][o] = s->gapCoarse;
                        if (dg->snr[0]) dg->snr[0][o] = (float)s->snr;
                        if (dg->fineGap[0]) dg->fineGap[0][o] = s->gapFine;
                    }
                }
            }
            int nTries = 0;
            while (!good && haveData && maskVal > 0 && nTries < par->maxTries) {
                int large = (nTries != 0);
                int ampWScale = large ? LA : 1;
                double cThresh = large ? 0.028 : 0.07;
                int a1a = par->aStart + i * par->deltaA - par->wA / 2 * ampWScale;
                int r1a = par->rStart + j * par->deltaR - par->wR / 2 * ampWScale;
                int r2a = r1a + (int)rShiftAmp, a2a = a1a + (int)aShiftAmp;
                if (bounds_check(s, r1a, a1a, r2a, a2a, par->wRa * ampWScale, par->wAa * ampWScale) && maskVal > 0) {
                    haveData = get_amp_patches(s, r1a, a1a, r2a, a2a, large);
                    if (haveData) {
                        int pI, pJ;
                        amp_match(s, large, i, j, &pI, &pJ, &iMax, &jMax, &cMax);
                        rShift1 = (float)(((float)jMax - par->wRa * 0.5 * OS * NOVER * ampWScale) * invNOVER / OS);
                        aShift1 = (float)(((float)iMax - par->wAa * 0.5 * OS * NOVER * ampWScale) * invNOVER / OS);
                        if (dg) {
                            int t = large ? 2 : 1;
                            if (dg->peakI[t]) dg->peakI[t][o] = pI;
                            if (dg->peakJ[t]) dg->peakJ[t][o] = pJ;
                            if (dg->fineI[t]) dg->fineI[t][o] = iMax;
                            if (dg->fineJ[t]) dg->fineJ[t][o] = jMax;
                            if (dg->cmax[t]) dg->cmax[t][o] = cMax;
                            if (dg->sigRg[t]) dg->sigRg[t][o] = s->sigRgLast;
                            if (dg->sigAz[t]) dg->sigAz[t][o] = s->sigAzLast;
                            if (dg->peakGap[t]) dg->peakGap[t][o] = s->gapCoarse;
                            if (dg->snr[t]) dg->snr[t][o] = (float)s->snr;
                            if (dg->fineGap[t]) dg->fineGap[t][o] = s->gapFine;
                        }
                    } else
                        cMax = 0.0;
                } else
                    cMax = 0.0;
                if (fabs((double)rShift1) < 0.15 * par->wRa * OS * ampWScale &&
                    fabs((double)aShift1) < 0.15 * par->wAa * OS * ampWScale && cMax > cThresh) {
                    good = 1;
                    if (large)
                        s->nAmpL++;
                    else
                        s->nAmp++;
                    type = large ? AMPMATCHLARGE : AMPMATCH;
                }
                nTries++;
            }
            float rShift, aShift;
            if (good && maskVal > 0) {
                rShift = rShiftAmp - rShift1;
                aShift = aShiftAmp - aShift1;
            } else {
                rShift = (float)-ORC_LARGEINT;
                aShift = (float)-ORC_LARGEINT;
                if (haveData)
                    s->nFail++;
            }
            out->offR[o] = rShift;
            out->offA[o] = aShift;
            if (rShift > -100000) {
                out->offR[o] *= (float)par->scaleFactor;
                out->offA[o] *= (float)par->scaleFactor;
            }
            out->corr[o] = (float)cMax;
            out->type[o] = (uint8_t)type;
        }
    }
    out->nComplex = s->nComplex;
    out->nAmp = s->nAmp;
    out->nAmpL = s->nAmpL;
    out->nFail = s->nFail;
    free(s->hanning);
    free(s->patch1);
    free(s->patch2);
    free(s->ps);
    free(s->cbuf);
    free(s->mag);
    free(s->im1in);
    free(s->im2in);
    free(s->f1a);
    free(s->f2a);
    free(s->im1);
    free(s->im2);
    free(s->cFastOver);
    free(azdLocal);
    return ST_OK;
}

/* ------------------------------------------------------------------ */
/* Test hooks: internals exposed so the oracle can be pinned against
 * closed forms (tests/test_oracle.py). */
void orc_oversample8(const double *P, double *Qmag)
{
    orc_st s;
    memset(&s, 0, sizeof(s));
    s.cFastOver = (ocplx *)calloc((size_t)NFAST * NOVER * NFAST * NOVER, sizeof(ocplx));
    for (int k = 0; k < NFAST * NFAST; k++) {
        s.cFast[k].re = (oreal)P[k];
        s.cFast[k].im = 0;
    }
    over_sample_c(&s);
    for (int k = 0; k < NFAST * NOVER * NFAST * NOVER; k++)
        Qmag[k] = (double)absc(s.cFastOver[k]);
    free(s.cFastOver);
}

float orc_gauss_sigma(const float *y, float x0, int azOrRg)
{
    oreal mag[NGAUSS * NGAUSS];
    memset(mag, 0, sizeof(mag));
    /* place the samples on row / column 9 of an 18 x 18 surface */
    for (int t = 0; t < NGAUSS; t++) {
        if (azOrRg == AZ)
            mag[t * NGAUSS + NGAUSS / 2] = (oreal)y[t];
        else
            mag[(NGAUSS / 2) * NGAUSS + t] = (oreal)y[t];
    }
    return gauss_peak_1d(mag, NGAUSS, NGAUSS / 2, NGAUSS / 2, azOrRg) + 0.0f * x0;
}

void orc_hanning2(int wA, int wR, float *out)
{
    orc_st s;
    memset(&s, 0, sizeof(s));
    s.wA = wA;
    s.wR = wR;
    s.hanning = (oreal *)calloc((size_t)wA * wR, sizeof(oreal));
    compute_hanning(&s);
    for (int k = 0; k < wA * wR; k++)
        out[k] = (float)s.hanning[k];
    free(s.hanning);
}

void orc_hanning(int w, float *out) { orc_hanning2(w, w, out); }

/* cmpTrackGauss alone, on a float magnitude surface of row stride W_R. */
void orc_gauss_fit_2d(const float *magIn, int W_R, int W_A, int iMax, int jMax, double cMax, int wR, int wA,
                      double p1, double p2, double hc, int *oI, int *oJ, double *oC)
{
    oreal *mag = (oreal *)malloc(sizeof(oreal) * (size_t)W_R * W_A);
    for (int k = 0; k < W_R * W_A; k++)
        mag[k] = (oreal)magIn[k];
    gauss_fit_2d(mag, W_R, iMax, jMax, cMax, wR, wA, p1, p2, hc, oI, oJ, oC);
    free(mag);
}

/* phaseCorrectInt alone: patch (wA x wR, in place) as patch 1 of the
 * point whose window starts at (r1, a1). */
void orc_int_correct(const st_track_params *par, int r1, int a1, ocplx *patch)
{
    orc_st s;
    memset(&s, 0, sizeof(s));
    s.par = par;
    s.wA = par->wA;
    s.wR = par->wR;
    s.patch1 = patch;
    phase_correct_int(&s, r1, a1);
}

/* cmpPSWithPad + the 2wA x 2wR inverse + findCPeak's fftshifted magnitude,
 * for two wA x wR patches (rows of wR) as getPatches leaves them. */
void orc_corr_surface2(const ocplx *p1, const ocplx *p2, int wA, int wR, int azShift, oreal *magOut)
{
    orc_st s;
    memset(&s, 0, sizeof(s));
    st_track_params par;
    memset(&par, 0, sizeof(par));
    par.nR = par.nA = 1;
    par.hanningFlag = 0;
    s.par = &par;
    s.wA = wA;
    s.wR = wR;
    s.azShift = azShift;
    s.p1 = s.p2 = 1.0;
    const size_t n = (size_t)wA * wR;
    s.patch1 = (ocplx *)calloc(n, sizeof(ocplx));
    s.patch2 = (ocplx *)calloc(n, sizeof(ocplx));
    s.ps = (ocplx *)calloc(4 * n, sizeof(ocplx));
    s.cbuf = (ocplx *)calloc(4 * n, sizeof(ocplx));
    s.mag = magOut;
    s.cFastOver = (ocplx *)calloc((size_t)NFAST * NOVER * NFAST * NOVER, sizeof(ocplx));
    float azd;
    s.azd = &azd;
    memcpy(s.patch1, p1, sizeof(ocplx) * n);
    memcpy(s.patch2, p2, sizeof(ocplx) * n);
    int pI, pJ, fI, fJ;
    double cMax;
    float r1, a1;
    complex_match(&s, 0, 0, 0, 0, &pI, &pJ, &fI, &fJ, &cMax, &r1, &a1);
    free(s.patch1);
    free(s.patch2);
    free(s.ps);
    free(s.cbuf);
    free(s.cFastOver);
}

void orc_corr_surface(const ocplx *p1, const ocplx *p2, int w, int azShift, oreal *magOut)
{
    orc_corr_surface2(p1, p2, w, w, azShift, magOut);
}

/* findCPeak alone on a W_A x W_R complex surface in the inverse
 * transform's (unshifted) layout; mag receives the fftshifted |C|. */
void orc_find_cpeak(const ocplx *C, int W_R, int W_A, oreal *mag, int *iMax, int *jMax, double *cMax, double *snr)
{
    orc_st s;
    memset(&s, 0, sizeof(s));
    find_cpeak(&s, mag, C, W_R, W_A, iMax, jMax, cMax);
    *snr = s.snr;
}

/* getShifts alone: the initial offset at (range, azimuth) from the grids
 * of par (initDr / initDa). */
void orc_get_shifts(const st_track_params *par, double range, double azimuth, double *dr, double *da)
{
    get_shifts(par, range, azimuth, dr, da);
}
