/*
 * ORACLE -- test infrastructure (see orc_common.h).  Not the product.
 *
 * CPU restatement of cullst's culling stages, Cullst/cullSTData.c,
 * cullStats.c, cullSmooth.c and cullIslands.c of fastice/speckleSource.
 *
 * - cullSTData: box median test, in place, raster order.  The median is
 *   the floor(n/2)-th smallest of the 1-based list (selectForCull,
 *   cullSTData.c:132-192); here it is taken by sorting, which selects the
 *   same order statistic.
 * - cullStats: plane fit {mean, row, col} by truncated-SVD least squares
 *   (NR svdfit contract: singular values below 1e-5 * max are zeroed;
 *   cRecipes is not vendored, the tolerance is NR's published one).  The
 *   SVD is one-sided Jacobi in double.
 * - cullSmooth: weighted box mean with effective looks.
 * - cullIslands: labelRegions (unwrapSource) is not vendored; components
 *   are 4-connected and labelled 1.. in raster order of their first pixel
 *   ("parity unpinned"); the last label is skipped as in the reference
 *   (cullIslands.c:152,179).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc_common.h"
#include "../include/speckle_amd.h"

#define NARROWREGION 4
#define NEG_LARGE (-2000000000.0f)

static inline int valid2(float r, float a)
{
    /* offR > (1 - LARGEINT) with LARGEINT = 2e9 (cullSTData.c:63) */
    return r > (1 - ORC_LARGEINT) && a > (1 - ORC_LARGEINT);
}

static int cmpf(const void *x, const void *y)
{
    float a = *(const float *)x, b = *(const float *)y;
    return (a > b) - (a < b);
}

int orc_cull_st_data(st_cull_params *cp)
{
    int nA = cp->nA, nR = cp->nR;
    float *offR = cp->offR, *offA = cp->offA;
    int maskFlag = cp->mask != NULL;
    for (int i = 0; i < nA; i++)
        for (int j = 0; j < nR; j++) {
            size_t o = (size_t)i * nR + j;
            if (cp->corr[o] < cp->corrThresh) {
                offR[o] = NEG_LARGE;
                offA[o] = NEG_LARGE;
                cp->type[o] = 0;
            }
            if (maskFlag) {
                unsigned char mv = (unsigned char)cp->mask[o] & NARROWREGION;
                if (mv > 0 && cp->type[o] > 2) {
                    offR[o] = NEG_LARGE;
                    offA[o] = NEG_LARGE;
                    cp->type[o] = 0;
                }
                if (cp->singleMT > 0 && cp->type[o] != cp->singleMT) {
                    offR[o] = NEG_LARGE;
                    offA[o] = NEG_LARGE;
                    cp->type[o] = 0;
                }
            }
        }
    size_t nlist = 1000 + (size_t)(cp->bR + 1) * (cp->bA + 1);
    float *listR = (float *)malloc(sizeof(float) * nlist), *listA = (float *)malloc(sizeof(float) * nlist);
    int nCulled = 0;
    for (int i = 0; i < nA; i++) {
        int i1 = i - cp->bA / 2 > 0 ? i - cp->bA / 2 : 0;
        int i2 = i + cp->bA / 2 < nA - 1 ? i + cp->bA / 2 : nA - 1;
        for (int j = 0; j < nR; j++) {
            int j1 = j - cp->bR / 2 > 0 ? j - cp->bR / 2 : 0;
            int j2 = j + cp->bR / 2 < nR - 1 ? j + cp->bR / 2 : nR - 1;
            unsigned ngood = 0;
            for (int ii = i1; ii <= i2; ii++)
                for (int jj = j1; jj <= j2; jj++) {
                    size_t q = (size_t)ii * nR + jj;
                    if (valid2(offR[q], offA[q])) {
                        listR[ngood] = offR[q];
                        listA[ngood] = offA[q];
                        ngood++;
                    }
                }
            size_t o = (size_t)i * nR + j;
            if (ngood > (unsigned)cp->nGood) {
                unsigned k = ngood / 2; /* 1-based index k -> 0-based k-1 */
                if (k == 0)
                    k = 1; /* ngood == 1: the reference reads the unused slot 0 */
                qsort(listR, ngood, sizeof(float), cmpf);
                qsort(listA, ngood, sizeof(float), cmpf);
                float medianR = listR[k - 1], medianA = listA[k - 1];
                double diffA = (double)(offA[o] - medianA);
                double diffR = (double)(offR[o] - medianR);
                if (fabs(diffA) > cp->maxA || fabs(diffR) > cp->maxR) {
                    offR[o] = NEG_LARGE;
                    offA[o] = NEG_LARGE;
                    nCulled++;
                }
            } else {
                offR[o] = NEG_LARGE;
                offA[o] = NEG_LARGE;
                nCulled++;
            }
        }
    }
    free(listR);
    free(listA);
    cp->nCulled = nCulled;
    return ST_OK;
}

/* Least squares y ~ A x, A = n x 3 with rows (c0, xi[k], xj[k]), by
 * one-sided Jacobi SVD with the svdfit truncation rule (singular values
 * below 1e-5 * max dropped).  The rotated columns U = A V are never
 * stored: every dot product recomputes u = a_k . v from the rows, in a
 * fixed order, so the GPU restatement can follow the same arithmetic.
 * Sweeps stop once every column pair is orthogonal to 1e-13 (cosine);
 * the rounding floor of these 81-row dot products is ~1e-14, so a
 * tighter limit only spins. */
static void lsq3_svd(double c0, const int *xi, const int *xj, const double *y, int n, double x[3])
{
    double V[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int sweep = 0; sweep < 30; sweep++) {
        double off = 0.0;
        for (int p = 0; p < 2; p++)
            for (int q = p + 1; q < 3; q++) {
                double alpha = 0, beta = 0, gamma = 0;
                for (int i = 0; i < n; i++) {
                    const double a0 = c0, a1 = (double)xi[i], a2 = (double)xj[i];
                    const double up = a0 * V[0 * 3 + p] + a1 * V[1 * 3 + p] + a2 * V[2 * 3 + p];
                    const double uq = a0 * V[0 * 3 + q] + a1 * V[1 * 3 + q] + a2 * V[2 * 3 + q];
                    alpha += up * up;
                    beta += uq * uq;
                    gamma += up * uq;
                }
                if (gamma == 0.0)
                    continue;
                double conv = fabs(gamma) / sqrt(alpha * beta);
                if (conv > off)
                    off = conv;
                if (conv < 1e-13)
                    continue;
                double zeta = (beta - alpha) / (2.0 * gamma);
                double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
                for (int r = 0; r < 3; r++) {
                    double vp = V[r * 3 + p], vq = V[r * 3 + q];
                    V[r * 3 + p] = c * vp - s * vq;
                    V[r * 3 + q] = s * vp + c * vq;
                }
            }
        if (off < 1e-13)
            break;
    }
    double w2[3] = {0, 0, 0}, uy[3] = {0, 0, 0};
    for (int i = 0; i < n; i++) {
        const double a0 = c0, a1 = (double)xi[i], a2 = (double)xj[i];
        for (int k = 0; k < 3; k++) {
            const double u = a0 * V[0 * 3 + k] + a1 * V[1 * 3 + k] + a2 * V[2 * 3 + k];
            w2[k] += u * u;
            uy[k] += u * y[i];
        }
    }
    double w[3], wmax = 0;
    for (int k = 0; k < 3; k++) {
        w[k] = sqrt(w2[k]);
        if (w[k] > wmax)
            wmax = w[k];
    }
    const double thresh = 1.0e-5 * wmax;
    double coef[3];
    for (int k = 0; k < 3; k++)
        coef[k] = (w[k] < thresh || w[k] == 0.0) ? 0.0 : uy[k] / w2[k];
    for (int r = 0; r < 3; r++)
        x[r] = V[r * 3 + 0] * coef[0] + V[r * 3 + 1] * coef[1] + V[r * 3 + 2] * coef[2];
}

/* Test hook: the plane fit of cullStats on explicit points. */
void orc_lsq3(double c0, const int *xi, const int *xj, const double *y, int n, double *x)
{
    lsq3_svd(c0, xi, xj, y, n, x);
}

/* cullStats / cullLSStats body: sigmaR/A of offR/A; addQuant adds the
 * quantisation variance `quant` before the square root (cullLSStats.c:144-146). */
static void cull_stats_core(int nA, int nR, const float *offR, const float *offA, float *sigmaR, float *sigmaA,
                            int addQuant, double quant);

int orc_cull_stats(st_cull_params *cp)
{
    cp->bR = 9; /* cullStats.c:38-39 forces 9 x 9 */
    cp->bA = 9;
    cull_stats_core(cp->nA, cp->nR, cp->offR, cp->offA, cp->sigmaR, cp->sigmaA, 0, 0.0);
    return ST_OK;
}

static void cull_stats_core(int nA, int nR, const float *offR, const float *offA, float *sigmaR, float *sigmaA,
                            int addQuant, double quant)
{
    const int b = 9;
    double listR[81], listA[81];
    int xi[81], xj[81];
    double aR[3] = {0, 0, 0}, aA[3] = {0, 0, 0};
    for (int i = 0; i < nA; i++) {
        int i1 = i - b / 2 > 0 ? i - b / 2 : 0;
        int i2 = i + b / 2 < nA - 1 ? i + b / 2 : nA - 1;
        for (int j = 0; j < nR; j++) {
            int j1 = j - b / 2 > 0 ? j - b / 2 : 0;
            int j2 = j + b / 2 < nR - 1 ? j + b / 2 : nR - 1;
            int ngood = 0;
            double meanA = 0.0, meanR = 0.0;
            for (int ii = i1; ii <= i2; ii++)
                for (int jj = j1; jj <= j2; jj++) {
                    size_t q = (size_t)ii * nR + jj;
                    if (valid2(offR[q], offA[q])) {
                        listR[ngood] = offR[q];
                        listA[ngood] = offA[q];
                        meanA += offA[q];
                        meanR += offR[q];
                        xi[ngood] = ii;
                        xj[ngood] = jj;
                        ngood++;
                    }
                }
            meanA /= (double)ngood;
            meanR /= (double)ngood;
            if (ngood > 8) {
                /* the basis value is stored in a float field (cullStats.c:80) */
                lsq3_svd((double)(float)meanR, xi, xj, listR, ngood, aR);
                lsq3_svd((double)(float)meanA, xi, xj, listA, ngood, aA);
                aR[0] *= meanR;
                aA[0] *= meanA;
            }
            int ng = 0;
            double mA = 0.0, mR = 0.0, varR = 0.0, varA = 0.0;
            for (int ii = i1; ii <= i2; ii++)
                for (int jj = j1; jj <= j2; jj++) {
                    size_t q = (size_t)ii * nR + jj;
                    if (valid2(offR[q], offA[q])) {
                        ng++;
                        float oR = offR[q], oA = offA[q];
                        oR = (float)(oR - (aR[0] + aR[1] * (float)ii + aR[2] * (float)jj));
                        oA = (float)(oA - (aA[0] + aA[1] * (float)ii + aA[2] * (float)jj));
                        mA += oA;
                        mR += oR;
                        varA += oA * oA;
                        varR += oR * oR;
                    }
                }
            size_t o = (size_t)i * nR + j;
            if (ng > 8 && offR[o] > (1 - ORC_LARGEINT)) {
                float sA = (float)(varA / (float)(ng - 3));
                float sR = (float)(varR / (float)(ng - 3));
                sA = (float)(sA - mA * mA / (float)((ng - 3) * ng));
                sR = (float)(sR - mR * mR / (float)((ng - 3) * ng));
                sA = (float)(0.00001 > sA ? 0.00001 : sA);
                sR = (float)(0.00001 > sR ? 0.00001 : sR);
                if (addQuant) {
                    sR = (float)(sR + quant);
                    sA = (float)(sA + quant);
                }
                sigmaA[o] = (float)sqrt((double)sA);
                sigmaR[o] = (float)sqrt((double)sR);
            } else {
                sigmaA[o] = NEG_LARGE;
                sigmaR[o] = NEG_LARGE;
            }
        }
    }
}

/* sWeights, cullSmooth.c:10-27 (allocated at its full size here) */
static void s_weights(int s, float *w /* 2*(s/2)+1 entries, centred */)
{
    float w2 = (float)((double)s / 2.0 - 0.01);
    for (int m = -s / 2; m <= s / 2; m++)
        w[m + s / 2] = (m < -w2 || m > w2) ? 0.5f : 1.0f;
}

int orc_cull_smooth(st_cull_params *cp)
{
    int nA = cp->nA, nR = cp->nR;
    float nlooks[4] = {0.0f, 1.0f, 0.25f, 0.015f};
    float wA[513], wR[513];
    if (cp->sA > 1000 || cp->sR > 1000)
        return ST_ERR_ARG;
    s_weights(cp->sA, wA);
    s_weights(cp->sR, wR);
    for (int i = 0; i < nA; i++) {
        int i1 = i - cp->sA / 2 > 0 ? i - cp->sA / 2 : 0;
        int i2 = i + cp->sA / 2 < nA - 1 ? i + cp->sA / 2 : nA - 1;
        int iwStart = i1 - i;
        for (int j = 0; j < nR; j++) {
            int j1 = j - cp->sR / 2 > 0 ? j - cp->sR / 2 : 0;
            int j2 = j + cp->sR / 2 < nR - 1 ? j + cp->sR / 2 : nR - 1;
            int jwStart = j1 - j;
            double ngood = 0, meanA = 0.0, meanR = 0.0;
            float nLooksEff = 0.0f;
            for (int ii = i1, iw = iwStart; ii <= i2; ii++, iw++)
                for (int jj = j1, jw = jwStart; jj <= j2; jj++, jw++) {
                    size_t q = (size_t)ii * nR + jj;
                    if (valid2(cp->offR[q], cp->offA[q])) {
                        double w = wA[iw + cp->sA / 2] * wR[jw + cp->sR / 2];
                        ngood += w;
                        nLooksEff = (float)(nLooksEff + w * nlooks[(int)cp->type[q] & 3]);
                        meanA += w * cp->offA[q];
                        meanR += w * cp->offR[q];
                    }
                }
            size_t o = (size_t)i * nR + j;
            if (((ngood > 1) || (cp->sA == 1) || (cp->sR == 1)) && cp->offR[o] > (1 - ORC_LARGEINT) &&
                cp->sigmaA[o] > (1 - ORC_LARGEINT)) {
                nLooksEff = (float)(nLooksEff > 1. ? nLooksEff : 1.);
                cp->sigmaA[o] = (float)(cp->sigmaA[o] / sqrt((double)nLooksEff));
                cp->sigmaR[o] = (float)(cp->sigmaR[o] / sqrt((double)nLooksEff));
                cp->offRS[o] = (float)(meanR / (float)ngood);
                cp->offAS[o] = (float)(meanA / (float)ngood);
            } else {
                cp->offRS[o] = NEG_LARGE;
                cp->offAS[o] = NEG_LARGE;
            }
        }
    }
    return ST_OK;
}

int orc_cull_islands(st_cull_params *cp)
{
    int nA = cp->nA, nR = cp->nR;
    size_t n = (size_t)nA * nR;
    int32_t *lab = (int32_t *)calloc(n, sizeof(int32_t));
    int32_t *stack = (int32_t *)malloc(n * sizeof(int32_t));
    int nlab = 0;
    for (size_t p0 = 0; p0 < n; p0++) {
        if (lab[p0] || !(cp->offRS[p0] >= (float)(-ORC_LARGEINT + 1000)))
            continue;
        nlab++;
        size_t sp = 0;
        stack[sp++] = (int32_t)p0;
        lab[p0] = nlab;
        while (sp) {
            int32_t p = stack[--sp];
            int i = p / nR, j = p % nR;
            const int di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
            for (int k = 0; k < 4; k++) {
                int ii = i + di[k], jj = j + dj[k];
                if (ii < 0 || ii >= nA || jj < 0 || jj >= nR)
                    continue;
                size_t q = (size_t)ii * nR + jj;
                if (!lab[q] && cp->offRS[q] >= (float)(-ORC_LARGEINT + 1000)) {
                    lab[q] = nlab;
                    stack[sp++] = (int32_t)q;
                }
            }
        }
    }
    int *minX = (int *)malloc(sizeof(int) * (nlab + 1)), *maxX = (int *)malloc(sizeof(int) * (nlab + 1));
    int *minY = (int *)malloc(sizeof(int) * (nlab + 1)), *maxY = (int *)malloc(sizeof(int) * (nlab + 1));
    for (int l = 0; l <= nlab; l++) {
        minX[l] = minY[l] = 2000000000;
        maxX[l] = maxY[l] = -2000000000;
    }
    for (size_t p = 0; p < n; p++)
        if (lab[p]) {
            int l = lab[p], i = (int)(p / nR), j = (int)(p % nR);
            if (j < minX[l]) minX[l] = j;
            if (j > maxX[l]) maxX[l] = j;
            if (i < minY[l]) minY[l] = i;
            if (i > maxY[l]) maxY[l] = i;
        }
    for (size_t p = 0; p < n; p++) {
        int l = lab[p];
        if (l >= 1 && l < nlab && (maxX[l] - minX[l]) <= cp->islandThresh && (maxY[l] - minY[l]) <= cp->islandThresh) {
            cp->offRS[p] = NEG_LARGE;
            if (cp->offAS)
                cp->offAS[p] = NEG_LARGE;
        }
    }
    free(lab);
    free(stack);
    free(minX);
    free(maxX);
    free(minY);
    free(maxY);
    return ST_OK;
}

int orc_cull_run(st_cull_params *cp)
{
    orc_cull_st_data(cp);
    orc_cull_st_data(cp);
    orc_cull_st_data(cp);
    orc_cull_stats(cp);
    orc_cull_smooth(cp);
    if (cp->islandThresh > 0)
        orc_cull_islands(cp);
    return ST_OK;
}

/* ------------------------------------------------------------------ */
/* Cullls (Landsat offsets): Cullls/cullLSData.c, cullLSStats.c,
 * cullLSSmooth.c, cullLSIslands.c.  X / Y take the places of offR / offA.
 * The SLOWVEL / NODATATOUSE bits, NODATA and the Landsat NOVER come from
 * un-vendored headers and are parameters. */

/* cullLSData.c:13-134 */
int orc_cull_ls_data(st_cull_ls_params *cp)
{
    int ny = cp->ny, nx = cp->nx;
    float *X = cp->X, *Y = cp->Y;
    int nGoodPts = 0, nWithData = 0;
    for (size_t o = 0; o < (size_t)nx * ny; o++) {
        if (valid2(X[o], Y[o]))
            nGoodPts++;
        if ((cp->type[o] & cp->noDataToUseMask) == 0)
            nWithData++;
    }
    if (cp->nMatch < 0) {
        cp->nAttempt = nWithData;
        cp->nMatch = nGoodPts;
    }
    size_t nlist = 1000 + (size_t)(cp->bR + 1) * (cp->bA + 1);
    float *listX = (float *)malloc(sizeof(float) * nlist), *listY = (float *)malloc(sizeof(float) * nlist);
    int nCulled = 0;
    for (int i = 0; i < ny; i++) {
        int i1 = i - cp->bA / 2 > 0 ? i - cp->bA / 2 : 0;
        int i2 = i + cp->bA / 2 < ny - 1 ? i + cp->bA / 2 : ny - 1;
        for (int j = 0; j < nx; j++) {
            int j1 = j - cp->bR / 2 > 0 ? j - cp->bR / 2 : 0;
            int j2 = j + cp->bR / 2 < nx - 1 ? j + cp->bR / 2 : nx - 1;
            int slowFlag = 1;
            unsigned ngood = 0;
            for (int ii = i1; ii <= i2; ii++)
                for (int jj = j1; jj <= j2; jj++) {
                    size_t q = (size_t)ii * nx + jj;
                    if (valid2(X[q], Y[q])) {
                        listX[ngood] = X[q];
                        listY[ngood] = Y[q];
                        ngood++;
                        if (!(cp->type[q] & cp->slowVelMask))
                            slowFlag = 0;
                    }
                }
            size_t o = (size_t)i * nx + j;
            if (ngood > (unsigned)cp->nGood && (double)Y[o] > (double)cp->noData + 1.0) {
                unsigned k = ngood / 2;
                if (k == 0)
                    k = 1;
                qsort(listX, ngood, sizeof(float), cmpf);
                qsort(listY, ngood, sizeof(float), cmpf);
                float medianX = listX[k - 1], medianY = listY[k - 1];
                double diffY = (double)(Y[o] - medianY);
                double diffX = (double)(X[o] - medianX);
                double maxX, maxY;
                if (slowFlag) {
                    /* 40 m/yr over deltaT days, in pixels, at least 0.2 */
                    maxX = 40.0 / 365. * cp->deltaT;
                    maxY = maxX;
                    maxX /= cp->dx;
                    maxY /= cp->dy;
                    maxX = maxX > 0.2 ? maxX : 0.2;
                    maxY = maxY > 0.2 ? maxY : 0.2;
                } else {
                    maxX = cp->maxX;
                    maxY = cp->maxY;
                }
                if (fabs(diffY) > maxY || fabs(diffX) > maxX) {
                    X[o] = NEG_LARGE;
                    Y[o] = NEG_LARGE;
                    nCulled++;
                }
            } else {
                X[o] = NEG_LARGE;
                Y[o] = NEG_LARGE;
                nCulled++;
            }
        }
    }
    free(listX);
    free(listY);
    cp->nCulled = nCulled;
    return ST_OK;
}

/* cullLSStats.c:21-160: cullStats plus the quantisation variance
 * (1/NOVER)^2 / 12 */
int orc_cull_ls_stats(st_cull_ls_params *cp)
{
    if (cp->nover <= 0)
        return ST_ERR_ARG;
    cp->bR = 9;
    cp->bA = 9;
    const double q = ((1.0 / cp->nover) * (1.0 / cp->nover)) / 12.;
    cull_stats_core(cp->ny, cp->nx, cp->X, cp->Y, cp->sigmaX, cp->sigmaY, 1, q);
    return ST_OK;
}

/* cullLSSmooth.c:11-73: unweighted box mean, effective looks = count */
int orc_cull_ls_smooth(st_cull_ls_params *cp)
{
    int ny = cp->ny, nx = cp->nx;
    for (int i = 0; i < ny; i++) {
        int i1 = i - cp->sY / 2 > 0 ? i - cp->sY / 2 : 0;
        int i2 = i + cp->sY / 2 < ny - 1 ? i + cp->sY / 2 : ny - 1;
        for (int j = 0; j < nx; j++) {
            int j1 = j - cp->sX / 2 > 0 ? j - cp->sX / 2 : 0;
            int j2 = j + cp->sX / 2 < nx - 1 ? j + cp->sX / 2 : nx - 1;
            unsigned ngood = 0;
            float nLooksEff = 0.0f;
            double meanY = 0.0, meanX = 0.0;
            for (int ii = i1; ii <= i2; ii++)
                for (int jj = j1; jj <= j2; jj++) {
                    size_t q = (size_t)ii * nx + jj;
                    if (valid2(cp->X[q], cp->Y[q])) {
                        ngood++;
                        nLooksEff++;
                        meanY += cp->Y[q];
                        meanX += cp->X[q];
                    }
                }
            size_t o = (size_t)i * nx + j;
            if (((ngood > 1) || (cp->sY == 1) || (cp->sX == 1)) && cp->X[o] > (1 - ORC_LARGEINT) &&
                cp->sigmaY[o] > (1 - ORC_LARGEINT)) {
                nLooksEff = (float)(nLooksEff > 1. ? nLooksEff : 1.);
                cp->sigmaY[o] = (float)(cp->sigmaY[o] / sqrt((double)nLooksEff));
                cp->sigmaX[o] = (float)(cp->sigmaX[o] / sqrt((double)nLooksEff));
                cp->XS[o] = (float)(meanX / (float)ngood);
                cp->YS[o] = (float)(meanY / (float)ngood);
            } else {
                cp->XS[o] = NEG_LARGE;
                cp->YS[o] = NEG_LARGE;
            }
        }
    }
    return ST_OK;
}

/* cullLSIslands.c:17-177: cullIslands on XS / YS */
int orc_cull_ls_islands(st_cull_ls_params *cp)
{
    st_cull_params c;
    memset(&c, 0, sizeof(c));
    c.nR = cp->nx;
    c.nA = cp->ny;
    c.offRS = cp->XS;
    c.offAS = cp->YS;
    c.islandThresh = cp->islandThresh;
    return orc_cull_islands(&c);
}

/* cullls.c:23-58 */
int orc_cull_ls_run(st_cull_ls_params *cp)
{
    cp->nMatch = -1;
    orc_cull_ls_data(cp);
    orc_cull_ls_data(cp);
    orc_cull_ls_data(cp);
    int rc = orc_cull_ls_stats(cp);
    if (rc != ST_OK)
        return rc;
    orc_cull_ls_smooth(cp);
    if (cp->islandThresh > 0)
        orc_cull_ls_islands(cp);
    return ST_OK;
}
