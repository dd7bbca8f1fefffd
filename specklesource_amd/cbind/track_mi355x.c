/*
 * track_mi355x.c -- speckleTrack (Strack/strack.h:170, speckleTrack.c:159-346)
 * and corrTrackFast (Strackw/strackw.h:1, corrTrackFast.c:89-231) on the
 * MI355X library: the translation unit that replaces speckleTrack.c and
 * corrTrackFast.c in the strack / strackw link lines.  The SLC files are
 * read whole (no strip buffers), swapped and tracked on the GPU; the
 * result grids are allocated here, as the reference's callee allocates
 * them (speckleTrack.c:1777-1781).  Failures print a message and exit(1),
 * as the reference's error() does.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "grimp_track.h"
#include "speckle_amd.h"

static st_ctx *gCtx;

static void fail(const char *what, const char *msg)
{
    fprintf(stderr, "%s: %s\n", what, msg ? msg : "");
    exit(1);
}

static st_ctx *ctx(void)
{
    if (!gCtx) {
        const char *e = getenv("SPECKLE_AMD_DEVICE");
        if (st_ctx_create(e ? atoi(e) : 0, &gCtx) != ST_OK)
            fail("speckleTrack", "no usable MI355X");
    }
    return gCtx;
}

float **grimp_grid_f(int32_t nRows, int32_t nCols)
{
    float **r = (float **)malloc(sizeof(float *) * (size_t)nRows);
    float *b = (float *)calloc((size_t)nRows * nCols, sizeof(float));
    if (!r || !b)
        fail("grimp_grid_f", "out of memory");
    for (int32_t i = 0; i < nRows; i++)
        r[i] = b + (size_t)i * nCols;
    return r;
}

char **grimp_grid_c(int32_t nRows, int32_t nCols)
{
    char **r = (char **)malloc(sizeof(char *) * (size_t)nRows);
    char *b = (char *)calloc((size_t)nRows * nCols, 1);
    if (!r || !b)
        fail("grimp_grid_c", "out of memory");
    for (int32_t i = 0; i < nRows; i++)
        r[i] = b + (size_t)i * nCols;
    return r;
}

void grimp_grid_free(void *rows)
{
    if (rows) {
        free(((void **)rows)[0]);
        free(rows);
    }
}

static void *read_image(const char *file, const SARData *p, size_t sampleBytes)
{
    const size_t bytes = (size_t)p->nSlpR * p->nSlpA * sampleBytes;
    void *buf = malloc(bytes);
    FILE *f = fopen(file, "rb");
    if (!buf || !f)
        fail("speckleTrack: could not open", file);
    if (fread(buf, 1, bytes, f) != bytes)
        fail("speckleTrack: short SLC file", file);
    fclose(f);
    return buf;
}

/* TrackParams -> st_track_params (the fields of the GPU path) */
static void to_params(const TrackParams *t, st_track_params *p)
{
    memset(p, 0, sizeof(*p));
    p->rStart = t->rStart;
    p->aStart = t->aStart;
    p->deltaR = t->deltaR;
    p->deltaA = t->deltaA;
    p->nR = t->nR;
    p->nA = t->nA;
    p->wR = t->wR;
    p->wA = t->wA;
    p->wRa = t->wRa;
    p->wAa = t->wAa;
    p->edgePadR = t->edgePadR;
    p->edgePadA = t->edgePadA;
    p->navgR = t->navgR;
    p->navgA = t->navgA;
    p->scaleFactor = t->scaleFactor > 0 ? t->scaleFactor : 1;
    p->noComplex = t->noComplex;
    p->hanningFlag = t->hanningFlag;
    p->gaussFlag = t->gaussFlag;
    p->maxTries = t->maxTries;
    p->checkAzFocus = t->checkAzFocus;
    p->azDefocusThresh = t->azDefocusThresh;
    p->precision = ST_PREC_FP32;
    p->polyShift = t->polyShift;
    memcpy(p->rShiftPoly, t->rShiftPoly, sizeof(p->rShiftPoly));
    memcpy(p->aShiftPoly, t->aShiftPoly, sizeof(p->aShiftPoly));
    if (!t->polyShift) { /* getShifts (speckleTrack.c:1857-1940) */
        p->initDr = t->initOffsets.dr[0];
        p->initDa = t->initOffsets.da[0];
        p->initNr = t->initOffsets.nr;
        p->initNa = t->initOffsets.na;
        p->initR0 = t->initOffsets.rO;
        p->initA0 = t->initOffsets.aO;
        p->initDeltaR = t->initOffsets.deltaR;
        p->initDeltaA = t->initOffsets.deltaA;
    }
    if (t->maskFlag) { /* maskValue (speckleTrack.c:1951-1965) */
        if (t->maskType != OFFSETMASK)
            fail("speckleTrack", "only offset masks (maskType OFFSETMASK) reach the GPU path");
        p->mask = (const uint8_t *)t->maskDat.mask[0];
        p->maskNr = t->maskDat.nr;
        p->maskNa = t->maskDat.na;
        p->maskR0 = t->maskDat.r0;
        p->maskA0 = t->maskDat.a0;
        p->maskNrl = t->maskDat.nrl;
        p->maskNal = t->maskDat.nal;
    }
    if (t->intFlag && !t->noComplex) { /* phaseCorrectInt (speckleTrack.c:909-1000) */
        p->intf = t->intDat.intf[0];
        p->intNr = t->intDat.nr;
        p->intNa = t->intDat.na;
        p->intNrl = t->intDat.nrl;
        p->intNal = t->intDat.nal;
    }
    /* The baseline ramp needs the look-angle tables of initPhaseCorrect,
     * built from SAR geometry in mosaicSource (not in the reference tree). */
    if (t->baseFlag && !t->noComplex)
        fail("speckleTrack", "baseline correction: pass the range tables through st_speckle_track");
}

static void run(TrackParams *t, int strackw)
{
    st_ctx *c = ctx();
    const size_t sb = t->floatFlag ? 8 : 4;
    void *img1 = read_image(t->imageFile1, &t->imageP1, sb);
    void *img2 = read_image(t->imageFile2, &t->imageP2, sb);
    int32_t pair = -1;
    if (st_pair_upload_raw(c, img1, t->imageP1.nSlpR, t->imageP1.nSlpA, img2, t->imageP2.nSlpR, t->imageP2.nSlpA,
                           t->floatFlag ? ST_SAMPLE_C64 : ST_SAMPLE_CI16, t->byteOrder == LSB ? ST_LSB : ST_MSB,
                           &pair) != ST_OK)
        fail("speckleTrack: image upload", st_last_error(c));
    free(img1);
    free(img2);
    st_track_params p;
    to_params(t, &p);
    t->offR = grimp_grid_f(t->nA, t->nR);
    t->offA = grimp_grid_f(t->nA, t->nR);
    t->corr = grimp_grid_f(t->nA, t->nR);
    t->type = grimp_grid_c(t->nA, t->nR);
    t->aZDefocus = strackw ? NULL : grimp_grid_f(t->nA, t->nR);
    st_track_out o;
    memset(&o, 0, sizeof(o));
    o.offR = t->offR[0];
    o.offA = t->offA[0];
    o.corr = t->corr[0];
    o.type = (uint8_t *)t->type[0];
    o.azd = strackw ? NULL : t->aZDefocus[0];
    const int rc = strackw ? st_corr_track_fast(c, pair, &p, &o, NULL) : st_speckle_track(c, pair, &p, &o, NULL);
    if (rc != ST_OK)
        fail(strackw ? "corrTrackFast" : "speckleTrack", st_last_error(c));
    st_pair_release(c, pair);
    t->nComplex = o.nComplex;
    t->nAmp = o.nAmp;
    t->nAmpL = o.nAmpL;
    t->nFail = o.nFail;
}

void speckleTrack(TrackParams *trackPar) { run(trackPar, 0); }

void corrTrackFast(TrackParams *trackPar) { run(trackPar, 1); }
