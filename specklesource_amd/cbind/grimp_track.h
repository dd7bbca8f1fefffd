/*
 * grimp_track.h -- the fields of the reference's TrackParams
 * (Strack/strack.h:56-165), CullParams (Cullst/cullst.h:3-58) and the
 * mosaicSource types they embed that the MI355X bindings read or write,
 * under the reference's names.  The real structs also hold FFTW plans,
 * GDAL handles and file names the GPU path does not use; a maintainer
 * builds the bindings against Strack/strack.h and Cullst/cullst.h
 * instead, and these declarations let them compile and be tested here.
 */
#ifndef GRIMP_TRACK_H
#define GRIMP_TRACK_H

#include <stdint.h>

#ifndef LSB
#define LSB 0 /* strack.h:7 */
#define MSB 1
#endif
#ifndef TRUE
#define TRUE 1
#define FALSE 0
#endif
#define OFFSETMASK 1 /* strack.h:6 */

/* mosaicSource/common SARData: the two fields speckleTrack.c reads
 * (speckleTrack.c:1377, 1825-1826) */
typedef struct {
    int32_t nSlpR, nSlpA;
} SARData;

/* mosaicSource/common Offsets (getShifts, speckleTrack.c:1870-1882) */
typedef struct {
    int32_t nr, na;
    double rO, aO, deltaR, deltaA;
    float **dr, **da;
} Offsets;

typedef struct { /* strack.h:24-34 without the FFTW plans; intf as (re, im) float pairs */
    int32_t nr, na, nrl, nal, patchSize;
    float **intf;
} intData;

typedef struct { /* strack.h:36-45 */
    int32_t r0, a0, nr, na, nrl, nal;
    char **mask;
} maskData;

typedef struct {
    char *imageFile1, *imageFile2; /* SLC image files (read whole: no strip buffers) */
    int32_t maskFlag, maskType;
    SARData imageP1, imageP2;
    int32_t noComplex;
    intData intDat;
    maskData maskDat;
    int32_t rStart, aStart, deltaR, deltaA, nR, nA;
    int32_t wR, wA, wRa, wAa;
    int32_t edgePad, edgePadA, edgePadR, scaleFactor, navgR, navgA;
    int32_t intFlag, baseFlag, floatFlag, polyShift;
    double rShiftPoly[3], aShiftPoly[3];
    float **offR, **offA, **corr, **aZDefocus; /* results: allocated by the callee */
    char **type;
    int32_t nComplex, nAmp, nAmpL, nFail;
    Offsets initOffsets;
    int32_t hanningFlag, gaussFlag;
    float azDefocusThresh;
    int32_t maxTries, byteOrder, checkAzFocus;
} TrackParams;

typedef struct { /* Cullst/cullst.h:3-58, the grids and knobs of the stages */
    int32_t rStart, aStart, deltaR, deltaA, nR, nA;
    int32_t maskFlag, useSim, singleMT;
    float **offR, **offA, **offRS, **offAS, **corr, **sigmaA, **sigmaR;
    char **type, **mask;
    int32_t bR, bA, nGood;
    float maxA, maxR;
    int32_t sR, sA, islandThresh, ignoreOffsets;
    float corrThresh;
} CullParams;

/* The reference's entry points (strack.h:170, strackw.h:1, cullst.h:60-67),
 * here on the MI355X library. */
void speckleTrack(TrackParams *trackPar);
void corrTrackFast(TrackParams *trackPar);
void cullSTData(CullParams *cullPar);
void cullStats(CullParams *cullPar);
void cullSmooth(CullParams *cullPar);
void cullIslands(CullParams *cullPar);

/* Row-pointer arrays over one contiguous nRows x nCols block, as the
 * reference allocates its grids (speckleTrack.c:1967-1976). */
float **grimp_grid_f(int32_t nRows, int32_t nCols);
char **grimp_grid_c(int32_t nRows, int32_t nCols);
void grimp_grid_free(void *rows);

#endif /* GRIMP_TRACK_H */
