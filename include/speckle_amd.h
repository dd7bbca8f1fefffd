/*
 * speckle_amd.h -- C ABI of the MI355X speckle-tracking library.
 *
 * This is the drop-in boundary for the GrIMP speckle-tracking hot path
 * (fastice/speckleSource).  The reference exposes three in-process entry
 * points that all take one fat struct:
 *
 *   void speckleTrack(TrackParams *)      Strack/strack.h:170,  called at Strack/strack.c:134
 *   void corrTrackFast(TrackParams *)     Strackw/strackw.h:1,  called at Strackw/strackw.c:126
 *   void cullSTData/cullStats/cullSmooth/cullIslands(CullParams *)
 *                                         Cullst/cullst.h:60-67, called at Cullst/cullst.c:75-100
 *
 * TrackParams embeds FFTW plan and GDAL handle types (Strack/strack.h:84-85,
 * 135-142), so it cannot be cloned as-is.  st_track_params carries the same
 * fields as plain C types; images live on the device as "pairs" that the
 * caller uploads once and tracks many times.  Output grids keep the
 * reference's layout (row-major nA x nR, float32 / byte) and sentinels
 * (-2e9 offsets, type 0, Strack/speckleTrack.c:306-323).
 *
 * Every function returns ST_OK (0) or a negative ST_ERR_* code; none exits
 * the process (the reference's error() does).  st_last_error() returns a
 * message for the last failing call on a context.
 *
 * Threading: one st_ctx per device; calls on one context are serialised by
 * the library.
 */
#ifndef SPECKLE_AMD_H
#define SPECKLE_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ST_OK 0
#define ST_ERR_ARG (-1)         /* invalid argument or size */
#define ST_ERR_HIP (-2)         /* HIP runtime failure */
#define ST_ERR_UNSUPPORTED (-3) /* option not implemented on this path */
#define ST_ERR_NOMEM (-4)       /* device allocation failed / over budget */
#define ST_ERR_NODEV (-5)       /* no usable GPU */
#define ST_ERR_CANCELLED (-7)   /* the job was cancelled (st_cancel); -6 is speckle_io.h's ST_ERR_IO */

/* SLC sample types (Strack/strack.h:10-16, speckleTrack.c:1221-1232) */
#define ST_SAMPLE_C64 0  /* complex float32, fftw_complex {re, im} */
#define ST_SAMPLE_CI16 1 /* complex int16, strackComplex {r, i} */

/* Match types written to .mt (Strack/speckleTrack.c:19-22) */
#define ST_BAD 0
#define ST_CMATCH 1
#define ST_AMPMATCH 2
#define ST_AMPMATCHLARGE 3

/* "no value" sentinel, -LARGEINT (Strack/writeVrt.c:69) */
#define ST_NODATA (-2.0e9f)

/* Precision of the patch correlations.  ST_PREC_FP32 is the plain FP32 FFT
 * path.  ST_PREC_MFMA forms the small window DFTs as BF16 hi+lo split
 * contractions on the matrix cores (FP32 accumulate). */
#define ST_PREC_FP32 0
#define ST_PREC_MFMA 1
#define ST_PREC_AUTO 2 /* the precision st_autotune found faster for the window (FP32 without a tune entry) */

typedef struct st_ctx st_ctx;

/* Tracking parameters: the fields of TrackParams (Strack/strack.h:56-165)
 * that the strack / strackw hot path reads. */
typedef struct {
    /* estimate grid, already divided by scaleFactor (parseTrack.c:207-210) */
    int32_t rStart, aStart;   /* strack.h:91-92 */
    int32_t deltaR, deltaA;   /* strack.h:94-95 */
    int32_t nR, nA;           /* strack.h:97-98 */
    /* windows */
    int32_t wR, wA;           /* complex window, strack.h:100-101 */
    int32_t wRa, wAa;         /* amplitude window, strack.h:103-104 */
    /* strackw only */
    int32_t edgePadR, edgePadA; /* strack.h:107-108 (caller value; +5 is added internally, corrTrackFast.c:116-117) */
    int32_t navgR, navgA;       /* strack.h:110-111 */
    int32_t scaleFactor;        /* strack.h:109 */
    /* flags, defaults in Strack/strack.c:143-154 */
    int32_t noComplex;          /* -noComplex */
    int32_t hanningFlag;        /* default 1; -noHanning clears it */
    int32_t gaussFlag;          /* -gauss: 2-D Gaussian peak fit (cmpTrackGauss) on the staged path */
    int32_t maxTries;           /* 2, -singleAmp gives 1 */
    int32_t checkAzFocus;       /* -checkAzFocus */
    float azDefocusThresh;      /* 1.8 (strack.c:45) */
    int32_t precision;          /* ST_PREC_* */
    /* initial offsets (strack.h:115,131-133,154; speckleTrack.c:810-847) */
    int32_t polyShift;          /* 1: use the polynomials only */
    double rShiftPoly[3];       /* rShift = p0 + p1*r + p2*a */
    double aShiftPoly[3];
    const float *initDr;        /* optional initial-offset grids, initNa x initNr */
    const float *initDa;
    int32_t initNr, initNa;
    double initR0, initA0, initDeltaR, initDeltaA;
    /* optional NN mask (strack.h:36-45, speckleTrack.c:1951-1965) */
    const uint8_t *mask;        /* maskNa x maskNr, NULL = no mask */
    int32_t maskNr, maskNa, maskR0, maskA0, maskNrl, maskNal;
    /* baseline phase ramp on patch 1 (speckleTrack.c:868-907) */
    int32_t baseFlag;           /* 0: the ramp is the identity and is skipped */
    double Bn, Bp, dBn, dBp, dBnQ, dBpQ, lambda;
    int32_t nSlpA1;             /* image-1 azimuth size used by the ramp */
    const double *sinThetaD;    /* per-range tables, length nr of image 1 */
    const double *cosThetaD;
    const double *overRange;
    /* interferogram phase correction on patch 1 (phaseCorrectInt,
     * speckleTrack.c:909-1000; getInt.c:11-77): intNa x intNr complex
     * samples as (re, im) float pairs, intNrl x intNal looks; NULL = off.
     * Complex path only; not on the matrix-core path. */
    const float *intf;
    int32_t intNr, intNa, intNrl, intNal;
} st_track_params;

/* Output grids, caller-owned host arrays of nA*nR elements (row-major). */
typedef struct {
    float *offR;   /* .dr */
    float *offA;   /* .da */
    float *corr;   /* .cc */
    float *azd;    /* .azd (strack only; may be NULL) */
    uint8_t *type; /* .mt */
    int32_t nComplex, nAmp, nAmpL, nFail; /* strack.h:150-153 */
} st_track_out;

/* Optional per-point diagnostics (any pointer may be NULL).  Attempt index
 * 0 = complex match, 1 = amplitude small window, 2 = amplitude large window.
 * Peaks are in the fftshifted 2x-oversampled correlation surface
 * (findCPeak, speckleTrack.c:744-805); -1000 = rejected by findCPeak,
 * -2000 = attempt not run. */
typedef struct {
    int32_t *r2, *a2;          /* integer position of window 2 (findImage2Pos) */
    int32_t *azShift;          /* azimuth spectral shift used by the complex match */
    int32_t *peakI[3], *peakJ[3];
    int32_t *fineI[3], *fineJ[3]; /* peak in 1/(2*NOVER) px units after oversampling */
    double *cmax[3];           /* normalised correlation of the attempt */
    float *sigRg[3], *sigAz[3];/* 1-D Gaussian widths (gaussPeak1D) */
    /* (best - runner-up) / best over the coarse surface and over the
     * oversampled search box: small values mark near-tied peaks whose
     * argmax may legitimately differ between FP32 implementations. */
    float *peakGap[3], *fineGap[3];
    /* coarse peak over the mean |C| of the correlation surface (a new
     * diagnostic: the reference computes none) */
    float *snr[3];
} st_track_diag;

/* ---- contexts and image pairs ---------------------------------------- */

int st_ctx_create(int32_t device, st_ctx **ctx);
int st_ctx_destroy(st_ctx *ctx);
const char *st_last_error(st_ctx *ctx);

/* Upload an SLC pair (image 1 na1 x nr1, image 2 na2 x nr2 samples,
 * native byte order) and keep it resident.  Replaces the strip buffers of
 * loadBuffer / updateSLCBuffer (speckleTrack.c:1369-1399,
 * corrTrackFast.c:554-602): the whole scene stays in HBM. */
int st_pair_upload(st_ctx *ctx, const void *img1, int32_t nr1, int32_t na1,
                   const void *img2, int32_t nr2, int32_t na2,
                   int32_t sample_type, int32_t *pair);
/* The same from samples as the reference's files store them: c64 samples
 * byte-swapped on the GPU when byteOrder is ST_MSB, ci16 samples always
 * (loadBuffer reads int16 SLCs big-endian whatever -LSB says,
 * speckleTrack.c:1380-1390).  The swap is an HBM-bound pass over the
 * uploaded pair. */
#define ST_LSB 0
#define ST_MSB 1
int st_pair_upload_raw(st_ctx *ctx, const void *img1, int32_t nr1, int32_t na1,
                       const void *img2, int32_t nr2, int32_t na2,
                       int32_t sample_type, int32_t byteOrder, int32_t *pair);
int st_pair_release(st_ctx *ctx, int32_t pair);
/* Copy rows [row0, row0+nrows) of both images of a resident pair back to
 * the host (complex float32). */
int st_pair_download(st_ctx *ctx, int32_t pair, int32_t row0, int32_t nrows,
                     void *img1, void *img2);

/* ---- tracking ---------------------------------------------------------- */

/* strack: complex match, then amplitude small / large retries.
 * Replaces speckleTrack (Strack/strack.h:170, speckleTrack.c:159-346). */
int st_speckle_track(st_ctx *ctx, int32_t pair, const st_track_params *par,
                     st_track_out *out, st_track_diag *diag);

/* Several strack calls (pairs, grids and parameters of their own) as one
 * batch: the calls whose complex matches run on the fused FP32 kernels
 * (square windows 16, 32, 64, 128) share one persistent launch per window
 * class, the
 * rest of each call runs as st_speckle_track does.  Each call's outputs are
 * the ones it gives alone.  pars[k] and outs[k] describe call k; no
 * diagnostics.  No counterpart in the reference (one call per process). */
int st_speckle_track_batch(st_ctx *ctx, int32_t n, const int32_t *pairs, const st_track_params *pars,
                           st_track_out *outs);

/* strackw: edge-padded normalised cross-correlation tracker.
 * Replaces corrTrackFast (Strackw/strackw.h:1, corrTrackFast.c:89-231). */
int st_corr_track_fast(st_ctx *ctx, int32_t pair, const st_track_params *par,
                       st_track_out *out, st_track_diag *diag);

/* ---- autotuning (no counterpart in the reference) -------------------- */
/* Times the complex-match stage of a synthetic job per window (16, 32,
 * 64, 128) over the launch choices -- precision for ST_PREC_AUTO and the
 * fused kernel's workgroups per CU (the cache's staged-path batch cap is
 * written as 0) -- keeps
 * the fastest for this process and device, and writes them to cacheFile
 * (text, one line per window: arch window precision wg_per_cu staged_batch
 * us_per_patch; NULL: not written).  A context reads its architecture's
 * cache at creation ($SPECKLE_AMD_TUNE, else tune/<arch>.txt beside the
 * library directory). */
int st_autotune(st_ctx *ctx, const int32_t *windows, int32_t n, const char *cacheFile);
/* The tuned choices of a window (ST_ERR_ARG when it has none). */
int st_tune_get(st_ctx *ctx, int32_t window, int32_t *precision, int32_t *wgPerCu, int32_t *stagedBatch,
                float *usPerPatch);
/* Read a tune cache into the process's table for ctx's device. */
int st_tune_load(st_ctx *ctx, const char *cacheFile);

/* Work counters and device timings of the last tracking call on a context.
 * Not in the reference (it prints line summaries and stage timers instead,
 * speckleTrack.c:1426-1447, corrTrackFast.c:219-224).  A "patch
 * correlation" is a window pair that reaches the FFT stage: every complex
 * match whose patches pass the data tests, every amplitude try whose
 * windows are in bounds, every strackw patch. */
#define ST_KERNEL_SLOTS 16
typedef struct {
    int64_t nPoints;
    int64_t nCorr[3];      /* patch correlations: complex, amp small, amp large */
    double stageMs[4];     /* device time: geometry, complex, amp small, amp large */
    double totalMs;        /* device time of the whole call */
    double kernelMs[ST_KERNEL_SLOTS]; /* summed duration of the dominant kernels (see kernelNames) */
    int64_t kernelLaunches[ST_KERNEL_SLOTS];
    char kernelNames[ST_KERNEL_SLOTS][32];
} st_track_stats;

int st_last_stats(st_ctx *ctx, st_track_stats *stats);
/* Enable per-kernel timing with HIP events (adds host synchronisation;
 * off by default). */
int st_set_profiling(st_ctx *ctx, int32_t on);

/* ---- culling (Cullst/cullst.h:3-67) ------------------------------------ */

typedef struct {
    int32_t nR, nA;
    float *offR, *offA;     /* in/out, culled in place */
    float *corr;            /* in */
    uint8_t *type;          /* in/out */
    const uint8_t *mask;    /* offsets.mask, may be NULL (maskFlag = 0) */
    float *offRS, *offAS;   /* out: smoothed offsets */
    float *sigmaR, *sigmaA; /* out: plane-fit sigmas */
    int32_t bR, bA;         /* median box (boxSize) */
    int32_t nGood;          /* cullst.c:144 default 17 */
    float maxA, maxR;       /* 1.0, 0.75 */
    int32_t sR, sA;         /* smoothing window (2, 8) */
    int32_t islandThresh;   /* -1 = off */
    int32_t singleMT;       /* -1 = off */
    float corrThresh;       /* 0 */
    int32_t nCulled;        /* out: points culled by the last st_cull_st_data pass */
} st_cull_params;

/* One pass of cullSTData (Cullst/cullSTData.c:13-126). */
int st_cull_st_data(st_ctx *ctx, st_cull_params *cp);
/* cullStats (Cullst/cullStats.c:19-144). */
int st_cull_stats(st_ctx *ctx, st_cull_params *cp);
/* cullSmooth (Cullst/cullSmooth.c:29-103). */
int st_cull_smooth(st_ctx *ctx, st_cull_params *cp);
/* cullIslands (Cullst/cullIslands.c:13-198). */
int st_cull_islands(st_ctx *ctx, st_cull_params *cp);
/* The cullst call order (Cullst/cullst.c:75-100): cullSTData x3,
 * cullStats, cullSmooth, cullIslands when islandThresh > 0. */
int st_cull_run(st_ctx *ctx, st_cull_params *cp);
/* st_cull_run without any device allocation: no counterpart in the
 * reference; it is for callers that cannot allocate device memory (code
 * under stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Every device grid of the chain lives in one fixed device
 * array of about 37 MB per GPU, held from library load until the process
 * ends; grids above 524 288 points (nR * nA) fail with ST_ERR_ARG.  The
 * array is shared by every context on a GPU: the library serialises such
 * calls within a process, and each call initialises what it reads. */
int st_cull_run_noalloc(st_ctx *ctx, st_cull_params *cp);

/* ---- Landsat culling (Cullls/cullls.h:2-35) ---------------------------- */
/* CullLSParams with the matchResult / lsFit fields the stages read.  The
 * type bits, the no-data value and the oversampling factor come from
 * landsatSource64 / mosaicSource headers that are not part of the
 * reference tree, so the caller passes them. */
typedef struct {
    int32_t nx, ny;            /* matches.nx / ny */
    float *X, *Y;              /* in/out: matches.X / Y, culled in place */
    const uint8_t *type;       /* matches.type */
    float *XS, *YS;            /* out: smoothed offsets */
    float *sigmaX, *sigmaY;    /* out: matches.sigmaX / sigmaY */
    int32_t bR, bA;            /* median box (boxSize, default 9) */
    int32_t nGood;             /* default 17 */
    float maxX, maxY;          /* default 0.75, 1.0 */
    int32_t sX, sY;            /* smoothing window (default 3, 3) */
    int32_t islandThresh;      /* -1 = off */
    double deltaT;             /* fitDat.deltaT, days */
    double dx, dy;             /* matches.dx / dy */
    uint8_t slowVelMask;       /* SLOWVEL bit of matches.type */
    uint8_t noDataToUseMask;   /* NODATATOUSE bit of matches.type */
    float noData;              /* NODATA */
    int32_t nover;             /* NOVER of the Landsat matcher (sigma quantisation term) */
    int32_t nMatch, nAttempt;  /* in/out: set by the first data pass when nMatch < 0 */
    int32_t nCulled;           /* out: points culled by the last data pass */
} st_cull_ls_params;

/* One pass of cullLSData (Cullls/cullLSData.c:13-134). */
int st_cull_ls_data(st_ctx *ctx, st_cull_ls_params *cp);
/* cullLSStats (Cullls/cullLSStats.c:21-160). */
int st_cull_ls_stats(st_ctx *ctx, st_cull_ls_params *cp);
/* cullLSSmooth (Cullls/cullLSSmooth.c:11-73). */
int st_cull_ls_smooth(st_ctx *ctx, st_cull_ls_params *cp);
/* cullLSIslands (Cullls/cullLSIslands.c:17-177) on XS / YS. */
int st_cull_ls_islands(st_ctx *ctx, st_cull_ls_params *cp);
/* The cullls call order (Cullls/cullls.c:23-58): nMatch = -1, cullLSData
 * x3, cullLSStats, cullLSSmooth, cullLSIslands when islandThresh > 0. */
int st_cull_ls_run(st_ctx *ctx, st_cull_ls_params *cp);

/* ---- serving --------------------------------------------------------- */
/* A server runs tracking jobs of different scene and window sizes
 * concurrently on one GPU (continuous batching).  Each of its slots owns a
 * stream and a scratch arena capped at the scratch the server charged for
 * it.  A job that names a resident pair is ready at once; a job that
 * brings host images is staged first: a stager thread uploads its images
 * on its own stream, as soon as they fit in the HBM budget, while the
 * slots compute other jobs.  Idle slots take ready jobs in arrival order,
 * so jobs that arrive while others run start as soon as a slot frees up
 * and their kernels overlap on the device.  Admission charges every slot's
 * arena (which holds a job's per-point state and device output grids),
 * the resident pairs and the staged images.  Results do not depend on what
 * else runs.  For strack jobs with a cull chain the offset grids stay on
 * the device from tracking into cullst.  No counterpart in the reference
 * (single-threaded, one job per process: Strack/strack.c:119-135); it is
 * the request loop the function-level entry points sit under. */
typedef struct st_server st_server;

#define ST_JOB_STRACK 0  /* st_speckle_track */
#define ST_JOB_STRACKW 1 /* st_corr_track_fast */
#define ST_JOB_CULL 2    /* cullst alone (st_cull_run) on the host grids of `cull`: no pair, no images */

#define ST_JOB_QUEUED 0
#define ST_JOB_RUNNING 1
#define ST_JOB_DONE 2
#define ST_JOB_FAILED 3
#define ST_JOB_CANCELLED 4

typedef struct {
    int32_t kind;            /* ST_JOB_* */
    int32_t pair;            /* resident pair of the server, or -1: upload img1/img2 for this job */
    const void *img1, *img2; /* host images (pair = -1); must stay valid until the job is done */
    int32_t nr1, na1, nr2, na2, sampleType;
    st_track_params par;     /* copied at submit; its pointer members must stay valid until done */
    st_track_out out;        /* caller-owned output arrays, written when the job is done */
    st_cull_params *cull;    /* optional: the cullst chain (st_cull_run's order) on the job's output
                                grids after tracking; its offR/offA/corr/type must be the out
                                arrays, which then hold the culled grids.  strack jobs cull the
                                grids on the device before they are copied out. */
} st_job;

typedef struct {
    int32_t state;       /* ST_JOB_* */
    int32_t status;      /* ST_OK or the error code of a failed job */
    int32_t slot;        /* slot that ran the job (-1 while queued) */
    int32_t pad;
    double queueMs;      /* submit -> start (wall clock) */
    double runMs;        /* start -> done */
    double latencyMs;    /* submit -> done */
    int64_t patches;     /* patch correlations the job computed */
    int32_t nComplex, nAmp, nAmpL, nFail; /* match counts of the job (st_track_out) */
    char error[128];     /* message of a failed job */
} st_job_info;

/* Create a server on `device` with `slots` concurrent job slots (1..8) and
 * an HBM budget in bytes for resident pairs plus admitted jobs. */
int st_server_create(int32_t device, int32_t slots, uint64_t hbmBudget, st_server **srv);
/* Waits for every submitted job, then frees everything. */
int st_server_destroy(st_server *srv);
/* Resident pairs shared by all jobs of the server (counted in the budget). */
int st_server_pair_upload(st_server *srv, const void *img1, int32_t nr1, int32_t na1, const void *img2,
                          int32_t nr2, int32_t na2, int32_t sampleType, int32_t *pair);
/* Marks the pair as leaving (new jobs naming it are refused), waits until
 * every submitted job that names it has ended, then frees it.  A second
 * release of the same pair fails with ST_ERR_ARG. */
int st_server_pair_release(st_server *srv, int32_t pair);
/* Up to maxJobs ready strack jobs whose complex matches run on the fused
 * FP32 kernel (windows 16, 32, 64) are taken by one slot as a batch and
 * share one persistent launch per window class (st_speckle_track_batch);
 * each job's results are the ones it gives alone, and all of them end
 * together.  1 (the default) turns batching off. */
int st_server_set_batch(st_server *srv, int32_t maxJobs);
/* Queue a job; returns its id at once. */
int st_submit(st_server *srv, const st_job *job, int64_t *id);
/* State of a job without waiting (info may be NULL). */
int st_poll(st_server *srv, int64_t id, st_job_info *info);
/* Block until the job is done, failed or cancelled. */
int st_wait(st_server *srv, int64_t id, st_job_info *info);
/* Cancel a job: a queued job ends at once, a running one at its next stage
 * or batch boundary (its outputs are then unspecified); either way its
 * state becomes ST_JOB_CANCELLED with status ST_ERR_CANCELLED.  ST_ERR_ARG
 * for a job that has already ended. */
int st_cancel(st_server *srv, int64_t id);
/* Forget an ended job (done, failed or cancelled) so the server's job
 * table does not grow without bound; its id is invalid afterwards.
 * ST_ERR_ARG while the job is queued or running. */
int st_job_release(st_server *srv, int64_t id);

#ifdef __cplusplus
}
#endif

#endif /* SPECKLE_AMD_H */
