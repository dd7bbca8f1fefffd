/*
 * strack / strackw on the MI355X library: the command-line programs of
 * Strack/strack.c:26-183 and Strackw/strackw.c:26-155 over the C ABI.
 *
 *   strack  [-noComplex -singleAmp -gauss -integerComplex -noHanning -checkAzFocus -LSB] parFile
 *   strackw [-integerComplex -LSB] parFile
 *
 * The parameter file is read by st_read_track_par (parseTrack's keywords),
 * the SLC images through the VRT beside each image (parseSLCVrtNew), the
 * initial offsets (parseInitialOffsets, readBothOffsetsStrack) and offset
 * masks (getMask) through the library's readers, and the products are
 * written by st_write_offsets (sTrackOut, writeDatFile, writeVrtFile).
 * A baseline file is read (parseBase) and must hold a zero baseline.  Not
 * available here, because the readers live in GrIMP libraries outside the
 * reference tree: SAR .par image headers (readOldPar), geodat masks, the
 * interferogram correction (getInt's geodat) and non-zero baselines (the
 * look-angle tables of initPhaseCorrect); the program stops with a message
 * when the parameter file asks for them.  The library's C ABI takes the
 * interferogram and the baseline tables directly.
 * Built as both programs; the name it is run under picks the tracker.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "speckle_amd.h"
#include "speckle_io.h"

static void die(const char *what, const char *msg)
{
    fprintf(stderr, "%s: %s\n", what, msg ? msg : "");
    exit(1);
}

static void usage(int strackw)
{
    if (strackw)
        die("usage", "sTrack -integerComplex -LSB parFile\n\tLSB  use LSB for both output and flaoting point input");
    die("usage", "sTrack -noComplex -singleAmp -gauss -integerComplex -noHanning -checkAzFocus -LSB parFile\n\tLSB  "
                 "use LSB for both output and floating point input");
}

typedef struct {
    const char *parFile;
    int noComplex, floatFlag, hanningFlag, gaussFlag, maxTries, byteOrder, checkAzFocus;
} Args;

/* readArgs (strack.c:137-178, strackw.c:130-151): the flags are matched as
 * substrings of the text from the first '-' on */
static void read_args(int argc, char **argv, int strackw, Args *a)
{
    memset(a, 0, sizeof(*a));
    a->floatFlag = 1;
    a->hanningFlag = 1;
    a->maxTries = 2;
    a->byteOrder = ST_MSB;
    if (argc < 2 || argc > (strackw ? 4 : 6))
        usage(strackw);
    a->parFile = argv[argc - 1];
    for (int i = 1; i <= argc - 2; i++) {
        const char *s = strchr(argv[i], '-');
        if (!s)
            usage(strackw);
        if (!strackw && strstr(s, "noComplex"))
            a->noComplex = 1;
        else if (strstr(s, "integerComplex"))
            a->floatFlag = 0;
        else if (!strackw && strstr(s, "-noHanning"))
            a->hanningFlag = 0;
        else if (!strackw && strstr(s, "-gauss"))
            a->gaussFlag = 1;
        else if (!strackw && strstr(s, "-singleAmp"))
            a->maxTries = 1;
        else if (strstr(s, "-LSB"))
            a->byteOrder = ST_LSB;
        else if (!strackw && strstr(s, "-checkAzFocus"))
            a->checkAzFocus = 1;
        else
            usage(strackw);
    }
}

/* one SLC image through its VRT: size and samples as (re, im) floats */
static float *read_slc(const char *image, const char *vrtFile, int32_t *nr, int32_t *na)
{
    if (!vrtFile || !vrtFile[0])
        die(image, "no VRT beside the image: SAR .par headers (readOldPar) are not supported");
    st_vrt *v = NULL;
    int32_t nb = 0;
    if (st_vrt_open(vrtFile, &v) != ST_OK || st_vrt_size(v, nr, na, &nb) != ST_OK)
        die(vrtFile, st_vrt_last_error());
    float *buf = (float *)malloc(sizeof(float) * 2 * (size_t)*nr * *na);
    if (!buf)
        die(image, "out of memory");
    if (st_vrt_read_c64(v, 1, buf) != ST_OK)
        die(vrtFile, st_vrt_last_error());
    st_vrt_close(v);
    return buf;
}

/* the first data line of a baseline file (blank lines and lines starting
 * with ';' skipped: GrIMP's getDataString, outside the reference tree) */
static int read_base(const char *file, double b[4])
{
    FILE *f = fopen(file, "r");
    if (!f)
        die("parseBase: could not open", file);
    char line[1024];
    int n = 0;
    while (fgets(line, sizeof(line), f)) {
        const char *c = line;
        while (*c == ' ' || *c == '\t')
            c++;
        if (*c == '\0' || *c == '\n' || *c == ';')
            continue;
        n = sscanf(c, "%lf%lf%lf%lf", &b[0], &b[1], &b[2], &b[3]);
        break;
    }
    fclose(f);
    return n;
}

static const char *base_name(const char *p)
{
    const char *s = strrchr(p, '/');
    return s ? s + 1 : p;
}

int main(int argc, char **argv)
{
    const int strackw = strstr(base_name(argv[0]), "strackw") != NULL;
    Args a;
    read_args(argc, argv, strackw, &a);
    st_track_par_file pf;
    if (st_read_track_par(a.parFile, a.noComplex, &pf) != ST_OK)
        die(a.parFile, st_io_last_error());
    if (pf.intFlag)
        die("parseTrack", "intfile: the interferogram correction needs getInt's geodat reader (not in the reference "
                          "tree); pass the interferogram through st_speckle_track");
    if (pf.baseFlag && !a.noComplex && !strackw) {
        /* parseBase (Strack/parseBase.c:11-44): Bn Bp dBn dBp on the first
         * data line.  A zero baseline makes phaseCorrectBaseline the
         * identity; any other needs the look-angle tables of
         * initPhaseCorrect, built from SAR geometry outside the reference
         * tree. */
        double b[4];
        if (read_base(pf.baseParams, b) != 4)
            die("parseBase -- Missing baseline at line", pf.baseParams);
        if (b[0] != 0 || b[1] != 0 || b[2] != 0 || b[3] != 0)
            die("parseBase", "a non-zero baseline needs the SAR geometry of initPhaseCorrect (not in the reference "
                             "tree); pass the range tables through st_speckle_track");
    }

    st_track_params p;
    memset(&p, 0, sizeof(p));
    p.rStart = pf.rStart;
    p.aStart = pf.aStart;
    p.deltaR = pf.deltaR;
    p.deltaA = pf.deltaA;
    p.nR = pf.nR;
    p.nA = pf.nA;
    p.wR = pf.wR;
    p.wA = pf.wA;
    p.wRa = pf.wRa;
    p.wAa = pf.wAa;
    p.edgePadR = pf.edgePadR;
    p.edgePadA = pf.edgePadA;
    p.navgR = pf.navgR;
    p.navgA = pf.navgA;
    p.scaleFactor = pf.scaleFactor > 0 ? pf.scaleFactor : 1;
    p.noComplex = strackw ? 1 : a.noComplex;
    p.hanningFlag = a.hanningFlag;
    p.gaussFlag = a.gaussFlag;
    p.maxTries = a.maxTries;
    p.checkAzFocus = a.checkAzFocus;
    p.azDefocusThresh = 1.8f; /* strack.c:45 */
    p.precision = ST_PREC_FP32;
    p.polyShift = pf.polyShift;
    /* parseInitialOffsets: the polynomials, also the fallback of the grids */
    if (pf.initialOffsetFile[0] && st_read_offset_polys(pf.initialOffsetFile, p.rShiftPoly, p.aShiftPoly) != ST_OK)
        die(pf.initialOffsetFile, st_io_last_error());
    float *initDr = NULL, *initDa = NULL;
    if (!pf.polyShift) { /* readBothOffsetsStrack: <initShift>.da / .dr with <initShift>.da.dat */
        char da[ST_PAR_PATH + 8], dr[ST_PAR_PATH + 8];
        snprintf(da, sizeof(da), "%s.da", pf.initShift);
        snprintf(dr, sizeof(dr), "%s.dr", pf.initShift);
        int32_t r0, a0, nr, na;
        float dR, dA;
        if (st_read_initial_offsets_header(da, &r0, &a0, &nr, &na, &dR, &dA) != ST_OK)
            die(da, st_io_last_error());
        initDr = (float *)malloc(sizeof(float) * (size_t)nr * na);
        initDa = (float *)malloc(sizeof(float) * (size_t)nr * na);
        if (st_read_initial_offsets(da, dr, nr, na, initDa, initDr) != ST_OK)
            die(da, st_io_last_error());
        p.initDr = initDr;
        p.initDa = initDa;
        p.initNr = nr;
        p.initNa = na;
        p.initR0 = r0;
        p.initA0 = a0;
        p.initDeltaR = dR;
        p.initDeltaA = dA;
    }
    uint8_t *mask = NULL;
    if (pf.maskFlag) { /* getMask, offset-mask branch */
        if (pf.maskType != ST_OFFSETMASK)
            die(pf.maskFile, "geodat masks need the geodat reader (not in the reference tree)");
        int32_t h[6];
        if (st_read_mask_header(pf.maskGeodat, h) != ST_OK)
            die(pf.maskGeodat, st_io_last_error());
        mask = (uint8_t *)malloc((size_t)h[2] * h[3]);
        if (st_read_mask(pf.maskFile, h[2], h[3], mask) != ST_OK)
            die(pf.maskFile, st_io_last_error());
        p.mask = mask;
        p.maskR0 = h[0];
        p.maskA0 = h[1];
        p.maskNr = h[2];
        p.maskNa = h[3];
        p.maskNrl = h[4];
        p.maskNal = h[5];
    }

    int32_t nr1, na1, nr2, na2;
    float *img1 = read_slc(pf.image1, pf.vrt1, &nr1, &na1);
    float *img2 = read_slc(pf.image2, pf.vrt2, &nr2, &na2);
    st_ctx *ctx = NULL;
    const char *dev = getenv("SPECKLE_AMD_DEVICE");
    if (st_ctx_create(dev ? atoi(dev) : 0, &ctx) != ST_OK)
        die(argv[0], "no usable MI355X");
    int32_t pair = -1;
    if (st_pair_upload(ctx, img1, nr1, na1, img2, nr2, na2, ST_SAMPLE_C64, &pair) != ST_OK)
        die("image upload", st_last_error(ctx));
    free(img1);
    free(img2);

    const size_t n = (size_t)p.nR * p.nA;
    st_track_out o;
    memset(&o, 0, sizeof(o));
    o.offR = (float *)malloc(sizeof(float) * n);
    o.offA = (float *)malloc(sizeof(float) * n);
    o.corr = (float *)malloc(sizeof(float) * n);
    o.type = (uint8_t *)malloc(n);
    o.azd = strackw ? NULL : (float *)malloc(sizeof(float) * n);
    const int rc = strackw ? st_corr_track_fast(ctx, pair, &p, &o, NULL) : st_speckle_track(ctx, pair, &p, &o, NULL);
    if (rc != ST_OK)
        die(strackw ? "corrTrackFast" : "speckleTrack", st_last_error(ctx));
    fprintf(stderr, "nComplex %i nAmp %i nAmpL %i nFail %i\n", o.nComplex, o.nAmp, o.nAmpL, o.nFail);

    st_offsets_meta meta;
    memset(&meta, 0, sizeof(meta));
    meta.rStart = p.rStart;
    meta.aStart = p.aStart;
    meta.deltaR = p.deltaR;
    meta.deltaA = p.deltaA;
    meta.scaleFactor = p.scaleFactor;
    meta.wR = p.wR;
    meta.wA = p.wA;
    meta.wRa = p.wRa;
    meta.wAa = p.wAa;
    meta.intGeodat = pf.intGeodat[0] ? pf.intGeodat : NULL;
    meta.imageFile1 = pf.image1;
    meta.imageFile2 = pf.image2;
    meta.maskFile = pf.maskFlag ? pf.maskFile : NULL;
    if (st_write_offsets(pf.outputFile, &meta, o.offR, o.offA, o.corr, o.azd, o.type, p.nR, p.nA, a.byteOrder) != ST_OK)
        die(pf.outputFile, st_io_last_error());
    st_pair_release(ctx, pair);
    st_ctx_destroy(ctx);
    return 0;
}
