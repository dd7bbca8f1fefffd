/*
 * cullst on the MI355X library: the command-line program of
 * Cullst/cullst.c:20-371 over the C ABI.
 *
 *   cullst [-islandThresh n -singleMT n -maxR x -maxA x -nGood n -boxSize n
 *           -sr n -sa n -corrThresh x -ignoreOffsets] inbase outbase
 *
 * inbase.dat / .dr / .da / .cc / .mt in (loadCullData's .dat path, MSB),
 * offsets.mask when present and the same size (loadCullMask), the cull
 * chain (cullSTData x3, cullStats, cullSmooth, cullIslands), and the
 * products of writeCullData.  -useSim needs the simulated-offset loader
 * of loadSimData and stops with a message.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "speckle_amd.h"
#include "speckle_io.h"

static void die(const char *what, const char *msg)
{
    fprintf(stderr, "%s: %s\n", what, msg ? msg : "");
    exit(1);
}

static void usage(void)
{
    die("usage", "cullst -useSim simBase -ignoreOffsets -islandThresh islandThresh -singleMT singleMT -maxR maxR "
                 "-maxA maxA -nGood nGood -boxSize boxSize -sr sr -sa sa -corrThresh corrThresh inbase outbase");
}

int main(int argc, char **argv)
{
    /* readArgs (cullst.c:123-250): defaults, then flags matched as
     * substrings of the text from the first '-' on, in this order */
    int sr = 2, sa = 8, boxSize = 9, nGood = 17, singleMT = -1, islandThresh = -1, ignoreOffsets = 0;
    float maxA = 1.0f, maxR = 0.75f, corrThresh = 0.0f;
    if (argc < 3)
        usage();
    const int n = argc - 3;
    for (int i = 1; i <= n; i++) {
        const char *s = strchr(argv[i], '-');
        if (!s)
            usage();
        if (strstr(s, "sr")) {
            sscanf(argv[i + 1], "%i", &sr);
            i++;
        } else if (strstr(s, "sa")) {
            sscanf(argv[i + 1], "%i", &sa);
            i++;
        } else if (strstr(s, "corrThresh")) {
            sscanf(argv[i + 1], "%f", &corrThresh);
            i++;
            if (corrThresh < 0 || corrThresh >= 1) {
                fprintf(stderr, "\n\ncoorThresh must be in interval (0,1)\n\n");
                usage();
            }
        } else if (strstr(s, "boxSize")) {
            sscanf(argv[i + 1], "%i", &boxSize);
            i++;
        } else if (strstr(s, "nGood")) {
            sscanf(argv[i + 1], "%i", &nGood);
            i++;
        } else if (strstr(s, "singleMT")) {
            sscanf(argv[i + 1], "%i", &singleMT);
            i++;
            if (singleMT < 1 || singleMT > 3) {
                fprintf(stderr, "singleMT should be in range 0 to 3\n");
                usage();
            }
        } else if (strstr(s, "maxA")) {
            sscanf(argv[i + 1], "%f", &maxA);
            i++;
        } else if (strstr(s, "maxR")) {
            sscanf(argv[i + 1], "%f", &maxR);
            i++;
        } else if (strstr(s, "islandThresh")) {
            sscanf(argv[i + 1], "%i", &islandThresh);
            i++;
        } else if (strstr(s, "useSim")) {
            die("cullst", "-useSim needs the simulated-offset loader (loadSimData), which this build does not have");
        } else if (strstr(s, "ignoreOffsets")) {
            ignoreOffsets = 1;
        } else {
            usage();
        }
    }
    const char *inBase = argv[argc - 2], *outBase = argv[argc - 1];

    char datFile[4096];
    snprintf(datFile, sizeof(datFile), "%s.dat", inBase);
    int32_t hdr[6];
    char geo1[1024], geo2[1024];
    if (st_read_dat(datFile, hdr, geo1, geo2, sizeof(geo1)) != ST_OK)
        die(datFile, st_io_last_error());
    const int32_t nR = hdr[2], nA = hdr[3];
    const size_t cnt = (size_t)nR * nA;
    float *offR = malloc(4 * cnt), *offA = malloc(4 * cnt), *corr = malloc(4 * cnt);
    float *offRS = malloc(4 * cnt), *offAS = malloc(4 * cnt), *sigR = malloc(4 * cnt), *sigA = malloc(4 * cnt);
    uint8_t *type = malloc(cnt), *mask = NULL;
    if (!offR || !offA || !corr || !offRS || !offAS || !sigR || !sigA || !type)
        die("cullst", "out of memory");
    if (st_read_offsets(inBase, nR, nA, ST_MSB, offR, offA, corr, type) != ST_OK)
        die(inBase, st_io_last_error());
    /* loadCullMask (loadCullData.c:231-256): offsets.mask with offsets.dat
     * of the same size, unless -ignoreOffsets */
    if (access("offsets.mask", F_OK) != -1 && !ignoreOffsets) {
        int32_t mh[6];
        if (st_read_dat("offsets.dat", mh, NULL, NULL, 0) != ST_OK)
            die("offsets.dat", st_io_last_error());
        if (mh[2] == nR && mh[3] == nA) {
            FILE *f = fopen("offsets.mask", "rb");
            mask = malloc(cnt);
            if (!f || !mask || fread(mask, 1, cnt, f) != cnt)
                die("offsets.mask", "short file");
            fclose(f);
        } else {
            fprintf(stderr, "Ignoring mask because it is not the same size: regoffsets ?\n");
        }
    }

    st_cull_params cp;
    memset(&cp, 0, sizeof(cp));
    cp.nR = nR;
    cp.nA = nA;
    cp.offR = offR;
    cp.offA = offA;
    cp.corr = corr;
    cp.type = type;
    cp.mask = mask;
    cp.offRS = offRS;
    cp.offAS = offAS;
    cp.sigmaR = sigR;
    cp.sigmaA = sigA;
    cp.bR = cp.bA = boxSize;
    cp.nGood = nGood;
    cp.maxA = maxA;
    cp.maxR = maxR;
    cp.sR = sr;
    cp.sA = sa;
    cp.islandThresh = islandThresh;
    cp.singleMT = singleMT;
    cp.corrThresh = corrThresh;
    st_ctx *ctx = NULL;
    const char *dev = getenv("SPECKLE_AMD_DEVICE");
    if (st_ctx_create(dev ? atoi(dev) : 0, &ctx) != ST_OK)
        die("cullst", "no usable MI355X");
    if (st_cull_run(ctx, &cp) != ST_OK)
        die("cullst", st_last_error(ctx));

    st_cull_meta meta;
    memset(&meta, 0, sizeof(meta));
    meta.rStart = hdr[0];
    meta.aStart = hdr[1];
    meta.deltaR = hdr[4];
    meta.deltaA = hdr[5];
    meta.sR = sr;
    meta.sA = sa;
    meta.boxSize = boxSize;
    meta.nGood = nGood;
    meta.singleMT = singleMT;
    meta.maxR = maxR;
    meta.maxA = maxA;
    meta.corrThresh = corrThresh;
    meta.geo1 = geo1[0] ? geo1 : NULL;
    meta.geo2 = geo2[0] ? geo2 : NULL;
    if (st_write_cull(outBase, inBase, &meta, offRS, offAS, sigR, sigA, corr, type, nR, nA, ST_MSB) != ST_OK)
        die(outBase, st_io_last_error());
    st_ctx_destroy(ctx);
    return 0;
}
