/*
 * Plain-C test of the reference-side bindings (track_mi355x.c,
 * cull_mi355x.c): the reference's own entry points, called as strack.c,
 * strackw.c and cullst.c call them, on SLC files written by the test
 * harness.
 *
 *   test_bindings dir
 *
 * dir/p32_1.slc, dir/p32_2.slc: 256 x 256 MSB complex64, image 2 = image 1
 *   shifted by (3, -2) px, both periodic with period 32 (every 32 x 32
 *   window of image 2 is a circular shift of image 1's);
 * dir/p64_1.slc, dir/p64_2.slc: 400 x 400, period 64, shift (-5, 4).
 * speckleTrack must find the exact shift at every point as complex matches
 * (closed form), corrTrackFast within one 1/32 px quantum; cullSTData keeps
 * the flat field and culls a planted outlier; cullStats / cullSmooth /
 * cullIslands run on it.  Exit 0 and "bindings ok" on success.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "grimp_track.h"

static int failures;

#define CHECK(cond, ...)                          \
    do {                                          \
        if (!(cond)) {                            \
            fprintf(stderr, "FAIL: " __VA_ARGS__); \
            fprintf(stderr, "\n");                \
            failures++;                           \
        }                                         \
    } while (0)

static char *path(const char *dir, const char *name)
{
    char *p = malloc(strlen(dir) + strlen(name) + 2);
    sprintf(p, "%s/%s", dir, name);
    return p;
}

static void base_params(TrackParams *t, const char *dir, const char *f1, const char *f2, int n)
{
    memset(t, 0, sizeof(*t));
    t->imageFile1 = path(dir, f1);
    t->imageFile2 = path(dir, f2);
    t->imageP1.nSlpR = t->imageP1.nSlpA = n;
    t->imageP2.nSlpR = t->imageP2.nSlpA = n;
    t->floatFlag = TRUE;
    t->byteOrder = MSB;
    t->scaleFactor = 1;
    t->polyShift = TRUE;
    t->hanningFlag = FALSE;
    t->maxTries = 2;
    t->azDefocusThresh = 1.8f;
}

int main(int argc, char **argv)
{
    if (argc != 2) {
        fprintf(stderr, "usage: test_bindings dir\n");
        return 2;
    }
    const char *dir = argv[1];

    /* speckleTrack: complex matches of circularly shifted windows */
    TrackParams t;
    base_params(&t, dir, "p32_1.slc", "p32_2.slc", 256);
    t.wR = t.wA = t.wRa = t.wAa = 32;
    t.rStart = t.aStart = 72;
    t.deltaR = t.deltaA = 16;
    t.nR = t.nA = 8;
    t.rShiftPoly[0] = 3;
    t.aShiftPoly[0] = -2;
    speckleTrack(&t);
    CHECK(t.nComplex == 64 && t.nFail == 0, "speckleTrack counts %d %d %d %d", t.nComplex, t.nAmp, t.nAmpL, t.nFail);
    for (int i = 0; i < t.nA; i++)
        for (int j = 0; j < t.nR; j++) {
            CHECK(t.type[i][j] == 1, "type[%d][%d] = %d", i, j, t.type[i][j]);
            CHECK(t.offR[i][j] == 3.0f && t.offA[i][j] == -2.0f, "offset[%d][%d] = %g %g", i, j, t.offR[i][j],
                  t.offA[i][j]);
            CHECK(fabsf(t.corr[i][j] - 1.0f) < 2e-4f, "corr[%d][%d] = %g", i, j, t.corr[i][j]);
        }

    /* corrTrackFast on the period-64 pair */
    TrackParams w;
    base_params(&w, dir, "p64_1.slc", "p64_2.slc", 400);
    w.wR = w.wA = w.wRa = w.wAa = 64;
    w.edgePadR = w.edgePadA = 10;
    w.navgR = w.navgA = 2;
    w.noComplex = TRUE;
    w.rStart = w.aStart = 136;
    w.deltaR = w.deltaA = 48;
    w.nR = w.nA = 3;
    corrTrackFast(&w);
    CHECK(w.nAmp == 9 && w.nFail == 0, "corrTrackFast counts %d %d", w.nAmp, w.nFail);
    for (int i = 0; i < w.nA; i++)
        for (int j = 0; j < w.nR; j++) {
            CHECK(w.type[i][j] == 2, "strackw type[%d][%d] = %d", i, j, w.type[i][j]);
            CHECK(fabsf(w.offR[i][j] + 5.0f) <= 1.0f / 32 + 1e-6f && fabsf(w.offA[i][j] - 4.0f) <= 1.0f / 32 + 1e-6f,
                  "strackw offset[%d][%d] = %g %g", i, j, w.offR[i][j], w.offA[i][j]);
        }

    /* cullst's order on the strack grid with one planted outlier */
    CullParams c;
    memset(&c, 0, sizeof(c));
    c.nR = t.nR;
    c.nA = t.nA;
    c.offR = t.offR;
    c.offA = t.offA;
    c.corr = t.corr;
    c.type = t.type;
    c.offR[4][4] += 20.0f;
    c.bR = c.bA = 5;
    c.nGood = 5;
    c.maxA = 1.0f;
    c.maxR = 0.75f;
    c.sR = 2;
    c.sA = 2;
    c.islandThresh = 2;
    c.singleMT = -1;
    cullSTData(&c);
    cullSTData(&c);
    cullSTData(&c);
    CHECK(c.offR[4][4] < -1e9f, "planted outlier kept: %g", c.offR[4][4]);
    CHECK(c.offR[3][3] == 3.0f && c.offA[5][5] == -2.0f, "flat field culled");
    cullStats(&c);
    CHECK(c.bR == 9 && c.bA == 9, "cullStats box %d %d", c.bR, c.bA);
    cullSmooth(&c);
    CHECK(fabsf(c.offRS[2][2] - 3.0f) < 1e-5f, "smoothed offset %g", c.offRS[2][2]);
    cullIslands(&c);
    if (failures) {
        fprintf(stderr, "%d failures\n", failures);
        return 1;
    }
    printf("bindings ok\n");
    return 0;
}
