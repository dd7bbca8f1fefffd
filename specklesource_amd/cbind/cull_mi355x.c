/*
 * cull_mi355x.c -- cullSTData, cullStats, cullSmooth and cullIslands
 * (Cullst/cullst.h:60-67; cullSTData.c, cullStats.c, cullSmooth.c,
 * cullIslands.c) on the MI355X library: the translation unit that replaces
 * those four files in the cullst link line.  CullParams' grids are
 * row-pointer arrays over contiguous blocks, so the GPU stages work on
 * grid[0] in place.  Failures print a message and exit(1), as the
 * reference's error() does.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "grimp_track.h"
#include "speckle_amd.h"

static st_ctx *gCtx;

static st_ctx *ctx(void)
{
    if (!gCtx) {
        const char *e = getenv("SPECKLE_AMD_DEVICE");
        if (st_ctx_create(e ? atoi(e) : 0, &gCtx) != ST_OK) {
            fprintf(stderr, "cullst: no usable MI355X\n");
            exit(1);
        }
    }
    return gCtx;
}

static float *row0(float **g) { return g ? g[0] : NULL; }

static void to_params(const CullParams *c, st_cull_params *p)
{
    memset(p, 0, sizeof(*p));
    p->nR = c->nR;
    p->nA = c->nA;
    p->offR = row0(c->offR);
    p->offA = row0(c->offA);
    p->corr = row0(c->corr);
    p->type = c->type ? (uint8_t *)c->type[0] : NULL;
    p->mask = (c->maskFlag && c->mask) ? (const uint8_t *)c->mask[0] : NULL;
    p->offRS = row0(c->offRS);
    p->offAS = row0(c->offAS);
    p->sigmaR = row0(c->sigmaR);
    p->sigmaA = row0(c->sigmaA);
    p->bR = c->bR;
    p->bA = c->bA;
    p->nGood = c->nGood;
    p->maxA = c->maxA;
    p->maxR = c->maxR;
    p->sR = c->sR;
    p->sA = c->sA;
    p->islandThresh = c->islandThresh;
    p->singleMT = c->singleMT;
    p->corrThresh = c->corrThresh;
}

static void call(const char *what, int (*fn)(st_ctx *, st_cull_params *), CullParams *c)
{
    st_cull_params p;
    to_params(c, &p);
    if (fn(ctx(), &p) != ST_OK) {
        fprintf(stderr, "%s: %s\n", what, st_last_error(ctx()));
        exit(1);
    }
    /* cullStats forces the 9 x 9 box (cullStats.c:38-39) */
    c->bR = p.bR;
    c->bA = p.bA;
}

void cullSTData(CullParams *cullPar) { call("cullSTData", st_cull_st_data, cullPar); }

void cullStats(CullParams *cullPar)
{
    if (!cullPar->sigmaR)
        cullPar->sigmaR = grimp_grid_f(cullPar->nA, cullPar->nR);
    if (!cullPar->sigmaA)
        cullPar->sigmaA = grimp_grid_f(cullPar->nA, cullPar->nR);
    call("cullStats", st_cull_stats, cullPar);
}

void cullSmooth(CullParams *cullPar)
{
    if (!cullPar->offRS)
        cullPar->offRS = grimp_grid_f(cullPar->nA, cullPar->nR);
    if (!cullPar->offAS)
        cullPar->offAS = grimp_grid_f(cullPar->nA, cullPar->nR);
    call("cullSmooth", st_cull_smooth, cullPar);
}

void cullIslands(CullParams *cullPar) { call("cullIslands", st_cull_islands, cullPar); }
