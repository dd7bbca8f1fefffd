/*
 * speckle_synth.h -- synthetic SLC pairs for tests and benchmarks.
 *
 * Not part of the reference interface (fastice/speckleSource ships no data
 * generator).  Image 1 is circular complex Gaussian speckle, CN(0,1),
 * band-limited like an oversampled SLC (17-tap Kaiser low-pass per axis,
 * azimuth band shifted to the Doppler centroid),
 * modulated by a log-normal backscatter field (a unit-variance Gaussian
 * random field with Gaussian covariance of length corrLen, scaled by
 * sigmaLn).  Image 2 is image 1 resampled at x - d(x) by a 16-tap
 * Kaiser-windowed sinc, mixed with fresh speckle to coherence gamma(x):
 *
 *   img2(x) = g(x) R_d[img1](x) + sqrt(1 - g(x)^2) sqrt(s0(x - d)) s2(x)
 *
 * so a tracker reading image 2 at image-1 position x + d recovers offset
 * d.  Everything is evaluated in double from a counter-based hash, so a
 * given spec always produces the same scene on the host; the device path
 * evaluates the same formulas (last-ulp differences of libm are possible
 * between host and device).
 */
#ifndef SPECKLE_SYNTH_H
#define SPECKLE_SYNTH_H

#include <stdint.h>
#include "speckle_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int32_t nr, na;       /* range samples per line, azimuth lines */
    uint64_t seed;
    /* offset field d(r, a) in pixels (r = range, a = azimuth):
     *   dr = dr0 + drR*r + drA*a + ridgeR*G + shearR*S
     *   da = da0 + daR*r + daA*a + ridgeA*G + shearA*S
     * with G = exp(-(r - ridgeC)^2 / (2 ridgeW^2)) and S = 2 r / nr - 1 */
    double dr0, drR, drA, da0, daR, daA;
    double ridgeR, ridgeA, ridgeC, ridgeW;
    double shearR, shearA;
    /* coherence g = gmin + (gmax-gmin) (0.5 + 0.5 sin(2 pi r/gLr) sin(2 pi a/gLa)) */
    double gmin, gmax, gLr, gLa;
    /* backscatter */
    double sigmaLn, corrLen;
    /* speckle spectrum: band-limited to bwR / bwA of the sampling rate
     * (1 = white), azimuth band centred on the Doppler centroid fdc
     * (cycles per sample), as in an oversampled SLC */
    double bwR, bwA, fdc;
    /* zero-filled rectangle [maskR0, maskR1) x [maskA0, maskA1) in both images */
    int32_t maskR0, maskR1, maskA0, maskA1;
} st_synth_spec;

/* Host generator: writes na*nr complex64 samples per image. */
int st_synth_host(const st_synth_spec *spec, float *img1, float *img2);
/* Device generator: creates a resident pair in ctx (see st_pair_upload). */
int st_synth_device(st_ctx *ctx, const st_synth_spec *spec, int32_t *pair);
/* The true offset field at (r, a). */
void st_synth_offset(const st_synth_spec *spec, double r, double a, double *dr, double *da);
/* Generate a pair in HBM as a resident pair of a server (counted in its
 * budget). */
int st_server_pair_synth(st_server *srv, const st_synth_spec *spec, int32_t *pair);

#ifdef __cplusplus
}
#endif

#endif
