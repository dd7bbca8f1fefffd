/*
 * ORACLE -- test infrastructure, never shipped or measured as the product.
 *
 * CPU restatement of the fastice/speckleSource hot path (strack, strackw,
 * cullst).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker / CPU baseline.
 *
 * The same sources are compiled twice:
 *   ORC_REAL=float   "FP32-faithful" oracle: fftw_real is float in the
 *                    reference (FFTW 2.1.5 single precision), the scalar
 *                    bookkeeping that the reference does in double stays
 *                    double.
 *   ORC_REAL=double  "FP64" oracle: every fftw_real quantity in double; it
 *                    is the yardstick for sub-pixel offset error and
 *                    correlation / SNR relative error.
 *
 * The FFT is this project's own (orc_fft.c), with FFTW 2's conventions:
 * FORWARD uses exp(-2 pi i jk/n), BACKWARD exp(+2 pi i jk/n), neither is
 * normalised.  Its rounding is not FFTW 2.1.5's (which is not vendored and
 * cannot be built here), so last-ulp behaviour is "parity unpinned".
 */
#ifndef ORC_COMMON_H
#define ORC_COMMON_H

#include <stdint.h>
#include <stddef.h>

#ifndef ORC_REAL
#define ORC_REAL float
#endif
typedef ORC_REAL oreal;

typedef struct {
    oreal re, im;
} ocplx;

#define ORC_FORWARD (-1)
#define ORC_BACKWARD (+1)

/* GrIMP's PI (clib/standard.h is not vendored; the value is the double
 * nearest pi). */
#define ORC_PI 3.14159265358979323846

/* -LARGEINT sentinel (Strack/writeVrt.c:69); LARGEINT = 2e9 in the
 * clib headers, -2e9 in Cullst/cullst.c:11 only. */
#define ORC_LARGEINT 2000000000.0

/* 1-D in-place FFT of n points with stride 1. */
void orc_fft1d(ocplx *x, int n, int sign);
/* 2-D in-place FFT of an na x nr row-major array (rows = azimuth). */
void orc_fft2d(ocplx *x, int na, int nr, int sign);
/* Release cached plans. */
void orc_fft_cleanup(void);

#endif
