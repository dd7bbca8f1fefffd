/*
 * ORACLE -- test infrastructure (see orc_common.h).
 *
 * Mixed-radix Stockham autosort FFT used by the CPU restatement in place
 * of FFTW 2.1.5 (Makefile:46,147-183 of the reference; not vendored).
 * Conventions follow FFTW 2: FORWARD = exp(-2 pi i jk/n), BACKWARD =
 * exp(+2 pi i jk/n), unnormalised, in place for the caller.
 *
 * Radices 4, 2, 3, 5, 7 and a generic O(p^2) butterfly for other primes.
 * Twiddles are evaluated in double and rounded to oreal once.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc_common.h"

#define MAX_STAGES 32

typedef struct {
    int n, sign;
    int nst;
    int radix[MAX_STAGES];
    int ns[MAX_STAGES];       /* sub-transform size before the stage */
    ocplx *tw[MAX_STAGES];    /* per stage: ns * radix twiddles */
    ocplx *root[MAX_STAGES];  /* per stage: radix roots exp(sign 2 pi i q/p) */
    ocplx *work;
} fft_plan;

#define MAX_PLANS 64
static fft_plan *g_plans[MAX_PLANS];
static int g_nplans;

static fft_plan *make_plan(int n, int sign)
{
    fft_plan *p = (fft_plan *)calloc(1, sizeof(fft_plan));
    int m = n, k = 0, ns = 1;
    p->n = n;
    p->sign = sign;
    while (m > 1) {
        int r;
        if (m % 4 == 0)
            r = 4;
        else if (m % 2 == 0)
            r = 2;
        else if (m % 3 == 0)
            r = 3;
        else if (m % 5 == 0)
            r = 5;
        else if (m % 7 == 0)
            r = 7;
        else {
            r = 11;
            while (m % r)
                r += 2;
        }
        p->radix[k] = r;
        p->ns[k] = ns;
        p->tw[k] = (ocplx *)malloc(sizeof(ocplx) * (size_t)ns * r);
        for (int j = 0; j < ns; j++)
            for (int q = 0; q < r; q++) {
                double ang = sign * 2.0 * ORC_PI * (double)(j * q) / (double)(ns * r);
                p->tw[k][j * r + q].re = (oreal)cos(ang);
                p->tw[k][j * r + q].im = (oreal)sin(ang);
            }
        p->root[k] = (ocplx *)malloc(sizeof(ocplx) * (size_t)r);
        for (int q = 0; q < r; q++) {
            double ang = sign * 2.0 * ORC_PI * (double)q / (double)r;
            p->root[k][q].re = (oreal)cos(ang);
            p->root[k][q].im = (oreal)sin(ang);
        }
        ns *= r;
        m /= r;
        k++;
    }
    p->nst = k;
    p->work = (ocplx *)malloc(sizeof(ocplx) * (size_t)(n > 0 ? n : 1));
    return p;
}

static fft_plan *get_plan(int n, int sign)
{
    for (int i = 0; i < g_nplans; i++)
        if (g_plans[i]->n == n && g_plans[i]->sign == sign)
            return g_plans[i];
    fft_plan *p = make_plan(n, sign);
    if (g_nplans < MAX_PLANS)
        g_plans[g_nplans++] = p;
    return p;
}

void orc_fft_cleanup(void)
{
    for (int i = 0; i < g_nplans; i++) {
        fft_plan *p = g_plans[i];
        for (int k = 0; k < p->nst; k++) {
            free(p->tw[k]);
            free(p->root[k]);
        }
        free(p->work);
        free(p);
    }
    g_nplans = 0;
}

static inline ocplx cmul(ocplx a, ocplx b)
{
    ocplx c;
    c.re = a.re * b.re - a.im * b.im;
    c.im = a.re * b.im + a.im * b.re;
    return c;
}

/* One Stockham stage: x (n points, sub-transforms of size ns, p of them
 * interleaved with stride n/p) -> y (sub-transforms of size ns*p). */
static void stage(const ocplx *x, ocplx *y, int n, int p, int ns, const ocplx *tw,
                  const ocplx *root, int sign)
{
    const int stride = n / p;
    ocplx v[64];
    ocplx *vv = v;
    ocplx *heap = NULL;
    if (p > 64) {
        heap = (ocplx *)malloc(sizeof(ocplx) * (size_t)p * 2);
        vv = heap;
    }
    for (int j = 0; j < stride; j++) {
        const int k = j % ns;
        const ocplx *t = tw + (size_t)k * p;
        for (int r = 0; r < p; r++) {
            ocplx a = x[j + r * stride];
            vv[r] = (r == 0) ? a : cmul(a, t[r]);
        }
        const int base = (j / ns) * ns * p + k;
        if (p == 2) {
            ocplx a = vv[0], b = vv[1];
            y[base].re = a.re + b.re;
            y[base].im = a.im + b.im;
            y[base + ns].re = a.re - b.re;
            y[base + ns].im = a.im - b.im;
        } else if (p == 4) {
            ocplx a = vv[0], b = vv[1], c = vv[2], d = vv[3];
            oreal s0r = a.re + c.re, s0i = a.im + c.im;
            oreal d0r = a.re - c.re, d0i = a.im - c.im;
            oreal s1r = b.re + d.re, s1i = b.im + d.im;
            oreal d1r = b.re - d.re, d1i = b.im - d.im;
            /* multiply (b - d) by sign*i */
            oreal er, ei;
            if (sign < 0) {
                er = d1i;
                ei = -d1r;
            } else {
                er = -d1i;
                ei = d1r;
            }
            y[base].re = s0r + s1r;
            y[base].im = s0i + s1i;
            y[base + ns].re = d0r + er;
            y[base + ns].im = d0i + ei;
            y[base + 2 * ns].re = s0r - s1r;
            y[base + 2 * ns].im = s0i - s1i;
            y[base + 3 * ns].re = d0r - er;
            y[base + 3 * ns].im = d0i - ei;
        } else {
            for (int q = 0; q < p; q++) {
                oreal sr = 0, si = 0;
                for (int r = 0; r < p; r++) {
                    ocplx w = root[(q * r) % p];
                    sr += vv[r].re * w.re - vv[r].im * w.im;
                    si += vv[r].re * w.im + vv[r].im * w.re;
                }
                y[base + q * ns].re = sr;
                y[base + q * ns].im = si;
            }
        }
    }
    free(heap);
}

void orc_fft1d(ocplx *x, int n, int sign)
{
    if (n <= 1)
        return;
    fft_plan *p = get_plan(n, sign);
    ocplx *a = x, *b = p->work;
    for (int k = 0; k < p->nst; k++) {
        stage(a, b, n, p->radix[k], p->ns[k], p->tw[k], p->root[k], sign);
        ocplx *t = a;
        a = b;
        b = t;
    }
    if (a != x)
        memcpy(x, a, sizeof(ocplx) * (size_t)n);
}

void orc_fft2d(ocplx *x, int na, int nr, int sign)
{
    for (int i = 0; i < na; i++)
        orc_fft1d(x + (size_t)i * nr, nr, sign);
    ocplx *col = (ocplx *)malloc(sizeof(ocplx) * (size_t)na);
    for (int j = 0; j < nr; j++) {
        for (int i = 0; i < na; i++)
            col[i] = x[(size_t)i * nr + j];
        orc_fft1d(col, na, sign);
        for (int i = 0; i < na; i++)
            x[(size_t)i * nr + j] = col[i];
    }
    free(col);
}
