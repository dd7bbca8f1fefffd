// Device pieces shared by the strack stages: sample access, the
// correlation-peak search with the reference scan order, the FP64
// sub-pixel refinement and the 1-D Gaussian Levenberg-Marquardt fit.
#pragma once

#include "st_common.h"

namespace sttrack {

constexpr int NOVER = 12;  // Strack/speckleTrack.c:15
constexpr int NFAST = 8;   // speckleTrack.c:16
constexpr int LA = 2;      // speckleTrack.c:18
constexpr int NGAUSS = 18; // speckleTrack.c:29
constexpr int kFine = 2 * NOVER * 2;  // 48: side of the searched oversampled box
// |C| block around a coarse peak (I, J): rows and columns I - kNb / 2 ..
// I + kNb / 2 - 1 cover the 8 x 8 refinement box and both 18-sample
// Gaussian fits at the refined peak (within 2 samples of the coarse one)
constexpr int kNb = 22;

// The cells of that block the refinement can read, stored compactly: the
// 8 x 8 box (block rows and columns 7..14) and the two fit crosses.  The
// refined peak (cI, cJ) = (I, J) + floor((f - 24) / 12), f in 0..47, lies
// within -2..+1 of the coarse one, so each cross runs along block row (or
// column) 9..12 over block columns (rows) 0..20.  168 of the 484 cells.
struct NbMap {
    short off[kNb * kNb];  // block cell -> compact index, -1 if never read
    unsigned char ri[kNb * kNb], ci[kNb * kNb];  // compact index -> block cell
    int n;
};
constexpr bool nb_needed(int i, int j)
{
    const int b0 = kNb / 2 - NFAST / 2, c0 = kNb / 2 - 2, c1 = kNb / 2 + 1;
    const int lo = c0 - NGAUSS / 2, hi = c1 + NGAUSS / 2 - 1;
    const bool box = i >= b0 && i < b0 + NFAST && j >= b0 && j < b0 + NFAST;
    const bool rowX = i >= c0 && i <= c1 && j >= lo && j <= hi;
    const bool colX = j >= c0 && j <= c1 && i >= lo && i <= hi;
    return box || rowX || colX;
}
constexpr NbMap make_nb_map()
{
    NbMap m{};
    m.n = 0;
    for (int i = 0; i < kNb; i++)
        for (int j = 0; j < kNb; j++) {
            m.off[i * kNb + j] = -1;
            if (nb_needed(i, j)) {
                m.off[i * kNb + j] = (short)m.n;
                m.ri[m.n] = (unsigned char)i;
                m.ci[m.n] = (unsigned char)j;
                m.n++;
            }
        }
    return m;
}
constexpr NbMap kNbMap = make_nb_map();
constexpr int kNbc = kNbMap.n;  // floats per patch in CpxBatch::nb
// every cell k_cpx_refine can read is stored: the box, and the crosses at
// every refined peak offset floor((f - kFine / 2) / NOVER), f in 0..kFine-1
constexpr bool nb_map_covers_reads()
{
    for (int m = 0; m < NFAST; m++)
        for (int n = 0; n < NFAST; n++)
            if (kNbMap.off[(kNb / 2 - NFAST / 2 + m) * kNb + kNb / 2 - NFAST / 2 + n] < 0)
                return false;
    const int I = NFAST / 2;  // the smallest accepted coarse peak
    for (int fi = 0; fi < kFine; fi++)
        for (int fj = 0; fj < kFine; fj += kFine / 4) {
            const int bi = kNb / 2 + (I * NOVER + fi + kFine / 2 - (NOVER * NFAST) / 2) / NOVER - I;
            const int bj = kNb / 2 + (I * NOVER + fj + kFine / 2 - (NOVER * NFAST) / 2) / NOVER - I;
            for (int k = 0; k < NGAUSS; k++)
                if (kNbMap.off[bi * kNb + bj - NGAUSS / 2 + k] < 0 || kNbMap.off[(bi - NGAUSS / 2 + k) * kNb + bj] < 0)
                    return false;
        }
    return true;
}
static_assert(kNbc == 168 && nb_map_covers_reads(), "compact peak block");
constexpr int NT = 256;

struct Images {
    const void *img1, *img2;
    int nr1, na1, nr2, na2, stype;
};

__device__ inline float2 load_sample(const void *img, int stype, int nr, int a, int r)
{
    const size_t o = (size_t)a * nr + r;
    if (stype == ST_SAMPLE_C64)
        return ((const float2 *)img)[o];
    const short2 s = ((const short2 *)img)[o];
    return make_float2((float)s.x, (float)s.y);
}

// Per-attempt result of the peak stage.
struct PeakRes {
    int I, J;         // coarse peak in the fftshifted surface, -1000 = rejected
    float raw;        // |C| at the coarse peak (findCPeak's cMax), -1 when rejected
    float gap;        // (best - runner-up) / best on the coarse surface
    float snr;        // coarse peak / mean |C| over the surface
    int fi, fj;       // oversampled-box argmax (0..47) or -1
    double fineMax;   // |Q| at the refined peak
    float fineGap;
    int lmOk;         // 1 when the 18-sample Gaussian fits are defined
    float lmX0[2];    // first abscissa of the range / azimuth samples
    float lmY[2][NGAUSS];
    int gI, gJ;       // -gauss mode: fitted peak in 1/NOVER units (-LARGEINT: failed)
    double gC;        // -gauss mode: fitted amplitude, capped at 1
};

// Sub-pixel refinement: overSampleC (speckleTrack.c:852-863, 2014-2038)
// zero-pads the 8 x 8 neighbourhood P of the coarse peak to 96 x 96, so
// the oversampled surface is Q = A P A^T with A[u][m] = D(u + 24 - 12 m)
// on the searched 48 x 48 box (:686-700; u, v = 24 .. 71 of the 96 grid)
// and D(t) = sum_{k=-4}^{3} exp(2 pi i k t / 96) = exp(-i pi t / 96) R(t),
// R(t) = sum_{j=1,3,5,7} 2 cos(pi j t / 96) real.  The phases factor out of
// |Q|: |Q[u][v]| = |sum_m R_um sum_n b_{m+n} P[m][n] R_vn| with
// R_um = R(u + 24 - 12 m) and b_k = exp(i pi k / 8), two real-by-complex
// products instead of complex-by-complex ones.
// exp(2 pi i j / 256), j = 0..255 (doppler_round's roots of unity)
extern __constant__ double2 c_root256[256];

// The Doppler shift of the acceptance test (estDopCarrier1 per window,
// dopplerEstimation.c:42-88, then speckleTrack.c's rounding):
//   k = round(W (a0 + a1) / (4 pi)), a_g = atan2(zi_g, zr_g),
// half away from zero, + W when negative, 0 when either lag sum is zero;
// found without evaluating an angle.  Each a_g is split as m_g d + r_g on
// the grid d = 2 pi / W (r_g in [0, d], m_g an integer) by a binary search
// of sign tests against the roots of unity in tab; then
// a0 + a1 = M d + (r0 + r1), and round((a0 + a1) / 2d) is (M + 1) / 2 for
// odd M, and M / 2 or M / 2 + 1 for even M as r0 + r1 stays below d or
// not (one more sign test).  Each test is exact up to the rounding of one
// product, so the result differs from rounding the atan2 sum only for
// angles within about 1e-16 of a rounding boundary.
__host__ __device__ inline int doppler_round(double zr0, double zi0, double zr1, double zi1, int W, const double2 *tab)
{
    if ((zr0 == 0.0 && zi0 == 0.0) || (zr1 == 0.0 && zi1 == 0.0))
        return 0;
    const int step = 256 / W;  // table stride of d
    double pr[2], pi[2];
    int M = 0;
    for (int g = 0; g < 2; g++) {
        const double zr = g ? zr1 : zr0, zi = g ? zi1 : zi0;
        // a in [0, pi] (zi > 0 or +0), else a = -b, b in [0, pi]
        const bool upper = zi > 0.0 || (zi == 0.0 && !__builtin_signbit(zi));
        const double yi = upper ? zi : -zi;
        // largest m in [0, W / 2] with m d <= b: b - m d in [0, pi] <=> cross(e_m, (zr, yi)) >= 0
        int lo = 0, hi = W / 2;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            const double2 e = tab[mid * step];
            if (e.x * yi - e.y * zr >= 0.0)
                lo = mid;
            else
                hi = mid - 1;
        }
        // upper: m = lo, rotate by -m d; lower: m = -(lo + 1), rotate by +(lo + 1) d
        const int j = upper ? lo : lo + 1;
        const double2 e = tab[(j * step) & 255];
        const double es = upper ? -e.y : e.y;
        double rr = zr * e.x - zi * es, ri = zr * es + zi * e.x;
        // scaled by a power of two (exact) so that the product below can
        // neither underflow nor overflow
        const int ex = ilogb(fmax(fabs(rr), fabs(ri)));
        pr[g] = ldexp(rr, -ex);
        pi[g] = ldexp(ri, -ex);
        M += upper ? lo : -(lo + 1);
    }
    int k;
    if (M & 1) {
        k = (M + 1) / 2;
    } else {
        // r0 + r1 >= d <=> cross(e_1, rho0 rho1) >= 0 (r0 + r1 in [0, 2d] with 2d < pi)
        const double qr = pr[0] * pr[1] - pi[0] * pi[1], qi = pr[0] * pi[1] + pi[0] * pr[1];
        const double2 e1 = tab[step];
        k = M / 2 + ((e1.x * qi - e1.y * qr >= 0.0) ? 1 : 0);
    }
    return k < 0 ? k + W : k;
}

// c_refR[u * 8 + m] = R_um (u < 48, m < 8); c_refB[k] = b_k (k < 15).
extern __constant__ double c_refR[kFine * NFAST];
extern __constant__ double2 c_refB[16];

// Workgroup best / runner-up of refine_fine's per-thread results (value,
// smaller rank first among equal values); every thread leaves with the
// result.  sd: 2 NTH / 64 doubles, sr: NTH / 64 ints of shared memory.
template <int NTH>
__device__ inline void block_top2_d(double &f1, double &f2, int &fr, double *sd, int *sr)
{
#pragma unroll
    for (int o = 1; o < (NTH < 64 ? NTH : 64); o <<= 1) {
        const double b1 = __shfl_xor(f1, o), b2 = __shfl_xor(f2, o);
        const int br = __shfl_xor(fr, o);
        if (b1 > f1 || (b1 == f1 && br < fr)) {
            f2 = fmax(f1, b2);
            f1 = b1;
            fr = br;
        } else
            f2 = fmax(f2, b1);
    }
    if constexpr (NTH > 64) {
        constexpr int NW = NTH / 64;
        const int t = threadIdx.x;
        __syncthreads();
        if ((t & 63) == 0) {
            sd[t >> 6] = f1;
            sd[NW + (t >> 6)] = f2;
            sr[t >> 6] = fr;
        }
        __syncthreads();
        f1 = sd[0];
        f2 = sd[NW];
        fr = sr[0];
        for (int k = 1; k < NW; k++) {
            const double b1 = sd[k], b2 = sd[NW + k];
            const int br = sr[k];
            if (b1 > f1 || (b1 == f1 && br < fr)) {
                f2 = fmax(f1, b2);
                f1 = b1;
                fr = br;
            } else
                f2 = fmax(f2, b1);
        }
        __syncthreads();
    }
}

// LDS scratch of refine_fine: P (64 doubles, written by the caller), b P
// (64 double2), T = (b P) R^T (8 x 48 double2) and, with RLDS, a copy of R
// (384 doubles; without it R is read from constant memory).
constexpr int kRefPm = 0, kRefPc = 512;
template <bool RLDS>
struct RefLayout {
    static constexpr int T = 1536, R = 1536 + 6144, BYTES = RLDS ? 1536 + 6144 + 3072 : 1536 + 6144;
};

// T = (b P) R^T from Pc = b P (8 x 8), then |Q|^2 = |R T|^2 over the box
// in TU x TU tiles; Rt: R (48 x 8) in LDS or constant memory; UM: unroll of
// the inner product loop.  Each thread leaves with the best (f1, rank fr)
// and runner-up f2 of its own tiles.
template <int NTH, int TU, int UM = 1, class Sync>
__device__ inline void refine_core(const double2 *Pc, double2 *Tq, const double *Rt, int t, Sync sync, double &f1,
                                   double &f2, int &fr)
{
    // T[m][v] = sum_n (b P)[m][n] R_vn
#pragma unroll 1
    for (int e = t; e < NFAST * kFine; e += NTH) {
        const int m = e / kFine, v = e - m * kFine;
        double re = 0, im = 0;
#pragma unroll
        for (int n = 0; n < NFAST; n++) {
            const double r = Rt[v * NFAST + n];
            const double2 pc = Pc[m * NFAST + n];
            re = fma(pc.x, r, re);
            im = fma(pc.y, r, im);
        }
        Tq[e] = make_double2(re, im);
    }
    sync();
    f1 = 0.0;
    f2 = 0.0;
    fr = 0x7fffffff;
    static_assert(kFine % TU == 0, "tile side");
    constexpr int NTS = kFine / TU;
#pragma unroll 1
    for (int ti = t; ti < NTS * NTS; ti += NTH) {
        const int u0 = (ti / NTS) * TU, v0 = (ti % NTS) * TU;
        double ar[TU][TU], ai[TU][TU];
#pragma unroll
        for (int a = 0; a < TU; a++)
#pragma unroll
            for (int b = 0; b < TU; b++)
                ar[a][b] = ai[a][b] = 0.0;
#pragma unroll(UM)
        for (int m = 0; m < NFAST; m++) {
            double r[TU];
            double2 tv[TU];
#pragma unroll
            for (int a = 0; a < TU; a++)
                r[a] = Rt[(u0 + a) * NFAST + m];
#pragma unroll
            for (int b = 0; b < TU; b++)
                tv[b] = Tq[m * kFine + v0 + b];
#pragma unroll
            for (int a = 0; a < TU; a++)
#pragma unroll
                for (int b = 0; b < TU; b++) {
                    ar[a][b] = fma(r[a], tv[b].x, ar[a][b]);
                    ai[a][b] = fma(r[a], tv[b].y, ai[a][b]);
                }
        }
#pragma unroll
        for (int a = 0; a < TU; a++)
#pragma unroll
            for (int b = 0; b < TU; b++) {
                const double q = ar[a][b] * ar[a][b] + ai[a][b] * ai[a][b];  // |Q|^2: same order as |Q|
                const int e = (u0 + a) * kFine + v0 + b;
                if (q > f1 || (q == f1 && e < fr)) {
                    f2 = fmax(f2, f1);
                    f1 = q;
                    fr = e;
                } else
                    f2 = fmax(f2, q);
            }
    }
}

// refine_core's product without the LDS tiles: the calling lane owns
// column v < kFine of one attempt's search box.  It forms T[m][v] (m < 8)
// in registers from pc(e) = (b P)[e / 8][e % 8] and its own row of R, then
// walks the rows u of the box with R[u][m] uniform across the workgroup
// (scalar loads), keeping its best and runner-up.  The products and their
// order are refine_core's, so each |Q|^2 is the same number.
template <class PC>
__device__ inline void refine_cols(PC pc, int v, double &f1, double &f2, int &fr)
{
    f1 = 0.0;
    f2 = 0.0;
    fr = 0x7fffffff;
    double tr[NFAST], ti[NFAST];
    {
        double rv[NFAST];
#pragma unroll
        for (int n = 0; n < NFAST; n++)
            rv[n] = c_refR[v * NFAST + n];
#pragma unroll
        for (int m = 0; m < NFAST; m++) {
            double re = 0, im = 0;
#pragma unroll
            for (int n = 0; n < NFAST; n++) {
                const double2 c = pc(m * NFAST + n);
                re = fma(c.x, rv[n], re);
                im = fma(c.y, rv[n], im);
            }
            tr[m] = re;
            ti[m] = im;
        }
    }
#pragma unroll 2
    for (int u = 0; u < kFine; u++) {
        double ar = 0.0, ai = 0.0;
#pragma unroll
        for (int m = 0; m < NFAST; m++) {
            const double r = c_refR[u * NFAST + m];
            ar = fma(r, tr[m], ar);
            ai = fma(r, ti[m], ai);
        }
        const double q = ar * ar + ai * ai;  // |Q|^2: refine_core's expression
        // ranks u kFine + v rise with u, so a later equal value never wins;
        // the runner-up is the larger of the old one and the smaller of
        // (best, q) (all values are finite and non-negative)
        fr = q > f1 ? u * kFine + v : fr;
        f2 = __builtin_fmax(f2, __builtin_fmin(f1, q));
        f1 = __builtin_fmax(f1, q);
    }
}

// Best / runner-up merge: the larger value wins, the smaller rank among
// equal values; the runner-up is the larger of everything else.
__device__ inline void top2_merge(double &f1, double &f2, int &fr, double b1, double b2, int br)
{
    if (b1 > f1 || (b1 == f1 && br < fr)) {
        f2 = fmax(f1, b2);
        f1 = b1;
        fr = br;
    } else {
        f2 = fmax(f2, b1);
    }
}

// The whole refinement for a caller-written P at scr + kRefPm (layout
// RefLayout<RLDS>): every thread leaves with its own best (value f1, rank
// fr, smaller rank first among equal values) and runner-up value f2, for
// the caller's reduction.  All NTH threads call it, t = 0 .. NTH - 1;
// sync() is the caller's barrier.  TU x TU tiles ((48 / TU)^2 of them;
// TU = 2 for kernels short of registers).
template <int NTH, bool RLDS, int TU = 3, class Sync>
__device__ inline void refine_fine(unsigned char *scr, int t, Sync sync, double &f1, double &f2, int &fr)
{
    using LY = RefLayout<RLDS>;
    const double *Pm = (const double *)(scr + kRefPm);
    double2 *Pc = (double2 *)(scr + kRefPc);
    double2 *Tq = (double2 *)(scr + LY::T);
    const double *Rt = c_refR;
    for (int e = t; e < NFAST * NFAST; e += NTH) {
        const double2 b = c_refB[e / NFAST + e % NFAST];
        Pc[e] = make_double2(b.x * Pm[e], b.y * Pm[e]);
    }
    if constexpr (RLDS) {
        double *Rl = (double *)(scr + LY::R);
        for (int e = t; e < kFine * NFAST; e += NTH)
            Rl[e] = c_refR[e];
        Rt = Rl;
    }
    sync();
    refine_core<NTH, TU>(Pc, Tq, Rt, t, sync, f1, f2, fr);
}

#pragma clang fp contract(off)
// 1-D Gaussian LM fit (gaussPeak1D / mrqminMod / mrqcofMod / myGauss1D,
// speckleTrack.c:2137-2202, 2309-2465), all in float like the reference.
__device__ inline void lm_gauss(float x, const float a[3], float *y, float dyda[3])
{
    float sig = a[2];
    float t = x - a[1];
    float argx = (float)(0.5 * (double)t * (double)t / (double)(sig * sig));
    float ex = (float)exp(-(double)argx);
    *y = a[0] * ex;
    dyda[0] = ex;
    dyda[1] = *y * t / (sig * sig);
    dyda[2] = *y * t * t / (sig * sig * sig);
}

__device__ inline void lm_coef(const float *x, const float *y, const float a[3], float alpha[3][3], float beta[3],
                               float *chisq)
{
    for (int j = 0; j < 3; j++) {
        for (int k = 0; k <= j; k++)
            alpha[j][k] = 0.0f;
        beta[j] = 0.0f;
    }
    *chisq = 0.0f;
    const float sg = 0.1f;
    const float sig2i = (float)(1.0 / (double)(sg * sg));
    for (int i = 0; i < NGAUSS; i++) {
        float ymod, dyda[3];
        lm_gauss(x[i], a, &ymod, dyda);
        float dy = y[i] - ymod;
        for (int l = 0; l < 3; l++) {
            float wt = dyda[l] * sig2i;
            for (int m = 0; m <= l; m++)
                alpha[l][m] += wt * dyda[m];
            beta[l] += dy * wt;
        }
        *chisq += dy * dy * sig2i;
    }
    for (int j = 1; j < 3; j++)
        for (int k = 0; k < j; k++)
            alpha[k][j] = alpha[j][k];
}

// Gauss-Jordan with full pivoting on a 3x3 system; 0 when singular.
__device__ inline int solve3(float m[3][3], float b[3])
{
    int used[3] = {0, 0, 0};
    for (int it = 0; it < 3; it++) {
        float big = 0;
        int r = -1, c = -1;
        for (int j = 0; j < 3; j++) {
            if (used[j])
                continue;
            for (int k = 0; k < 3; k++)
                if (!used[k] && fabsf(m[j][k]) >= big) {
                    big = fabsf(m[j][k]);
                    r = j;
                    c = k;
                }
        }
        if (r < 0)
            return 0;
        used[c] = 1;
        if (r != c) {
            for (int l = 0; l < 3; l++) {
                float t = m[r][l];
                m[r][l] = m[c][l];
                m[c][l] = t;
            }
            float t = b[r];
            b[r] = b[c];
            b[c] = t;
        }
        if (m[c][c] == 0.0f)
            return 0;
        float pinv = (float)(1.0 / (double)m[c][c]);
        m[c][c] = 1.0f;
        for (int l = 0; l < 3; l++)
            m[c][l] *= pinv;
        b[c] *= pinv;
        for (int ll = 0; ll < 3; ll++)
            if (ll != c) {
                float d = m[ll][c];
                m[ll][c] = 0.0f;
                for (int l = 0; l < 3; l++)
                    m[ll][l] -= m[c][l] * d;
                b[ll] -= b[c] * d;
            }
    }
    return 1;
}

// gaussPeak1D: returns sigma = width / 2 of the fit through y[0..17] at
// abscissae x0, x0+1, ...
__device__ inline float gauss_peak_1d(const float *yin, float x0)
{
    float x[NGAUSS], y[NGAUSS];
    float maxC = 0.0f;
    for (int i = 0; i < NGAUSS; i++) {
        x[i] = x0 + (float)i;
        y[i] = yin[i];
        maxC = y[i] > maxC ? y[i] : maxC;
    }
    for (int i = 0; i < NGAUSS; i++)
        y[i] /= maxC;
    float a[3] = {1.0f, x[NGAUSS / 2], 2.0f};
    float alpha[3][3], beta[3], ochisq, chisq;
    float lambda = 0.001f;
    lm_coef(x, y, a, alpha, beta, &chisq);
    ochisq = chisq;
    float lasti = 2e9f;
    for (int it = 0; it < 5; it++) {
        float covar[3][3], da[3];
        for (int j = 0; j < 3; j++) {
            for (int k = 0; k < 3; k++)
                covar[j][k] = alpha[j][k];
            covar[j][j] = (float)((double)alpha[j][j] * (1.0 + (double)lambda));
            da[j] = beta[j];
        }
        if (!solve3(covar, da))
            break;
        float atry[3];
        for (int l = 0; l < 3; l++)
            atry[l] = a[l] + da[l];
        float nalpha[3][3], nbeta[3];
        lm_coef(x, y, atry, nalpha, nbeta, &chisq);
        if (chisq < ochisq) {
            lambda = (float)((double)lambda * 0.1);
            ochisq = chisq;
            for (int j = 0; j < 3; j++) {
                for (int k = 0; k < 3; k++)
                    alpha[j][k] = nalpha[j][k];
                beta[j] = nbeta[j];
            }
            for (int l = 0; l < 3; l++)
                a[l] = atry[l];
        } else {
            lambda = (float)((double)lambda * 10.0);
        }
        if (fabs((double)(a[1] - lasti)) < 0.001)
            break;
        lasti = a[1];
    }
    return (float)(a[2] * 0.5f);
}
// 2-D Gaussian LM fit of the -gauss peak mode (cmpTrackGauss / mrqminMod /
// mrqcofMod / myGauss, speckleTrack.c:2205-2303, 2309-2443), in float like
// the reference, one thread per fit.  Parameters a = {amplitude, mu_i,
// sigma_i, mu_j, sigma_j}; data: the 18 x 18 magnitudes around the coarse
// peak times `scale`, weights 1.
__device__ inline void lm_gauss2d(float xx, float yy, const float a[5], float *y, float dyda[5])
{
    const float dx = xx - a[1], dy = yy - a[3];
    const float argx = 0.5f * (dx * dx) / (a[2] * a[2]);
    const float argy = 0.5f * (dy * dy) / (a[4] * a[4]);
    const float ex = (float)exp((double)(-argx - argy));
    *y = a[0] * ex;
    dyda[0] = ex;
    dyda[1] = (*y) * (xx - a[1]) / (a[2] * a[2]);
    dyda[2] = *y * (xx - a[1]) * (xx - a[1]) / (a[2] * a[2] * a[2]);
    dyda[3] = *y * (yy - a[3]) / (a[4] * a[4]);
    dyda[4] = *y * (yy - a[3]) * (yy - a[3]) / (a[4] * a[4] * a[4]);
}

__device__ inline void lm2d_coef(const float *M, int W, int i0, int j0, double scale, const float a[5],
                                 float alpha[5][5], float beta[5], float *chisq)
{
    for (int j = 0; j < 5; j++) {
        for (int k = 0; k <= j; k++)
            alpha[j][k] = 0.0f;
        beta[j] = 0.0f;
    }
    *chisq = 0.0f;
    for (int i = 0; i < NGAUSS; i++)
        for (int j = 0; j < NGAUSS; j++) {
            const float g = (float)((double)M[(i0 + i) * W + j0 + j] * scale);
            float ymod, dyda[5];
            lm_gauss2d((float)(i0 + i), (float)(j0 + j), a, &ymod, dyda);
            const float dy = g - ymod;
            for (int l = 0; l < 5; l++) {
                const float wt = dyda[l];
                for (int m = 0; m <= l; m++)
                    alpha[l][m] += wt * dyda[m];
                beta[l] += dy * wt;
            }
            *chisq += dy * dy;
        }
    for (int j = 1; j < 5; j++)
        for (int k = 0; k < j; k++)
            alpha[k][j] = alpha[j][k];
}

// Gauss-Jordan with full pivoting on a 5x5 system; 0 when singular.
__device__ inline int solve5(float m[5][5], float b[5])
{
    int used[5] = {0, 0, 0, 0, 0};
    for (int it = 0; it < 5; it++) {
        float big = 0;
        int r = -1, c = -1;
        for (int j = 0; j < 5; j++) {
            if (used[j])
                continue;
            for (int k = 0; k < 5; k++)
                if (!used[k] && fabsf(m[j][k]) >= big) {
                    big = fabsf(m[j][k]);
                    r = j;
                    c = k;
                }
        }
        if (r < 0)
            return 0;
        used[c] = 1;
        if (r != c) {
            for (int l = 0; l < 5; l++) {
                float t = m[r][l];
                m[r][l] = m[c][l];
                m[c][l] = t;
            }
            float t = b[r];
            b[r] = b[c];
            b[c] = t;
        }
        if (m[c][c] == 0.0f)
            return 0;
        float pinv = (float)(1.0 / (double)m[c][c]);
        m[c][c] = 1.0f;
        for (int l = 0; l < 5; l++)
            m[c][l] *= pinv;
        b[c] *= pinv;
        for (int ll = 0; ll < 5; ll++)
            if (ll != c) {
                float d = m[ll][c];
                m[ll][c] = 0.0f;
                for (int l = 0; l < 5; l++)
                    m[ll][l] -= m[c][l] * d;
                b[ll] -= b[c] * d;
            }
    }
    return 1;
}

// cmpTrackGauss on the fftshifted magnitude M (row stride W) around the
// coarse peak (I, J) with findCPeak's raw maximum cMax; wR / wA: the window
// of the surface; scale = hanningCorrection / (sqrt(p1 p2) (wR wA)^2).
// Returns the peak in 1/NOVER units and the fitted amplitude capped at 1, or
// -LARGEINT / 0 near the edge or when a width exceeds 2000.
__device__ inline void gauss_fit_2d(const float *M, int W, int I, int J, double cMax, int wR, int wA, double scale,
                                    int *oI, int *oJ, double *oC)
{
    constexpr int ng = NGAUSS;
    if ((I - ng / 2) < 0 || (I + ng / 2) >= (wA * 2 - 1) || (J - ng / 2) < 0 || (J + ng / 2) >= (wR * 2 - 1)) {
        *oI = *oJ = -2000000000;
        *oC = 0;
        return;
    }
    const int i0 = I - ng / 2, j0 = J - ng / 2;
    float a[5] = {(float)(cMax * scale), (float)I, 3.0f, (float)J, 3.0f};
    float alpha[5][5], beta[5], ochisq, chisq;
    float lambda = 0.001f;
    lm2d_coef(M, W, i0, j0, scale, a, alpha, beta, &chisq);
    ochisq = chisq;
    float lasti = 2e9f, lastj = 2.0e9f;
    for (int it = 0; it < 8; it++) {
        float covar[5][5], da[5];
        for (int j = 0; j < 5; j++) {
            for (int k = 0; k < 5; k++)
                covar[j][k] = alpha[j][k];
            covar[j][j] = (float)((double)alpha[j][j] * (1.0 + (double)lambda));
            da[j] = beta[j];
        }
        if (!solve5(covar, da))
            break;
        float atry[5];
        for (int l = 0; l < 5; l++)
            atry[l] = a[l] + da[l];
        float nalpha[5][5], nbeta[5];
        lm2d_coef(M, W, i0, j0, scale, atry, nalpha, nbeta, &chisq);
        if (chisq < ochisq) {
            lambda = (float)((double)lambda * 0.1);
            ochisq = chisq;
            for (int j = 0; j < 5; j++) {
                for (int k = 0; k < 5; k++)
                    alpha[j][k] = nalpha[j][k];
                beta[j] = nbeta[j];
            }
            for (int l = 0; l < 5; l++)
                a[l] = atry[l];
        } else {
            lambda = (float)((double)lambda * 10.0);
        }
        if (fabs((double)(a[1] - lasti)) < 0.001 && fabs((double)(a[3] - lastj)) < 0.001)
            break;
        lasti = a[1];
        lastj = a[3];
    }
    *oI = (int)round((double)(a[1] * (float)NOVER));
    *oJ = (int)round((double)(a[3] * (float)NOVER));
    *oC = (double)a[0] < 1.0 ? (double)a[0] : 1.0;
    if ((0.5 * (double)a[2]) > 1000.0 || (0.5 * (double)a[4]) > 1000.0) {
        *oI = *oJ = -2000000000;
        *oC = 0;
    }
}
#pragma clang fp contract(on)

// Doppler estimate of a w x w patch pair (estDopCarrier1,
// Strack/dopplerEstimation.c:25-88): lag-1 azimuth autocorrelation after
// complex mean removal, in double.
template <class Get>
__device__ inline int doppler_shift(Get get, int wA, int wR, int nAz, int wAnorm, double *sh)
{
    const int N = wA * wR;
    double m[4] = {0, 0, 0, 0};
    for (int e = threadIdx.x; e < N; e += NT) {
        float2 a, b;
        get(e, a, b);
        m[0] += a.x;
        m[1] += a.y;
        m[2] += b.x;
        m[3] += b.y;
    }
    for (int k = 0; k < 4; k++)
        m[k] = block_sum_d<NT>(m[k], sh) / ((double)wA * (double)wR);
    double z[4] = {0, 0, 0, 0};
    for (int e = threadIdx.x; e < N - wR; e += NT) {
        float2 a0, b0, a1, b1;
        get(e, a0, b0);
        get(e + wR, a1, b1);
        double ar = a1.x - m[0], ai = a1.y - m[1], br = a0.x - m[0], bi = a0.y - m[1];
        z[0] += ar * br + ai * bi;
        z[1] += ai * br - ar * bi;
        ar = b1.x - m[2], ai = b1.y - m[3], br = b0.x - m[2], bi = b0.y - m[3];
        z[2] += ar * br + ai * bi;
        z[3] += ai * br - ar * bi;
    }
    for (int k = 0; k < 4; k++)
        z[k] = block_sum_d<NT>(z[k], sh);
    double ang1 = (z[0] == 0.0 && z[1] == 0.0) ? NAN : atan2(z[1], z[0]);
    double ang2 = (z[2] == 0.0 && z[3] == 0.0) ? NAN : atan2(z[3], z[2]);
    double v = (double)(nAz / 2) * (ang1 + ang2) * .5 / 3.14159265358979323846;
    int shift = isnan(v) ? 0 : ((v >= 0.0) ? (int)(v + 0.5) : (int)(v - 0.5));
    if (shift < 0)
        shift += nAz;
    return (int)((float)shift * (float)wAnorm / (float)nAz);
}

}  // namespace sttrack
