// strack on the GPU: the decision cascade of speckleTrack
// (Strack/speckleTrack.c:159-346) run as batched stages over work lists.
//
//   geometry   findImage2Pos / boundsCheck / maskValue for every point
//   complex    getPatches -> cmpPSWithPad -> cmpTrackFast -> gaussPeak1D x2
//              -> finalizeCmpxMatch                      (speckleTrack.c:349-394)
//   amp small  getAmpPatches -> ampMatch -> ampTrackFast  (speckleTrack.c:398-597)
//   amp large  the same with the window doubled (LA = 2)
//   compose    offsets, corr, type, nFail                 (speckleTrack.c:301-323)
//
// Each attempt stage works on a compacted list of the points that reach it,
// in batches sized to the scratch budget.  Per-patch FFTs go through the
// batched LDS FFT engine; the peak stage fuses the fftshifted magnitude,
// the scan-order argmax, the FP64 8x8 -> 96x96 refinement and the sample
// extraction for the Gaussian fits.
#include <math.h>
#include <string.h>

#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

#include "st_ctx.h"
#include "st_fftlds.h"
#include "st_track_common.h"
#include "st_fftreg.h"

using namespace sttrack;

namespace sttrack {
__constant__ double c_refR[kFine * NFAST];
__constant__ double2 c_refB[16];
__constant__ double2 c_root256[256];
}
namespace fftreg {
__constant__ float2 c_tw128[128];
}  // namespace fftreg

// global-address-space views of pointers held in parameter blocks (loads
// through them are global loads, not flat ones that also count against
// the LDS wait counter)
typedef __attribute__((address_space(1))) const char gchar;
typedef __attribute__((address_space(1))) const float gfloat;
typedef __attribute__((address_space(1))) const unsigned long long gu64;
typedef __attribute__((address_space(1))) const unsigned gu32;

namespace {

// Kept bins of one point's interferogram spectrum (phaseCorrectInt): the
// first kIntInline sit in the per-point record, a point with more keeps all
// of them in a slot of ps^2 entries of the overflow area.
struct IntBin {
    int kp, lp;  // position in the Pa x Pr zero pad (azimuth, range)
    double v;    // kept value: the real part of the bin
};
constexpr int kIntInline = 16;
struct IntSet {
    int n;    // kept non-zero bins (0: no correction)
    int ovf;  // overflow slot when n > kIntInline
    IntBin b[kIntInline];
};

struct KPar {
    int nR, nA, rStart, aStart, deltaR, deltaA, wR, wA, wRa, wAa, sf;
    int noComplex, hanning, maxTries, checkAz, gauss;
    float azThresh;
    double rPoly[3], aPoly[3];
    const uint8_t *mask;
    int maskNr, maskNa, maskR0, maskA0, maskNrl, maskNal;
    int baseFlag;
    double Bn, Bp, dBn, dBp, dBnQ, dBpQ, lambda;
    int nSlpA1;
    const double *sinT, *cosT, *overR;
    // interferogram phase correction: intSet is indexed by list position
    int intOn, intNr, intNa, intNrl, intNal, intPs, intPa, intPr;
    const float *intf;
    const IntSet *intSet;
    const IntBin *intOvf;
    const float *hann;
    Images im;
};

// Per-point state, one entry per grid point (row-major nA x nR).
struct Pts {
    int *r1, *a1, *r2, *a2;
    uint8_t *maskv, *haveData, *good, *type, *azValid;
    int *azShift, *azUse;
    float *corr, *rShift1, *aShift1, *azd;
    int *peakI, *peakJ, *fineI, *fineJ; // 3 x npts
    double *cmax;
    float *sigRg, *sigAz, *peakGap, *fineGap, *snr;
    int *counts;  // nComplex, nAmp, nAmpL, nFail, list sizes...
};

enum { CNT_COMPLEX = 0, CNT_AMP, CNT_AMPL, CNT_FAIL, CNT_LIST, CNT_CORR0, CNT_CORR1, CNT_CORR2, CNT_N };

__device__ inline bool in_bounds(int r1, int a1, int wR, int wA, int sR, int sA)
{
    return r1 > 0 && (r1 + wR) < sR && a1 > 0 && (a1 + wA) < sA;
}

// ---------------------------------------------------------------------
// Interferogram phase correction (phaseCorrectInt, speckleTrack.c:909-1000;
// patch size speckleTrack.c:1576-1580, osF = 1).  The ps x ps interferogram
// patch at ((a1 + wA/2)/nal - ps/2, (r1 + wR/2)/nrl - ps/2) has spectrum X;
// the bins whose power reaches 0.99 of the maximum keep v = Re X (the
// reference clears the imaginary part of every bin) and move to their place
// in an Pa x Pr zero pad (Pa = ps nal, Pr = ps nrl; with odd ps the last row
// and column are not copied).  y, the unnormalised inverse of that pad, is a
// sum over the kept bins and is evaluated only on the w x w window that
// starts at (Pa/2 - wA/2, Pr/2 - wR/2); patch 1 is multiplied there by
// conj(y) / |y|, or left as it is where |Re y| and |Im y| are both below
// 1e-6.  Spectrum and correction run in FP64.
constexpr int kIntThreads = 256;
constexpr int kIntMaxPs = 64;    // k_int_bins: the patch and its transforms in LDS
constexpr int kIntMaxPsG = 128;  // k_int_bins_g: in a global workspace per workgroup
constexpr int kIntPer = kIntMaxPs * kIntMaxPs / kIntThreads;

__device__ inline void int_correct(const KPar &P, int it, int i, int j, double &cr, double &ci)
{
    const IntSet *I = P.intSet + it;
    const int nb = I->n;
    if (nb == 0)
        return;
    const IntBin *B = nb <= kIntInline ? I->b : P.intOvf + (size_t)I->ovf * P.intPs * P.intPs;
    const int m = P.intPa / 2 - P.wA / 2 + i, n = P.intPr / 2 - P.wR / 2 + j;
    double yr = 0, yi = 0;
    for (int q = 0; q < nb; q++) {
        const IntBin b = B[q];
        const double f = (double)((b.kp * m) % P.intPa) / P.intPa + (double)((b.lp * n) % P.intPr) / P.intPr;
        double sn, cs;
        sincospi(2.0 * f, &sn, &cs);
        yr += b.v * cs;
        yi += b.v * sn;
    }
    const double pr = yr, pim = -yi;
    if (fabs(pr) < 1.0e-6 && fabs(pim) < 1.0e-6)
        return;
    const double scale = 1.0 / sqrt(pr * pr + pim * pim);
    const double nr = (cr * pr - ci * pim) * scale, ni = (cr * pim + ci * pr) * scale;
    cr = nr;
    ci = ni;
}

// Points with at most kIntTab kept bins (nearly all: a fringe has one
// spectral peak) use the separable form of each bin's term,
// exp(2 pi i (kp m / Pa + lp n / Pr)) = A[m] B[n]: the workgroup tabulates A
// over the window's rows and B over its columns once per point, and an
// element costs a few FP64 multiplies instead of a sincos.  Other points
// (and windows wider than TW) evaluate every term directly.
constexpr int kIntTab = 2;

template <int TW>
struct IntTab {
    double2 a[kIntTab][TW], b[kIntTab][TW];
    double v[kIntTab];
    int nb;  // kept bins, or -1: direct evaluation
};

// All threads of the workgroup call it (one barrier inside).
template <int TW>
__device__ inline void int_tables(const KPar &P, int it, int w, IntTab<TW> &T, int tid, int nth)
{
    const IntSet *I = P.intSet + it;
    const int nb = I->n;
    if (nb <= kIntTab && w <= TW) {
        const int m0 = P.intPa / 2 - P.wA / 2, n0 = P.intPr / 2 - P.wR / 2;
        for (int e = tid; e < nb * w; e += nth) {
            const int q = e / w, k = e - (e / w) * w;
            const IntBin b = I->b[q];
            double sn, cs;
            sincospi(2.0 * (double)((b.kp * (m0 + k)) % P.intPa) / P.intPa, &sn, &cs);
            T.a[q][k] = make_double2(cs, sn);
            sincospi(2.0 * (double)((b.lp * (n0 + k)) % P.intPr) / P.intPr, &sn, &cs);
            T.b[q][k] = make_double2(cs, sn);
        }
        if (tid < nb)
            T.v[tid] = I->b[tid].v;
    }
    if (tid == 0)
        T.nb = (nb <= kIntTab && w <= TW) ? nb : -1;
    __syncthreads();
}

template <int TW>
__device__ inline void int_apply(const KPar &P, const IntTab<TW> &T, int it, int i, int j, double &cr, double &ci)
{
    const int nb = T.nb;
    if (nb == 0)
        return;
    if (nb < 0) {
        int_correct(P, it, i, j, cr, ci);
        return;
    }
    double yr = 0, yi = 0;
    for (int q = 0; q < nb; q++) {
        const double2 a = T.a[q][i], b = T.b[q][j];
        yr += T.v[q] * (a.x * b.x - a.y * b.y);
        yi += T.v[q] * (a.x * b.y + a.y * b.x);
    }
    const double pr = yr, pim = -yi;
    if (fabs(pr) < 1.0e-6 && fabs(pim) < 1.0e-6)
        return;
    const double scale = 1.0 / sqrt(pr * pr + pim * pim);
    const double nr = (cr * pr - ci * pim) * scale, ni = (cr * pim + ci * pr) * scale;
    cr = nr;
    ci = ni;
}

// One workgroup per listed point.  Pass 0 writes the record (n, the first
// kIntInline bins, an overflow slot); pass 1 runs only for points with more
// bins and writes all of them to their slot.
__global__ void __launch_bounds__(kIntThreads) k_int_bins(KPar P, Pts S, const int *list, IntSet *set, IntBin *ovf,
                                                          int *ovfCount, int pass)
{
    extern __shared__ double2 ldsd[];
    __shared__ double red[kIntThreads / 64];
    __shared__ int nKept;
    const int it = blockIdx.x;
    IntSet *me = set + it;
    if (pass == 1 && me->n <= kIntInline)
        return;
    const int ps = P.intPs, ps2 = ps * ps, h = ps / 2;
    const int pt = list[it];
    const int ia1 = (S.a1[pt] + P.wA / 2) / P.intNal - ps / 2;
    const int ir1 = (S.r1[pt] + P.wR / 2) / P.intNrl - ps / 2;
    if (!in_bounds(ir1, ia1, ps, ps, P.intNr, P.intNa)) {
        if (threadIdx.x == 0) {
            me->n = 0;
            me->ovf = 0;
        }
        return;
    }
    double2 *T = ldsd, *tw = ldsd + ps2;
    float2 *xin = (float2 *)(ldsd + ps2 + ps);
    const float2 *src = (const float2 *)P.intf;
    for (int e = threadIdx.x; e < ps2; e += kIntThreads) {
        const int i = e / ps, j = e - (e / ps) * ps;
        xin[e] = src[(size_t)(ia1 + i) * P.intNr + ir1 + j];
    }
    for (int t = threadIdx.x; t < ps; t += kIntThreads) {
        double sn, cs;
        sincospi(-2.0 * t / ps, &sn, &cs);
        tw[t] = make_double2(cs, sn);
    }
    if (threadIdx.x == 0)
        nKept = 0;
    __syncthreads();
    // forward DFT along range: T[i][l] = sum_j x[i][j] exp(-2 pi i j l / ps)
    for (int e = threadIdx.x; e < ps2; e += kIntThreads) {
        const int i = e / ps, l = e - (e / ps) * ps;
        double sr = 0, si = 0;
        int idx = 0;
        for (int j = 0; j < ps; j++) {
            const float2 x = xin[i * ps + j];
            const double2 w = tw[idx];
            sr += (double)x.x * w.x - (double)x.y * w.y;
            si += (double)x.x * w.y + (double)x.y * w.x;
            idx += l;
            idx -= idx >= ps ? ps : 0;
        }
        T[e] = make_double2(sr, si);
    }
    __syncthreads();
    // along azimuth, kept in registers: X[k][l] = sum_i T[i][l] exp(-2 pi i i k / ps)
    double xr[kIntPer], pw[kIntPer];
    double mx = 0;
#pragma unroll
    for (int q = 0; q < kIntPer; q++) {
        const int e = threadIdx.x + q * kIntThreads;
        xr[q] = 0;
        pw[q] = -1;
        if (e < ps2) {
            const int k = e / ps, l = e - (e / ps) * ps;
            double sr = 0, si = 0;
            int idx = 0;
            for (int i = 0; i < ps; i++) {
                const double2 t = T[i * ps + l], w = tw[idx];
                sr += t.x * w.x - t.y * w.y;
                si += t.x * w.y + t.y * w.x;
                idx += k;
                idx -= idx >= ps ? ps : 0;
            }
            xr[q] = sr;
            pw[q] = sr * sr + si * si;
            mx = fmax(mx, pw[q]);
        }
    }
    mx = block_max_d<kIntThreads>(mx, red);
    const double thr = 0.99 * mx;
    IntBin *dst = pass == 0 ? me->b : ovf + (size_t)me->ovf * ps2;
    const int cap = pass == 0 ? kIntInline : ps2;
#pragma unroll
    for (int q = 0; q < kIntPer; q++) {
        const int e = threadIdx.x + q * kIntThreads;
        if (e < ps2 && !(pw[q] < thr) && xr[q] != 0.0) {
            const int k = e / ps, l = e - (e / ps) * ps;
            if (k < 2 * h && l < 2 * h) {
                const int pos = atomicAdd(&nKept, 1);
                if (pos < cap)
                    dst[pos] = IntBin{k < h ? k : P.intPa - 2 * h + k, l < h ? l : P.intPr - 2 * h + l, xr[q]};
            }
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && pass == 0) {
        me->n = nKept;
        me->ovf = nKept > kIntInline ? atomicAdd(ovfCount, 1) : 0;
    }
}

size_t int_lds_bytes(int ps) { return (size_t)ps * ps * sizeof(double2) + ps * sizeof(double2) + (size_t)ps * ps * sizeof(float2); }

// k_int_bins for patches above kIntMaxPs samples a side (up to kIntMaxPsG,
// e.g. w = 128 without looks): the same transforms, sums and selection, with
// the patch, the range transform and the spectrum in a workspace of
// int_ws_bytes(ps) per workgroup (they do not fit the LDS), the workgroups
// walking the attempts as a persistent grid.  Each thread reads back only
// the spectrum cells it wrote; the other cells cross threads after a barrier.
// (each array starts on a 256-byte boundary: ps may be odd)
__host__ __device__ inline size_t int_ws_al(size_t b) { return (b + 255) & ~(size_t)255; }
__host__ __device__ inline size_t int_ws_bytes(int ps)
{
    const size_t n = (size_t)ps * ps;
    return int_ws_al(n * sizeof(float2)) + int_ws_al(n * sizeof(double2)) + 2 * int_ws_al(n * sizeof(double));
}

__global__ void __launch_bounds__(kIntThreads) k_int_bins_g(KPar P, Pts S, const int *list, int count, IntSet *set,
                                                            IntBin *ovf, int *ovfCount, int pass, unsigned char *ws)
{
    __shared__ double2 tw[kIntMaxPsG];
    __shared__ double red[kIntThreads / 64];
    __shared__ int nKept;
    const int ps = P.intPs, ps2 = ps * ps, h = ps / 2;
    unsigned char *mine = ws + (size_t)blockIdx.x * int_ws_bytes(ps);
    float2 *xin = (float2 *)mine;
    double2 *T = (double2 *)(mine + int_ws_al((size_t)ps2 * sizeof(float2)));
    double *Xr = (double *)((unsigned char *)T + int_ws_al((size_t)ps2 * sizeof(double2)));
    double *Pw = (double *)((unsigned char *)Xr + int_ws_al((size_t)ps2 * sizeof(double)));
    for (int t = threadIdx.x; t < ps; t += kIntThreads) {
        double sn, cs;
        sincospi(-2.0 * t / ps, &sn, &cs);
        tw[t] = make_double2(cs, sn);
    }
    const float2 *src = (const float2 *)P.intf;
    for (int it = blockIdx.x; it < count; it += gridDim.x) {
        IntSet *me = set + it;
        if (pass == 1 && me->n <= kIntInline)
            continue;
        const int pt = list[it];
        const int ia1 = (S.a1[pt] + P.wA / 2) / P.intNal - ps / 2;
        const int ir1 = (S.r1[pt] + P.wR / 2) / P.intNrl - ps / 2;
        if (!in_bounds(ir1, ia1, ps, ps, P.intNr, P.intNa)) {
            if (threadIdx.x == 0) {
                me->n = 0;
                me->ovf = 0;
            }
            continue;
        }
        for (int e = threadIdx.x; e < ps2; e += kIntThreads) {
            const int i = e / ps, j = e - (e / ps) * ps;
            xin[e] = src[(size_t)(ia1 + i) * P.intNr + ir1 + j];
        }
        if (threadIdx.x == 0)
            nKept = 0;
        __syncthreads();
        for (int e = threadIdx.x; e < ps2; e += kIntThreads) {
            const int i = e / ps, l = e - (e / ps) * ps;
            double sr = 0, si = 0;
            int idx = 0;
            for (int j = 0; j < ps; j++) {
                const float2 x = xin[i * ps + j];
                const double2 w = tw[idx];
                sr += (double)x.x * w.x - (double)x.y * w.y;
                si += (double)x.x * w.y + (double)x.y * w.x;
                idx += l;
                idx -= idx >= ps ? ps : 0;
            }
            T[e] = make_double2(sr, si);
        }
        __syncthreads();
        double mx = 0;
        for (int e = threadIdx.x; e < ps2; e += kIntThreads) {
            const int k = e / ps, l = e - (e / ps) * ps;
            double sr = 0, si = 0;
            int idx = 0;
            for (int i = 0; i < ps; i++) {
                const double2 t = T[i * ps + l], w = tw[idx];
                sr += t.x * w.x - t.y * w.y;
                si += t.x * w.y + t.y * w.x;
                idx += k;
                idx -= idx >= ps ? ps : 0;
            }
            Xr[e] = sr;
            Pw[e] = sr * sr + si * si;
            mx = fmax(mx, Pw[e]);
        }
        mx = block_max_d<kIntThreads>(mx, red);
        const double thr = 0.99 * mx;
        IntBin *dst = pass == 0 ? me->b : ovf + (size_t)me->ovf * ps2;
        const int cap = pass == 0 ? kIntInline : ps2;
        for (int e = threadIdx.x; e < ps2; e += kIntThreads) {
            const double xr = Xr[e];
            if (!(Pw[e] < thr) && xr != 0.0) {
                const int k = e / ps, l = e - (e / ps) * ps;
                if (k < 2 * h && l < 2 * h) {
                    const int pos = atomicAdd(&nKept, 1);
                    if (pos < cap)
                        dst[pos] = IntBin{k < h ? k : P.intPa - 2 * h + k, l < h ? l : P.intPr - 2 * h + l, xr};
                }
            }
        }
        __syncthreads();
        if (threadIdx.x == 0 && pass == 0) {
            me->n = nKept;
            me->ovf = nKept > kIntInline ? atomicAdd(ovfCount, 1) : 0;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------
// geometry: findImage2Pos (polynomial) or host-computed positions,
// maskValue, complex eligibility; initialises the point state.
__global__ void k_geometry(KPar P, Pts S, const int *hostR2, const int *hostA2, int *list)
{
    const int npts = P.nR * P.nA;
    const int pt = blockIdx.x * blockDim.x + threadIdx.x;
    if (pt >= npts)
        return;
    const int i = pt / P.nR, j = pt - i * P.nR;
    const int a1 = P.aStart + i * P.deltaA - P.wA / 2;
    const int r1 = P.rStart + j * P.deltaR - P.wR / 2;
    int r2, a2;
    if (hostR2) {
        r2 = hostR2[pt];
        a2 = hostA2[pt];
    } else {
        const double sc = (double)P.sf;
        double rs = P.rPoly[0] + (double)(r1 * sc) * P.rPoly[1] + (double)(a1 * sc) * P.rPoly[2];
        double as = P.aPoly[0] + (double)(r1 * sc) * P.aPoly[1] + (double)(a1 * sc) * P.aPoly[2];
        rs /= sc;
        as /= sc;
        r2 = (rs >= 0) ? r1 + (int)(rs + 0.5) : r1 + (int)(rs - 0.5);
        a2 = (as >= 0) ? a1 + (int)(as + 0.5) : a1 + (int)(as - 0.5);
    }
    int mv = 1;
    if (P.mask) {
        const int rr = r1 * P.sf, aa = a1 * P.sf;
        const int ia1 = (aa + P.wA / 2 - P.maskA0) / P.maskNal;
        const int ir1 = (rr + P.wR / 2 - P.maskR0) / P.maskNrl;
        if (ia1 < 0 || ia1 >= P.maskNa || ir1 < 0 || ir1 >= P.maskNr)
            mv = 0;
        else
            mv = P.mask[(size_t)ia1 * P.maskNr + ir1];
        mv = mv < 1 ? mv : 1;
    }
    S.r1[pt] = r1;
    S.a1[pt] = a1;
    S.r2[pt] = r2;
    S.a2[pt] = a2;
    S.maskv[pt] = (uint8_t)mv;
    S.haveData[pt] = 1;
    S.good[pt] = 0;
    S.type[pt] = 0;
    S.azValid[pt] = 0;
    S.azShift[pt] = 0;
    S.azUse[pt] = 0;
    S.corr[pt] = 0.f;
    S.rShift1[pt] = 0.f;
    S.aShift1[pt] = 0.f;
    S.azd[pt] = 0.f;
    for (int t = 0; t < 3; t++) {
        const size_t o = (size_t)t * npts + pt;
        S.peakI[o] = S.peakJ[o] = S.fineI[o] = S.fineJ[o] = -2000;
        S.cmax[o] = 0;
        S.sigRg[o] = S.sigAz[o] = S.peakGap[o] = S.fineGap[o] = S.snr[o] = 0.f;
    }
    const bool ok = in_bounds(r1, a1, P.wR, P.wA, P.im.nr1, P.im.na1) && in_bounds(r2, a2, P.wR, P.wA, P.im.nr2, P.im.na2);
    if (ok && mv > 0 && !P.noComplex) {
        int k = atomicAdd(&S.counts[CNT_LIST], 1);
        list[k] = pt;
    }
}

// ---------------------------------------------------------------------
// complex stage: getPatches (speckleTrack.c:1194-1330) + Doppler +
// phaseCorrectBaseline (speckleTrack.c:868-907).  One block per point.
struct CpxBatch {
    float2 *P;        // [2][B][w*w]: patch 1 then patch 2
    float2 *S;        // [B][4w^2] padded cross spectrum / correlation
    float2 *tmp;      // transpose scratch
    float *mag;       // [B][4w^2]
    PeakRes *res;     // [B]
    float *sig;       // [B][2]
    uint8_t *have;    // [B]
    int *az;          // [B]
    double *p;        // [B][2]
    const int *list;  // fused and matrix-core paths: item -> grid point
    float *nb;        // fused path: [B][kNb^2] |C| around each coarse peak
};

__global__ void __launch_bounds__(NT) k_cpx_gather(KPar P, Pts S, const int *list, int off, int B, CpxBatch C)
{
    __shared__ double sh[NT / 64];
    __shared__ int shi[NT / 64];
    const int b = blockIdx.x;
    const int pt = list[off + b];
    const int wa = P.wA, wr = P.wR, N = wa * wr;  // wA rows of wR samples
    const int r1 = S.r1[pt], a1 = S.a1[pt], r2 = S.r2[pt], a2 = S.a2[pt];
    const bool isf = P.im.stype == ST_SAMPLE_C64;
    const float thr = isf ? 1e-15f : 1e-9f;
    const float thrAll = thr * P.wA * P.wR;
    auto get = [&](int e, float2 &x, float2 &y) {
        const int i = e / wr, j = e - (e / wr) * wr;
        x = load_sample(P.im.img1, P.im.stype, P.im.nr1, a1 + i, r1 + j);
        y = load_sample(P.im.img2, P.im.stype, P.im.nr2, a2 + i, r2 + j);
        if (isf) {
            x.x *= 1000;
            x.y *= 1000;
            y.x *= 1000;
            y.y *= 1000;
        }
        if (P.hanning) {
            y.x *= P.hann[e];
            y.y *= P.hann[e];
        }
    };
    float2 *P1 = C.P + (size_t)b * N, *P2 = C.P + ((size_t)B + b) * N;
    __shared__ IntTab<128> itab;
    if (P.intOn)
        int_tables<128>(P, off + b, imax(wa, wr), itab, threadIdx.x, NT);
    int nZero = 0;
    double p1 = 0, p2 = 0;
    for (int e = threadIdx.x; e < N; e += NT) {
#pragma clang fp contract(off)
        float2 x, y;
        get(e, x, y);
        const float pw1 = x.x * x.x + x.y * x.y, pw2 = y.x * y.x + y.y * y.y;
        nZero += (pw1 < thr || pw2 < thr);
        p1 += pw1;
        p2 += pw2;
        P2[e] = y;
        if (P.baseFlag) {
            const int i = e / wr, j = e - (e / wr) * wr;
            double xx = (a1 - P.wA / 2 - (P.nSlpA1 / 2)) / (double)P.nSlpA1;
            xx += i * (1. / (double)P.nSlpA1);
            double x0 = (a1 - P.wA / 2 - (P.nSlpA1 / 2)) / (double)P.nSlpA1;
            double bn = P.Bn + x0 * P.dBn + x0 * x0 * P.dBnQ;
            double bp = P.Bp + xx * P.dBp + xx * xx * P.dBpQ;
            double bsq = bn * bn + bp * bp;
            int j1 = (r1 - P.wR / 2 > 0 ? r1 - P.wR / 2 : 0) + j;
            double ph = -bn * P.sinT[j1] - bp * P.cosT[j1] + (0.5 * bsq * P.overR[j1]);
            ph *= -(4.0 * 3.14159265358979323846) / P.lambda;
            double pr = cos(ph), pim = sin(ph);
            double cr = x.x, ci = x.y;
            x = make_float2((float)(cr * pr - ci * pim), (float)(cr * pim + ci * pr));
        }
        if (P.intOn) {
            const int i = e / wr, j = e - (e / wr) * wr;
            double cr = x.x, ci = x.y;
            int_apply<128>(P, itab, off + b, i, j, cr, ci);
            x = make_float2((float)cr, (float)ci);
        }
        P1[e] = x;
    }
    nZero = block_sum_i<NT>(nZero, shi);
    p1 = block_sum_d<NT>(p1, sh);
    p2 = block_sum_d<NT>(p2, sh);
    if (nZero > 20 || p1 < thrAll || p2 < thrAll) {
        if (threadIdx.x == 0)
            C.have[b] = 0;
        return;
    }
    const int az = doppler_shift(get, wa, wr, wa, P.wA, sh);
    if (threadIdx.x == 0) {
        C.have[b] = 1;
        C.az[b] = az;
        C.p[2 * b] = p1 / (double)(P.wA * P.wR);
        C.p[2 * b + 1] = p2 / (double)(P.wA * P.wR);
    }
}


// ---------------------------------------------------------------------
// Peak stage (findCPeak speckleTrack.c:744-805, overSampleC :852-863,
// argmax over the central box :686-700, sample extraction for
// gaussPeak1D :2158-2180).  mode 0 = complex (Gaussian fits at the
// refined peak / NOVER), 1 = amplitude (Gaussian fits at the coarse peak).
// Gaussian-mode arguments of the peak stage: the window wWin of the surface,
// the per-item powers p (p1, p2) and their factor pScale, the Hanning
// correction hc (cmpTrackGauss's scale, speckleTrack.c:2249-2254).
struct GaussArg {
    int on, wWinA, wWinR;
    double hc, pScale;
    const double *p;
};

// The part of the peak stage after the coarse scan: border check, FP64
// refinement of the 8 x 8 neighbourhood on the searched 48 x 48 box (or,
// in -gauss mode, the 2-D Gaussian fit), and the samples of the two 1-D
// Gaussian fits, from the fftshifted WA x WR magnitude surface M.
template <bool GAUSS>
__device__ void peak_finish(const float *M, int WA, int WR, int mode, int I, int J, float raw, float sec, float snr,
                            PeakRes *out, const GaussArg &G, int item)
{
    __shared__ double shd[2 * (NT / 64)];
    __shared__ int shr[NT / 64];
    __shared__ __align__(16) unsigned char scr[RefLayout<true>::BYTES];
    PeakRes r;
    r.gap = raw > 0 ? (raw - sec) / raw : 0.f;
    r.snr = snr;
    r.lmOk = 0;
    r.fi = r.fj = -1;
    r.fineMax = 0;
    r.fineGap = 0;
    r.gI = r.gJ = -2000000000;
    r.gC = 0;
    if (I < NFAST / 2 || I > WA - NFAST / 2 || J < NFAST / 2 || J > WR - NFAST / 2) {
        r.I = r.J = -1000;
        r.raw = -1.f;
        if (threadIdx.x == 0)
            *out = r;
        return;
    }
    r.I = I;
    r.J = J;
    r.raw = raw;
    if constexpr (GAUSS) {
        // the fit itself runs in k_gauss_fit, one thread per attempt; the
        // amplitude mode's 1-D fits sit at the coarse peak and are known now
        if (threadIdx.x == 0) {
            if (mode == 1 && I > NGAUSS / 2 && I < WA - NGAUSS / 2 && J > NGAUSS / 2 && J < WR - NGAUSS / 2) {
                r.lmOk = 1;
                for (int t = 0; t < NGAUSS; t++) {
                    r.lmY[0][t] = M[I * WR + (J - NGAUSS / 2 + t)];
                    r.lmY[1][t] = M[(I - NGAUSS / 2 + t) * WR + J];
                }
                r.lmX0[0] = (float)(J - NGAUSS / 2);
                r.lmX0[1] = (float)(I - NGAUSS / 2);
            }
            *out = r;
        }
        (void)G;
        (void)item;
        return;
    }
    // FP64 refinement on the searched 48 x 48 box
    double *Pm = (double *)(scr + kRefPm);
    if (threadIdx.x < NFAST * NFAST)
        Pm[threadIdx.x] = (double)M[(I - NFAST / 2 + threadIdx.x / NFAST) * WR + (J - NFAST / 2 + threadIdx.x % NFAST)];
    __syncthreads();
    double f1, f2;
    int fr;
    refine_fine<NT, true>(scr, threadIdx.x, [] { __syncthreads(); }, f1, f2, fr);
    block_top2_d<NT>(f1, f2, fr, shd, shr);
    if (fr != 0x7fffffff) {
        r.fi = fr / kFine;
        r.fj = fr - r.fi * kFine;
        r.fineMax = sqrt(f1);
        r.fineGap = (float)((sqrt(f1) - sqrt(f2)) / sqrt(f1));
    } else {
        r.fi = r.fj = 0;
        r.fineMax = 0;
    }
    // samples for the two Gaussian fits
    int cI, cJ;
    bool ok;
    if (mode == 0) {
        const int iF = I * NOVER + (r.fi + 24) - (NOVER * NFAST) / 2;
        const int jF = J * NOVER + (r.fj + 24) - (NOVER * NFAST) / 2;
        cI = iF / NOVER;
        cJ = jF / NOVER;
        ok = (cI - NGAUSS / 2) >= 0 && (cI + NGAUSS / 2) < WA && (cJ - NGAUSS / 2) >= 0 && (cJ + NGAUSS / 2) < WR;
    } else {
        cI = I;
        cJ = J;
        ok = I > NGAUSS / 2 && I < WA - NGAUSS / 2 && J > NGAUSS / 2 && J < WR - NGAUSS / 2;
    }
    r.lmOk = ok ? 1 : 0;
    if (ok) {
        for (int t = 0; t < NGAUSS; t++) {
            r.lmY[0][t] = M[cI * WR + (cJ - NGAUSS / 2 + t)];
            r.lmY[1][t] = M[(cI - NGAUSS / 2 + t) * WR + cJ];
        }
        r.lmX0[0] = (float)(cJ - NGAUSS / 2);
        r.lmX0[1] = (float)(cI - NGAUSS / 2);
    }
    if (threadIdx.x == 0)
        *out = r;
}


// -gauss: the 2-D fit of each attempt on one thread (cmpTrackGauss,
// speckleTrack.c:2205-2303), so a wave runs 64 fits at once in the
// reference's sequential operation order.  In the complex mode the 1-D fit
// samples follow at the fitted peak / NOVER (speckleTrack.c:377-386).
__global__ void __launch_bounds__(128) k_gauss_fit(const float *mag, int WA, int WR, int mode, PeakRes *res, int B,
                                                   GaussArg G)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B)
        return;
    PeakRes &r = res[b];
    if (r.I < -999)
        return;  // rejected coarse peak: the fit's edge test fails (defaults stand)
    const float *M = mag + (size_t)b * WA * WR;
    const double p1 = G.p[2 * b] * G.pScale, p2 = G.p[2 * b + 1] * G.pScale;
    const double scale = G.hc / (sqrt(p1 * p2) * pow((double)(G.wWinR * G.wWinA), 2));
    int gI, gJ;
    double gC;
    gauss_fit_2d(M, WR, r.I, r.J, (double)r.raw, G.wWinR, G.wWinA, scale, &gI, &gJ, &gC);
    r.gI = gI;
    r.gJ = gJ;
    r.gC = gC;
    if (mode != 0)
        return;
    const int cI = gI / NOVER, cJ = gJ / NOVER;
    const bool ok = (cI - NGAUSS / 2) >= 0 && (cI + NGAUSS / 2) < WA && (cJ - NGAUSS / 2) >= 0 &&
                    (cJ + NGAUSS / 2) < WR;
    r.lmOk = ok ? 1 : 0;
    if (ok) {
        for (int t = 0; t < NGAUSS; t++) {
            r.lmY[0][t] = M[cI * WR + (cJ - NGAUSS / 2 + t)];
            r.lmY[1][t] = M[(cI - NGAUSS / 2 + t) * WR + cJ];
        }
        r.lmX0[0] = (float)(cJ - NGAUSS / 2);
        r.lmX0[1] = (float)(cI - NGAUSS / 2);
    }
}

// Two fits per attempt: [0] along range, [1] along azimuth.  Launched in
// 128-thread blocks: the bound lets a fit keep its samples, the 3 x 3 system
// and the trial parameters in registers (at the default 1024-thread bound
// they spill to scratch).
__global__ void __launch_bounds__(128) k_lm(const PeakRes *res, int B, float *sig)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * B)
        return;
    const int b = t >> 1, k = t & 1;
    const PeakRes &r = res[b];
    sig[t] = r.lmOk ? gauss_peak_1d(r.lmY[k], r.lmX0[k]) : 0.f;
}

// complexMatch tail + finalizeCmpxMatch (speckleTrack.c:377-394,714-739).
__global__ void k_cpx_finalize(KPar P, Pts S, const int *list, int off, int B, CpxBatch C)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B)
        return;
    const int pt = list[off + b];
    const int npts = P.nR * P.nA;
    if (!C.have[b]) {
        S.haveData[pt] = 0;
        return;
    }
    S.azValid[pt] = 1;
    S.azShift[pt] = C.az[b];
    atomicAdd(&S.counts[CNT_CORR0], 1);
    const PeakRes &r = C.res[b];
    const double p1 = C.p[2 * b], p2 = C.p[2 * b + 1];
    double cMax;
    int iMax, jMax;
    if (P.gauss) {
        // complexMatch's -gauss branch (speckleTrack.c:367-372)
        cMax = r.gC;
        iMax = r.gI;
        jMax = r.gJ;
    } else if (r.I < -999) {
        cMax = -1.0;
        iMax = jMax = -1000;
    } else {
        const double hc = P.hanning ? 1.5 : 1.0;
        cMax = r.fineMax / (NFAST * NFAST);
        cMax = hc * cMax / (sqrt(p1 * p2) * pow((double)(P.wR * P.wA), 2));
        iMax = r.I * NOVER + (r.fi + 24) - (NOVER * NFAST) / 2;
        jMax = r.J * NOVER + (r.fj + 24) - (NOVER * NFAST) / 2;
    }
    float azd = -2.0e9f;
    int good = 0;
    float sRg = C.sig[2 * b], sAz = C.sig[2 * b + 1];
    if (r.lmOk) {
        azd = sAz / sRg;
        if (fabsf(azd) < P.azThresh || !P.checkAz) {
            const double invNOVER = 1.0 / (float)NOVER;
            const double wRo = P.wR / 2. * NOVER * 2, wAo = P.wA / 2. * NOVER * 2;
            if (cMax > .185 && !P.noComplex) {
                float rs = (float)(((float)jMax - wRo) * invNOVER / 2);
                float as = (float)(((float)iMax - wAo) * invNOVER / 2);
                if (fabs((double)rs) < 0.125 * P.wR * 2 && fabs((double)as) < 0.125 * P.wA * 2) {
                    good = 1;
                    S.rShift1[pt] = rs;
                    S.aShift1[pt] = as;
                    atomicAdd(&S.counts[CNT_COMPLEX], 1);
                }
            }
        }
    }
    S.azd[pt] = azd;
    S.corr[pt] = (float)cMax;
    if (good) {
        S.good[pt] = 1;
        S.type[pt] = ST_CMATCH;
    }
    S.peakI[pt] = r.I;
    S.peakJ[pt] = r.J;
    S.fineI[pt] = iMax;
    S.fineJ[pt] = jMax;
    S.cmax[pt] = cMax;
    S.sigRg[pt] = r.lmOk ? sRg : 0.f;
    S.sigAz[pt] = r.lmOk ? sAz : 0.f;
    S.peakGap[pt] = r.gap;
    S.fineGap[pt] = r.fineGap;
    S.snr[pt] = r.snr;
    (void)npts;
}

// Carry the last estimated azShift forward in raster order (the
// reference's trackPar->azShift persists between points).
// The persistent azimuth shift (speckleTrack.c:246-262): a point uses the
// Doppler shift of the last point at or before it in raster order that
// estimated one, 0 before the first.  With lastIdx(p) = max{q <= p : valid
// q}, this is a max-scan: k_carry_blocks leaves each block's largest valid
// index, k_carry_az scans its block and takes the carry from the blocks
// before it.
constexpr int kCarryBlk = 1024;

__device__ inline int wave_max_scan(int v)
{
    const int lane = threadIdx.x & 63;
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v, o);
        if (lane >= o)
            v = max(v, u);
    }
    return v;
}

__global__ void __launch_bounds__(kCarryBlk) k_carry_blocks(Pts S, int npts, int *blkMax)
{
    __shared__ int sh[kCarryBlk / 64];
    const int p = blockIdx.x * kCarryBlk + threadIdx.x;
    int v = (p < npts && S.azValid[p]) ? p : -1;
    for (int o = 32; o > 0; o >>= 1)
        v = max(v, __shfl_xor(v, o));
    if ((threadIdx.x & 63) == 0)
        sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int m = -1;
        for (int w = 0; w < kCarryBlk / 64; w++)
            m = max(m, sh[w]);
        blkMax[blockIdx.x] = m;
    }
}

__global__ void __launch_bounds__(kCarryBlk) k_carry_az(Pts S, int npts, const int *blkMax)
{
    __shared__ int sh[kCarryBlk / 64];
    __shared__ int carry;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // carry from the blocks before this one
    int c = -1;
    for (int k = threadIdx.x; k < (int)blockIdx.x; k += kCarryBlk)
        c = max(c, blkMax[k]);
    for (int o = 32; o > 0; o >>= 1)
        c = max(c, __shfl_xor(c, o));
    if (lane == 0)
        sh[wv] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        int m = -1;
        for (int w = 0; w < kCarryBlk / 64; w++)
            m = max(m, sh[w]);
        carry = m;
    }
    __syncthreads();
    const int p = blockIdx.x * kCarryBlk + threadIdx.x;
    int v = wave_max_scan((p < npts && S.azValid[p]) ? p : -1);
    if (lane == 63)
        sh[wv] = v;
    __syncthreads();
    int m = carry;
    for (int w = 0; w < wv; w++)
        m = max(m, sh[w]);
    v = max(v, m);
    if (p < npts)
        S.azUse[p] = v >= 0 ? S.azShift[v] : 0;
}

// Work list of the amplitude tries (the while loop of speckleTrack.c:246).
__global__ void k_amp_list(KPar P, Pts S, int npts, int *list)
{
    const int pt = blockIdx.x * blockDim.x + threadIdx.x;
    if (pt >= npts)
        return;
    if (!S.good[pt] && S.haveData[pt] && S.maskv[pt] > 0) {
        int k = atomicAdd(&S.counts[CNT_LIST], 1);
        list[k] = pt;
    }
}

// ---------------------------------------------------------------------
// amplitude stages (getAmpPatches speckleTrack.c:1006-1189, ampMatch
// :398-484, ampTrackFast :489-597)
struct AmpPart;
struct Top2;
// ws = wsA x wsR: the patch (wAa x wRa, doubled on the large try);
// W = WA x WR = 2 wsA x 2 wsR: the interpolated surfaces.
struct AmpBatch {
    float2 *I;       // [2][B][ws] patches, then their range spectra
    float2 *F;       // [2][B][W]: interpolation columns [2][B][WA][wsR], then range spectra Z [B][W]
    float2 *S;       // [B][W]: amplitudes [2][B][W] (float), then the correlation's azimuth half
    float *mag;      // [B][W]
    PeakRes *res;
    float *sig;
    uint8_t *state;  // 0 = bounds failed, 1 = no data, 2 = matched
    int *az;
    double *p;
    AmpPart *dpart;  // detection partials [B][nPd]
    double2 *ppart;  // power partials [B][nPd]
    double2 *mean;   // amplitude means [B]
    Top2 *part;      // coarse top two [B][nPm]
};

__global__ void __launch_bounds__(NT) k_amp_gather(KPar P, Pts S, const int *list, int off, int B, int large,
                                                    AmpBatch A)
{
    __shared__ double sh[NT / 64];
    const int b = blockIdx.x;
    const int pt = list[off + b];
    const int sc = large ? LA : 1;
    const int wsA = P.wAa * sc, wsR = P.wRa * sc;
    const int i = pt / P.nR, j = pt - (pt / P.nR) * P.nR;
    const int a1 = P.aStart + i * P.deltaA - P.wA / 2 * sc;
    const int r1 = P.rStart + j * P.deltaR - P.wR / 2 * sc;
    const int r2 = r1 + (S.r2[pt] - S.r1[pt]), a2 = a1 + (S.a2[pt] - S.a1[pt]);
    const bool ok = in_bounds(r1, a1, P.wRa * sc, P.wAa * sc, P.im.nr1, P.im.na1) &&
                    in_bounds(r2, a2, P.wRa * sc, P.wAa * sc, P.im.nr2, P.im.na2);
    if (!ok) {
        if (threadIdx.x == 0)
            A.state[b] = 0;
        return;
    }
    const int N = wsA * wsR;
    float2 *I1 = A.I + (size_t)b * N, *I2 = A.I + ((size_t)B + b) * N;
    for (int e = threadIdx.x; e < N; e += NT) {
        const int ii = e / wsR, jj = e - (e / wsR) * wsR;
        I1[e] = load_sample(P.im.img1, P.im.stype, P.im.nr1, a1 + ii, r1 + jj);
        I2[e] = load_sample(P.im.img2, P.im.stype, P.im.nr2, a2 + ii, r2 + jj);
    }
    int azUse = S.azUse[pt];
    if (P.noComplex && !large) {
        // the small try's patch: wAa rows of wRa samples
        const int rl = P.wRa;
        auto get = [&](int e, float2 &x, float2 &y) {
            const int ii = e / rl, jj = e - (e / rl) * rl;
            x = load_sample(P.im.img1, P.im.stype, P.im.nr1, a1 + ii, r1 + jj);
            y = load_sample(P.im.img2, P.im.stype, P.im.nr2, a2 + ii, r2 + jj);
        };
        azUse = doppler_shift(get, P.wAa, P.wRa, P.wAa, P.wA, sh);
        if (threadIdx.x == 0)
            S.azUse[pt] = azUse;
    }
    if (threadIdx.x == 0) {
        A.state[b] = 2;
        A.az[b] = (int)((float)azUse * (float)wsA / (float)P.wA);
    }
}

// ampTrackFast tail and the driver's acceptance test
// (speckleTrack.c:537-597, 272-298).
__global__ void k_amp_finalize(KPar P, Pts S, const int *list, int off, int B, int large, AmpBatch A)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B)
        return;
    const int pt = list[off + b];
    const int npts = P.nR * P.nA;
    const int t = large ? 2 : 1;
    const size_t o = (size_t)t * npts + pt;
    if (A.state[b] == 0) {
        S.corr[pt] = 0.f;
        return;
    }
    atomicAdd(&S.counts[CNT_CORR0 + t], 1);
    if (A.state[b] == 1) {
        S.haveData[pt] = 0;
        S.corr[pt] = 0.f;
        return;
    }
    const int sc = large ? LA : 1;
    const int WA = 2 * P.wAa * sc, WR = 2 * P.wRa * sc;
    const PeakRes &r = A.res[b];
    float sRg = A.sig[2 * b], sAz = A.sig[2 * b + 1];
    if (r.lmOk)
        S.azd[pt] = sAz / sRg;
    const float azdNow = S.azd[pt];
    double cMax;
    int iMax, jMax;
    if (r.I < -999 || (fabsf(azdNow) >= P.azThresh && P.checkAz)) {
        cMax = r.I < -999 ? -1.0 : (double)r.raw;
        iMax = r.I;
        jMax = r.J;
    } else if (P.gauss) {
        // ampTrackFast's -gauss branch (speckleTrack.c:553-557)
        cMax = r.gC;
        iMax = r.gI;
        jMax = r.gJ;
    } else {
        const double p1 = A.p[2 * b], p2 = A.p[2 * b + 1];
        cMax = r.fineMax / (pow((double)WA, 2) * pow((double)WR, 2) * pow((double)NFAST, 2));
        cMax = cMax / sqrt(p1 * p2);
        iMax = r.I * NOVER + (r.fi + 24) - (NOVER * NFAST) / 2;
        jMax = r.J * NOVER + (r.fj + 24) - (NOVER * NFAST) / 2;
    }
    const double invNOVER = 1.0 / (float)NOVER;
    const float rs = (float)(((float)jMax - P.wRa * 0.5 * 2 * NOVER * sc) * invNOVER / 2);
    const float as = (float)(((float)iMax - P.wAa * 0.5 * 2 * NOVER * sc) * invNOVER / 2);
    const double thr = large ? 0.028 : 0.07;
    S.corr[pt] = (float)cMax;
    if (fabs((double)rs) < 0.15 * P.wRa * 2 * sc && fabs((double)as) < 0.15 * P.wAa * 2 * sc && cMax > thr) {
        S.good[pt] = 1;
        S.type[pt] = large ? ST_AMPMATCHLARGE : ST_AMPMATCH;
        S.rShift1[pt] = rs;
        S.aShift1[pt] = as;
        atomicAdd(&S.counts[large ? CNT_AMPL : CNT_AMP], 1);
    }
    S.peakI[o] = r.I;
    S.peakJ[o] = r.J;
    S.fineI[o] = iMax;
    S.fineJ[o] = jMax;
    S.cmax[o] = cMax;
    S.sigRg[o] = r.lmOk ? sRg : 0.f;
    S.sigAz[o] = r.lmOk ? sAz : 0.f;
    S.peakGap[o] = r.gap;
    S.fineGap[o] = r.fineGap;
    S.snr[o] = r.snr;
}

// Final offsets (speckleTrack.c:301-323).
__global__ void k_compose(KPar P, Pts S, float *offR, float *offA, int npts)
{
    const int pt = blockIdx.x * blockDim.x + threadIdx.x;
    if (pt >= npts)
        return;
    float rs, as;
    if (S.good[pt] && S.maskv[pt] > 0) {
        rs = (float)(S.r2[pt] - S.r1[pt]) - S.rShift1[pt];
        as = (float)(S.a2[pt] - S.a1[pt]) - S.aShift1[pt];
    } else {
        rs = -2.0e9f;
        as = -2.0e9f;
        if (S.haveData[pt])
            atomicAdd(&S.counts[CNT_FAIL], 1);
    }
    if (rs > -100000) {
        rs *= (float)P.sf;
        as *= (float)P.sf;
    }
    offR[pt] = rs;
    offA[pt] = as;
}

// ---------------------------------------------------------------------
// Fused complex match for w in {16, 32, 64}: one workgroup of 4w threads
// per patch pair, persistent over the work list.  The chain
// windows -> getPatches' sums and acceptance test -> cmpPSWithPad -> 2w x 2w
// inverse -> findCPeak -> overSampleC stays on chip:
//   * the window sums and the Doppler lag sum are formed from the loaded
//     samples (lane_sums) and reduced across the workgroup while the
//     forward transforms run; one thread applies the acceptance test and
//     finds the shift (patch_accept) before the inverse needs it;
//   * the w x w forward FFTs run in registers, each length-w transform
//     split over a lane pair (st_fftreg.h), with one LDS transpose;
//   * the zero-padded 2w x 2w inverse is done as its four polyphase
//     components C[2n+p][2m+q] = IFFT_w(X[(k+az) mod w][l] exp(i pi (k p +
//     l q) / w)) (k, l signed), so no padded array is ever formed;
//   * the fftshifted magnitude surface stays in LDS for the argmax; the
//     cells of the 22 x 22 block around the coarse peak that k_cpx_refine
//     reads (kNbMap) go to it, which refines the peak and takes the
//     Gaussian-fit samples.
// Thread t: group g = t / 2w (patch g in the forward part, column parity q
// = g in the inverse part), index r = (t / 2) mod w, lane half h = t & 1.
struct Top2 {
    float v1;
    int r1;
    float v2;
    float sum;  // sum of |C| over the values seen (the SNR diagnostic's mean)
};

__device__ inline Top2 top2_combine(Top2 a, Top2 b)
{
    Top2 o;
    o.sum = a.sum + b.sum;
    if (b.v1 > a.v1 || (b.v1 == a.v1 && b.r1 < a.r1)) {
        o.v1 = b.v1;
        o.r1 = b.r1;
        o.v2 = fmaxf(a.v1, b.v2);
    } else {
        o.v1 = a.v1;
        o.r1 = a.r1;
        o.v2 = fmaxf(b.v1, a.v2);
    }
    return o;
}

template <int NTH>
__device__ inline Top2 wg_top2(Top2 a, Top2 *sh)
{
#pragma unroll
    for (int o = 1; o < (NTH < 64 ? NTH : 64); o <<= 1) {
        Top2 b;
        b.v1 = __shfl_xor(a.v1, o);
        b.r1 = __shfl_xor(a.r1, o);
        b.v2 = __shfl_xor(a.v2, o);
        b.sum = __shfl_xor(a.sum, o);
        a = top2_combine(a, b);
    }
    if constexpr (NTH <= 64) {
        return a;
    } else {
        const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
        __syncthreads();
        if (lane == 0)
            sh[wv] = a;
        __syncthreads();
        Top2 r = sh[0];
        for (int k = 1; k < NTH / 64; k++)
            r = top2_combine(r, sh[k]);
        __syncthreads();
        return r;
    }
}

// sqrtf for x >= 2^-96 (finite): the hardware square root, then the two
// one-ulp corrections of the correctly rounded sequence the compiler emits
// for sqrtf (its input scaling below 2^-96 and its zero / infinity pass-
// through change nothing in this range), so the result is sqrtf(x) bit for
// bit.
__device__ inline float sqrt_rn_normal(float x)
{
    const float s = __builtin_amdgcn_sqrtf(x);
    const float sm = __int_as_float(__float_as_int(s) - 1), sp = __int_as_float(__float_as_int(s) + 1);
    const float rm = __builtin_fmaf(-sm, s, x), rp = __builtin_fmaf(-sp, s, x);
    float r = rm <= 0.0f ? sm : s;
    r = rp > 0.0f ? sp : r;
    return r;
}

// The peak scan without ranks: each thread keeps the largest value, the
// runner-up (the larger of all others) and the sum; the merge gives the
// same three numbers as top2_combine, in the same order of additions.
__device__ inline Top2 max2_combine(Top2 a, Top2 b)
{
    Top2 o;
    o.sum = a.sum + b.sum;
    o.v1 = fmaxf(a.v1, b.v1);
    o.v2 = fmaxf(fmaxf(a.v2, b.v2), fminf(a.v1, b.v1));
    o.r1 = 0x7fffffff;
    return o;
}

__device__ inline void max2_add(Top2 &a, float v)
{
    a.sum = a.sum + v;
    a.v2 = fmaxf(a.v2, fminf(a.v1, v));
    a.v1 = fmaxf(a.v1, v);
}

template <int NTH>
__device__ inline Top2 wg_max2(Top2 a, Top2 *sh)
{
#pragma unroll
    for (int o = 1; o < (NTH < 64 ? NTH : 64); o <<= 1) {
        Top2 b;
        b.v1 = __shfl_xor(a.v1, o);
        b.v2 = __shfl_xor(a.v2, o);
        b.sum = __shfl_xor(a.sum, o);
        a = max2_combine(a, b);
    }
    if constexpr (NTH <= 64) {
        return a;
    } else {
        const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
        __syncthreads();
        if (lane == 0)
            sh[wv] = a;
        __syncthreads();
        Top2 r = sh[0];
        for (int k = 1; k < NTH / 64; k++)
            r = max2_combine(r, sh[k]);
        __syncthreads();
        return r;
    }
}

// findCPeak's rank of a surface cell (I, J) of the 2w x 2w fftshifted
// surface: row-major within the unshifted quadrant, quadrant order 0-3.
template <int W>
__device__ inline int surf_rank(int I, int J)
{
    const int i0 = I < W ? I : I - W, j0 = J < W ? J : J - W;
    const int q = (I < W) ? (J < W ? 0 : 3) : (J < W ? 2 : 1);
    return (i0 * W + j0) * 4 + q;
}

// Sums of getPatches (speckleTrack.c:1194-1330) and estDopCarrier1
// (dopplerEstimation.c:42-88) over one window: [0] sum |x|^2, [1..2] sum x,
// [3..4] the lag-1 sum along azimuth sum x[a] conj(x[a - 1]), [5..6] the
// sum over row 0, [7..8] the sum over row W - 1.
struct PatchSums {
    double v[2][9];
};

// The acceptance test (more than 20 near-zero samples or a near-zero
// window rejects the pair) and the Doppler shift from the sums, with the
// lag-1 sum expanded around the patch mean m:
//   sum (x1 - m) conj(x0 - m) = sum x1 conj(x0) - conj(m) sum x1
//                               - m sum conj(x0) + (W - 1) W |m|^2
// (the row-0 / row-(W-1) sums taken out of the totals).  Writes the item's
// have / az / p; returns have.
template <int W, int STYPE>
__device__ inline int patch_accept(const KPar &P, const CpxBatch &C, int it, const PatchSums &s, int nz, int &az)
{
    const float thr = STYPE == ST_SAMPLE_C64 ? 1e-15f : 1e-9f;
    const float thrAll = thr * W * W;
    az = 0;
    if (nz > 20 || s.v[0][0] < thrAll || s.v[1][0] < thrAll) {
        C.have[it] = 0;
        return 0;
    }
    const double cnt = (double)W * (double)W;
    double zr[2], zi[2];
    for (int g = 0; g < 2; g++) {
        const double *T = s.v[g] + 1, *Lg = s.v[g] + 3, *F0 = s.v[g] + 5, *FL = s.v[g] + 7;
        const double mr = T[0] / cnt, mi = T[1] / cnt;
        // sums over the lag pairs: x1 over rows 1..W-1, x0 over rows 0..W-2
        const double s1r = T[0] - F0[0], s1i = T[1] - F0[1];
        const double s0r = T[0] - FL[0], s0i = T[1] - FL[1];
        // - conj(m) s1 - m conj(s0) + (W - 1) W |m|^2
        zr[g] = Lg[0] - (mr * s1r + mi * s1i) - (mr * s0r + mi * s0i) +
                (double)(W - 1) * (double)W * (mr * mr + mi * mi);
        zi[g] = Lg[1] - (mr * s1i - mi * s1r) - (mi * s0r - mr * s0i);
    }
    const int shift = doppler_round(zr[0], zi[0], zr[1], zi[1], W, c_root256);
    az = (int)((float)shift * (float)P.wA / (float)W);
    C.have[it] = 1;
    C.az[it] = az;
    C.p[2 * it] = s.v[0][0] / cnt;
    C.p[2 * it + 1] = s.v[1][0] / cnt;
    return 1;
}

// Value of lane h - 1 of the lane's group of L (lane L - 1 for h = 0).
template <int L>
__device__ inline float2 lane_prev(float2 v)
{
    constexpr int ctl = L == 2 ? 0xB1 : 0x93;  // quad_perm [1,0,3,2] / [3,0,1,2]
    return make_float2(__int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v.x), ctl, 0xF, 0xF, false)),
                       __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v.y), ctl, 0xF, 0xF, false)));
}

// One lane's part of the window sums: it holds rows a = L m + h (m < H) of
// a column in x (scaled and tapered as the transforms take them).  The
// previous row of the lag sum comes from lane h - 1 of the group, or from
// lane L - 1 one step earlier; the per-sample arithmetic is k_cpx_prep_cols'.
// Bit m of mask: |x[m]|^2 below the near-zero threshold.
template <int H, int L>
__device__ inline void lane_sums(const float2 (&x)[H], int h, float thr, double (&acc)[9], unsigned &mask)
{
    float2 carry = make_float2(0.f, 0.f);
#pragma unroll
    for (int k = 0; k < 9; k++)
        acc[k] = 0.0;
    mask = 0u;
#pragma unroll
    for (int m = 0; m < H; m++) {
#pragma clang fp contract(off)
        const float2 v = x[m];
        const float2 nb = lane_prev<L>(v);
        const float2 pv = h ? nb : carry;
        carry = nb;
        const float pw = v.x * v.x + v.y * v.y;
        acc[0] += pw;
        acc[1] += v.x;
        acc[2] += v.y;
        if (m == 0 && h == 0) {
            acc[5] += v.x;
            acc[6] += v.y;
        } else {
            acc[3] = fma((double)v.y, (double)pv.y, fma((double)v.x, (double)pv.x, acc[3]));
            acc[4] = fma(-(double)v.x, (double)pv.y, fma((double)v.y, (double)pv.x, acc[4]));
        }
        if (m == H - 1 && h == L - 1) {
            acc[7] += v.x;
            acc[8] += v.y;
        }
        mask |= (pw < thr ? 1u : 0u) << m;
    }
}

// Lane values moved by DPP control CTRL within the rows of row_mask RM
// (lanes of other rows, and lanes whose source lies outside the row, get 0).
template <int CTRL, int RM>
__device__ inline double dpp_d(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, RM, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, RM, 0xF, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

template <int CTRL, int RM>
__device__ inline int dpp_i(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, RM, 0xF, true);
}

// Sum over aligned groups of SPAN (32 or 64) lanes, complete in the
// group's last lane: prefix sums within each 16-lane row (row_shr 1, 2, 4,
// 8), then row 0's total into row 1 and row 2's into row 3 (row_bcast:15),
// and for SPAN = 64 lane 31's into rows 2 and 3 (row_bcast:31).
template <int SPAN, class T, int N>
__device__ inline void group_sum_last(T (&acc)[N])
{
    static_assert(SPAN == 32 || SPAN == 64, "group span");
    auto step = [&](auto f) {
#pragma unroll
        for (int k = 0; k < N; k++)
            acc[k] += f(acc[k]);
    };
    if constexpr (sizeof(T) == 8) {
        step([](T v) { return dpp_d<0x111, 0xF>(v); });
        step([](T v) { return dpp_d<0x112, 0xF>(v); });
        step([](T v) { return dpp_d<0x114, 0xF>(v); });
        step([](T v) { return dpp_d<0x118, 0xF>(v); });
        step([](T v) { return dpp_d<0x142, 0xA>(v); });
        if constexpr (SPAN == 64)
            step([](T v) { return dpp_d<0x143, 0xC>(v); });
    } else {
        step([](T v) { return dpp_i<0x111, 0xF>(v); });
        step([](T v) { return dpp_i<0x112, 0xF>(v); });
        step([](T v) { return dpp_i<0x114, 0xF>(v); });
        step([](T v) { return dpp_i<0x118, 0xF>(v); });
        step([](T v) { return dpp_i<0x142, 0xA>(v); });
        if constexpr (SPAN == 64)
            step([](T v) { return dpp_i<0x143, 0xC>(v); });
    }
}

// Per-patch statistics of getPatches (speckleTrack.c:1194-1330) with one
// lane per window column (two at w = 128): 64 / W patch pairs per wave (one
// at w = 128), each lane walks its columns down both windows (coalesced row
// reads), so every sum is a lane sum followed by a segmented wave
// reduction -- no LDS, no barriers.  The
// Doppler lag-1 sum of estDopCarrier1 (dopplerEstimation.c:42-88) is
// expanded around the patch mean:
//   sum (x1 - m) conj(x0 - m) = sum x1 conj(x0) - conj(m) sum x1
//                               - m sum conj(x0) + (W - 1) W |m|^2
// with the first / last row sums taken out of the totals.
// Row segments of the prep kernel (one wave each): at w = 128 a single
// wave per patch leaves too few waves in flight to cover the loads.
template <int W>
constexpr int prep_segs()
{
    return W == 128 ? 4 : 1;
}

template <int W, int STYPE>
__global__ void __launch_bounds__(64 * prep_segs<W>()) k_cpx_prep_cols(const KPar *__restrict__ Pp, const Pts *__restrict__ Sp,
                                                      const int *list, int count, const CpxBatch *__restrict__ Cp)
{
    const KPar &P = *Pp;
    const Pts &S = *Sp;
    const CpxBatch &C = *Cp;
    // lanes per patch, patches per wave, columns per lane
    constexpr int LPP = W < 64 ? W : 64, PPW = 64 / LPP, CPL = W / LPP;
    // row segments of RS rows, one wave each (the wave of segment g > 0
    // starts its lag sum from the last row of segment g - 1)
    constexpr int SEG = prep_segs<W>(), RS = W / SEG;
    constexpr bool isf = STYPE == ST_SAMPLE_C64;
    const float thr = isf ? 1e-15f : 1e-9f;
    const int lane = threadIdx.x & 63, seg = threadIdx.x >> 6;
    // consecutive items (overlapping patches) on one XCD, so their shared
    // rows hit that XCD's L2: workgroups are dealt round-robin over the 8 XCDs
    const int nb = gridDim.x, per = nb / 8, b = blockIdx.x;
    const int blk = b < per * 8 ? (b % 8) * per + b / 8 : b;
    const int it = blk * PPW + lane / LPP, c0 = lane % LPP;
    const bool act = it < count;
    // per column of the lane (CPL independent accumulation chains), summed after the loop
    double ps[2] = {0, 0}, T[2][2] = {{0, 0}, {0, 0}}, Lg[2][2] = {{0, 0}, {0, 0}};
    double F0[2][2] = {{0, 0}, {0, 0}}, FL[2][2] = {{0, 0}, {0, 0}};
    int nz = 0;
    if (act) {
        const int pt = list[it];
        const int r1 = S.r1[pt], a1 = S.a1[pt], r2 = S.r2[pt], a2 = S.a2[pt];
        const bool useH = P.hanning != 0;
        double cps[CPL][2], cT[CPL][2][2], cLg[CPL][2][2];
        float2 prev[CPL][2];
        // sample (a, c) of window g, scaled and tapered as the transforms take it
        auto sample = [&](int g, int a, int c) {
#pragma clang fp contract(off)
            const void *img = g ? P.im.img2 : P.im.img1;
            const int nr = g ? P.im.nr2 : P.im.nr1;
            const size_t o = (size_t)((g ? a2 : a1) + a) * nr + (g ? r2 : r1) + c;
            float2 x;
            if constexpr (isf) {
                x = __builtin_bit_cast(float2, ((const gu64 *)img)[o]);
                x.x *= 1000;
                x.y *= 1000;
            } else {
                const unsigned q = ((const gu32 *)img)[o];
                x = make_float2((float)(short)(q & 0xffffu), (float)(short)(q >> 16));
            }
            if (g == 1 && useH) {
                const float hw = ((const gfloat *)P.hann)[a * W + c];
                x.x *= hw;
                x.y *= hw;
            }
            return x;
        };
        const int a0 = seg * RS;
#pragma unroll
        for (int cc = 0; cc < CPL; cc++)
#pragma unroll
            for (int g = 0; g < 2; g++) {
                cps[cc][g] = 0;
                cT[cc][g][0] = cT[cc][g][1] = 0;
                cLg[cc][g][0] = cLg[cc][g][1] = 0;
                prev[cc][g] = a0 > 0 ? sample(g, a0 - 1, c0 + cc * LPP) : make_float2(0.f, 0.f);
            }
#pragma unroll(CPL > 1 ? 2 : 4)
        for (int a = a0; a < a0 + RS; a++) {
#pragma unroll
            for (int cc = 0; cc < CPL; cc++) {
#pragma clang fp contract(off)
                const int c = c0 + cc * LPP;
                float pw[2];
#pragma unroll
                for (int g = 0; g < 2; g++) {
                    const float2 x = sample(g, a, c);
                    pw[g] = x.x * x.x + x.y * x.y;
                    cps[cc][g] += pw[g];
                    cT[cc][g][0] += x.x;
                    cT[cc][g][1] += x.y;
                    if (a == 0) {
                        F0[g][0] += x.x;
                        F0[g][1] += x.y;
                    } else {
                        // x conj(prev): re = x.re p.re + x.im p.im, im = x.im p.re - x.re p.im.
                        // A product of two floats is exact in double, so each fma
                        // rounds once, like the add it replaces.
                        const float2 pv = prev[cc][g];
                        cLg[cc][g][0] = fma((double)x.y, (double)pv.y, fma((double)x.x, (double)pv.x, cLg[cc][g][0]));
                        cLg[cc][g][1] = fma(-(double)x.x, (double)pv.y, fma((double)x.y, (double)pv.x, cLg[cc][g][1]));
                    }
                    if (a == W - 1) {
                        FL[g][0] += x.x;
                        FL[g][1] += x.y;
                    }
                    prev[cc][g] = x;
                }
                nz += (pw[0] < thr || pw[1] < thr) ? 1 : 0;
            }
        }
#pragma unroll
        for (int cc = 0; cc < CPL; cc++)
#pragma unroll
            for (int g = 0; g < 2; g++) {
                ps[g] += cps[cc][g];
                T[g][0] += cT[cc][g][0];
                T[g][1] += cT[cc][g][1];
                Lg[g][0] += cLg[cc][g][0];
                Lg[g][1] += cLg[cc][g][1];
            }
    }
    // segmented reduction over the LPP lanes of each patch pair
#pragma unroll
    for (int o = 1; o < LPP; o <<= 1) {
        for (int g = 0; g < 2; g++) {
            ps[g] += __shfl_xor(ps[g], o);
            for (int k = 0; k < 2; k++) {
                T[g][k] += __shfl_xor(T[g][k], o);
                Lg[g][k] += __shfl_xor(Lg[g][k], o);
                F0[g][k] += __shfl_xor(F0[g][k], o);
                FL[g][k] += __shfl_xor(FL[g][k], o);
            }
        }
        nz += __shfl_xor(nz, o);
    }
    if constexpr (SEG > 1) {
        // the segments' sums, added in segment order by wave 0 (PPW = 1 here)
        static_assert(PPW == 1, "one patch per wave when rows are split");
        __shared__ double segs[SEG][19];
        if (c0 == 0) {
            double *d = segs[seg];
            for (int g = 0; g < 2; g++) {
                d[g * 9] = ps[g];
                for (int k = 0; k < 2; k++) {
                    d[g * 9 + 1 + k] = T[g][k];
                    d[g * 9 + 3 + k] = Lg[g][k];
                    d[g * 9 + 5 + k] = F0[g][k];
                    d[g * 9 + 7 + k] = FL[g][k];
                }
            }
            d[18] = nz;
        }
        __syncthreads();
        if (seg != 0 || c0 != 0 || !act)
            return;
        for (int q = 1; q < SEG; q++) {
            const double *d = segs[q];
            for (int g = 0; g < 2; g++) {
                ps[g] += d[g * 9];
                for (int k = 0; k < 2; k++) {
                    T[g][k] += d[g * 9 + 1 + k];
                    Lg[g][k] += d[g * 9 + 3 + k];
                    F0[g][k] += d[g * 9 + 5 + k];
                    FL[g][k] += d[g * 9 + 7 + k];
                }
            }
            nz += (int)d[18];
        }
    }
    if (!act || c0 != 0)
        return;
    PatchSums ss;
    for (int g = 0; g < 2; g++) {
        ss.v[g][0] = ps[g];
        for (int k = 0; k < 2; k++) {
            ss.v[g][1 + k] = T[g][k];
            ss.v[g][3 + k] = Lg[g][k];
            ss.v[g][5 + k] = F0[g][k];
            ss.v[g][7 + k] = FL[g][k];
        }
    }
    int az;
    patch_accept<W, STYPE>(P, C, it, ss, nz, az);
}

// ---------------------------------------------------------------------
// Staged complex path in two passes (windows outside the fused kernels'
// sizes): the forward range transforms of both patches are an engine row
// pass; k_cpx_colx and k_cpx_rowmag then do the rest of cmpPSWithPad
// (speckleTrack.c:604-643) and the coarse scan of cmpTrackFast (:648-709)
// without forming the 2w x 2w padded spectrum: the zero pad is separable,
// so the azimuth half (forward azimuth transforms, F1 conj(F2), rotation by
// the Doppler shift, pad, inverse) runs on spectrum columns and the range
// half (pad, inverse, |C|) on rows.
constexpr int kStElems = 4096;  // elements per workgroup (LDS and register staging)

// Azimuth half: C adjacent spectrum columns la (of wr) of one patch per
// workgroup, each wa long; out T[b][n][la], n < 2wa.  Workgroups are
// numbered so that the column blocks sharing cache lines run on the same XCD.
__global__ void __launch_bounds__(NT) k_cpx_colx(const float2 *P, int B, int wa, int wr, const int *az, float2 *T,
                                                  int lc, stfft::PlanArgs Pf, stfft::PlanArgs Pi, stfft::FastDiv dW)
{
    extern __shared__ float2 ldsx[];
    constexpr int IT = kStElems / NT;
    const int C = 1 << lc, W = 2 * wa, h = wa / 2;
    const int nbx = gridDim.x, total = nbx * gridDim.y, flat = blockIdx.y * nbx + blockIdx.x;
    const int per = total / 8;
    const int Lb = flat < per * 8 ? (flat % 8) * per + flat / 8 : flat;
    const int b = Lb / nbx, c0 = (Lb - (Lb / nbx) * nbx) * C;
    const int nc = imin(C, wr - c0);
    const int px = W + 1, py = wa + 1;
    // F1 [C][py] and F2 [C][py] back to back, so one transform takes both;
    // the padded columns [C][px] then overwrite them from the start of the
    // buffer (every read of the pad pass precedes its first write)
    float2 *X = ldsx, *Y = ldsx + (size_t)C * py;
    const size_t N = (size_t)wa * wr;
    const float2 *P1 = P + (size_t)b * N + c0, *P2 = P + ((size_t)B + b) * N + c0;
    {
        const int cnt = wa << lc;
        float2 u[IT], v[IT];
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = threadIdx.x + it * NT, a = t >> lc, c = t & (C - 1);
            const bool ok = t < cnt && c < nc;
            u[it] = ok ? P1[(size_t)a * wr + c] : make_float2(0.f, 0.f);
            v[it] = ok ? P2[(size_t)a * wr + c] : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = threadIdx.x + it * NT, a = t >> lc, c = t & (C - 1);
            if (t < cnt) {
                X[c * py + a] = u[it];
                Y[c * py + a] = v[it];
            }
        }
    }
    __syncthreads();
    stfft::lds_fft<kStElems, NT>(X, py, 2 * C, Pf, -1);
    {
        const int azb = az[b], cnt = C * W;
        float2 v[IT];
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = threadIdx.x + it * NT;
            v[it] = make_float2(0.f, 0.f);
            if (t < cnt) {
                const int c = stfft::fdiv(t, dW), n = t - c * W;
                const int k = n < h ? n : (n >= W - h ? n - W : (1 << 20));
                if (k != (1 << 20)) {
                    int ka = k + wa + azb;  // in [0, 3wa)
                    ka -= ka >= wa ? wa : 0;
                    ka -= ka >= wa ? wa : 0;
                    v[it] = cmulconj(X[c * py + ka], Y[c * py + ka]);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = threadIdx.x + it * NT;
            if (t < cnt) {
                const int c = stfft::fdiv(t, dW), n = t - c * W;
                X[c * px + n] = v[it];
            }
        }
        __syncthreads();
    }
    stfft::lds_fft<kStElems, NT>(X, px, C, Pi, +1);
    float2 *Tb = T + (size_t)b * W * wr + c0;
    const int cnt = W << lc;
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int t = threadIdx.x + it * NT, n = t >> lc, c = t & (C - 1);
        if (t < cnt && c < nc)
            Tb[(size_t)n * wr + c] = X[c * px + n];
    }
}

// Range half: R rows (of 2wa) of one patch per workgroup: the wr range
// frequencies of T spread over 2wr (Nyquist on the negative side, zeros
// between), inverse transform, |C| into the fftshifted 2wa x 2wr surface M,
// and this block's top two in findCPeak's scan-rank order
// (speckleTrack.c:744-805).
__global__ void __launch_bounds__(NT) k_cpx_rowmag(const float2 *T, int wa, int wr, int R, float *mag, Top2 *part,
                                                    stfft::PlanArgs Pi, stfft::FastDiv dW)
{
    extern __shared__ float2 ldsx[];
    __shared__ Top2 sht[NT / 64];
    constexpr int IT = kStElems / NT;
    const int WA = 2 * wa, WR = 2 * wr, h = wr / 2, HA = wa, HR = wr;
    const int b = blockIdx.y, n0 = blockIdx.x * R, nrow = imin(R, WA - n0);
    const float2 *Tb = T + ((size_t)b * WA + n0) * wr;
    const int cnt = nrow * WR;
    {
        float2 v[IT];
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = threadIdx.x + it * NT;
            v[it] = make_float2(0.f, 0.f);
            if (t < cnt) {
                const int i = stfft::fdiv(t, dW), m = t - i * WR;
                const int la = m < h ? m : (m >= WR - h ? m - wr : -1);
                if (la >= 0)
                    v[it] = Tb[(size_t)i * wr + la];
            }
        }
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = threadIdx.x + it * NT;
            if (t < cnt)
                ldsx[t] = v[it];
        }
    }
    __syncthreads();
    stfft::lds_fft<kStElems, NT>(ldsx, WR, nrow, Pi, +1);
    float *M = mag + (size_t)b * WA * WR;
    Top2 best{0.0f, 0x7fffffff, 0.0f, 0.0f};
    for (int t = threadIdx.x; t < cnt; t += NT) {
#pragma clang fp contract(off)
        const int i = stfft::fdiv(t, dW), m = t - i * WR, n = n0 + i;
        const float2 z = ldsx[t];
        const float x2 = z.x * z.x + z.y * z.y;
        const float v = sqrtf(x2);
        const int I = n < HA ? n + HA : n - HA, J = m < HR ? m + HR : m - HR;
        M[(size_t)I * WR + J] = v;
        const int i0 = I < HA ? I : I - HA, j0 = J < HR ? J : J - HR;
        const int q = (I < HA) ? (J < HR ? 0 : 3) : (J < HR ? 2 : 1);
        const int rank = (i0 * HR + j0) * 4 + q;
        if (v > 0.0f)
            best = top2_combine(best, Top2{v, rank, 0.0f, v});
        else
            best.v2 = fmaxf(best.v2, v);
    }
    best = wg_top2<NT>(best, sht);
    if (threadIdx.x == 0)
        part[(size_t)b * gridDim.x + blockIdx.x] = best;
}

// Top two over the row blocks of each patch (in block order), then the
// refinement and the Gaussian-fit samples (peak_finish).
template <bool GAUSS>
__global__ void __launch_bounds__(NT) k_peak_merge(const Top2 *part, int nPart, const float *mag, int WA, int WR,
                                                    int mode, PeakRes *res, GaussArg G)
{
    const int b = blockIdx.x, HA = WA / 2, HR = WR / 2;
    Top2 best = part[(size_t)b * nPart];
    for (int k = 1; k < nPart; k++)
        best = top2_combine(best, part[(size_t)b * nPart + k]);
    int I = -1000, J = -1000;
    float raw = 0.f;
    if (best.r1 != 0x7fffffff) {
        const int q = best.r1 & 3, ij = best.r1 >> 2;
        const int i = ij / HR, j = ij - (ij / HR) * HR;
        I = (q == 0 || q == 3) ? i : i + HA;
        J = (q == 0 || q == 2) ? j : j + HR;
        raw = best.v1;
    }
    const float snr = best.sum > 0.0f ? best.v1 * (float)(WA * WR) / best.sum : 0.0f;
    peak_finish<GAUSS>(mag + (size_t)b * WA * WR, WA, WR, mode, I, J, raw, best.v2, snr, res + b, G, b);
}

// ---------------------------------------------------------------------
// Amplitude correlation (getAmpPatches speckleTrack.c:1006-1189, ampMatch
// :398-484) in five passes that never form a padded array: after the range
// transforms of the two ws x ws patches,
//   k_amp_colx    azimuth transforms, Doppler rotation, pad to W = 2 ws,
//                 inverse azimuth transforms               -> T [2][B][W][ws]
//   k_amp_rowdet  range pad, inverse range transforms, detection |z|^2,
//                 amplitude, partial sums of the zero test and the means
//                                                          -> amplitudes [2][B][W][W]
//   k_amp_stat    zero test and means from the partials (one wave per patch)
//   k_amp_rowfwd  mean removal, partial powers, forward range transforms
//                 of v1 + i v2                             -> Z [B][W][W]
//   k_amp_colx2   forward azimuth transforms, separation of the two
//                 spectra, F1 conj(F2), inverse azimuth transforms of the
//                 columns l <= W/2 (the correlation is real)
//                                                          -> T [B][W][W/2 + 1 of W]
//   k_amp_rowmag  inverse range transforms two rows at a time, |s|,
//                 coarse top two
// so each patch moves about a quarter of the bytes of the padded form.

// partial sums of one k_amp_rowdet workgroup
struct AmpPart {
    double s1, s2, m1, m2;
    int nz, pad;
};

// workgroup -> (item, block) with neighbouring blocks of one item on one
// XCD (dispatch deals consecutive workgroups round-robin over 8 XCDs)
__device__ inline void xcd_item_block(int nbx, int &item, int &blk)
{
    const int total = nbx * gridDim.y, flat = blockIdx.y * nbx + blockIdx.x;
    const int per = total / 8;
    const int L = flat < per * 8 ? (flat % 8) * per + flat / 8 : flat;
    item = L / nbx;
    blk = L - item * nbx;
}

// zeroPadFFT's azimuth half (speckleTrack.c:2014-2038): C adjacent range
// frequencies la (of wsR) of one image's row spectra per workgroup, each
// column wsA long, padded to WA = 2 wsA; item z = image * B + patch.
__global__ void __launch_bounds__(NT) k_amp_colx(const float2 *I, int B, int wsA, int wsR, const int *az,
                                                  const uint8_t *state, float2 *T, int lc, stfft::PlanArgs Pf,
                                                  stfft::PlanArgs Pi, stfft::FastDiv dW)
{
    extern __shared__ float2 ldsx[];
    constexpr int IT = kStElems / NT;
    const int C = 1 << lc, W = 2 * wsA, h = wsA / 2;
    int z, cb;
    xcd_item_block(gridDim.x, z, cb);
    const int b = z % B, c0 = cb * C;
    if (state[b] != 2)
        return;
    const int nc = imin(C, wsR - c0), px = C > 1 ? W + 1 : W;
    float2 *X = ldsx;  // [C][px]
    const float2 *Ib = I + (size_t)z * wsA * wsR + c0;
    {
        const int cnt = wsA << lc;
        float2 u[IT];
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = threadIdx.x + it * NT, a = t >> lc, c = t & (C - 1);
            u[it] = t < cnt && c < nc ? Ib[(size_t)a * wsR + c] : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = threadIdx.x + it * NT, a = t >> lc, c = t & (C - 1);
            if (t < cnt)
                X[c * px + a] = u[it];
        }
    }
    __syncthreads();
    stfft::lds_fft<kStElems, NT>(X, px, C, Pf, -1);
    {
        const int azb = az[b], cnt = C * W;
        float2 v[IT];
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = threadIdx.x + it * NT;
            v[it] = make_float2(0.f, 0.f);
            if (t < cnt) {
                const int c = stfft::fdiv(t, dW), n = t - c * W;
                const int k = n < h ? n : (n >= W - h ? n - W : (1 << 20));
                if (k != (1 << 20))
                    v[it] = X[c * px + (k + wsA + azb) % wsA];
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = threadIdx.x + it * NT;
            if (t < cnt) {
                const int c = stfft::fdiv(t, dW), n = t - c * W;
                X[c * px + n] = v[it];
            }
        }
        __syncthreads();
    }
    stfft::lds_fft<kStElems, NT>(X, px, C, Pi, +1);
    float2 *Tb = T + (size_t)z * W * wsR + c0;
    const int cnt = W << lc;
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int t = threadIdx.x + it * NT, n = t >> lc, c = t & (C - 1);
        if (t < cnt && c < nc)
            Tb[(size_t)n * wsR + c] = X[c * px + n];
    }
}

// zeroPadFFT's range half, then detection and amplitude
// (speckleTrack.c:1166-1188): R rows (of WA = 2 wsA) of both images of one
// patch per workgroup, each row's wsR frequencies padded to WR = 2 wsR (the
// zero test counts cells where either image is below the threshold).
__global__ void __launch_bounds__(NT) k_amp_rowdet(KPar P, const float2 *T, int B, int wsA, int wsR, int R,
                                                    const uint8_t *state, float *amp, AmpPart *part,
                                                    stfft::PlanArgs Pi, stfft::FastDiv dW)
{
    extern __shared__ float2 ldsx[];
    __shared__ double sh[4 * (NT / 64)];
    __shared__ int shi[NT / 64];
    constexpr int IT = kStElems / NT;
    const int WA = 2 * wsA, W = 2 * wsR, h = wsR / 2;
    int b, blk;
    xcd_item_block(gridDim.x, b, blk);
    if (state[b] != 2)
        return;
    const int n0 = blk * R, nrow = imin(R, WA - n0), cnt = nrow * W;
    // the second image's rows follow the first's, so one transform can take
    // both (2 nrow sequences) where they fit the register staging
    float2 *X1 = ldsx, *X2 = ldsx + (size_t)nrow * W;
    const float2 *T1 = T + ((size_t)b * WA + n0) * wsR, *T2 = T + (((size_t)B + b) * WA + n0) * wsR;
    for (int t0 = 0; t0 < cnt; t0 += kStElems) {
        float2 u[IT], v[IT];
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = t0 + threadIdx.x + it * NT;
            u[it] = v[it] = make_float2(0.f, 0.f);
            if (t < cnt) {
                const int i = stfft::fdiv(t, dW), m = t - i * W;
                const int la = m < h ? m : (m >= W - h ? m - wsR : -1);
                if (la >= 0) {
                    u[it] = T1[(size_t)i * wsR + la];
                    v[it] = T2[(size_t)i * wsR + la];
                }
            }
        }
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = t0 + threadIdx.x + it * NT;
            if (t < cnt) {
                X1[t] = u[it];
                X2[t] = v[it];
            }
        }
    }
    __syncthreads();
    if (2 * cnt <= kStElems) {
        stfft::lds_fft<kStElems, NT>(X1, W, 2 * nrow, Pi, +1);
    } else {
        stfft::lds_fft<kStElems, NT>(X1, W, nrow, Pi, +1);
        stfft::lds_fft<kStElems, NT>(X2, W, nrow, Pi, +1);
    }
    const size_t swR = (size_t)wsR, swA = (size_t)wsA;
    const float scale = (float)(1. / (double)(swR * swA * swR * swA * 8));
    const float thr = P.im.stype == ST_SAMPLE_C64 ? 1.0e-15f : 1.0e-6f;
    float *A1 = amp + ((size_t)b * WA + n0) * W, *A2 = amp + (((size_t)B + b) * WA + n0) * W;
    int nz = 0;
    double s1 = 0, s2 = 0, m1 = 0, m2 = 0;
    for (int t = threadIdx.x; t < cnt; t += NT) {
#pragma clang fp contract(off)
        const float2 u = X1[t], v = X2[t];
        const float x1 = (u.x * u.x + u.y * u.y) * scale, x2 = (v.x * v.x + v.y * v.y) * scale;
        nz += (x1 < thr || x2 < thr);
        s1 += x1;
        s2 += x2;
        // the reference's sqrt in double, rounded to float: sqrtf gives the same
        // bits for every non-negative finite float (all of them checked on
        // gfx950), without the FP64 square-root sequence
        const float a1 = sqrtf(x1), a2 = sqrtf(x2);
        m1 += a1;
        m2 += a2;
        A1[t] = a1;
        A2[t] = a2;
    }
    // the five sums in one reduction: wave sums, then the first thread over
    // the waves (one barrier)
    for (int o = 32; o > 0; o >>= 1) {
        s1 += __shfl_xor(s1, o);
        s2 += __shfl_xor(s2, o);
        m1 += __shfl_xor(m1, o);
        m2 += __shfl_xor(m2, o);
        nz += __shfl_xor(nz, o);
    }
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[4 * wid] = s1;
        sh[4 * wid + 1] = s2;
        sh[4 * wid + 2] = m1;
        sh[4 * wid + 3] = m2;
        shi[wid] = nz;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        AmpPart a{sh[0], sh[1], sh[2], sh[3], shi[0], 0};
        for (int w = 1; w < NT / 64; w++) {
            a.s1 += sh[4 * w];
            a.s2 += sh[4 * w + 1];
            a.m1 += sh[4 * w + 2];
            a.m2 += sh[4 * w + 3];
            a.nz += shi[w];
        }
        part[(size_t)b * gridDim.x + blk] = a;
    }
}

// The zero test and the means of ampMatch (speckleTrack.c:437-467) from
// the detection partials: one wave per patch, lanes over the partials, then
// a fixed-order wave sum; a rejected patch is marked "no data" (state 1).
__global__ void __launch_bounds__(64) k_amp_stat(const AmpPart *dpart, int nPd, int WA, int WR, uint8_t *state,
                                                 double2 *mean)
{
    const int b = blockIdx.x, l = threadIdx.x;
    if (state[b] != 2)
        return;
    double s1 = 0, s2 = 0, m1 = 0, m2 = 0;
    int nz = 0;
    for (int k = l; k < nPd; k += 64) {
        const AmpPart q = dpart[(size_t)b * nPd + k];
        s1 += q.s1;
        s2 += q.s2;
        m1 += q.m1;
        m2 += q.m2;
        nz += q.nz;
    }
    for (int o = 32; o > 0; o >>= 1) {
        s1 += __shfl_xor(s1, o);
        s2 += __shfl_xor(s2, o);
        m1 += __shfl_xor(m1, o);
        m2 += __shfl_xor(m2, o);
        nz += __shfl_xor(nz, o);
    }
    if (l == 0) {
        if (nz > 20 || s1 < 1.0e-6 || s2 < 1.0e-6)
            state[b] = 1;
        else
            mean[b] = make_double2(m1 / (double)((size_t)WA * WR), m2 / (double)((size_t)WA * WR));
    }
}

// Mean removal (ampMatch's rules) and the forward range transforms of the
// correlation.  Both images are real, so one complex transform of
// v1 + i v2 carries both spectra (k_amp_colx2 separates them); out
// Z [B][WA][W] and this block's partial powers.
__global__ void __launch_bounds__(NT) k_amp_rowfwd(const float *amp, int B, int WA, int W, int R,
                                                    const double2 *mean, const uint8_t *state, float2 *Z,
                                                    double2 *ppart, stfft::PlanArgs Pf)
{
    extern __shared__ float2 ldsx[];
    __shared__ double sh[2 * (NT / 64)];
    int b, blk;
    xcd_item_block(gridDim.x, b, blk);
    if (state[b] != 2)
        return;
    const double m1 = mean[b].x, m2 = mean[b].y;
    const int n0 = blk * R, nrow = imin(R, WA - n0), cnt = nrow * W;
    float2 *X = ldsx;
    const float *A1 = amp + ((size_t)b * WA + n0) * W, *A2 = amp + (((size_t)B + b) * WA + n0) * W;
    constexpr int IT = kStElems / NT;
    double p1 = 0, p2 = 0;
    for (int t0 = 0; t0 < cnt; t0 += kStElems) {
        float u[IT], v[IT];
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = t0 + threadIdx.x + it * NT;
            u[it] = t < cnt ? A1[t] : 0.f;
            v[it] = t < cnt ? A2[t] : 0.f;
        }
#pragma unroll
        for (int it = 0; it < IT; it++) {
#pragma clang fp contract(off)
            const int t = t0 + threadIdx.x + it * NT;
            if (t < cnt) {
                float v1 = u[it], v2 = v[it];
                if (v1 > 1.0e-20)
                    v1 = (float)(v1 - m1);
                v2 = (float)(v2 - m2);
                p1 += v1 * v1;
                p2 += v2 * v2;
                X[t] = make_float2(v1, v2);
            }
        }
    }
    // partial powers: wave sums, then the first thread over the waves; the
    // one barrier also orders the LDS stores before the transform
    for (int o = 32; o > 0; o >>= 1) {
        p1 += __shfl_xor(p1, o);
        p2 += __shfl_xor(p2, o);
    }
    const int wid = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        sh[2 * wid] = p1;
        sh[2 * wid + 1] = p2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double q1 = sh[0], q2 = sh[1];
        for (int w = 1; w < NT / 64; w++) {
            q1 += sh[2 * w];
            q2 += sh[2 * w + 1];
        }
        ppart[(size_t)b * gridDim.x + blk] = make_double2(q1, q2);
    }
    stfft::lds_fft<kStElems, NT>(X, W, nrow, Pf, -1);
    float2 *Zb = Z + ((size_t)b * WA + n0) * W;
    for (int t = threadIdx.x; t < cnt; t += NT)
        Zb[t] = X[t];
}

// The two spectra of Z = F(v1) + i F(v2) at (k, l) from Z(k, l) and
// Z(-k, -l), and their cross spectrum F1 conj(F2) (ampMatch's product).
__device__ inline float2 amp_cross_sep(float2 za, float2 zb)
{
    const float2 x1 = make_float2(0.5f * (za.x + zb.x), 0.5f * (za.y - zb.y));  // (za + conj zb) / 2
    const float2 x2 = make_float2(0.5f * (za.y + zb.y), 0.5f * (zb.x - za.x));  // (za - conj zb) / 2i
    return cmulconj(x1, x2);
}

// The correlation's azimuth half: slot c of a workgroup holds column
// l = c0 + c (l < WR/2) in X and its mirror -l in Y (column WR/2 for l = 0);
// forward transforms (WA long), separation and cross spectra of both
// columns, inverse transforms; out S [B][WA][WR].
__global__ void __launch_bounds__(NT) k_amp_colx2(const float2 *Z, int WA, int WR, const uint8_t *state, float2 *S,
                                                   int lc, stfft::PlanArgs Pf, stfft::PlanArgs Pi, stfft::FastDiv dW)
{
    extern __shared__ float2 ldsx[];
    constexpr int IT = kStElems / NT;
    const int C = 1 << lc, H = WR / 2, W = WA;
    int b, cb;
    xcd_item_block(gridDim.x, b, cb);
    if (state[b] != 2)
        return;
    const int c0 = cb * C, nc = imin(C, H - c0), px = C > 1 ? W + 1 : W;
    float2 *X = ldsx, *Y = ldsx + (size_t)C * px;
    const float2 *Zb = Z + (size_t)b * WA * WR;
    const int cnt = W << lc;
    for (int t0 = 0; t0 < cnt; t0 += kStElems) {
        float2 u[IT], v[IT];
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = t0 + threadIdx.x + it * NT, a = t >> lc, c = t & (C - 1), l = c0 + c;
            const bool ok = t < cnt && c < nc;
            u[it] = ok ? Zb[(size_t)a * WR + l] : make_float2(0.f, 0.f);
            v[it] = ok ? Zb[(size_t)a * WR + (l == 0 ? H : WR - l)] : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = t0 + threadIdx.x + it * NT, a = t >> lc, c = t & (C - 1);
            if (t < cnt) {
                X[c * px + a] = u[it];
                Y[c * px + a] = v[it];
            }
        }
    }
    __syncthreads();
    // Y follows X at the same pitch: one transform of 2C sequences where
    // they fit the register staging
    if (2 * C * W <= kStElems) {
        stfft::lds_fft<kStElems, NT>(X, px, 2 * C, Pf, -1);
    } else {
        stfft::lds_fft<kStElems, NT>(X, px, C, Pf, -1);
        stfft::lds_fft<kStElems, NT>(Y, px, C, Pf, -1);
    }
    for (int t0 = 0; t0 < cnt; t0 += kStElems) {
        float2 u[IT], v[IT];
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = t0 + threadIdx.x + it * NT;
            u[it] = v[it] = make_float2(0.f, 0.f);
            if (t < cnt) {
                const int c = stfft::fdiv(t, dW), k = t - c * W, km = k == 0 ? 0 : W - k;
                if (c0 + c == 0) {  // columns 0 and WR/2 are their own mirrors
                    u[it] = amp_cross_sep(X[k], X[km]);
                    v[it] = amp_cross_sep(Y[k], Y[km]);
                } else {
                    u[it] = amp_cross_sep(X[c * px + k], Y[c * px + km]);
                    v[it] = amp_cross_sep(Y[c * px + k], X[c * px + km]);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = t0 + threadIdx.x + it * NT;
            if (t < cnt) {
                const int c = stfft::fdiv(t, dW), k = t - c * W;
                X[c * px + k] = u[it];
                Y[c * px + k] = v[it];
            }
        }
        __syncthreads();
    }
    // s = F^-1(F1 conj(F2)) is real, so T(n, -l) = conj T(n, l): only the
    // columns l <= WR/2 are transformed back and stored
    if (c0 != 0) {
        stfft::lds_fft<kStElems, NT>(X, px, C, Pi, +1);
    } else if ((C + 1) * W <= kStElems) {
        stfft::lds_fft<kStElems, NT>(X, px, C + 1, Pi, +1);  // Y's slot 0 (column WR/2) is sequence C
    } else {
        stfft::lds_fft<kStElems, NT>(X, px, C, Pi, +1);
        stfft::lds_fft<kStElems, NT>(Y, px, 1, Pi, +1);
    }
    float2 *Sb = S + (size_t)b * WA * WR;
    for (int t = threadIdx.x; t < cnt; t += NT) {
        const int n = t >> lc, c = t & (C - 1);
        if (c < nc)
            Sb[(size_t)n * WR + c0 + c] = X[c * px + n];
    }
    if (c0 == 0)
        for (int n = threadIdx.x; n < W; n += NT)
            Sb[(size_t)n * WR + H] = Y[n];
}

// The correlation's range half: from the stored columns l <= W/2 of T
// (WA rows of W = WR), two rows per complex inverse transform (row n0 + 2q in the real part,
// n0 + 2q + 1 in the imaginary part, each transform real), then |s| into
// the fftshifted surface and this block's top two, as k_cpx_rowmag.
__global__ void __launch_bounds__(NT) k_amp_rowmag(const float2 *T, int WA, int W, int R, float *mag, Top2 *part,
                                                    stfft::PlanArgs Pi, stfft::FastDiv dW, const uint8_t *state)
{
    extern __shared__ float2 ldsx[];
    __shared__ Top2 sht[NT / 64];
    constexpr int IT = kStElems / NT;
    const int H = W / 2, HA = WA / 2;
    const int b = blockIdx.y, n0 = blockIdx.x * R, nrow = imin(R, WA - n0), npair = (nrow + 1) / 2;
    if (state[b] != 2) {
        if (threadIdx.x == 0)
            part[(size_t)b * gridDim.x + blockIdx.x] = Top2{0.0f, 0x7fffffff, 0.0f, 0.0f};
        return;
    }
    const float2 *Tb = T + ((size_t)b * WA + n0) * W;
    const int cnt = npair * W;
    {
        float2 v[IT];
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = threadIdx.x + it * NT;
            v[it] = make_float2(0.f, 0.f);
            if (t < cnt) {
                const int q = stfft::fdiv(t, dW), m = t - q * W, la = m <= H ? m : W - m;
                float2 a = Tb[(size_t)(2 * q) * W + la];
                float2 c = 2 * q + 1 < nrow ? Tb[(size_t)(2 * q + 1) * W + la] : make_float2(0.f, 0.f);
                if (m > H) {
                    a.y = -a.y;
                    c.y = -c.y;
                }
                v[it] = make_float2(a.x - c.y, a.y + c.x);
            }
        }
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = threadIdx.x + it * NT;
            if (t < cnt)
                ldsx[t] = v[it];
        }
    }
    __syncthreads();
    stfft::lds_fft<kStElems, NT>(ldsx, W, npair, Pi, +1);
    float *M = mag + (size_t)b * WA * W;
    Top2 best{0.0f, 0x7fffffff, 0.0f, 0.0f};
    for (int t = threadIdx.x; t < nrow * W; t += NT) {
        const int i = stfft::fdiv(t, dW), m = t - i * W, n = n0 + i;
        const float2 z = ldsx[(i >> 1) * W + m];
        const float v = fabsf((i & 1) ? z.y : z.x);
        const int I = n < HA ? n + HA : n - HA, J = m < H ? m + H : m - H;
        M[(size_t)I * W + J] = v;
        const int i0 = I < HA ? I : I - HA, j0 = J < H ? J : J - H;
        const int q = (I < HA) ? (J < H ? 0 : 3) : (J < H ? 2 : 1);
        const int rank = (i0 * H + j0) * 4 + q;
        if (v > 0.0f)
            best = top2_combine(best, Top2{v, rank, 0.0f, v});
        else
            best.v2 = fmaxf(best.v2, v);
    }
    best = wg_top2<NT>(best, sht);
    if (threadIdx.x == 0)
        part[(size_t)b * gridDim.x + blockIdx.x] = best;
}

// The mean powers of ampMatch from the partials, in block order.
__global__ void k_amp_psum(const double2 *ppart, int nPf, int B, int WA, int WR, const uint8_t *state, double *p)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B || state[b] != 2)
        return;
    double p1 = 0, p2 = 0;
    for (int k = 0; k < nPf; k++) {
        p1 += ppart[(size_t)b * nPf + k].x;
        p2 += ppart[(size_t)b * nPf + k].y;
    }
    const double N = (double)((size_t)WA * WR);
    p[2 * b] = p1 / N;
    p[2 * b + 1] = p2 / N;
}

// Sub-pixel refinement and Gaussian-fit samples of the fused path's
// attempts (overSampleC and the central-box argmax, speckleTrack.c:686-700,
// 852-863; the samples of complexMatch's fits, :377-386), from the block
// cells k_cpx_fused left around each coarse peak.  Four attempts per
// workgroup of 192 threads: thread 48 p + v owns column v of attempt p's
// 48 x 48 search box (refine_cols), so every lane of the three waves works
// (one attempt per wave would leave 16 lanes idle); the Dirichlet rows come
// as scalars, the 8 x 8 boxes (times b) through LDS, and wave 0 merges each
// attempt's 48 column results and writes them.
constexpr int kRefPer = 4, kRefThreads = kRefPer * kFine;
__global__ void __launch_bounds__(kRefThreads) k_cpx_refine(CpxBatch C, int count, int WW)
{
    __shared__ double2 Pc[kRefPer][NFAST * NFAST];
    __shared__ double s1[kRefThreads], s2[kRefThreads];
    __shared__ int sr[kRefThreads];
    const int tid = threadIdx.x, base = blockIdx.x * kRefPer;
    // an attempt is refined when the fused kernel accepted it and found a coarse peak
    auto live = [&](int it) { return it < count && C.have[it] && C.res[it].I >= -999; };
    for (int e2 = tid; e2 < kRefPer * NFAST * NFAST; e2 += kRefThreads) {
        const int pp = e2 / (NFAST * NFAST), e = e2 % (NFAST * NFAST), it = base + pp;
        double2 c = make_double2(0.0, 0.0);
        if (live(it)) {
            const int m = e / NFAST, n = e % NFAST;
            const double pm = (double)C.nb[(size_t)it * kNbc +
                                           kNbMap.off[(kNb / 2 - NFAST / 2 + m) * kNb + (kNb / 2 - NFAST / 2 + n)]];
            const double2 b = c_refB[m + n];
            c = make_double2(b.x * pm, b.y * pm);
        }
        Pc[pp][e] = c;
    }
    __syncthreads();
    {
        const int p = tid / kFine, v = tid % kFine;
        double f1, f2;
        int fr;
        refine_cols([&](int e) { return Pc[p][e]; }, v, f1, f2, fr);
        s1[tid] = f1;
        s2[tid] = f2;
        sr[tid] = fr;
    }
    __syncthreads();
    if (tid >= 64)
        return;
    // wave 0: 16 lanes per attempt, three columns each, then a butterfly over the 16
    const int pp = tid >> 4, l = tid & 15, it = base + pp;
    double f1 = 0.0, f2 = 0.0;
    int fr = 0x7fffffff;
#pragma unroll
    for (int k = 0; k < kFine / 16; k++) {
        const int c = pp * kFine + l * (kFine / 16) + k;
        top2_merge(f1, f2, fr, s1[c], s2[c], sr[c]);
    }
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
        const double b1 = __shfl_xor(f1, o), b2 = __shfl_xor(f2, o);
        const int br = __shfl_xor(fr, o);
        top2_merge(f1, f2, fr, b1, b2, br);
    }
    if (!live(it))
        return;
    PeakRes *res = &C.res[it];
    const int I = res->I, J = res->J;
    const float *nb = C.nb + (size_t)it * kNbc;
    int fi = 0, fj = 0;
    if (fr != 0x7fffffff) {
        fi = fr / kFine;
        fj = fr - fi * kFine;
    }
    const int iF = I * NOVER + (fi + 24) - (NOVER * NFAST) / 2;
    const int jF = J * NOVER + (fj + 24) - (NOVER * NFAST) / 2;
    const int cI = iF / NOVER, cJ = jF / NOVER;
    const bool ok = (cI - NGAUSS / 2) >= 0 && (cI + NGAUSS / 2) < WW && (cJ - NGAUSS / 2) >= 0 &&
                    (cJ + NGAUSS / 2) < WW;
    if (l == 0) {
        res->fi = fi;
        res->fj = fj;
        res->fineMax = fr != 0x7fffffff ? sqrt(f1) : 0.0;
        res->fineGap = fr != 0x7fffffff ? (float)((sqrt(f1) - sqrt(f2)) / sqrt(f1)) : 0.f;
        res->lmOk = ok ? 1 : 0;
        res->lmX0[0] = (float)(cJ - NGAUSS / 2);
        res->lmX0[1] = (float)(cI - NGAUSS / 2);
    }
    if (ok) {
        // cI, cJ within 2 samples of (I, J): inside the block
        const int bi = cI - I + kNb / 2, bj = cJ - J + kNb / 2;
        for (int q = l; q < 2 * NGAUSS; q += 16) {
            const int k = q % NGAUSS;
            res->lmY[q / NGAUSS][k] = q < NGAUSS ? nb[kNbMap.off[bi * kNb + (bj - NGAUSS / 2 + k)]]
                                                 : nb[kNbMap.off[(bi - NGAUSS / 2 + k) * kNb + bj]];
        }
    }
}

// Lanes per transform of the fused kernel.  Lane pairs at every window
// size: lane quads at w = 64 (16 points per lane, 512-thread blocks) were
// measured slower on the full-scene job (55.6 ms vs 43.6 ms per launch) --
// the second DPP exchange stage, the per-lane twiddle selection and the
// transposes' bank conflicts cost more than the occupancy gained.
template <int W>
struct FusedLanes {
    static constexpr int L = 2;
    // LDS transpose buffers [W][LD] float2: LD = 2 (mod 32) spreads the 16
    // rows x 2 lanes of a half-wave row access over all banks; for w >= 32
    // the rows of the second half start 16 elements later so that the
    // column access of a lane pair (rows k and w/2 + k, 256-byte aligned
    // apart) falls into the other half of the banks (at w = 16 the pair is
    // already 128 bytes apart).
    static constexpr int LD = W <= 16 ? 34 : W + 2;
    static constexpr int SH = W >= 32 ? 16 : 0;
    static constexpr int BUF = W * LD + SH;
    __device__ static inline int at(int row, int col) { return row * LD + col + (row >= W / 2 ? SH : 0); }
};

// BASE bit 0 adds the baseline phase ramp on patch 1, bit 1 the
// interferogram correction (separate instantiations: their FP64 sin/cos
// would otherwise raise the register allocation of every launch).  OCC is the minimum number of waves per
// SIMD the register allocation must allow: 2 (at w = 64 this caps the
// allocation at 256 VGPRs with a few spilled values instead of one wave
// per SIMD).
//
// One launch serves the complex attempts of njobs tracking calls (pairs
// and grids of their own, one window class): items [jobStart[j],
// jobStart[j + 1]) are job j's, and every item reads its job's parameter,
// point-state and batch blocks.
template <int W, int STYPE, int BASE, int OCC>
__global__ void __launch_bounds__(2 * FusedLanes<W>::L * W) __attribute__((amdgpu_waves_per_eu(OCC, 8))) k_cpx_fused(const KPar *__restrict__ Pp, const Pts *__restrict__ Sp,
                                                     const int *__restrict__ jobStart, int njobs,
                                                     const CpxBatch *__restrict__ Cp)
{
    const int count = jobStart[njobs];
    using namespace ::fftreg;
    constexpr int L = FusedLanes<W>::L, H = W / L, NTH = 2 * L * W, WW = 2 * W, BUF = FusedLanes<W>::BUF;
    using FL = FusedLanes<W>;
    using LF = LaneFFT<W, L>;
    extern __shared__ float2 lds[];  // [2][BUF]
    __shared__ Top2 sht[16];
    // window sums: one partial per group of SPAN lanes (NP / 2 per window),
    // near-zero counts per wave, the sample masks, the acceptance result
    constexpr int SPAN = 2 * W < 64 ? 2 * W : 64, NP = NTH / SPAN;
    __shared__ double stp[NP][9];
    __shared__ int stz[NTH / 64], sflag[2];
    __shared__ unsigned smask[NTH];
    __shared__ int srank;  // rank of the surface maximum (atomicMin over its holders)
    constexpr bool isf = STYPE == ST_SAMPLE_C64;
    const float thr = isf ? 1e-15f : 1e-9f;

    // Items in XCD-contiguous order: at every round the workgroups of one
    // XCD (dealt round-robin, b % 8) take a contiguous run of items, so the
    // overlapping patches of neighbouring points share that XCD's L2.
    const int G = gridDim.x, Q = G / 8;
    const int base = (G % 8 == 0) ? (blockIdx.x % 8) * Q + blockIdx.x / 8 : blockIdx.x;
    for (int k0 = 0; k0 < count; k0 += G) {
        const int itg = k0 + base;
        if (itg >= count)
            break;
        // the item's job (workgroup-uniform: a scalar search), its index there
        int jb = 0;
        for (int lo = 0, hi = njobs - 1; lo < hi;) {
            const int mid = (lo + hi + 1) >> 1;
            if (jobStart[mid] <= itg)
                lo = jb = mid;
            else
                hi = mid - 1;
        }
        const int it = itg - jobStart[jb];
        const KPar &P = Pp[jb];
        const Pts &S = Sp[jb];
        const CpxBatch &C = Cp[jb];
        // Lane indices are re-derived behind an opaque copy every item so
        // the compiler does not hoist per-lane addresses out of the work
        // loop (they would stay live in registers for the whole kernel).
        int t = threadIdx.x;
        asm volatile("" : "+v"(t));
        const int g = t / (L * W);
        const int r = (t / L) % W;
        const int h = t & (L - 1);
        const int blk = LF::blk(h);
        const int pt = C.list[it];
        const int r1 = S.r1[pt], a1 = S.a1[pt], r2 = S.r2[pt], a2 = S.a2[pt];
        // ---------------- load column r of patch g: samples a = 2m + h
        float2 x[H];
        const void *img = g ? P.im.img2 : P.im.img1;
        const int nr = g ? P.im.nr2 : P.im.nr1;
        const int ra = (g ? r2 : r1) + r, aa = (g ? a2 : a1) + h;
        const bool useH = (g == 1) && P.hanning;
#pragma unroll
        for (int m = 0; m < H; m++) {
            const size_t o = (size_t)(aa + L * m) * nr + ra;
            if constexpr (isf) {
                x[m] = __builtin_bit_cast(float2, ((const gu64 *)img)[o]);
            } else {
                const unsigned q = ((const gu32 *)img)[o];
                x[m] = make_float2((float)(short)(q & 0xffffu), (float)(short)(q >> 16));
            }
        }
#pragma unroll
        for (int m = 0; m < H; m++) {
#pragma clang fp contract(off)
            float2 v = x[m];
            if constexpr (isf) {
                v.x *= 1000;
                v.y *= 1000;
            }
            const float hw = ((const gfloat *)P.hann)[(L * m + h) * W + r];
            v.x = useH ? v.x * hw : v.x;
            v.y = useH ? v.y * hw : v.y;
            x[m] = v;
        }
        // ---------------- window sums (getPatches' statistics and the
        // Doppler lag sum) from the loaded samples, reduced to one partial
        // per SPAN lanes; the pair's acceptance test and shift follow once
        // both windows' near-zero masks are in LDS
        {
            double acc[9];
            unsigned mk;
            lane_sums<H, L>(x, h, thr, acc, mk);
            group_sum_last<SPAN>(acc);
            if ((t & (SPAN - 1)) == SPAN - 1) {
#pragma unroll
                for (int k = 0; k < 9; k++)
                    stp[t / SPAN][k] = acc[k];
            }
            smask[t] = mk;
        }
        __builtin_amdgcn_sched_barrier(0);  // the sums finish before the transforms start
        // ---------------- phase corrections on patch 1 (in LDS)
        if constexpr (BASE != 0) {
            float2 *R = lds + (size_t)g * BUF;
#pragma unroll
            for (int m = 0; m < H; m++)
                R[FL::at(L * m + h, r)] = x[m];
            __syncthreads();
            float2 *R0 = lds;
            if constexpr ((BASE & 1) != 0) {
            const double x0 = (a1 - P.wA / 2 - (P.nSlpA1 / 2)) / (double)P.nSlpA1;
            const double bn = P.Bn + x0 * P.dBn + x0 * x0 * P.dBnQ;
            const double ptod = -(4.0 * 3.14159265358979323846) / P.lambda;
            const int j0 = r1 - P.wR / 2 > 0 ? r1 - P.wR / 2 : 0;
#pragma unroll 1
            for (int e = t; e < W * W; e += NTH) {
                const int i = e / W, j = e - (e / W) * W;
                const double xx = x0 + i * (1. / (double)P.nSlpA1);
                const double bp = P.Bp + xx * P.dBp + xx * xx * P.dBpQ;
                const double bsq = bn * bn + bp * bp;
                double ph = -bn * P.sinT[j0 + j] - bp * P.cosT[j0 + j] + (0.5 * bsq * P.overR[j0 + j]);
                ph *= ptod;
                const double pr = cos(ph), pim = sin(ph);
                const float2 v = R0[FL::at(i, j)];
                const double cr = v.x, ci = v.y;
                R0[FL::at(i, j)] = make_float2((float)(cr * pr - ci * pim), (float)(cr * pim + ci * pr));
            }
            }
            if constexpr ((BASE & 2) != 0) {
                __shared__ IntTab<W> itab;
                int_tables<W>(P, it, W, itab, t, NTH);
#pragma unroll 1
                for (int e = t; e < W * W; e += NTH) {
                    const int i = e / W, j = e - (e / W) * W;
                    const float2 v = R0[FL::at(i, j)];
                    double cr = v.x, ci = v.y;
                    int_apply<W>(P, itab, it, i, j, cr, ci);
                    R0[FL::at(i, j)] = make_float2((float)cr, (float)ci);
                }
            }
            __syncthreads();
#pragma unroll
            for (int m = 0; m < H; m++)
                x[m] = R[FL::at(L * m + h, r)];
            __syncthreads();
        }
        // ---------------- forward FFT along azimuth (column r), then LDS
        LF::fwd(x, h);
        float2 *F = lds + (size_t)g * BUF;
#pragma unroll
        for (int k = 0; k < H; k++)
            F[FL::at(blk * H + k, r)] = x[k];
        __syncthreads();
        {
            // samples near zero in either window (thread t + 2w holds the
            // same samples of window 2)
            int nzl[1] = {g == 0 ? (int)__popc(smask[t] | smask[t + 2 * W]) : 0};
            group_sum_last<64>(nzl);
            if ((t & 63) == 63)
                stz[t >> 6] = nzl[0];
        }
        // ---------------- forward FFT along range (row r)
#pragma unroll
        for (int m = 0; m < H; m++)
            x[m] = F[FL::at(r, L * m + h)];
        __syncthreads();
        if (t == 0) {
            PatchSums ss;
#pragma unroll
            for (int gg = 0; gg < 2; gg++)
#pragma unroll
                for (int k = 0; k < 9; k++) {
                    double a = 0.0;
#pragma unroll
                    for (int q = 0; q < NP / 2; q++)
                        a += stp[gg * (NP / 2) + q][k];
                    ss.v[gg][k] = a;
                }
            int nz = 0;
#pragma unroll
            for (int q = 0; q < NTH / 64; q++)
                nz += stz[q];
            int azv;
            sflag[0] = patch_accept<W, STYPE>(P, C, it, ss, nz, azv);
            sflag[1] = azv;
        }
        LF::fwd(x, h);
#pragma unroll
        for (int k = 0; k < H; k++)
            F[FL::at(r, blk * H + k)] = x[k];
        __syncthreads();
        // a pair that failed the acceptance test (have = 0, az = 0) runs on;
        // nothing downstream reads its results (skipping here costs registers)
        const int az = sflag[1];
        // ---------------- cross spectrum X = F1 conj(F2), row r, half h
        {
            const float2 *Fo = lds + (size_t)(1 - g) * BUF;
#pragma unroll
            for (int k = 0; k < H; k++) {
#pragma clang fp contract(off)
                const float2 o = Fo[FL::at(r, blk * H + k)];
                x[k] = g == 0 ? cmulconj(x[k], o) : cmulconj(o, x[k]);
            }
        }
        __syncthreads();
        // ---------------- polyphase inverse: components (pp, q = g)
        const int krow = ((r - az) % W + W) % W;
        float2 *Z = lds + (size_t)g * BUF;
        Top2 best{0.0f, 0x7fffffff, 0.0f, 0.0f};
        if (t == 0)
            srank = 0x7fffffff;  // the previous item's last reads of it are behind barriers
        // exp(i pi (ks pp + ls q) / w) = row factor (ks: signed frequency of
        // the row, pp) x column factor (per element, q = g only).  The column
        // factor is applied before the row transform; the row factor is
        // constant along a row, so one row transform serves both components
        // and the factor of pp = 1 multiplies its output on the column pass
        // (element k of lane h is row blk H + k there).
        if (g) {
#pragma unroll
            for (int k = 0; k < H; k++) {
                // column l = blk H + k, signed frequency (Nyquist negative)
                float2 cf = c_tw128[(k * (64 / W)) & 127];
#pragma unroll
                for (int b2 = 1; b2 < L; b2++) {
                    const int l = b2 * H + k;
                    const int ls = l < W / 2 ? l : l - W;
                    const float2 cb = c_tw128[(ls * (64 / W)) & 127];
                    cf = blk == b2 ? cb : cf;
                }
                x[k] = cmul_f(x[k], cf);
            }
        }
        LF::inv(x, h);
        // x[n'] = Z[krow][L n' + h]
#pragma unroll
        for (int n = 0; n < H; n++)
            Z[FL::at(krow, L * n + h)] = x[n];
        __syncthreads();
        // surface row n = 2 (L j + h) + pp, column m = 2 r + g, fftshifted
        const int Jc = (2 * r + g + W) & (WW - 1);
        float mg0[H];  // |C| of component pp = 0, kept until the LDS is free
        auto scan = [&](const float2 *y, int pp, float *mg) {
            // |C|^2 first; a wave with no value below 2^-96 (nearly always)
            // takes the short form of the correctly rounded square root
            bool tiny = false;
#pragma unroll
            for (int jj = 0; jj < H; jj++) {
#pragma clang fp contract(off)
                mg[jj] = y[jj].x * y[jj].x + y[jj].y * y[jj].y;
                tiny |= !(mg[jj] >= 0x1p-96f);
            }
            if (__builtin_amdgcn_ballot_w64(tiny) == 0) {
#pragma unroll
                for (int jj = 0; jj < H; jj++)
                    mg[jj] = sqrt_rn_normal(mg[jj]);
            } else {
#pragma unroll
                for (int jj = 0; jj < H; jj++)
                    mg[jj] = sqrtf(mg[jj]);
            }
#pragma unroll
            for (int jj = 0; jj < H; jj++)
                max2_add(best, mg[jj]);
        };
        {
            float2 y[H];
#pragma unroll
            for (int k = 0; k < H; k++)
                y[k] = Z[FL::at(blk * H + k, r)];
            LF::inv(y, h);
            scan(y, 0, mg0);
            // keep the second component's loads behind this scan (hoisted
            // above it, both components' values would be live at once)
            __builtin_amdgcn_sched_barrier(0);
        }
        {
            float2 y[H];
#pragma unroll
            for (int k = 0; k < H; k++) {
                // row blk H + k, signed frequency k - blk H (L = 2)
                float2 rf = c_tw128[(k * (64 / W)) & 127];
#pragma unroll
                for (int b2 = 1; b2 < L; b2++) {
                    const int rho = b2 * H + k;
                    const int rs = rho < W / 2 ? rho : rho - W;
                    const float2 cb = c_tw128[(rs * (64 / W)) & 127];
                    rf = blk == b2 ? cb : rf;
                }
                y[k] = cmul_f(Z[FL::at(blk * H + k, r)], rf);
            }
            __syncthreads();  // every read of Z done: the LDS takes the surface
            LF::inv(y, h);
            float mg1[H];
            scan(y, 1, mg1);
            // the 2w x 2w |C| surface in LDS (16 w^2 bytes <= the two transpose buffers)
            float *Ms = (float *)lds;
#pragma unroll
            for (int jj = 0; jj < H; jj++) {
                const int n = 2 * (L * jj + h);
                Ms[((n + W) & (WW - 1)) * WW + Jc] = mg0[jj];
                Ms[((n + 1 + W) & (WW - 1)) * WW + Jc] = mg1[jj];
            }
        }
        {
            // the maximum's rank: its holders (normally one thread) look it
            // up among their own cells of the LDS surface
            const float mine = best.v1;
            best = wg_max2<NTH>(best, sht);
            if (best.v1 > 0.0f && mine == best.v1) {
                const float *Ms = (const float *)lds;
                int rk = 0x7fffffff;
                for (int e = 0; e < 2 * H; e++) {
                    const int I = (2 * (L * (e >> 1) + h) + (e & 1) + W) & (WW - 1);
                    if (Ms[I * WW + Jc] == best.v1)
                        rk = min(rk, surf_rank<W>(I, Jc));
                }
                atomicMin(&srank, rk);
            }
            __syncthreads();
            best.r1 = srank;
        }
        PeakRes *res = &C.res[it];
        int I = -1000, J = -1000;
        float raw = 0.f;
        if (best.r1 != 0x7fffffff) {
            const int q = best.r1 & 3, ij = best.r1 >> 2;
            const int i0 = ij / W, j0 = ij - (ij / W) * W;
            I = (q == 0 || q == 3) ? i0 : i0 + W;
            J = (q == 0 || q == 2) ? j0 : j0 + W;
            raw = best.v1;
        }
        if (t == 0) {
            res->gap = raw > 0 ? (raw - best.v2) / raw : 0.f;
            res->snr = best.sum > 0.0f ? best.v1 * (float)(WW * WW) / best.sum : 0.0f;
            res->lmOk = 0;
            res->fi = res->fj = -1;
            res->fineMax = 0;
            res->fineGap = 0;
        }
        if (I < NFAST / 2 || I > WW - NFAST / 2 || J < NFAST / 2 || J > WW - NFAST / 2) {
            if (t == 0) {
                res->I = res->J = -1000;
                res->raw = -1.f;
            }
        } else {
            if (t == 0) {
                res->I = I;
                res->J = J;
                res->raw = raw;
            }
            // the |C| block around the peak out of the LDS surface, for
            // k_cpx_refine (cells off the surface are never read: the fits'
            // edge tests reject them)
            const float *Ms = (const float *)lds;
            float *nb = C.nb + (size_t)it * kNbc;
            for (int e = t; e < kNbc; e += NTH) {
                const int ri = I - kNb / 2 + kNbMap.ri[e], ci = J - kNb / 2 + kNbMap.ci[e];
                nb[e] = (ri >= 0 && ri < WW && ci >= 0 && ci < WW) ? Ms[ri * WW + ci] : 0.0f;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------
// Fused complex match at w = 128: one workgroup of 512 threads per patch
// pair (thread t: index r = t / 4, lane quarter h = t & 3), persistent over
// the work list.  A 128 x 128 complex array fills most of a CU's LDS, so
// the two windows go through one transpose buffer in turn: window 1's
// spectrum waits in registers (64 per thread) while window 2's is formed,
// the cross spectrum stays there for both column phases of the polyphase
// inverse, and the 256 x 256 |C| surface goes to a per-workgroup global
// slot (phase-major, so every store fills whole cache lines), from which
// the block around the coarse peak is taken for k_cpx_refine.  Otherwise as
// k_cpx_fused, with each 128-point transform on a lane quad
// (fwd_quad128 / inv_quad128): the components C[2n+p][2m+q], one row
// transform per column phase, findCPeak's rank order.
constexpr int kF128 = 128;
constexpr size_t kF128SlotM = (size_t)4 * kF128 * kF128;  // float per workgroup: |C|

// exp(i pi (blk 32 + k - (blk >= 2 ? 128 : 0)) / 128) = s_blk exp(i pi k / 128):
// the polyphase factor of the signed frequency of element k of block blk
__device__ inline float2 f128_factor(float2 sblk, int k)
{
    return ::fftreg::cmul_f(sblk, make_float2(::fftreg::kTw256Re[k], ::fftreg::kTw256Im[k]));
}

template <int STYPE>
__global__ void __launch_bounds__(512) k_cpx_fused128(const KPar *__restrict__ Pp, const Pts *__restrict__ Sp,
                                                      const int *__restrict__ jobStart, int njobs,
                                                      const CpxBatch *__restrict__ Cp, float *slotM)
{
    constexpr int W = kF128, L = 4, H = W / L, NTH = 512, WW = 2 * W;
    using FL = FusedLanes<W>;
    extern __shared__ float2 lds[];  // [BUF]: the transpose buffer
    __shared__ Top2 sht[NTH / 64];
    constexpr bool isf = STYPE == ST_SAMPLE_C64;
    const int count = jobStart[njobs];
    // items in XCD-contiguous order (as k_cpx_fused)
    const int G = gridDim.x, Q = G / 8;
    const int base = (G % 8 == 0) ? (blockIdx.x % 8) * Q + blockIdx.x / 8 : blockIdx.x;
    float *M = slotM + (size_t)blockIdx.x * kF128SlotM;
    for (int k0 = 0; k0 < count; k0 += G) {
        const int itg = k0 + base;
        if (itg >= count)
            break;
        int jb = 0;
        for (int lo = 0, hi = njobs - 1; lo < hi;) {
            const int mid = (lo + hi + 1) >> 1;
            if (jobStart[mid] <= itg)
                lo = jb = mid;
            else
                hi = mid - 1;
        }
        const int it = itg - jobStart[jb];
        const KPar &P = Pp[jb];
        const Pts &S = Sp[jb];
        const CpxBatch &C = Cp[jb];
        int t = threadIdx.x;
        asm volatile("" : "+v"(t));
        const int r = t >> 2;
        // the item's values are the same on every lane: scalar registers
        const int pt = __builtin_amdgcn_readfirstlane(C.list[it]);
        if (!__builtin_amdgcn_readfirstlane(C.have[it]))
            continue;
        const int r1 = __builtin_amdgcn_readfirstlane(S.r1[pt]), a1 = __builtin_amdgcn_readfirstlane(S.a1[pt]);
        const int r2 = __builtin_amdgcn_readfirstlane(S.r2[pt]), a2 = __builtin_amdgcn_readfirstlane(S.a2[pt]);
        const int az = __builtin_amdgcn_readfirstlane(C.az[it]);
        // ---------------- forward 2-D transforms, window 1 then window 2
        // fwd2d(g, x): window g's spectrum, row r, columns blk H + k in x[k]
        auto fwd2d = [&](int g, float2 (&x)[H]) {
            // per-lane values behind an opaque copy in every body, so the
            // compiler does not hoist the lane-selected twiddles and factors
            // out of it (they would hold registers throughout)
            int h = t & 3;
            asm volatile("" : "+v"(h));
            const int blk = ::fftreg::qblk(h);
            const int nr = g ? P.im.nr2 : P.im.nr1;
            const bool useH = (g == 1) && P.hanning;
            // window origin (uniform) + row step L m nr (uniform) + the
            // lane's 32-bit offset h nr + r: one address register per lane
            constexpr size_t ES = isf ? sizeof(float2) : sizeof(short2);
            const char *org = (const char *)(g ? P.im.img2 : P.im.img1) +
                              ((size_t)(g ? a2 : a1) * nr + (g ? r2 : r1)) * ES;
            const unsigned lo = (unsigned)((h * nr + r) * ES);
#pragma unroll
            for (int m = 0; m < H; m++) {
                const gchar *pm = (const gchar *)org + (size_t)(L * m) * nr * ES;
                float2 v;
                if constexpr (isf) {
                    v = __builtin_bit_cast(float2, *(const gu64 *)(pm + lo));
                } else {
                    const unsigned q = *(const gu32 *)(pm + lo);
                    v = make_float2((float)(short)(q & 0xffffu), (float)(short)(q >> 16));
                }
                x[m] = v;
            }
            // the taper is read on both windows and selected (a uniform
            // branch per element would wait for every load in flight)
            const gfloat *hn = (const gfloat *)P.hann + h * W + r;
#pragma unroll
            for (int m = 0; m < H; m++) {
#pragma clang fp contract(off)
                float2 v = x[m];
                if constexpr (isf) {
                    v.x *= 1000;
                    v.y *= 1000;
                }
                const float hw = hn[L * m * W];
                v.x = useH ? v.x * hw : v.x;
                v.y = useH ? v.y * hw : v.y;
                x[m] = v;
            }
            ::fftreg::fwd_quad128(x, h);  // column r: outputs blk H + k
#pragma unroll
            for (int k = 0; k < H; k++)
                lds[FL::at(blk * H + k, r)] = x[k];
            __syncthreads();
#pragma unroll
            for (int m = 0; m < H; m++)
                x[m] = lds[FL::at(r, L * m + h)];
            __syncthreads();
            ::fftreg::fwd_quad128(x, h);  // row r: x[k] = F_g[r][blk H + k]
        };
        // window 1's spectrum stays in registers while window 2's is formed,
        // and the cross spectrum X = F1 conj(F2) stays there for both
        // column phases of the inverse
        float2 Xr[H];
        fwd2d(0, Xr);
        {
            float2 x[H];
            fwd2d(1, x);
#pragma unroll
            for (int k = 0; k < H; k++) {
#pragma clang fp contract(off)
                Xr[k] = cmulconj(Xr[k], x[k]);
            }
        }
        // ---------------- polyphase inverse, one column phase q at a time
        const int krow = ((r - az) % W + W) % W;
        // s_blk = exp(i pi (blk 32 - (blk >= 2 ? 128 : 0)) / 128): 1, exp(i pi / 4), -i, -exp(i 3 pi / 4)
        auto sblk_of = [](int b) {
            const float c45 = 0x1.6a09e6p-1f;
            return b == 0 ? make_float2(1.f, 0.f)
                   : b == 1 ? make_float2(c45, c45) : b == 2 ? make_float2(0.f, -1.f) : make_float2(c45, -c45);
        };
        const int Jq = (2 * r + W) & (WW - 1);  // surface column of q = 0 (q = 1: + 1)
        Top2 best{0.0f, 0x7fffffff, 0.0f, 0.0f};
#pragma unroll 1
        for (int q = 0; q < 2; q++) {
            {
                int h = t & 3;
                asm volatile("" : "+v"(h));
                const int blk = ::fftreg::qblk(h);
                const float2 sblk = sblk_of(blk);
                float2 x[H];
#pragma unroll
                for (int k = 0; k < H; k++)  // column l = blk 32 + k: exp(i pi ls / 128) on phase 1
                    x[k] = q ? ::fftreg::cmul_f(Xr[k], f128_factor(sblk, k)) : Xr[k];
                ::fftreg::inv_quad128(x, h);  // x[n'] = Z[krow][L n' + h]
#pragma unroll
                for (int n = 0; n < H; n++)
                    lds[FL::at(krow, L * n + h)] = x[n];
            }
            __syncthreads();
            const int J = Jq + q;
#pragma unroll 1
            for (int pp = 0; pp < 2; pp++) {
                int h = t & 3;
                asm volatile("" : "+v"(h));
                const int blk = ::fftreg::qblk(h);
                const float2 sblk = sblk_of(blk);
                float2 y[H];
#pragma unroll
                for (int k = 0; k < H; k++) {
                    y[k] = lds[FL::at(blk * H + k, r)];
                    if (pp)  // row rho = blk 32 + k: exp(i pi rs / 128)
                        y[k] = ::fftreg::cmul_f(y[k], f128_factor(sblk, k));
                }
                ::fftreg::inv_quad128(y, h);  // y[j] = C_pp,q[L j + h][r]
                // |C|^2 first; a wave with no value below 2^-96 (nearly always)
                // takes the short form of the correctly rounded square root
                float mg[H];
                bool tiny = false;
#pragma unroll
                for (int jj = 0; jj < H; jj++) {
#pragma clang fp contract(off)
                    mg[jj] = y[jj].x * y[jj].x + y[jj].y * y[jj].y;
                    tiny |= !(mg[jj] >= 0x1p-96f);
                }
                if (__builtin_amdgcn_ballot_w64(tiny) == 0) {
#pragma unroll
                    for (int jj = 0; jj < H; jj++)
                        mg[jj] = sqrt_rn_normal(mg[jj]);
                } else {
#pragma unroll
                    for (int jj = 0; jj < H; jj++)
                        mg[jj] = sqrtf(mg[jj]);
                }
#pragma unroll
                for (int jj = 0; jj < H; jj++) {
                    const int I = (2 * (L * jj + h) + pp + W) & (WW - 1);
                    const float v = mg[jj];
                    M[((pp * 2 + q) * H + jj) * NTH + t] = v;  // full cache lines per store
                    const int i0 = I < W ? I : I - W, j0 = J < W ? J : J - W;
                    const int qd = (I < W) ? (J < W ? 0 : 3) : (J < W ? 2 : 1);
                    const int rank = (i0 * W + j0) * 4 + qd;
                    if (v > 0.0f) {
                        Top2 c{v, rank, 0.0f, v};
                        best = top2_combine(best, c);
                    } else {
                        best.v2 = fmaxf(best.v2, v);
                    }
                }
            }
            __syncthreads();  // every read of Z done before the next phase writes it
        }
        best = wg_top2<NTH>(best, sht);
        __syncthreads();
        PeakRes *res = &C.res[it];
        int I = -1000, J = -1000;
        float raw = 0.f;
        if (best.r1 != 0x7fffffff) {
            const int qd = best.r1 & 3, ij = best.r1 >> 2;
            const int i0 = ij / W, j0 = ij - (ij / W) * W;
            I = (qd == 0 || qd == 3) ? i0 : i0 + W;
            J = (qd == 0 || qd == 2) ? j0 : j0 + W;
            raw = best.v1;
        }
        if (t == 0) {
            res->gap = raw > 0 ? (raw - best.v2) / raw : 0.f;
            res->snr = best.sum > 0.0f ? best.v1 * (float)(WW * WW) / best.sum : 0.0f;
            res->lmOk = 0;
            res->fi = res->fj = -1;
            res->fineMax = 0;
            res->fineGap = 0;
        }
        if (I < NFAST / 2 || I > WW - NFAST / 2 || J < NFAST / 2 || J > WW - NFAST / 2) {
            if (t == 0) {
                res->I = res->J = -1000;
                res->raw = -1.f;
            }
        } else {
            if (t == 0) {
                res->I = I;
                res->J = J;
                res->raw = raw;
            }
            // the |C| block around the peak, for k_cpx_refine (cells off the
            // surface are never read: the fits' edge tests reject them)
            float *nb = C.nb + (size_t)it * kNbc;
            for (int e = t; e < kNbc; e += NTH) {
                const int ri = I - kNb / 2 + kNbMap.ri[e], ci = J - kNb / 2 + kNbMap.ci[e];
                // cell (ri, ci): phase (pp, q), row index L jj + h, column r
                const int u = (ri - W) & (WW - 1), v = (ci - W) & (WW - 1);
                const int ip = u >> 1;
                const size_t o = ((size_t)(((u & 1) * 2 + (v & 1)) * H + (ip >> 2)) * NTH) + (v >> 1) * L + (ip & 3);
                nb[e] = (ri >= 0 && ri < WW && ci >= 0 && ci < WW) ? M[o] : 0.0f;
            }
        }
        __syncthreads();  // block reads of M done before the next item writes it
    }
}

// ---------------------------------------------------------------------
// host side

void host_hanning(int wA, int wR, std::vector<float> &h)
{
    // computeHanning, speckleTrack.c:1830-1852, on the reference's storage:
    // one wA x wR block with rows of wR (speckleTrack.c:1775) whose columns
    // are indexed with wA / 2 (:1843-1844).  For wR != wA an element
    // [i][j] lands at flat offset i wR + j, in a neighbouring row or outside
    // the block: writes inside it land where the reference's do, in its
    // loop order; writes outside it are dropped (the reference's corrupt
    // its heap) and cells never written stay 0 (indeterminate there).
    const long n = (long)wA * wR;
    h.assign((size_t)n, 0.f);
    auto put = [&](int i, int j, float v) {
        const long f = (long)i * wR + j;
        if (f >= 0 && f < n)
            h[(size_t)f] = v;
    };
    for (int i = 0; i < wA / 2; i++) {
        int i1 = wA / 2 + i, i2 = wA / 2 - i - 1;
        double wi = 0.5 * (1.0 + cos((2.0 * M_PI) * (double)i / (double)(wA - 1)));
        for (int j = 0; j < wR / 2; j++) {
            int j1 = wA / 2 + j, j2 = wA / 2 - j - 1;
            double wj = 0.5 * (1.0 + cos((2.0 * M_PI) * (double)j / (double)(wR - 1)));
            float v = (float)(wi * wj);
            put(i1, j1, v);
            put(i2, j2, v);
            put(i2, j1, v);
            put(i1, j2, v);
        }
    }
}

// findImage2Pos with an initial-offset grid (speckleTrack.c:810-847,
// 1857-1940), evaluated on the host because points without a value keep
// the previous point's position.
void host_positions(const st_track_params &p, std::vector<int> &r2v, std::vector<int> &a2v)
{
    const int npts = p.nR * p.nA;
    r2v.resize(npts);
    a2v.resize(npts);
    int r2 = 0, a2 = 0;
    const float minvalue = (float)-ST_LARGEINT * 0.9f;
    for (int i = 0; i < p.nA; i++) {
        int a1 = p.aStart + i * p.deltaA - p.wA / 2;
        for (int j = 0; j < p.nR; j++) {
            int r1 = p.rStart + j * p.deltaR - p.wR / 2;
            double sc = (double)p.scaleFactor;
            double dr = -ST_LARGEINT, da = -ST_LARGEINT;
            bool have = false;
            {
                double range = (r1 - p.initR0) / p.initDeltaR, azimuth = (a1 - p.initA0) / p.initDeltaA;
                if (!(range < 0.0 || azimuth < 0.0 || (int)range >= p.initNr || (int)azimuth >= p.initNa)) {
                    const float *rim = p.initDr, *aim = p.initDa;
                    int nr = p.initNr, jj = (int)range, ii = (int)azimuth;
                    if (jj == p.initNr - 1 || ii == p.initNa - 1) {
                        dr = rim[ii * nr + jj];
                        da = aim[ii * nr + jj];
                        if (dr < minvalue || da < minvalue)
                            dr = da = -ST_LARGEINT;
                    } else {
                        float t = (float)(range - (double)jj), u = (float)(azimuth - (double)ii);
                        float q1 = rim[ii * nr + jj], q2 = rim[ii * nr + jj + 1], q3 = rim[(ii + 1) * nr + jj + 1],
                              q4 = rim[(ii + 1) * nr + jj];
                        bool bad = false;
                        if (q1 <= minvalue || q2 <= minvalue || q3 <= minvalue || q4 <= minvalue) {
                            dr = fmaxf(fmaxf(fmaxf(q1, q2), q3), q4);
                            if (dr < minvalue) {
                                dr = da = -ST_LARGEINT;
                                bad = true;
                            }
                        } else
                            dr = (float)((1.0 - t) * (1.0 - u) * q1 + t * (1.0 - u) * q2 + t * u * q3 + (1.0 - t) * u * q4);
                        if (!bad) {
                            q1 = aim[ii * nr + jj];
                            q2 = aim[ii * nr + jj + 1];
                            q3 = aim[(ii + 1) * nr + jj + 1];
                            q4 = aim[(ii + 1) * nr + jj];
                            if (q1 <= minvalue || q2 <= minvalue || q3 <= minvalue || q4 <= minvalue) {
                                da = fmaxf(fmaxf(fmaxf(q1, q2), q3), q4);
                                if (da < minvalue)
                                    dr = da = -ST_LARGEINT;
                            } else
                                da = (float)((1.0 - t) * (1.0 - u) * q1 + t * (1.0 - u) * q2 + t * u * q3 +
                                             (1.0 - t) * u * q4);
                        }
                    }
                }
            }
            if (dr > (-ST_LARGEINT + 1))
                have = true;
            if (have) {
                dr /= sc;
                da /= sc;
                r2 = (dr >= 0) ? r1 + (int)(dr + 0.5) : r1 + (int)(dr - 0.5);
                a2 = (da >= 0) ? a1 + (int)(da + 0.5) : a1 + (int)(da - 0.5);
            }
            r2v[(size_t)i * p.nR + j] = r2;
            a2v[(size_t)i * p.nR + j] = a2;
        }
    }
}

void init_dirichlet(st_ctx *ctx)
{
    // constant tables are per device and per process; contexts of several
    // threads may arrive here together, and none may launch before the
    // upload is complete
    static std::mutex mu;
    static bool done[64] = {};
    std::lock_guard<std::mutex> lk(mu);
    ST_REQUIRE(ctx->device >= 0 && ctx->device < 64, ST_ERR_ARG, "device index");
    if (done[ctx->device])
        return;
    // R(t) = sum_{j = 1, 3, 5, 7} 2 cos(pi j t / 96) at t = u + 24 - 12 m, and
    // b_k = exp(i pi k / 8) (refine_fine)
    double rt[kFine * NFAST];
    for (int u = 0; u < kFine; u++)
        for (int m = 0; m < NFAST; m++) {
            const double tt = (double)(u + 24 - 12 * m);
            double v = 0;
            for (int j = 1; j <= 7; j += 2)
                v += 2.0 * cos(M_PI * j * tt / 96.0);
            rt[u * NFAST + m] = v;
        }
    ST_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_refR), rt, sizeof(rt)));
    double2 bt[16];
    for (int k = 0; k < 16; k++)
        bt[k] = make_double2(cos(M_PI * k / 8.0), sin(M_PI * k / 8.0));
    ST_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_refB), bt, sizeof(bt)));
    float2 t128[128];
    for (int j = 0; j < 128; j++)
        t128[j] = make_float2((float)cos(2.0 * M_PI * j / 128.0), (float)sin(2.0 * M_PI * j / 128.0));
    ST_HIP(hipMemcpyToSymbol(HIP_SYMBOL(fftreg::c_tw128), t128, sizeof(t128)));
    double2 r256[256];
    for (int j = 0; j < 256; j++)
        r256[j] = make_double2(cos(2.0 * M_PI * j / 256.0), sin(2.0 * M_PI * j / 256.0));
    ST_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_root256), r256, sizeof(r256)));

    done[ctx->device] = true;
}

template <class T>
T *upload(st_ctx *ctx, const T *h, size_t n)
{
    T *d = ctx->arena.take<T>(n);
    ST_HIP(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return d;
}

#include "st_cpx_mfma.h"

// Dynamic LDS of the fused kernel: the two transpose buffers, which later
// hold the |C| surface (16 w^2 bytes).
template <int W>
constexpr size_t fused_lds_bytes()
{
    constexpr size_t t = (size_t)2 * FusedLanes<W>::BUF * sizeof(float2);
    static_assert(t >= (size_t)16 * W * W, "surface does not fit the transpose buffers");
    return t;
}

template <int W, int STYPE, int BASE, int OCC>
void launch_fused_kernel_o(st_ctx *ctx, int slot, const KPar *dP, const Pts *dS, const int *dStart, int njobs,
                           int total, const CpxBatch *dC, size_t lds)
{
    int perCU = 0, ncu = 0;
    constexpr int NTH = 2 * FusedLanes<W>::L * W;
    ST_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k_cpx_fused<W, STYPE, BASE, OCC>, NTH, lds));
    ST_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device));
    TuneEntry te;
    if (tune_find(ctx->device, W, &te) && te.wgPerCu > 0 && te.wgPerCu < perCU)
        perCU = te.wgPerCu;
    const int grid = std::max(1, std::min(total, std::max(1, perCU) * ncu));
    ctx->time_begin(slot);
    hipLaunchKernelGGL((k_cpx_fused<W, STYPE, BASE, OCC>), dim3(grid), dim3(NTH), lds, ctx->stream, dP, dS, dStart,
                       njobs, dC);
    ctx->time_end(slot);
    ST_HIP(hipGetLastError());
}

// Minimum waves per SIMD each instantiation is compiled for: what its LDS
// allows (w = 16: about 17 one-wave workgroups per CU; w = 32: 9 two-wave
// ones; w = 64: 2 four-wave ones), or less where the phase corrections
// need the registers.
template <int W, int BASE>
constexpr int fused_occ()
{
    return W == 16 ? 4 : W == 32 ? (BASE == 0 ? 4 : 3) : 2;
}

template <int W, int STYPE, int BASE>
void launch_fused_kernel(st_ctx *ctx, int slot, const KPar *dP, const Pts *dS, const int *dStart, int njobs,
                         int total, const CpxBatch *dC, size_t lds)
{
    launch_fused_kernel_o<W, STYPE, BASE, fused_occ<W, BASE>()>(ctx, slot, dP, dS, dStart, njobs, total, dC, lds);
}

// The complex attempts of one or more tracking calls of window class W.
struct FusedJob {
    KPar P;
    Pts S;
    CpxBatch C;  // C.list: the call's items
    int count;   // its complex attempts
};

// w = 128: one workgroup per CU (the transpose buffer fills the LDS), with
// its two global slots.
constexpr size_t kF128Lds = (size_t)FusedLanes<kF128>::BUF * sizeof(float2);
constexpr size_t kF128SlotBytes = kF128SlotM * sizeof(float);

template <int STYPE>
int fused128_per_cu()
{
    int perCU = 0;
    ST_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k_cpx_fused128<STYPE>, 512, kF128Lds));
    return std::max(1, perCU);
}

template <int STYPE>
void launch_fused128_kernel(st_ctx *ctx, int slot, const KPar *dP, const Pts *dS, const int *dStart, int njobs,
                            int total, const CpxBatch *dC)
{
    int ncu = 0;
    ST_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const int grid = std::max(1, std::min(total, fused128_per_cu<STYPE>() * ncu));
    float *sm = ctx->arena.take<float>((size_t)grid * kF128SlotM);
    ctx->time_begin(slot);
    hipLaunchKernelGGL((k_cpx_fused128<STYPE>), dim3(grid), dim3(512), kF128Lds, ctx->stream, dP, dS, dStart, njobs,
                       dC, sm);
    ctx->time_end(slot);
    ST_HIP(hipGetLastError());
}

template <int W, int STYPE>
void launch_cpx_fused_t(st_ctx *ctx, int slot, const std::vector<FusedJob> &jobs, const KPar *dP, const Pts *dS,
                        const CpxBatch *dC, const int *dStart, int total, int base)
{
    if constexpr (W == kF128) {
        // the window sums and acceptance test in their own kernel (the fused
        // kernels for w <= 64 form them from the windows they load)
        const int prepSlot = ctx->kernel_slot("k_cpx_prep");
        ctx->time_begin(prepSlot);
        for (size_t j = 0; j < jobs.size(); j++) {
            const int count = jobs[j].count;
            hipLaunchKernelGGL((k_cpx_prep_cols<W, STYPE>), dim3(count), dim3(64 * prep_segs<W>()), 0, ctx->stream,
                               dP + j, dS + j, jobs[j].C.list, count, dC + j);
        }
        ctx->time_end(prepSlot);
    }
    const int n = (int)jobs.size();
    if constexpr (W == kF128) {
        (void)base;  // the w = 128 kernel takes no phase corrections (plan() sends those calls to the staged path)
        launch_fused128_kernel<STYPE>(ctx, slot, dP, dS, dStart, n, total, dC);
    } else {
        const size_t lds = fused_lds_bytes<W>();
        switch (base) {
        case 0: launch_fused_kernel<W, STYPE, 0>(ctx, slot, dP, dS, dStart, n, total, dC, lds); break;
        case 1: launch_fused_kernel<W, STYPE, 1>(ctx, slot, dP, dS, dStart, n, total, dC, lds); break;
        case 2: launch_fused_kernel<W, STYPE, 2>(ctx, slot, dP, dS, dStart, n, total, dC, lds); break;
        default: launch_fused_kernel<W, STYPE, 3>(ctx, slot, dP, dS, dStart, n, total, dC, lds); break;
        }
    }
}

template <int W, int OCC>
int fused_per_cu_o(int stype)
{
    const size_t lds = fused_lds_bytes<W>();
    constexpr int NTH = 2 * FusedLanes<W>::L * W;
    int perCU = 0;
    if (stype == ST_SAMPLE_C64)
        ST_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k_cpx_fused<W, ST_SAMPLE_C64, 0, OCC>, NTH, lds));
    else
        ST_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k_cpx_fused<W, ST_SAMPLE_CI16, 0, OCC>, NTH, lds));
    return std::max(1, perCU);
}

// Resident workgroups per CU of the fused kernel (sizes its |C| scratch).
template <int W>
int fused_per_cu(int stype)
{
    return fused_per_cu_o<W, fused_occ<W, 0>()>(stype);
}

int fused_per_cu_w(int w, int stype)
{
    if (w == kF128)
        return stype == ST_SAMPLE_C64 ? fused128_per_cu<ST_SAMPLE_C64>() : fused128_per_cu<ST_SAMPLE_CI16>();
    return w == 16 ? fused_per_cu<16>(stype) : (w == 32 ? fused_per_cu<32>(stype) : fused_per_cu<64>(stype));
}

// One persistent launch over the complex attempts of every job (all of
// window class W, one sample type and correction set).  The parameter
// blocks live in device memory, read through the scalar cache instead of
// occupying SGPRs for the whole kernel.
template <int W>
void launch_cpx_fused(st_ctx *ctx, int slot, const std::vector<FusedJob> &jobs)
{
    const int n = (int)jobs.size();
    std::vector<KPar> hp(n);
    std::vector<Pts> hs(n);
    std::vector<CpxBatch> hc(n);
    std::vector<int> start(n + 1, 0);
    for (int j = 0; j < n; j++) {
        hp[j] = jobs[j].P;
        hs[j] = jobs[j].S;
        hc[j] = jobs[j].C;
        start[j + 1] = start[j] + jobs[j].count;
    }
    KPar *dP = ctx->arena.take<KPar>(n);
    Pts *dS = ctx->arena.take<Pts>(n);
    CpxBatch *dC = ctx->arena.take<CpxBatch>(n);
    int *dStart = ctx->arena.take<int>(n + 1);
    ST_HIP(hipMemcpyAsync(dP, hp.data(), sizeof(KPar) * n, hipMemcpyHostToDevice, ctx->stream));
    ST_HIP(hipMemcpyAsync(dS, hs.data(), sizeof(Pts) * n, hipMemcpyHostToDevice, ctx->stream));
    ST_HIP(hipMemcpyAsync(dC, hc.data(), sizeof(CpxBatch) * n, hipMemcpyHostToDevice, ctx->stream));
    ST_HIP(hipMemcpyAsync(dStart, start.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, ctx->stream));
    const KPar &P0 = jobs[0].P;
    const int base = (P0.baseFlag != 0 ? 1 : 0) | (P0.intOn ? 2 : 0);
    if (P0.im.stype == ST_SAMPLE_C64)
        launch_cpx_fused_t<W, ST_SAMPLE_C64>(ctx, slot, jobs, dP, dS, dC, dStart, start[n], base);
    else
        launch_cpx_fused_t<W, ST_SAMPLE_CI16>(ctx, slot, jobs, dP, dS, dC, dStart, start[n], base);
}

// bf16 of a float, round to nearest even (host side of the operand split)
uint16_t bf16_bits(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

float bf16_value(uint16_t b)
{
    const uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// A-operand tables of the MFMA complex match for window W (built once per
// device and window, kept for the life of the process).
MfmaTab mfma_tables(st_ctx *ctx, int W)
{
    static std::mutex mu;
    static std::map<std::pair<int, int>, MfmaTab> cache;
    std::lock_guard<std::mutex> lk(mu);
    const auto key = std::make_pair(ctx->device, W);
    auto f = cache.find(key);
    if (f != cache.end())
        return f->second;
    const int K2 = 2 * W;
    const double pi = 3.14159265358979323846;
    // [matrix][re/im output][hi/lo] host planes
    std::vector<uint16_t> h[2][2][2];
    const int rows[2] = {W, 2 * W};
    for (int mtx = 0; mtx < 2; mtx++)
        for (int o = 0; o < 2; o++)
            for (int hl = 0; hl < 2; hl++)
                h[mtx][o][hl].assign((size_t)rows[mtx] * K2, 0);
    auto put = [&](int mtx, int o, int row, int col, double v) {
        const float hi = bf16_value(bf16_bits((float)v));
        h[mtx][o][0][(size_t)row * K2 + col] = bf16_bits((float)v);
        h[mtx][o][1][(size_t)row * K2 + col] = bf16_bits((float)(v - (double)hi));
    };
    // (re, im) interleaved along K: row entries [2j] / [2j + 1] are
    // (Mr, -Mi) for the real part of a product and (Mi, Mr) for the
    // imaginary part
    for (int k = 0; k < W; k++)
        for (int j = 0; j < W; j++) {
            // A_W[k][j] = exp(-2 pi i k j / W)
            const double ang = -2.0 * pi * (double)((long)k * j % W) / W;
            const double ar = cos(ang), ai = sin(ang);
            put(0, 0, k, 2 * j, ar);
            put(0, 0, k, 2 * j + 1, -ai);
            put(0, 1, k, 2 * j, ai);
            put(0, 1, k, 2 * j + 1, ar);
        }
    for (int m = 0; m < 2 * W; m++)
        for (int u = 0; u < W; u++) {
            // B[m][u] = exp(+2 pi i m u_s / 2W), u_s signed (Nyquist negative)
            const int us = u < W / 2 ? u : u - W;
            const long e = (((long)m * us) % (2 * W) + 2 * W) % (2 * W);
            const double ang = 2.0 * pi * (double)e / (2.0 * W);
            const double br = cos(ang), bi = sin(ang);
            put(1, 0, m, 2 * u, br);
            put(1, 0, m, 2 * u + 1, -bi);
            put(1, 1, m, 2 * u, bi);
            put(1, 1, m, 2 * u + 1, br);
        }
    MfmaTab t{};
    for (int mtx = 0; mtx < 2; mtx++)
        for (int o = 0; o < 2; o++)
            for (int hl = 0; hl < 2; hl++) {
                void *d = nullptr;
                const size_t bytes = h[mtx][o][hl].size() * 2;
                ST_HIP(hipMalloc(&d, bytes));
                ST_HIP(hipMemcpy(d, h[mtx][o][hl].data(), bytes, hipMemcpyHostToDevice));
                (mtx == 0 ? t.fa : t.ib)[o][hl] = (const __bf16 *)d;
            }
    cache[key] = t;
    return t;
}

template <int W, int STYPE>
void launch_mfma_kernel(st_ctx *ctx, int slot, const KPar *dP, const Pts *dS, const int *list, int count,
                        const CpxBatch *dC, const MfmaTab &tab)
{
    using G = MfGeom<W>;
    const size_t lds = (size_t)G::PPB * G::BYTES;
    int perCU = 0, ncu = 0;
    ST_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, k_cpx_mfma<W, STYPE>, 256, lds));
    ST_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const int blocksNeeded = (count + G::PPB - 1) / G::PPB;
    const int grid = std::max(1, std::min(blocksNeeded, std::max(1, perCU) * ncu));
    ctx->time_begin(slot);
    hipLaunchKernelGGL((k_cpx_mfma<W, STYPE>), dim3(grid), dim3(256), lds, ctx->stream, dP, dS, list, count, dC,
                       tab);
    ctx->time_end(slot);
    ST_HIP(hipGetLastError());
}

// Complex attempts on the matrix cores (windows 16 and 32).
template <int W>
void launch_cpx_mfma(st_ctx *ctx, int slot, const KPar &P, const Pts &S, const int *list, int count,
                     const CpxBatch &C)
{
    const MfmaTab tab = mfma_tables(ctx, W);
    KPar *dP = ctx->arena.take<KPar>(1);
    Pts *dS = ctx->arena.take<Pts>(1);
    CpxBatch *dC = ctx->arena.take<CpxBatch>(1);
    ST_HIP(hipMemcpyAsync(dP, &P, sizeof(KPar), hipMemcpyHostToDevice, ctx->stream));
    ST_HIP(hipMemcpyAsync(dS, &S, sizeof(Pts), hipMemcpyHostToDevice, ctx->stream));
    ST_HIP(hipMemcpyAsync(dC, &C, sizeof(CpxBatch), hipMemcpyHostToDevice, ctx->stream));
    const int prepSlot = ctx->kernel_slot("k_cpx_prep");
    ctx->time_begin(prepSlot);
    const unsigned pb = (unsigned)((count + 64 / W - 1) / (64 / W));
    if (P.im.stype == ST_SAMPLE_C64)
        hipLaunchKernelGGL((k_cpx_prep_cols<W, ST_SAMPLE_C64>), dim3(pb), dim3(64), 0, ctx->stream, dP, dS, list,
                           count, dC);
    else
        hipLaunchKernelGGL((k_cpx_prep_cols<W, ST_SAMPLE_CI16>), dim3(pb), dim3(64), 0, ctx->stream, dP, dS, list,
                           count, dC);
    ctx->time_end(prepSlot);
    if (P.im.stype == ST_SAMPLE_C64)
        launch_mfma_kernel<W, ST_SAMPLE_C64>(ctx, slot, dP, dS, list, count, dC, tab);
    else
        launch_mfma_kernel<W, ST_SAMPLE_CI16>(ctx, slot, dP, dS, list, count, dC, tab);
}

bool staged_only()
{
    const char *e = getenv("SPECKLE_AMD_PATH");
    return e && strcmp(e, "staged") == 0;
}

int read_count(st_ctx *ctx, int *dcounts, int idx)
{
    int v = 0;
    ST_HIP(hipMemcpyAsync(&v, dcounts + idx, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    ST_HIP(hipStreamSynchronize(ctx->stream));
    return v;
}

}  // namespace

namespace {

// One strack call on the device: its parameters, the per-point state it
// carves from the context arena and the stages that run for it alone.  A
// batch of calls carves all of them first; the calls whose complex
// matches run on the fused kernel then share one launch per window class.
struct StrackJob {
    st_ctx *ctx = nullptr;
    StPair pr;
    const st_track_params *pp = nullptr;
    st_track_out *out = nullptr;
    st_track_diag *diag = nullptr;
    st_cull_params *cull = nullptr;
    bool intOn = false, gauss = false, mfmaPath = false, fusedPath = false;
    int kPeak = 0, kAmpPeak = 0, kGather = 0, kFft = 0, kAmpDet = 0;
    int npts = 0, w = 0, wa = 0, intPs = 0;  // w: range window, wa: azimuth window
    std::vector<float> hann;
    std::vector<int> hr2, ha2;
    size_t fixed = 0;
    size_t perCpx = 0, perAmp = 0;
    int Bc = 1, Ba = 1;
    Pts S{};
    KPar P{};
    int *carryMax = nullptr, *list = nullptr;
    float *dOffR = nullptr, *dOffA = nullptr;
    const int *dR2 = nullptr, *dA2 = nullptr;
    int nCpx = 0;
    size_t mark = 0;  // arena position shared by the batch's stage scratch
    int64_t patches = 0;  // patch correlations of this call
    CpxBatch C{};     // fused path: the per-item results of this call
    struct OvfBuf {
        void *p = nullptr;
        OvfBuf() = default;
        OvfBuf(const OvfBuf &) = delete;
        OvfBuf &operator=(const OvfBuf &) = delete;
        ~OvfBuf()
        {
            if (p)
                (void)hipFree(p);
        }
    } intOvfBuf, intWsBuf;

    StrackJob(st_ctx *c, StPair &p, const st_track_params &par, st_track_out *o, st_track_diag *d,
              st_cull_params *cu)
        : ctx(c), pr(p), pp(&par), out(o), diag(d), cull(cu)
    {
    }
    StrackJob(const StrackJob &) = delete;
    StrackJob &operator=(const StrackJob &) = delete;

    // Validation, host tables and the bytes this call needs.
    void plan()
    {
        const st_track_params &par = *pp;
        // range and azimuth sizes are independent (parseTrack.c:167-171)
        ST_REQUIRE(par.wR >= 8 && par.wA >= 8 && par.wRa >= 8 && par.wAa >= 8 && !(par.wR & 1) && !(par.wA & 1) &&
                       !(par.wRa & 1) && !(par.wAa & 1),
                   ST_ERR_ARG, "window sizes");
        for (int ws : {par.wR, par.wA})
            ST_REQUIRE(stfft::supported_size(ws) && stfft::supported_size(2 * ws), ST_ERR_UNSUPPORTED,
                       "window size not supported by the FFT engine");
        for (int ws : {par.wRa, par.wAa})
            ST_REQUIRE(stfft::supported_size(ws) && stfft::supported_size(2 * ws) && stfft::supported_size(ws * LA) &&
                           stfft::supported_size(2 * ws * LA),
                       ST_ERR_UNSUPPORTED, "window size not supported by the FFT engine");
        ST_REQUIRE(par.nR > 0 && par.nA > 0 && par.scaleFactor > 0, ST_ERR_ARG, "grid");
        ST_REQUIRE(par.polyShift || (par.initDr && par.initDa && par.initNr > 0 && par.initNa > 0), ST_ERR_ARG,
                   "initial offsets: polyShift=0 needs initDr/initDa");
        // the matrix-core complex match covers windows 16 and 32 without the
        // baseline ramp; other cases run the FP32 paths
        intOn = par.intf != nullptr && !par.noComplex;
        // -gauss fits its peak on the staged path's magnitude surfaces
        gauss = par.gaussFlag != 0;
        // the matrix-core and fused kernels take square windows; the staged
        // kernels take any supported wA x wR
        const bool square = par.wR == par.wA;
        int precision = par.precision;
        if (precision == ST_PREC_AUTO) {
            TuneEntry te;
            precision = square && tune_find(ctx->device, par.wR, &te) ? te.precision : ST_PREC_FP32;
        }
        mfmaPath = precision == ST_PREC_MFMA && square && (par.wR == 16 || par.wR == 32) && !par.baseFlag &&
                       !intOn && !gauss && !staged_only();
        fusedPath = square && (par.wR == 16 || par.wR == 32 || par.wR == 64 ||
                               (par.wR == kF128 && !par.baseFlag && !intOn)) &&
                    !gauss && !staged_only();
        kPeak = ctx->kernel_slot(mfmaPath ? "k_cpx_mfma" : (fusedPath ? "k_cpx_fused" : "peak_refine"));
        kAmpPeak = ctx->kernel_slot("amp_peak");
        kGather = ctx->kernel_slot("cpx_gather");
        kFft = ctx->kernel_slot("fft2d");
        kAmpDet = ctx->kernel_slot("amp_detect");
        npts = par.nR * par.nA;
        w = par.wR;
        wa = par.wA;

        // per-point state + inputs
        host_hanning(par.wA, par.wR, hann);
        if (!par.polyShift)
            host_positions(par, hr2, ha2);
        fixed = (size_t)npts * 200 + hann.size() * 4 + (hr2.size() * 8) + 4096 +
                       ((size_t)npts / kCarryBlk + 1) * 4 + 256;
        if (par.mask)
            fixed += (size_t)par.maskNr * par.maskNa;
        if (par.baseFlag)
            fixed += (size_t)pr.nr1 * 24;
        intPs = 0;
        if (intOn) {
            ST_REQUIRE(par.intNr > 0 && par.intNa > 0 && par.intNrl > 0 && par.intNal > 0, ST_ERR_ARG,
                       "interferogram size and looks");
            intPs = std::max(std::max(par.wA / par.intNal, par.wR / par.intNrl), 8);
            ST_REQUIRE(intPs <= kIntMaxPsG, ST_ERR_UNSUPPORTED,
                       "interferogram patch above 128 x 128 samples (window / looks)");
            ST_REQUIRE((intPs * par.intNal) / 2 >= par.wA / 2 && (intPs * par.intNrl) / 2 >= par.wR / 2, ST_ERR_ARG,
                       "window not covered by the oversampled interferogram patch (window not a multiple of the looks)");
            ST_REQUIRE((long)intPs * par.intNal <= 32768 && (long)intPs * par.intNrl <= 32768, ST_ERR_ARG,
                       "interferogram looks");
            fixed += (size_t)par.intNr * par.intNa * 8 + (size_t)npts * sizeof(IntSet) + 512;
        }
        // the cull chain works on the output grids in the same arena
        if (cull) {
            ST_REQUIRE(cull->nR == par.nR && cull->nA == par.nA && cull->offR == out->offR && cull->offA == out->offA &&
                           cull->corr == out->corr && cull->type == out->type,
                       ST_ERR_ARG, "cull grids must be the tracking outputs");
            fixed += cull_device_bytes((size_t)npts);
        }
        const int sMax = par.maxTries >= 2 ? LA : 1, wsA = par.wAa * sMax, wsR = par.wRa * sMax;
        perCpx = (size_t)w * wa * (16 + 32 + 32 + 16) + 2048;
        perAmp = (size_t)wsA * wsR * 128 + (size_t)(wsA + wsR) * 96 + 4096;
    }

    // Bytes of the fused path's per-item results and |C| blocks (an upper
    // bound: every point an attempt; the kernel keeps its |C| surfaces in
    // LDS).
    size_t fused_bytes() const
    {
        if (!fusedPath)
            return 0;
        const size_t perPt =
            sizeof(PeakRes) + 2 * sizeof(float) + 1 + sizeof(int) + 2 * sizeof(double) + kNbc * sizeof(float);
        size_t slots = 0;
        if (w == kF128) {
            // the w = 128 kernel's per-workgroup slots
            int ncu = 0;
            ST_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device));
            slots = (size_t)fused_per_cu_w(w, pr.stype) * ncu * kF128SlotBytes + 512;
        }
        return (size_t)npts * perPt + 8 * 256 + slots;
    }

    // Per-point state, parameter block, uploads, geometry, the count of
    // complex attempts and the interferogram spectra.
    void setup()
    {
        const st_track_params &par = *pp;
        hipStream_t st = ctx->stream;
        hipEvent_t *sev = ctx->ev + 64;
        Arena &ar = ctx->arena;
        S = Pts{};
        S.r1 = ar.take<int>(npts);
        S.a1 = ar.take<int>(npts);
        S.r2 = ar.take<int>(npts);
        S.a2 = ar.take<int>(npts);
        S.maskv = ar.take<uint8_t>(npts);
        S.haveData = ar.take<uint8_t>(npts);
        S.good = ar.take<uint8_t>(npts);
        S.type = ar.take<uint8_t>(npts);
        S.azValid = ar.take<uint8_t>(npts);
        S.azShift = ar.take<int>(npts);
        S.azUse = ar.take<int>(npts);
        carryMax = ar.take<int>((npts + kCarryBlk - 1) / kCarryBlk);
        S.corr = ar.take<float>(npts);
        S.rShift1 = ar.take<float>(npts);
        S.aShift1 = ar.take<float>(npts);
        S.azd = ar.take<float>(npts);
        S.peakI = ar.take<int>(3 * (size_t)npts);
        S.peakJ = ar.take<int>(3 * (size_t)npts);
        S.fineI = ar.take<int>(3 * (size_t)npts);
        S.fineJ = ar.take<int>(3 * (size_t)npts);
        S.cmax = ar.take<double>(3 * (size_t)npts);
        S.sigRg = ar.take<float>(3 * (size_t)npts);
        S.sigAz = ar.take<float>(3 * (size_t)npts);
        S.peakGap = ar.take<float>(3 * (size_t)npts);
        S.fineGap = ar.take<float>(3 * (size_t)npts);
        S.snr = ar.take<float>(3 * (size_t)npts);
        S.counts = ar.take<int>(CNT_N);
        list = ar.take<int>(npts);
        dOffR = ar.take<float>(npts);
        dOffA = ar.take<float>(npts);

        P = KPar{};
        P.nR = par.nR;
        P.nA = par.nA;
        P.rStart = par.rStart;
        P.aStart = par.aStart;
        P.deltaR = par.deltaR;
        P.deltaA = par.deltaA;
        P.wR = par.wR;
        P.wA = par.wA;
        P.wRa = par.wRa;
        P.wAa = par.wAa;
        P.sf = par.scaleFactor;
        P.noComplex = par.noComplex;
        P.hanning = par.hanningFlag;
        P.maxTries = par.maxTries;
        P.checkAz = par.checkAzFocus;
        P.gauss = gauss ? 1 : 0;
        P.azThresh = par.azDefocusThresh;
        for (int k = 0; k < 3; k++) {
            P.rPoly[k] = par.rShiftPoly[k];
            P.aPoly[k] = par.aShiftPoly[k];
        }
        P.mask = par.mask ? upload(ctx, par.mask, (size_t)par.maskNr * par.maskNa) : nullptr;
        P.maskNr = par.maskNr;
        P.maskNa = par.maskNa;
        P.maskR0 = par.maskR0;
        P.maskA0 = par.maskA0;
        P.maskNrl = par.maskNrl > 0 ? par.maskNrl : 1;
        P.maskNal = par.maskNal > 0 ? par.maskNal : 1;
        P.baseFlag = par.baseFlag;
        if (par.baseFlag) {
            ST_REQUIRE(par.sinThetaD && par.cosThetaD && par.overRange && par.nSlpA1 > 0 && par.lambda != 0, ST_ERR_ARG,
                       "baseline correction needs the range tables");
            P.Bn = par.Bn;
            P.Bp = par.Bp;
            P.dBn = par.dBn;
            P.dBp = par.dBp;
            P.dBnQ = par.dBnQ;
            P.dBpQ = par.dBpQ;
            P.lambda = par.lambda;
            P.nSlpA1 = par.nSlpA1;
            P.sinT = upload(ctx, par.sinThetaD, (size_t)pr.nr1);
            P.cosT = upload(ctx, par.cosThetaD, (size_t)pr.nr1);
            P.overR = upload(ctx, par.overRange, (size_t)pr.nr1);
        }
        P.intOn = intOn ? 1 : 0;
        if (intOn) {
            P.intNr = par.intNr;
            P.intNa = par.intNa;
            P.intNrl = par.intNrl;
            P.intNal = par.intNal;
            P.intPs = intPs;
            P.intPa = intPs * par.intNal;
            P.intPr = intPs * par.intNrl;
            P.intf = upload(ctx, par.intf, (size_t)2 * par.intNr * par.intNa);
        }
        P.hann = upload(ctx, hann.data(), hann.size());
        P.im = Images{pr.d1, pr.d2, pr.nr1, pr.na1, pr.nr2, pr.na2, pr.stype};
        dR2 = nullptr;
        dA2 = nullptr;
        if (!par.polyShift) {
            dR2 = upload(ctx, hr2.data(), hr2.size());
            dA2 = upload(ctx, ha2.data(), ha2.size());
        }
        ST_HIP(hipMemsetAsync(S.counts, 0, CNT_N * sizeof(int), st));
        hipLaunchKernelGGL(k_geometry, dim3((npts + 255) / 256), dim3(256), 0, st, P, S, dR2, dA2, list);
        ST_HIP(hipGetLastError());
        ST_HIP(hipEventRecord(sev[1], st));
        nCpx = read_count(ctx, S.counts, CNT_LIST);

        // ---- interferogram spectra of the complex attempts
        if (intOn && nCpx > 0) {
            IntSet *set = ar.take<IntSet>(nCpx);
            int *nOvfDev = ar.take<int>(1);
            ST_HIP(hipMemsetAsync(nOvfDev, 0, sizeof(int), st));
            P.intSet = set;
            const size_t lds = int_lds_bytes(intPs);
            const int kInt = ctx->kernel_slot("k_int_bins");
            // patches above kIntMaxPs: k_int_bins_g over a persistent grid with its workspace
            const bool big = intPs > kIntMaxPs;
            const int gridG = std::min(nCpx, 512);
            if (big)
                ST_HIP(hipMalloc(&intWsBuf.p, (size_t)gridG * int_ws_bytes(intPs)));
            auto bins = [&](IntBin *ovfp, int pass) {
                if (big)
                    hipLaunchKernelGGL(k_int_bins_g, dim3(gridG), dim3(kIntThreads), 0, st, P, S, list, nCpx, set, ovfp,
                                       nOvfDev, pass, (unsigned char *)intWsBuf.p);
                else
                    hipLaunchKernelGGL(k_int_bins, dim3(nCpx), dim3(kIntThreads), lds, st, P, S, list, set, ovfp,
                                       nOvfDev, pass);
            };
            ctx->time_begin(kInt);
            bins((IntBin *)nullptr, 0);
            ctx->time_end(kInt);
            ST_HIP(hipGetLastError());
            const int nOvf = read_count(ctx, nOvfDev, 0);
            if (nOvf > 0) {
                // spectra with more than kIntInline bins within 1% of their peak
                ST_HIP(hipMalloc(&intOvfBuf.p, (size_t)nOvf * intPs * intPs * sizeof(IntBin)));
                P.intOvf = (const IntBin *)intOvfBuf.p;
                ctx->time_begin(kInt);
                bins((IntBin *)intOvfBuf.p, 1);
                ctx->time_end(kInt);
                ST_HIP(hipGetLastError());
            }
        }

    }

    // This call's complex attempts on the matrix-core or the staged path
    // (the fused path runs batched, run_fused_group).
    void complex_alone()
    {
        const st_track_params &par = *pp;
        hipStream_t st = ctx->stream;
        Arena &ar = ctx->arena;
        if (nCpx > 0 && mfmaPath) {
            CpxBatch C{};
            C.res = ar.take<PeakRes>(nCpx);
            C.sig = ar.take<float>(2 * (size_t)nCpx);
            C.have = ar.take<uint8_t>(nCpx);
            C.az = ar.take<int>(nCpx);
            C.p = ar.take<double>(2 * (size_t)nCpx);
            C.nb = ar.take<float>((size_t)nCpx * kNbc);
            if (w == 16)
                launch_cpx_mfma<16>(ctx, kPeak, P, S, list, nCpx, C);
            else
                launch_cpx_mfma<32>(ctx, kPeak, P, S, list, nCpx, C);
            const int kRef = ctx->kernel_slot("k_cpx_refine");
            ctx->time_begin(kRef);
            hipLaunchKernelGGL(k_cpx_refine, dim3((nCpx + kRefPer - 1) / kRefPer), dim3(kRefThreads), 0, st, C,
                               nCpx, 2 * w);
            ctx->time_end(kRef);
            hipLaunchKernelGGL(k_lm, dim3((2 * nCpx + 127) / 128), dim3(128), 0, st, C.res, nCpx, C.sig);
            hipLaunchKernelGGL(k_cpx_finalize, dim3((nCpx + 127) / 128), dim3(128), 0, st, P, S, list, 0, nCpx, C);
            ST_HIP(hipGetLastError());
            ar.used = mark;
        } else if (nCpx > 0 && !fusedPath) {
            int B = std::min(Bc, nCpx);
            TuneEntry te;
            if (w == wa && tune_find(ctx->device, w, &te) && te.stagedBatch > 0)
                B = std::min(B, te.stagedBatch);
            const size_t N = (size_t)wa * w;  // wA rows of wR samples
            CpxBatch C;
            C.P = ar.take<float2>(2 * B * N);
            C.S = ar.take<float2>(B * 4 * N);
            C.tmp = ar.take<float2>(B * 4 * N);
            C.mag = ar.take<float>(B * 4 * N);
            C.res = ar.take<PeakRes>(B);
            C.sig = ar.take<float>(2 * (size_t)B);
            C.have = ar.take<uint8_t>(B);
            C.az = ar.take<int>(B);
            C.p = ar.take<double>(2 * (size_t)B);
            // blocking of the two passes after the range transforms: Cx spectrum
            // columns (each 2 wA long) per k_cpx_colx workgroup, Rr rows (each
            // 2 wR long) per k_cpx_rowmag workgroup
            const int WA = 2 * wa, WR = 2 * w;
            ST_REQUIRE(WA <= kStElems && WR <= kStElems, ST_ERR_UNSUPPORTED,
                       "window too large for the staged complex path");
            int lcx = 0;
            while ((2 << lcx) <= 16 && (2 << lcx) <= w && (2 << lcx) * WA <= kStElems)
                lcx++;
            const int Cx = 1 << lcx, Rr = kStElems / WR, nPart = (WA + Rr - 1) / Rr;
            const size_t ldsX = (size_t)Cx * std::max(2 * (wa + 1), WA + 1) * sizeof(float2),
                         ldsR = (size_t)Rr * WR * sizeof(float2);
            const stfft::PlanArgs PfA = ctx->fft.plan(wa, -1), PiA2 = ctx->fft.plan(WA, +1),
                                  PiR2 = ctx->fft.plan(WR, +1);
            const stfft::FastDiv dA2 = stfft::make_fastdiv((uint32_t)WA), dR2 = stfft::make_fastdiv((uint32_t)WR);
            Top2 *part = (Top2 *)C.tmp;  // [B][nPart]
            for (int off = 0; off < nCpx; off += B) {
                ctx->check_cancel();
                const int nb = std::min(B, nCpx - off);
                ctx->time_begin(kGather);
                hipLaunchKernelGGL(k_cpx_gather, dim3(nb), dim3(NT), 0, st, P, S, list, off, nb, C);
                ctx->time_end(kGather);
                ctx->time_begin(kFft);
                ctx->fft.rows(C.P, (long)2 * nb * wa, w, -1, st);
                ctx->time_end(kFft);
                ctx->time_begin(kPeak);
                hipLaunchKernelGGL(k_cpx_colx, dim3((w + Cx - 1) / Cx, nb), dim3(NT), ldsX, st, C.P, nb, wa, w, C.az, C.S,
                                   lcx, PfA, PiA2, dA2);
                hipLaunchKernelGGL(k_cpx_rowmag, dim3(nPart, nb), dim3(NT), ldsR, st, C.S, wa, w, Rr, C.mag, part, PiR2,
                                   dR2);
                const GaussArg G{gauss ? 1 : 0, wa, w, par.hanningFlag ? 1.5 : 1.0, 1.0, C.p};
                if (gauss) {
                    hipLaunchKernelGGL(k_peak_merge<true>, dim3(nb), dim3(NT), 0, st, part, nPart, C.mag, WA, WR, 0,
                                       C.res, G);
                    hipLaunchKernelGGL(k_gauss_fit, dim3((nb + 127) / 128), dim3(128), 0, st, C.mag, WA, WR, 0, C.res, nb,
                                       G);
                } else
                    hipLaunchKernelGGL(k_peak_merge<false>, dim3(nb), dim3(NT), 0, st, part, nPart, C.mag, WA, WR, 0,
                                       C.res, G);
                ctx->time_end(kPeak);
                hipLaunchKernelGGL(k_lm, dim3((2 * nb + 127) / 128), dim3(128), 0, st, C.res, nb, C.sig);
                hipLaunchKernelGGL(k_cpx_finalize, dim3((nb + 127) / 128), dim3(128), 0, st, P, S, list, off, nb, C);
                ST_HIP(hipGetLastError());
            }
            ar.used = mark;
        }
    }

    // The fused path: this call's share of a batched launch.
    FusedJob fused_part()
    {
        Arena &ar = ctx->arena;
        C = CpxBatch{};
        C.res = ar.take<PeakRes>(nCpx);
        C.sig = ar.take<float>(2 * (size_t)nCpx);
        C.have = ar.take<uint8_t>(nCpx);
        C.az = ar.take<int>(nCpx);
        C.p = ar.take<double>(2 * (size_t)nCpx);
        C.list = list;
        C.nb = ar.take<float>((size_t)nCpx * kNbc);
        return FusedJob{P, S, C, nCpx};
    }

    // Width fits and acceptance of the fused path's attempts.
    void fused_finish()
    {
        hipStream_t st = ctx->stream;
        const int kRef = ctx->kernel_slot("k_cpx_refine");
        ctx->time_begin(kRef);
        hipLaunchKernelGGL(k_cpx_refine, dim3((nCpx + kRefPer - 1) / kRefPer), dim3(kRefThreads), 0, st, C, nCpx,
                           2 * w);
        ctx->time_end(kRef);
        hipLaunchKernelGGL(k_lm, dim3((2 * nCpx + 127) / 128), dim3(128), 0, st, C.res, nCpx, C.sig);
        hipLaunchKernelGGL(k_cpx_finalize, dim3((nCpx + 127) / 128), dim3(128), 0, st, P, S, list, 0, nCpx, C);
        ST_HIP(hipGetLastError());
    }

    // Carry of the azimuth shift, amplitude retries, composition, the
    // optional cull chain and the copies out.
    void finish()
    {
        const st_track_params &par = *pp;
        hipStream_t st = ctx->stream;
        hipEvent_t *sev = ctx->ev + 64;
        Arena &ar = ctx->arena;
        ctx->check_cancel();
        {
            const int nCarry = (npts + kCarryBlk - 1) / kCarryBlk;
            hipLaunchKernelGGL(k_carry_blocks, dim3(nCarry), dim3(kCarryBlk), 0, st, S, npts, carryMax);
            hipLaunchKernelGGL(k_carry_az, dim3(nCarry), dim3(kCarryBlk), 0, st, S, npts, carryMax);
        }
        ST_HIP(hipEventRecord(sev[2], st));
        ST_HIP(hipEventRecord(sev[3], st));
        ST_HIP(hipEventRecord(sev[4], st));

        // ---- amplitude attempts
        for (int t = 0; t < par.maxTries && t < 2; t++) {
            const int large = t;
            ST_HIP(hipMemsetAsync(S.counts + CNT_LIST, 0, sizeof(int), st));
            hipLaunchKernelGGL(k_amp_list, dim3((npts + 255) / 256), dim3(256), 0, st, P, S, npts, list);
            const int nA = read_count(ctx, S.counts, CNT_LIST);
            if (nA == 0)
                break;
            // patches wsA x wsR, interpolated surfaces WA x WR
            const int wsA = par.wAa * (large ? LA : 1), wsR = par.wRa * (large ? LA : 1), WA = 2 * wsA, WR = 2 * wsR;
            ST_REQUIRE(WA <= kStElems && WR <= kStElems, ST_ERR_UNSUPPORTED,
                       "window too large for the staged amplitude path");
            // blocking: Cx interpolation columns per k_amp_colx workgroup, Rd
            // rows per k_amp_rowdet / k_amp_rowfwd workgroup, Cc correlation
            // columns per k_amp_colx2 workgroup, Rm rows per k_amp_rowmag one
            int lcx = 0, lcc = 0;
            while ((2 << lcx) <= wsR && (2 << lcx) * WA <= kStElems)
                lcx++;
            while ((2 << lcc) <= wsR && (2 << lcc) * WA * 2 <= kStElems)
                lcc++;
            const int Cx = 1 << lcx, Cc = 1 << lcc;
            const int Rd = std::max(1, kStElems / (2 * WR)), nPd = (WA + Rd - 1) / Rd;
            const int Rf = kStElems / WR, nPf = (WA + Rf - 1) / Rf;
            const int Rm = 2 * (kStElems / WR), nPm = (WA + Rm - 1) / Rm;
            const size_t ldsX = (size_t)Cx * (Cx > 1 ? WA + 1 : WA) * sizeof(float2);
            const size_t ldsD = (size_t)2 * Rd * WR * sizeof(float2), ldsF = (size_t)Rf * WR * sizeof(float2);
            const size_t ldsC = (size_t)2 * Cc * (Cc > 1 ? WA + 1 : WA) * sizeof(float2);
            const size_t ldsM = (size_t)(Rm / 2) * WR * sizeof(float2);
            const size_t per = (size_t)wsA * wsR * 128 + (size_t)nPd * (sizeof(AmpPart) + sizeof(double2)) +
                               (size_t)nPm * sizeof(Top2) + 2048;
            // the amplitude kernels put 2 x batch on grid y (at most 65535)
            const size_t kMaxAmpBatch = 16384;
            const int B = (int)std::max<size_t>(1, std::min<size_t>({(size_t)nA, (ar.cap - ar.used) / per, kMaxAmpBatch}));
            AmpBatch A;
            A.I = ar.take<float2>((size_t)2 * B * wsA * wsR);
            A.F = ar.take<float2>((size_t)2 * B * WA * WR);
            A.S = ar.take<float2>((size_t)B * WA * WR);
            A.mag = ar.take<float>((size_t)B * WA * WR);
            A.res = ar.take<PeakRes>(B);
            A.sig = ar.take<float>(2 * (size_t)B);
            A.state = ar.take<uint8_t>(B);
            A.az = ar.take<int>(B);
            A.p = ar.take<double>(2 * (size_t)B);
            A.dpart = ar.take<AmpPart>((size_t)B * nPd);
            A.ppart = ar.take<double2>((size_t)B * nPd);
            A.mean = ar.take<double2>(B);
            A.part = ar.take<Top2>((size_t)B * nPm);
            const stfft::PlanArgs PfsA = ctx->fft.plan(wsA, -1), PfA = ctx->fft.plan(WA, -1),
                                  PiA = ctx->fft.plan(WA, +1), PfR = ctx->fft.plan(WR, -1),
                                  PiR = ctx->fft.plan(WR, +1);
            const stfft::FastDiv dA = stfft::make_fastdiv((uint32_t)WA), dR = stfft::make_fastdiv((uint32_t)WR);
            for (int off = 0; off < nA; off += B) {
                ctx->check_cancel();
                const int nb = std::min(B, nA - off);
                hipLaunchKernelGGL(k_amp_gather, dim3(nb), dim3(NT), 0, st, P, S, list, off, nb, large, A);
                ctx->fft.rows(A.I, (long)2 * nb * wsA, wsR, -1, st);
                hipLaunchKernelGGL(k_amp_colx, dim3((wsR + Cx - 1) / Cx, 2 * nb), dim3(NT), ldsX, st, A.I, nb, wsA, wsR,
                                   A.az, A.state, A.F, lcx, PfsA, PiA, dA);
                ctx->time_begin(kAmpDet);
                hipLaunchKernelGGL(k_amp_rowdet, dim3(nPd, nb), dim3(NT), ldsD, st, P, A.F, nb, wsA, wsR, Rd, A.state,
                                   (float *)A.S, A.dpart, PiR, dR);
                ctx->time_end(kAmpDet);
                hipLaunchKernelGGL(k_amp_stat, dim3(nb), dim3(64), 0, st, A.dpart, nPd, WA, WR, A.state, A.mean);
                hipLaunchKernelGGL(k_amp_rowfwd, dim3(nPf, nb), dim3(NT), ldsF, st, (const float *)A.S, nb, WA, WR, Rf,
                                   A.mean, A.state, A.F, A.ppart, PfR);
                hipLaunchKernelGGL(k_amp_colx2, dim3((wsR + Cc - 1) / Cc, nb), dim3(NT), ldsC, st, A.F, WA, WR, A.state,
                                   A.S, lcc, PfA, PiA, dA);
                hipLaunchKernelGGL(k_amp_psum, dim3((nb + 127) / 128), dim3(128), 0, st, A.ppart, nPf, nb, WA, WR,
                                   A.state, A.p);
                ctx->time_begin(kAmpPeak);
                hipLaunchKernelGGL(k_amp_rowmag, dim3(nPm, nb), dim3(NT), ldsM, st, A.S, WA, WR, Rm, A.mag, A.part, PiR,
                                   dR, A.state);
                const GaussArg G{gauss ? 1 : 0, wsA, wsR, 1.0, 16.0, A.p};
                if (gauss) {
                    hipLaunchKernelGGL(k_peak_merge<true>, dim3(nb), dim3(NT), 0, st, A.part, nPm, A.mag, WA, WR, 1,
                                       A.res, G);
                    hipLaunchKernelGGL(k_gauss_fit, dim3((nb + 127) / 128), dim3(128), 0, st, A.mag, WA, WR, 1, A.res, nb,
                                       G);
                } else
                    hipLaunchKernelGGL(k_peak_merge<false>, dim3(nb), dim3(NT), 0, st, A.part, nPm, A.mag, WA, WR, 1,
                                       A.res, G);
                ctx->time_end(kAmpPeak);
                hipLaunchKernelGGL(k_lm, dim3((2 * nb + 127) / 128), dim3(128), 0, st, A.res, nb, A.sig);
                hipLaunchKernelGGL(k_amp_finalize, dim3((nb + 127) / 128), dim3(128), 0, st, P, S, list, off, nb, large, A);
                ST_HIP(hipGetLastError());
            }
            ar.used = mark;
            if (!large)
                ST_HIP(hipEventRecord(sev[3], st));
        }
        // the doubled-window stage ends here also when it did not run (no
        // failed small tries, or one try allowed), so its time reads zero
        ST_HIP(hipEventRecord(sev[4], st));

        hipLaunchKernelGGL(k_compose, dim3((npts + 255) / 256), dim3(256), 0, st, P, S, dOffR, dOffA, npts);
        ST_HIP(hipGetLastError());
        if (cull) {
            ctx->check_cancel();
            st_cull_chain_device(ctx, cull, dOffR, dOffA, S.corr, S.type);
        }
        int counts[CNT_N];
        ST_HIP(hipMemcpyAsync(counts, S.counts, sizeof(counts), hipMemcpyDeviceToHost, st));
        // results go through a pinned staging buffer (a device-to-pageable copy
        // takes the driver's slower staged path) and are spread into the
        // caller's arrays after the stream has drained
        const size_t f4 = (size_t)npts * sizeof(float);
        char *stg = (char *)ctx->pinned_reserve(4 * f4 + (size_t)npts);
        std::vector<std::tuple<void *, const void *, size_t>> spread = {
            {out->offR, stg, f4}, {out->offA, stg + f4, f4}, {out->corr, stg + 2 * f4, f4},
            {out->type, stg + 3 * f4, (size_t)npts}};
        ST_HIP(hipMemcpyAsync(stg, dOffR, f4, hipMemcpyDeviceToHost, st));
        ST_HIP(hipMemcpyAsync(stg + f4, dOffA, f4, hipMemcpyDeviceToHost, st));
        ST_HIP(hipMemcpyAsync(stg + 2 * f4, S.corr, f4, hipMemcpyDeviceToHost, st));
        ST_HIP(hipMemcpyAsync(stg + 3 * f4, S.type, npts, hipMemcpyDeviceToHost, st));
        if (out->azd) {
            ST_HIP(hipMemcpyAsync(stg + 3 * f4 + npts, S.azd, f4, hipMemcpyDeviceToHost, st));
            spread.emplace_back(out->azd, stg + 3 * f4 + npts, f4);
        }
        if (diag) {
            auto cp = [&](void *h, const void *d, size_t bytes) {
                if (h)
                    ST_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, st));
            };
            cp(diag->r2, S.r2, npts * 4);
            cp(diag->a2, S.a2, npts * 4);
            for (int t = 0; t < 3; t++) {
                size_t o = (size_t)t * npts;
                cp(diag->peakI[t], S.peakI + o, npts * 4);
                cp(diag->peakJ[t], S.peakJ + o, npts * 4);
                cp(diag->fineI[t], S.fineI + o, npts * 4);
                cp(diag->fineJ[t], S.fineJ + o, npts * 4);
                cp(diag->cmax[t], S.cmax + o, npts * 8);
                cp(diag->sigRg[t], S.sigRg + o, npts * 4);
                cp(diag->sigAz[t], S.sigAz + o, npts * 4);
                cp(diag->peakGap[t], S.peakGap + o, npts * 4);
                cp(diag->fineGap[t], S.fineGap + o, npts * 4);
                cp(diag->snr[t], S.snr + o, npts * 4);
            }
        }
        std::vector<uint8_t> azv;
        std::vector<int> azs;
        if (diag && diag->azShift) {
            azv.resize(npts);
            azs.resize(npts);
            ST_HIP(hipMemcpyAsync(azv.data(), S.azValid, npts, hipMemcpyDeviceToHost, st));
            ST_HIP(hipMemcpyAsync(azs.data(), S.azShift, npts * 4, hipMemcpyDeviceToHost, st));
        }
        ST_HIP(hipEventRecord(sev[5], st));
        ST_HIP(hipStreamSynchronize(st));
        parallel_copy(spread);
        ctx->flush_timings();
        {
            float ms;
            ST_HIP(hipEventElapsedTime(&ms, sev[0], sev[1]));
            ctx->stats.stageMs[0] = ms;
            ST_HIP(hipEventElapsedTime(&ms, sev[1], sev[2]));
            ctx->stats.stageMs[1] = ms;
            ST_HIP(hipEventElapsedTime(&ms, sev[2], sev[3]));
            ctx->stats.stageMs[2] = ms;
            ST_HIP(hipEventElapsedTime(&ms, sev[3], sev[4]));
            ctx->stats.stageMs[3] = ms;
            ST_HIP(hipEventElapsedTime(&ms, sev[0], sev[5]));
            ctx->stats.totalMs = ms;
            ctx->stats.nPoints += npts;
            for (int t = 0; t < 3; t++) {
                ctx->stats.nCorr[t] += counts[CNT_CORR0 + t];
                patches += counts[CNT_CORR0 + t];
            }
        }
        if (diag && diag->azShift)
            for (int k = 0; k < npts; k++)
                diag->azShift[k] = azv[k] ? azs[k] : -1;
        out->nComplex = counts[CNT_COMPLEX];
        out->nAmp = counts[CNT_AMP];
        out->nAmpL = counts[CNT_AMPL];
        out->nFail = counts[CNT_FAIL];
    }
};

// The fused launches of a batch: one per window class (and sample type,
// correction set), over every call of that class.
void run_fused_group(st_ctx *ctx, std::vector<StrackJob *> &group)
{
    std::vector<FusedJob> fj;
    for (StrackJob *j : group)
        fj.push_back(j->fused_part());
    const int w = group[0]->w, slot = group[0]->kPeak;
    if (w == 16)
        launch_cpx_fused<16>(ctx, slot, fj);
    else if (w == 32)
        launch_cpx_fused<32>(ctx, slot, fj);
    else if (w == 64)
        launch_cpx_fused<64>(ctx, slot, fj);
    else
        launch_cpx_fused<kF128>(ctx, slot, fj);
    for (StrackJob *j : group)
        j->fused_finish();
}

// Reserve the arena for a batch: every call's fixed part and fused
// results, plus stage scratch for the largest staged / amplitude batch; a
// budget the device cannot hold is halved until it fits.
void reserve_batch(st_ctx *ctx, std::vector<StrackJob *> &jobs)
{
    size_t budget = ctx->scratchBudget;
    const size_t kMaxBatch = 8192;  // keeps 2 x batch inside the 65535 grid-z limit of the transposes
    for (;;) {
        size_t need = 65536, stage = 0;
        bool all1 = true;
        for (StrackJob *j : jobs) {
            j->Bc = (int)std::max<size_t>(1, std::min<size_t>({(size_t)j->npts, budget / j->perCpx, kMaxBatch}));
            j->Ba = (int)std::max<size_t>(1, std::min<size_t>({(size_t)j->npts, budget / j->perAmp, kMaxBatch}));
            all1 = all1 && j->Bc == 1 && j->Ba == 1;
            need += j->fixed + j->fused_bytes() + 2048;
            stage = std::max(stage, std::max((size_t)j->Bc * j->perCpx, (size_t)j->Ba * j->perAmp));
        }
        need += stage;
        try {
            ctx->arena_reserve(need);
            return;
        } catch (const StError &e) {
            if (e.code != ST_ERR_NOMEM || all1)
                throw;
            budget /= 2;
        }
    }
}

// The calls of a batch, start to end.  Results of each call are those it
// gives alone: every kernel computes its items independently.
void run_strack_batch(st_ctx *ctx, std::vector<StrackJob *> &jobs)
{
    init_dirichlet(ctx);
    ctx->stats_reset();
    for (StrackJob *j : jobs)
        j->plan();
    reserve_batch(ctx, jobs);
    ST_HIP(hipEventRecord(ctx->ev[64], ctx->stream));
    for (StrackJob *j : jobs)
        j->setup();
    const size_t mark = ctx->arena.used;
    for (StrackJob *j : jobs)
        j->mark = mark;
    // the fused calls, grouped by window, sample type and corrections
    std::map<std::tuple<int, int, int>, std::vector<StrackJob *>> groups;
    for (StrackJob *j : jobs)
        if (j->fusedPath && !j->mfmaPath && j->nCpx > 0)
            groups[std::make_tuple(j->w, j->pr.stype, (j->P.baseFlag ? 1 : 0) | (j->P.intOn ? 2 : 0))].push_back(j);
    for (auto &g : groups)
        run_fused_group(ctx, g.second);
    ctx->arena.used = mark;
    for (StrackJob *j : jobs) {
        j->complex_alone();
        j->finish();
    }
}

}  // namespace

void st_speckle_track_impl(st_ctx *ctx, StPair &pr, const st_track_params &par, st_track_out *out,
                           st_track_diag *diag, st_cull_params *cull)
{
    StrackJob job(ctx, pr, par, out, diag, cull);
    std::vector<StrackJob *> jobs{&job};
    run_strack_batch(ctx, jobs);
}

void st_speckle_track_batch_impl(st_ctx *ctx, int n, StPair *const *prs, const st_track_params *const *pars,
                                 st_track_out *const *outs, st_cull_params *const *culls, int64_t *patches)
{
    std::vector<std::unique_ptr<StrackJob>> own;
    std::vector<StrackJob *> jobs;
    for (int k = 0; k < n; k++) {
        own.emplace_back(new StrackJob(ctx, *prs[k], *pars[k], outs[k], nullptr, culls ? culls[k] : nullptr));
        jobs.push_back(own.back().get());
    }
    run_strack_batch(ctx, jobs);
    if (patches)
        for (int k = 0; k < n; k++)
            patches[k] = jobs[k]->patches;
}

int fused_occupancy(int w, int stype) { return fused_per_cu_w(w, stype); }

bool strack_batchable(const st_track_params &p)
{
    return p.wR == p.wA && p.wRa == p.wAa && (p.wR == 16 || p.wR == 32 || p.wR == 64 || p.wR == kF128) &&
           p.precision == ST_PREC_FP32 && !p.gaussFlag && !p.intf && !p.baseFlag && !p.noComplex && !staged_only();
}
