// strackw on the GPU: the edge-padded normalised cross-correlation tracker
// of Strackw/corrTrackFast.c (corrTrackFast :89-231, getCorrPatchesFast
// :299-372, detectPatch :765-839, computeMeanSigma :407-477, ampMatchEdge
// :903-947, getPeakCorr :862-899, correlateFast + overSampleC :482-551,
// 951-990, corrMatch :376-404).
//
// Patches are processed in batches; every stage is a batched kernel or a
// batched FFT on the context's engine:
//   gather + Doppler -> FFT wa^2 (x2) -> rotated 2x spectral pad -> inverse
//   FFT (2wa)^2 (x2) -> detection with boxcar into one packed complex image
//   per patch (search + i chip) -> window statistics -> one forward FFT per
//   patch -> cross spectra of two patches packed into one array -> one
//   inverse FFT per two patches -> NCC over all lags + argmax -> 21 x 21 ->
//   64 x 64 Dirichlet refinement in FP64 (the 336 x 336 inverse FFT of the
//   reference evaluated only on the searched window).
// The search and chip images are real, so one complex transform carries
// both (spectra separated by Hermitian symmetry), and the real correlation
// surfaces of two patches share one inverse transform.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "st_ctx.h"
#include "st_fftlds.h"
#include "st_track_common.h"

using namespace sttrack;

namespace {

constexpr int WNOVER = 16;  // corrTrackFast.c:11
constexpr int WNFAST = 20;  // corrTrackFast.c:12
constexpr int WOS = 2;      // corrTrackFast.c:16
constexpr int kN21 = WNFAST + 1;
constexpr int kFineLo = (kN21 * WNOVER) / 2 - 2 * WNOVER;  // 136: searched oversampled window
constexpr int kFineN = 4 * WNOVER;                         // 64
constexpr int kTile = 32;
// Boxcar half-widths up to 24: k_w_detect's two haloed tiles,
// (32 + 2 * 24)^2 floats each, stay within 64 KB of LDS with its output tile.
constexpr int kMaxNavg = 24;
constexpr int kAmpMatch = 2;
// Dirichlet kernel of the 21 -> 336 zero pad, D(t), t = kDirLo .. kDirLo + kDirN - 1
constexpr int kDirLo = kFineLo - WNOVER * WNFAST;  // -184
constexpr int kDirN = kFineN + WNOVER * WNFAST;    // 384

__constant__ double c_dirw[kDirN];

struct WPar {
    Images im;
    int wRa, wAa, W_R, W_A, epR, epA, wR2, wA2, nLagR, nLagA;
    stfft::FastDiv dwRa, dWR, dWA;  // index splits in the interpolation kernels
    int navgR, navgA;
    float scale;
    double scaleW;
};

struct WRes {
    float cMax, rShift1, aShift1, gapCoarse, gapFine;
    int pI, pJ, fI, fJ, az;
};

struct RowBest {
    double v, second;
    int rank, pad;
};

struct WBatch {
    const int4 *pos;  // r1, a1, r2, a2 of every listed patch
    float2 *T;        // [2][B][wAa * W_R]: range-interpolated rows
    double *dop;      // [2][B][nblk][4]: Doppler partial sums per row block
    double *dopE;     // [2][B][4]: first- and last-row sums
    float2 *U;        // [2][B][W_A * W_R]; later X: [ceil(B/2)][W_A * W_R]
    float2 *Z;        // [B][W_A * W_R]: search + i chip
    float2 *Zs;       // [B][W_A * W_R]: its spectrum
    double *chip;     // [B][2]: chip mean, variance
    double *csum;     // [B][ceil(W_A/32)][2][W_R]: column sums of s, s^2 over 32-row tiles
    float *cr;        // [B][nLagA][nLagR]: NCC surface
    RowBest *rows;    // [B][nLagA]
    int *az;          // [B]
    WRes *res;        // [B]
};

__device__ inline float2 sample0(const Images &im, bool second, int a, int r)
{
    const int nr = second ? im.nr2 : im.nr1, na = second ? im.na2 : im.na1;
    if (a < 0 || a >= na || r < 0 || r >= nr)
        return make_float2(0.f, 0.f);
    return load_sample(second ? im.img2 : im.img1, im.stype, nr, a, r);
}

constexpr int kRowElems = 4096;  // register staging and LDS of k_w_rowinterp (32 KB)
constexpr int kColElems = 4096;  // same for k_w_colinterp (32 KB + one pad element per column)

// getCorrPatchesFast + zeroPad (corrTrackFast.c:299-372, 738-761), range
// half.  The reference transforms each window (wAa x wRa), moves the
// spectrum into the corners of a 2wAa x 2wRa array (azimuth rows rotated by
// the Doppler shift, Nyquist bins on the negative side) and inverse
// transforms it.  Transform, pad and inverse are separable, so each block
// takes R rows of one window (samples outside the image read as zero),
// transforms them (wRa), pads each row spectrum to 2wRa and inverse
// transforms it (2wRa); k_w_colinterp does the azimuth half.  The block also
// leaves the Doppler partial sums of its rows: sum of x, sum of
// x[i+1] conj(x[i]) over its rows i with a successor, and the first and last
// window rows' sums.
__global__ void __launch_bounds__(NT) k_w_rowinterp(WPar p, int off, WBatch Bt, int B, int R, stfft::PlanArgs Pf,
                                                     stfft::PlanArgs Pi)
{
    extern __shared__ float2 ldsf[];
    float2 *lds = ldsf;
    __shared__ double sh[NT / 64];
    const int z = blockIdx.y, img = z >= B ? 1 : 0, b = z - img * B;
    const int4 q = Bt.pos[off + b];
    const int rr = img ? q.z : q.x, ra = img ? q.w : q.y;
    const int i0 = blockIdx.x * R, nrow = imin(R, p.wAa - i0);
    const int pitch = p.W_R;
    double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // S, L, F0, FL (re, im)
    {
        constexpr int IT = kRowElems / NT;
        const int cnt = nrow * p.wRa;
        float2 x[IT];
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = threadIdx.x + it * NT, i = stfft::fdiv(t, p.dwRa), j = t - i * p.wRa;
            x[it] = t < cnt ? sample0(p.im, img, ra + i0 + i, rr + j) : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = threadIdx.x + it * NT, i = stfft::fdiv(t, p.dwRa), j = t - i * p.wRa;
            if (t < cnt) {
                lds[i * pitch + j] = x[it];
                s[0] += x[it].x;
                s[1] += x[it].y;
                if (i0 + i == 0) {
                    s[4] += x[it].x;
                    s[5] += x[it].y;
                }
                if (i0 + i == p.wAa - 1) {
                    s[6] += x[it].x;
                    s[7] += x[it].y;
                }
            }
        }
    }
    __syncthreads();
    const int nlag = imin(nrow, p.wAa - 1 - i0);
    for (int t = threadIdx.x; t < nlag * p.wRa; t += NT) {
        const int i = stfft::fdiv(t, p.dwRa), j = t - i * p.wRa;
        const float2 x0 = lds[i * pitch + j];
        const float2 x1 = i + 1 < nrow ? lds[(i + 1) * pitch + j] : sample0(p.im, img, ra + i0 + i + 1, rr + j);
        s[2] += (double)x1.x * x0.x + (double)x1.y * x0.y;
        s[3] += (double)x1.y * x0.x - (double)x1.x * x0.y;
    }
    for (int k = 0; k < 8; k++)
        s[k] = block_sum_d<NT>(s[k], sh);
    if (threadIdx.x == 0) {
        double *d = Bt.dop + ((size_t)z * gridDim.x + blockIdx.x) * 4;
        for (int k = 0; k < 4; k++)
            d[k] = s[k];
        if (i0 == 0) {
            Bt.dopE[4 * z + 0] = s[4];
            Bt.dopE[4 * z + 1] = s[5];
        }
        if (i0 + nrow == p.wAa) {
            Bt.dopE[4 * z + 2] = s[6];
            Bt.dopE[4 * z + 3] = s[7];
        }
    }
    __syncthreads();
    stfft::lds_fft<kRowElems, NT>(lds, pitch, nrow, Pf, -1);
    // pad each row spectrum wRa -> 2wRa: bins 0 .. wRa/2-1 stay, the Nyquist
    // bin and the negative frequencies move to the top, zeros between
    {
        constexpr int IT = kRowElems / NT;
        const int hR = p.wRa / 2;
        float2 v[IT];
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = threadIdx.x + it * NT;
            if (t < nrow * p.W_R) {
                const int i = stfft::fdiv(t, p.dWR), m = t - i * p.W_R;
                const int l = m < hR ? m : (m >= p.W_R - hR ? m - p.wRa : -1);
                v[it] = l >= 0 ? lds[i * pitch + l] : make_float2(0.f, 0.f);
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = threadIdx.x + it * NT;
            if (t < nrow * p.W_R)
                lds[t] = v[it];
        }
        __syncthreads();
    }
    stfft::lds_fft<kRowElems, NT>(lds, pitch, nrow, Pi, +1);
    float2 *Tz = Bt.T + ((size_t)z * p.wAa + i0) * p.W_R;
    for (int t = threadIdx.x; t < nrow * p.W_R; t += NT)
        Tz[t] = lds[t];
}

// The Doppler shift of each patch (getCorrPatchesFast, corrTrackFast.c:
// 340-350, estDopCarrier1, dopplerEstimation.c:25-88) from the row-block
// partial sums: sum (x[i+1] - m) conj(x[i] - m) over the lag pairs is
// L - conj(m) (S - F0) - m conj(S - FL) + (wAa - 1) wRa |m|^2.
__global__ void k_w_az(WPar p, WBatch Bt, int B, int nblk)
{
    const int b = blockIdx.x, lane = threadIdx.x;
    double ang[2];
    for (int img = 0; img < 2; img++) {
        const int z = img * B + b;
        double a[4] = {0, 0, 0, 0};
        for (int k = lane; k < nblk; k += 64)
            for (int e = 0; e < 4; e++)
                a[e] += Bt.dop[((size_t)z * nblk + k) * 4 + e];
        for (int o = 32; o > 0; o >>= 1)
            for (int e = 0; e < 4; e++)
                a[e] += __shfl_xor(a[e], o);
        const double cnt = (double)p.wAa * (double)p.wRa;
        const double mr = a[0] / cnt, mi = a[1] / cnt;
        const double *E = Bt.dopE + 4 * z;
        const double s1r = a[0] - E[0], s1i = a[1] - E[1];  // x over rows 1 .. wAa-1
        const double s0r = a[0] - E[2], s0i = a[1] - E[3];  // x over rows 0 .. wAa-2
        const double zr = a[2] - (mr * s1r + mi * s1i) - (mr * s0r + mi * s0i) +
                          (double)(p.wAa - 1) * (double)p.wRa * (mr * mr + mi * mi);
        const double zi = a[3] - (mr * s1i - mi * s1r) - (mi * s0r - mr * s0i);
        ang[img] = (zr == 0.0 && zi == 0.0) ? NAN : atan2(zi, zr);
    }
    if (lane != 0)
        return;
    const double v = (double)(p.wAa / 2) * (ang[0] + ang[1]) * .5 / 3.14159265358979323846;
    int shift = isnan(v) ? 0 : ((v >= 0.0) ? (int)(v + 0.5) : (int)(v - 0.5));
    if (shift < 0)
        shift += p.wAa;
    int az = (int)((float)shift * (float)p.wAa / (float)p.wAa);
    if (az < 0 || az >= p.wAa)
        az = 0;
    Bt.az[b] = az;
}

// Azimuth half of the interpolation: C adjacent columns of the
// range-interpolated rows of one window, transformed (wAa), rotated by the
// Doppler shift and zero-padded to 2wAa (Nyquist on the negative side),
// inverse transformed (2wAa).  Work items are laid out so that the column
// blocks sharing cache lines go to the same XCD (dispatch is round-robin
// over the 8 XCDs).
__global__ void __launch_bounds__(NT) k_w_colinterp(WPar p, WBatch Bt, int B, int lc, stfft::PlanArgs Pf,
                                                     stfft::PlanArgs Pi)
{
    const int C = 1 << lc;
    extern __shared__ float2 ldsf[];
    float2 *lds = ldsf;
    const int nbx = gridDim.x, total = nbx * gridDim.y, flat = blockIdx.y * nbx + blockIdx.x;
    const int per = total / 8;
    const int L = flat < per * 8 ? (flat % 8) * per + flat / 8 : flat;
    const int z = L / nbx, cb = L - (L / nbx) * nbx;
    const int b = z >= B ? z - B : z;
    const int c0 = cb * C, nc = imin(C, p.W_R - c0);
    const int pitch = p.W_A + 1;
    const float2 *Tz = Bt.T + (size_t)z * p.wAa * p.W_R + c0;
    {
        constexpr int IT = kColElems / NT;
        const int cnt = p.wAa << lc;
        float2 x[IT];
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = threadIdx.x + it * NT, i = t >> lc, c = t & (C - 1);
            x[it] = t < cnt && c < nc ? Tz[(size_t)i * p.W_R + c] : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = threadIdx.x + it * NT, i = t >> lc, c = t & (C - 1);
            if (t < cnt)
                lds[c * pitch + i] = x[it];
        }
    }
    __syncthreads();
    stfft::lds_fft<kColElems, NT>(lds, pitch, C, Pf, -1);
    {
        constexpr int IT = kColElems / NT;
        const int hA = p.wAa / 2, az = Bt.az[b];
        float2 v[IT];
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = threadIdx.x + it * NT;
            if (t < C * p.W_A) {
                const int c = stfft::fdiv(t, p.dWA), n = t - c * p.W_A;
                const int k = n < hA ? n : (n >= p.W_A - hA ? n - p.W_A : (1 << 20));
                int ka = k + p.wAa + az;  // in [0, 3 wAa)
                ka -= ka >= p.wAa ? p.wAa : 0;
                ka -= ka >= p.wAa ? p.wAa : 0;
                v[it] = k != (1 << 20) ? lds[c * pitch + ka] : make_float2(0.f, 0.f);
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = threadIdx.x + it * NT;
            if (t < C * p.W_A) {
                const int c = stfft::fdiv(t, p.dWA), n = t - c * p.W_A;
                lds[c * pitch + n] = v[it];
            }
        }
        __syncthreads();
    }
    stfft::lds_fft<kColElems, NT>(lds, pitch, C, Pi, +1);
    float2 *Uz = Bt.U + (size_t)z * p.W_A * p.W_R + c0;
    {
        constexpr int IT = kColElems / NT;
        const int cnt = p.W_A << lc;
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = threadIdx.x + it * NT, n = t >> lc, c = t & (C - 1);
            if (t < cnt && c < nc)
                Uz[(size_t)n * p.W_R + c] = lds[c * pitch + n];
        }
    }
}

#pragma clang fp contract(off)
// detectPatch (corrTrackFast.c:765-839) of the search (image 1, whole
// array, boxcar clamped to the array) and the chip (image 2, inner region,
// boxcar clamped as the reference does, with the azimuth bound written in
// terms of edgePadR); sqrt; packed as search + i chip with the chip's
// edge-pad frame zero.
__global__ void __launch_bounds__(256) k_w_detect(WPar p, WBatch Bt, int B)
{
    extern __shared__ float w_det_lds[];  // the two haloed tiles, sized by detect_lds_bytes
    __shared__ float sv[kTile][kTile + 1];
    const int b = blockIdx.z;
    const int NN = p.W_A * p.W_R;
    const float2 *U1 = Bt.U + (size_t)b * NN, *U2 = Bt.U + ((size_t)B + b) * NN;
    const bool box = p.navgA > 1 || p.navgR > 1;
    const int hA = box ? p.navgA : 0, hR = box ? p.navgR : 0;
    const int y0 = blockIdx.y * kTile, x0 = blockIdx.x * kTile;
    const int TH = kTile + 2 * hA, TW = kTile + 2 * hR;
    float *d1 = w_det_lds, *d2 = w_det_lds + TH * TW;
    const int c0 = p.epA * WOS, c1 = c0 + p.wA2;  // chip rows [c0, c1)
    const int e0 = p.epR * WOS, e1 = e0 + p.wR2;  // chip cols [e0, e1)
    const bool chipTile = y0 < c1 && y0 + kTile > c0 && x0 < e1 && x0 + kTile > e0;
    const int kmaxC = p.epR * WOS + p.wA2 + p.navgA - 1;
    const int lmaxC = p.epR * WOS + p.wR2 + p.navgR - 1;
    for (int t = threadIdx.x; t < TH * TW; t += blockDim.x) {
        const int ty = t / TW, tx = t - (t / TW) * TW;
        const int y = y0 - hA + ty, x = x0 - hR + tx;
        const int ky = imin(imax(y, 0), p.W_A - 1), kx = imin(imax(x, 0), p.W_R - 1);
        const float2 z1 = U1[(size_t)ky * p.W_R + kx];
        d1[t] = (z1.x * z1.x + z1.y * z1.y) * p.scale;
        if (chipTile) {
            const int ky2 = imin(imax(y, 0), kmaxC), kx2 = imin(imax(x, 0), lmaxC);
            const float2 z2 = U2[(size_t)imin(ky2, p.W_A - 1) * p.W_R + imin(kx2, p.W_R - 1)];
            d2[t] = (z2.x * z2.x + z2.y * z2.y) * p.scale;
        }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < kTile * kTile; t += blockDim.x) {
        const int oy = t / kTile, ox = t - (t / kTile) * kTile;
        const int y = y0 + oy, x = x0 + ox;
        sv[oy][ox] = 0.f;
        if (y >= p.W_A || x >= p.W_R)
            continue;
        float v1 = 0.f;
        for (int m = -hA; m <= hA; m++)
            for (int n = -hR; n <= hR; n++)
                v1 += d1[(oy + hA + m) * TW + ox + hR + n];
        float s2 = 0.f;
        if (chipTile && y >= c0 && y < c1 && x >= e0 && x < e1) {
            float v2 = 0.f;
            for (int m = -hA; m <= hA; m++)
                for (int n = -hR; n <= hR; n++)
                    v2 += d2[(oy + hA + m) * TW + ox + hR + n];
            s2 = sqrtf(v2);
        }
        // sqrtf: the double square root rounded to float, bit for bit, for
        // every non-negative finite float (all of them checked on gfx950)
        const float sr = sqrtf(v1);
        sv[oy][ox] = sr;
        Bt.Z[(size_t)b * NN + (size_t)y * p.W_R + x] = make_float2(sr, s2);
    }
    // column sums of s and s^2 over this tile's rows (FP64), for the window
    // sums of k_w_ncc
    __syncthreads();
    if (threadIdx.x < kTile && x0 + (int)threadIdx.x < p.W_R) {
        double a1 = 0.0, a2 = 0.0;
        for (int oy = 0; oy < kTile; oy++) {
            const float v = sv[oy][threadIdx.x];
            a1 += v;
            a2 += v * v;
        }
        double *cs = Bt.csum + ((size_t)b * gridDim.y + blockIdx.y) * 2 * p.W_R + x0 + threadIdx.x;
        cs[0] = a1;
        cs[p.W_R] = a2;
    }
}

// k_w_detect's dynamic LDS: the haloed search and chip tiles (the boxcar
// runs only when a half-width exceeds 1, detectPatch's test).
size_t detect_lds_bytes(const WPar &p)
{
    const bool box = p.navgA > 1 || p.navgR > 1;
    const size_t th = kTile + 2 * (box ? p.navgA : 0), tw = kTile + 2 * (box ? p.navgR : 0);
    return 2 * th * tw * sizeof(float);
}

// Chip mean and variance (computeMeanSigma, corrTrackFast.c:418-431):
// sigma = sum(d^2)/(N-1) - mean^2.
__global__ void __launch_bounds__(NT) k_w_chipstats(WPar p, WBatch Bt)
{
    __shared__ double sh[NT / 64];
    const int b = blockIdx.x;
    const float2 *Zb = Bt.Z + (size_t)b * p.W_A * p.W_R;
    const int n = p.wA2 * p.wR2;
    double s1 = 0.0, s2 = 0.0;
    for (int e = threadIdx.x; e < n; e += NT) {
        const int i = e / p.wR2, j = e - (e / p.wR2) * p.wR2;
        const float d = Zb[(size_t)(p.epA * WOS + i) * p.W_R + p.epR * WOS + j].y;
        s1 += d;
        s2 += d * d;
    }
    s1 = block_sum_d<NT>(s1, sh);
    s2 = block_sum_d<NT>(s2, sh);
    if (threadIdx.x == 0) {
        const double mean = s1 / (double)n;
        double sig = s2 / (double)(n - 1.);
        sig -= mean * mean;
        Bt.chip[2 * b] = mean;
        Bt.chip[2 * b + 1] = sig;
    }
}

// ampMatchEdge (corrTrackFast.c:903-947) cross spectra.  Z = s + i c holds
// the search and chip images of one patch: S(k) = (Z(k) + Z*(-k)) / 2,
// C(k) = (Z(k) - Z*(-k)) / 2i, G = S conj(C).  Patches 2q and 2q+1 are
// packed as G0 + i G1, whose inverse transform is g0 + i g1 (both real).
__global__ void k_w_cross(WPar p, WBatch Bt, int B)
{
    const int q = blockIdx.y;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int NN = p.W_A * p.W_R;
    if (e >= NN)
        return;
    const int ky = e / p.W_R, kx = e - (e / p.W_R) * p.W_R;
    const int ny = ky == 0 ? 0 : p.W_A - ky, nx = kx == 0 ? 0 : p.W_R - kx;
    const int en = ny * p.W_R + nx;
    float2 G[2];
    for (int h = 0; h < 2; h++) {
        const int b = 2 * q + h;
        if (b >= B) {
            G[h] = make_float2(0.f, 0.f);
            continue;
        }
        const float2 a = Bt.Zs[(size_t)b * NN + e];
        const float2 c = Bt.Zs[(size_t)b * NN + en];  // Z(-k); conj below
        const float2 S = make_float2(0.5f * (a.x + c.x), 0.5f * (a.y - c.y));
        const float2 Cc = make_float2(0.5f * (a.y + c.y), 0.5f * (c.x - a.x));
        G[h] = cmulconj(S, Cc);
    }
    Bt.U[(size_t)q * NN + e] = make_float2(G[0].x - G[1].y, G[0].y + G[1].x);
}

constexpr int kNccRows = 32;  // lag rows per k_w_ncc block (a multiple of the 32-row tiles)
constexpr int kNccStep = 4;   // lag rows per round of block barriers
static_assert(kNccRows % kTile == 0, "NCC runs start on tile rows");

// Top-2 merge with the first-rank tie-break.
__device__ inline void top2_merge(double &best, double &second, int &rank, double ob, double os, int orank)
{
    const bool takeOther = ob > best || (ob == best && orank < rank);
    const double lo = takeOther ? best : ob;
    const double hs = takeOther ? os : second;
    second = fmax(hs, lo);
    if (takeOther) {
        best = ob;
        rank = orank;
    }
}

// computeMeanSigma (corrTrackFast.c:407-477, search part) and getPeakCorr
// (:862-899) for a run of kNccRows lag rows of one patch, nStep <= kNccStep
// rows per round of block barriers.  Thread t owns columns t + 256 k,
// k < PER.  The vertical window sums of s and s^2 (FP64, wA2 rows) start
// from the 32-row tile column sums of k_w_detect (plus the rows of a last
// partial tile) and slide down one row at a time; each row's horizontal
// window sums come from an exclusive prefix over the columns (in
// LDS, one slice per row of the round).  NCC in FP64; best and runner-up
// within the searched lags (rank = row-major lag index, the reference's
// first-maximum rule) per row.
template <int PER>
__global__ void __launch_bounds__(256) k_w_ncc(WPar p, WBatch Bt, int nRun, int nStep)
{
    extern __shared__ double lds[];
    __shared__ double sh[4][2 * kNccStep];
    __shared__ RowBest rb[4][kNccStep];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int b = blockIdx.y, run = blockIdx.x;
    const int i0 = run * kNccRows, i1 = imin(i0 + kNccRows, p.nLagA);
    const int slice = 2 * (p.W_R + 1);
    const float2 *Zb = Bt.Z + (size_t)b * p.W_A * p.W_R;
    const int NN = p.W_A * p.W_R;
    const float2 *Xb = Bt.U + (size_t)(b >> 1) * NN;
    const bool odd = b & 1;
    const int nTile = (p.W_A + kTile - 1) / kTile;
    const double *cs = Bt.csum + (size_t)b * nTile * 2 * p.W_R;
    const int jA = i0 / kTile, jB = (i0 + p.wA2) / kTile;
    double v1[PER], v2[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int c = threadIdx.x + 256 * k;
        v1[k] = 0.0;
        v2[k] = 0.0;
        if (c < p.W_R) {
            for (int j = jA; j < jB; j++) {
                v1[k] += cs[(size_t)j * 2 * p.W_R + c];
                v2[k] += cs[(size_t)j * 2 * p.W_R + p.W_R + c];
            }
            for (int l = jB * kTile; l < i0 + p.wA2; l++) {
                const float d = Zb[(size_t)l * p.W_R + c].x;
                v1[k] += d;
                v2[k] += d * d;
            }
        }
    }
    const int per = (p.W_R + 255) / 256;
    const int cb = threadIdx.x * per, ce = imin(cb + per, p.W_R);
    const double N = (double)p.wR2 * p.wA2;
    const double meanR = Bt.chip[2 * b], sigmaR = Bt.chip[2 * b + 1];
    for (int ib = i0; ib < i1; ib += nStep) {
        const int nr = imin(nStep, i1 - ib);
        // slide the window sums to each row of the round and park them
#pragma unroll
        for (int r = 0; r < kNccStep; r++) {
            const int i = ib + r;
            if (r < nr) {
                if (i > i0) {
                    const float2 *top = Zb + (size_t)(i - 1) * p.W_R, *bot = Zb + (size_t)(i - 1 + p.wA2) * p.W_R;
#pragma unroll
                    for (int k = 0; k < PER; k++) {
                        const int c = threadIdx.x + 256 * k;
                        if (c < p.W_R) {
                            const float e = top[c].x, d = bot[c].x;
                            v1[k] -= e;
                            v2[k] -= e * e;
                            v1[k] += d;
                            v2[k] += d * d;
                        }
                    }
                }
                double *P1 = lds + r * slice, *P2 = P1 + (p.W_R + 1);
#pragma unroll
                for (int k = 0; k < PER; k++) {
                    const int c = threadIdx.x + 256 * k;
                    if (c < p.W_R) {
                        P1[c] = v1[k];
                        P2[c] = v2[k];
                    }
                }
            }
        }
        __syncthreads();
        // exclusive prefixes over the columns: P[c] = sum_{t < c} V[t]
        double a[2 * kNccStep], x[2 * kNccStep];
#pragma unroll
        for (int r = 0; r < kNccStep; r++) {
            const double *P1 = lds + r * slice, *P2 = P1 + (p.W_R + 1);
            a[2 * r] = 0.0;
            a[2 * r + 1] = 0.0;
            if (r < nr)
                for (int c = cb; c < ce; c++) {
                    a[2 * r] += P1[c];
                    a[2 * r + 1] += P2[c];
                }
            x[2 * r] = a[2 * r];
            x[2 * r + 1] = a[2 * r + 1];
        }
        for (int o = 1; o < 64; o <<= 1) {
#pragma unroll
            for (int q = 0; q < 2 * kNccStep; q++) {
                const double u = __shfl_up(x[q], o);
                if (lane >= o)
                    x[q] += u;
            }
        }
        if (lane == 63)
#pragma unroll
            for (int q = 0; q < 2 * kNccStep; q++)
                sh[wid][q] = x[q];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 2 * kNccStep; q++) {
            double pre = 0.0;
            for (int w = 0; w < wid; w++)
                pre += sh[w][q];
            x[q] += pre - a[q];  // exclusive prefix at the start of the chunk
        }
#pragma unroll
        for (int r = 0; r < kNccStep; r++) {
            if (r < nr) {
                double *P1 = lds + r * slice, *P2 = P1 + (p.W_R + 1);
                double r1 = x[2 * r], r2 = x[2 * r + 1];
                for (int c = cb; c < ce; c++) {
                    const double u1 = P1[c], u2 = P2[c];
                    P1[c] = r1;
                    P2[c] = r2;
                    r1 += u1;
                    r2 += u2;
                }
                if (threadIdx.x == 255) {
                    P1[p.W_R] = r1;
                    P2[p.W_R] = r2;
                }
            }
        }
        __syncthreads();
        double best[kNccStep], second[kNccStep];
        int rank[kNccStep];
#pragma unroll
        for (int r = 0; r < kNccStep; r++) {
            best[r] = -1e30;
            second[r] = -1e30;
            rank[r] = 0x7fffffff;
            const int i = ib + r;
            if (r < nr) {
                const double *P1 = lds + r * slice, *P2 = P1 + (p.W_R + 1);
                const int l = i - WOS * p.epA;
                const bool rowIn = l >= -WOS * (p.epA - WNFAST / 4) && l <= WOS * (p.epA - WNFAST / 4);
                const float2 *X = Xb + (size_t)((l + p.W_A) % p.W_A) * p.W_R;
                float *crRow = Bt.cr + ((size_t)b * p.nLagA + i) * p.nLagR;
                for (int s = threadIdx.x; s < p.nLagR; s += 256) {
                    const double m = (P1[s + p.wR2] - P1[s]) / N;
                    const double sig = (P2[s + p.wR2] - P2[s]) / (N - 1.) - m * m;
                    const int c = s - WOS * p.epR;
                    const float2 xv = X[(c + p.W_R) % p.W_R];
                    const float v = odd ? xv.y : xv.x;
                    const float mag = sqrtf(v * v);  // = (float)sqrt((double)(v * v)), bit for bit
                    double cr = (double)mag * p.scaleW;
                    cr = cr - meanR * m;
                    cr = cr / sqrt(sigmaR * sig);
                    crRow[s] = (float)cr;
                    if (rowIn && c >= -WOS * (p.epR - WNFAST / 4) && c <= WOS * (p.epR - WNFAST / 4)) {
                        if (cr > best[r]) {
                            second[r] = best[r];
                            best[r] = cr;
                            rank[r] = i * p.nLagR + s;
                        } else if (cr > second[r]) {
                            second[r] = cr;
                        }
                    }
                }
            }
        }
        // top-2 with first-rank tie-break across the block
        for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
            for (int r = 0; r < kNccStep; r++) {
                const double ob = __shfl_xor(best[r], o), os = __shfl_xor(second[r], o);
                const int orank = __shfl_xor(rank[r], o);
                top2_merge(best[r], second[r], rank[r], ob, os, orank);
            }
        }
        if (lane == 0)
#pragma unroll
            for (int r = 0; r < kNccStep; r++)
                rb[wid][r] = RowBest{best[r], second[r], rank[r], 0};
        __syncthreads();
        if (threadIdx.x < nr) {
            const int r = threadIdx.x;
            RowBest q = rb[0][r];
            for (int w = 1; w < 4; w++)
                top2_merge(q.v, q.second, q.rank, rb[w][r].v, rb[w][r].second, rb[w][r].rank);
            Bt.rows[(size_t)b * p.nLagA + ib + r] = q;
        }
    }
}

// Peak over all lag rows, then correlateFast's refinement
// (corrTrackFast.c:482-551, 951-990): the 21 x 21 NCC neighbourhood is
// interpolated by its 16x spectral zero pad, evaluated in FP64 on the
// searched 64 x 64 window only, Q = A C A^T with A[u][m] = D(136 + u - 16 m).
__global__ void __launch_bounds__(256) k_w_refine(WPar p, WBatch Bt)
{
    __shared__ double dir[kDirN];
    __shared__ float c[kN21 * kN21];
    __shared__ double T[kN21 * kFineN];
    __shared__ ArgMax am[4];
    __shared__ double s2[4];
    __shared__ RowBest best;
    const int b = blockIdx.x;
    for (int k = threadIdx.x; k < kDirN; k += blockDim.x)
        dir[k] = c_dirw[k];
    if (threadIdx.x == 0) {
        RowBest r{-1e30, -1e30, 0x7fffffff, 0};
        const RowBest *rows = Bt.rows + (size_t)b * p.nLagA;
        for (int i = 0; i < p.nLagA; i++) {
            const RowBest o = rows[i];
            const bool takeOther = o.v > r.v || (o.v == r.v && o.rank < r.rank);
            const double lo = takeOther ? r.v : o.v;
            const double hs = takeOther ? o.second : r.second;
            r.second = fmax(hs, lo);
            if (takeOther) {
                r.v = o.v;
                r.rank = o.rank;
            }
        }
        best = r;
    }
    __syncthreads();
    const RowBest r = best;
    const int iMax = r.rank == 0x7fffffff ? 0 : r.rank / p.nLagR;
    const int jMax = r.rank == 0x7fffffff ? 0 : r.rank - (r.rank / p.nLagR) * p.nLagR;
    WRes out{};
    out.pI = iMax;
    out.pJ = jMax;
    out.fI = out.fJ = -2000;
    out.az = Bt.az[b];
    out.gapCoarse = r.v > 0 ? (float)((r.v - r.second) / r.v) : 0.f;
    if (!(r.v >= 0.0)) {  // maxCorr < 0 (or nothing defined): no match
        if (threadIdx.x == 0) {
            out.cMax = 0.f;
            Bt.res[b] = out;
        }
        return;
    }
    const float *cr = Bt.cr + (size_t)b * p.nLagA * p.nLagR;
    for (int k = threadIdx.x; k < kN21 * kN21; k += blockDim.x) {
        const int m = k / kN21, n = k - (k / kN21) * kN21;
        c[k] = cr[(size_t)(iMax - WNFAST / 2 + m) * p.nLagR + jMax - WNFAST / 2 + n];
    }
    __syncthreads();
    for (int k = threadIdx.x; k < kN21 * kFineN; k += blockDim.x) {
        const int m = k / kFineN, v = k - (k / kFineN) * kFineN;
        double acc = 0.0;
        for (int n = 0; n < kN21; n++)
            acc += (double)c[m * kN21 + n] * dir[kFineLo + v - WNOVER * n - kDirLo];
        T[k] = acc;
    }
    __syncthreads();
    ArgMax a{-1.f, 0x7fffffff};
    double sec = -1.0;
    for (int k = threadIdx.x; k < kFineN * kFineN; k += blockDim.x) {
        const int u = k / kFineN, v = k - (k / kFineN) * kFineN;
        double acc = 0.0;
        for (int m = 0; m < kN21; m++)
            acc += dir[kFineLo + u - WNOVER * m - kDirLo] * T[m * kFineN + v];
        const float q = (float)acc;
        const float pw = q * q;
        if (pw > a.v || (pw == a.v && k < a.rank)) {
            if (a.v >= 0.f)
                sec = fmax(sec, (double)a.v);
            a.v = pw;
            a.rank = k;
        } else {
            sec = fmax(sec, (double)pw);
        }
    }
    // block argmax (first maximum) and runner-up
    {
        for (int o = 32; o > 0; o >>= 1) {
            ArgMax o2;
            o2.v = __shfl_xor(a.v, o);
            o2.rank = __shfl_xor(a.rank, o);
            const double os = __shfl_xor(sec, o);
            const bool takeOther = o2.v > a.v || (o2.v == a.v && o2.rank < a.rank);
            const double lo = takeOther ? (double)a.v : (double)o2.v;
            const double hs = takeOther ? os : sec;
            sec = fmax(hs, lo);
            if (takeOther)
                a = o2;
        }
        const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
        if (lane == 0) {
            am[wid] = a;
            s2[wid] = sec;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            ArgMax A = am[0];
            double S = s2[0];
            for (int w = 1; w < 4; w++) {
                const bool takeOther = am[w].v > A.v || (am[w].v == A.v && am[w].rank < A.rank);
                const double lo = takeOther ? (double)A.v : (double)am[w].v;
                const double hs = takeOther ? s2[w] : S;
                S = fmax(hs, lo);
                if (takeOther)
                    A = am[w];
            }
            const int u = A.rank / kFineN, v = A.rank - (A.rank / kFineN) * kFineN;
            out.gapFine = A.v > 0.f ? (float)(((double)A.v - S) / (double)A.v) : 0.f;
            out.cMax = (float)(sqrt((double)A.v) / (double)(kN21 * kN21));
            const int iF = iMax * WNOVER + kFineLo + u - (WNOVER * WNFAST) / 2;
            const int jF = jMax * WNOVER + kFineLo + v - (WNOVER * WNFAST) / 2;
            out.fI = iF;
            out.fJ = jF;
            out.rShift1 = ((float)jF - (float)(p.epR * WOS * WNOVER)) / (float)(WNOVER * WOS);
            out.aShift1 = ((float)iF - (float)(p.epA * WOS * WNOVER)) / (float)(WNOVER * WOS);
            Bt.res[b] = out;
        }
    }
}
#pragma clang fp contract(on)

// ---------------------------------------------------------------------
// host side

// Plain bilinear interpolation of the initial-offset grid (the reference
// calls mosaicSource's bilinearInterp, corrTrackFast.c:851-852).
float bilinear(const float *g, double x, double y, int nx, int ny, float minvalue, float nodata)
{
    if (!(x >= 0.0 && y >= 0.0 && x <= nx - 1 && y <= ny - 1))
        return nodata;
    int ix = (int)x, iy = (int)y;
    if (ix > nx - 2)
        ix = nx - 2;
    if (iy > ny - 2)
        iy = ny - 2;
    if (ix < 0 || iy < 0)
        return nodata;
    const double t = x - ix, u = y - iy;
    const float p1 = g[(size_t)iy * nx + ix], p2 = g[(size_t)iy * nx + ix + 1];
    const float p3 = g[(size_t)(iy + 1) * nx + ix + 1], p4 = g[(size_t)(iy + 1) * nx + ix];
    if (p1 <= minvalue || p2 <= minvalue || p3 <= minvalue || p4 <= minvalue)
        return nodata;
    return (float)((1.0 - t) * (1.0 - u) * p1 + t * (1.0 - u) * p2 + t * u * p3 + (1.0 - t) * u * p4);
}

// findImage2Pos + getShifts of strackw (corrTrackFast.c:257-292, 843-858).
void image2_pos(const st_track_params &par, int r1, int a1, int *r2, int *a2)
{
    double aShift, rShift;
    const double sc = (double)par.scaleFactor;
    auto poly = [&]() {
        rShift = par.rShiftPoly[0] + (double)(r1 * sc) * par.rShiftPoly[1] + (double)(a1 * sc) * par.rShiftPoly[2];
        aShift = par.aShiftPoly[0] + (double)(r1 * sc) * par.aShiftPoly[1] + (double)(a1 * sc) * par.aShiftPoly[2];
    };
    if (par.polyShift || par.initDr == nullptr) {
        poly();
    } else {
        const float minvalue = (float)-ST_LARGEINT * 0.9f;
        const double range = (r1 - par.initR0) / par.initDeltaR;
        const double azimuth = (a1 - par.initA0) / par.initDeltaA;
        rShift = bilinear(par.initDr, range, azimuth, par.initNr, par.initNa, minvalue, (float)-ST_LARGEINT);
        aShift = bilinear(par.initDa, range, azimuth, par.initNr, par.initNa, minvalue, (float)-ST_LARGEINT);
        if (rShift <= minvalue || aShift <= minvalue) {
            rShift = minvalue;
            aShift = minvalue;
        }
        if (rShift < (-ST_LARGEINT + 1))
            poly();
    }
    rShift /= sc;
    aShift /= sc;
    *r2 = (rShift >= 0) ? r1 + (int)(rShift + 0.5) : r1 + (int)(rShift - 0.5);
    *a2 = (aShift >= 0) ? a1 + (int)(aShift + 0.5) : a1 + (int)(aShift - 0.5);
}

// maskValue of strackw (corrTrackFast.c:1141-1160)
int mask_value(const st_track_params &p, int r1, int a1)
{
    if (p.mask == nullptr)
        return 1;
    const int dr = (p.wRa / 2 - p.maskR0) * p.scaleFactor;
    const int da = (p.wAa / 2 - p.maskA0) * p.scaleFactor;
    const int nrl = p.maskNrl > 0 ? p.maskNrl : 1, nal = p.maskNal > 0 ? p.maskNal : 1;
    const int ia1 = (a1 + da) / nal, ir1 = (r1 + dr) / nrl;
    if (ia1 < 0 || ia1 >= p.maskNa || ir1 < 0 || ir1 >= p.maskNr)
        return 0;
    return (int)p.mask[(size_t)ia1 * p.maskNr + ir1];
}

void init_dirichlet_w(st_ctx *ctx)
{
    static std::mutex mu;
    static bool done[64] = {};
    std::lock_guard<std::mutex> lk(mu);
    ST_REQUIRE(ctx->device >= 0 && ctx->device < 64, ST_ERR_ARG, "device index");
    if (done[ctx->device])
        return;
    const int N = kN21 * WNOVER;
    double h[kDirN];
    for (int k = 0; k < kDirN; k++) {
        const int t = kDirLo + k;
        long double s = 1.0L;
        for (int f = 1; f <= WNFAST / 2; f++)
            s += 2.0L * cosl(2.0L * 3.141592653589793238462643383279502884L * (long double)f * (long double)t /
                             (long double)N);
        h[k] = (double)s;
    }
    ST_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_dirw), h, sizeof(h)));
    done[ctx->device] = true;
}

}  // namespace

void st_corr_track_fast_impl(st_ctx *ctx, StPair &pr, const st_track_params &par, st_track_out *out,
                             st_track_diag *diag)
{
    ST_REQUIRE(par.nR > 0 && par.nA > 0 && par.scaleFactor > 0, ST_ERR_ARG, "grid");
    ST_REQUIRE(par.wRa > 0 && par.wAa > 0 && !(par.wRa & 1) && !(par.wAa & 1), ST_ERR_ARG, "window sizes");
    WPar p{};
    p.wRa = par.wRa;
    p.wAa = par.wAa;
    p.W_R = 2 * par.wRa;
    p.W_A = 2 * par.wAa;
    p.epR = par.edgePadR + WNFAST / 4;  // corrTrackFast.c:116-117
    p.epA = par.edgePadA + WNFAST / 4;
    p.wR2 = (p.wRa - 2 * p.epR) * WOS;
    p.wA2 = (p.wAa - 2 * p.epA) * WOS;
    p.nLagR = 2 * p.epR * WOS + 1;
    p.nLagA = 2 * p.epA * WOS + 1;
    p.navgR = par.navgR;
    p.navgA = par.navgA;
    p.dwRa = stfft::make_fastdiv((uint32_t)p.wRa);
    p.dWR = stfft::make_fastdiv((uint32_t)p.W_R);
    p.dWA = stfft::make_fastdiv((uint32_t)p.W_A);
    ST_REQUIRE(par.edgePadR >= 0 && par.edgePadA >= 0 && p.wR2 > 0 && p.wA2 > 0, ST_ERR_ARG,
               "edge pad leaves no chip");
    ST_REQUIRE(stfft::supported_size(p.wRa) && stfft::supported_size(p.wAa) && stfft::supported_size(p.W_R) &&
                   stfft::supported_size(p.W_A),
               ST_ERR_UNSUPPORTED, "window size not supported by the FFT engine");
    ST_REQUIRE(p.navgR >= 0 && p.navgA >= 0 && p.navgR <= kMaxNavg && p.navgA <= kMaxNavg, ST_ERR_UNSUPPORTED,
               "boxcar half-widths above 24");
    ST_REQUIRE(par.polyShift || (par.initDr && par.initDa && par.initNr > 1 && par.initNa > 1), ST_ERR_ARG,
               "initial offsets: polyShift=0 needs initDr/initDa");
    // corrTrackFast.c:333-334
    p.scale = (float)(1. / ((float)p.wRa * (float)p.wAa * (float)p.wRa * (float)p.wAa *
                            (float)(WOS * WOS * WOS * WOS) * (WOS * WOS * (p.navgA + 1) * (p.navgR + 1))));
    p.scaleW = 1.0 / ((WOS * WOS * WOS * WOS) * (double)p.wAa * (double)(p.wAa - 2 * p.epA) * (double)p.wRa *
                      (double)(p.wRa - 2 * p.epR));
    p.im = Images{pr.d1, pr.d2, pr.nr1, pr.na1, pr.nr2, pr.na2, pr.stype};

    hipStream_t st = ctx->stream;
    init_dirichlet_w(ctx);
    ctx->stats_reset();
    hipEvent_t *sev = ctx->ev + 64;
    ST_HIP(hipEventRecord(sev[0], st));
    const int kFftSlot = ctx->kernel_slot("fft2d"), kDet = ctx->kernel_slot("w_detect"),
              kNcc = ctx->kernel_slot("w_ncc"), kRef = ctx->kernel_slot("w_refine"),
              kStat = ctx->kernel_slot("w_stats");

    // geometry on the host: grid walk, initial offsets, mask, corner-only
    // bounds (corrTrackFast.c:160-231)
    const int npts = par.nR * par.nA;
    std::vector<int4> pos;
    std::vector<int> listPt;
    std::vector<int> r2v(npts), a2v(npts);
    std::vector<uint8_t> typ(npts, 0);
    for (int i = 0; i < par.nA; i++) {
        const int a1 = par.aStart + i * par.deltaA - par.wAa / 2;
        for (int j = 0; j < par.nR; j++) {
            const int o = i * par.nR + j;
            const int r1 = par.rStart + j * par.deltaR - par.wRa / 2;
            int r2, a2;
            image2_pos(par, r1, a1, &r2, &a2);
            r2v[o] = r2;
            a2v[o] = a2;
            const int mv = mask_value(par, r1 * par.scaleFactor, a1 * par.scaleFactor);
            if (mv == 1 && r1 > 0 && r1 < pr.nr1 && a1 > 0 && a1 < pr.na1 && r2 > 0 && r2 < pr.nr2 && a2 > 0 &&
                a2 < pr.na2) {
                pos.push_back(make_int4(r1, a1, r2, a2));
                listPt.push_back(o);
            }
        }
    }
    const int nList = (int)pos.size();
    std::vector<WRes> res(nList);

    if (nList > 0) {
        const size_t NN = (size_t)p.W_A * p.W_R;
        // interpolation blocking: R rows per k_w_rowinterp block, C columns
        // per k_w_colinterp block
        const int R = std::max(1, std::min(p.wAa, kRowElems / p.W_R));
        ST_REQUIRE(p.W_R <= kRowElems && p.W_A <= kColElems, ST_ERR_UNSUPPORTED, "window too large for the FFT blocks");
        int lc = 0;
        while ((2 << lc) <= 16 && (2 << lc) * p.W_A <= kColElems)
            lc++;
        const int C = 1 << lc;
        const int nblk = (p.wAa + R - 1) / R;
        const int nRun = (p.nLagA + kNccRows - 1) / kNccRows;
        const size_t per = 2 * NN * 8 + 2 * NN * 8 + (size_t)p.nLagA * p.nLagR * 4 +
                           (size_t)p.nLagA * sizeof(RowBest) + 2 * (size_t)(nblk + 1) * 32 +
                           (size_t)((p.W_A + kTile - 1) / kTile) * 2 * p.W_R * 8 + 4096;
        // a batch the device cannot hold is halved until it fits
        size_t budget = std::max(ctx->scratchBudget, per * 4);
        int B = 1;
        for (;;) {
            B = (int)std::max<size_t>(1, std::min<size_t>({(size_t)nList, budget / per, (size_t)4096}));
            try {
                ctx->arena_reserve((size_t)B * per + (size_t)nList * (16 + sizeof(WRes)) + 65536);
                break;
            } catch (const StError &e) {
                if (e.code != ST_ERR_NOMEM || B == 1)
                    throw;
                budget = (size_t)B * per / 2;
            }
        }
        Arena &ar = ctx->arena;
        int4 *dPos = ar.take<int4>(nList);
        WRes *dRes = ar.take<WRes>(nList);
        ST_HIP(hipMemcpyAsync(dPos, pos.data(), nList * sizeof(int4), hipMemcpyHostToDevice, st));
        WBatch Bt{};
        Bt.pos = dPos;
        Bt.U = ar.take<float2>(2 * (size_t)B * NN);
        float2 *TMP = ar.take<float2>(2 * (size_t)B * NN);
        Bt.Z = TMP;
        Bt.Zs = TMP + (size_t)B * NN;  // Z stays intact for the window sums of k_w_ncc
        Bt.T = TMP;  // [2][B][wAa][W_R] fits: consumed before k_w_detect writes Z
        Bt.dop = ar.take<double>(2 * (size_t)B * nblk * 4);
        Bt.dopE = ar.take<double>(2 * (size_t)B * 4);
        Bt.chip = ar.take<double>(2 * (size_t)B);
        Bt.csum = ar.take<double>((size_t)B * ((p.W_A + kTile - 1) / kTile) * 2 * p.W_R);
        Bt.cr = ar.take<float>((size_t)B * p.nLagA * p.nLagR);
        Bt.rows = ar.take<RowBest>((size_t)B * p.nLagA);
        Bt.az = ar.take<int>(B);
        const int nStep = (int)std::min<size_t>(kNccStep, 65536 / (2 * (size_t)(p.W_R + 1) * sizeof(double)));
        ST_REQUIRE(nStep >= 1 && p.W_R <= 256 * 8, ST_ERR_UNSUPPORTED, "window too large for the NCC kernel");
        const size_t ldsNcc = (size_t)nStep * 2 * (p.W_R + 1) * sizeof(double);
        const stfft::PlanArgs PfR = ctx->fft.plan(p.wRa, -1), PiR = ctx->fft.plan(p.W_R, +1);
        const stfft::PlanArgs PfA = ctx->fft.plan(p.wAa, -1), PiA = ctx->fft.plan(p.W_A, +1);
        const size_t ldsRow = (size_t)R * p.W_R * sizeof(float2), ldsCol = (size_t)C * (p.W_A + 1) * sizeof(float2);
        for (int off = 0; off < nList; off += B) {
            ctx->check_cancel();
            const int nb = std::min(B, nList - off);
            const int np = (nb + 1) / 2;
            Bt.res = dRes + off;
            ctx->time_begin(kFftSlot);
            hipLaunchKernelGGL(k_w_rowinterp, dim3(nblk, 2 * nb), dim3(NT), ldsRow, st, p, off, Bt, nb, R, PfR, PiR);
            hipLaunchKernelGGL(k_w_az, dim3(nb), dim3(64), 0, st, p, Bt, nb, nblk);
            hipLaunchKernelGGL(k_w_colinterp, dim3((p.W_R + C - 1) / C, 2 * nb), dim3(NT), ldsCol, st, p, Bt, nb, lc,
                               PfA, PiA);
            ctx->time_end(kFftSlot);
            ctx->time_begin(kDet);
            hipLaunchKernelGGL(k_w_detect, dim3((p.W_R + kTile - 1) / kTile, (p.W_A + kTile - 1) / kTile, nb),
                               dim3(256), detect_lds_bytes(p), st, p, Bt, nb);
            ctx->time_end(kDet);
            ctx->time_begin(kStat);
            hipLaunchKernelGGL(k_w_chipstats, dim3(nb), dim3(NT), 0, st, p, Bt);
            ctx->time_end(kStat);
            ctx->time_begin(kFftSlot);
            ctx->fft.fft2d_to(Bt.Z, Bt.Zs, nb, p.W_A, p.W_R, -1, st);
            ctx->time_end(kFftSlot);
            hipLaunchKernelGGL(k_w_cross, dim3((unsigned)((NN + 255) / 256), np), dim3(256), 0, st, p, Bt, nb);
            ctx->time_begin(kFftSlot);
            ctx->fft.fft2d(Bt.U, TMP, np, p.W_A, p.W_R, +1, st);
            ctx->time_end(kFftSlot);
            ctx->time_begin(kNcc);
            {
                const dim3 g(nRun, nb);
                if (p.W_R <= 256 * 4)
                    hipLaunchKernelGGL(k_w_ncc<4>, g, dim3(256), ldsNcc, st, p, Bt, nRun, nStep);
                else
                    hipLaunchKernelGGL(k_w_ncc<8>, g, dim3(256), ldsNcc, st, p, Bt, nRun, nStep);
            }
            ctx->time_end(kNcc);
            ctx->time_begin(kRef);
            hipLaunchKernelGGL(k_w_refine, dim3(nb), dim3(256), 0, st, p, Bt);
            ctx->time_end(kRef);
            ST_HIP(hipGetLastError());
        }
        ST_HIP(hipMemcpyAsync(res.data(), dRes, nList * sizeof(WRes), hipMemcpyDeviceToHost, st));
    }
    ST_HIP(hipEventRecord(sev[1], st));
    ST_HIP(hipStreamSynchronize(st));
    ctx->flush_timings();

    // corrMatch (corrTrackFast.c:376-404) and the output grids
    for (int k = 0; k < npts; k++) {
        out->offR[k] = (float)-ST_LARGEINT;
        out->offA[k] = (float)-ST_LARGEINT;
        out->corr[k] = 0.f;
        out->type[k] = 0;
        if (out->azd)
            out->azd[k] = 0.f;
    }
    if (diag) {
        for (int k = 0; k < npts; k++) {
            if (diag->r2)
                diag->r2[k] = r2v[k];
            if (diag->a2)
                diag->a2[k] = a2v[k];
            if (diag->azShift)
                diag->azShift[k] = -1;
            for (int t = 0; t < 3; t++) {
                if (diag->peakI[t])
                    diag->peakI[t][k] = -2000;
                if (diag->peakJ[t])
                    diag->peakJ[t][k] = -2000;
                if (diag->fineI[t])
                    diag->fineI[t][k] = -2000;
                if (diag->fineJ[t])
                    diag->fineJ[t][k] = -2000;
                if (diag->cmax[t])
                    diag->cmax[t][k] = 0;
                if (diag->sigRg[t])
                    diag->sigRg[t][k] = 0;
                if (diag->sigAz[t])
                    diag->sigAz[t][k] = 0;
                if (diag->peakGap[t])
                    diag->peakGap[t][k] = 0;
                if (diag->fineGap[t])
                    diag->fineGap[t][k] = 0;
                if (diag->snr[t])
                    diag->snr[t][k] = 0;
            }
        }
    }
    int nAmp = 0, nFail = 0;
    for (int q = 0; q < nList; q++) {
        const int o = listPt[q];
        const int4 ps = pos[q];
        const WRes &r = res[q];
        float rShift = (float)(ps.z - ps.x), aShift = (float)(ps.w - ps.y);
        if (r.cMax > 0.0f) {
            rShift -= r.rShift1;
            aShift -= r.aShift1;
            nAmp++;
        } else {
            rShift = (float)-ST_LARGEINT;
            aShift = (float)-ST_LARGEINT;
            nFail++;
        }
        out->offR[o] = rShift;
        out->offA[o] = aShift;
        out->corr[o] = r.cMax;
        out->type[o] = kAmpMatch;
        if (diag) {
            if (diag->azShift)
                diag->azShift[o] = r.az;
            if (diag->peakI[0])
                diag->peakI[0][o] = r.pI;
            if (diag->peakJ[0])
                diag->peakJ[0][o] = r.pJ;
            if (diag->fineI[0])
                diag->fineI[0][o] = r.fI;
            if (diag->fineJ[0])
                diag->fineJ[0][o] = r.fJ;
            if (diag->cmax[0])
                diag->cmax[0][o] = r.cMax;
            if (diag->peakGap[0])
                diag->peakGap[0][o] = r.gapCoarse;
            if (diag->fineGap[0])
                diag->fineGap[0][o] = r.gapFine;
        }
    }
    out->nComplex = 0;
    out->nAmp = nAmp;
    out->nAmpL = 0;
    out->nFail = nFail;
    {
        float ms;
        ST_HIP(hipEventElapsedTime(&ms, sev[0], sev[1]));
        ctx->stats.totalMs = ms;
        ctx->stats.stageMs[2] = ms;
        ctx->stats.nPoints = npts;
        ctx->stats.nCorr[1] = nList;
    }
}
