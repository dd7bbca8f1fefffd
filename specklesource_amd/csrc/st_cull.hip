// cullst stages on the GPU (Cullst/cullSTData.c, cullStats.c,
// cullSmooth.c, cullIslands.c).
//
// cullSTData culls in place in raster order, so a cell's decision sees the
// decisions already taken for earlier cells of the same pass.  The GPU
// solves that recurrence exactly as a fixpoint: every iteration re-decides
// all cells in parallel, earlier cells taking their decision from the
// previous iteration; the sequential result is the unique fixpoint and is
// reached when an iteration changes nothing (a dependency chain of length
// k is settled after k iterations).  A single-workgroup wavefront kernel
// (cell (i, j) runs at step i*(bR/2+1) + j) finishes the pass if the
// fixpoint has not settled after kMaxIter iterations.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "st_ctx.h"

// The cull arithmetic mirrors the reference operation by operation; no FMA
// contraction anywhere in this file.
#pragma clang fp contract(off)

namespace {

constexpr float kNoData = -2.0e9f;
constexpr int kMaxIter = 48;
// The median kernels keep each wave's box values (range and azimuth) in
// dynamic LDS sized for the box of the call, at most kCullLds bytes per
// workgroup: box sides up to 45 (2025 cells; 4 waves x 2 x 4 B per cell).
constexpr int kMaxBox = 45;
constexpr size_t kCullLds = 65536;
constexpr int cull_box_cells(int bR, int bA) { return ((bR / 2) * 2 + 1) * ((bA / 2) * 2 + 1); }

__device__ inline bool valid2(float r, float a)
{
    return r > (float)(1 - ST_LARGEINT) && a > (float)(1 - ST_LARGEINT);
}

struct CullDev {
    int nR, nA;
    float *offR, *offA, *corr;
    uint8_t *type;
    const uint8_t *mask;
    float *offRS, *offAS, *sigmaR, *sigmaA;
    float *R0, *A0;       // state at the start of a cullSTData pass
    uint8_t *cPrev, *cNew; // cull decisions of the pass
    int *counters;        // [0] changed, [1] nCulled, [2] max root, ...
    int *parent, *root, *bbox;   // islands
};

// Pre-pass of cullSTData (cullSTData.c:34-68).
__global__ void k_cull_prepass(CullDev D, float corrThresh, int singleMT)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.nR * D.nA)
        return;
    if (D.corr[c] < corrThresh) {
        D.offR[c] = kNoData;
        D.offA[c] = kNoData;
        D.type[c] = 0;
    }
    if (D.mask) {
        const unsigned char mv = (unsigned char)D.mask[c] & 4;
        if (mv > 0 && D.type[c] > 2) {
            D.offR[c] = kNoData;
            D.offA[c] = kNoData;
            D.type[c] = 0;
        }
        if (singleMT > 0 && D.type[c] != singleMT) {
            D.offR[c] = kNoData;
            D.offA[c] = kNoData;
            D.type[c] = 0;
        }
    }
    D.R0[c] = D.offR[c];
    D.A0[c] = D.offA[c];
    D.cPrev[c] = 0;
}

// Start of a cullLSData pass (cullLSData.c:37-58): the pass state, and the
// counts of good cells ([3]) and cells with data ([4]).
__global__ void k_cull_ls_begin(CullDev D, const uint8_t *type, uint8_t noDataToUse)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = c < D.nR * D.nA;
    bool good = false, data = false;
    if (in) {
        D.R0[c] = D.offR[c];
        D.A0[c] = D.offA[c];
        D.cPrev[c] = 0;
        good = valid2(D.offR[c], D.offA[c]);
        data = (type[c] & noDataToUse) == 0;
    }
    const unsigned long long g = __ballot(good), d = __ballot(data);
    if ((threadIdx.x & 63) == 0) {
        if (g)
            atomicAdd(&D.counters[3], __popcll(g));
        if (d)
            atomicAdd(&D.counters[4], __popcll(d));
    }
}

// Landsat variant of the median test (cullLSData.c:66-121): the box's
// good points all carrying the slow-velocity bit switch the limits to
// 40 m/yr over deltaT in pixels (at least 0.2), and a cell whose Y is not
// above NODATA + 1 is culled.
struct LsArg {
    const uint8_t *type;
    uint8_t slow;
    double noData1;  // NODATA + 1
    double deltaT, dx, dy;
};

// Median test of one cell by one wavefront (cullSTData.c:73-122): the box
// values are ranked by counting, the floor(n/2)-th smallest is the median
// (selectForCull's k = n/2 on a 1-based list).  LS selects the Landsat
// limits (cullLSData.c); maxA / maxR are then maxY / maxX.
template <bool LS>
__device__ inline bool cull_decide(const CullDev &D, int c, int bR, int bA, int nGood, float maxA, float maxR,
                                   const uint8_t *earlier, float *lr, float *la, const LsArg &L)
{
    const int lane = threadIdx.x & 63;
    const int i = c / D.nR, j = c - (c / D.nR) * D.nR;
    const int i1 = imax(0, i - bA / 2), i2 = imin(D.nA - 1, i + bA / 2);
    const int j1 = imax(0, j - bR / 2), j2 = imin(D.nR - 1, j + bR / 2);
    const int bw = j2 - j1 + 1, cnt = bw * (i2 - i1 + 1);
    int ngood = 0;
    bool anyFast = false;
    for (int base = 0; base < cnt; base += 64) {
        const int s = base + lane;
        bool ok = false, fast = false;
        float r = 0, a = 0;
        if (s < cnt) {
            const int q = (i1 + s / bw) * D.nR + j1 + s % bw;
            const bool culled = q < c && earlier[q];
            r = culled ? kNoData : D.R0[q];
            a = culled ? kNoData : D.A0[q];
            ok = valid2(r, a);
            if constexpr (LS)
                fast = ok && !(L.type[q] & L.slow);
        }
        if constexpr (LS)
            anyFast = anyFast || __ballot(fast) != 0ull;
        const unsigned long long m = __ballot(ok);
        const int pos = ngood + __popcll(m & ((1ull << lane) - 1ull));
        if (ok) {
            lr[pos] = r;
            la[pos] = a;
        }
        ngood += __popcll(m);
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (ngood <= nGood)
        return true;
    if constexpr (LS) {
        if (!((double)D.A0[c] > L.noData1))
            return true;
    }
    int k = ngood / 2;
    if (k == 0)
        k = 1;
    float medR = 0, medA = 0;
    for (int base = 0; base < ngood; base += 64) {
        const int s = base + lane;
        bool isR = false, isA = false;
        float vr = 0, va = 0;
        if (s < ngood) {
            vr = lr[s];
            va = la[s];
            int lessR = 0, eqR = 0, lessA = 0, eqA = 0;
            for (int u = 0; u < ngood; u++) {
                const float xr = lr[u], xa = la[u];
                lessR += xr < vr;
                eqR += xr == vr;
                lessA += xa < va;
                eqA += xa == va;
            }
            isR = lessR < k && k <= lessR + eqR;
            isA = lessA < k && k <= lessA + eqA;
        }
        const unsigned long long mR = __ballot(isR), mA = __ballot(isA);
        if (mR)
            medR = __shfl(vr, __ffsll((long long)mR) - 1);
        if (mA)
            medA = __shfl(va, __ffsll((long long)mA) - 1);
    }
    const double diffA = (double)(D.A0[c] - medA);
    const double diffR = (double)(D.R0[c] - medR);
    if constexpr (LS) {
        double mX = maxR, mY = maxA;
        if (!anyFast) {
            mX = 40.0 / 365. * L.deltaT;
            mY = mX;
            mX /= L.dx;
            mY /= L.dy;
            mX = mX > 0.2 ? mX : 0.2;
            mY = mY > 0.2 ? mY : 0.2;
        }
        return fabs(diffA) > mY || fabs(diffR) > mX;
    }
    return fabs(diffA) > maxA || fabs(diffR) > maxR;
}

template <bool LS>
__global__ void __launch_bounds__(256) k_cull_iter(CullDev D, int bR, int bA, int nGood, float maxA, float maxR,
                                                   LsArg L)
{
    extern __shared__ float cull_lds[];  // per wave: range values, then azimuth values
    const int w = threadIdx.x >> 6, cells = cull_box_cells(bR, bA);
    const int c = blockIdx.x * 4 + w;
    if (c >= D.nR * D.nA)
        return;
    float *lr = cull_lds + (size_t)w * 2 * cells, *la = lr + cells;
    const bool cull = cull_decide<LS>(D, c, bR, bA, nGood, maxA, maxR, D.cPrev, lr, la, L);
    if ((threadIdx.x & 63) == 0) {
        D.cNew[c] = cull ? 1 : 0;
        if ((D.cPrev[c] != 0) != cull)
            atomicAdd(&D.counters[0], 1);
    }
}

// Exact sequential order for the leftover case: one workgroup (up to 16
// waves, fewer for large boxes) sweeps the wavefront steps, one wave per
// cell of a step.
template <bool LS>
__global__ void __launch_bounds__(1024) k_cull_wavefront(CullDev D, int bR, int bA, int nGood, float maxA, float maxR,
                                                         LsArg L)
{
    extern __shared__ float cull_lds[];
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6, cells = cull_box_cells(bR, bA);
    float *lr = cull_lds + (size_t)w * 2 * cells, *la = lr + cells;
    const int s = bR / 2 + 1;
    const int steps = (D.nA - 1) * s + D.nR;
    for (int T = 0; T < steps; T++) {
        // cells with i*s + j == T
        const int iLo = imax(0, (T - (D.nR - 1) + s - 1) / s), iHi = imin(D.nA - 1, T / s);
        for (int i = iLo + w; i <= iHi; i += nw) {
            const int j = T - i * s;
            const int c = i * D.nR + j;
            const bool cull = cull_decide<LS>(D, c, bR, bA, nGood, maxA, maxR, D.cNew, lr, la, L);
            if ((threadIdx.x & 63) == 0)
                D.cNew[c] = cull ? 1 : 0;
        }
        __syncthreads();
    }
}

__global__ void k_cull_apply(CullDev D)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.nR * D.nA)
        return;
    if (D.cNew[c]) {
        D.offR[c] = kNoData;
        D.offA[c] = kNoData;
        atomicAdd(&D.counters[1], 1);
    } else {
        D.offR[c] = D.R0[c];
        D.offA[c] = D.A0[c];
    }
}

// cullStats (cullStats.c:19-144) with the plane fit done by a one-sided
// Jacobi SVD whose rotated columns are recomputed from the box rows (the
// oracle follows the same arithmetic).  Sweeps stop at column-pair cosines
// below 1e-13, just above the rounding floor of the 81-row sums.
__device__ void lsq3(double c0, const CullDev &D, int i1, int i2, int j1, int j2, bool useR, double x[3])
{
    double V[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    for (int sweep = 0; sweep < 30; sweep++) {
        double off = 0.0;
        for (int p = 0; p < 2; p++)
            for (int q = p + 1; q < 3; q++) {
                double alpha = 0, beta = 0, gamma = 0;
                for (int ii = i1; ii <= i2; ii++)
                    for (int jj = j1; jj <= j2; jj++) {
                        const int o = ii * D.nR + jj;
                        if (!valid2(D.offR[o], D.offA[o]))
                            continue;
                        const double a0 = c0, a1 = (double)ii, a2 = (double)jj;
                        const double up = a0 * V[0 * 3 + p] + a1 * V[1 * 3 + p] + a2 * V[2 * 3 + p];
                        const double uq = a0 * V[0 * 3 + q] + a1 * V[1 * 3 + q] + a2 * V[2 * 3 + q];
                        alpha += up * up;
                        beta += uq * uq;
                        gamma += up * uq;
                    }
                if (gamma == 0.0)
                    continue;
                const double conv = fabs(gamma) / sqrt(alpha * beta);
                if (conv > off)
                    off = conv;
                if (conv < 1e-13)
                    continue;
                const double zeta = (beta - alpha) / (2.0 * gamma);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
                for (int r = 0; r < 3; r++) {
                    const double vp = V[r * 3 + p], vq = V[r * 3 + q];
                    V[r * 3 + p] = c * vp - s * vq;
                    V[r * 3 + q] = s * vp + c * vq;
                }
            }
        if (off < 1e-13)
            break;
    }
    double w2[3] = {0, 0, 0}, uy[3] = {0, 0, 0};
    for (int ii = i1; ii <= i2; ii++)
        for (int jj = j1; jj <= j2; jj++) {
            const int o = ii * D.nR + jj;
            if (!valid2(D.offR[o], D.offA[o]))
                continue;
            const double yv = useR ? (double)D.offR[o] : (double)D.offA[o];
            const double a0 = c0, a1 = (double)ii, a2 = (double)jj;
            for (int k = 0; k < 3; k++) {
                const double u = a0 * V[0 * 3 + k] + a1 * V[1 * 3 + k] + a2 * V[2 * 3 + k];
                w2[k] += u * u;
                uy[k] += u * yv;
            }
        }
    double w[3], wmax = 0;
    for (int k = 0; k < 3; k++) {
        w[k] = sqrt(w2[k]);
        if (w[k] > wmax)
            wmax = w[k];
    }
    const double thresh = 1.0e-5 * wmax;
    double coef[3];
    for (int k = 0; k < 3; k++)
        coef[k] = (w[k] < thresh || w[k] == 0.0) ? 0.0 : uy[k] / w2[k];
    for (int r = 0; r < 3; r++)
        x[r] = V[r * 3 + 0] * coef[0] + V[r * 3 + 1] * coef[1] + V[r * 3 + 2] * coef[2];
}

// addQuant: cullLSStats adds the quantisation variance `quant` before the
// square root (cullLSStats.c:144-146).
__global__ void k_cull_stats(CullDev D, int addQuant, double quant)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.nR * D.nA)
        return;
    const int b = 9;
    const int i = c / D.nR, j = c - (c / D.nR) * D.nR;
    const int i1 = imax(0, i - b / 2), i2 = imin(D.nA - 1, i + b / 2);
    const int j1 = imax(0, j - b / 2), j2 = imin(D.nR - 1, j + b / 2);
    int ngood = 0;
    double meanA = 0.0, meanR = 0.0;
    for (int ii = i1; ii <= i2; ii++)
        for (int jj = j1; jj <= j2; jj++) {
            const int o = ii * D.nR + jj;
            if (valid2(D.offR[o], D.offA[o])) {
                ngood++;
                meanA += D.offA[o];
                meanR += D.offR[o];
            }
        }
    meanA /= (double)ngood;
    meanR /= (double)ngood;
    if (!(ngood > 8 && D.offR[c] > (float)(1 - ST_LARGEINT))) {
        D.sigmaA[c] = kNoData;
        D.sigmaR[c] = kNoData;
        return;
    }
    double aR[3], aA[3];
    lsq3((double)(float)meanR, D, i1, i2, j1, j2, true, aR);
    lsq3((double)(float)meanA, D, i1, i2, j1, j2, false, aA);
    aR[0] *= meanR;
    aA[0] *= meanA;
    int ng = 0;
    double mA = 0.0, mR = 0.0, varR = 0.0, varA = 0.0;
    for (int ii = i1; ii <= i2; ii++)
        for (int jj = j1; jj <= j2; jj++) {
            const int o = ii * D.nR + jj;
            if (valid2(D.offR[o], D.offA[o])) {
                ng++;
                float oR = D.offR[o], oA = D.offA[o];
                oR = (float)(oR - (aR[0] + aR[1] * (float)ii + aR[2] * (float)jj));
                oA = (float)(oA - (aA[0] + aA[1] * (float)ii + aA[2] * (float)jj));
                mA += oA;
                mR += oR;
                varA += oA * oA;
                varR += oR * oR;
            }
        }
    float sA = (float)(varA / (float)(ng - 3));
    float sR = (float)(varR / (float)(ng - 3));
    sA = (float)(sA - mA * mA / (float)((ng - 3) * ng));
    sR = (float)(sR - mR * mR / (float)((ng - 3) * ng));
    sA = (float)(0.00001 > sA ? 0.00001 : sA);
    sR = (float)(0.00001 > sR ? 0.00001 : sR);
    if (addQuant) {
        sR = (float)(sR + quant);
        sA = (float)(sA + quant);
    }
    D.sigmaA[c] = (float)sqrt((double)sA);
    D.sigmaR[c] = (float)sqrt((double)sR);
}

// cullSmooth (cullSmooth.c:29-103)
__device__ inline float s_weight(int m, int s)
{
    const float w2 = (float)((double)s / 2.0 - 0.01);
    return (m < -w2 || m > w2) ? 0.5f : 1.0f;
}

__global__ void k_cull_smooth(CullDev D, int sR, int sA)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.nR * D.nA)
        return;
    const float nlooks[4] = {0.0f, 1.0f, 0.25f, 0.015f};
    const int i = c / D.nR, j = c - (c / D.nR) * D.nR;
    const int i1 = imax(0, i - sA / 2), i2 = imin(D.nA - 1, i + sA / 2);
    const int j1 = imax(0, j - sR / 2), j2 = imin(D.nR - 1, j + sR / 2);
    double ngood = 0, meanA = 0.0, meanR = 0.0;
    float nLooksEff = 0.0f;
    for (int ii = i1; ii <= i2; ii++)
        for (int jj = j1; jj <= j2; jj++) {
            const int o = ii * D.nR + jj;
            if (valid2(D.offR[o], D.offA[o])) {
                const double w = (double)(s_weight(ii - i, sA) * s_weight(jj - j, sR));
                ngood += w;
                nLooksEff = (float)(nLooksEff + w * nlooks[(int)D.type[o] & 3]);
                meanA += w * D.offA[o];
                meanR += w * D.offR[o];
            }
        }
    if (((ngood > 1) || (sA == 1) || (sR == 1)) && D.offR[c] > (float)(1 - ST_LARGEINT) &&
        D.sigmaA[c] > (float)(1 - ST_LARGEINT)) {
        nLooksEff = (float)(nLooksEff > 1. ? nLooksEff : 1.);
        D.sigmaA[c] = (float)(D.sigmaA[c] / sqrt((double)nLooksEff));
        D.sigmaR[c] = (float)(D.sigmaR[c] / sqrt((double)nLooksEff));
        D.offRS[c] = (float)(meanR / (float)ngood);
        D.offAS[c] = (float)(meanA / (float)ngood);
    } else {
        D.offRS[c] = kNoData;
        D.offAS[c] = kNoData;
    }
}

// cullLSSmooth (cullLSSmooth.c:11-73): unweighted box mean, effective
// looks = the count of good cells.
__global__ void k_cull_ls_smooth(CullDev D, int sX, int sY)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.nR * D.nA)
        return;
    const int i = c / D.nR, j = c - (c / D.nR) * D.nR;
    const int i1 = imax(0, i - sY / 2), i2 = imin(D.nA - 1, i + sY / 2);
    const int j1 = imax(0, j - sX / 2), j2 = imin(D.nR - 1, j + sX / 2);
    unsigned ngood = 0;
    float nLooksEff = 0.0f;
    double meanY = 0.0, meanX = 0.0;
    for (int ii = i1; ii <= i2; ii++)
        for (int jj = j1; jj <= j2; jj++) {
            const int o = ii * D.nR + jj;
            if (valid2(D.offR[o], D.offA[o])) {
                ngood++;
                nLooksEff++;
                meanY += D.offA[o];
                meanX += D.offR[o];
            }
        }
    if (((ngood > 1) || (sY == 1) || (sX == 1)) && D.offR[c] > (float)(1 - ST_LARGEINT) &&
        D.sigmaA[c] > (float)(1 - ST_LARGEINT)) {
        nLooksEff = (float)(nLooksEff > 1. ? nLooksEff : 1.);
        D.sigmaA[c] = (float)(D.sigmaA[c] / sqrt((double)nLooksEff));
        D.sigmaR[c] = (float)(D.sigmaR[c] / sqrt((double)nLooksEff));
        D.offRS[c] = (float)(meanX / (float)ngood);
        D.offAS[c] = (float)(meanY / (float)ngood);
    } else {
        D.offRS[c] = kNoData;
        D.offAS[c] = kNoData;
    }
}

// cullIslands (cullIslands.c:13-198): 4-connected components of valid
// smoothed cells by atomic union-find.  Roots are the smallest member
// index, i.e. the component's first cell in raster order, which is the
// order a raster-scan labeller numbers components in; the component with
// the largest root therefore carries the last label, which the reference
// leaves alone (the loops stop at maxLabel - 1).
__device__ inline bool isl_valid(const CullDev &D, int c) { return D.offRS[c] >= (float)(-ST_LARGEINT + 1000); }

// Parent links only ever point to a smaller index, so every value a thread
// reads for par[x] is an ancestor of x, and path halving (par[x] = par[par[x]])
// may race with other finds and with the linking CAS without breaking a tree:
// it writes only to nodes that are no longer roots.  Loads and stores are
// agent-scope atomics so a retried find sees links made by other waves.
__device__ inline int uf_ld(const int *par, int x) { return __hip_atomic_load(par + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

__device__ inline int uf_find(int *par, int x)
{
    while (true) {
        const int p = uf_ld(par, x);
        if (p == x)
            return x;
        const int g = uf_ld(par, p);
        if (g == p)
            return p;
        __hip_atomic_store(par + x, g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = g;
    }
}

__device__ inline void uf_unite(int *par, int a, int b)
{
    while (true) {
        a = uf_find(par, a);
        b = uf_find(par, b);
        if (a == b)
            return;
        if (a < b) {
            const int t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(&par[a], a, b) == a)
            return;
    }
}

// Each valid cell starts linked to its left neighbour when that is valid,
// else to the one above, so a row run is one chain before any merge.
__global__ void k_isl_init(CullDev D)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.nR * D.nA)
        return;
    int p = c;
    if (isl_valid(D, c)) {
        const int j = c % D.nR;
        if (j > 0 && isl_valid(D, c - 1))
            p = c - 1;
        else if (c >= D.nR && isl_valid(D, c - D.nR))
            p = c - D.nR;
    }
    D.parent[c] = p;
}

// The only edge the initial links leave out: up, when left was taken.
__global__ void k_isl_merge(CullDev D)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.nR * D.nA || c < D.nR || !isl_valid(D, c))
        return;
    const int j = c % D.nR;
    if (j > 0 && isl_valid(D, c - 1) && isl_valid(D, c - D.nR))
        uf_unite(D.parent, c, c - D.nR);
}

// Roots and bounding boxes.  Most cells of a wave share a root (one large
// component covers most of a grid), so the min/max are first reduced over
// the lanes with the same root and one lane per root issues the atomics.
__global__ void k_isl_bbox(CullDev D)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const bool v = c < D.nR * D.nA && isl_valid(D, c);
    int r = -1, i = 0, j = 0;
    if (v) {
        r = uf_find(D.parent, c);
        D.root[c] = r;
        i = c / D.nR;
        j = c - i * D.nR;
    }
    int rmax = r;
    for (int o = 32; o > 0; o >>= 1)
        rmax = max(rmax, __shfl_xor(rmax, o));
    const int lane = threadIdx.x & 63;
    if (lane == 0 && rmax >= 0)
        atomicMax(&D.counters[2], rmax);
    unsigned long long pending = __ballot(v);
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int rl = __shfl(r, leader);
        const bool mine = v && r == rl;
        int jmin = mine ? j : 0x7fffffff, jmax = mine ? j : -0x7fffffff;
        int imin = mine ? i : 0x7fffffff, imax = mine ? i : -0x7fffffff;
        for (int o = 32; o > 0; o >>= 1) {
            jmin = min(jmin, __shfl_xor(jmin, o));
            jmax = max(jmax, __shfl_xor(jmax, o));
            imin = min(imin, __shfl_xor(imin, o));
            imax = max(imax, __shfl_xor(imax, o));
        }
        if (lane == leader) {
            atomicMin(&D.bbox[4 * rl + 0], jmin);
            atomicMax(&D.bbox[4 * rl + 1], jmax);
            atomicMin(&D.bbox[4 * rl + 2], imin);
            atomicMax(&D.bbox[4 * rl + 3], imax);
        }
        pending &= ~__ballot(mine);
    }
}

__global__ void k_isl_bbox_init(CullDev D)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.nR * D.nA)
        return;
    D.bbox[4 * c + 0] = 0x7fffffff;
    D.bbox[4 * c + 1] = -0x7fffffff;
    D.bbox[4 * c + 2] = 0x7fffffff;
    D.bbox[4 * c + 3] = -0x7fffffff;
}

__global__ void k_isl_remove(CullDev D, int thresh)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= D.nR * D.nA || !isl_valid(D, c))
        return;
    const int r = D.root[c];
    if (r == D.counters[2])
        return;
    const int *bb = D.bbox + 4 * r;
    if ((bb[1] - bb[0]) <= thresh && (bb[3] - bb[2]) <= thresh) {
        D.offRS[c] = kNoData;
        if (D.offAS)
            D.offAS[c] = kNoData;
    }
}

inline unsigned nb(int n) { return (unsigned)((n + 255) / 256); }

// Fixpoint iterations before the wavefront kernel takes over;
// SPECKLE_AMD_CULL_ITERS lowers it (0 runs every pass on the wavefront
// kernel), which the tests use to cover both schedules.
int max_iterations()
{
    const char *e = getenv("SPECKLE_AMD_CULL_ITERS");
    if (!e || !*e)
        return kMaxIter;
    const int v = atoi(e);
    return v < 0 ? 0 : (v > kMaxIter ? kMaxIter : v);
}

struct CullRun {
    st_ctx *ctx;
    st_cull_params *cp;
    CullDev D{};
    int n;
    Arena *ar = nullptr;  // where the grids are carved from
    Arena fixedBuf;       // a caller-given buffer (the no-allocation entry point)

    CullRun(st_ctx *c, st_cull_params *p) : ctx(c), cp(p)
    {
        ST_REQUIRE(cp && cp->nR > 0 && cp->nA > 0 && cp->offR && cp->offA, ST_ERR_ARG, "cull grids");
        n = cp->nR * cp->nA;
        ctx->arena_reserve(cull_device_bytes((size_t)n) + (size_t)4 * 4 * n + 4 * 256);
        ar = &ctx->arena;
        alloc(cp->nR, cp->nA);
        take_mask();
    }
    // Every grid in `bytes` of device memory at `base`: allocates nothing.
    CullRun(st_ctx *c, st_cull_params *p, void *base, size_t bytes) : ctx(c), cp(p)
    {
        ST_REQUIRE(cp && cp->nR > 0 && cp->nA > 0 && cp->offR && cp->offA, ST_ERR_ARG, "cull grids");
        n = cp->nR * cp->nA;
        fixedBuf.base = (char *)base;
        fixedBuf.cap = bytes;
        ar = &fixedBuf;
        alloc(cp->nR, cp->nA);
        take_mask();
    }
    void take_mask()
    {
        if (cp->mask) {
            uint8_t *m = ar->take<uint8_t>(n);
            put(m, cp->mask, n);
            D.mask = m;
        }
    }
    // Device copies of every grid of the stages (n x 72 bytes).
    void alloc(int nR, int nA)
    {
        D.offR = ar->take<float>(n);
        D.offA = ar->take<float>(n);
        D.corr = ar->take<float>(n);
        D.type = ar->take<uint8_t>(n);
        carve(nR, nA);
    }
    // The stage grids other than offR / offA / corr / type, after what is
    // in use.
    void carve(int nR, int nA)
    {
        Arena &a = *ar;
        D.nR = nR;
        D.nA = nA;
        D.mask = nullptr;
        D.offRS = a.take<float>(n);
        D.offAS = a.take<float>(n);
        D.sigmaR = a.take<float>(n);
        D.sigmaA = a.take<float>(n);
        D.R0 = a.take<float>(n);
        D.A0 = a.take<float>(n);
        D.cPrev = a.take<uint8_t>(n);
        D.cNew = a.take<uint8_t>(n);
        D.counters = a.take<int>(16);
        D.parent = a.take<int>(n);
        D.root = a.take<int>(n);
        D.bbox = a.take<int>(4 * (size_t)n);
    }
    CullRun(st_ctx *c, int nR, int nA) : ctx(c), cp(nullptr)
    {
        n = nR * nA;
        ctx->arena_reserve(cull_device_bytes((size_t)n) + (size_t)4 * 4 * n + 4 * 256);
        ar = &ctx->arena;
        alloc(nR, nA);
    }
    // On grids already on the device (tracking outputs): no arena reset.
    CullRun(st_ctx *c, st_cull_params *p, float *dOffR, float *dOffA, float *dCorr, uint8_t *dType)
        : ctx(c), cp(p)
    {
        ST_REQUIRE(cp && cp->nR > 0 && cp->nA > 0, ST_ERR_ARG, "cull grids");
        n = cp->nR * cp->nA;
        ar = &ctx->arena;
        D.offR = dOffR;
        D.offA = dOffA;
        D.corr = dCorr;
        D.type = dType;
        carve(cp->nR, cp->nA);
        take_mask();
    }
    template <class T>
    void put(T *d, const T *h, size_t cnt)
    {
        if (h)
            ST_HIP(hipMemcpyAsync(d, h, cnt * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    }
    template <class T>
    void get(T *h, const T *d, size_t cnt)
    {
        if (h)
            ST_HIP(hipMemcpyAsync(h, d, cnt * sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    }
    int counter(int k)
    {
        int v = 0;
        ST_HIP(hipMemcpyAsync(&v, D.counters + k, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        ST_HIP(hipStreamSynchronize(ctx->stream));
        return v;
    }

    void st_data()
    {
        hipLaunchKernelGGL(k_cull_prepass, dim3(nb(n)), dim3(256), 0, ctx->stream, D, cp->corrThresh, cp->singleMT);
        cp->nCulled = median_pass<false>(cp->bR, cp->bA, cp->nGood, cp->maxA, cp->maxR, LsArg{});
    }
    // The in-place raster-order median pass on the state left in R0 / A0:
    // fixpoint iterations, the wavefront kernel if they have not settled,
    // then the culls are applied.  Returns the number of culled cells.
    template <bool LS>
    int median_pass(int bR, int bA, int nGood, float maxA, float maxR, const LsArg &L)
    {
        ST_REQUIRE(bR >= 1 && bA >= 1 && (bR / 2) * 2 + 1 <= kMaxBox && (bA / 2) * 2 + 1 <= kMaxBox,
                   ST_ERR_UNSUPPORTED, "box sizes above 45");
        hipStream_t st = ctx->stream;
        const size_t perWave = 2 * sizeof(float) * (size_t)cull_box_cells(bR, bA);
        const int wfWaves = (int)std::min<size_t>(16, kCullLds / perWave);
        bool settled = false;
        const int maxIter = max_iterations();
        for (int it = 0; it < maxIter; it++) {
            ST_HIP(hipMemsetAsync(D.counters, 0, sizeof(int), st));
            hipLaunchKernelGGL(k_cull_iter<LS>, dim3((n + 3) / 4), dim3(256), 4 * perWave, st, D, bR, bA, nGood, maxA,
                               maxR, L);
            ST_HIP(hipGetLastError());
            std::swap(D.cPrev, D.cNew);
            if (counter(0) == 0) {
                settled = true;
                break;
            }
        }
        std::swap(D.cPrev, D.cNew);  // cNew holds the last decisions
        if (!settled)
            hipLaunchKernelGGL(k_cull_wavefront<LS>, dim3(1), dim3(64 * wfWaves), wfWaves * perWave, st, D, bR, bA,
                               nGood, maxA, maxR, L);
        ST_HIP(hipMemsetAsync(D.counters + 1, 0, sizeof(int), st));
        hipLaunchKernelGGL(k_cull_apply, dim3(nb(n)), dim3(256), 0, st, D);
        ST_HIP(hipGetLastError());
        return counter(1);
    }
    void stats()
    {
        cp->bR = cp->bA = 9;  // cullStats.c:38-39
        hipLaunchKernelGGL(k_cull_stats, dim3(nb(n)), dim3(256), 0, ctx->stream, D, 0, 0.0);
        ST_HIP(hipGetLastError());
    }
    void smooth()
    {
        ST_REQUIRE(cp->sR >= 1 && cp->sA >= 1, ST_ERR_ARG, "smoothing window");
        hipLaunchKernelGGL(k_cull_smooth, dim3(nb(n)), dim3(256), 0, ctx->stream, D, cp->sR, cp->sA);
        ST_HIP(hipGetLastError());
    }
    void islands() { islands_t(cp->islandThresh); }
    void islands_t(int thresh)
    {
        hipStream_t st = ctx->stream;
        ST_HIP(hipMemsetAsync(D.counters + 2, 0xff, sizeof(int), st));  // max root = -1
        hipLaunchKernelGGL(k_isl_init, dim3(nb(n)), dim3(256), 0, st, D);
        hipLaunchKernelGGL(k_isl_bbox_init, dim3(nb(n)), dim3(256), 0, st, D);
        hipLaunchKernelGGL(k_isl_merge, dim3(nb(n)), dim3(256), 0, st, D);
        hipLaunchKernelGGL(k_isl_bbox, dim3(nb(n)), dim3(256), 0, st, D);
        hipLaunchKernelGGL(k_isl_remove, dim3(nb(n)), dim3(256), 0, st, D, thresh);
        ST_HIP(hipGetLastError());
    }
};

}  // namespace

// Arena bytes the stage grids take beyond offR / offA / corr / type,
// including the optional mask and the 256-byte alignment of each grid.
size_t cull_device_bytes(size_t n) { return n * 56 + n + 16 * 256 + 65536; }

void st_cull_chain_device(st_ctx *c, st_cull_params *cp, float *dOffR, float *dOffA, float *dCorr, uint8_t *dType)
{
    CullRun R(c, cp, dOffR, dOffA, dCorr, dType);
    ST_REQUIRE(cp->sigmaR && cp->sigmaA && cp->offRS && cp->offAS, ST_ERR_ARG, "cull arrays");
    R.st_data();
    R.st_data();
    R.st_data();
    R.stats();
    R.smooth();
    if (cp->islandThresh > 0)
        R.islands();
    R.get(cp->sigmaR, R.D.sigmaR, R.n);
    R.get(cp->sigmaA, R.D.sigmaA, R.n);
    R.get(cp->offRS, R.D.offRS, R.n);
    R.get(cp->offAS, R.D.offAS, R.n);
}

extern "C" int st_cull_st_data(st_ctx *c, st_cull_params *cp)
{
    return st_guard(c, [&]() {
        CullRun R(c, cp);
        ST_REQUIRE(cp->corr && cp->type, ST_ERR_ARG, "corr/type");
        R.put(R.D.offR, cp->offR, R.n);
        R.put(R.D.offA, cp->offA, R.n);
        R.put(R.D.corr, cp->corr, R.n);
        R.put(R.D.type, cp->type, R.n);
        R.st_data();
        R.get(cp->offR, R.D.offR, R.n);
        R.get(cp->offA, R.D.offA, R.n);
        R.get(cp->type, R.D.type, R.n);
        ST_HIP(hipStreamSynchronize(c->stream));
    });
}

extern "C" int st_cull_stats(st_ctx *c, st_cull_params *cp)
{
    return st_guard(c, [&]() {
        CullRun R(c, cp);
        ST_REQUIRE(cp->sigmaR && cp->sigmaA, ST_ERR_ARG, "sigma outputs");
        R.put(R.D.offR, cp->offR, R.n);
        R.put(R.D.offA, cp->offA, R.n);
        R.stats();
        R.get(cp->sigmaR, R.D.sigmaR, R.n);
        R.get(cp->sigmaA, R.D.sigmaA, R.n);
        ST_HIP(hipStreamSynchronize(c->stream));
    });
}

extern "C" int st_cull_smooth(st_ctx *c, st_cull_params *cp)
{
    return st_guard(c, [&]() {
        CullRun R(c, cp);
        ST_REQUIRE(cp->sigmaR && cp->sigmaA && cp->offRS && cp->offAS && cp->type, ST_ERR_ARG, "smooth arrays");
        R.put(R.D.offR, cp->offR, R.n);
        R.put(R.D.offA, cp->offA, R.n);
        R.put(R.D.type, cp->type, R.n);
        R.put(R.D.sigmaR, cp->sigmaR, R.n);
        R.put(R.D.sigmaA, cp->sigmaA, R.n);
        R.smooth();
        R.get(cp->sigmaR, R.D.sigmaR, R.n);
        R.get(cp->sigmaA, R.D.sigmaA, R.n);
        R.get(cp->offRS, R.D.offRS, R.n);
        R.get(cp->offAS, R.D.offAS, R.n);
        ST_HIP(hipStreamSynchronize(c->stream));
    });
}

extern "C" int st_cull_islands(st_ctx *c, st_cull_params *cp)
{
    return st_guard(c, [&]() {
        CullRun R(c, cp);
        ST_REQUIRE(cp->offRS, ST_ERR_ARG, "offRS");
        R.put(R.D.offRS, cp->offRS, R.n);
        if (cp->offAS)
            R.put(R.D.offAS, cp->offAS, R.n);
        else
            R.D.offAS = nullptr;
        R.islands();
        R.get(cp->offRS, R.D.offRS, R.n);
        if (cp->offAS)
            R.get(cp->offAS, R.D.offAS, R.n);
        ST_HIP(hipStreamSynchronize(c->stream));
    });
}

namespace {

// cullst's order on host grids: upload, cullSTData x3, cullStats,
// cullSmooth, cullIslands, download.
void cull_chain_host(st_ctx *c, st_cull_params *cp, CullRun &R)
{
    R.put(R.D.offR, cp->offR, R.n);
    R.put(R.D.offA, cp->offA, R.n);
    R.put(R.D.corr, cp->corr, R.n);
    R.put(R.D.type, cp->type, R.n);
    R.st_data();
    R.st_data();
    R.st_data();
    R.stats();
    R.smooth();
    if (cp->islandThresh > 0)
        R.islands();
    R.get(cp->offR, R.D.offR, R.n);
    R.get(cp->offA, R.D.offA, R.n);
    R.get(cp->type, R.D.type, R.n);
    R.get(cp->sigmaR, R.D.sigmaR, R.n);
    R.get(cp->sigmaA, R.D.sigmaA, R.n);
    R.get(cp->offRS, R.D.offRS, R.n);
    R.get(cp->offAS, R.D.offAS, R.n);
    ST_HIP(hipStreamSynchronize(c->stream));
}

bool cull_arrays(const st_cull_params *cp)
{
    return cp->corr && cp->type && cp->sigmaR && cp->sigmaA && cp->offRS && cp->offAS;
}

constexpr size_t kNoAllocCells = (size_t)1 << 19;  // 524 288 grid points
// the stage grids (50 B per point), the four input grids (13 B), the mask
// (1 B), with room to spare, and 256-byte alignment of each of the 18 grids
constexpr size_t kNoAllocBytes = kNoAllocCells * (56 + 1 + 13 + 1) + 18 * 256 + 4096;

}  // namespace

// ---- the entry point that allocates nothing ------------------------------
// Every grid of the chain lives in this file-scope device array.  In host
// code `scratch` is only a stand-in name: its device address on each GPU
// comes from hipGetSymbolAddress, resolved once per device (st_ctx_create
// calls cull_noalloc_init), and only that address is handed to the
// kernels and copies.
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex gScratchMu;               // guards gScratch and serialises the calls
void *gScratch[kMaxDevices] = {};    // device address of `scratch` per device

}  // namespace

void cull_noalloc_init(int device)
{
    std::lock_guard<std::mutex> lk(gScratchMu);
    ST_REQUIRE(device >= 0 && device < kMaxDevices, ST_ERR_ARG, "device index");
    if (gScratch[device])
        return;
    void *p = nullptr;
    ST_HIP(hipSetDevice(device));
    ST_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    ST_REQUIRE(p != nullptr, ST_ERR_HIP, "no device address for the cull scratch array");
    gScratch[device] = p;
}

extern "C" int st_cull_run(st_ctx *c, st_cull_params *cp)
{
    return st_guard(c, [&]() {
        CullRun R(c, cp);
        ST_REQUIRE(cull_arrays(cp), ST_ERR_ARG, "cull arrays");
        cull_chain_host(c, cp, R);
    });
}

extern "C" int st_cull_run_noalloc(st_ctx *c, st_cull_params *cp)
{
    return st_guard(c, [&]() {
        ST_REQUIRE(cp && cp->nR > 0 && cp->nA > 0 && cp->offR && cp->offA && cull_arrays(cp), ST_ERR_ARG,
                   "cull grids");
        ST_REQUIRE((size_t)cp->nR * cp->nA <= kNoAllocCells, ST_ERR_ARG,
                   "grid of " + std::to_string((size_t)cp->nR * cp->nA) + " points above the " +
                       std::to_string(kNoAllocCells) + " the no-allocation cull holds");
        // one call at a time per process: every context on a device shares
        // the device's array
        std::lock_guard<std::mutex> lk(gScratchMu);
        void *base = (c->device >= 0 && c->device < kMaxDevices) ? gScratch[c->device] : nullptr;
        ST_REQUIRE(base, ST_ERR_HIP, "cull scratch array not set up on this device");
        CullRun R(c, cp, base, kNoAllocBytes);
        cull_chain_host(c, cp, R);
    });
}

namespace {

// Cullls on the cullst machinery: X / Y in the offR / offA slots, XS / YS
// in offRS / offAS, sigmaX / sigmaY in sigmaR / sigmaA.
struct CullLSRun {
    st_cull_ls_params *cp;
    CullRun R;
    LsArg L{};

    static st_cull_ls_params *check(st_cull_ls_params *p)
    {
        ST_REQUIRE(p && p->nx > 0 && p->ny > 0 && p->X && p->Y, ST_ERR_ARG, "cull grids");
        return p;
    }
    CullLSRun(st_ctx *c, st_cull_ls_params *p) : cp(check(p)), R(c, p->nx, p->ny) {}

    void ls_data()
    {
        ST_REQUIRE(cp->dx != 0 && cp->dy != 0, ST_ERR_ARG, "pixel sizes");
        hipStream_t st = R.ctx->stream;
        ST_HIP(hipMemsetAsync(R.D.counters + 3, 0, 2 * sizeof(int), st));
        hipLaunchKernelGGL(k_cull_ls_begin, dim3(nb(R.n)), dim3(256), 0, st, R.D, R.D.type, cp->noDataToUseMask);
        ST_HIP(hipGetLastError());
        const int nGoodPts = R.counter(3), nWithData = R.counter(4);
        if (cp->nMatch < 0) {
            cp->nAttempt = nWithData;
            cp->nMatch = nGoodPts;
        }
        L.type = R.D.type;
        L.slow = cp->slowVelMask;
        L.noData1 = (double)cp->noData + 1.0;
        L.deltaT = cp->deltaT;
        L.dx = cp->dx;
        L.dy = cp->dy;
        cp->nCulled = R.median_pass<true>(cp->bR, cp->bA, cp->nGood, cp->maxY, cp->maxX, L);
    }
    void ls_stats()
    {
        ST_REQUIRE(cp->nover > 0, ST_ERR_ARG, "nover");
        cp->bR = cp->bA = 9;  // cullLSStats.c:39-40
        const double q = ((1.0 / cp->nover) * (1.0 / cp->nover)) / 12.;
        hipLaunchKernelGGL(k_cull_stats, dim3(nb(R.n)), dim3(256), 0, R.ctx->stream, R.D, 1, q);
        ST_HIP(hipGetLastError());
    }
    void ls_smooth()
    {
        ST_REQUIRE(cp->sX >= 1 && cp->sY >= 1, ST_ERR_ARG, "smoothing window");
        hipLaunchKernelGGL(k_cull_ls_smooth, dim3(nb(R.n)), dim3(256), 0, R.ctx->stream, R.D, cp->sX, cp->sY);
        ST_HIP(hipGetLastError());
    }
};

}  // namespace

extern "C" int st_cull_ls_data(st_ctx *c, st_cull_ls_params *cp)
{
    return st_guard(c, [&]() {
        CullLSRun L(c, cp);
        ST_REQUIRE(cp->type, ST_ERR_ARG, "type");
        L.R.put(L.R.D.offR, cp->X, L.R.n);
        L.R.put(L.R.D.offA, cp->Y, L.R.n);
        L.R.put(L.R.D.type, cp->type, L.R.n);
        L.ls_data();
        L.R.get(cp->X, L.R.D.offR, L.R.n);
        L.R.get(cp->Y, L.R.D.offA, L.R.n);
        ST_HIP(hipStreamSynchronize(c->stream));
    });
}

extern "C" int st_cull_ls_stats(st_ctx *c, st_cull_ls_params *cp)
{
    return st_guard(c, [&]() {
        CullLSRun L(c, cp);
        ST_REQUIRE(cp->sigmaX && cp->sigmaY, ST_ERR_ARG, "sigma outputs");
        L.R.put(L.R.D.offR, cp->X, L.R.n);
        L.R.put(L.R.D.offA, cp->Y, L.R.n);
        L.ls_stats();
        L.R.get(cp->sigmaX, L.R.D.sigmaR, L.R.n);
        L.R.get(cp->sigmaY, L.R.D.sigmaA, L.R.n);
        ST_HIP(hipStreamSynchronize(c->stream));
    });
}

extern "C" int st_cull_ls_smooth(st_ctx *c, st_cull_ls_params *cp)
{
    return st_guard(c, [&]() {
        CullLSRun L(c, cp);
        ST_REQUIRE(cp->sigmaX && cp->sigmaY && cp->XS && cp->YS, ST_ERR_ARG, "smooth arrays");
        L.R.put(L.R.D.offR, cp->X, L.R.n);
        L.R.put(L.R.D.offA, cp->Y, L.R.n);
        L.R.put(L.R.D.sigmaR, cp->sigmaX, L.R.n);
        L.R.put(L.R.D.sigmaA, cp->sigmaY, L.R.n);
        L.ls_smooth();
        L.R.get(cp->sigmaX, L.R.D.sigmaR, L.R.n);
        L.R.get(cp->sigmaY, L.R.D.sigmaA, L.R.n);
        L.R.get(cp->XS, L.R.D.offRS, L.R.n);
        L.R.get(cp->YS, L.R.D.offAS, L.R.n);
        ST_HIP(hipStreamSynchronize(c->stream));
    });
}

extern "C" int st_cull_ls_islands(st_ctx *c, st_cull_ls_params *cp)
{
    return st_guard(c, [&]() {
        CullLSRun L(c, cp);
        ST_REQUIRE(cp->XS, ST_ERR_ARG, "XS");
        L.R.put(L.R.D.offRS, cp->XS, L.R.n);
        if (cp->YS)
            L.R.put(L.R.D.offAS, cp->YS, L.R.n);
        else
            L.R.D.offAS = nullptr;
        L.R.islands_t(cp->islandThresh);
        L.R.get(cp->XS, L.R.D.offRS, L.R.n);
        if (cp->YS)
            L.R.get(cp->YS, L.R.D.offAS, L.R.n);
        ST_HIP(hipStreamSynchronize(c->stream));
    });
}

extern "C" int st_cull_ls_run(st_ctx *c, st_cull_ls_params *cp)
{
    return st_guard(c, [&]() {
        CullLSRun L(c, cp);
        ST_REQUIRE(cp->type && cp->sigmaX && cp->sigmaY && cp->XS && cp->YS, ST_ERR_ARG, "cull arrays");
        L.R.put(L.R.D.offR, cp->X, L.R.n);
        L.R.put(L.R.D.offA, cp->Y, L.R.n);
        L.R.put(L.R.D.type, cp->type, L.R.n);
        cp->nMatch = -1;
        L.ls_data();
        L.ls_data();
        L.ls_data();
        L.ls_stats();
        L.ls_smooth();
        if (cp->islandThresh > 0)
            L.R.islands_t(cp->islandThresh);
        L.R.get(cp->X, L.R.D.offR, L.R.n);
        L.R.get(cp->Y, L.R.D.offA, L.R.n);
        L.R.get(cp->sigmaX, L.R.D.sigmaR, L.R.n);
        L.R.get(cp->sigmaY, L.R.D.sigmaA, L.R.n);
        L.R.get(cp->XS, L.R.D.offRS, L.R.n);
        L.R.get(cp->YS, L.R.D.offAS, L.R.n);
        ST_HIP(hipStreamSynchronize(c->stream));
    });
}
