// Complex match with the transforms as dense contractions on the matrix
// cores (included by st_strack.hip inside its anonymous namespace, after
// KPar / Pts / CpxBatch / PeakRes).
//
// Per patch pair (window W in {16, 32}):
//   S1  L_g  = A_W P_g              (DFT along azimuth, both patches)
//   S2  F_g^T = A_W L_g^T           (DFT along range)
//   X~[u][v] = F_1 conj(F_2) at ((u + az) mod W, v)   (cmpPSWithPad's
//             rotated corners, Nyquist on the negative side)
//   S3  G^T  = B X~^T               (padded inverse along range)
//   S4  C    = B G                  (padded inverse along azimuth)
// with A_W[k][j] = exp(-2 pi i k j / W) and B[m][u] = exp(+2 pi i m u_s /
// 2W), u_s the signed frequency of u.  Each complex product runs as real
// MFMA products on the stacked [re; im] operands; every operand is split
// into bf16 hi + lo parts and a product is Ah Bh + Ah Bl + Al Bh with FP32
// accumulation (about 16 significant bits per operand).  |C| goes to LDS;
// the peak search, the FP64 8x8 -> 48x48 refinement and the Gaussian-fit
// samples follow as in the FP32 path.
#pragma once

typedef __bf16 mf_bf16x8 __attribute__((ext_vector_type(8)));
typedef float mf_f4 __attribute__((ext_vector_type(4)));

// Operand tables of one window size in global memory: for each of the two
// complex matrices (A_W: W x W, B: 2W x W) the real-form A operands of the
// real and imaginary outputs, [row][2W] bf16, hi and lo parts.
struct MfmaTab {
    const __bf16 *fa[2][2];  // [re/im output][hi/lo], W rows
    const __bf16 *ib[2][2];  // [re/im output][hi/lo], 2W rows
};

__device__ inline mf_f4 mf_mma(mf_bf16x8 a, mf_bf16x8 b, mf_f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ inline void mf_split(float x, __bf16 &hi, __bf16 &lo)
{
    hi = (__bf16)x;
    lo = (__bf16)(x - (float)hi);
}

template <int W>
struct MfGeom {
    static constexpr int LDK = 2 * W + 8;                   // row pitch (bf16) of a [n][2W] operand
    static constexpr int PLANE = W * LDK;                   // one W-row operand plane
    // bytes: P (2 patches, hi+lo), then G (2W rows, hi+lo), then the
    // refinement rows T (8 x 48 double2)
    static constexpr int R0 = 8 * W * LDK > 8 * 48 * 16 ? 8 * W * LDK : 8 * 48 * 16;
    // L (2 patches, hi+lo), then |C| (2W x 2W float)
    static constexpr int R1 = 8 * W * LDK > 16 * W * W ? 8 * W * LDK : 16 * W * W;
    static constexpr int R2 = 4 * W * LDK;                  // X~ (hi+lo), then the 8 x 8 doubles
    static constexpr int BYTES = R0 + R1 + R2;
    static constexpr int WPP = W == 16 ? 1 : 4;             // waves per patch
    static constexpr int PPB = 4 / WPP;                     // patches per 256-thread block
};

// One 16 x 16 output tile pair (re, im) of  Y = A X  over KC chunks of 32
// real K values: A from the global table (rows m0..m0+15), X from an LDS
// operand [n][k] (n0..n0+15).
template <int KC>
__device__ inline void mf_tile(const __bf16 *__restrict__ ArH, const __bf16 *__restrict__ ArL,
                               const __bf16 *__restrict__ AiH, const __bf16 *__restrict__ AiL, int lda, int m0,
                               const __bf16 *XH, const __bf16 *XL, int ldx, int n0, mf_f4 &accR, mf_f4 &accI)
{
    const int lane = threadIdx.x & 63;
    const int row = m0 + (lane & 15), kq = 8 * (lane >> 4);
    const int col = n0 + (lane & 15);
#pragma unroll
    for (int kc = 0; kc < KC; kc++) {
        const int ka = row * lda + kc * 32 + kq;
        const mf_bf16x8 arh = *(const mf_bf16x8 *)(ArH + ka), arl = *(const mf_bf16x8 *)(ArL + ka);
        const mf_bf16x8 aih = *(const mf_bf16x8 *)(AiH + ka), ail = *(const mf_bf16x8 *)(AiL + ka);
        const int kx = col * ldx + kc * 32 + kq;
        const mf_bf16x8 xh = *(const mf_bf16x8 *)(XH + kx), xl = *(const mf_bf16x8 *)(XL + kx);
        accR = mf_mma(arh, xh, accR);
        accR = mf_mma(arh, xl, accR);
        accR = mf_mma(arl, xh, accR);
        accI = mf_mma(aih, xh, accI);
        accI = mf_mma(aih, xl, accI);
        accI = mf_mma(ail, xh, accI);
    }
}

// Store a complex value into an LDS operand [n][2W] (re at k, im at W + k).
template <int W>
__device__ inline void mf_put(__bf16 *H, __bf16 *L, int n, int k, float re, float im)
{
    constexpr int LDK = MfGeom<W>::LDK;
    __bf16 h, l;
    mf_split(re, h, l);
    H[n * LDK + k] = h;
    L[n * LDK + k] = l;
    mf_split(im, h, l);
    H[n * LDK + W + k] = h;
    L[n * LDK + W + k] = l;
}

template <int WPP>
__device__ inline void mf_sync()
{
    if constexpr (WPP == 1) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

template <int W, int STYPE>
__global__ void __launch_bounds__(256) k_cpx_mfma(const KPar *__restrict__ Pp, const Pts *__restrict__ Sp,
                                                  const int *list, int count, const CpxBatch *__restrict__ Cp,
                                                  MfmaTab tab)
{
    using G = MfGeom<W>;
    constexpr int WPP = G::WPP, PPB = G::PPB, LDK = G::LDK, WW = 2 * W, TPP = 64 * WPP;
    const KPar &P = *Pp;
    const Pts &S = *Sp;
    const CpxBatch &C = *Cp;
    extern __shared__ __attribute__((aligned(16))) unsigned char mf_lds[];
    __shared__ double shd[4][16];
    __shared__ int shi[4][8];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int slot = wave / WPP;      // patch slot of this wave in the block
    const int wv = wave % WPP;        // wave index within the patch
    const int tp = wv * 64 + lane;    // thread index within the patch
    unsigned char *base = mf_lds + (size_t)slot * G::BYTES;
    __bf16 *bP = (__bf16 *)base;                  // P: [g][hi/lo] planes; later G: [hi/lo] 2W rows
    double2 *T = (double2 *)base;                 // refinement rows (8 x 48), after G is dead
    __bf16 *bL = (__bf16 *)(base + G::R0);        // L: [g][hi/lo]
    float *Mg = (float *)(base + G::R0);          // |C| (2W x 2W), after L is dead
    __bf16 *bX = (__bf16 *)(base + G::R0 + G::R1);  // X~: [hi/lo]
    constexpr bool isf = STYPE == ST_SAMPLE_C64;

    for (int it = blockIdx.x * PPB + slot; it < count + (PPB - 1); it += gridDim.x * PPB) {
        // every wave of a multi-wave patch takes the same branch
        const bool live = it < count && C.have[it];
        if constexpr (WPP > 1) {
            if (!__syncthreads_or(live ? 1 : 0))
                continue;
        } else {
            if (!live)
                continue;
        }
        const int pt = list[it];
        const int az = C.az[it];
        // ---------------- gather both windows into the S1 operand [r][a]
        {
            const int r1 = S.r1[pt], a1 = S.a1[pt], r2 = S.r2[pt], a2 = S.a2[pt];
#pragma unroll 1
            for (int e = tp; e < 2 * W * W; e += TPP) {
                const int g = e / (W * W), a = (e / W) % W, r = e % W;
                const void *img = g ? P.im.img2 : P.im.img1;
                const int nr = g ? P.im.nr2 : P.im.nr1;
                const size_t o = (size_t)((g ? a2 : a1) + a) * nr + (g ? r2 : r1) + r;
                float2 v;
                if constexpr (isf) {
                    v = ((const float2 *)img)[o];
                    v.x *= 1000;
                    v.y *= 1000;
                } else {
                    const short2 q = ((const short2 *)img)[o];
                    v = make_float2((float)q.x, (float)q.y);
                }
                if (g == 1 && P.hanning) {
                    const float hw = P.hann[a * W + r];
                    v.x *= hw;
                    v.y *= hw;
                }
                mf_put<W>(bP + (2 * g) * G::PLANE, bP + (2 * g + 1) * G::PLANE, r, a, v.x, v.y);
            }
        }
        mf_sync<WPP>();
        // ---------------- S1: L_g = A_W P_g
#pragma unroll 1
        for (int u = wv; u < 2 * (W / 16) * (W / 16); u += WPP) {
            const int g = u / ((W / 16) * (W / 16)), mt = (u / (W / 16)) % (W / 16), nt = u % (W / 16);
            mf_f4 ar = {0, 0, 0, 0}, ai = {0, 0, 0, 0};
            mf_tile<W / 16>(tab.fa[0][0], tab.fa[0][1], tab.fa[1][0], tab.fa[1][1], WW, mt * 16,
                            bP + (2 * g) * G::PLANE, bP + (2 * g + 1) * G::PLANE, LDK, nt * 16, ar, ai);
            const int col = nt * 16 + (lane & 15);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int row = mt * 16 + 4 * (lane >> 4) + q;
                mf_put<W>(bL + (2 * g) * G::PLANE, bL + (2 * g + 1) * G::PLANE, row, col, ar[q], ai[q]);
            }
        }
        mf_sync<WPP>();
        // ---------------- S2: F_g^T = A_W L_g^T, then X~ = F_1 conj(F_2) rotated
#pragma unroll 1
        for (int u = wv; u < (W / 16) * (W / 16); u += WPP) {
            const int mt = u / (W / 16), nt = u % (W / 16);
            mf_f4 r0 = {0, 0, 0, 0}, i0 = {0, 0, 0, 0}, r1 = {0, 0, 0, 0}, i1 = {0, 0, 0, 0};
            mf_tile<W / 16>(tab.fa[0][0], tab.fa[0][1], tab.fa[1][0], tab.fa[1][1], WW, mt * 16, bL, bL + G::PLANE,
                            LDK, nt * 16, r0, i0);
            mf_tile<W / 16>(tab.fa[0][0], tab.fa[0][1], tab.fa[1][0], tab.fa[1][1], WW, mt * 16,
                            bL + 2 * G::PLANE, bL + 3 * G::PLANE, LDK, nt * 16, r1, i1);
            const int k = nt * 16 + (lane & 15);  // azimuth frequency
            const int un = (k - az + W) % W;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int l = mt * 16 + 4 * (lane >> 4) + q;  // range frequency
                // F1 conj(F2): re = ar br + ai bi, im = ai br - ar bi
                const float xr = r0[q] * r1[q] + i0[q] * i1[q];
                const float xi = i0[q] * r1[q] - r0[q] * i1[q];
                mf_put<W>(bX, bX + G::PLANE, un, l, xr, xi);
            }
        }
        mf_sync<WPP>();
        // ---------------- S3: G^T = B X~^T   (2W x W)
        __bf16 *bG = bP;
#pragma unroll 1
        for (int u = wv; u < (WW / 16) * (W / 16); u += WPP) {
            const int mt = u / (W / 16), nt = u % (W / 16);
            mf_f4 ar = {0, 0, 0, 0}, ai = {0, 0, 0, 0};
            mf_tile<W / 16>(tab.ib[0][0], tab.ib[0][1], tab.ib[1][0], tab.ib[1][1], WW, mt * 16, bX, bX + G::PLANE,
                            LDK, nt * 16, ar, ai);
            const int col = nt * 16 + (lane & 15);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int row = mt * 16 + 4 * (lane >> 4) + q;
                mf_put<W>(bG, bG + 2 * G::PLANE, row, col, ar[q], ai[q]);
            }
        }
        mf_sync<WPP>();
        // ---------------- S4: C = B G   (2W x 2W) -> fftshifted |C|, top-2
        Top2 best{0.0f, 0x7fffffff, 0.0f};
#pragma unroll 1
        for (int u = wv; u < (WW / 16) * (WW / 16); u += WPP) {
            const int mt = u / (WW / 16), nt = u % (WW / 16);
            mf_f4 ar = {0, 0, 0, 0}, ai = {0, 0, 0, 0};
            mf_tile<W / 16>(tab.ib[0][0], tab.ib[0][1], tab.ib[1][0], tab.ib[1][1], WW, mt * 16, bG, bG + 2 * G::PLANE,
                            LDK, nt * 16, ar, ai);
            const int n = nt * 16 + (lane & 15);
            const int J = (n + W) & (WW - 1);
#pragma unroll
            for (int q = 0; q < 4; q++) {
#pragma clang fp contract(off)
                const int m = mt * 16 + 4 * (lane >> 4) + q;
                const int I = (m + W) & (WW - 1);
                const float x2 = ar[q] * ar[q] + ai[q] * ai[q];
                const float v = (float)sqrt((double)x2);
                Mg[I * WW + J] = v;
                const int i0 = I < W ? I : I - W, j0 = J < W ? J : J - W;
                const int qd = (I < W) ? (J < W ? 0 : 3) : (J < W ? 2 : 1);
                const int rank = (i0 * W + j0) * 4 + qd;
                if (v > 0.0f) {
                    Top2 c{v, rank, 0.0f};
                    best = top2_combine(best, c);
                } else {
                    best.v2 = fmaxf(best.v2, v);
                }
            }
        }
        // top-2 over the patch's threads
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            Top2 b;
            b.v1 = __shfl_xor(best.v1, o);
            b.r1 = __shfl_xor(best.r1, o);
            b.v2 = __shfl_xor(best.v2, o);
            best = top2_combine(best, b);
        }
        if constexpr (WPP > 1) {
            __syncthreads();
            if (lane == 0) {
                shd[wv][0] = best.v1;
                shd[wv][1] = best.v2;
                shi[wv][0] = best.r1;
            }
            __syncthreads();
            best = Top2{(float)shd[0][0], shi[0][0], (float)shd[0][1]};
            for (int k = 1; k < WPP; k++)
                best = top2_combine(best, Top2{(float)shd[k][0], shi[k][0], (float)shd[k][1]});
        }
        mf_sync<WPP>();
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
        if (tp == 0) {
            res->gap = raw > 0 ? (raw - best.v2) / raw : 0.f;
            res->lmOk = 0;
            res->fi = res->fj = -1;
            res->fineMax = 0;
            res->fineGap = 0;
        }
        if (I < NFAST / 2 || I > WW - NFAST / 2 || J < NFAST / 2 || J > WW - NFAST / 2) {
            if (tp == 0) {
                res->I = res->J = -1000;
                res->raw = -1.f;
            }
            mf_sync<WPP>();
            continue;
        }
        if (tp == 0) {
            res->I = I;
            res->J = J;
            res->raw = raw;
        }
        // ---------------- FP64 refinement of the 8 x 8 neighbourhood
        // (overSampleC, speckleTrack.c:852-863): rows first, T[mm][v]
        double *Pm = (double *)bX;
        if (tp < NFAST * NFAST)
            Pm[tp] = (double)Mg[(I - NFAST / 2 + tp / NFAST) * WW + (J - NFAST / 2 + tp % NFAST)];
        mf_sync<WPP>();
#pragma unroll 1
        for (int e = tp; e < NFAST * kFine; e += TPP) {
            const int mm = e / kFine, v = e - mm * kFine;
            double re = 0, im = 0;
            for (int nn = 0; nn < NFAST; nn++) {
                const double2 d = c_dir96[(v + 24) - 12 * nn + 96];
                const double pv = Pm[mm * NFAST + nn];
                re += pv * d.x;
                im += pv * d.y;
            }
            T[e] = make_double2(re, im);
        }
        mf_sync<WPP>();
        double f1 = 0.0, f2 = 0.0;
        int fr = 0x7fffffff;
#pragma unroll 1
        for (int e = tp; e < kFine * kFine; e += TPP) {
            const int uu = e / kFine, v = e - uu * kFine;
            double re = 0, im = 0;
            for (int mm = 0; mm < NFAST; mm++) {
                const double2 d = c_dir96[(uu + 24) - 12 * mm + 96];
                const double2 tv = T[mm * kFine + v];
                re += d.x * tv.x - d.y * tv.y;
                im += d.x * tv.y + d.y * tv.x;
            }
            const double a = sqrt(re * re + im * im);
            if (a > f1 || (a == f1 && e < fr)) {
                f2 = fmax(f2, f1);
                f1 = a;
                fr = e;
            } else
                f2 = fmax(f2, a);
        }
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const double b1 = __shfl_xor(f1, o), b2 = __shfl_xor(f2, o);
            const int br = __shfl_xor(fr, o);
            if (b1 > f1 || (b1 == f1 && br < fr)) {
                f2 = fmax(f1, b2);
                f1 = b1;
                fr = br;
            } else
                f2 = fmax(f2, b1);
        }
        if constexpr (WPP > 1) {
            __syncthreads();
            if (lane == 0) {
                shd[wv][2] = f1;
                shd[wv][3] = f2;
                shi[wv][1] = fr;
            }
            __syncthreads();
            f1 = shd[0][2];
            f2 = shd[0][3];
            fr = shi[0][1];
            for (int k = 1; k < WPP; k++) {
                const double b1 = shd[k][2], b2 = shd[k][3];
                const int br = shi[k][1];
                if (b1 > f1 || (b1 == f1 && br < fr)) {
                    f2 = fmax(f1, b2);
                    f1 = b1;
                    fr = br;
                } else
                    f2 = fmax(f2, b1);
            }
        }
        int fi = 0, fj = 0;
        if (fr != 0x7fffffff) {
            fi = fr / kFine;
            fj = fr - fi * kFine;
        }
        if (tp == 0) {
            res->fi = fi;
            res->fj = fj;
            res->fineMax = fr != 0x7fffffff ? f1 : 0.0;
            res->fineGap = fr != 0x7fffffff ? (float)((f1 - f2) / f1) : 0.f;
        }
        const int iF = I * NOVER + (fi + 24) - (NOVER * NFAST) / 2;
        const int jF = J * NOVER + (fj + 24) - (NOVER * NFAST) / 2;
        const int cI = iF / NOVER, cJ = jF / NOVER;
        const bool ok = (cI - NGAUSS / 2) >= 0 && (cI + NGAUSS / 2) < WW && (cJ - NGAUSS / 2) >= 0 &&
                        (cJ + NGAUSS / 2) < WW;
        if (tp == 0) {
            res->lmOk = ok ? 1 : 0;
            res->lmX0[0] = (float)(cJ - NGAUSS / 2);
            res->lmX0[1] = (float)(cI - NGAUSS / 2);
        }
        if (ok && tp < 2 * NGAUSS) {
            const int k = tp % NGAUSS;
            res->lmY[tp / NGAUSS][k] = tp < NGAUSS ? Mg[cI * WW + (cJ - NGAUSS / 2 + k)]
                                                   : Mg[(cI - NGAUSS / 2 + k) * WW + cJ];
        }
        mf_sync<WPP>();
    }
}
