// Complex match with the transforms as dense contractions on the matrix
// cores (included by st_strack.hip inside its anonymous namespace, after
// KPar / Pts / CpxBatch / PeakRes / Top2).
//
// Per patch pair (window W in {16, 32}), with A_W[k][j] = exp(-2 pi i k j
// / W) and B[m][u] = exp(+2 pi i m u_s / 2W) (u_s the signed frequency of
// u, Nyquist on the negative side):
//   S1  L_g^T[r][k] = sum_a P_g^T[r][a] A_W[a][k]      (along azimuth)
//   S2  F_g^T[l][k] = sum_r A_W[l][r] L_g^T[r][k]      (along range)
//   X~[u][v] = F_1 conj(F_2) at ((u + az) mod W, v)    (cmpPSWithPad's
//             rotated corners)
//   S3  G[u][n]     = sum_v X~[u][v] B[n][v]           (padded inverse, range)
//   S4  C[m][n]     = sum_u B[m][u] G[u][n]            (padded inverse, azimuth)
// The operand roles alternate so that every stage contracts over the row
// index of the previous stage's output tile: each lane's four accumulator
// rows then land as one contiguous 16-byte run of the next operand.
// Complex values are interleaved (re, im) along K; a complex product is two
// real MFMA products against the [re, -im] / [im, re] table rows.  Every
// operand is split into bf16 hi + lo and a product is Ah Bh + Ah Bl + Al Bh
// with FP32 accumulation (about 16 significant bits per operand).  |C| goes
// to LDS; the peak search, the FP64 8x8 -> 48x48 refinement and the
// Gaussian-fit samples follow as in the FP32 path.
#pragma once

typedef __bf16 mf_bf16x8 __attribute__((ext_vector_type(8)));
typedef float mf_f4 __attribute__((ext_vector_type(4)));

// Table operands of one window size in global memory, [row][2W] bf16 with
// (re, im) interleaved along K: fa for A_W (W rows), ib for B (2W rows);
// [0] gives the real part of a product, [1] the imaginary part; hi and lo.
struct MfmaTab {
    const __bf16 *fa[2][2];
    const __bf16 *ib[2][2];
};

__device__ inline mf_f4 mf_mma(mf_bf16x8 a, mf_bf16x8 b, mf_f4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// 8 consecutive K values of row `row` (chunk kc) for this lane.
__device__ inline mf_bf16x8 mf_frag(const __bf16 *p, int pitch, int row, int kc)
{
    return *(const mf_bf16x8 *)(p + row * pitch + kc * 32 + 8 * ((threadIdx.x & 63) >> 4));
}

// hi/lo split of four complex values, packed as (re, im) pairs
__device__ inline void mf_split4(const mf_f4 &re, const mf_f4 &im, mf_bf16x8 &hi, mf_bf16x8 &lo)
{
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const __bf16 hr = (__bf16)re[q], hi_ = (__bf16)im[q];
        hi[2 * q] = hr;
        hi[2 * q + 1] = hi_;
        lo[2 * q] = (__bf16)(re[q] - (float)hr);
        lo[2 * q + 1] = (__bf16)(im[q] - (float)hi_);
    }
}

template <int W>
struct MfGeom {
    static constexpr int LDK = 2 * W + 8;  // row pitch (bf16) of an LDS operand
    static constexpr int PLANE = W * LDK;  // W rows
    // bytes of one patch slot: P (2 patches, hi+lo), later G (2W rows,
    // hi+lo); X~ (hi+lo); L (2 patches, hi+lo), later |C| (2W x 2W float,
    // transposed).  The refinement rows T (8 x 48 double2) overlay P/G and
    // the head of X~; the 8 x 8 doubles Pm sit at the end of X~.
    static constexpr int RP = 8 * W * LDK;
    static constexpr int RX = 4 * W * LDK;
    static constexpr int RL = 8 * W * LDK > 16 * W * W ? 8 * W * LDK : 16 * W * W;
    static constexpr int BYTES = RP + RX + RL;
    static constexpr int WPP = W == 16 ? 1 : 4;  // waves per patch
    static constexpr int PPB = 4 / WPP;          // patches per 256-thread block
};

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

// Table fragments of one 16-row block: [chunk][re/im][hi/lo]
template <int KC>
struct MfTab {
    mf_bf16x8 f[KC][2][2];
};

template <int KC>
__device__ inline void mf_load_tab(MfTab<KC> &t, const __bf16 *const (&tab)[2][2], int pitch, int row)
{
#pragma unroll
    for (int kc = 0; kc < KC; kc++)
#pragma unroll
        for (int o = 0; o < 2; o++)
#pragma unroll
            for (int hl = 0; hl < 2; hl++)
                t.f[kc][o][hl] = mf_frag(tab[o][hl], pitch, row, kc);
}

// acc{R,I} += (data x table) with the data as the A operand (rows of the
// output), or (table x data) with the table as the A operand.
template <int KC, bool DATA_A>
__device__ inline void mf_mul(const MfTab<KC> &t, const __bf16 *dH, const __bf16 *dL, int pitch, int row,
                              mf_f4 &accR, mf_f4 &accI)
{
#pragma unroll
    for (int kc = 0; kc < KC; kc++) {
        const mf_bf16x8 xh = mf_frag(dH, pitch, row, kc), xl = mf_frag(dL, pitch, row, kc);
        if constexpr (DATA_A) {
            accR = mf_mma(xh, t.f[kc][0][0], accR);
            accR = mf_mma(xl, t.f[kc][0][0], accR);
            accR = mf_mma(xh, t.f[kc][0][1], accR);
            accI = mf_mma(xh, t.f[kc][1][0], accI);
            accI = mf_mma(xl, t.f[kc][1][0], accI);
            accI = mf_mma(xh, t.f[kc][1][1], accI);
        } else {
            accR = mf_mma(t.f[kc][0][0], xh, accR);
            accR = mf_mma(t.f[kc][0][0], xl, accR);
            accR = mf_mma(t.f[kc][0][1], xh, accR);
            accI = mf_mma(t.f[kc][1][0], xh, accI);
            accI = mf_mma(t.f[kc][1][0], xl, accI);
            accI = mf_mma(t.f[kc][1][1], xh, accI);
        }
    }
}

template <int W, int STYPE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W == 16 ? 3 : 2, 8))) k_cpx_mfma(const KPar *__restrict__ Pp, const Pts *__restrict__ Sp,
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
    __bf16 *bX = (__bf16 *)(base + G::RP);        // X~: [hi/lo]
    __bf16 *bL = (__bf16 *)(base + G::RP + G::RX);  // L: [g][hi/lo]
    float *Mg = (float *)(base + G::RP + G::RX);    // |C| (2W x 2W), after L is dead
    constexpr bool isf = STYPE == ST_SAMPLE_C64;

    // Window samples of a patch: NG per thread (element e = tp + k TPP of
    // [g][a][r]), all loads issued before the first conversion.
    constexpr int NG = 2 * W * W / TPP;
    static_assert(NG * TPP == 2 * W * W, "gather split");
    using Raw = typename std::conditional<isf, float2, short2>::type;
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
        // ---------------- both windows into the S1 operand [r][2a + c]
        {
            const int r1 = S.r1[pt], a1 = S.a1[pt], r2 = S.r2[pt], a2 = S.a2[pt];
            Raw pre[NG];
#pragma unroll
            for (int k = 0; k < NG; k++) {
                const int e = tp + k * TPP;
                const int g = e / (W * W), a = (e / W) % W, r = e % W;
                const void *img = g ? P.im.img2 : P.im.img1;
                const int nr = g ? P.im.nr2 : P.im.nr1;
                const size_t o = (size_t)((g ? a2 : a1) + a) * nr + (g ? r2 : r1) + r;
                pre[k] = ((const Raw *)img)[o];
            }
#pragma unroll
            for (int k = 0; k < NG; k++) {
                const int e = tp + k * TPP;
                const int g = e / (W * W), a = (e / W) % W, r = e % W;
                float2 v;
                if constexpr (isf) {
                    v = pre[k];
                    v.x *= 1000;
                    v.y *= 1000;
                } else {
                    v = make_float2((float)pre[k].x, (float)pre[k].y);
                }
                if (g == 1 && P.hanning) {
                    const float hw = P.hann[a * W + r];
                    v.x *= hw;
                    v.y *= hw;
                }
                const __bf16 hr = (__bf16)v.x, hi_ = (__bf16)v.y;
                __bf16 *H = bP + (2 * g) * G::PLANE + r * LDK + 2 * a;
                __bf16 *L = bP + (2 * g + 1) * G::PLANE + r * LDK + 2 * a;
                H[0] = hr;
                H[1] = hi_;
                L[0] = (__bf16)(v.x - (float)hr);
                L[1] = (__bf16)(v.y - (float)hi_);
            }
        }
        mf_sync<WPP>();
        const int lrow = lane & 15, lq = 4 * (lane >> 4);
        // ---------------- S1: L_g^T = P_g^T A_W   (tiles: table k-block x (g, r-block))
        {
            constexpr int NTT = W / 16, ND = 2 * (W / 16), NU = NTT * ND, PER = NU / WPP;
            MfTab<W / 16> tb;
            int tcur = -1;
#pragma unroll 1
            for (int u = wv * PER; u < (wv + 1) * PER; u++) {
                const int tt = u / ND, dt = u % ND, g = dt / (W / 16), mt = dt % (W / 16);
                if (tt != tcur) {
                    mf_load_tab<W / 16>(tb, tab.fa, WW, tt * 16 + lrow);
                    tcur = tt;
                }
                mf_f4 ar = {0, 0, 0, 0}, ai = {0, 0, 0, 0};
                mf_mul<W / 16, true>(tb, bP + (2 * g) * G::PLANE, bP + (2 * g + 1) * G::PLANE, LDK, mt * 16 + lrow,
                                     ar, ai);
                // rows r = mt*16 + lq + 0..3, column k = tt*16 + lrow -> L operand [k][2r + c]
                mf_bf16x8 hv, lv;
                mf_split4(ar, ai, hv, lv);
                const int k = tt * 16 + lrow, r0 = mt * 16 + lq;
                *(mf_bf16x8 *)(bL + (2 * g) * G::PLANE + k * LDK + 2 * r0) = hv;
                *(mf_bf16x8 *)(bL + (2 * g + 1) * G::PLANE + k * LDK + 2 * r0) = lv;
            }
        }
        mf_sync<WPP>();
        // ---------------- S2: F_g^T = A_W L_g^T, then X~ = F_1 conj(F_2) rotated
        {
            constexpr int NTT = W / 16, ND = W / 16, NU = NTT * ND, PER = (NU + WPP - 1) / WPP;
            MfTab<W / 16> tb;
            int tcur = -1;
#pragma unroll 1
            for (int u = wv * PER; u < (wv + 1) * PER && u < NU; u++) {
                const int tt = u / ND, nt = u % ND;
                if (tt != tcur) {
                    mf_load_tab<W / 16>(tb, tab.fa, WW, tt * 16 + lrow);
                    tcur = tt;
                }
                mf_f4 r0 = {0, 0, 0, 0}, i0 = {0, 0, 0, 0}, r1 = {0, 0, 0, 0}, i1 = {0, 0, 0, 0};
                mf_mul<W / 16, false>(tb, bL, bL + G::PLANE, LDK, nt * 16 + lrow, r0, i0);
                mf_mul<W / 16, false>(tb, bL + 2 * G::PLANE, bL + 3 * G::PLANE, LDK, nt * 16 + lrow, r1, i1);
                // rows l = tt*16 + lq + 0..3 (range frequency), column k = nt*16 + lrow
                mf_f4 xr, xi;
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    // F1 conj(F2): re = ar br + ai bi, im = ai br - ar bi
                    xr[q] = r0[q] * r1[q] + i0[q] * i1[q];
                    xi[q] = i0[q] * r1[q] - r0[q] * i1[q];
                }
                mf_bf16x8 hv, lv;
                mf_split4(xr, xi, hv, lv);
                const int un = (nt * 16 + lrow - az + W) % W, l0 = tt * 16 + lq;
                *(mf_bf16x8 *)(bX + un * LDK + 2 * l0) = hv;
                *(mf_bf16x8 *)(bX + G::PLANE + un * LDK + 2 * l0) = lv;
            }
        }
        mf_sync<WPP>();
        // ---------------- S3: G = X~ B^T   (2W columns: table n-block x u-block)
        __bf16 *bG = bP;
        {
            constexpr int NTT = WW / 16, ND = W / 16, NU = NTT * ND, PER = NU / WPP;
            MfTab<W / 16> tb;
            int tcur = -1;
#pragma unroll 1
            for (int u = wv * PER; u < (wv + 1) * PER; u++) {
                const int tt = u / ND, mt = u % ND;
                if (tt != tcur) {
                    mf_load_tab<W / 16>(tb, tab.ib, WW, tt * 16 + lrow);
                    tcur = tt;
                }
                mf_f4 ar = {0, 0, 0, 0}, ai = {0, 0, 0, 0};
                mf_mul<W / 16, true>(tb, bX, bX + G::PLANE, LDK, mt * 16 + lrow, ar, ai);
                // rows u = mt*16 + lq + 0..3, column n = tt*16 + lrow -> G operand [n][2u + c]
                mf_bf16x8 hv, lv;
                mf_split4(ar, ai, hv, lv);
                const int n = tt * 16 + lrow, u0 = mt * 16 + lq;
                *(mf_bf16x8 *)(bG + n * LDK + 2 * u0) = hv;
                *(mf_bf16x8 *)(bG + 2 * G::PLANE + n * LDK + 2 * u0) = lv;
            }
        }
        mf_sync<WPP>();
        // ---------------- S4: C = B G   (table m-block x n-block) -> |C|, top-2
        Top2 best{0.0f, 0x7fffffff, 0.0f, 0.0f};
        {
            constexpr int NTT = WW / 16, ND = WW / 16, NU = NTT * ND, PER = NU / WPP;
            MfTab<W / 16> tb;
            int tcur = -1;
#pragma unroll 1
            for (int u = wv * PER; u < (wv + 1) * PER; u++) {
                const int tt = u / ND, nt = u % ND;
                if (tt != tcur) {
                    mf_load_tab<W / 16>(tb, tab.ib, WW, tt * 16 + lrow);
                    tcur = tt;
                }
                mf_f4 ar = {0, 0, 0, 0}, ai = {0, 0, 0, 0};
                mf_mul<W / 16, false>(tb, bG, bG + 2 * G::PLANE, LDK, nt * 16 + lrow, ar, ai);
                // rows m = tt*16 + lq + 0..3, column n = nt*16 + lrow
                const int n = nt * 16 + lrow;
                const int J = (n + W) & (WW - 1);
                const int I0 = (tt * 16 + lq + W) & (WW - 1);
                mf_f4 mv;
#pragma unroll
                for (int q = 0; q < 4; q++) {
#pragma clang fp contract(off)
                    const int I = I0 + q;
                    const float x2 = ar[q] * ar[q] + ai[q] * ai[q];
                    const float v = sqrtf(x2);
                    mv[q] = v;
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
                *(mf_f4 *)(Mg + J * WW + I0) = mv;  // transposed: Mg[J][I]
            }
        }
        // top-2 over the patch's threads
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            Top2 b;
            b.v1 = __shfl_xor(best.v1, o);
            b.r1 = __shfl_xor(best.r1, o);
            b.v2 = __shfl_xor(best.v2, o);
            b.sum = __shfl_xor(best.sum, o);
            best = top2_combine(best, b);
        }
        if constexpr (WPP > 1) {
            __syncthreads();
            if (lane == 0) {
                shd[wv][0] = best.v1;
                shd[wv][1] = best.v2;
                shd[wv][4] = best.sum;
                shi[wv][0] = best.r1;
            }
            __syncthreads();
            best = Top2{(float)shd[0][0], shi[0][0], (float)shd[0][1], (float)shd[0][4]};
            for (int k = 1; k < WPP; k++)
                best = top2_combine(best, Top2{(float)shd[k][0], shi[k][0], (float)shd[k][1], (float)shd[k][4]});
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
            res->snr = best.sum > 0.0f ? best.v1 * (float)(WW * WW) / best.sum : 0.0f;
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
        // the block cells k_cpx_refine reads (the refinement and the fit
        // samples run there, as on the FP32 fused path); Mg is transposed
        float *nb = C.nb + (size_t)it * kNbc;
        for (int e = tp; e < kNbc; e += TPP) {
            const int ri = I - kNb / 2 + kNbMap.ri[e], ci = J - kNb / 2 + kNbMap.ci[e];
            nb[e] = (ri >= 0 && ri < WW && ci >= 0 && ci < WW) ? Mg[ci * WW + ri] : 0.0f;
        }
        mf_sync<WPP>();
    }
}
