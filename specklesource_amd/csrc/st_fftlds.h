// Device building blocks of the FFT engine: radix butterflies and an
// in-place mixed-radix Stockham transform of several sequences held in LDS.
//
// lds_fft<MAXE, NTH> transforms `nseq` sequences of length P.n that sit in
// LDS at buf + s * pitch (element stride 1).  Every stage reads each
// butterfly's inputs into registers, synchronises the block, and writes the
// Stockham outputs back into the same buffer, so one buffer of
// nseq * pitch elements suffices.  MAXE bounds nseq * P.n (it sizes the
// register staging), NTH is the block size; all NTH threads must call it.
// Conventions are those of st_fft.h (sign -1 forward, no normalisation).
#pragma once

#include "st_fft.h"

namespace stfft {

__device__ inline float2 mul_si(float2 v, int sign)
{
    // v * (sign * i)
    return sign < 0 ? make_float2(v.y, -v.x) : make_float2(-v.y, v.x);
}

__device__ inline void dft2(float2 *v)
{
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

__device__ inline void dft4(float2 &a, float2 &b, float2 &c, float2 &d, int sign)
{
    float2 s0 = cadd(a, c), d0 = csub(a, c), s1 = cadd(b, d), d1 = mul_si(csub(b, d), sign);
    a = cadd(s0, s1);
    b = cadd(d0, d1);
    c = csub(s0, s1);
    d = csub(d0, d1);
}

__device__ inline void dft8(float2 *v, int sign)
{
    const float h = 0.70710678118654752440f;
    float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4(e0, e1, e2, e3, sign);
    dft4(o0, o1, o2, o3, sign);
    // W8^1 = (1 + s i)/sqrt2, W8^2 = s i, W8^3 = (-1 + s i)/sqrt2
    float2 w1 = make_float2(h, sign * h), w3 = make_float2(-h, sign * h);
    o1 = cmul(o1, w1);
    o2 = mul_si(o2, sign);
    o3 = cmul(o3, w3);
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1);
    v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2);
    v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3);
    v[7] = csub(e3, o3);
}

__device__ inline void dft3(float2 *v, int sign)
{
    const float c = -0.5f, s = 0.86602540378443864676f * sign;
    float2 a = v[0], b = v[1], d = v[2];
    float2 t1 = cadd(b, d), t2 = csub(b, d);
    float2 m = make_float2(a.x + c * t1.x, a.y + c * t1.y);
    float2 r = make_float2(-s * t2.y, s * t2.x);
    v[0] = cadd(a, t1);
    v[1] = cadd(m, r);
    v[2] = csub(m, r);
}

template <int P>
__device__ inline void dft_generic(float2 *v, int sign)
{
    float2 o[P];
#pragma unroll
    for (int q = 0; q < P; q++) {
        float sr = 0, si = 0;
#pragma unroll
        for (int r = 0; r < P; r++) {
            const int e = (q * r) % P;
            float cs, sn;
            sincospif(2.0f * (float)e / (float)P, &sn, &cs);
            sn *= sign;
            sr += v[r].x * cs - v[r].y * sn;
            si += v[r].x * sn + v[r].y * cs;
        }
        o[q] = make_float2(sr, si);
    }
#pragma unroll
    for (int q = 0; q < P; q++)
        v[q] = o[q];
}

template <int P>
__device__ inline void dft_any(float2 *v, int sign)
{
    if constexpr (P == 2)
        dft2(v);
    else if constexpr (P == 4)
        dft4(v[0], v[1], v[2], v[3], sign);
    else if constexpr (P == 8)
        dft8(v, sign);
    else if constexpr (P == 3)
        dft3(v, sign);
    else
        dft_generic<P>(v, sign);
}

// One Stockham stage.  Kept out of line: inlined into the stage loop of
// lds_fft, the six radix cases did not share registers and the caller
// needed twice the VGPRs of the largest case.
template <int R, int MAXE, int NTH, int SIGN>
__device__ __attribute__((noinline)) void lds_stage(float2 *buf, int pitch, int nseq, int n, int ns, FastDiv dstride,
                                                    FastDiv dns, const float2 *tw)
{
    constexpr int sign = SIGN;
    static_assert(MAXE <= 8192, "FastDiv needs indices below 2^13");
    constexpr int IT = (MAXE / R + NTH - 1) / NTH;
    const int stride = n / R, work = nseq * stride;
    float2 v[IT][R];
    int at[IT];
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int t = threadIdx.x + it * NTH;
        at[it] = -1;
        if (t < work) {
            const int sq = fdiv(t, dstride), j = t - sq * stride, jq = fdiv(j, dns), k = j - jq * ns;
            const float2 *src = buf + sq * pitch;
            v[it][0] = src[j];
#pragma unroll
            for (int r = 1; r < R; r++)
                v[it][r] = cmul(src[j + r * stride], tw[k * R + r]);
            dft_any<R>(v[it], sign);
            at[it] = sq * pitch + jq * ns * R + k;
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; it++)
        if (at[it] >= 0) {
#pragma unroll
            for (int q = 0; q < R; q++)
                buf[at[it] + q * ns] = v[it][q];
        }
    __syncthreads();
}

template <int MAXE, int NTH, int SIGN>
__device__ inline void lds_fft_s(float2 *buf, int pitch, int nseq, const PlanArgs &P)
{
    for (int s = 0; s < P.nst; s++) {
        const int ns = P.ns[s];
        const FastDiv ds = P.dstride[s], dn = P.dns[s];
        const float2 *tw = P.tw + P.twoff[s];
        switch (P.radix[s]) {
        case 8: lds_stage<8, MAXE, NTH, SIGN>(buf, pitch, nseq, P.n, ns, ds, dn, tw); break;
        case 4: lds_stage<4, MAXE, NTH, SIGN>(buf, pitch, nseq, P.n, ns, ds, dn, tw); break;
        case 2: lds_stage<2, MAXE, NTH, SIGN>(buf, pitch, nseq, P.n, ns, ds, dn, tw); break;
        case 3: lds_stage<3, MAXE, NTH, SIGN>(buf, pitch, nseq, P.n, ns, ds, dn, tw); break;
        case 5: lds_stage<5, MAXE, NTH, SIGN>(buf, pitch, nseq, P.n, ns, ds, dn, tw); break;
        default: lds_stage<7, MAXE, NTH, SIGN>(buf, pitch, nseq, P.n, ns, ds, dn, tw); break;
        }
    }
}

// The sign is a template parameter of the stages (the butterflies' +-i
// rotations fold into adds and subtracts); one branch picks the instance.
template <int MAXE, int NTH>
__device__ inline void lds_fft(float2 *buf, int pitch, int nseq, const PlanArgs &P, int sign)
{
    if (sign < 0)
        lds_fft_s<MAXE, NTH, -1>(buf, pitch, nseq, P);
    else
        lds_fft_s<MAXE, NTH, +1>(buf, pitch, nseq, P);
}

}  // namespace stfft
