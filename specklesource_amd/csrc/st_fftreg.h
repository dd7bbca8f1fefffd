// Register-resident radix-2 FFTs for the fused per-patch kernels.
//
// A length-W transform is split over a pair of adjacent lanes (lane ^ 1):
// each lane runs a W/2-point FFT in registers and one exchange of W/2
// values combines the halves.  Forward transforms split the input into
// even / odd samples (decimation in time) and leave the first / second half
// of the output in lanes 0 / 1; inverse transforms take the first / second
// half of the input (decimation in frequency) and leave the even / odd
// outputs.  Twiddles are compile-time literals.
#pragma once

#include "st_common.h"

namespace fftreg {

// exp(-2 pi i j / 64), j = 0..63: (float)cos and (float)sin of the double
// angle -2 pi j / 64, written as literals so the compiler folds them into
// instructions instead of holding them in SGPRs.
constexpr float kTw64Re[64] = {0x1p+0f, 0x1.fd88dap-1f, 0x1.f6297cp-1f, 0x1.e9f416p-1f, 0x1.d906bcp-1f, 0x1.c38b3p-1f, 0x1.a9b662p-1f, 0x1.8bc806p-1f, 0x1.6a09e6p-1f, 0x1.44cf32p-1f, 0x1.1c73b4p-1f, 0x1.e2b5d4p-2f, 0x1.87de2ap-2f, 0x1.294062p-2f, 0x1.8f8b84p-3f, 0x1.917a6cp-4f, 0x1.1a6264p-54f, -0x1.917a6cp-4f, -0x1.8f8b84p-3f, -0x1.294062p-2f, -0x1.87de2ap-2f, -0x1.e2b5d4p-2f, -0x1.1c73b4p-1f, -0x1.44cf32p-1f, -0x1.6a09e6p-1f, -0x1.8bc806p-1f, -0x1.a9b662p-1f, -0x1.c38b3p-1f, -0x1.d906bcp-1f, -0x1.e9f416p-1f, -0x1.f6297cp-1f, -0x1.fd88dap-1f, -0x1p+0f, -0x1.fd88dap-1f, -0x1.f6297cp-1f, -0x1.e9f416p-1f, -0x1.d906bcp-1f, -0x1.c38b3p-1f, -0x1.a9b662p-1f, -0x1.8bc806p-1f, -0x1.6a09e6p-1f, -0x1.44cf32p-1f, -0x1.1c73b4p-1f, -0x1.e2b5d4p-2f, -0x1.87de2ap-2f, -0x1.294062p-2f, -0x1.8f8b84p-3f, -0x1.917a6cp-4f, -0x1.a79394p-53f, 0x1.917a6cp-4f, 0x1.8f8b84p-3f, 0x1.294062p-2f, 0x1.87de2ap-2f, 0x1.e2b5d4p-2f, 0x1.1c73b4p-1f, 0x1.44cf32p-1f, 0x1.6a09e6p-1f, 0x1.8bc806p-1f, 0x1.a9b662p-1f, 0x1.c38b3p-1f, 0x1.d906bcp-1f, 0x1.e9f416p-1f, 0x1.f6297cp-1f, 0x1.fd88dap-1f};
constexpr float kTw64Im[64] = {-0x0p+0f, -0x1.917a6cp-4f, -0x1.8f8b84p-3f, -0x1.294062p-2f, -0x1.87de2ap-2f, -0x1.e2b5d4p-2f, -0x1.1c73b4p-1f, -0x1.44cf32p-1f, -0x1.6a09e6p-1f, -0x1.8bc806p-1f, -0x1.a9b662p-1f, -0x1.c38b3p-1f, -0x1.d906bcp-1f, -0x1.e9f416p-1f, -0x1.f6297cp-1f, -0x1.fd88dap-1f, -0x1p+0f, -0x1.fd88dap-1f, -0x1.f6297cp-1f, -0x1.e9f416p-1f, -0x1.d906bcp-1f, -0x1.c38b3p-1f, -0x1.a9b662p-1f, -0x1.8bc806p-1f, -0x1.6a09e6p-1f, -0x1.44cf32p-1f, -0x1.1c73b4p-1f, -0x1.e2b5d4p-2f, -0x1.87de2ap-2f, -0x1.294062p-2f, -0x1.8f8b84p-3f, -0x1.917a6cp-4f, -0x1.1a6264p-53f, 0x1.917a6cp-4f, 0x1.8f8b84p-3f, 0x1.294062p-2f, 0x1.87de2ap-2f, 0x1.e2b5d4p-2f, 0x1.1c73b4p-1f, 0x1.44cf32p-1f, 0x1.6a09e6p-1f, 0x1.8bc806p-1f, 0x1.a9b662p-1f, 0x1.c38b3p-1f, 0x1.d906bcp-1f, 0x1.e9f416p-1f, 0x1.f6297cp-1f, 0x1.fd88dap-1f, 0x1p+0f, 0x1.fd88dap-1f, 0x1.f6297cp-1f, 0x1.e9f416p-1f, 0x1.d906bcp-1f, 0x1.c38b3p-1f, 0x1.a9b662p-1f, 0x1.8bc806p-1f, 0x1.6a09e6p-1f, 0x1.44cf32p-1f, 0x1.1c73b4p-1f, 0x1.e2b5d4p-2f, 0x1.87de2ap-2f, 0x1.294062p-2f, 0x1.8f8b84p-3f, 0x1.917a6cp-4f};
// exp(+2 pi i j / 128), j = 0..127
extern __constant__ float2 c_tw128[128];

__device__ inline float2 cmul_f(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// exp(SIGN 2 pi i k / len) for len dividing 64
template <int SIGN>
__device__ inline float2 tw(int k, int len)
{
    float2 w = make_float2(kTw64Re[k * (64 / len)], kTw64Im[k * (64 / len)]);
    if (SIGN > 0)
        w.y = -w.y;
    return w;
}

template <int N>
constexpr int bitrev(int i)
{
    int j = 0;
    for (int b = 0; (1 << b) < N; b++)
        if (i & (1 << b))
            j |= (N >> (b + 1));
    return j;
}

// In-place N-point FFT, natural order in and out.
template <int N, int SIGN>
__device__ inline void fft(float2 (&x)[N])
{
#pragma unroll
    for (int i = 0; i < N; i++) {
        const int j = bitrev<N>(i);
        if (j > i) {
            float2 t = x[i];
            x[i] = x[j];
            x[j] = t;
        }
    }
#pragma unroll
    for (int len = 2; len <= N; len <<= 1) {
#pragma unroll
        for (int i = 0; i < N; i += len) {
#pragma unroll
            for (int k = 0; k < len / 2; k++) {
                const float2 a = x[i + k], b = x[i + k + len / 2];
                float2 tb;
                if (k == 0)
                    tb = b;
                else if (4 * k == len)
                    tb = SIGN < 0 ? make_float2(b.y, -b.x) : make_float2(-b.y, b.x);
                else
                    tb = cmul_f(b, tw<SIGN>(k, len));
                x[i + k] = make_float2(a.x + tb.x, a.y + tb.y);
                x[i + k + len / 2] = make_float2(a.x - tb.x, a.y - tb.y);
            }
        }
    }
}

// Value of the partner lane (lane ^ 1) via DPP quad_perm [1,0,3,2].
__device__ inline float dpp_swap1(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
}

__device__ inline float2 shfl_pair(float2 v)
{
    return make_float2(dpp_swap1(v.x), dpp_swap1(v.y));
}

// Forward W-point FFT over a lane pair.  Lane h holds samples 2m + h
// (m = 0..W/2-1); on return it holds outputs h*W/2 + k.  The odd lane
// applies the twiddle to its own half before the exchange (the even lane
// multiplies by 1, exactly), so both lanes finish with one sign-selected
// add: out = partner + (h ? -1 : 1) * own.  The results are bit-identical
// to forming E + tw O and E - tw O after the exchange.
template <int W>
__device__ inline void fwd_pair(float2 (&x)[W / 2], int h)
{
    constexpr int H = W / 2;
    fft<H, -1>(x);
    const float hf = (float)h, sg = h ? -1.f : 1.f;
#pragma unroll
    for (int k = 0; k < H; k++) {
        float2 t;
        if (k == 0) {
            t = x[k];
        } else if (4 * k == W) {
            t = h ? make_float2(x[k].y, -x[k].x) : x[k];
        } else {
            const float2 w = tw<-1>(k, W);
            t = cmul_f(x[k], make_float2(h ? w.x : 1.f, hf * w.y));
        }
        const float2 y = shfl_pair(t);
        x[k] = make_float2(fmaf(sg, t.x, y.x), fmaf(sg, t.y, y.y));
    }
}

// Inverse (exp(+2 pi i jk/W)) W-point FFT over a lane pair.  Lane h holds
// inputs h*W/2 + n; on return it holds outputs 2k + h.  (Forming the
// butterfly as in fwd_pair, one sign-selected add and a per-lane twiddle,
// raised register pressure here: 72 VGPRs spilled at W = 64.)
template <int W>
__device__ inline void inv_pair(float2 (&x)[W / 2], int h)
{
    constexpr int H = W / 2;
#pragma unroll
    for (int n = 0; n < H; n++) {
        const float2 y = shfl_pair(x[n]);
        const float2 a = make_float2(x[n].x + y.x, x[n].y + y.y);
        const float2 d = make_float2(y.x - x[n].x, y.y - x[n].y);
        float2 b;
        if (n == 0)
            b = d;
        else if (4 * n == W)
            b = make_float2(-d.y, d.x);
        else
            b = cmul_f(d, tw<+1>(n, W));
        x[n] = h ? b : a;
    }
    fft<H, +1>(x);
}

// Lane-group FFTs (lane pairs): the block of outputs (forward) / inputs
// (inverse) held by lane h, and the transforms.
template <int W, int L>
struct LaneFFT {
    static_assert(L == 2, "lane pairs");
    __device__ static inline int blk(int h) { return h; }
    __device__ static inline void fwd(float2 (&x)[W / L], int h) { fwd_pair<W>(x, h); }
    __device__ static inline void inv(float2 (&x)[W / L], int h) { inv_pair<W>(x, h); }
};

}  // namespace fftreg
