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

// exp(-2 pi i j / 128), j = 0..127: (float)cos and (float)sin of the double
// angle -2 pi j / 128, written as literals so the compiler folds them into
// instructions instead of holding them in SGPRs.
constexpr float kTw128Re[128] = {0x1p+0f, 0x1.ff621ep-1f, 0x1.fd88dap-1f, 0x1.fa7558p-1f, 0x1.f6297cp-1f, 0x1.f0a7fp-1f, 0x1.e9f416p-1f, 0x1.e2121p-1f, 0x1.d906bcp-1f, 0x1.ced7bp-1f, 0x1.c38b3p-1f, 0x1.b72834p-1f, 0x1.a9b662p-1f, 0x1.9b3e04p-1f, 0x1.8bc806p-1f, 0x1.7b5df2p-1f, 0x1.6a09e6p-1f, 0x1.57d694p-1f, 0x1.44cf32p-1f, 0x1.30ff8p-1f, 0x1.1c73b4p-1f, 0x1.07387ap-1f, 0x1.e2b5d4p-2f, 0x1.b5d1p-2f, 0x1.87de2ap-2f, 0x1.58f9a8p-2f, 0x1.294062p-2f, 0x1.f19f98p-3f, 0x1.8f8b84p-3f, 0x1.2c8106p-3f, 0x1.917a6cp-4f, 0x1.91f66p-5f, 0x1.1a6264p-54f, -0x1.91f66p-5f, -0x1.917a6cp-4f, -0x1.2c8106p-3f, -0x1.8f8b84p-3f, -0x1.f19f98p-3f, -0x1.294062p-2f, -0x1.58f9a8p-2f, -0x1.87de2ap-2f, -0x1.b5d1p-2f, -0x1.e2b5d4p-2f, -0x1.07387ap-1f, -0x1.1c73b4p-1f, -0x1.30ff8p-1f, -0x1.44cf32p-1f, -0x1.57d694p-1f, -0x1.6a09e6p-1f, -0x1.7b5df2p-1f, -0x1.8bc806p-1f, -0x1.9b3e04p-1f, -0x1.a9b662p-1f, -0x1.b72834p-1f, -0x1.c38b3p-1f, -0x1.ced7bp-1f, -0x1.d906bcp-1f, -0x1.e2121p-1f, -0x1.e9f416p-1f, -0x1.f0a7fp-1f, -0x1.f6297cp-1f, -0x1.fa7558p-1f, -0x1.fd88dap-1f, -0x1.ff621ep-1f, -0x1p+0f, -0x1.ff621ep-1f, -0x1.fd88dap-1f, -0x1.fa7558p-1f, -0x1.f6297cp-1f, -0x1.f0a7fp-1f, -0x1.e9f416p-1f, -0x1.e2121p-1f, -0x1.d906bcp-1f, -0x1.ced7bp-1f, -0x1.c38b3p-1f, -0x1.b72834p-1f, -0x1.a9b662p-1f, -0x1.9b3e04p-1f, -0x1.8bc806p-1f, -0x1.7b5df2p-1f, -0x1.6a09e6p-1f, -0x1.57d694p-1f, -0x1.44cf32p-1f, -0x1.30ff8p-1f, -0x1.1c73b4p-1f, -0x1.07387ap-1f, -0x1.e2b5d4p-2f, -0x1.b5d1p-2f, -0x1.87de2ap-2f, -0x1.58f9a8p-2f, -0x1.294062p-2f, -0x1.f19f98p-3f, -0x1.8f8b84p-3f, -0x1.2c8106p-3f, -0x1.917a6cp-4f, -0x1.91f66p-5f, -0x1.a79394p-53f, 0x1.91f66p-5f, 0x1.917a6cp-4f, 0x1.2c8106p-3f, 0x1.8f8b84p-3f, 0x1.f19f98p-3f, 0x1.294062p-2f, 0x1.58f9a8p-2f, 0x1.87de2ap-2f, 0x1.b5d1p-2f, 0x1.e2b5d4p-2f, 0x1.07387ap-1f, 0x1.1c73b4p-1f, 0x1.30ff8p-1f, 0x1.44cf32p-1f, 0x1.57d694p-1f, 0x1.6a09e6p-1f, 0x1.7b5df2p-1f, 0x1.8bc806p-1f, 0x1.9b3e04p-1f, 0x1.a9b662p-1f, 0x1.b72834p-1f, 0x1.c38b3p-1f, 0x1.ced7bp-1f, 0x1.d906bcp-1f, 0x1.e2121p-1f, 0x1.e9f416p-1f, 0x1.f0a7fp-1f, 0x1.f6297cp-1f, 0x1.fa7558p-1f, 0x1.fd88dap-1f, 0x1.ff621ep-1f};
constexpr float kTw128Im[128] = {-0x0p+0f, -0x1.91f66p-5f, -0x1.917a6cp-4f, -0x1.2c8106p-3f, -0x1.8f8b84p-3f, -0x1.f19f98p-3f, -0x1.294062p-2f, -0x1.58f9a8p-2f, -0x1.87de2ap-2f, -0x1.b5d1p-2f, -0x1.e2b5d4p-2f, -0x1.07387ap-1f, -0x1.1c73b4p-1f, -0x1.30ff8p-1f, -0x1.44cf32p-1f, -0x1.57d694p-1f, -0x1.6a09e6p-1f, -0x1.7b5df2p-1f, -0x1.8bc806p-1f, -0x1.9b3e04p-1f, -0x1.a9b662p-1f, -0x1.b72834p-1f, -0x1.c38b3p-1f, -0x1.ced7bp-1f, -0x1.d906bcp-1f, -0x1.e2121p-1f, -0x1.e9f416p-1f, -0x1.f0a7fp-1f, -0x1.f6297cp-1f, -0x1.fa7558p-1f, -0x1.fd88dap-1f, -0x1.ff621ep-1f, -0x1p+0f, -0x1.ff621ep-1f, -0x1.fd88dap-1f, -0x1.fa7558p-1f, -0x1.f6297cp-1f, -0x1.f0a7fp-1f, -0x1.e9f416p-1f, -0x1.e2121p-1f, -0x1.d906bcp-1f, -0x1.ced7bp-1f, -0x1.c38b3p-1f, -0x1.b72834p-1f, -0x1.a9b662p-1f, -0x1.9b3e04p-1f, -0x1.8bc806p-1f, -0x1.7b5df2p-1f, -0x1.6a09e6p-1f, -0x1.57d694p-1f, -0x1.44cf32p-1f, -0x1.30ff8p-1f, -0x1.1c73b4p-1f, -0x1.07387ap-1f, -0x1.e2b5d4p-2f, -0x1.b5d1p-2f, -0x1.87de2ap-2f, -0x1.58f9a8p-2f, -0x1.294062p-2f, -0x1.f19f98p-3f, -0x1.8f8b84p-3f, -0x1.2c8106p-3f, -0x1.917a6cp-4f, -0x1.91f66p-5f, -0x1.1a6264p-53f, 0x1.91f66p-5f, 0x1.917a6cp-4f, 0x1.2c8106p-3f, 0x1.8f8b84p-3f, 0x1.f19f98p-3f, 0x1.294062p-2f, 0x1.58f9a8p-2f, 0x1.87de2ap-2f, 0x1.b5d1p-2f, 0x1.e2b5d4p-2f, 0x1.07387ap-1f, 0x1.1c73b4p-1f, 0x1.30ff8p-1f, 0x1.44cf32p-1f, 0x1.57d694p-1f, 0x1.6a09e6p-1f, 0x1.7b5df2p-1f, 0x1.8bc806p-1f, 0x1.9b3e04p-1f, 0x1.a9b662p-1f, 0x1.b72834p-1f, 0x1.c38b3p-1f, 0x1.ced7bp-1f, 0x1.d906bcp-1f, 0x1.e2121p-1f, 0x1.e9f416p-1f, 0x1.f0a7fp-1f, 0x1.f6297cp-1f, 0x1.fa7558p-1f, 0x1.fd88dap-1f, 0x1.ff621ep-1f, 0x1p+0f, 0x1.ff621ep-1f, 0x1.fd88dap-1f, 0x1.fa7558p-1f, 0x1.f6297cp-1f, 0x1.f0a7fp-1f, 0x1.e9f416p-1f, 0x1.e2121p-1f, 0x1.d906bcp-1f, 0x1.ced7bp-1f, 0x1.c38b3p-1f, 0x1.b72834p-1f, 0x1.a9b662p-1f, 0x1.9b3e04p-1f, 0x1.8bc806p-1f, 0x1.7b5df2p-1f, 0x1.6a09e6p-1f, 0x1.57d694p-1f, 0x1.44cf32p-1f, 0x1.30ff8p-1f, 0x1.1c73b4p-1f, 0x1.07387ap-1f, 0x1.e2b5d4p-2f, 0x1.b5d1p-2f, 0x1.87de2ap-2f, 0x1.58f9a8p-2f, 0x1.294062p-2f, 0x1.f19f98p-3f, 0x1.8f8b84p-3f, 0x1.2c8106p-3f, 0x1.917a6cp-4f, 0x1.91f66p-5f};
// exp(+2 pi i j / 128), j = 0..127
extern __constant__ float2 c_tw128[128];

__device__ inline float2 cmul_f(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// exp(SIGN 2 pi i k / len) for len dividing 128
template <int SIGN>
__device__ inline float2 tw(int k, int len)
{
    float2 w = make_float2(kTw128Re[k * (128 / len)], kTw128Im[k * (128 / len)]);
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

// ---- 128 points over a lane quad (lanes 4i .. 4i + 3, h = lane & 3), 32
// values per lane.  Forward: lane h holds samples 4m + h; each lane runs a
// 32-point transform, lane pairs (h, h ^ 2) combine them into the 64-point
// transforms of the even and odd samples, and pairs (h, h ^ 1) into the
// 128-point one; lane h leaves with outputs qblk(h) * 32 + k.  Inverse:
// lane h holds inputs qblk(h) * 32 + n, the same two exchanges in the
// opposite order (decimation in frequency), then the 32-point transforms;
// lane h leaves with outputs 4j + h.
// exp(+2 pi i j / 256), j = 0..31 (the polyphase factors at w = 128)
constexpr float kTw256Re[32] = {0x1p+0f, 0x1.ffd886p-1f, 0x1.ff621ep-1f, 0x1.fe9cdap-1f, 0x1.fd88dap-1f, 0x1.fc2648p-1f, 0x1.fa7558p-1f, 0x1.f8765p-1f, 0x1.f6297cp-1f, 0x1.f38f3ap-1f, 0x1.f0a7fp-1f, 0x1.ed740ep-1f, 0x1.e9f416p-1f, 0x1.e6288ep-1f, 0x1.e2121p-1f, 0x1.ddb13cp-1f, 0x1.d906bcp-1f, 0x1.d4134ep-1f, 0x1.ced7bp-1f, 0x1.c954b2p-1f, 0x1.c38b3p-1f, 0x1.bd7c0ap-1f, 0x1.b72834p-1f, 0x1.b090a6p-1f, 0x1.a9b662p-1f, 0x1.a29a7ap-1f, 0x1.9b3e04p-1f, 0x1.93a224p-1f, 0x1.8bc806p-1f, 0x1.83b0ep-1f, 0x1.7b5df2p-1f, 0x1.72d084p-1f};
constexpr float kTw256Im[32] = {0x0p+0f, 0x1.92156p-6f, 0x1.91f66p-5f, 0x1.2d520ap-4f, 0x1.917a6cp-4f, 0x1.f564e6p-4f, 0x1.2c8106p-3f, 0x1.5e2144p-3f, 0x1.8f8b84p-3f, 0x1.c0b826p-3f, 0x1.f19f98p-3f, 0x1.111d26p-2f, 0x1.294062p-2f, 0x1.4135cap-2f, 0x1.58f9a8p-2f, 0x1.708854p-2f, 0x1.87de2ap-2f, 0x1.9ef794p-2f, 0x1.b5d1p-2f, 0x1.cc66eap-2f, 0x1.e2b5d4p-2f, 0x1.f8ba4ep-2f, 0x1.07387ap-1f, 0x1.11eb36p-1f, 0x1.1c73b4p-1f, 0x1.26d054p-1f, 0x1.30ff8p-1f, 0x1.3affa2p-1f, 0x1.44cf32p-1f, 0x1.4e6cacp-1f, 0x1.57d694p-1f, 0x1.610b76p-1f};

__device__ inline int qblk(int h) { return ((h & 1) << 1) | (h >> 1); }

// Value of lane ^ 2 via DPP quad_perm [2,3,0,1].
__device__ inline float dpp_swap2(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));
}

template <int XOR>
__device__ inline float2 shfl_quad(float2 v)
{
    if constexpr (XOR == 1)
        return make_float2(dpp_swap1(v.x), dpp_swap1(v.y));
    else
        return make_float2(dpp_swap2(v.x), dpp_swap2(v.y));
}

// exp(SIGN 2 pi i k / len) with the exact values at k = 0 and k = len / 4.
template <int SIGN>
__device__ inline float2 twx(int k, int len)
{
    if (k == 0)
        return make_float2(1.f, 0.f);
    if (4 * k == len)
        return make_float2(0.f, SIGN < 0 ? -1.f : 1.f);
    return tw<SIGN>(k, len);
}

// One radix-2 combine between lane h and lane h ^ XOR: role 0 holds the
// even-sample transform E, role 1 the odd-sample one O, with twiddle
// exp(SIGN 2 pi i (k + off) / len) where off = len / 4 on lanes with `rot`
// set (applied as an exact multiplication by SIGN i, so every twiddle stays
// a compile-time operand instead of a per-lane register).  Role 0 leaves
// with E + w O, role 1 with E - w O.
template <int XOR, int SIGN, int LEN>
__device__ inline void quad_fwd_stage(float2 (&x)[32], int role, int rot)
{
#pragma unroll
    for (int k = 0; k < 32; k++) {
        float2 m = cmul_f(x[k], twx<SIGN>(k, LEN));
        if (rot)  // times exp(SIGN i pi / 2) = SIGN i
            m = SIGN < 0 ? make_float2(m.y, -m.x) : make_float2(-m.y, m.x);
        const float2 t = role ? m : x[k];
        const float2 y = shfl_quad<XOR>(t);
        x[k] = role ? make_float2(y.x - t.x, y.y - t.y) : make_float2(y.x + t.x, y.y + t.y);
    }
}

// Inverse combine: role 0 holds the first half a, role 1 the second half
// b (same index n); role 0 leaves with a + b, role 1 with (a - b) w (w as
// in quad_fwd_stage).
template <int XOR, int SIGN, int LEN>
__device__ inline void quad_inv_stage(float2 (&x)[32], int role, int rot)
{
#pragma unroll
    for (int n = 0; n < 32; n++) {
        const float2 y = shfl_quad<XOR>(x[n]);
        const float2 a = make_float2(x[n].x + y.x, x[n].y + y.y);
        const float2 d = make_float2(y.x - x[n].x, y.y - x[n].y);
        float2 m = cmul_f(d, twx<SIGN>(n, LEN));
        if (rot)
            m = SIGN < 0 ? make_float2(m.y, -m.x) : make_float2(-m.y, m.x);
        x[n] = role ? m : a;
    }
}

__device__ inline void fwd_quad128(float2 (&x)[32], int h)
{
    fft<32, -1>(x);
    // (h, h ^ 2): 64-point transforms of the even (h = 0, 2) and odd (1, 3) samples
    quad_fwd_stage<2, -1, 64>(x, (h >> 1) & 1, 0);
    // (h, h ^ 1): lanes 0 / 1 hold outputs 0..31 of E / O, lanes 2 / 3
    // outputs 32..63 (twiddle exp(-2 pi i (32 + k) / 128))
    quad_fwd_stage<1, -1, 128>(x, h & 1, (h >> 1) & 1);
}

__device__ inline void inv_quad128(float2 (&x)[32], int h)
{
    // (h, h ^ 1): inputs n and n + 64 (blocks 0 / 2 on lanes 0 / 1, 1 / 3 on 2 / 3)
    quad_inv_stage<1, +1, 128>(x, h & 1, (h >> 1) & 1);
    // (h, h ^ 2): within each half, inputs n and n + 32
    quad_inv_stage<2, +1, 64>(x, (h >> 1) & 1, 0);
    fft<32, +1>(x);
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
