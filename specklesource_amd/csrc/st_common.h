// Shared host/device helpers of the MI355X speckle-tracking library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <string>

#include "../../include/speckle_amd.h"

#define ST_LARGEINT 2000000000.0

// Error propagation inside the library: every HIP call is checked and the
// first failure is turned into an ST_ERR_HIP return with a message.
struct StError {
    int code;
    std::string msg;
};

#define ST_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            throw StError{ST_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)}; \
    } while (0)

#define ST_REQUIRE(cond, code, text)                  \
    do {                                              \
        if (!(cond))                                  \
            throw StError{(code), std::string(text)}; \
    } while (0)

__host__ __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b) with the reference's operand order (speckleTrack.c:631-632):
// re = ar*br + ai*bi, im = ai*br - ar*bi
__host__ __device__ inline float2 cmulconj(float2 a, float2 b)
{
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

__host__ __device__ inline int imin(int a, int b) { return a < b ? a : b; }
__host__ __device__ inline int imax(int a, int b) { return a > b ? a : b; }

// Block-wide reductions for 256-thread (4 x wave64) blocks.
template <int NT>
__device__ inline double block_sum_d(double v, double *sh)
{
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0)
        sh[wid] = v;
    __syncthreads();
    double r = 0;
    for (int w = 0; w < NT / 64; w++)
        r += sh[w];
    __syncthreads();
    return r;
}

template <int NT>
__device__ inline int block_sum_i(int v, int *sh)
{
    for (int o = 32; o > 0; o >>= 1)
        v += __shfl_xor(v, o);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0)
        sh[wid] = v;
    __syncthreads();
    int r = 0;
    for (int w = 0; w < NT / 64; w++)
        r += sh[w];
    __syncthreads();
    return r;
}

// Argmax with tie-break on the smallest rank: (value, rank) pairs, larger
// value wins, equal values keep the smaller rank.  Values that are not
// > 0 never win (the reference scans with a strict '>' from 0).
struct ArgMax {
    float v;
    int rank;
};

__device__ inline ArgMax argmax_combine(ArgMax a, ArgMax b)
{
    if (b.v > a.v || (b.v == a.v && b.rank < a.rank))
        return b;
    return a;
}

template <int NT>
__device__ inline ArgMax block_argmax(ArgMax a, ArgMax *sh)
{
    for (int o = 32; o > 0; o >>= 1) {
        ArgMax b;
        b.v = __shfl_xor(a.v, o);
        b.rank = __shfl_xor(a.rank, o);
        a = argmax_combine(a, b);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0)
        sh[wid] = a;
    __syncthreads();
    ArgMax r = sh[0];
    for (int w = 1; w < NT / 64; w++)
        r = argmax_combine(r, sh[w]);
    __syncthreads();
    return r;
}

struct ArgMaxD {
    double v;
    int rank;
};

__device__ inline ArgMaxD argmaxd_combine(ArgMaxD a, ArgMaxD b)
{
    if (b.v > a.v || (b.v == a.v && b.rank < a.rank))
        return b;
    return a;
}

template <int NT>
__device__ inline ArgMaxD block_argmax_d(ArgMaxD a, ArgMaxD *sh)
{
    for (int o = 32; o > 0; o >>= 1) {
        ArgMaxD b;
        b.v = __shfl_xor(a.v, o);
        b.rank = __shfl_xor(a.rank, o);
        a = argmaxd_combine(a, b);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0)
        sh[wid] = a;
    __syncthreads();
    ArgMaxD r = sh[0];
    for (int w = 1; w < NT / 64; w++)
        r = argmaxd_combine(r, sh[w]);
    __syncthreads();
    return r;
}

template <int NT>
__device__ inline float block_max_f(float v, float *sh)
{
    for (int o = 32; o > 0; o >>= 1)
        v = fmaxf(v, __shfl_xor(v, o));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0)
        sh[wid] = v;
    __syncthreads();
    float r = sh[0];
    for (int w = 1; w < NT / 64; w++)
        r = fmaxf(r, sh[w]);
    __syncthreads();
    return r;
}

template <int NT>
__device__ inline double block_max_d(double v, double *sh)
{
    for (int o = 32; o > 0; o >>= 1)
        v = fmax(v, __shfl_xor(v, o));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0)
        sh[wid] = v;
    __syncthreads();
    double r = sh[0];
    for (int w = 1; w < NT / 64; w++)
        r = fmax(r, sh[w]);
    __syncthreads();
    return r;
}

// Device scratch arena: one growing allocation per context, carved into
// aligned sub-buffers per call.
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    void reset() { used = 0; }
    template <class T>
    T *take(size_t n)
    {
        size_t bytes = (n * sizeof(T) + 255) & ~(size_t)255;
        ST_REQUIRE(used + bytes <= cap, ST_ERR_NOMEM, "scratch arena exhausted");
        T *p = (T *)(base + used);
        used += bytes;
        return p;
    }
};
