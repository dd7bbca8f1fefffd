// Batched FP32 complex FFT engine for CDNA4 (no rocFFT).
//
// Rows of length n (n = 2^a 3^b 5^c 7^d <= 4096) are transformed in LDS by
// a mixed-radix Stockham autosort (radices 8, 4, 2, 3, 5, 7), several rows
// per workgroup.  2-D transforms are a row pass plus a column pass that
// transforms blocks of adjacent columns in LDS (no transposes).
// Conventions follow FFTW 2, which the reference links: sign -1 is
// exp(-2 pi i jk/n) ("forward"), +1 is exp(+2 pi i jk/n) ("backward"),
// no normalisation.
#pragma once

#include <map>
#include <utility>

#include "st_common.h"

namespace stfft {

constexpr int kMaxStages = 14;

// Exact q = n / d for 0 <= n < 2^13 and 1 <= d <= 4096 by one multiply and
// a shift: s = 13 + ceil(log2 d), m = ceil(2^s / d).  With e = m d - 2^s < d,
// n m / 2^s = n / d + n e / (d 2^s) and n e < 2^13 d <= 2^s, so the floor is
// exact; n m < 2^13 (2^14 + 1) fits 32 bits.
struct FastDiv {
    uint32_t m, s;
};

inline FastDiv make_fastdiv(uint32_t d)
{
    uint32_t l = 0;
    while ((1u << l) < d)
        l++;
    const uint32_t s = 13 + l;
    return FastDiv{(uint32_t)(((1ull << s) + d - 1) / d), s};
}

__device__ inline int fdiv(int n, FastDiv f) { return (int)(((uint32_t)n * f.m) >> f.s); }

struct PlanArgs {
    int n, nst, rowsPerBlock;
    int radix[kMaxStages], ns[kMaxStages], twoff[kMaxStages];
    FastDiv dstride[kMaxStages], dns[kMaxStages];  // n / radix and ns of each stage
    const float2 *tw;
};

class Engine {
  public:
    ~Engine();
    const PlanArgs &plan(int n, int sign);
    // In-place batched 1-D FFT of nrows contiguous rows of length n.
    void rows(float2 *data, long nrows, int n, int sign, hipStream_t st);
    // The same, reading src and writing data (src == data is in place).
    void rows(const float2 *src, float2 *data, long nrows, int n, int sign, hipStream_t st);
    // out[b][c][r] = in[b][r][c] for batch b of rows x cols arrays.
    static void transpose(const float2 *in, float2 *out, int batch, int rows, int cols, hipStream_t st);
    // In-place column transforms (length na) of batch arrays na x nr.
    void cols(float2 *data, int batch, int na, int nr, int sign, hipStream_t st);
    // In-place batched 2-D FFT of batch arrays na x nr (row-major); tmp is
    // not used (kept so callers need not change).
    void fft2d(float2 *data, float2 *tmp, int batch, int na, int nr, int sign, hipStream_t st);
    // Out of place: in is left untouched.
    void fft2d_to(const float2 *in, float2 *out, int batch, int na, int nr, int sign, hipStream_t st);

  private:
    std::map<std::pair<int, int>, PlanArgs> plans_;
};

bool supported_size(int n);

}  // namespace stfft
