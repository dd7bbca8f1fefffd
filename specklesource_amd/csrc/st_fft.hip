// Batched mixed-radix Stockham FFT in LDS (see st_fft.h).
#include <math.h>

#include <vector>

#include "st_fftlds.h"

namespace stfft {

namespace {

constexpr int kPointsPerBlock = 1024;  // LDS: 2 x 1024 x 8 B per block (indices stay below 2^13)
constexpr int kThreads = 256;
constexpr int kColElems = 4096;  // per block: LDS kColElems x 8 B (+ one pad element per column)

template <int P, int sign>
__device__ inline void stage_p(const float2 *A, float2 *B, int n, int ns, FastDiv ds, FastDiv dn, const float2 *tw,
                               int t)
{
    const int stride = n / P;
    const int row = fdiv(t, ds), j = t - row * stride;
    const float2 *src = A + row * n;
    float2 *dst = B + row * n;
    const int jq = fdiv(j, dn), k = j - jq * ns;
    float2 v[P];
    v[0] = src[j];
#pragma unroll
    for (int r = 1; r < P; r++)
        v[r] = cmul(src[j + r * stride], tw[k * P + r]);
    dft_any<P>(v, sign);
    const int base = jq * ns * P + k;
#pragma unroll
    for (int q = 0; q < P; q++)
        dst[base + q * ns] = v[q];
}

template <int SIGN>
__global__ void __launch_bounds__(kThreads) k_fft_rows(const float2 *src, float2 *data, long nrows, PlanArgs P)
{
    extern __shared__ float2 lds[];
    const int n = P.n, R = P.rowsPerBlock;
    const long row0 = (long)blockIdx.x * R;
    const int nrow = (int)((nrows - row0) < R ? (nrows - row0) : R);
    float2 *A = lds, *B = lds + R * n;
    const long base = row0 * n;
    // per chunk of kPointsPerBlock points, all of a thread's loads are
    // issued before its first LDS store
    constexpr int IT = kPointsPerBlock / kThreads;
    const int cnt = nrow * n;
    for (int t0 = 0; t0 < cnt; t0 += kPointsPerBlock) {
        float2 u[IT];
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = t0 + threadIdx.x + it * kThreads;
            u[it] = t < cnt ? src[base + t] : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int it = 0; it < IT; it++) {
            const int t = t0 + threadIdx.x + it * kThreads;
            if (t < cnt)
                A[t] = u[it];
        }
    }
    __syncthreads();
    for (int s = 0; s < P.nst; s++) {
        const int p = P.radix[s], ns = P.ns[s];
        const FastDiv ds = P.dstride[s], dn = P.dns[s];
        const float2 *tw = P.tw + P.twoff[s];
        const int work = nrow * (n / p);
        for (int t = threadIdx.x; t < work; t += blockDim.x) {
            switch (p) {
            case 8: stage_p<8, SIGN>(A, B, n, ns, ds, dn, tw, t); break;
            case 4: stage_p<4, SIGN>(A, B, n, ns, ds, dn, tw, t); break;
            case 2: stage_p<2, SIGN>(A, B, n, ns, ds, dn, tw, t); break;
            case 3: stage_p<3, SIGN>(A, B, n, ns, ds, dn, tw, t); break;
            case 5: stage_p<5, SIGN>(A, B, n, ns, ds, dn, tw, t); break;
            default: stage_p<7, SIGN>(A, B, n, ns, ds, dn, tw, t); break;
            }
        }
        __syncthreads();
        float2 *T = A;
        A = B;
        B = T;
    }
    for (int t = threadIdx.x; t < nrow * n; t += blockDim.x)
        data[base + t] = A[t];
}

__global__ void k_transpose(const float2 *in, float2 *out, int rows, int cols)
{
    __shared__ float2 tile[32][33];
    const long b = blockIdx.z;
    const float2 *src = in + b * (long)rows * cols;
    float2 *dst = out + b * (long)rows * cols;
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    for (int dy = threadIdx.y; dy < 32; dy += 8) {
        int r = r0 + dy, c = c0 + threadIdx.x;
        if (r < rows && c < cols)
            tile[dy][threadIdx.x] = src[(long)r * cols + c];
    }
    __syncthreads();
    for (int dy = threadIdx.y; dy < 32; dy += 8) {
        int c = c0 + dy, r = r0 + threadIdx.x;
        if (r < rows && c < cols)
            dst[(long)c * rows + r] = tile[threadIdx.x][dy];
    }
}

// Column transforms of a batch of na x nr row-major arrays: each block
// loads C adjacent columns (rows of C * 8 contiguous bytes), transforms them
// in LDS as C sequences of length na, and writes them back in place.  All
// of a thread's loads are issued before the first LDS store, and blocks are
// numbered so that neighbouring column blocks (which share cache lines) run
// on the same XCD: dispatch deals consecutive workgroups round-robin over
// the 8 XCDs.
__global__ void __launch_bounds__(kThreads) k_fft_cols(float2 *data, int na, int nr, int lc, PlanArgs P, int sign)
{
    extern __shared__ float2 lds[];
    constexpr int IT = kColElems / kThreads;
    const int C = 1 << lc;
    const int nbx = gridDim.x, total = nbx * gridDim.y, flat = blockIdx.y * nbx + blockIdx.x;
    const int per = total / 8;
    const int L = flat < per * 8 ? (flat % 8) * per + flat / 8 : flat;
    const int z = L / nbx, cb = L - (L / nbx) * nbx;
    const int pitch = na + 1;
    const int c0 = cb * C;
    const int nc = nr - c0 < C ? nr - c0 : C;
    const int cnt = na * C;
    float2 *base = data + (size_t)z * na * nr + c0;
    float2 v[IT];
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int t = threadIdx.x + it * kThreads, r = t >> lc, c = t & (C - 1);
        v[it] = t < cnt && c < nc ? base[(size_t)r * nr + c] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int t = threadIdx.x + it * kThreads, r = t >> lc, c = t & (C - 1);
        if (t < cnt)
            lds[c * pitch + r] = v[it];
    }
    __syncthreads();
    lds_fft<kColElems, kThreads>(lds, pitch, C, P, sign);
#pragma unroll
    for (int it = 0; it < IT; it++) {
        const int t = threadIdx.x + it * kThreads, r = t >> lc, c = t & (C - 1);
        if (t < cnt && c < nc)
            base[(size_t)r * nr + c] = lds[c * pitch + r];
    }
}

}  // namespace

bool supported_size(int n)
{
    if (n < 1 || n > 4096)
        return false;
    int m = n;
    const int rs[] = {2, 3, 5, 7};
    for (int r : rs)
        while (m % r == 0)
            m /= r;
    return m == 1;
}

Engine::~Engine()
{
    for (auto &kv : plans_)
        hipFree((void *)kv.second.tw);
}

const PlanArgs &Engine::plan(int n, int sign)
{
    auto key = std::make_pair(n, sign);
    auto it = plans_.find(key);
    if (it != plans_.end())
        return it->second;
    ST_REQUIRE(supported_size(n), ST_ERR_UNSUPPORTED, "FFT size " + std::to_string(n) + " not supported");
    PlanArgs P{};
    P.n = n;
    int m = n, ns = 1, k = 0;
    std::vector<float2> tw;
    int twos = 0;
    for (int t = n; t % 2 == 0; t /= 2)
        twos++;
    while (m > 1) {
        // powers of two as radix 8, with 4 x 4 in place of 8 x 2 when the
        // count of twos leaves one over (a radix-2 stage only for n = 2 * odd)
        int r;
        if (twos >= 3 && twos != 4)
            r = 8;
        else if (twos >= 2)
            r = 4;
        else if (twos == 1)
            r = 2;
        else
            r = (m % 3 == 0) ? 3 : (m % 5 == 0) ? 5 : 7;
        twos -= r == 8 ? 3 : r == 4 ? 2 : r == 2 ? 1 : 0;
        ST_REQUIRE(k < kMaxStages, ST_ERR_UNSUPPORTED, "FFT plan too deep");
        P.radix[k] = r;
        P.ns[k] = ns;
        P.dstride[k] = make_fastdiv((uint32_t)(n / r));
        P.dns[k] = make_fastdiv((uint32_t)ns);
        P.twoff[k] = (int)tw.size();
        for (int j = 0; j < ns; j++)
            for (int q = 0; q < r; q++) {
                double ang = sign * 2.0 * M_PI * (double)(j * q) / (double)(ns * r);
                tw.push_back(make_float2((float)cos(ang), (float)sin(ang)));
            }
        ns *= r;
        m /= r;
        k++;
    }
    P.nst = k;
    P.rowsPerBlock = n >= kPointsPerBlock ? 1 : kPointsPerBlock / n;
    if (tw.empty())
        tw.push_back(make_float2(1.f, 0.f));
    float2 *d = nullptr;
    ST_HIP(hipMalloc(&d, tw.size() * sizeof(float2)));
    ST_HIP(hipMemcpy(d, tw.data(), tw.size() * sizeof(float2), hipMemcpyHostToDevice));
    P.tw = d;
    return plans_.emplace(key, P).first->second;
}

void Engine::rows(float2 *data, long nrows, int n, int sign, hipStream_t st) { rows(data, data, nrows, n, sign, st); }

void Engine::rows(const float2 *src, float2 *data, long nrows, int n, int sign, hipStream_t st)
{
    if (nrows <= 0 || n <= 1)
        return;
    const PlanArgs &P = plan(n, sign);
    long blocks = (nrows + P.rowsPerBlock - 1) / P.rowsPerBlock;
    size_t lds = 2 * (size_t)P.rowsPerBlock * n * sizeof(float2);
    if (sign < 0)
        hipLaunchKernelGGL(k_fft_rows<-1>, dim3((unsigned)blocks), dim3(kThreads), lds, st, src, data, nrows, P);
    else
        hipLaunchKernelGGL(k_fft_rows<+1>, dim3((unsigned)blocks), dim3(kThreads), lds, st, src, data, nrows, P);
    ST_HIP(hipGetLastError());
}

void Engine::transpose(const float2 *in, float2 *out, int batch, int rows, int cols, hipStream_t st)
{
    dim3 grid((cols + 31) / 32, (rows + 31) / 32, batch);
    hipLaunchKernelGGL(k_transpose, grid, dim3(32, 8), 0, st, in, out, rows, cols);
    ST_HIP(hipGetLastError());
}

int col_block_log2(int na)
{
    int lc = 0;
    while ((2 << lc) * na <= kColElems && (2 << lc) <= 32)
        lc++;
    return lc;
}

void Engine::cols(float2 *data, int batch, int na, int nr, int sign, hipStream_t st)
{
    if (batch <= 0 || na <= 1)
        return;
    const PlanArgs &P = plan(na, sign);
    ST_REQUIRE(na <= kColElems, ST_ERR_UNSUPPORTED, "column FFT length above the LDS block");
    const int lc = col_block_log2(na), C = 1 << lc;
    const size_t lds = (size_t)C * (na + 1) * sizeof(float2);
    hipLaunchKernelGGL(k_fft_cols, dim3((unsigned)((nr + C - 1) / C), (unsigned)batch), dim3(kThreads), lds, st, data,
                       na, nr, lc, P, sign);
    ST_HIP(hipGetLastError());
}

void Engine::fft2d(float2 *data, float2 *tmp, int batch, int na, int nr, int sign, hipStream_t st)
{
    (void)tmp;
    rows(data, (long)batch * na, nr, sign, st);
    cols(data, batch, na, nr, sign, st);
}

void Engine::fft2d_to(const float2 *in, float2 *out, int batch, int na, int nr, int sign, hipStream_t st)
{
    rows(in, out, (long)batch * na, nr, sign, st);
    cols(out, batch, na, nr, sign, st);
}

}  // namespace stfft
