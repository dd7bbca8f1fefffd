// Batched mixed-radix Stockham FFT in LDS (see st_fft.h).
#include <math.h>

#include <vector>

#include "st_fft.h"

namespace stfft {

namespace {

constexpr int kPointsPerBlock = 1024;  // LDS: 2 x 1024 x 8 B per block
constexpr int kThreads = 256;

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
__device__ inline void stage_p(const float2 *A, float2 *B, int n, int ns, const float2 *tw, int sign, int t)
{
    const int stride = n / P;
    const int row = t / stride, j = t - row * stride;
    const float2 *src = A + row * n;
    float2 *dst = B + row * n;
    const int k = j % ns;
    float2 v[P];
    v[0] = src[j];
#pragma unroll
    for (int r = 1; r < P; r++)
        v[r] = cmul(src[j + r * stride], tw[k * P + r]);
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
    const int base = (j / ns) * ns * P + k;
#pragma unroll
    for (int q = 0; q < P; q++)
        dst[base + q * ns] = v[q];
}

__global__ void __launch_bounds__(kThreads) k_fft_rows(float2 *data, long nrows, PlanArgs P, int sign)
{
    extern __shared__ float2 lds[];
    const int n = P.n, R = P.rowsPerBlock;
    const long row0 = (long)blockIdx.x * R;
    const int nrow = (int)((nrows - row0) < R ? (nrows - row0) : R);
    float2 *A = lds, *B = lds + R * n;
    const long base = row0 * n;
    for (int t = threadIdx.x; t < nrow * n; t += blockDim.x)
        A[t] = data[base + t];
    __syncthreads();
    for (int s = 0; s < P.nst; s++) {
        const int p = P.radix[s], ns = P.ns[s];
        const float2 *tw = P.tw + P.twoff[s];
        const int work = nrow * (n / p);
        for (int t = threadIdx.x; t < work; t += blockDim.x) {
            switch (p) {
            case 8: stage_p<8>(A, B, n, ns, tw, sign, t); break;
            case 4: stage_p<4>(A, B, n, ns, tw, sign, t); break;
            case 2: stage_p<2>(A, B, n, ns, tw, sign, t); break;
            case 3: stage_p<3>(A, B, n, ns, tw, sign, t); break;
            case 5: stage_p<5>(A, B, n, ns, tw, sign, t); break;
            default: stage_p<7>(A, B, n, ns, tw, sign, t); break;
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
    while (m > 1) {
        int r = (m % 8 == 0) ? 8 : (m % 4 == 0) ? 4 : (m % 2 == 0) ? 2 : (m % 3 == 0) ? 3 : (m % 5 == 0) ? 5 : 7;
        ST_REQUIRE(k < kMaxStages, ST_ERR_UNSUPPORTED, "FFT plan too deep");
        P.radix[k] = r;
        P.ns[k] = ns;
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

void Engine::rows(float2 *data, long nrows, int n, int sign, hipStream_t st)
{
    if (nrows <= 0 || n <= 1)
        return;
    const PlanArgs &P = plan(n, sign);
    long blocks = (nrows + P.rowsPerBlock - 1) / P.rowsPerBlock;
    size_t lds = 2 * (size_t)P.rowsPerBlock * n * sizeof(float2);
    hipLaunchKernelGGL(k_fft_rows, dim3((unsigned)blocks), dim3(kThreads), lds, st, data, nrows, P, sign);
    ST_HIP(hipGetLastError());
}

void Engine::transpose(const float2 *in, float2 *out, int batch, int rows, int cols, hipStream_t st)
{
    dim3 grid((cols + 31) / 32, (rows + 31) / 32, batch);
    hipLaunchKernelGGL(k_transpose, grid, dim3(32, 8), 0, st, in, out, rows, cols);
    ST_HIP(hipGetLastError());
}

void Engine::fft2d(float2 *data, float2 *tmp, int batch, int na, int nr, int sign, hipStream_t st)
{
    rows(data, (long)batch * na, nr, sign, st);
    transpose(data, tmp, batch, na, nr, st);
    rows(tmp, (long)batch * nr, na, sign, st);
    transpose(tmp, data, batch, nr, na, st);
}

}  // namespace stfft
