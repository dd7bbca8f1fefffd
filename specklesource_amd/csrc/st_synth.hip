// Device path of the synthetic SLC generator (st_synth_device): the same
// per-sample formulas as st_synth_host (synth_common.h), evaluated by
// kernels so that full-size scenes (20000 x 14000) are generated in HBM
// without a host round trip.  Intermediate speckle is kept in float.
#include "st_ctx.h"
#include "synth_common.h"

using namespace synth;

namespace {

__global__ void k_lattice(st_synth_spec s, Lattice L, double *N)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)L.nx * L.ny)
        return;
    N[t] = lattice_noise(s, L, (int)(t / L.nx), (int)(t % L.nx));
}

__global__ void k_lattice_rows(st_synth_spec s, Lattice L, const double *N, double *H, double *hn)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)L.ny * s.nr)
        return;
    const int ky = (int)(t / s.nr), x = (int)(t % s.nr);
    const int c = x / kLatStep + L.pad;
    double acc = 0, nrm = 0;
    for (int kx = c - L.reach; kx <= c + L.reach; kx++) {
        if (kx < 0 || kx >= L.nx)
            continue;
        double w = lat_weight(s, x, kx, L);
        acc += w * N[(size_t)ky * L.nx + kx];
        nrm += w * w;
    }
    H[t] = acc;
    if (ky == 0)
        hn[x] = nrm;
}

__global__ void k_backscatter(st_synth_spec s, Lattice L, const double *H, const double *hn, float *g)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)s.na * s.nr)
        return;
    const int y = (int)(t / s.nr), x = (int)(t % s.nr);
    const int c = y / kLatStep + L.pad;
    double acc = 0, nrm = 0;
    for (int ky = c - L.reach; ky <= c + L.reach; ky++)
        if (ky >= 0 && ky < L.ny) {
            double w = lat_weight(s, y, ky, L);
            acc += w * H[(size_t)ky * s.nr + x];
            nrm += w * w;
        }
    g[t] = (float)(acc / sqrt(hn[x] * nrm));
}

struct Fir {
    double hr[2 * kBandHalf + 1], hri[2 * kBandHalf + 1], ha[2 * kBandHalf + 1], hai[2 * kBandHalf + 1];
};

// range pass of the band-limiting FIR over white noise (extended rows)
__global__ void k_speckle_range(st_synth_spec s, Fir f, uint64_t stream, float2 *T)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int nae = s.na + 2 * kBandHalf;
    if (t >= (long)nae * s.nr)
        return;
    const int ae = (int)(t / s.nr), r = (int)(t % s.nr);
    const int a = ae - kBandHalf;
    double re = 0, im = 0;
    for (int k = -kBandHalf; k <= kBandHalf; k++) {
        double xr, xi;
        white(s, stream, a, r - k, &xr, &xi);
        double cr = f.hr[k + kBandHalf], ci = f.hri[k + kBandHalf];
        re += cr * xr - ci * xi;
        im += cr * xi + ci * xr;
    }
    T[t] = make_float2((float)re, (float)im);
}

__global__ void k_speckle_azimuth(st_synth_spec s, Fir f, const float2 *T, float2 *out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)s.na * s.nr)
        return;
    const int a = (int)(t / s.nr), r = (int)(t % s.nr);
    double re = 0, im = 0;
    for (int k = -kBandHalf; k <= kBandHalf; k++) {
        const float2 x = T[(size_t)(a - k + kBandHalf) * s.nr + r];
        double cr = f.ha[k + kBandHalf], ci = f.hai[k + kBandHalf];
        re += cr * x.x - ci * x.y;
        im += cr * x.y + ci * x.x;
    }
    out[t] = make_float2((float)re, (float)im);
}

__global__ void k_img1(st_synth_spec s, const float *g, const float2 *S1, float2 *img1)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)s.na * s.nr)
        return;
    const int a = (int)(t / s.nr), r = (int)(t % s.nr);
    double amp = masked(s, r, a) ? 0.0 : sqrt(exp(s.sigmaLn * (double)g[t]));
    img1[t] = make_float2((float)(amp * S1[t].x), (float)(amp * S1[t].y));
}

__global__ void k_img2(st_synth_spec s, const float *g, const float2 *img1, const float2 *S2, float2 *img2)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)s.na * s.nr)
        return;
    const int a = (int)(t / s.nr), r = (int)(t % s.nr);
    if (masked(s, r, a)) {
        img2[t] = make_float2(0.f, 0.f);
        return;
    }
    double dr, da;
    offset_field(s, r, a, &dr, &da);
    const double pr = r - dr, pa = a - da;
    const int r0 = (int)floor(pr), a0 = (int)floor(pa);
    double wr[kTaps], wa[kTaps];
    resample_weights(pr - r0, wr);
    resample_weights(pa - a0, wa);
    double sr = 0, si = 0;
    for (int ta = 0; ta < kTaps; ta++) {
        const int aa = a0 - 7 + ta;
        if (aa < 0 || aa >= s.na)
            continue;
        double rr = 0, ri = 0;
        for (int tr = 0; tr < kTaps; tr++) {
            const int rc = r0 - 7 + tr;
            if (rc < 0 || rc >= s.nr)
                continue;
            const float2 v = img1[(size_t)aa * s.nr + rc];
            rr += wr[tr] * v.x;
            ri += wr[tr] * v.y;
        }
        sr += wa[ta] * rr;
        si += wa[ta] * ri;
    }
    const double gam = coherence(s, r, a);
    int rn = (int)floor(pr + 0.5), an = (int)floor(pa + 0.5);
    rn = rn < 0 ? 0 : (rn >= s.nr ? s.nr - 1 : rn);
    an = an < 0 ? 0 : (an >= s.na ? s.na - 1 : an);
    const double amp = sqrt(exp(s.sigmaLn * (double)g[(size_t)an * s.nr + rn]));
    const double f = sqrt(1.0 - gam * gam) * amp;
    img2[t] = make_float2((float)(gam * sr + f * S2[t].x), (float)(gam * si + f * S2[t].y));
}

inline unsigned nblk(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

extern "C" int st_synth_device(st_ctx *c, const st_synth_spec *spec, int32_t *pair)
{
    return st_guard(c, [&]() {
        ST_REQUIRE(spec && pair && spec->nr > 0 && spec->na > 0 && spec->corrLen > 0, ST_ERR_ARG, "bad synth spec");
        const st_synth_spec s = *spec;
        hipStream_t st = c->stream;
        const long npx = (long)s.na * s.nr;
        Lattice L = lattice(s);
        double *N = nullptr, *H = nullptr, *hn = nullptr;
        float *g = nullptr;
        float2 *T = nullptr, *S1 = nullptr, *S2 = nullptr, *img1 = nullptr, *img2 = nullptr;
        ST_HIP(hipMalloc(&N, sizeof(double) * L.nx * L.ny));
        ST_HIP(hipMalloc(&H, sizeof(double) * (size_t)L.ny * s.nr));
        ST_HIP(hipMalloc(&hn, sizeof(double) * s.nr));
        ST_HIP(hipMalloc(&g, sizeof(float) * npx));
        ST_HIP(hipMalloc(&T, sizeof(float2) * (size_t)(s.na + 2 * kBandHalf) * s.nr));
        ST_HIP(hipMalloc(&S1, sizeof(float2) * npx));
        ST_HIP(hipMalloc(&S2, sizeof(float2) * npx));
        ST_HIP(hipMalloc(&img1, sizeof(float2) * npx));
        ST_HIP(hipMalloc(&img2, sizeof(float2) * npx));
        hipLaunchKernelGGL(k_lattice, dim3(nblk((long)L.nx * L.ny)), dim3(256), 0, st, s, L, N);
        hipLaunchKernelGGL(k_lattice_rows, dim3(nblk((long)L.ny * s.nr)), dim3(256), 0, st, s, L, N, H, hn);
        hipLaunchKernelGGL(k_backscatter, dim3(nblk(npx)), dim3(256), 0, st, s, L, H, hn, g);
        Fir f;
        band_fir(s.bwR, 0.0, f.hr, f.hri);
        band_fir(s.bwA, s.fdc, f.ha, f.hai);
        const long next = (long)(s.na + 2 * kBandHalf) * s.nr;
        hipLaunchKernelGGL(k_speckle_range, dim3(nblk(next)), dim3(256), 0, st, s, f, (uint64_t)1, T);
        hipLaunchKernelGGL(k_speckle_azimuth, dim3(nblk(npx)), dim3(256), 0, st, s, f, T, S1);
        hipLaunchKernelGGL(k_speckle_range, dim3(nblk(next)), dim3(256), 0, st, s, f, (uint64_t)2, T);
        hipLaunchKernelGGL(k_speckle_azimuth, dim3(nblk(npx)), dim3(256), 0, st, s, f, T, S2);
        hipLaunchKernelGGL(k_img1, dim3(nblk(npx)), dim3(256), 0, st, s, g, S1, img1);
        hipLaunchKernelGGL(k_img2, dim3(nblk(npx)), dim3(256), 0, st, s, g, img1, S2, img2);
        ST_HIP(hipGetLastError());
        ST_HIP(hipStreamSynchronize(st));
        hipFree(N);
        hipFree(H);
        hipFree(hn);
        hipFree(g);
        hipFree(T);
        hipFree(S1);
        hipFree(S2);
        StPair p;
        p.d1 = img1;
        p.d2 = img2;
        p.nr1 = p.nr2 = s.nr;
        p.na1 = p.na2 = s.na;
        p.stype = ST_SAMPLE_C64;
        int id = c->nextPair++;
        c->pairs[id] = p;
        *pair = id;
    });
}
