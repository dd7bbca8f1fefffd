// Per-sample math of the synthetic SLC generator, shared by the host loop
// (st_synth_host) and the device kernels (st_synth_device).  Everything is
// double precision; randomness comes from a counter-based hash so any
// sample can be generated independently.
#pragma once

#include <stdint.h>
#include <math.h>

#include "../../include/speckle_synth.h"

#if defined(__HIPCC__)
#define SYN_HD __host__ __device__
#else
#define SYN_HD
#endif

namespace synth {

constexpr double kPi = 3.14159265358979323846;
constexpr int kLatStep = 16;     // lattice spacing of the backscatter noise (px)
constexpr int kTaps = 16;        // resampler taps per axis
constexpr double kKaiserBeta = 6.0;

SYN_HD inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Uniform in (0, 1).
SYN_HD inline double u01(uint64_t seed, uint64_t stream, uint64_t idx)
{
    uint64_t h = mix64(mix64(seed ^ (stream * 0xD1B54A32D192ED03ULL)) ^ (idx * 0x9E3779B97F4A7C15ULL));
    return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// Two independent N(0,1) values (Box-Muller).
SYN_HD inline void gauss2(uint64_t seed, uint64_t stream, uint64_t idx, double *g0, double *g1)
{
    double u1 = u01(seed, stream, 2 * idx), u2 = u01(seed, stream, 2 * idx + 1);
    double r = sqrt(-2.0 * log(u1));
    double t = 2.0 * kPi * u2;
    *g0 = r * cos(t);
    *g1 = r * sin(t);
}

// Circular complex Gaussian CN(0,1): re, im ~ N(0, 1/2).
SYN_HD inline void cgauss(uint64_t seed, uint64_t stream, uint64_t idx, double *re, double *im)
{
    double a, b;
    gauss2(seed, stream, idx, &a, &b);
    *re = a * 0.70710678118654752440;
    *im = b * 0.70710678118654752440;
}

SYN_HD inline void offset_field(const st_synth_spec &s, double r, double a, double *dr, double *da)
{
    double G = 0.0;
    if (s.ridgeW > 0) {
        double t = (r - s.ridgeC) / s.ridgeW;
        G = exp(-0.5 * t * t);
    }
    double S = 2.0 * r / (double)s.nr - 1.0;
    *dr = s.dr0 + s.drR * r + s.drA * a + s.ridgeR * G + s.shearR * S;
    *da = s.da0 + s.daR * r + s.daA * a + s.ridgeA * G + s.shearA * S;
}

SYN_HD inline double coherence(const st_synth_spec &s, double r, double a)
{
    double v = 0.5;
    if (s.gLr > 0 && s.gLa > 0)
        v = 0.5 + 0.5 * sin(2.0 * kPi * r / s.gLr) * sin(2.0 * kPi * a / s.gLa);
    double g = s.gmin + (s.gmax - s.gmin) * v;
    return g < 0 ? 0 : (g > 1 ? 1 : g);
}

SYN_HD inline bool masked(const st_synth_spec &s, int r, int a)
{
    return r >= s.maskR0 && r < s.maskR1 && a >= s.maskA0 && a < s.maskA1;
}

// Lattice geometry of the backscatter noise.
struct Lattice {
    int pad;     // lattice points beyond each edge
    int nx, ny;  // lattice points per axis
    int reach;   // kernel half-width in lattice steps
};

SYN_HD inline Lattice lattice(const st_synth_spec &s)
{
    Lattice L;
    L.reach = (int)ceil(3.0 * s.corrLen / kLatStep) + 1;
    L.pad = L.reach + 1;
    L.nx = s.nr / kLatStep + 2 * L.pad + 2;
    L.ny = s.na / kLatStep + 2 * L.pad + 2;
    return L;
}

// Lattice noise value for lattice index (iy, ix) with ix, iy >= 0.
SYN_HD inline double lattice_noise(const st_synth_spec &s, const Lattice &L, int iy, int ix)
{
    double g0, g1;
    gauss2(s.seed, 3, (uint64_t)iy * (uint64_t)L.nx + (uint64_t)ix, &g0, &g1);
    return g0;
}

// Kernel weight of lattice point with coordinate c (px) at position x.
SYN_HD inline double lat_weight(const st_synth_spec &s, double x, int k, const Lattice &L)
{
    double c = (double)(k - L.pad) * kLatStep;
    double d = (x - c) / s.corrLen;
    return exp(-d * d);
}

// Modified Bessel I0 (power series).
SYN_HD inline double bessel_i0(double x)
{
    double sum = 1.0, term = 1.0, q = 0.25 * x * x;
    for (int k = 1; k < 40; k++) {
        term *= q / ((double)k * (double)k);
        sum += term;
        if (term < 1e-17 * sum)
            break;
    }
    return sum;
}

// 16-tap Kaiser-windowed sinc weights for fractional position f in [0,1):
// taps at integer offsets t = -7 .. 8 (w[t + 7]); normalised to sum 1.
SYN_HD inline void resample_weights(double f, double w[kTaps])
{
    const double half = kTaps / 2;
    const double i0b = bessel_i0(kKaiserBeta);
    double sum = 0.0;
    for (int t = -7; t <= 8; t++) {
        double u = (double)t - f;
        double sinc = (fabs(u) < 1e-12) ? 1.0 : sin(kPi * u) / (kPi * u);
        double q = u / half;
        double win = (q * q < 1.0) ? bessel_i0(kKaiserBeta * sqrt(1.0 - q * q)) / i0b : 0.0;
        w[t + 7] = sinc * win;
        sum += w[t + 7];
    }
    for (int t = 0; t < kTaps; t++)
        w[t] /= sum;
}

constexpr int kBandHalf = 8;  // band-limiting FIR taps at -8 .. 8

// Band-limiting FIR of the speckle: low-pass of relative bandwidth bw
// (Kaiser-windowed sinc), modulated to centre frequency fc (cycles per
// sample), scaled so that sum |h|^2 = 1 (unit-variance output).
SYN_HD inline void band_fir(double bw, double fc, double hre[2 * kBandHalf + 1], double him[2 * kBandHalf + 1])
{
    const double i0b = bessel_i0(kKaiserBeta);
    double e = 0.0;
    for (int t = -kBandHalf; t <= kBandHalf; t++) {
        double lp;
        if (bw >= 1.0)
            lp = (t == 0) ? 1.0 : 0.0;
        else {
            double x = bw * t;
            double sinc = (t == 0) ? 1.0 : sin(kPi * x) / (kPi * x);
            double q = (double)t / (kBandHalf + 0.5);
            lp = bw * sinc * bessel_i0(kKaiserBeta * sqrt(1.0 - q * q)) / i0b;
        }
        hre[t + kBandHalf] = lp * cos(2.0 * kPi * fc * t);
        him[t + kBandHalf] = lp * sin(2.0 * kPi * fc * t);
        e += lp * lp;
    }
    double sc = 1.0 / sqrt(e);
    for (int t = 0; t <= 2 * kBandHalf; t++) {
        hre[t] *= sc;
        him[t] *= sc;
    }
}

// White CN(0,1) sample of stream `stream` at (a, r) of the extended grid
// [-kBandHalf, na + kBandHalf) x [-kBandHalf, nr + kBandHalf).
SYN_HD inline void white(const st_synth_spec &s, uint64_t stream, int a, int r, double *re, double *im)
{
    uint64_t W = (uint64_t)s.nr + 2 * kBandHalf;
    uint64_t idx = (uint64_t)(a + kBandHalf) * W + (uint64_t)(r + kBandHalf);
    cgauss(s.seed, stream, idx, re, im);
}

}  // namespace synth
