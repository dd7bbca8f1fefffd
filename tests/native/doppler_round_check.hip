// Host check of doppler_round (st_track_common.h) against the rounding it
// replaces: round(W (atan2(zi0, zr0) + atan2(zi1, zr1)) / (4 pi)), half
// away from zero, + W when negative, 0 for a zero lag sum
// (dopplerEstimation.c:42-88 per window, speckleTrack.c's rounding).
// Prints "random <bad> <n> special <bad> <n> near <bad> <n>"; the near
// cases put the angle sum 1e-9 .. 3e-9 rad either side of a rounding
// boundary.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <random>
#include "st_track_common.h"
using namespace sttrack;
namespace sttrack {
__constant__ double2 c_root256[256];
__constant__ double c_refR[kFine * NFAST];
__constant__ double2 c_refB[16];
}  // namespace sttrack

static int by_atan2(double zr0, double zi0, double zr1, double zi1, int W)
{
    const double a0 = (zr0 == 0.0 && zi0 == 0.0) ? NAN : atan2(zi0, zr0);
    const double a1 = (zr1 == 0.0 && zi1 == 0.0) ? NAN : atan2(zi1, zr1);
    const double v = (double)(W / 2) * (a0 + a1) * .5 / M_PI;
    int k = std::isnan(v) ? 0 : ((v >= 0.0) ? (int)(v + 0.5) : (int)(v - 0.5));
    return k < 0 ? k + W : k;
}

int main()
{
    double2 tab[256];
    for (int j = 0; j < 256; j++)
        tab[j] = make_double2(cos(2.0 * M_PI * j / 256.0), sin(2.0 * M_PI * j / 256.0));
    const int Ws[4] = {16, 32, 64, 128};
    std::mt19937_64 rng(5);
    std::normal_distribution<double> nd(0, 1);
    long bad1 = 0, n1 = 0;
    for (int it = 0; it < 400000; it++) {
        const int W = Ws[it & 3];
        const double s0 = std::pow(10.0, (int)(rng() % 40) - 20), s1 = std::pow(10.0, (int)(rng() % 40) - 20);
        const double a = nd(rng) * s0, b = nd(rng) * s0, c = nd(rng) * s1, d = nd(rng) * s1;
        bad1 += doppler_round(a, b, c, d, W, tab) != by_atan2(a, b, c, d, W);
        n1++;
    }
    // signed zeros, axes, extreme magnitudes
    const double sp[] = {0.0, -0.0, 1.0, -1.0, 2.5, -3.0, 1e-300, -1e300, 4e-320};
    long bad2 = 0, n2 = 0;
    for (int W : Ws)
        for (double a : sp)
            for (double b : sp)
                for (double c : sp)
                    for (double d : sp) {
                        bad2 += doppler_round(a, b, c, d, W, tab) != by_atan2(a, b, c, d, W);
                        n2++;
                    }
    long bad3 = 0, n3 = 0;
    for (int W : Ws)
        for (int j = -W / 2; j < W / 2; j++)
            for (int e = -3; e <= 3; e++) {
                if (e == 0)
                    continue;  // on the boundary itself either answer is a rounding of the same tie
                const double S = (j + 0.5) * 4 * M_PI / W + e * 1e-9;
                const double a0 = S * 0.37, a1 = S - a0;
                if (a0 <= -M_PI || a0 > M_PI || a1 <= -M_PI || a1 > M_PI)
                    continue;
                const double zr0 = 3.3 * cos(a0), zi0 = 3.3 * sin(a0), zr1 = 0.7 * cos(a1), zi1 = 0.7 * sin(a1);
                bad3 += doppler_round(zr0, zi0, zr1, zi1, W, tab) != by_atan2(zr0, zi0, zr1, zi1, W);
                n3++;
            }
    printf("random %ld %ld special %ld %ld near %ld %ld\n", bad1, n1, bad2, n2, bad3, n3);
    return 0;
}
