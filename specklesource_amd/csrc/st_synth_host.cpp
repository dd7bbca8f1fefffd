// Host path of the synthetic SLC generator (st_synth_host).  Rows are
// generated on a few std::threads; every sample depends only on the spec,
// so the result does not depend on the thread count.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "synth_common.h"

using namespace synth;

namespace {

template <class F>
void parallel_rows(int n, F f)
{
    int nt = (int)std::min<unsigned>(8u, std::max(1u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++)
        th.emplace_back([=]() {
            for (int i = t; i < n; i += nt)
                f(i);
        });
    for (auto &x : th)
        x.join();
}

}  // namespace

// Backscatter field g (unit-variance Gaussian random field), na x nr.
static void backscatter_field(const st_synth_spec &s, std::vector<float> &g)
{
    Lattice L = lattice(s);
    std::vector<double> N((size_t)L.ny * L.nx);
    for (int iy = 0; iy < L.ny; iy++)
        for (int ix = 0; ix < L.nx; ix++)
            N[(size_t)iy * L.nx + ix] = lattice_noise(s, L, iy, ix);
    // H[ky][x] and the row normalisation
    std::vector<double> H((size_t)L.ny * s.nr), hn(s.nr), vn(s.na);
    parallel_rows(L.ny, [&](int ky) {
        for (int x = 0; x < s.nr; x++) {
            int c = x / kLatStep + L.pad;
            double acc = 0, nrm = 0;
            for (int kx = c - L.reach; kx <= c + L.reach; kx++) {
                if (kx < 0 || kx >= L.nx)
                    continue;
                double w = lat_weight(s, x, kx, L);
                acc += w * N[(size_t)ky * L.nx + kx];
                nrm += w * w;
            }
            H[(size_t)ky * s.nr + x] = acc;
            if (ky == 0)
                hn[x] = nrm;
        }
    });
    for (int y = 0; y < s.na; y++) {
        int c = y / kLatStep + L.pad;
        double nrm = 0;
        for (int ky = c - L.reach; ky <= c + L.reach; ky++)
            if (ky >= 0 && ky < L.ny) {
                double w = lat_weight(s, y, ky, L);
                nrm += w * w;
            }
        vn[y] = nrm;
    }
    g.assign((size_t)s.na * s.nr, 0.f);
    parallel_rows(s.na, [&](int y) {
        int c = y / kLatStep + L.pad;
        for (int x = 0; x < s.nr; x++) {
            double acc = 0;
            for (int ky = c - L.reach; ky <= c + L.reach; ky++)
                if (ky >= 0 && ky < L.ny)
                    acc += lat_weight(s, y, ky, L) * H[(size_t)ky * s.nr + x];
            g[(size_t)y * s.nr + x] = (float)(acc / sqrt(hn[x] * vn[y]));
        }
    });
}

// Band-limited CN(0,1) speckle of stream `stream` (na x nr, interleaved
// re/im doubles): separable FIR over white noise on the extended grid.
static void speckle(const st_synth_spec &s, uint64_t stream, std::vector<double> &out)
{
    const int K = kBandHalf, nr = s.nr, na = s.na;
    double hr[2 * K + 1], hri[2 * K + 1], ha[2 * K + 1], hai[2 * K + 1];
    band_fir(s.bwR, 0.0, hr, hri);
    band_fir(s.bwA, s.fdc, ha, hai);
    const int nae = na + 2 * K;
    std::vector<double> T((size_t)nae * nr * 2);
    parallel_rows(nae, [&](int ae) {
        int a = ae - K;
        std::vector<double> w((size_t)(nr + 2 * K) * 2);
        for (int r = -K; r < nr + K; r++)
            white(s, stream, a, r, &w[2 * (r + K)], &w[2 * (r + K) + 1]);
        for (int r = 0; r < nr; r++) {
            double re = 0, im = 0;
            for (int t = -K; t <= K; t++) {
                const double *x = &w[2 * (r - t + K)];
                double cr = hr[t + K], ci = hri[t + K];
                re += cr * x[0] - ci * x[1];
                im += cr * x[1] + ci * x[0];
            }
            T[2 * ((size_t)ae * nr + r)] = re;
            T[2 * ((size_t)ae * nr + r) + 1] = im;
        }
    });
    out.assign((size_t)na * nr * 2, 0.0);
    parallel_rows(na, [&](int a) {
        for (int r = 0; r < nr; r++) {
            double re = 0, im = 0;
            for (int t = -K; t <= K; t++) {
                const double *x = &T[2 * ((size_t)(a - t + K) * nr + r)];
                double cr = ha[t + K], ci = hai[t + K];
                re += cr * x[0] - ci * x[1];
                im += cr * x[1] + ci * x[0];
            }
            out[2 * ((size_t)a * nr + r)] = re;
            out[2 * ((size_t)a * nr + r) + 1] = im;
        }
    });
}

extern "C" int st_synth_host(const st_synth_spec *spec, float *img1, float *img2)
{
    if (!spec || !img1 || !img2 || spec->nr <= 0 || spec->na <= 0 || spec->corrLen <= 0)
        return ST_ERR_ARG;
    const st_synth_spec s = *spec;
    std::vector<float> g;
    backscatter_field(s, g);
    std::vector<double> s1, s2;
    speckle(s, 1, s1);
    speckle(s, 2, s2);
    const size_t nr = s.nr;
    parallel_rows(s.na, [&](int a) {
        for (int r = 0; r < s.nr; r++) {
            size_t o = (size_t)a * nr + r;
            double re = s1[2 * o], im = s1[2 * o + 1];
            double amp = sqrt(exp(s.sigmaLn * (double)g[o]));
            if (masked(s, r, a))
                amp = 0.0;
            img1[2 * o] = (float)(amp * re);
            img1[2 * o + 1] = (float)(amp * im);
        }
    });
    parallel_rows(s.na, [&](int a) {
        double wr[kTaps], wa[kTaps];
        for (int r = 0; r < s.nr; r++) {
            size_t o = (size_t)a * nr + r;
            if (masked(s, r, a)) {
                img2[2 * o] = img2[2 * o + 1] = 0.f;
                continue;
            }
            double dr, da;
            offset_field(s, r, a, &dr, &da);
            double pr = r - dr, pa = a - da;
            int r0 = (int)floor(pr), a0 = (int)floor(pa);
            resample_weights(pr - r0, wr);
            resample_weights(pa - a0, wa);
            double sr = 0, si = 0;
            for (int ta = 0; ta < kTaps; ta++) {
                int aa = a0 - 7 + ta;
                if (aa < 0 || aa >= s.na)
                    continue;
                double rr = 0, ri = 0;
                for (int tr = 0; tr < kTaps; tr++) {
                    int rc = r0 - 7 + tr;
                    if (rc < 0 || rc >= s.nr)
                        continue;
                    size_t q = (size_t)aa * nr + rc;
                    rr += wr[tr] * img1[2 * q];
                    ri += wr[tr] * img1[2 * q + 1];
                }
                sr += wa[ta] * rr;
                si += wa[ta] * ri;
            }
            double gam = coherence(s, r, a);
            int rn = (int)floor(pr + 0.5), an = (int)floor(pa + 0.5);
            rn = rn < 0 ? 0 : (rn >= s.nr ? s.nr - 1 : rn);
            an = an < 0 ? 0 : (an >= s.na ? s.na - 1 : an);
            double amp = sqrt(exp(s.sigmaLn * (double)g[(size_t)an * nr + rn]));
            double nre = s2[2 * o], nim = s2[2 * o + 1];
            double f = sqrt(1.0 - gam * gam) * amp;
            img2[2 * o] = (float)(gam * sr + f * nre);
            img2[2 * o + 1] = (float)(gam * si + f * nim);
        }
    });
    return ST_OK;
}

extern "C" void st_synth_offset(const st_synth_spec *spec, double r, double a, double *dr, double *da)
{
    offset_field(*spec, r, a, dr, da);
}
