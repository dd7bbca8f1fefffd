// Autotuner of the strack launch choices, and its cache.
//
// Per window class the tuner times a synthetic job on the device, run as
// four concurrent copies (one per context, as a loaded server runs them),
// and keeps the fastest of:
//   - the precision ST_PREC_AUTO runs (FP32 fused kernel, or the split-bf16
//     matrix-core kernel at windows 16 and 32);
//   - the persistent grid of the fused kernel, as workgroups per CU (from
//     the occupancy limit down to half of it).
// The cache format also carries a batch cap for the staged path; the
// tuner leaves it at 0 (scratch-budget batches), since every tuned window
// now runs a fused kernel.
// The choices are kept per device in a process-wide table and in a text
// cache, one line per window:
//
//   # speckle_amd tune cache 1: arch window precision wg_per_cu staged_batch us_per_patch
//   gfx950 64 fp32 2 0 0.0476
//
// A context loads the cache of its device's architecture when the first
// context on that device is created: $SPECKLE_AMD_TUNE if set, else
// tune/<arch>.txt beside the library directory.  Windows without an entry
// use the defaults (FP32, occupancy limit, scratch-budget batches).
#include <dlfcn.h>
#include <string.h>

#include <chrono>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "st_ctx.h"
#include "../../include/speckle_synth.h"

namespace {

struct TuneState {
    std::mutex mu;
    std::map<int, std::string> arch;                  // device -> architecture name
    std::map<std::string, std::map<int, TuneEntry>> t;  // arch -> window -> entry
    std::map<int, bool> loaded;                        // device -> cache read
};

TuneState &state()
{
    static TuneState s;
    return s;
}

std::string device_arch(int device)
{
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess)
        return "unknown";
    std::string a(p.gcnArchName);
    const size_t c = a.find(':');
    return c == std::string::npos ? a : a.substr(0, c);
}

const char *prec_name(int p) { return p == ST_PREC_MFMA ? "mfma" : "fp32"; }

constexpr int kLoad = 4;  // concurrent copies of a tuning job (the server's default slots)
// smallest gain over the default choice that the tuner acts on: the
// served-pass noise is about 1-2 %
constexpr float kMinGain = 0.03f;

// Parse cache lines into the table of `arch`; lines of other
// architectures are skipped.  Returns the number of entries read.
int parse_cache(std::istream &in, const std::string &arch, std::map<int, TuneEntry> &out)
{
    std::string line;
    int n = 0;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#')
            continue;
        std::istringstream ss(line);
        std::string a, prec;
        TuneEntry e;
        if (!(ss >> a >> e.window >> prec >> e.wgPerCu >> e.stagedBatch >> e.usPerPatch))
            continue;
        if (a != arch || e.window <= 0)
            continue;
        e.precision = prec == "mfma" ? ST_PREC_MFMA : ST_PREC_FP32;
        out[e.window] = e;
        n++;
    }
    return n;
}

std::string default_cache_path(const std::string &arch)
{
    if (const char *e = getenv("SPECKLE_AMD_TUNE"))
        return e;
    Dl_info info;
    if (dladdr((void *)&device_arch, &info) && info.dli_fname) {
        std::string lib(info.dli_fname);
        const size_t s = lib.rfind('/');
        const std::string dir = s == std::string::npos ? "." : lib.substr(0, s);
        return dir + "/../tune/" + arch + ".txt";
    }
    return "tune/" + arch + ".txt";
}

}  // namespace

void tune_init(int device)
{
    TuneState &S = state();
    std::lock_guard<std::mutex> lk(S.mu);
    if (S.loaded[device])
        return;
    S.loaded[device] = true;
    const std::string arch = device_arch(device);
    S.arch[device] = arch;
    std::ifstream f(default_cache_path(arch));
    if (f)
        parse_cache(f, arch, S.t[arch]);
}

bool tune_find(int device, int window, TuneEntry *out)
{
    TuneState &S = state();
    std::lock_guard<std::mutex> lk(S.mu);
    auto a = S.arch.find(device);
    if (a == S.arch.end())
        return false;
    auto &tab = S.t[a->second];
    auto it = tab.find(window);
    if (it == tab.end())
        return false;
    *out = it->second;
    return true;
}

namespace {

void tune_set(int device, const TuneEntry &e)
{
    TuneState &S = state();
    std::lock_guard<std::mutex> lk(S.mu);
    S.t[S.arch[device]][e.window] = e;
}

void tune_clear(int device, int window)
{
    TuneState &S = state();
    std::lock_guard<std::mutex> lk(S.mu);
    S.t[S.arch[device]].erase(window);
}

// A synthetic job of window w with about `target` complex attempts.
void tune_job(int w, int target, st_synth_spec &s, st_track_params &p)
{
    const int step = w / 2, pad = 2 * w + 40;
    int side = 1;
    while (side * side < target)
        side++;
    const int n = 2 * pad + 2 * w + side * step;
    memset(&s, 0, sizeof(s));
    s.nr = s.na = n;
    s.seed = 9000 + w;
    s.dr0 = 0.4;
    s.da0 = -0.3;
    s.gmin = 0.5;
    s.gmax = 0.9;
    s.gLr = n / 1.3;
    s.gLa = n / 1.7;
    s.sigmaLn = 0.5;
    s.corrLen = 64;
    s.bwR = 0.85;
    s.bwA = 0.8;
    s.fdc = 0.12;
    memset(&p, 0, sizeof(p));
    p.rStart = p.aStart = pad + w;
    p.deltaR = p.deltaA = step;
    p.nR = p.nA = side;
    p.wR = p.wA = p.wRa = p.wAa = w;
    p.scaleFactor = 1;
    p.hanningFlag = 1;
    p.maxTries = 2;
    p.azDefocusThresh = 1.8f;
    p.polyShift = 1;
    p.rShiftPoly[0] = 0.4;
    p.aShiftPoly[0] = -0.3;
}

}  // namespace

extern "C" int st_autotune(st_ctx *ctx, const int32_t *windows, int32_t n, const char *cacheFile)
{
    if (!ctx || !windows || n < 1)
        return ST_ERR_ARG;
    std::vector<TuneEntry> best;
    int rc = ST_OK;
    // the other contexts of the concurrent measurement
    std::vector<st_ctx *> load;
    struct Release {
        std::vector<st_ctx *> &v;
        ~Release()
        {
            for (st_ctx *c : v)
                st_ctx_destroy(c);
        }
    } release{load};
    for (int q = 1; q < kLoad && rc == ST_OK; q++) {
        st_ctx *c = nullptr;
        rc = st_ctx_create(ctx->device, &c);
        if (rc == ST_OK) {
            c->scratchBudget = ctx->scratchBudget / kLoad;
            load.push_back(c);
        }
    }
    if (rc != ST_OK)
        return rc;
    for (int k = 0; k < n && rc == ST_OK; k++) {
        const int w = windows[k];
        if (w != 16 && w != 32 && w != 64 && w != 128)
            return st_guard(ctx, [&]() { ST_REQUIRE(false, ST_ERR_ARG, "tuned windows are 16, 32, 64, 128"); });
        st_synth_spec spec;
        st_track_params par;
        tune_job(w, w >= 128 ? 2048 : 16384, spec, par);
        int32_t pair = -1;
        rc = st_synth_device(ctx, &spec, &pair);
        if (rc != ST_OK)
            break;
        const size_t npts = (size_t)par.nR * par.nA;
        std::vector<float> o0(npts), o1(npts), o2(npts);
        std::vector<uint8_t> ot(npts);
        st_track_out out;
        memset(&out, 0, sizeof(out));
        out.offR = o0.data();
        out.offA = o1.data();
        out.corr = o2.data();
        out.type = ot.data();
        // candidates
        std::vector<TuneEntry> cand;
        int maxPerCu = 0;
        rc = st_guard(ctx, [&]() { maxPerCu = fused_occupancy(w, ST_SAMPLE_C64); });
        for (int g = maxPerCu; g >= std::max(1, maxPerCu / 2); g--)
            cand.push_back(TuneEntry{w, ST_PREC_FP32, g, 0, 0});
        if (w <= 32)
            cand.push_back(TuneEntry{w, ST_PREC_MFMA, 0, 0, 0});
        // cand[0] is the default choice (FP32, occupancy limit, budget
        // batches); another one must beat it by kMinGain to replace it
        // every candidate runs as kLoad copies of the job at once, one per
        // context, as a loaded server runs them: a kernel that is fastest
        // alone is not always fastest beside others
        StPair pr;
        {
            std::lock_guard<std::mutex> lk(ctx->mu);
            pr = ctx->pairs[pair];
        }
        std::vector<std::vector<float>> bufs(3 * kLoad, std::vector<float>(npts));
        std::vector<std::vector<uint8_t>> tbufs(kLoad, std::vector<uint8_t>(npts));
        std::vector<st_track_out> outs(kLoad);
        for (int q = 0; q < kLoad; q++) {
            memset(&outs[q], 0, sizeof(st_track_out));
            outs[q].offR = bufs[3 * q].data();
            outs[q].offA = bufs[3 * q + 1].data();
            outs[q].corr = bufs[3 * q + 2].data();
            outs[q].type = tbufs[q].data();
        }
        (void)out;
        TuneEntry win{};
        float winUs = 0;
        par.precision = ST_PREC_AUTO;
        for (TuneEntry &c : cand) {
            tune_set(ctx->device, c);
            float us = 0;
            for (int rep = 0; rep < 3 && rc == ST_OK; rep++) {
                std::vector<int> rcs(kLoad, ST_OK);
                std::vector<std::thread> th;
                const auto t0 = std::chrono::steady_clock::now();
                for (int q = 0; q < kLoad; q++)
                    th.emplace_back([&, q]() {
                        st_ctx *cq = q == 0 ? ctx : load[q - 1];
                        StPair pq = pr;
                        rcs[q] = st_guard(cq, [&]() { st_speckle_track_impl(cq, pq, par, &outs[q], nullptr); });
                    });
                for (auto &t : th)
                    t.join();
                const double wallUs =
                    std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
                for (int q = 0; q < kLoad; q++)
                    rc = rc == ST_OK ? rcs[q] : rc;
                const double corr = ctx->stats.nCorr[0] > 0 ? (double)ctx->stats.nCorr[0] : 1.0;
                if (rep > 0)  // the first run warms up
                    us += (float)(wallUs / (kLoad * corr) / 2);
            }
            c.usPerPatch = us;
            if (rc == ST_OK && (winUs == 0 || us < winUs)) {
                win = c;
                winUs = us;
            }
        }
        if (rc == ST_OK && !cand.empty() && win.usPerPatch > (1.0f - kMinGain) * cand[0].usPerPatch)
            win = cand[0];
        st_pair_release(ctx, pair);
        if (rc != ST_OK)
            break;
        tune_set(ctx->device, win);
        best.push_back(win);
    }
    if (rc != ST_OK) {
        for (int k = 0; k < n; k++)
            tune_clear(ctx->device, windows[k]);
        return rc;
    }
    if (cacheFile && *cacheFile) {
        std::ofstream f(cacheFile);
        if (!f)
            return st_guard(ctx, [&]() { ST_REQUIRE(false, ST_ERR_ARG, "cannot write the tune cache"); });
        const std::string arch = device_arch(ctx->device);
        f << "# speckle_amd tune cache 1: arch window precision wg_per_cu staged_batch us_per_patch\n";
        for (const TuneEntry &e : best) {
            char line[160];
            snprintf(line, sizeof(line), "%s %d %s %d %d %.5f\n", arch.c_str(), e.window, prec_name(e.precision),
                     e.wgPerCu, e.stagedBatch, e.usPerPatch);
            f << line;
        }
    }
    return ST_OK;
}

extern "C" int st_tune_get(st_ctx *ctx, int32_t window, int32_t *precision, int32_t *wgPerCu,
                           int32_t *stagedBatch, float *usPerPatch)
{
    if (!ctx)
        return ST_ERR_ARG;
    TuneEntry e;
    if (!tune_find(ctx->device, window, &e))
        return ST_ERR_ARG;
    if (precision)
        *precision = e.precision;
    if (wgPerCu)
        *wgPerCu = e.wgPerCu;
    if (stagedBatch)
        *stagedBatch = e.stagedBatch;
    if (usPerPatch)
        *usPerPatch = e.usPerPatch;
    return ST_OK;
}

extern "C" int st_tune_load(st_ctx *ctx, const char *cacheFile)
{
    if (!ctx || !cacheFile)
        return ST_ERR_ARG;
    const std::string arch = device_arch(ctx->device);
    std::ifstream f(cacheFile);
    if (!f)
        return st_guard(ctx, [&]() { ST_REQUIRE(false, ST_ERR_ARG, std::string("cannot read ") + cacheFile); });
    std::map<int, TuneEntry> tab;
    parse_cache(f, arch, tab);
    for (auto &kv : tab)
        tune_set(ctx->device, kv.second);
    return ST_OK;
}
