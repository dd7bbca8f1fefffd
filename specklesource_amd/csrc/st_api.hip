// Contexts, resident pairs and the tracking entry points of the C ABI
// (include/speckle_amd.h).
#include <string.h>

#include <algorithm>
#include <thread>

#include "st_ctx.h"

void *st_ctx::pinned_reserve(size_t bytes)
{
    if (pinnedCap < bytes) {
        if (pinned)
            ST_HIP(hipHostFree(pinned));
        pinned = nullptr;
        pinnedCap = 0;
        const size_t cap = bytes + (bytes >> 3);
        ST_HIP(hipHostMalloc(&pinned, cap, hipHostMallocDefault));
        pinnedCap = cap;
    }
    return pinned;
}

void parallel_copy(const std::vector<std::tuple<void *, const void *, size_t>> &jobs)
{
    constexpr size_t kPiece = (size_t)1 << 20;
    std::vector<std::tuple<char *, const char *, size_t>> pieces;
    size_t total = 0;
    for (const auto &j : jobs) {
        char *d = (char *)std::get<0>(j);
        const char *s = (const char *)std::get<1>(j);
        const size_t n = std::get<2>(j);
        for (size_t o = 0; o < n; o += kPiece)
            pieces.emplace_back(d + o, s + o, std::min(kPiece, n - o));
        total += n;
    }
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>({(size_t)4, (size_t)hw, pieces.size()});
    if (nt <= 1 || total < ((size_t)2 << 20)) {
        for (const auto &p : pieces)
            memcpy(std::get<0>(p), std::get<1>(p), std::get<2>(p));
        return;
    }
    std::vector<std::thread> th;
    for (size_t t = 0; t < nt; t++)
        th.emplace_back([&, t]() {
            for (size_t k = t; k < pieces.size(); k += nt)
                memcpy(std::get<0>(pieces[k]), std::get<1>(pieces[k]), std::get<2>(pieces[k]));
        });
    for (auto &x : th)
        x.join();
}

void st_ctx::arena_reserve(size_t bytes)
{
    ST_REQUIRE(!arenaLimit || bytes <= arenaLimit, ST_ERR_NOMEM,
               "scratch need of " + std::to_string(bytes >> 20) + " MB above the context's arena limit");
    if (arena.cap < bytes) {
        if (arena.base)
            ST_HIP(hipFree(arena.base));
        arena.base = nullptr;
        arena.cap = 0;
        size_t cap = bytes + (bytes >> 3);
        if (arenaLimit)
            cap = std::min(cap, arenaLimit);
        const hipError_t e = hipMalloc(&arena.base, cap);
        if (e == hipErrorOutOfMemory) {
            (void)hipGetLastError();
            arena.base = nullptr;
            throw StError{ST_ERR_NOMEM, "device memory: no room for " + std::to_string(cap >> 20) + " MB of scratch"};
        }
        ST_HIP(e);
        arena.cap = cap;
    }
    arena.reset();
}

void st_ctx::stats_reset()
{
    stats = st_track_stats{};
    npend = 0;
    if (!evInit) {
        for (int k = 0; k < 72; k++)
            ST_HIP(hipEventCreate(&ev[k]));
        evInit = true;
    }
}

int st_ctx::kernel_slot(const char *name)
{
    for (int k = 0; k < ST_KERNEL_SLOTS; k++) {
        if (stats.kernelNames[k][0] == 0) {
            strncpy(stats.kernelNames[k], name, 31);
            return k;
        }
        if (strncmp(stats.kernelNames[k], name, 31) == 0)
            return k;
    }
    return ST_KERNEL_SLOTS - 1;
}

void st_ctx::time_begin(int slot)
{
    if (!profiling)
        return;
    if (npend >= 31)
        flush_timings();
    pendSlot[npend] = slot;
    ST_HIP(hipEventRecord(ev[2 * npend], stream));
}

void st_ctx::time_end(int slot)
{
    (void)slot;
    if (!profiling)
        return;
    ST_HIP(hipEventRecord(ev[2 * npend + 1], stream));
    npend++;
}

void st_ctx::flush_timings()
{
    if (!npend)
        return;
    ST_HIP(hipEventSynchronize(ev[2 * npend - 1]));
    for (int k = 0; k < npend; k++) {
        float ms = 0;
        ST_HIP(hipEventElapsedTime(&ms, ev[2 * k], ev[2 * k + 1]));
        stats.kernelMs[pendSlot[k]] += ms;
        stats.kernelLaunches[pendSlot[k]] += 1;
    }
    npend = 0;
}

StPair &st_ctx::pair(int id)
{
    auto it = pairs.find(id);
    ST_REQUIRE(it != pairs.end(), ST_ERR_ARG, "unknown pair id " + std::to_string(id));
    return it->second;
}

extern "C" int st_ctx_create(int32_t device, st_ctx **out)
{
    if (!out)
        return ST_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return ST_ERR_NODEV;
    if (device < 0 || device >= n)
        return ST_ERR_ARG;
    st_ctx *c = new st_ctx();
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return ST_ERR_HIP;
    }
    size_t freeB = 0, totalB = 0;
    if (hipMemGetInfo(&freeB, &totalB) == hipSuccess && freeB > 0)
        c->scratchBudget = std::min(c->scratchBudget, std::max<size_t>(freeB / 4, (size_t)64 << 20));
    try {
        cull_noalloc_init(device);
        tune_init(device);
    } catch (const StError &e) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return e.code;
    }
    *out = c;
    return ST_OK;
}

extern "C" int st_ctx_destroy(st_ctx *c)
{
    if (!c)
        return ST_ERR_ARG;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    for (auto &kv : c->pairs) {
        hipFree(kv.second.d1);
        hipFree(kv.second.d2);
    }
    if (c->arena.base)
        hipFree(c->arena.base);
    if (c->pinned)
        hipHostFree(c->pinned);
    if (c->evInit)
        for (int k = 0; k < 72; k++)
            hipEventDestroy(c->ev[k]);
    hipStreamDestroy(c->stream);
    delete c;
    return ST_OK;
}

extern "C" const char *st_last_error(st_ctx *c)
{
    return c ? c->err.c_str() : "null context";
}

namespace {

// In-place byte swap of an uploaded image: 32-bit words (c64 samples) or
// the two 16-bit halves of each word (ci16 samples); 16 bytes per lane.
__global__ void k_bswap(uint4 *d, size_t n16, uint32_t *tail, size_t nTail, int half)
{
    auto sw = [half](uint32_t x) {
        return half ? (((x & 0x00FF00FFu) << 8) | ((x >> 8) & 0x00FF00FFu)) : __builtin_bswap32(x);
    };
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) {
        uint4 v = d[i];
        v.x = sw(v.x);
        v.y = sw(v.y);
        v.z = sw(v.z);
        v.w = sw(v.w);
        d[i] = v;
    }
    if (blockIdx.x == 0 && threadIdx.x < nTail)
        tail[threadIdx.x] = sw(tail[threadIdx.x]);
}

void bswap_device(void *d, size_t bytes, int half, hipStream_t st)
{
    const size_t n16 = bytes / 16, nTail = (bytes - n16 * 16) / 4;
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>((n16 + 255) / 256, 8192));
    hipLaunchKernelGGL(k_bswap, dim3(blocks), dim3(256), 0, st, (uint4 *)d, n16, (uint32_t *)((char *)d + n16 * 16),
                       nTail, half);
    ST_HIP(hipGetLastError());
}

int upload_pair(st_ctx *c, const void *img1, int32_t nr1, int32_t na1, const void *img2, int32_t nr2, int32_t na2,
                int32_t stype, int swap, int32_t *pair)
{
    return st_guard(c, [&]() {
        ST_REQUIRE(img1 && img2 && pair && nr1 > 0 && na1 > 0 && nr2 > 0 && na2 > 0, ST_ERR_ARG, "bad pair arguments");
        ST_REQUIRE(stype == ST_SAMPLE_C64 || stype == ST_SAMPLE_CI16, ST_ERR_ARG, "bad sample type");
        size_t es = stype == ST_SAMPLE_C64 ? 8 : 4;
        StPair p;
        p.nr1 = nr1;
        p.na1 = na1;
        p.nr2 = nr2;
        p.na2 = na2;
        p.stype = stype;
        size_t b1 = es * (size_t)nr1 * na1, b2 = es * (size_t)nr2 * na2;
        ST_HIP(hipMalloc(&p.d1, b1));
        ST_HIP(hipMalloc(&p.d2, b2));
        ST_HIP(hipMemcpyAsync(p.d1, img1, b1, hipMemcpyHostToDevice, c->stream));
        ST_HIP(hipMemcpyAsync(p.d2, img2, b2, hipMemcpyHostToDevice, c->stream));
        if (swap) {
            const int half = stype == ST_SAMPLE_CI16;
            c->stats_reset();
            const int slot = c->kernel_slot("k_bswap");
            c->time_begin(slot);
            bswap_device(p.d1, b1, half, c->stream);
            bswap_device(p.d2, b2, half, c->stream);
            c->time_end(slot);
        }
        ST_HIP(hipStreamSynchronize(c->stream));
        c->flush_timings();
        int id = c->nextPair++;
        c->pairs[id] = p;
        *pair = id;
    });
}

}  // namespace

extern "C" int st_pair_upload(st_ctx *c, const void *img1, int32_t nr1, int32_t na1, const void *img2, int32_t nr2,
                              int32_t na2, int32_t stype, int32_t *pair)
{
    return upload_pair(c, img1, nr1, na1, img2, nr2, na2, stype, 0, pair);
}

extern "C" int st_pair_upload_raw(st_ctx *c, const void *img1, int32_t nr1, int32_t na1, const void *img2,
                                  int32_t nr2, int32_t na2, int32_t stype, int32_t byteOrder, int32_t *pair)
{
    if (byteOrder != ST_MSB && byteOrder != ST_LSB)
        return st_guard(c, [&]() { ST_REQUIRE(false, ST_ERR_ARG, "byte order"); });
    const int swap = stype == ST_SAMPLE_CI16 || byteOrder == ST_MSB;
    return upload_pair(c, img1, nr1, na1, img2, nr2, na2, stype, swap, pair);
}

extern "C" int st_pair_release(st_ctx *c, int32_t id)
{
    return st_guard(c, [&]() {
        StPair &p = c->pair(id);
        ST_HIP(hipStreamSynchronize(c->stream));
        ST_HIP(hipFree(p.d1));
        ST_HIP(hipFree(p.d2));
        c->pairs.erase(id);
    });
}

extern "C" int st_pair_download(st_ctx *c, int32_t id, int32_t row0, int32_t nrows, void *img1, void *img2)
{
    return st_guard(c, [&]() {
        StPair &p = c->pair(id);
        ST_REQUIRE(p.stype == ST_SAMPLE_C64, ST_ERR_UNSUPPORTED, "download of complex int16 pairs");
        ST_REQUIRE(row0 >= 0 && nrows >= 0 && row0 + nrows <= p.na1 && row0 + nrows <= p.na2, ST_ERR_ARG,
                   "row range outside the pair");
        if (img1)
            ST_HIP(hipMemcpyAsync(img1, (char *)p.d1 + (size_t)row0 * p.nr1 * 8, (size_t)nrows * p.nr1 * 8,
                                  hipMemcpyDeviceToHost, c->stream));
        if (img2)
            ST_HIP(hipMemcpyAsync(img2, (char *)p.d2 + (size_t)row0 * p.nr2 * 8, (size_t)nrows * p.nr2 * 8,
                                  hipMemcpyDeviceToHost, c->stream));
        ST_HIP(hipStreamSynchronize(c->stream));
    });
}

extern "C" int st_speckle_track(st_ctx *c, int32_t pair, const st_track_params *par, st_track_out *out,
                                st_track_diag *diag)
{
    return st_guard(c, [&]() {
        ST_REQUIRE(par && out && out->offR && out->offA && out->corr && out->type, ST_ERR_ARG, "null argument");
        st_speckle_track_impl(c, c->pair(pair), *par, out, diag);
    });
}

extern "C" int st_speckle_track_batch(st_ctx *c, int32_t n, const int32_t *pairs, const st_track_params *pars,
                                      st_track_out *outs)
{
    return st_guard(c, [&]() {
        ST_REQUIRE(n >= 1 && pairs && pars && outs, ST_ERR_ARG, "null argument");
        std::vector<StPair *> prs(n);
        std::vector<const st_track_params *> ps(n);
        std::vector<st_track_out *> os(n);
        for (int k = 0; k < n; k++) {
            ST_REQUIRE(outs[k].offR && outs[k].offA && outs[k].corr && outs[k].type, ST_ERR_ARG, "null output");
            prs[k] = &c->pair(pairs[k]);
            ps[k] = &pars[k];
            os[k] = &outs[k];
        }
        st_speckle_track_batch_impl(c, n, prs.data(), ps.data(), os.data());
    });
}

extern "C" int st_corr_track_fast(st_ctx *c, int32_t pair, const st_track_params *par, st_track_out *out,
                                  st_track_diag *diag)
{
    return st_guard(c, [&]() {
        ST_REQUIRE(par && out && out->offR && out->offA && out->corr && out->type, ST_ERR_ARG, "null argument");
        st_corr_track_fast_impl(c, c->pair(pair), *par, out, diag);
    });
}

extern "C" int st_last_stats(st_ctx *c, st_track_stats *s)
{
    return st_guard(c, [&]() {
        ST_REQUIRE(s, ST_ERR_ARG, "null stats");
        *s = c->stats;
    });
}

extern "C" int st_set_profiling(st_ctx *c, int32_t on)
{
    return st_guard(c, [&]() { c->profiling = on != 0; });
}
