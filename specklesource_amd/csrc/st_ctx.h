// Library context: one per device, owning the stream, the FFT engine, the
// scratch arena and the resident SLC pairs.
#pragma once

#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "st_common.h"
#include "st_fft.h"

struct StPair {
    void *d1 = nullptr, *d2 = nullptr;
    int nr1 = 0, na1 = 0, nr2 = 0, na2 = 0, stype = ST_SAMPLE_C64;
};

struct st_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    stfft::Engine fft;
    Arena arena;
    std::string err;
    std::map<int, StPair> pairs;
    int nextPair = 1;
    std::mutex mu;
    size_t scratchBudget = (size_t)12 << 30;  // bytes of scratch per tracking batch
    st_track_stats stats{};
    bool profiling = false;
    hipEvent_t ev[72];  // [0, 64) kernel timing pairs, [64, 72) stage boundaries
    bool evInit = false;

    void stats_reset();
    // Event-timed region on the context stream: returns an index, call
    // tick_end(index) after the launches; durations are summed into
    // stats at sync points (only when profiling is on).
    int kernel_slot(const char *name);
    void time_begin(int slot);
    void time_end(int slot);
    void flush_timings();
    int pendSlot[32];
    int npend = 0;

    // Make the arena at least `bytes` large and empty it.
    void arena_reserve(size_t bytes);
    // Pinned host staging for result copies, at least `bytes` large.
    void *pinned = nullptr;
    size_t pinnedCap = 0;
    void *pinned_reserve(size_t bytes);
    StPair &pair(int id);
};

// Run fn, translating StError into a return code and st_last_error text.
template <class F>
int st_guard(st_ctx *ctx, F fn)
{
    if (!ctx)
        return ST_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    try {
        ST_HIP(hipSetDevice(ctx->device));
        fn();
        ctx->err.clear();
        return ST_OK;
    } catch (const StError &e) {
        ctx->err = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        ctx->err = e.what();
        return ST_ERR_NOMEM;
    }
}

// Host copies (dst, src, bytes) split over a few threads: the step from a
// pinned staging buffer into the caller's arrays.
void parallel_copy(const std::vector<std::tuple<void *, const void *, size_t>> &jobs);

// Entry points implemented in the per-tool translation units.
void st_speckle_track_impl(st_ctx *ctx, StPair &pr, const st_track_params &par, st_track_out *out,
                           st_track_diag *diag);
void st_corr_track_fast_impl(st_ctx *ctx, StPair &pr, const st_track_params &par, st_track_out *out,
                             st_track_diag *diag);
