// Library context: one per device, owning the stream, the FFT engine, the
// scratch arena and the resident SLC pairs.
#pragma once

#include <atomic>
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
    // bytes of scratch per tracking batch: min(12 GB, a quarter of the free
    // device memory at creation); a batch that cannot get its arena is
    // retried with half the budget
    size_t scratchBudget = (size_t)12 << 30;
    // hard cap on the arena (0 = none): a server slot's arena never grows
    // past the scratch the server charged for it
    size_t arenaLimit = 0;
    // set by st_cancel for the job running on this context; the tracking
    // calls poll it between stages and batches
    std::atomic<bool> cancel{false};
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

    // Make the arena at least `bytes` large and empty it.  Throws
    // ST_ERR_NOMEM when the device (or arenaLimit) cannot give that much.
    void arena_reserve(size_t bytes);
    // Throws ST_ERR_CANCELLED once st_cancel has asked for it.
    void check_cancel() const
    {
        if (cancel.load(std::memory_order_relaxed))
            throw StError{ST_ERR_CANCELLED, "job cancelled"};
    }
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

// Entry points implemented in the per-tool translation units.  With
// `cull` set, strack runs the cullst chain on its output grids while they
// are still on the device (st_cull_chain_device) before copying them out.
void st_speckle_track_impl(st_ctx *ctx, StPair &pr, const st_track_params &par, st_track_out *out,
                           st_track_diag *diag, st_cull_params *cull = nullptr);
void st_corr_track_fast_impl(st_ctx *ctx, StPair &pr, const st_track_params &par, st_track_out *out,
                             st_track_diag *diag);
// n strack calls as one batch (shared fused launches per window class);
// culls may be null, or hold a cull chain (or null) per call.
void st_speckle_track_batch_impl(st_ctx *ctx, int n, StPair *const *prs, const st_track_params *const *pars,
                                 st_track_out *const *outs, st_cull_params *const *culls = nullptr,
                                 int64_t *patches = nullptr);
// Whether a strack call's complex matches run on the fused FP32 kernel
// (what a batch shares).
bool strack_batchable(const st_track_params &p);

// The cullst call order (Cullst/cullst.c:75-100) on grids already on the
// device: offR / offA / type are culled in place, corr is read; the other
// grids of cp (sigma, smoothed offsets) are copied to its host arrays.
// Takes its scratch from the context arena after what is in use (the
// caller reserved cull_device_bytes(n) for it).
size_t cull_device_bytes(size_t n);
void st_cull_chain_device(st_ctx *ctx, st_cull_params *cp, float *dOffR, float *dOffA, float *dCorr,
                          uint8_t *dType);

// Resolves the device address of the no-allocation cull's scratch array on
// `device` (once per device; st_ctx_create calls it).
void cull_noalloc_init(int device);

// Launch choices of one window class (st_autotune and its cache).
struct TuneEntry {
    int window = 0;
    int precision = ST_PREC_FP32;  // what ST_PREC_AUTO runs
    int wgPerCu = 0;               // fused kernel's persistent grid per CU (0: occupancy limit)
    int stagedBatch = 0;           // staged path's batch cap (0: the scratch budget decides)
    float usPerPatch = 0;          // measured complex-stage time per attempt
};
// Reads the device's tune cache once (st_ctx_create calls it).
void tune_init(int device);
// The entry of `window` on `device`; false when there is none.
bool tune_find(int device, int window, TuneEntry *out);
// Resident workgroups per CU of the fused kernel at window w (16, 32, 64).
int fused_occupancy(int w, int stype);
