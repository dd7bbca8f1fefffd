// Serving: continuous batching of tracking jobs on one GPU.
//
// A server owns `slots` contexts (one stream, FFT plans and a capped
// scratch arena each), one worker thread per slot and one stager thread.
// Jobs that name a resident pair are ready at once.  Jobs that bring host
// images wait in the staging queue: the stager takes the first one whose
// images fit in what the HBM budget has left (so small jobs overtake a
// large one that must wait for memory), uploads them on its own stream
// while the slots compute, and hands the job on as ready.  Idle workers
// take ready jobs in arrival order, so jobs that arrive while others run
// start as soon as a slot frees up.  Kernels of different slots overlap on
// the device (all streams at the default priority).  A job's results
// depend only on its inputs: every slot runs the same code on its own
// stream and scratch.
#include <string.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "st_ctx.h"
#include "../../include/speckle_synth.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Job {
    st_job j{};
    size_t bytes = 0;  // admission charge: the job's own images (pair = -1)
    int state = ST_JOB_QUEUED, status = ST_OK, slot = -1;
    bool staging = false;  // the stager is uploading its images
    bool cancel = false;   // st_cancel while staging or running
    StPair img;            // staged images of a pair = -1 job
    Clock::time_point tSubmit, tStart, tEnd;
    int64_t patches = 0;
    std::string err;
};

struct ResPair {
    StPair p;
    size_t bytes = 0;
    int refs = 0;  // submitted jobs naming the pair that have not ended
    bool leaving = false;
};

double ms(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

size_t sample_bytes(int stype) { return stype == ST_SAMPLE_C64 ? 8 : 4; }

bool ended(int state) { return state == ST_JOB_DONE || state == ST_JOB_FAILED || state == ST_JOB_CANCELLED; }

void free_pair(StPair &p)
{
    if (p.d1)
        (void)hipFree(p.d1);
    if (p.d2)
        (void)hipFree(p.d2);
    p.d1 = p.d2 = nullptr;
}

}  // namespace

struct st_server {
    int device = 0;
    uint64_t budget = 0;
    size_t slotScratch = 0;
    uint64_t used = 0;  // slot arenas + resident pairs + staged images
    std::vector<st_ctx *> ctxs;
    std::vector<std::thread> workers;
    std::thread stagerThread;
    std::mutex mu;
    std::condition_variable work;   // ready queue changed, stop
    std::condition_variable stage;  // staging queue changed, memory freed, stop
    std::condition_variable done;   // a job ended
    std::deque<int64_t> queue;      // waiting for the stager
    std::deque<int64_t> ready;      // waiting for a slot
    std::map<int64_t, Job> jobs;
    std::map<int, ResPair> pairs;
    int nextPair = 1;
    int64_t nextId = 1;
    bool stop = false;
    hipStream_t up = nullptr;
    int maxBatch = 1;  // strack jobs per fused batch (st_server_set_batch)
    static constexpr size_t kMaxBatchPoints = (size_t)4 << 20;  // grid points per batch

    // Called with mu held: the job has ended (any state); drop its pair
    // reference and wake whoever waits.
    void finish(Job &jb, int state)
    {
        jb.state = state;
        jb.tEnd = Clock::now();
        if (jb.j.pair >= 0) {
            auto it = pairs.find(jb.j.pair);
            if (it != pairs.end())
                it->second.refs--;
        }
        done.notify_all();
    }

    void stager()
    {
        for (;;) {
            Job *jb = nullptr;
            int64_t id = -1;
            {
                std::unique_lock<std::mutex> lk(mu);
                stage.wait(lk, [&]() {
                    if (stop && queue.empty())
                        return true;
                    // stage ahead of the slots, not without bound
                    if (ready.size() >= ctxs.size())
                        return false;
                    for (auto it = queue.begin(); it != queue.end(); ++it) {
                        Job &c = jobs.at(*it);
                        if (used + c.bytes <= budget) {
                            id = *it;
                            jb = &c;
                            queue.erase(it);
                            return true;
                        }
                    }
                    return false;
                });
                if (!jb)
                    return;
                used += jb->bytes;
                jb->staging = true;
            }
            const st_job &J = jb->j;
            StPair p;
            p.nr1 = J.nr1;
            p.na1 = J.na1;
            p.nr2 = J.nr2;
            p.na2 = J.na2;
            p.stype = J.sampleType;
            const size_t b1 = sample_bytes(J.sampleType) * (size_t)J.nr1 * J.na1,
                         b2 = sample_bytes(J.sampleType) * (size_t)J.nr2 * J.na2;
            const bool ok = hipSetDevice(device) == hipSuccess && hipMalloc(&p.d1, b1) == hipSuccess &&
                            hipMalloc(&p.d2, b2) == hipSuccess &&
                            hipMemcpyAsync(p.d1, J.img1, b1, hipMemcpyHostToDevice, up) == hipSuccess &&
                            hipMemcpyAsync(p.d2, J.img2, b2, hipMemcpyHostToDevice, up) == hipSuccess &&
                            hipStreamSynchronize(up) == hipSuccess;
            std::lock_guard<std::mutex> lk(mu);
            jb->staging = false;
            if (!ok || jb->cancel) {
                free_pair(p);
                used -= jb->bytes;
                if (jb->cancel) {
                    jb->status = ST_ERR_CANCELLED;
                    jb->err = "job cancelled";
                    finish(*jb, ST_JOB_CANCELLED);
                } else {
                    jb->status = ST_ERR_HIP;
                    jb->err = "image upload failed";
                    finish(*jb, ST_JOB_FAILED);
                }
                stage.notify_all();
                work.notify_all();  // idle workers may be waiting to stop
                continue;
            }
            jb->img = p;
            ready.push_back(id);
            work.notify_one();
        }
    }

    void run(int slot, Job &jb)
    {
        st_ctx *c = ctxs[slot];
        StPair pr;
        int rc = ST_OK;
        if (jb.j.kind == ST_JOB_CULL) {
            rc = c->cancel.load() ? ST_ERR_CANCELLED : st_cull_run(c, jb.j.cull);
            if (rc == ST_ERR_CANCELLED)
                jb.err = "job cancelled";
            else if (rc != ST_OK)
                jb.err = c->err.empty() ? "job failed" : c->err;
            jb.patches = 0;
            jb.status = rc;
            return;
        }
        if (jb.j.pair >= 0) {
            std::lock_guard<std::mutex> lk(mu);
            auto it = pairs.find(jb.j.pair);
            if (it == pairs.end())
                rc = ST_ERR_ARG;  // cannot happen: a pair with references is never freed
            else
                pr = it->second.p;
        } else {
            pr = jb.img;
        }
        if (rc == ST_OK) {
            rc = st_guard(c, [&]() {
                const bool cullOnDevice = jb.j.cull && jb.j.kind == ST_JOB_STRACK;
                if (jb.j.kind == ST_JOB_STRACK)
                    st_speckle_track_impl(c, pr, jb.j.par, &jb.j.out, nullptr, cullOnDevice ? jb.j.cull : nullptr);
                else
                    st_corr_track_fast_impl(c, pr, jb.j.par, &jb.j.out, nullptr);
                jb.patches = (int64_t)c->stats.nCorr[0] + c->stats.nCorr[1] + c->stats.nCorr[2];
            });
        } else {
            c->err = "the job's pair is no longer resident";
        }
        // strackw composes its grids on the host: they go through the
        // host-grid cull entry point
        if (rc == ST_OK && jb.j.cull && jb.j.kind != ST_JOB_STRACK)
            rc = c->cancel.load() ? ST_ERR_CANCELLED : st_cull_run(c, jb.j.cull);
        if (rc == ST_ERR_CANCELLED)
            jb.err = "job cancelled";
        else if (rc != ST_OK)
            jb.err = c->err.empty() ? "job failed" : c->err;
        jb.status = rc;
    }

    static bool batchable(const Job &jb)
    {
        return jb.j.kind == ST_JOB_STRACK && strack_batchable(jb.j.par);
    }
    static size_t points(const Job &jb) { return (size_t)std::max(0, jb.j.par.nR) * std::max(0, jb.j.par.nA); }

    // Several ready strack jobs on the fused path run as one batch: their
    // complex matches share one persistent launch per window class.  If
    // the batch fails (a job's own error, or a cancel), every job that was
    // not cancelled runs again alone, so each gets its own status.
    void run_batch(int slot, std::vector<Job *> &batch)
    {
        st_ctx *c = ctxs[slot];
        const int n = (int)batch.size();
        std::vector<StPair> prs(n);
        std::vector<StPair *> pp(n);
        std::vector<const st_track_params *> ps(n);
        std::vector<st_track_out *> os(n);
        std::vector<st_cull_params *> cs(n);
        std::vector<int64_t> pat(n, 0);
        {
            std::lock_guard<std::mutex> lk(mu);
            for (int k = 0; k < n; k++) {
                Job &jb = *batch[k];
                prs[k] = jb.j.pair >= 0 ? pairs.at(jb.j.pair).p : jb.img;
                pp[k] = &prs[k];
                ps[k] = &jb.j.par;
                os[k] = &jb.j.out;
                cs[k] = jb.j.cull;
            }
        }
        const int rc = st_guard(c, [&]() {
            st_speckle_track_batch_impl(c, n, pp.data(), ps.data(), os.data(), cs.data(), pat.data());
        });
        if (rc == ST_OK) {
            for (int k = 0; k < n; k++) {
                batch[k]->status = ST_OK;
                batch[k]->patches = pat[k];
            }
            return;
        }
        for (Job *jb : batch) {
            bool cancelled;
            {
                std::lock_guard<std::mutex> lk(mu);
                cancelled = jb->cancel;
                c->cancel.store(cancelled);
            }
            if (cancelled) {
                jb->status = ST_ERR_CANCELLED;
                jb->err = "job cancelled";
            } else {
                run(slot, *jb);
            }
        }
    }

    void worker(int slot)
    {
        for (;;) {
            std::vector<Job *> batch;
            {
                std::unique_lock<std::mutex> lk(mu);
                work.wait(lk, [&]() { return !ready.empty() || (stop && queue.empty() && !any_staging()); });
                if (ready.empty())
                    return;
                batch.push_back(&jobs.at(ready.front()));
                ready.pop_front();
                if (maxBatch > 1 && batchable(*batch[0])) {
                    size_t pts = points(*batch[0]);
                    for (auto it = ready.begin(); it != ready.end() && (int)batch.size() < maxBatch;) {
                        Job &o = jobs.at(*it);
                        if (batchable(o) && pts + points(o) <= kMaxBatchPoints) {
                            pts += points(o);
                            batch.push_back(&o);
                            it = ready.erase(it);
                        } else {
                            ++it;
                        }
                    }
                }
                const auto now = Clock::now();
                for (Job *jb : batch) {
                    jb->state = ST_JOB_RUNNING;
                    jb->slot = slot;
                    jb->tStart = now;
                }
                ctxs[slot]->cancel.store(false);
            }
            stage.notify_all();  // ready places freed up
            if (batch.size() == 1)
                run(slot, *batch[0]);
            else
                run_batch(slot, batch);
            {
                std::lock_guard<std::mutex> lk(mu);
                ctxs[slot]->cancel.store(false);
                for (Job *jb : batch) {
                    if (jb->j.pair < 0 && jb->j.kind != ST_JOB_CULL) {
                        free_pair(jb->img);
                        used -= jb->bytes;
                    }
                    const bool cancelled = jb->status == ST_ERR_CANCELLED;
                    finish(*jb, jb->status == ST_OK ? ST_JOB_DONE : (cancelled ? ST_JOB_CANCELLED : ST_JOB_FAILED));
                }
            }
            stage.notify_all();
            work.notify_all();
        }
    }

    bool any_staging() const
    {
        for (const auto &kv : jobs)
            if (kv.second.staging)
                return true;
        return false;
    }

    void fill(const Job &jb, st_job_info *info)
    {
        if (!info)
            return;
        memset(info, 0, sizeof(*info));
        info->state = jb.state;
        info->status = jb.status;
        info->slot = jb.slot;
        const auto now = Clock::now();
        if (jb.state == ST_JOB_QUEUED) {
            info->queueMs = ms(jb.tSubmit, now);
        } else if (jb.slot < 0) {
            // ended before it reached a slot (cancelled, or its upload failed)
            info->queueMs = ms(jb.tSubmit, jb.tEnd);
            info->latencyMs = info->queueMs;
        } else {
            info->queueMs = ms(jb.tSubmit, jb.tStart);
            const auto end = jb.state == ST_JOB_RUNNING ? now : jb.tEnd;
            info->runMs = ms(jb.tStart, end);
            info->latencyMs = ms(jb.tSubmit, end);
        }
        info->patches = jb.patches;
        info->nComplex = jb.j.out.nComplex;
        info->nAmp = jb.j.out.nAmp;
        info->nAmpL = jb.j.out.nAmpL;
        info->nFail = jb.j.out.nFail;
        strncpy(info->error, jb.err.c_str(), sizeof(info->error) - 1);
    }
};

extern "C" int st_server_create(int32_t device, int32_t slots, uint64_t hbmBudget, st_server **out)
{
    if (!out || slots < 1 || slots > 8)
        return ST_ERR_ARG;
    *out = nullptr;
    st_server *s = new st_server();
    s->device = device;
    s->budget = hbmBudget;
    s->slotScratch = std::min<size_t>((size_t)12 << 30, hbmBudget / (2 * (size_t)slots));
    auto fail = [s](int rc) {
        for (st_ctx *o : s->ctxs)
            st_ctx_destroy(o);
        if (s->up)
            (void)hipStreamDestroy(s->up);
        delete s;
        return rc;
    };
    for (int k = 0; k < slots; k++) {
        st_ctx *c = nullptr;
        const int rc = st_ctx_create(device, &c);
        if (rc != ST_OK)
            return fail(rc);
        // the arena is charged in full up front and never grows past it
        c->scratchBudget = std::min(c->scratchBudget, s->slotScratch);
        c->arenaLimit = s->slotScratch;
        s->ctxs.push_back(c);
    }
    // each slot's arena is held at its charged size from the start, so no
    // job frees and reallocates it (a hipFree waits for the whole device,
    // every other slot included)
    for (st_ctx *c : s->ctxs) {
        try {
            if (hipSetDevice(device) != hipSuccess)
                return fail(ST_ERR_HIP);
            c->arena_reserve(s->slotScratch);
        } catch (const StError &e) {
            return fail(e.code);
        }
    }
    s->used = (uint64_t)slots * s->slotScratch;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&s->up, hipStreamNonBlocking) != hipSuccess)
        return fail(ST_ERR_HIP);
    for (int k = 0; k < slots; k++)
        s->workers.emplace_back([s, k]() { s->worker(k); });
    s->stagerThread = std::thread([s]() { s->stager(); });
    *out = s;
    return ST_OK;
}

extern "C" int st_server_destroy(st_server *s)
{
    if (!s)
        return ST_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->stage.notify_all();
    s->work.notify_all();
    s->stagerThread.join();
    // the stager has handed on or ended every job it took
    s->work.notify_all();
    for (auto &t : s->workers)
        t.join();
    for (st_ctx *c : s->ctxs)
        st_ctx_destroy(c);
    (void)hipSetDevice(s->device);
    for (auto &kv : s->pairs)
        free_pair(kv.second.p);
    (void)hipStreamDestroy(s->up);
    delete s;
    return ST_OK;
}

extern "C" int st_server_pair_upload(st_server *s, const void *img1, int32_t nr1, int32_t na1, const void *img2,
                                     int32_t nr2, int32_t na2, int32_t stype, int32_t *pair)
{
    if (!s || !img1 || !img2 || !pair || nr1 <= 0 || na1 <= 0 || nr2 <= 0 || na2 <= 0 ||
        (stype != ST_SAMPLE_C64 && stype != ST_SAMPLE_CI16))
        return ST_ERR_ARG;
    const size_t b1 = sample_bytes(stype) * (size_t)nr1 * na1, b2 = sample_bytes(stype) * (size_t)nr2 * na2;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->used + b1 + b2 > s->budget)
            return ST_ERR_NOMEM;
        s->used += b1 + b2;  // reserved while uploading
    }
    ResPair rp;
    rp.p.nr1 = nr1;
    rp.p.na1 = na1;
    rp.p.nr2 = nr2;
    rp.p.na2 = na2;
    rp.p.stype = stype;
    rp.bytes = b1 + b2;
    // a dedicated stream: the stager's uploads keep running meanwhile
    hipStream_t st = nullptr;
    bool ok = hipSetDevice(s->device) == hipSuccess && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess &&
              hipMalloc(&rp.p.d1, b1) == hipSuccess && hipMalloc(&rp.p.d2, b2) == hipSuccess &&
              hipMemcpyAsync(rp.p.d1, img1, b1, hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(rp.p.d2, img2, b2, hipMemcpyHostToDevice, st) == hipSuccess &&
              hipStreamSynchronize(st) == hipSuccess;
    if (st)
        (void)hipStreamDestroy(st);
    std::lock_guard<std::mutex> lk(s->mu);
    if (!ok) {
        free_pair(rp.p);
        s->used -= b1 + b2;
        return ST_ERR_HIP;
    }
    const int id = s->nextPair++;
    s->pairs[id] = rp;
    *pair = id;
    return ST_OK;
}

extern "C" int st_server_pair_release(st_server *s, int32_t pair)
{
    if (!s)
        return ST_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->pairs.find(pair);
    if (it == s->pairs.end() || it->second.leaving)
        return ST_ERR_ARG;
    // from here on st_submit refuses the pair, and no other release call
    // touches it: this call alone erases it
    it->second.leaving = true;
    s->done.wait(lk, [&]() { return it->second.refs == 0; });
    (void)hipSetDevice(s->device);
    free_pair(it->second.p);
    s->used -= it->second.bytes;
    s->pairs.erase(it);
    s->stage.notify_all();
    return ST_OK;
}

extern "C" int st_server_set_batch(st_server *s, int32_t maxJobs)
{
    if (!s || maxJobs < 1 || maxJobs > 64)
        return ST_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    s->maxBatch = maxJobs;
    return ST_OK;
}

extern "C" int st_server_pair_synth(st_server *s, const st_synth_spec *spec, int32_t *pair)
{
    if (!s || !spec || !pair || spec->nr <= 0 || spec->na <= 0)
        return ST_ERR_ARG;
    const size_t bytes = 16 * (size_t)spec->nr * spec->na;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->used + bytes > s->budget)
            return ST_ERR_NOMEM;
        s->used += bytes;
    }
    // generated on slot 0's context (waits for its running job), then
    // handed over to the server's pair table
    st_ctx *c = s->ctxs[0];
    int32_t pid = -1;
    int rc = st_synth_device(c, spec, &pid);
    ResPair rp;
    if (rc == ST_OK) {
        std::lock_guard<std::mutex> lk(c->mu);
        rp.p = c->pairs[pid];
        c->pairs.erase(pid);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (rc != ST_OK) {
        s->used -= bytes;
        return rc;
    }
    rp.bytes = bytes;
    const int id = s->nextPair++;
    s->pairs[id] = rp;
    *pair = id;
    return ST_OK;
}

extern "C" int st_submit(st_server *s, const st_job *job, int64_t *id)
{
    if (!s || !job || !id || (job->kind != ST_JOB_STRACK && job->kind != ST_JOB_STRACKW && job->kind != ST_JOB_CULL))
        return ST_ERR_ARG;
    if (job->kind == ST_JOB_CULL ? !job->cull
                                 : (!job->out.offR || !job->out.offA || !job->out.corr || !job->out.type))
        return ST_ERR_ARG;
    Job jb;
    jb.j = *job;
    jb.tSubmit = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop)
        return ST_ERR_ARG;
    if (job->kind == ST_JOB_CULL) {
        jb.j.pair = -1;
        const int64_t k = s->nextId++;
        s->jobs[k] = jb;
        s->ready.push_back(k);
        s->work.notify_one();
        *id = k;
        return ST_OK;
    }
    if (job->pair >= 0) {
        auto it = s->pairs.find(job->pair);
        if (it == s->pairs.end() || it->second.leaving)
            return ST_ERR_ARG;
    } else {
        if (!job->img1 || !job->img2 || job->nr1 <= 0 || job->na1 <= 0 || job->nr2 <= 0 || job->na2 <= 0 ||
            (job->sampleType != ST_SAMPLE_C64 && job->sampleType != ST_SAMPLE_CI16))
            return ST_ERR_ARG;
        jb.bytes = sample_bytes(job->sampleType) * ((size_t)job->nr1 * job->na1 + (size_t)job->nr2 * job->na2);
    }
    // a job that could never be admitted fails now instead of waiting forever
    uint64_t fixed = (uint64_t)s->ctxs.size() * s->slotScratch;
    for (auto &kv : s->pairs)
        fixed += kv.second.bytes;
    if (fixed + jb.bytes > s->budget)
        return ST_ERR_NOMEM;
    const int64_t k = s->nextId++;
    s->jobs[k] = jb;
    if (job->pair >= 0) {
        s->pairs[job->pair].refs++;
        s->ready.push_back(k);
        s->work.notify_one();
    } else {
        s->queue.push_back(k);
        s->stage.notify_one();
    }
    *id = k;
    return ST_OK;
}

extern "C" int st_poll(st_server *s, int64_t id, st_job_info *info)
{
    if (!s)
        return ST_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end())
        return ST_ERR_ARG;
    s->fill(it->second, info);
    return ST_OK;
}

extern "C" int st_wait(st_server *s, int64_t id, st_job_info *info)
{
    if (!s)
        return ST_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end())
        return ST_ERR_ARG;
    s->done.wait(lk, [&]() { return ended(it->second.state); });
    s->fill(it->second, info);
    return ST_OK;
}

extern "C" int st_cancel(st_server *s, int64_t id)
{
    if (!s)
        return ST_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end() || ended(it->second.state))
        return ST_ERR_ARG;
    Job &jb = it->second;
    jb.cancel = true;
    if (jb.state == ST_JOB_RUNNING) {
        // the slot's tracking call stops at its next stage or batch boundary
        s->ctxs[jb.slot]->cancel.store(true);
        return ST_OK;
    }
    if (jb.staging)
        return ST_OK;  // the stager ends it once the upload returns
    for (auto q = s->queue.begin(); q != s->queue.end(); ++q)
        if (*q == id) {
            s->queue.erase(q);
            break;
        }
    for (auto q = s->ready.begin(); q != s->ready.end(); ++q)
        if (*q == id) {
            s->ready.erase(q);
            if (jb.j.pair < 0 && jb.j.kind != ST_JOB_CULL) {  // staged images
                (void)hipSetDevice(s->device);
                free_pair(jb.img);
                s->used -= jb.bytes;
            }
            break;
        }
    jb.status = ST_ERR_CANCELLED;
    jb.err = "job cancelled";
    s->finish(jb, ST_JOB_CANCELLED);
    s->stage.notify_all();
    return ST_OK;
}

extern "C" int st_job_release(st_server *s, int64_t id)
{
    if (!s)
        return ST_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(id);
    if (it == s->jobs.end() || !ended(it->second.state))
        return ST_ERR_ARG;
    s->jobs.erase(it);
    return ST_OK;
}
