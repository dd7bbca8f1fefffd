// Serving: continuous batching of tracking jobs on one GPU.
//
// A server owns `slots` contexts (one stream, FFT plans and scratch arena
// each) and one worker thread per slot.  Submitted jobs wait in a queue;
// an idle worker takes the first queued job whose device memory fits in
// what the budget has left, so small jobs overtake a large one that must
// wait for memory, and jobs that arrive while others run start as soon as
// a slot frees up.  Kernels of different slots overlap on the device (all
// streams at the default priority).  A job's results depend only on its
// inputs: every slot runs the same code on its own stream and scratch.
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
    size_t bytes = 0;  // admission charge: per-job image upload
    int state = ST_JOB_QUEUED, status = ST_OK, slot = -1;
    Clock::time_point tSubmit, tStart, tEnd;
    int64_t patches = 0;
    std::string err;
};

struct ResPair {
    StPair p;
    size_t bytes = 0;
    int refs = 0;
};

double ms(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

size_t sample_bytes(int stype) { return stype == ST_SAMPLE_C64 ? 8 : 4; }

}  // namespace

struct st_server {
    int device = 0;
    uint64_t budget = 0;
    size_t slotScratch = 0;
    uint64_t used = 0;  // slot arenas + resident pairs + running jobs' images
    std::vector<st_ctx *> ctxs;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable work;  // queue changed, memory freed, stop
    std::condition_variable done;  // a job finished
    std::deque<int64_t> queue;
    std::map<int64_t, Job> jobs;
    std::map<int, ResPair> pairs;
    int nextPair = 1;
    int64_t nextId = 1;
    bool stop = false;
    hipStream_t up = nullptr;

    Job *pick()
    {
        for (auto it = queue.begin(); it != queue.end(); ++it) {
            Job &jb = jobs[*it];
            if (used + jb.bytes <= budget) {
                queue.erase(it);
                return &jb;
            }
        }
        return nullptr;
    }

    void run(int slot, Job &jb)
    {
        st_ctx *c = ctxs[slot];
        StPair pr;
        int pid = -1;
        int rc = ST_OK;
        if (jb.j.pair >= 0) {
            std::lock_guard<std::mutex> lk(mu);
            pr = pairs[jb.j.pair].p;
        } else {
            rc = st_pair_upload(c, jb.j.img1, jb.j.nr1, jb.j.na1, jb.j.img2, jb.j.nr2, jb.j.na2, jb.j.sampleType,
                                &pid);
            if (rc == ST_OK)
                pr = c->pairs[pid];
        }
        if (rc == ST_OK) {
            rc = st_guard(c, [&]() {
                if (jb.j.kind == ST_JOB_STRACK)
                    st_speckle_track_impl(c, pr, jb.j.par, &jb.j.out, nullptr);
                else
                    st_corr_track_fast_impl(c, pr, jb.j.par, &jb.j.out, nullptr);
                jb.patches = (int64_t)c->stats.nCorr[0] + c->stats.nCorr[1] + c->stats.nCorr[2];
            });
        }
        if (rc == ST_OK && jb.j.cull)
            rc = st_cull_run(c, jb.j.cull);
        if (rc != ST_OK)
            jb.err = c->err;
        if (pid >= 0)
            st_pair_release(c, pid);
        jb.status = rc;
    }

    void worker(int slot)
    {
        for (;;) {
            Job *jb = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu);
                work.wait(lk, [&]() {
                    jb = pick();
                    return jb != nullptr || (stop && queue.empty());
                });
                if (!jb)
                    return;
                jb->state = ST_JOB_RUNNING;
                jb->slot = slot;
                jb->tStart = Clock::now();
                used += jb->bytes;
                if (jb->j.pair >= 0)
                    pairs[jb->j.pair].refs++;
            }
            run(slot, *jb);
            {
                std::lock_guard<std::mutex> lk(mu);
                jb->tEnd = Clock::now();
                jb->state = jb->status == ST_OK ? ST_JOB_DONE : ST_JOB_FAILED;
                used -= jb->bytes;
                if (jb->j.pair >= 0)
                    pairs[jb->j.pair].refs--;
            }
            done.notify_all();
            work.notify_all();
        }
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
    for (int k = 0; k < slots; k++) {
        st_ctx *c = nullptr;
        const int rc = st_ctx_create(device, &c);
        if (rc != ST_OK) {
            for (st_ctx *o : s->ctxs)
                st_ctx_destroy(o);
            delete s;
            return rc;
        }
        c->scratchBudget = s->slotScratch;
        s->ctxs.push_back(c);
    }
    s->used = (uint64_t)slots * s->slotScratch;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&s->up, hipStreamNonBlocking) != hipSuccess) {
        for (st_ctx *o : s->ctxs)
            st_ctx_destroy(o);
        delete s;
        return ST_ERR_HIP;
    }
    for (int k = 0; k < slots; k++)
        s->workers.emplace_back([s, k]() { s->worker(k); });
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
    s->work.notify_all();
    for (auto &t : s->workers)
        t.join();
    for (st_ctx *c : s->ctxs)
        st_ctx_destroy(c);
    hipSetDevice(s->device);
    for (auto &kv : s->pairs) {
        hipFree(kv.second.p.d1);
        hipFree(kv.second.p.d2);
    }
    hipStreamDestroy(s->up);
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
    bool ok = hipSetDevice(s->device) == hipSuccess && hipMalloc(&rp.p.d1, b1) == hipSuccess &&
              hipMalloc(&rp.p.d2, b2) == hipSuccess &&
              hipMemcpyAsync(rp.p.d1, img1, b1, hipMemcpyHostToDevice, s->up) == hipSuccess &&
              hipMemcpyAsync(rp.p.d2, img2, b2, hipMemcpyHostToDevice, s->up) == hipSuccess &&
              hipStreamSynchronize(s->up) == hipSuccess;
    std::lock_guard<std::mutex> lk(s->mu);
    if (!ok) {
        if (rp.p.d1)
            hipFree(rp.p.d1);
        if (rp.p.d2)
            hipFree(rp.p.d2);
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
    if (it == s->pairs.end())
        return ST_ERR_ARG;
    // queued jobs that name the pair would fail later: refuse while any exist
    for (int64_t id : s->queue)
        if (s->jobs[id].j.pair == pair)
            return ST_ERR_ARG;
    s->done.wait(lk, [&]() { return s->pairs[pair].refs == 0; });
    hipSetDevice(s->device);
    hipFree(it->second.p.d1);
    hipFree(it->second.p.d2);
    s->used -= it->second.bytes;
    s->pairs.erase(it);
    s->work.notify_all();
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
    if (!s || !job || !id || (job->kind != ST_JOB_STRACK && job->kind != ST_JOB_STRACKW))
        return ST_ERR_ARG;
    if (!job->out.offR || !job->out.offA || !job->out.corr || !job->out.type)
        return ST_ERR_ARG;
    Job jb;
    jb.j = *job;
    jb.tSubmit = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    if (job->pair >= 0) {
        if (!s->pairs.count(job->pair))
            return ST_ERR_ARG;
    } else {
        if (!job->img1 || !job->img2 || job->nr1 <= 0 || job->na1 <= 0 || job->nr2 <= 0 || job->na2 <= 0)
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
    s->queue.push_back(k);
    *id = k;
    s->work.notify_one();
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
    s->done.wait(lk, [&]() { return it->second.state == ST_JOB_DONE || it->second.state == ST_JOB_FAILED; });
    s->fill(it->second, info);
    return ST_OK;
}
