// hm_server.hip -- continuous batching of tracking jobs on one GPU.
//
// Jobs wait in a FIFO queue and are admitted into device-resident slots
// while a slot is free and the HBM budget allows.  Every tick advances all
// resident jobs together (one segmented launch per kernel per turn) and
// retires those that reached their turn count; freed memory and slots admit
// the next queued jobs at the following tick.  A job's tiling, reduction
// order and random stream depend on that job alone, so its result is the
// same alone and in any mix.
#include <hip/hip_runtime.h>

#include <chrono>
#include <deque>
#include <memory>

#include "hm_internal.hpp"

using namespace hm;

namespace {

struct Job {
    int64_t id = 0;
    hm_job spec{};
    std::vector<double> wake, res;   // copies of the caller's tables
    hm_bunch *b = nullptr;
    int slot = -1;
    size_t bytes = 0;
    double t_submit = 0, t_admit = 0;
    bool cancelled = false, failed_by_request = false;
};

size_t job_bytes(const hm_job &j) {
    const size_t n = (size_t)std::max<int64_t>(j.n, 1);
    const size_t nb = (size_t)std::max<int64_t>(j.nbins, 1);
    return 16 * n + 104 * nb + (size_t)(1 << 16) + 2 * (size_t)kMaxFFT * 16;
}

}  // namespace

struct hm_server {
    hm_ctx *ctx = nullptr;
    int64_t budget = 0;
    int max_slots = 0;
    int64_t next_id = 1;
    size_t used = 0;
    std::deque<Job> queue;
    std::vector<Job> resident;
    std::deque<hm_job_result> done;
    std::unique_ptr<Batch> batch;
    std::chrono::steady_clock::time_point t0;
    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

static void push_result(hm_server *s, const Job &j, hm_status st, const hm_result *r, double cz,
                        double cE) {
    hm_job_result out;
    std::memset(&out, 0, sizeof(out));
    out.job_id = j.id;
    out.status = st;
    if (r) out.result = *r;
    out.t_submit = j.t_submit;
    out.t_admit = j.t_admit;
    out.t_done = s->now();
    out.centroid_z = cz;
    out.centroid_E = cE;
    s->done.push_back(out);
}

static hm_status admit(hm_server *s) {
    std::vector<RunBunch> specs;
    std::vector<Job> admitted;
    while (!s->queue.empty() && (int)(s->resident.size() + admitted.size()) < s->max_slots) {
        Job &j = s->queue.front();
        if (s->used + j.bytes > (size_t)s->budget && !(s->resident.empty() && admitted.empty()))
            break;  // wait for memory (a job larger than the budget runs alone)
        Job job = std::move(j);
        s->queue.pop_front();
        hm_status st = hm_bunch_create(s->ctx, job.spec.n, &job.b);
        if (st == HM_OK)
            st = hm_bunch_generate(job.b, job.spec.particle_seed, 0.0, job.spec.sigma_z, 0.0,
                                   job.spec.sigma_E);
        if (st != HM_OK) {
            if (job.b) hm_bunch_destroy(job.b);
            push_result(s, job, st, nullptr, 0, 0);
            continue;
        }
        RunBunch q;
        q.b = job.b;
        q.p = job.spec.params;
        q.o = job.spec.opts;
        q.o.normals = nullptr;
        q.o.wake_table = job.wake.empty() ? nullptr : job.wake.data();
        q.o.resonators = job.res.empty() ? nullptr : job.res.data();
        q.nbins = job.spec.nbins;
        q.fail_setup = job.spec.fail_setup != 0;
        if (job.spec.fail_after_turns > 0 && job.spec.fail_after_turns < q.p.n_turns) {
            q.p.n_turns = job.spec.fail_after_turns;
            job.failed_by_request = true;
        }
        job.t_admit = s->now();
        s->used += job.bytes;
        specs.push_back(q);
        admitted.push_back(std::move(job));
    }
    if (specs.empty()) return HM_OK;
    std::vector<int> idx;
    if (s->batch->add(specs, &idx) == HM_OK) {
        for (size_t k = 0; k < admitted.size(); ++k) {
            admitted[k].slot = idx[k];
            s->resident.push_back(std::move(admitted[k]));
        }
        return HM_OK;
    }
    // Batch::add took nothing: admit the jobs one at a time, so that a job
    // whose set-up fails (device memory, a failure by request) fails alone
    for (size_t k = 0; k < admitted.size(); ++k) {
        Job &j = admitted[k];
        std::vector<int> one;
        const hm_status st = s->batch->add({specs[k]}, &one);
        if (st != HM_OK) {
            s->used -= j.bytes;
            hm_bunch_destroy(j.b);
            j.b = nullptr;
            push_result(s, j, st, nullptr, 0, 0);
            continue;
        }
        j.slot = one[0];
        s->resident.push_back(std::move(j));
    }
    return HM_OK;
}

static hm_status collect(hm_server *s) {
    std::vector<Job> keep;
    for (auto &j : s->resident) {
        RunBunch &q = s->batch->slot(j.slot);
        if (!q.finished) {
            keep.push_back(std::move(j));
            continue;
        }
        hm_result r;
        double cz = 0, cE = 0;
        hm_status st = s->batch->result(j.slot, &r, &cz, &cE);
        if (st == HM_OK && (j.cancelled || j.failed_by_request)) st = HM_ERR_CANCELLED;
        push_result(s, j, st, &r, cz, cE);
        s->batch->remove(j.slot);
        hm_bunch_destroy(j.b);
        s->used -= j.bytes;
    }
    s->resident = std::move(keep);
    return HM_OK;
}

extern "C" {

hm_status hm_server_create(hm_ctx *ctx, int64_t budget_bytes, int32_t max_slots, hm_server **out) {
    if (!ctx || !out || budget_bytes <= 0 || max_slots <= 0) return fail(HM_ERR_ARG, "bad arguments");
    auto s = std::make_unique<hm_server>();
    s->ctx = ctx;
    s->budget = budget_bytes;
    s->max_slots = max_slots;
    s->batch = std::make_unique<Batch>(ctx);
    s->t0 = std::chrono::steady_clock::now();
    *out = s.release();
    return HM_OK;
}

hm_status hm_server_destroy(hm_server *s) {
    if (!s) return HM_OK;
    std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
    for (auto &j : s->resident) hm_bunch_destroy(j.b);
    s->batch.reset();
    delete s;
    return HM_OK;
}

hm_status hm_server_submit(hm_server *s, const hm_job *job, int64_t *job_id) {
    if (!s || !job) return fail(HM_ERR_ARG, "bad arguments");
    if (job->opts.normals) return fail(HM_ERR_ARG, "served jobs draw their own normals");
    if (job->opts.monitor) return fail(HM_ERR_ARG, "served jobs keep no turn monitor");
    HM_TRY(validate(&job->params, &job->opts, job->n, job->nbins));
    if (job->n == 0) return fail(HM_ERR_ARG, "served jobs need particles");
    std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
    Job j;
    j.id = s->next_id++;
    j.spec = *job;
    if (job->opts.wake_table && job->params.use_wakefield)
        j.wake.assign(job->opts.wake_table, job->opts.wake_table + job->nbins);
    if (job->opts.n_resonators > 0)
        j.res.assign(job->opts.resonators, job->opts.resonators + 3 * job->opts.n_resonators);
    j.spec.opts.wake_table = nullptr;
    j.spec.opts.resonators = nullptr;
    j.bytes = job_bytes(*job);
    j.t_submit = s->now();
    if (job_id) *job_id = j.id;
    s->queue.push_back(std::move(j));
    return HM_OK;
}

hm_status hm_server_cancel(hm_server *s, int64_t job_id) {
    if (!s) return fail(HM_ERR_ARG, "server is NULL");
    std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
    for (auto it = s->queue.begin(); it != s->queue.end(); ++it) {
        if (it->id == job_id) {
            push_result(s, *it, HM_ERR_CANCELLED, nullptr, 0, 0);
            s->queue.erase(it);
            return HM_OK;
        }
    }
    for (auto &j : s->resident) {
        if (j.id == job_id) {
            j.cancelled = true;
            HM_TRY(s->batch->finish_now(j.slot));
            return collect(s);
        }
    }
    return fail(HM_ERR_ARG, "no queued or running job %lld", (long long)job_id);
}

hm_status hm_server_tick(hm_server *s, int32_t turns, int32_t *active) {
    if (!s || turns < 0) return fail(HM_ERR_ARG, "bad arguments");
    std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
    HM_TRY(set_device(s->ctx));
    HM_TRY(admit(s));
    if (!s->resident.empty()) HM_TRY(s->batch->advance(turns));
    HM_TRY(collect(s));
    if (active) *active = (int32_t)(s->queue.size() + s->resident.size());
    return HM_OK;
}

hm_status hm_server_poll(hm_server *s, hm_job_result *out, int32_t cap, int32_t *count) {
    if (!s || (cap > 0 && !out) || !count) return fail(HM_ERR_ARG, "bad arguments");
    std::lock_guard<std::recursive_mutex> lk(s->ctx->mu);
    int32_t c = 0;
    while (c < cap && !s->done.empty()) {
        out[c++] = s->done.front();
        s->done.pop_front();
    }
    *count = c;
    return HM_OK;
}

}  // extern "C"
