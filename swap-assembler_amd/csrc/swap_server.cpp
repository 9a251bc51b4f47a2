// Many assemblies on one GPU: a job queue served by a fixed number of slots.
//
// Each slot is a host thread that owns one context (its own HIP stream and
// grow-only device buffers) and runs whole assemblies on it: build, merge,
// output formatting.  The slots' kernels overlap on the device and each slot's
// host work (order replay, writers) overlaps the others' device work, so
// small, uneven samples keep the GPU busy where one context alone would leave
// it idle between launches and host synchronisations.
//
// Admission: a job starts when a slot is free and the estimated device
// footprint of the running jobs plus its own fits the HBM budget (a job alone
// is always admitted).  Jobs start in submission order.  Results are the same
// as running each job alone on a fresh context: a context keeps no state from
// one job to the next except buffer capacity.
//
// Fused submissions (swap_server_submit_fused): jobs of one k and c are cut
// into groups of up to a byte budget; a builder thread builds each group with
// one swap_build_batch on one of two batch contexts (group g + 1 is built
// while the slots merge group g), then queues the group's jobs, and the slot
// that runs a job takes its sample from the batch context (swap_batch_take)
// instead of building it.  A batch context is built again only after every
// job of its previous group has taken its sample.  A group whose batch build
// fails (a sample that is not plain line-pair FASTA, too many survivors to
// order) runs as ordinary jobs.
//
// The reference assembles one sample per run and has no counterpart of this
// module.
#include "swap_hip.h"
#include "swap_common.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

namespace swapk {
int format_file_string(swap_ctx *c, int which, std::string &out);
}

namespace {

using clk = std::chrono::steady_clock;

// Device bytes per FASTA byte a running job may hold: two 12-byte occurrence
// records per input k-mer (about one k-mer per input byte for short-read
// FASTA), survivor arrays reserved at the occurrence count, the packed input
// and the merge arrays.
constexpr double kBytesPerInputByte = 36.0;

struct JobRec {
    swap_job spec{};
    int group = -1;   // fused group, -1 for an ordinary job
    size_t bidx = 0;  // its sample in the group's batch build
    std::string files[7];
    int status = SWAP_JOB_QUEUED;
    std::string err;
    swap_stats stats{};
    double t_submit = 0, t_start = 0, t_end = 0;
    size_t est = 0;
};

}  // namespace

// One group of a fused submission.
struct FusedGroup {
    std::vector<int64_t> jobs;
    swap_ctx *bctx = nullptr;  // the batch context it was built on (null: ordinary jobs)
    size_t taken = 0;          // jobs that have taken (or failed to take) their sample
};

struct swap_server {
    int device = 0;
    int slots = 1;
    size_t budget = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<int64_t> queue;
    std::vector<std::unique_ptr<JobRec>> jobs;
    size_t in_flight = 0;
    int running = 0;
    bool stopping = false;
    std::vector<std::thread> workers;
    clk::time_point t0 = clk::now();
    // fused submissions
    std::vector<std::unique_ptr<FusedGroup>> groups;
    std::deque<size_t> group_queue;  // groups waiting for their batch build
    size_t pending_fused = 0;        // jobs of groups not yet queued
    std::condition_variable cv_fuse;
    std::thread fuser;
    swap_ctx *bctx[2] = {nullptr, nullptr};
    int64_t bctx_group[2] = {-1, -1};  // the last group built on each batch context

    double now_ms() const { return std::chrono::duration<double, std::milli>(clk::now() - t0).count(); }
};

namespace {

int run_job(swap_server *s, swap_ctx *&ctx, JobRec &j, std::string &err) {
    swap_ctx *batch = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (j.group >= 0) batch = s->groups[j.group]->bctx;
    }
    // a job of a built group counts as taken once it no longer reads the
    // batch context, whatever happens to it
    auto taken = [&] {
        if (!batch) return;
        std::lock_guard<std::mutex> lk(s->mu);
        s->groups[j.group]->taken++;
        s->cv_fuse.notify_all();
    };
    swap_cfg cfg;
    cfg.k = j.spec.k;
    cfg.c = j.spec.c;
    cfg.device = s->device;
    cfg.flags = j.spec.flags;
    int rc = ctx ? swap_reconfigure(ctx, &cfg) : swap_create(&cfg, &ctx);
    if (rc) {
        taken();
        err = ctx ? swap_last_error(ctx) : swap_strerror(rc);
        return rc;
    }
    if (batch) {
        rc = swap_batch_take(batch, j.bidx, ctx);
        taken();
    } else if (j.spec.job_flags & SWAP_JOB_DEVICE_INPUT) {
        rc = swap_build_device(ctx, j.spec.fasta, j.spec.n);
    } else {
        rc = swap_build(ctx, static_cast<const char *>(j.spec.fasta), j.spec.n);
    }
    if (rc == SWAP_OK && !(j.spec.flags & SWAP_FLAG_KMER_GRAPH)) rc = swap_merge(ctx);
    for (int w = 0; w < 7 && rc == SWAP_OK; w++)
        if (j.spec.outputs & (1u << w)) rc = swapk::format_file_string(ctx, w, j.files[w]);
    if (rc == SWAP_OK) rc = swap_get_stats(ctx, &j.stats);
    if (rc) err = swap_last_error(ctx);
    return rc;
}

void worker(swap_server *s) {
    swap_ctx *ctx = nullptr;
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_work.wait(lk, [&] {
            if (s->queue.empty()) return s->stopping && s->pending_fused == 0;
            const JobRec &j = *s->jobs[s->queue.front()];
            return s->running == 0 || s->in_flight + j.est <= s->budget;
        });
        if (s->queue.empty()) break;
        const int64_t id = s->queue.front();
        s->queue.pop_front();
        JobRec &j = *s->jobs[id];
        j.status = SWAP_JOB_RUNNING;
        j.t_start = s->now_ms();
        s->in_flight += j.est;
        s->running++;
        lk.unlock();
        std::string err;
        const int rc = run_job(s, ctx, j, err);
        lk.lock();
        j.t_end = s->now_ms();
        j.err = err;
        j.status = rc;
        s->in_flight -= j.est;
        s->running--;
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    }
    lk.unlock();
    if (ctx) swap_destroy(ctx);
}

// The builder thread of fused submissions.
void fuse_loop(swap_server *s) {
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_fuse.wait(lk, [&] { return !s->group_queue.empty() || s->stopping; });
        if (s->group_queue.empty()) break;
        const size_t g = s->group_queue.front();
        s->group_queue.pop_front();
        FusedGroup &grp = *s->groups[g];
        const JobRec &j0 = *s->jobs[grp.jobs[0]];
        // the batch context whose last group has been taken first
        const int b = (s->bctx_group[0] <= s->bctx_group[1]) ? 0 : 1;
        s->cv_fuse.wait(lk, [&] {
            const int64_t pg = s->bctx_group[b];
            return pg < 0 || s->groups[pg]->taken == s->groups[pg]->jobs.size();
        });
        std::vector<const void *> ptr;
        std::vector<size_t> len;
        for (int64_t id : grp.jobs) {
            ptr.push_back(s->jobs[id]->spec.fasta);
            len.push_back(s->jobs[id]->spec.n);
        }
        swap_cfg cfg;
        cfg.k = j0.spec.k;
        cfg.c = j0.spec.c;
        cfg.device = s->device;
        cfg.flags = j0.spec.flags;
        swap_ctx *&bc = s->bctx[b];
        lk.unlock();
        int rc = bc ? swap_reconfigure(bc, &cfg) : swap_create(&cfg, &bc);
        if (rc == SWAP_OK) rc = swap_build_batch(bc, ptr.data(), len.data(), ptr.size());
        lk.lock();
        if (rc == SWAP_OK) {
            grp.bctx = bc;
            s->bctx_group[b] = (int64_t)g;
        }
        // queued: with the batch context, or as ordinary jobs
        for (size_t i = 0; i < grp.jobs.size(); i++) {
            JobRec &j = *s->jobs[grp.jobs[i]];
            if (rc != SWAP_OK) j.group = -1;
            s->queue.push_back(grp.jobs[i]);
        }
        s->pending_fused -= grp.jobs.size();
        s->cv_work.notify_all();
        s->cv_done.notify_all();
    }
}

}  // namespace

extern "C" {

int swap_server_create(int device, int slots, size_t hbm_budget, swap_server **out) {
    if (!out || slots < 1 || slots > 64) return SWAP_E_INVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return SWAP_E_NOGPU;
    if (device < 0 || device >= ndev) return SWAP_E_INVAL;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return SWAP_E_NOGPU;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return SWAP_E_NOGPU;
    if (hbm_budget == 0) {
        size_t free_b = 0, total_b = 0;
        if (hipSetDevice(device) != hipSuccess || hipMemGetInfo(&free_b, &total_b) != hipSuccess)
            return SWAP_E_DEVICE;
        hbm_budget = free_b / 10 * 8;
    }
    swap_server *s = new swap_server();
    s->device = device;
    s->slots = slots;
    s->budget = hbm_budget;
    for (int i = 0; i < slots; i++) s->workers.emplace_back(worker, s);
    *out = s;
    return SWAP_OK;
}

int64_t swap_server_submit(swap_server *s, const swap_job *job) {
    if (!s || !job || (!job->fasta && job->n)) return SWAP_E_INVAL;
    if (job->k < 9 || job->k > 31 || (job->k % 2) == 0) return SWAP_E_INVAL;
    if (job->flags & SWAP_FLAG_KMER_GRAPH) {
        if (job->outputs & ~(1u << SWAP_FILE_KMER_GRAPH)) return SWAP_E_INVAL;
    } else if ((job->outputs & (1u << SWAP_FILE_KMER_GRAPH)) || job->k < swapk::kMergeMinK) {
        return SWAP_E_INVAL;  // a merge job needs k >= 19
    }
    auto j = std::make_unique<JobRec>();
    j->spec = *job;
    j->est = (size_t)(kBytesPerInputByte * (double)job->n);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) return SWAP_E_STATE;
    j->t_submit = s->now_ms();
    const int64_t id = (int64_t)s->jobs.size();
    s->jobs.push_back(std::move(j));
    s->queue.push_back(id);
    s->cv_work.notify_one();
    return id;
}

int swap_server_submit_batch(swap_server *s, const swap_job *jobs, size_t n, int64_t *ids) {
    if (!s || (!jobs && n) || (!ids && n)) return SWAP_E_INVAL;
    // largest input first: the long jobs start early and the short ones fill in
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), (size_t)0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return jobs[a].n > jobs[b].n; });
    for (size_t i : order) {
        const int64_t id = swap_server_submit(s, &jobs[i]);
        if (id < 0) return (int)id;
        ids[i] = id;
    }
    return SWAP_OK;
}

int swap_server_submit_fused(swap_server *s, const swap_job *jobs, size_t n, size_t group_bytes, int64_t *ids) {
    if (!s || (!jobs && n) || (!ids && n)) return SWAP_E_INVAL;
    if (n == 0) return SWAP_OK;
    // 256 MiB groups measured best on the 256-sample workload (128 MiB: more,
    // less efficient builds; 512 MiB: builds of wider partition layouts that
    // overlap fewer merges); above 512 MiB a group's survivors can exceed the
    // first-occurrence ordering's bins and the group would run as ordinary jobs
    if (group_bytes == 0) group_bytes = (size_t)256 << 20;
    group_bytes = std::min(group_bytes, (size_t)512 << 20);
    for (size_t i = 0; i < n; i++) {
        const swap_job &j = jobs[i];
        if (!j.fasta || j.n == 0 || !(j.job_flags & SWAP_JOB_DEVICE_INPUT)) return SWAP_E_INVAL;
        if (j.k != jobs[0].k || j.c != jobs[0].c || j.flags != jobs[0].flags) return SWAP_E_INVAL;
        if (j.flags & (SWAP_FLAG_KMER_GRAPH | SWAP_FLAG_HASH_ORDER)) return SWAP_E_INVAL;
        if (j.k < swapk::kMergeMinK || j.k > 31 || (j.k % 2) == 0) return SWAP_E_INVAL;
    }
    // largest input first, cut into groups of at most group_bytes (and the
    // batch build's sample limit); an input above the budget is a group alone
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), (size_t)0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return jobs[a].n > jobs[b].n; });
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) return SWAP_E_STATE;
    if (!s->fuser.joinable()) s->fuser = std::thread(fuse_loop, s);
    std::unique_ptr<FusedGroup> grp;
    size_t bytes = 0;
    auto close_group = [&] {
        if (!grp) return;
        s->group_queue.push_back(s->groups.size());
        s->groups.push_back(std::move(grp));
        bytes = 0;
    };
    for (size_t i : order) {
        if (grp && (bytes + jobs[i].n > group_bytes || grp->jobs.size() == swapk::kMaxBatchSamples)) close_group();
        if (!grp) grp.reset(new FusedGroup());
        auto j = std::make_unique<JobRec>();
        j->spec = jobs[i];
        j->est = (size_t)(kBytesPerInputByte * (double)jobs[i].n);
        j->t_submit = s->now_ms();
        j->group = (int)s->groups.size();
        j->bidx = grp->jobs.size();
        const int64_t id = (int64_t)s->jobs.size();
        s->jobs.push_back(std::move(j));
        grp->jobs.push_back(id);
        bytes += jobs[i].n;
        ids[i] = id;
        s->pending_fused++;
    }
    close_group();
    s->cv_fuse.notify_all();
    return SWAP_OK;
}

int swap_server_wait(swap_server *s, int64_t id) {
    if (!s) return SWAP_E_INVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    if (id < 0 || id >= (int64_t)s->jobs.size()) return SWAP_E_INVAL;
    JobRec &j = *s->jobs[id];
    s->cv_done.wait(lk, [&] { return j.status != SWAP_JOB_QUEUED && j.status != SWAP_JOB_RUNNING; });
    return j.status;
}

int swap_server_wait_all(swap_server *s) {
    if (!s) return SWAP_E_INVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->queue.empty() && s->running == 0 && s->pending_fused == 0; });
    for (const auto &j : s->jobs)
        if (j->status < 0) return j->status;
    return SWAP_OK;
}

int swap_server_job_info(swap_server *s, int64_t id, swap_job_info *info) {
    if (!s || !info) return SWAP_E_INVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    if (id < 0 || id >= (int64_t)s->jobs.size()) return SWAP_E_INVAL;
    const JobRec &j = *s->jobs[id];
    info->status = j.status;
    info->submit_ms = j.t_submit;
    info->start_ms = j.t_start;
    info->end_ms = j.t_end;
    info->stats = j.stats;
    return SWAP_OK;
}

const char *swap_server_job_error(swap_server *s, int64_t id) {
    if (!s) return "";
    std::lock_guard<std::mutex> lk(s->mu);
    if (id < 0 || id >= (int64_t)s->jobs.size()) return "";
    return s->jobs[id]->err.c_str();
}

int64_t swap_server_output(swap_server *s, int64_t id, int which, char *buf, size_t cap) {
    if (!s || which < 0 || which > 6) return SWAP_E_INVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    if (id < 0 || id >= (int64_t)s->jobs.size()) return SWAP_E_INVAL;
    const JobRec &j = *s->jobs[id];
    if (j.status != SWAP_OK) return j.status < 0 ? j.status : SWAP_E_STATE;
    if (!(j.spec.outputs & (1u << which))) return SWAP_E_STATE;
    const std::string &f = j.files[which];
    if (buf) {
        if (cap < f.size()) return SWAP_E_RANGE;
        std::memcpy(buf, f.data(), f.size());
    }
    return (int64_t)f.size();
}

int swap_server_release(swap_server *s, int64_t id) {
    if (!s) return SWAP_E_INVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    if (id < 0 || id >= (int64_t)s->jobs.size()) return SWAP_E_INVAL;
    JobRec &j = *s->jobs[id];
    if (j.status == SWAP_JOB_QUEUED || j.status == SWAP_JOB_RUNNING) return SWAP_E_STATE;
    for (auto &f : j.files) std::string().swap(f);
    return SWAP_OK;
}

void swap_server_destroy(swap_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->cv_work.notify_all();
    s->cv_fuse.notify_all();
    if (s->fuser.joinable()) s->fuser.join();
    for (auto &t : s->workers) t.join();
    for (swap_ctx *c : s->bctx)
        if (c) swap_destroy(c);
    delete s;
}

}  // extern "C"
