// C ABI of the k-mer graph path (declared in include/swap_hip.h) and the host
// writers of the reference's output formats.
#include "swap_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <deque>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "swap_device.h"
#include "swap_host.h"

// Host-to-device staging of FASTA bytes: two pinned buffers; while the DMA
// engine copies one to the device, host threads fill the other, so the
// transfer runs at the rate of the slower of the two instead of their sum.
struct Staging {
    static constexpr size_t kChunk = 32u << 20;
    static constexpr int kThreads = 4;  // host copy threads per chunk
    char *buf[2] = {nullptr, nullptr};
    hipEvent_t done[2] = {nullptr, nullptr};
    bool ready = false;

    hipError_t init() {
        if (ready) return hipSuccess;
        for (int b = 0; b < 2; b++) {
            hipError_t e = hipHostMalloc((void **)&buf[b], kChunk, hipHostMallocDefault);
            if (e != hipSuccess) return e;
            e = hipEventCreateWithFlags(&done[b], hipEventDisableTiming);
            if (e != hipSuccess) return e;
        }
        ready = true;
        return hipSuccess;
    }
    ~Staging() {
        for (int b = 0; b < 2; b++) {
            if (buf[b]) (void)hipHostFree(buf[b]);
            if (done[b]) (void)hipEventDestroy(done[b]);
        }
    }
    // copies n host bytes to dst on stream st (asynchronous with respect to the
    // host once the last chunk is queued)
    hipError_t copy(void *dst, const char *src, size_t n, hipStream_t st) {
        hipError_t e = init();
        if (e != hipSuccess) return e;
        for (int b = 0; b < 2; b++)  // a previous call's copies out of the buffers have ended
            if ((e = hipEventSynchronize(done[b])) != hipSuccess) return e;
        for (size_t off = 0, i = 0; off < n; off += kChunk, i++) {
            const size_t len = std::min(kChunk, n - off);
            const int b = (int)(i & 1);
            if (i >= 2 && (e = hipEventSynchronize(done[b])) != hipSuccess) return e;
            const size_t part = (len + kThreads - 1) / kThreads;
            std::vector<std::thread> team;
            for (int t = 1; t < kThreads; t++) {
                const size_t lo = std::min(len, part * (size_t)t), hi = std::min(len, lo + part);
                if (lo < hi) team.emplace_back([=] { std::memcpy(buf[b] + lo, src + off + lo, hi - lo); });
            }
            std::memcpy(buf[b], src + off, std::min(len, part));
            for (auto &th : team) th.join();
            if ((e = hipMemcpyAsync(static_cast<char *>(dst) + off, buf[b], len, hipMemcpyHostToDevice, st)) !=
                hipSuccess)
                return e;
            if ((e = hipEventRecord(done[b], st)) != hipSuccess) return e;
        }
        return hipSuccess;
    }
};

// Read-ahead of the next input: swap_prefetch copies host bytes to one of two
// device buffers on a copy stream of its own, driven by a host thread (the
// staging loop above), while the caller's build of the previous input runs
// on the context's stream.  Builds consume prefetches in the order they were
// made (same pointer and size), so a stream of inputs costs the longer of the
// transfer and the device path per input instead of their sum.
struct Prefetch {
    const char *src = nullptr;
    size_t n = 0;
    int buf = 0;        // which of the context's two read-ahead buffers
    std::thread th;     // the staging loop
    hipError_t rc = hipSuccess;
};

struct swap_ctx {
    swap_cfg cfg;
    std::unique_ptr<swapk::Pipeline> pipe;
    swapk::DevBuf input;
    Staging staging;
    // read-ahead (at most two inputs prefetched and not yet built)
    swapk::DevBuf ahead[2];
    hipStream_t copy_st = nullptr;
    hipEvent_t ahead_done[2] = {nullptr, nullptr};
    std::deque<std::unique_ptr<Prefetch>> fifo;
    int next_buf = 0;
    bool built = false, merged = false;
    double build_ms = 0, merge_ms = 0;
    std::string err;
    // host copies, filled on demand
    bool have_graph = false;
    swapk::HostKmerGraph graph;     // sorted by key
    std::vector<uint32_t> visit;    // survivor index -> merge visit rank
    bool have_compacted = false;
    swapk::HostCompacted comp;
    bool have_ce = false;
    int ce_rc = 0;
    swapk::CeFiles ce;
    bool from_static = false;  // the last build ran in the static scratch (swap_assemble_static)
    std::vector<uint64_t> batch_off;  // the samples' byte offsets of the last batch build, empty if none
    bool tune_fixed = false;   // swap_set_tuning: the table is not consulted

    // waits for every staging loop still running (the staging buffers and the
    // copy stream are shared by them); returns the first error
    hipError_t join_prefetches() {
        hipError_t e = hipSuccess;
        for (auto &p : fifo)
            if (p->th.joinable()) {
                p->th.join();
                if (p->rc != hipSuccess && e == hipSuccess) e = p->rc;
            }
        return e;
    }
    ~swap_ctx() {
        (void)join_prefetches();
        if (copy_st) (void)hipStreamSynchronize(copy_st);
        for (auto &e : ahead_done)
            if (e) (void)hipEventDestroy(e);
        if (copy_st) (void)hipStreamDestroy(copy_st);
    }
};

namespace {

int set_err(swap_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

using clk = std::chrono::steady_clock;

int ensure_graph(swap_ctx *c) {
    if (c->have_graph) return 0;
    if (!c->built) return set_err(c, SWAP_E_STATE, "no build");
    swapk::HostKmerGraph g;
    int rc = c->pipe->export_kmer_graph(g);
    if (rc) return set_err(c, rc, c->pipe->last_error());
    // sort by key (the device emits survivors in partition order)
    std::vector<uint32_t> idx(g.key.size());
    std::iota(idx.begin(), idx.end(), 0u);
    std::sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return g.key[a] < g.key[b]; });
    c->graph.key.resize(idx.size());
    c->graph.mult.resize(idx.size());
    c->graph.first.resize(g.first.empty() ? 0 : idx.size());
    for (size_t i = 0; i < idx.size(); i++) {
        c->graph.key[i] = g.key[idx[i]];
        c->graph.mult[i] = g.mult[idx[i]];
        if (!g.first.empty()) c->graph.first[i] = g.first[idx[i]];
    }
    c->graph.epoch_first = g.epoch_first;
    c->graph.distinct = g.distinct;
    c->graph.tail_lookup = g.tail_lookup;
    c->have_graph = true;
    return 0;
}

// Output order of kmerGraph.txt: the reference's table order after the gather
// (reference order), else ascending key.
void kmer_graph_order(swap_ctx *c, std::vector<uint32_t> &order) {
    order.clear();
    if (c->graph.first.empty()) return;
    uint64_t buckets = 0;
    std::vector<uint64_t> first(c->graph.first.begin(), c->graph.first.end());
    std::vector<uint32_t> visit = swapk::visit_order(c->graph.key, first, c->graph.epoch_first, c->graph.distinct,
                                                     c->graph.tail_lookup, &buckets);
    std::vector<uint32_t> pos(visit.size());
    for (size_t i = 0; i < visit.size(); i++) pos[visit[i]] = (uint32_t)i;
    order = swapk::gather_order(c->graph.key, pos, buckets);
}

}  // namespace

namespace swapk {

// Output file bytes of a context (the body of swap_format_file, without the
// caller-buffer copy).
int format_file_string(swap_ctx *c, int which, std::string &out) {
    out.clear();
    if (!c) return SWAP_E_INVAL;
    int rc = 0;
    if (which == SWAP_FILE_KMER_GRAPH) {
        rc = ensure_graph(c);
        if (rc) return rc;
        std::vector<uint32_t> order;
        kmer_graph_order(c, order);
        format_kmer_graph(c->cfg.k, c->graph, order, out);
    } else if (which >= SWAP_FILE_CONTIG_GRAPH && which <= SWAP_FILE_JUNG_MUL) {
        if (!c->merged) return set_err(c, SWAP_E_STATE, "no merge");
        if (!c->have_compacted) {
            rc = materialize_merge(*c->pipe, c->cfg, c->comp);
            if (rc) return set_err(c, rc, c->pipe->last_error());
            c->have_compacted = true;
        }
        if (which == SWAP_FILE_CONTIG_GRAPH) {
            format_contig_graph(c->comp, out);
        } else if (which == SWAP_FILE_NOCE) {
            format_noce(c->cfg.k, c->comp.finished ? *c->comp.finished : c->comp, out);
        } else if (which == SWAP_FILE_PERF_CONTIGS) {
            format_perf_contigs(c->cfg.k, c->comp, out);
        } else {
            if (!c->have_ce) {
                c->ce_rc = contig_extension(c->cfg.k, c->comp, c->ce);
                c->have_ce = true;
            }
            if (c->ce_rc)
                return set_err(c, c->ce_rc,
                               "contig extension: no edge base left to average the coverage over (the reference "
                               "divides by zero here)");
            out = which == SWAP_FILE_CE_CONTIGS ? c->ce.ce : which == SWAP_FILE_JUNG_ARC ? c->ce.jung_arc : c->ce.jung_mul;
        }
    } else {
        return set_err(c, SWAP_E_INVAL, "unknown file");
    }
    return 0;
}

}  // namespace swapk

extern "C" {

int swap_create(const swap_cfg *cfg, swap_ctx **out) {
    if (!cfg || !out) return SWAP_E_INVAL;
    *out = nullptr;
    if (cfg->k < 9 || cfg->k > 31 || (cfg->k % 2) == 0) return SWAP_E_INVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return SWAP_E_NOGPU;
    if (cfg->device < 0 || cfg->device >= ndev) return SWAP_E_INVAL;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, cfg->device) != hipSuccess) return SWAP_E_NOGPU;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return SWAP_E_NOGPU;
    swap_ctx *c = new swap_ctx();
    c->cfg = *cfg;
    c->pipe.reset(new swapk::Pipeline(cfg->device, cfg->k, cfg->c, !(cfg->flags & SWAP_FLAG_HASH_ORDER)));
    if (c->pipe->init_error() != hipSuccess) {
        delete c;
        return SWAP_E_DEVICE;
    }
    c->pipe->set_table_spill((cfg->flags & SWAP_FLAG_TABLE_SPILL) != 0);
    c->pipe->set_wide_parts((cfg->flags & SWAP_FLAG_WIDE_PARTITIONS) != 0);
    c->pipe->set_single_fill((cfg->flags & SWAP_FLAG_SINGLE_FILL) != 0);
    c->pipe->set_coarse_renumber((cfg->flags & SWAP_FLAG_COARSE_RENUMBER) != 0);
    // the static scratch's device address, once per device (the Pipeline made
    // this device current); a failure only disables swap_assemble_static
    (void)swapk::static_scratch_setup(cfg->device);
    *out = c;
    return SWAP_OK;
}

void swap_destroy(swap_ctx *c) { delete c; }

int swap_build_device(swap_ctx *c, const void *d_fasta, size_t n) {
    if (!c || (!d_fasta && n)) return SWAP_E_INVAL;
    c->built = c->merged = c->have_graph = c->have_compacted = c->have_ce = false;
    c->batch_off.clear();
    c->from_static = false;
    if (n == 0) return set_err(c, SWAP_E_EMPTY, "empty input");
    if (!c->tune_fixed) c->pipe->set_tuning(swapk::tuning_for(n));
    auto t0 = clk::now();
    int rc = c->pipe->build(static_cast<const uint8_t *>(d_fasta), n);
    c->build_ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
    if (rc) return set_err(c, rc, c->pipe->last_error());
    c->built = true;
    return SWAP_OK;
}

int swap_prefetch(swap_ctx *c, const char *fasta, size_t n) {
    if (!c || !fasta || n == 0) return SWAP_E_INVAL;
    if (c->fifo.size() >= 2) return set_err(c, SWAP_E_STATE, "two inputs are prefetched and not yet built");
    if (hipSetDevice(c->cfg.device) != hipSuccess) return set_err(c, SWAP_E_DEVICE, "hipSetDevice");
    // one staging loop at a time: the previous one has (nearly always) ended
    // during the build that followed it
    if (c->join_prefetches() != hipSuccess) {
        c->fifo.clear();
        return set_err(c, SWAP_E_DEVICE, "copy of a prefetched input to the device");
    }
    if (!c->copy_st) {
        if (hipStreamCreateWithFlags(&c->copy_st, hipStreamNonBlocking) != hipSuccess)
            return set_err(c, SWAP_E_DEVICE, "copy stream");
        for (auto &e : c->ahead_done)
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
                return set_err(c, SWAP_E_DEVICE, "copy events");
    }
    // the buffer no queued prefetch holds (builds are synchronous, so the
    // buffer of an input already built is free)
    int b = c->next_buf;
    if (!c->fifo.empty()) b = 1 - c->fifo.back()->buf;
    if (c->ahead[b].reserve(n + 16) != hipSuccess)
        return set_err(c, SWAP_E_DEVICE, "device allocation of the prefetched input");
    std::unique_ptr<Prefetch> p(new Prefetch());
    p->src = fasta;
    p->n = n;
    p->buf = b;
    Prefetch *q = p.get();
    void *dst = c->ahead[b].p;
    hipEvent_t done = c->ahead_done[b];
    Staging *stg = &c->staging;
    hipStream_t st = c->copy_st;
    // the copy starts after the context's queued device work (a build that
    // failed part way may still read the buffer)
    if (hipEventRecord(c->ahead_done[b], c->pipe->stream()) != hipSuccess ||
        hipStreamWaitEvent(st, c->ahead_done[b], 0) != hipSuccess)
        return set_err(c, SWAP_E_DEVICE, "ordering of the prefetch after the context's work");
    const int dev = c->cfg.device;
    q->th = std::thread([q, dst, done, stg, st, dev] {
        hipError_t e = hipSetDevice(dev);
        if (e == hipSuccess) e = stg->copy(dst, q->src, q->n, st);
        if (e == hipSuccess) e = hipEventRecord(done, st);
        q->rc = e;
    });
    c->fifo.push_back(std::move(p));
    c->next_buf = 1 - b;
    return SWAP_OK;
}

int swap_build(swap_ctx *c, const char *fasta, size_t n) {
    if (!c) return SWAP_E_INVAL;
    if (!c->fifo.empty()) {
        Prefetch &p = *c->fifo.front();
        if (p.src == fasta && p.n == n && fasta && n) {
            if (hipSetDevice(c->cfg.device) != hipSuccess) return set_err(c, SWAP_E_DEVICE, "hipSetDevice");
            // the oldest prefetch is this input: its copy (and only its copy)
            // is waited for on the device
            if (p.th.joinable()) p.th.join();
            const hipError_t e = p.rc;
            const int b = p.buf;
            c->fifo.pop_front();
            if (e != hipSuccess) return set_err(c, SWAP_E_DEVICE, "copy of a prefetched input to the device");
            if (hipStreamWaitEvent(c->pipe->stream(), c->ahead_done[b], 0) != hipSuccess)
                return set_err(c, SWAP_E_DEVICE, "wait for a prefetched input");
            return swap_build_device(c, c->ahead[b].p, n);
        }
        // not the input that was prefetched: the read-ahead is dropped
        const hipError_t e = c->join_prefetches();
        c->fifo.clear();
        if (e != hipSuccess) return set_err(c, SWAP_E_DEVICE, "copy of a prefetched input to the device");
    }
    if (!fasta && n) return SWAP_E_INVAL;
    if (n == 0) return set_err(c, SWAP_E_EMPTY, "empty input");
    if (hipSetDevice(c->cfg.device) != hipSuccess) return set_err(c, SWAP_E_DEVICE, "hipSetDevice");
    if (c->input.reserve(n + 16) != hipSuccess) return set_err(c, SWAP_E_DEVICE, "device allocation of the input");
    // the build's first kernels are queued behind the copies on the same stream
    if (c->staging.copy(c->input.p, fasta, n, c->pipe->stream()) != hipSuccess)
        return set_err(c, SWAP_E_DEVICE, "copy of the input to the device");
    return swap_build_device(c, c->input.p, n);
}

int swap_build_batch(swap_ctx *c, const void *const *d_fasta, const size_t *n, size_t nsamples) {
    if (!c || !d_fasta || !n || nsamples == 0) return SWAP_E_INVAL;
    c->built = c->merged = c->have_graph = c->have_compacted = c->have_ce = false;
    c->from_static = false;
    c->batch_off.clear();
    if (c->cfg.flags & (SWAP_FLAG_HASH_ORDER | SWAP_FLAG_KMER_GRAPH))
        return set_err(c, SWAP_E_INVAL, "a batch build is for merges in the reference's order");
    if (nsamples > swapk::kMaxBatchSamples) return set_err(c, SWAP_E_INVAL, "more than 512 samples in one batch");
    std::vector<uint64_t> off(nsamples + 1, 0);
    for (size_t i = 0; i < nsamples; i++) {
        if (!d_fasta[i] || n[i] == 0) return set_err(c, SWAP_E_EMPTY, "an empty sample");
        off[i + 1] = off[i] + n[i];
    }
    if (hipSetDevice(c->cfg.device) != hipSuccess) return set_err(c, SWAP_E_DEVICE, "hipSetDevice");
    auto t0 = clk::now();
    // the samples back to back in the context's input buffer
    if (c->input.reserve(off[nsamples] + 16) != hipSuccess)
        return set_err(c, SWAP_E_DEVICE, "device allocation of the batch input");
    for (size_t i = 0; i < nsamples; i++)
        if (hipMemcpyAsync(static_cast<char *>(c->input.p) + off[i], d_fasta[i], n[i], hipMemcpyDeviceToDevice,
                           c->pipe->stream()) != hipSuccess)
            return set_err(c, SWAP_E_DEVICE, "copy of a sample into the batch input");
    if (!c->tune_fixed) c->pipe->set_tuning(swapk::tuning_for(off[nsamples]));
    int rc = c->pipe->build_batch(static_cast<const uint8_t *>(c->input.p), off);
    c->build_ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
    if (rc) return set_err(c, rc, c->pipe->last_error());
    c->batch_off = off;
    return SWAP_OK;
}

int swap_batch_take(swap_ctx *batch, size_t i, swap_ctx *dst) {
    if (!batch || !dst || batch == dst) return SWAP_E_INVAL;
    dst->built = dst->merged = dst->have_graph = dst->have_compacted = dst->have_ce = false;
    dst->from_static = false;
    if (batch->batch_off.empty()) return set_err(dst, SWAP_E_STATE, "no batch build to take a sample from");
    if (i + 1 >= batch->batch_off.size()) return set_err(dst, SWAP_E_INVAL, "no such sample in the batch");
    if (dst->cfg.k != batch->cfg.k || dst->cfg.c != batch->cfg.c || dst->cfg.device != batch->cfg.device ||
        (dst->cfg.flags & (SWAP_FLAG_HASH_ORDER | SWAP_FLAG_KMER_GRAPH)))
        return set_err(dst, SWAP_E_INVAL, "the taking context needs the batch's k, c, device and order");
    if (hipSetDevice(dst->cfg.device) != hipSuccess) return set_err(dst, SWAP_E_DEVICE, "hipSetDevice");
    if (!dst->tune_fixed) dst->pipe->set_tuning(swapk::tuning_for(batch->batch_off[i + 1] - batch->batch_off[i]));
    auto t0 = clk::now();
    const int rc = dst->pipe->take_sample(*batch->pipe, i);
    dst->build_ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
    if (rc) return set_err(dst, rc, dst->pipe->last_error());
    dst->built = true;
    return SWAP_OK;
}

int swap_merge(swap_ctx *c) {
    if (!c) return SWAP_E_INVAL;
    if (!c->built) return set_err(c, SWAP_E_STATE, "merge before build");
    if (c->cfg.flags & SWAP_FLAG_KMER_GRAPH) return set_err(c, SWAP_E_STATE, "context was created for -s (no merge)");
    if (c->from_static && !swapk::current_arena())
        return set_err(c, SWAP_E_STATE, "the last build ran in the static scratch and is merged already");
    if (c->cfg.k < swapk::kMergeMinK)
        return set_err(c, SWAP_E_INVAL, "merge needs k >= 19 (the reference's range, sequence.cpp:44); k-mer graphs "
                                         "of shorter k are built with SWAP_FLAG_KMER_GRAPH");
    auto t0 = clk::now();
    bool mat = false;
    int rc = swapk::host_merge(*c->pipe, c->cfg, &mat, c->comp);
    c->merge_ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
    if (rc) return set_err(c, rc, c->pipe->last_error());
    c->merged = true;
    c->have_compacted = mat;
    c->have_ce = false;
    return SWAP_OK;
}

int swap_assemble_static(swap_ctx *c, const void *d_fasta, size_t n) {
    static std::atomic<uint64_t> epochs{0};
    if (!c || (!d_fasta && n)) return SWAP_E_INVAL;
    c->built = c->merged = c->have_graph = c->have_compacted = c->have_ce = false;
    c->from_static = false;
    if (n == 0) return set_err(c, SWAP_E_EMPTY, "empty input");
    if (n > swapk::kStaticMaxInput)
        return set_err(c, SWAP_E_TOOBIG,
                       "input of " + std::to_string(n) + " bytes exceeds the static scratch's limit of " +
                           std::to_string(swapk::kStaticMaxInput) + " bytes (use swap_build_device)");
    const bool merge = !(c->cfg.flags & SWAP_FLAG_KMER_GRAPH);
    if (merge && c->cfg.k < swapk::kMergeMinK)
        return set_err(c, SWAP_E_INVAL, "merge needs k >= 19 (the reference's range, sequence.cpp:44)");
    void *base = swapk::static_scratch(c->cfg.device);
    if (!base) return set_err(c, SWAP_E_DEVICE, "the static scratch's address could not be looked up on this device");
    swapk::Arena arena;
    arena.base = static_cast<uint8_t *>(base);
    arena.size = swapk::kStaticScratchBytes;
    arena.epoch = ++epochs;
    int rc;
    {
        swapk::ArenaScope scope(&arena);
        rc = swap_build_device(c, d_fasta, n);
        c->from_static = true;
        if (rc == SWAP_OK && merge) rc = swap_merge(c);
    }
    if (rc != SWAP_OK && arena.wanted > arena.size)
        return set_err(c, SWAP_E_TOOBIG,
                       "the build needs more device memory than the static scratch holds (" +
                           std::to_string(arena.wanted) + " bytes asked for, " +
                           std::to_string(swapk::kStaticScratchBytes) + " held)");
    return rc;
}

static swap_tuning to_c(const swapk::Tuning &t) {
    swap_tuning o;
    o.part_recs = t.part_recs;
    o.fill_recs = t.fill_recs;
    o.index_load = t.index_load;
    o.ruler_stride = t.ruler_stride;
    o.occ_groups = t.occ_groups;
    o.order_bin_keys = t.order_bin_keys;
    return o;
}

int swap_tuning_for(uint64_t input_bytes, swap_tuning *out) {
    if (!out) return SWAP_E_INVAL;
    *out = to_c(swapk::tuning_for(input_bytes));
    return SWAP_OK;
}

int swap_set_tuning(swap_ctx *c, const swap_tuning *t) {
    if (!c) return SWAP_E_INVAL;
    c->tune_fixed = t != nullptr;
    if (t) {
        swapk::Tuning x;
        x.part_recs = t->part_recs;
        x.fill_recs = t->fill_recs;
        x.index_load = t->index_load;
        x.ruler_stride = t->ruler_stride;
        x.occ_groups = t->occ_groups;
        x.order_bin_keys = t->order_bin_keys;
        c->pipe->set_tuning(x);
    }
    return SWAP_OK;
}

int swap_get_tuning(const swap_ctx *c, swap_tuning *out) {
    if (!c || !out) return SWAP_E_INVAL;
    *out = to_c(c->pipe->tuning());
    return SWAP_OK;
}

int swap_get_stats(const swap_ctx *c, swap_stats *s) {
    if (!c || !s) return SWAP_E_INVAL;
    s->reads = c->pipe->n_reads;
    s->kmers = c->pipe->n_occ;
    s->distinct = c->pipe->n_distinct;
    s->survivors = c->pipe->n_surv;
    s->kept = c->pipe->n_kept;
    s->overflow = c->pipe->n_overflow;
    s->special = c->pipe->n_special;
    s->build_ms = c->build_ms;
    s->merge_ms = c->merge_ms;
    for (int i = 0; i < 16; i++) s->phase_ms[i] = c->pipe->phase_ms[i];
    return SWAP_OK;
}

int64_t swap_export_kmer_graph(swap_ctx *c, uint64_t *keys, uint8_t *mult, size_t cap) {
    if (!c) return SWAP_E_INVAL;
    int rc = ensure_graph(c);
    if (rc) return rc;
    const size_t n = c->graph.key.size();
    if ((keys || mult) && cap < n) return set_err(c, SWAP_E_RANGE, "buffer too small");
    for (size_t i = 0; i < n && (keys || mult); i++) {
        if (keys) keys[i] = c->graph.key[i];
        if (mult) for (int t = 0; t < 8; t++) mult[8 * i + t] = (uint8_t)(c->graph.mult[i] >> (8 * t));
    }
    return (int64_t)n;
}

int64_t swap_format_file(swap_ctx *c, int which, char *buf, size_t cap) {
    if (!c) return SWAP_E_INVAL;
    std::string out;
    const int rc = swapk::format_file_string(c, which, out);
    if (rc) return rc;
    if (buf) {
        if (cap < out.size()) return set_err(c, SWAP_E_RANGE, "buffer too small");
        std::memcpy(buf, out.data(), out.size());
    }
    return (int64_t)out.size();
}

int64_t swap_write_file(swap_ctx *c, int which, const char *path) {
    if (!c || !path) return SWAP_E_INVAL;
    std::string b;
    const int rc = swapk::format_file_string(c, which, b);
    if (rc) return rc;
    FILE *f = std::fopen(path, "wb");
    if (!f) return set_err(c, SWAP_E_IO, std::string("cannot open ") + path);
    const bool ok = std::fwrite(b.data(), 1, b.size(), f) == b.size();
    std::fclose(f);
    return ok ? (int64_t)b.size() : set_err(c, SWAP_E_IO, std::string("short write to ") + path);
}

int swap_reconfigure(swap_ctx *c, const swap_cfg *cfg) {
    if (!c || !cfg) return SWAP_E_INVAL;
    if (cfg->k < 9 || cfg->k > 31 || (cfg->k % 2) == 0) return SWAP_E_INVAL;
    if (cfg->device != c->cfg.device) return set_err(c, SWAP_E_INVAL, "a context stays on its device");
    c->cfg = *cfg;
    c->pipe->reconfigure(cfg->k, cfg->c, !(cfg->flags & SWAP_FLAG_HASH_ORDER));
    c->pipe->set_table_spill((cfg->flags & SWAP_FLAG_TABLE_SPILL) != 0);
    c->pipe->set_wide_parts((cfg->flags & SWAP_FLAG_WIDE_PARTITIONS) != 0);
    c->pipe->set_single_fill((cfg->flags & SWAP_FLAG_SINGLE_FILL) != 0);
    c->pipe->set_coarse_renumber((cfg->flags & SWAP_FLAG_COARSE_RENUMBER) != 0);
    c->built = c->merged = c->have_graph = c->have_compacted = c->have_ce = false;
    return SWAP_OK;
}

int64_t swap_export_compacted(swap_ctx *c, uint64_t *keys, uint64_t *edge_off, uint32_t *edge_len,
                              uint8_t *edge_mult, size_t cap_nodes, char *pool, size_t cap_pool,
                              uint64_t *pool_len) {
    if (!c) return SWAP_E_INVAL;
    if (!c->merged) return set_err(c, SWAP_E_STATE, "no merge");
    if (!c->have_compacted) {
        const int rc = swapk::materialize_merge(*c->pipe, c->cfg, c->comp);
        if (rc) return set_err(c, rc, c->pipe->last_error());
        c->have_compacted = true;
    }
    const swapk::HostCompacted &g = c->comp;
    const size_t n = g.key.size();
    if (pool_len) *pool_len = g.pool.size();
    const bool want_nodes = keys || edge_off || edge_len || edge_mult;
    if (want_nodes && cap_nodes < n) return set_err(c, SWAP_E_RANGE, "node buffers too small");
    if (pool && cap_pool < g.pool.size()) return set_err(c, SWAP_E_RANGE, "pool buffer too small");
    if (keys) std::memcpy(keys, g.key.data(), n * 8);
    if (edge_off) std::memcpy(edge_off, g.off.data(), 8 * n * 8);
    if (edge_len) std::memcpy(edge_len, g.len.data(), 8 * n * 4);
    if (edge_mult) std::memcpy(edge_mult, g.mult.data(), 8 * n);
    if (pool && !g.pool.empty()) std::memcpy(pool, g.pool.data(), g.pool.size());
    return (int64_t)n;
}

int64_t swap_table_order(const uint64_t *keys, const uint64_t *ranks, size_t n, uint64_t n_total, int extra_resize,
                         uint32_t *order_out) {
    if ((!keys || !ranks || !order_out) && n) return SWAP_E_INVAL;
    std::vector<uint64_t> k(keys, keys + n), r(ranks, ranks + n);
    for (size_t i = 0; i < n; i++)
        if (r[i] >= n_total) return SWAP_E_INVAL;
    std::vector<uint32_t> o = swapk::table_order(k, r, n_total, extra_resize != 0, nullptr);
    std::copy(o.begin(), o.end(), order_out);
    return (int64_t)n;
}

int64_t swap_visit_order(const uint64_t *keys, const uint64_t *first, size_t n, const uint64_t *epoch_first,
                         size_t n_epochs, uint64_t distinct, int tail_lookup, uint32_t *order_out) {
    if ((!keys || !first || !order_out) && n) return SWAP_E_INVAL;
    std::vector<uint64_t> k(keys, keys + n), f(first, first + n), e(epoch_first, epoch_first + n_epochs);
    std::vector<uint32_t> o = swapk::visit_order(k, f, e, distinct, tail_lookup != 0, nullptr);
    std::copy(o.begin(), o.end(), order_out);
    return (int64_t)n;
}

int64_t swap_table_epochs(uint64_t n_total, int extra_resize, uint64_t *starts, size_t cap) {
    std::vector<uint64_t> s = swapk::table_epoch_starts(n_total, extra_resize != 0);
    for (size_t i = 0; i < s.size() && i < cap; i++) starts[i] = s[i];
    return (int64_t)s.size();
}

const char *swap_strerror(int code) {
    switch (code) {
        case SWAP_OK: return "ok";
        case SWAP_E_INVAL: return "invalid argument";
        case SWAP_E_EMPTY: return "empty input";
        case SWAP_E_FORMAT: return "input has no FASTA record ('>')";
        case SWAP_E_TOOBIG: return "too many k-mers for one build";
        case SWAP_E_DEVICE: return "HIP runtime error";
        case SWAP_E_NOGPU: return "no gfx950 device";
        case SWAP_E_IO: return "I/O error";
        case SWAP_E_STATE: return "call out of order";
        case SWAP_E_RANGE: return "buffer too small";
        default: return "unknown error";
    }
}

const char *swap_last_error(const swap_ctx *c) { return c ? c->err.c_str() : ""; }

}  // extern "C"
