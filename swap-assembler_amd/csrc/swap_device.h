// Device-side pipeline of the k-mer graph path (one sample on one GPU).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <memory>
#include <vector>

#include "swap_common.h"

namespace swapk {

// A bump allocator over a fixed device region (the static scratch of
// swap_assemble_static): while an ArenaScope is alive on a host thread, every
// DevBuf::reserve on that thread takes fresh space from it instead of calling
// hipMalloc; running out of space is reported, nothing is allocated.
struct Arena {
    uint8_t *base = nullptr;
    size_t size = 0;
    size_t used = 0;
    size_t wanted = 0;   // bytes asked for in all (more than size after an overflow)
    uint64_t epoch = 0;  // distinguishes the space of one call from the next
};
Arena *current_arena();
struct ArenaScope {
    explicit ArenaScope(Arena *a);
    ~ArenaScope();
    Arena *prev;
};

// Performance parameters of a build and merge; none of them changes a result.
//   part_recs     records per count partition the partition count aims at
//                 (sets the load of the on-chip counting table)
//   fill_recs     records a shared fill of the on-chip table aims at (small inputs)
//   index_load    largest load of a survivor-index region, in percent (50..85)
//   ruler_stride  one ruler per this many oriented nodes in the list ranking
//   occ_groups    workgroups per CU of the k-mer passes (histogram, level-1 partition)
//   order_bin_keys  average keys per bin of the visit-order sort (a bin is
//                 ordered in LDS up to 6144 keys, in device memory above)
struct Tuning {
    uint32_t part_recs = 8192;
    uint32_t fill_recs = 4096;
    uint32_t index_load = 67;
    uint32_t ruler_stride = 32;
    uint32_t occ_groups = 1;
    uint32_t order_bin_keys = 2048;
};
// Clamped to the ranges the kernels support.
Tuning clamp_tuning(const Tuning &t);
// The table entry for an input of n bytes (the table file found next to the
// library, or SWAP_TUNE_FILE; built-in defaults without one).
Tuning tuning_for(uint64_t n_bytes);

// Grow-only device buffer.  Outside an arena its space comes from hipMalloc
// (kept in heap); inside one, from the arena.  A buffer keeps pointing at
// arena space after the scope ends (results of a static call stay readable
// until the next reserve that needs more room) until leave_arena().
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    void *heap = nullptr;
    size_t heap_cap = 0;
    uint64_t arena_epoch = 0;  // nonzero: p is arena space of that call
    ~DevBuf();
    hipError_t reserve(size_t bytes);
    // back to the buffer's own device memory (nothing is freed or allocated)
    void leave_arena() {
        if (arena_epoch) { p = heap; cap = heap_cap; arena_epoch = 0; }
    }
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

struct Timing {
    float ms_build = 0, ms_merge = 0;
};

// Survivors of the cutoff: one record per k-molecule (canonical k-mer), with the
// 8 saturated edge multiplicities packed one byte per slot (slot t in byte t).
struct HostKmerGraph {
    std::vector<uint64_t> key;
    std::vector<uint64_t> mult;
    std::vector<uint32_t> first;  // first input k-mer index of the key (reference order mode)
    std::vector<uint64_t> epoch_first;  // first-occurrence index of the key of each table growth
    uint64_t distinct = 0;        // D
    bool tail_lookup = false;     // an input k-mer with edges follows the last first occurrence
};

// Compacted graph after the merge: the kept nodes, each with 8 edge slots
// (offset/length into a base pool, multiplicity).
struct HostCompacted {
    std::vector<uint64_t> key;       // kept node ids
    std::vector<uint64_t> off;       // 8 per node: offset of slot string in pool
    std::vector<uint32_t> len;       // 8 per node
    std::vector<uint8_t> mult;       // 8 per node
    std::vector<char> pool;          // edge strings, bases as 'A','C','G','T'
    // nodes are stored in the order of contigGraph.txt (the gather walks the
    // nodes table in visit order); noce_order lists them in the order of the
    // table the gather re-fills, which printContigs walks
    std::vector<uint32_t> noce_order;
    // bucket count of the reference's nodes table when the gather re-fills it
    uint64_t node_buckets = 0;
    // The table after the single-rank finisher (masterSimplifyNodeGraph), which
    // noCEcontigs.fasta is written from: same rows and order as above, a node
    // the finisher absorbed has no edges left.  Null when the finisher changes
    // nothing (no node of exactly one left and one right edge survived the merge).
    std::shared_ptr<const HostCompacted> finished;
};

// Outputs of the contig extension (host, after the merge).
struct CeFiles {
    std::string noce;      // noCEcontigs.fasta as printContigs writes it from the re-filled table
    std::string ce;        // CEContig.fasta
    std::string jung_arc;  // JungGraph_arc.txt
    std::string jung_mul;  // JungGraph_mul.txt
};

// Growth epochs of a reference hash table: epoch e starts at insert rank
// start[e] (or at the first occurrence of that rank's key) with P[e] buckets.
struct EpochTabHost {
    int n = 0;
    uint64_t start[32] = {0};
    uint64_t P[32] = {0};
    uint64_t M[32] = {0};  // floor((2^64 - 1) / P[e]): x mod P[e] by one high multiply (Barrett)
};
EpochTabHost epoch_table(uint64_t n_total, bool extra, const std::vector<uint64_t> *starts_override);

// Device merge results copied to the host.
struct MergeDev {
    uint64_t n_kept = 0;
    std::vector<uint32_t> kept_src;   // survivor index of each kept node
    std::vector<uint32_t> elen;       // 8 per kept node: edge string length (0 = no edge / replayed on host)
    std::vector<uint64_t> eoff;       // 8 per kept node: offset into pool
    std::vector<uint8_t> emult;       // 8 per kept node
    std::vector<uint8_t> especial;    // 8 per kept node: edge enters a cycle / fold-back component
    std::vector<char> pool;
    std::vector<uint32_t> special;    // (1,1) survivors of cycle / fold-back components
    std::vector<uint64_t> kept_prio;  // visit priority of each kept node
    std::vector<uint64_t> kept_key;   // key of each kept node
    // when special components exist, reference order: the special nodes' key,
    // multiplicities and priority, the kept nodes' multiplicities, and the
    // (1,1) nodes as a bitmap over visit ranks (priorities are ranks 0..S-1)
    std::vector<uint64_t> spec_key, spec_mult, spec_prio, kept_mult;
    std::vector<uint32_t> one_bits;
    // hash order (priorities are key hashes, not ranks): all survivors
    std::vector<uint64_t> all_key, all_mult, all_prio;
    uint64_t n11 = 0;
};

// One sample of a batch build (Pipeline::build_batch): what a build of the
// sample alone leaves for the merge and the exports.
struct BatchSample {
    uint64_t kbase = 0;       // its first input k-mer index in the batch
    uint64_t reads = 0, kmers = 0, distinct = 0, survivors = 0;
    uint64_t surv0 = 0;       // its survivors: [surv0, surv0 + survivors) of the batch's, in first-occurrence order
    uint64_t last_first = 0;  // largest first occurrence of its distinct keys (its own input indices)
    std::vector<uint64_t> epoch_first;  // first occurrence of the key of each growth rank of its k-mer table
};

class Pipeline {
public:
    Pipeline(int device, int k, int c, bool reference_order);
    ~Pipeline();
    // New k, c and order mode for the next build; device buffers are kept.
    void reconfigure(int k, int c, bool reference_order);
    // every count partition through the device-memory table (a testing aid)
    void set_table_spill(bool on) { table_spill_ = on; }
    // at least 2^18 partitions and per-table visit-order bin passes on any
    // input (a testing aid for large inputs)
    void set_wide_parts(bool on) { wide_parts_ = on; }
    // one partition per on-chip table fill on any input (a testing aid)
    void set_single_fill(bool on) { single_fill_ = on; }
    // survivors renumbered by coarse first-occurrence bins only, as inputs
    // beyond the exact ordering's range are (a testing aid)
    void set_coarse_renumber(bool on) { coarse_renumber_ = on; }
    void set_tuning(const Tuning &t) { tune_ = clamp_tuning(t); }
    const Tuning &tuning() const { return tune_; }
    int device() const { return dev_; }
    hipStream_t stream() const { return st_; }

    // Builds the k-mer graph from FASTA bytes already in device memory.
    // Returns 0 or a negative SWAP_E* code.
    int build(const uint8_t *d_fasta, uint64_t n);
    // Builds the k-mer graphs of several samples in one pass of the build
    // kernels (reference order only): the samples' FASTA bytes lie back to back
    // in d_fasta, sample i at [off[i], off[i + 1]).  Every sample must start
    // with '>', end with a newline and hold an even number of lines before the
    // last sample (SWAP_E_INVAL otherwise).  Results per sample in batch().
    int build_batch(const uint8_t *d_fasta, const std::vector<uint64_t> &off);
    const std::vector<BatchSample> &batch() const { return batch_; }
    // This context's state becomes that of a build of sample i of src's last
    // batch build alone (same k, c and order mode): ready for merge_device and
    // the exports.  Returns when the copies are done.
    int take_sample(const Pipeline &src, size_t i);
    // Compacts unitigs (k-molecule merge).  prio: per-survivor visit rank in
    // survivor order (device or host array, may be null for the built-in
    // hash order).
    int merge_device(bool reference_order, MergeDev &out);
    int fetch_merge(MergeDev &out);
    // on stream s (the side stream in a merge)
    int visit_priorities(uint64_t *d_prio, hipStream_t s);
    // exact_only: the exact first-occurrence order or none (batch builds slice by it)
    int renumber_survivors(bool *renumbered, bool exact_only = false);

    // Counters of the last build / merge.
    uint64_t n_reads = 0, n_occ = 0, n_distinct = 0, n_surv = 0, n_kept = 0, n_overflow = 0;
    uint64_t n_special = 0;  // (1,1) nodes in cycles / fold-backs, merged on the host

    int export_kmer_graph(HostKmerGraph &out);
    int export_compacted(HostCompacted &out);

    // Device pointers of the survivors (valid after build()).
    const uint64_t *d_surv_key() const { return surv_key_.as<uint64_t>(); }
    const uint64_t *d_surv_mult() const { return surv_mult_.as<uint64_t>(); }

    const std::string &last_error() const { return err_; }
    // hipSuccess, or the first HIP error met while setting up the context
    hipError_t init_error() const { return init_err_; }

    int k() const { return k_; }
    int c() const { return c_; }
    // Every device buffer back to its own memory (after calls that used an arena).
    void leave_arena();
    // Device time of the build phases of the last build, from events recorded
    // on the pipeline stream around the launches: 0 record scan, 1 pack,
    // 2 partition histogram, 3 level-1 partition, 4 LDS count + cutoff,
    // 5 overflow count, 6 growth-rank selection, 7 level-2 partition; merge
    // phases from 8.
    static constexpr int kPhases = 16;
    float phase_ms[kPhases] = {0};
    void phase_begin(int i, hipStream_t s = nullptr);
    void phase_end(int i, hipStream_t s = nullptr);
    // side: also wait for the side stream (phases recorded there)
    void collect_phases(bool side = true);

private:
    int fail(int code, const char *msg);
    int dev_;
    int k_, c_;
    bool ref_order_;
    bool table_spill_ = false;
    bool wide_parts_ = false;
    bool single_fill_ = false;
    bool coarse_renumber_ = false;
    Tuning tune_;
    int pbits_ = kPartBitsMin;  // partition bits of the last build
    // batch builds: the samples' byte offsets while build() runs for
    // build_batch, and the per-sample results
    const std::vector<uint64_t> *batch_off_ = nullptr;
    std::vector<BatchSample> batch_;
    DevBuf bdev_, bsel_, bbits_;
    hipStream_t st_ = nullptr;
    // Side stream (default priority): the build's growth-rank selection and
    // the merge's visit-order replay run there, overlapping the main stream's
    // work that does not need them.  sel_done_: the last selection has ended
    // (the epoch table is final); prio_done_: the visit priorities are written.
    hipStream_t st2_ = nullptr;
    hipEvent_t sel_done_ = nullptr, prio_done_ = nullptr, fork_ = nullptr;
    DevBuf scan_scratch2_;  // the side stream's scans
    hipError_t init_err_ = hipSuccess;
    Mixer mx_;
    std::string err_;
    // build buffers
    DevBuf fasta_, tile_nl_, tile_off_, nl_pos_, scan_scratch_, small_, rec_start_, rec_len_, occ_count_,
        word_count_, occ_base_, word_base_, block_rec_, packed_, wmeta_, rows_, hist_, bucket_, cursor_, hist1_, base1_, l1info_, rec1_, rec_, surv_key_, surv_mult_,
        surv_first_, dfo_, sel_, sel_target_, sel_out_, overflow_, part_surv_, part_first_, gtab_key_, gtab_cnt_, gtab_fo_;
    // merge buffers
    enum {
        M_IENT, M_IBASE, M_PRIO, M_IS11, M_NXT, M_EX, M_SELF, M_WALKA, M_WALKB, M_SPA, M_SPB, M_BCNT, M_BOFF,
        M_KEPTOF, M_KEPTSRC, M_ELEN, M_EMULT, M_EOTHER, M_CSIZE, M_ESPEC, M_EOFF, M_COFF, M_POOL, M_CPRI,
        M_CMUL, M_CEDGE, M_LEV, M_LEVOFF, M_PAR, M_NEED, M_NL, M_NR, M_INL, M_INR, M_CNT, M_SPEC, M_PROPS, M_SEG,
        M_COUNT
    };
    DevBuf mb_[M_COUNT];
    enum { O_ENT, O_R2, O_GSORT, O_BCNT, O_BOFF, O_BSTART, O_COUNT };
    DevBuf ob_[O_COUNT];
    // survivor renumbering
    enum { R_CHIST, R_CPRE, R_HEAD, R_HPRE, R_BINMAP, R_CNT, R_OFF, R_BSTART, R_KV, R_FI, R_GSRT, R_KEY, R_PERM, R_COUNT };
    DevBuf rn_[R_COUNT];
    // the survivor arrays are in first-occurrence order (renumbered by a merge
    // of this build): they no longer follow the count's partition runs
    bool surv_renumbered_ = false;
    // host mirrors of small results
    uint64_t *h_small_ = nullptr;
    uint64_t last_first_ = 0;
    uint64_t merged_ne_ = 0, merged_pool_ = 0, merged_M_ = 0;
    bool merged_ref_ = true;  // the last merge's visit priorities are reference-order ranks
    bool epoch_first_valid_ = false;
    EpochTabHost epoch1_;           // growth epochs of the k-mer table (starts on the device)
    DevBuf etab_, etabS_;           // device epoch tables: k-mer table, S-sized tables
    uint64_t *h_starts_ = nullptr;  // pinned staging
    EpochTabHost *h_etab_ = nullptr, *h_etabS_ = nullptr;
    uint64_t *hs_tail() { return h_small_ + 40; }
    uint64_t *d_small_ptr() { return small_.as<uint64_t>(); }
    std::vector<uint64_t> epoch_first_;
    hipEvent_t ev_[2 * kPhases] = {};
    bool ev_used_[kPhases] = {};
};

const char *error_string(int code);

// The static scratch (swap_static.hip): looked up once per device with that
// device current; null until then or if the lookup failed.
hipError_t static_scratch_setup(int device);
void *static_scratch(int device);

}  // namespace swapk
