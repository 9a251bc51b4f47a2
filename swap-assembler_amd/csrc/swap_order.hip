// The reference's merge visit order computed on the device.
//
// SWAP-Assembler 2 visits nodes in the iteration order of a libstdc++
// __gnu_cxx::hash_map (graph.cpp:1409) that was filled from two other such
// tables (graph.cpp:1058-1070, 1096-1112, 1123-1133).  Iteration order is
// bucket-major (bucket = hash % bucket_count, hash of graph.cpp:29-32); within a
// bucket it depends on the history of inserts (head insertion) and rehashes
// (which reverse each chain), and that history only depends on each key's
// insertion rank and on the ranks at which the table grew.  So each table
// order is: sort by final bucket, then order each bucket's few keys with a
// comparator that walks back through the growth epochs.
#include <algorithm>
#include <cstring>

#include <rocprim/device/device_segmented_radix_sort.hpp>

#include "swap_host.h"
#include "swap_scan.h"

namespace swapk {

struct EpochTab {
    int n;              // number of epochs
    uint64_t start[32]; // first insert rank of each epoch
    uint64_t P[32];     // bucket count during each epoch
    uint64_t M[32];     // floor((2^64 - 1) / P[e])
};

// x mod P with M = floor((2^64 - 1) / P): q = floor(x * M / 2^64) is the
// quotient or one less, so one conditional subtraction finishes the remainder
// (instead of a 64-bit division)
__device__ __forceinline__ uint64_t mod_p(uint64_t x, uint64_t P, uint64_t M) {
    const uint64_t q = __umul64hi(x, M);
    uint64_t r = x - q * P;
    if (r >= P) r -= P;
    return r;
}

__device__ __forceinline__ uint64_t dev_ref_hash(uint64_t x) {
    return (x >> 32) ^ (uint64_t)(int64_t)(int32_t)(uint32_t)(x & 0xFFFFFFFFull);
}

// What the order of one key in a table depends on, gathered in one 16-byte
// load: its hash, its insertion rank (< 2^32: an input k-mer index or a
// position in another table) and the growth epoch it was inserted in.
struct OrdKey {
    uint64_t h;
    uint32_t rank;
    uint32_t ins;
};

__device__ __forceinline__ int insert_epoch(const EpochTab &et, uint64_t r) {
    int e = 0;
    for (int j = 1; j < et.n; j++)
        if (et.start[j] <= r) e = j;
    return e;
}

__global__ void k_ord_prep(const uint64_t *__restrict__ key, const uint64_t *__restrict__ rank, uint64_t n,
                           const EpochTab *__restrict__ etp, OrdKey *__restrict__ ok, uint32_t *__restrict__ bucket,
                           uint32_t *__restrict__ idx) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const EpochTab &et = *etp;
    const uint64_t hv = dev_ref_hash(key[i]);
    const uint64_t r = rank[i];
    OrdKey o;
    o.h = hv;
    o.rank = (uint32_t)r;
    o.ins = (uint32_t)insert_epoch(et, r);
    ok[i] = o;
    bucket[i] = (uint32_t)mod_p(hv, et.P[et.n - 1], et.M[et.n - 1]);
    idx[i] = (uint32_t)i;
}

// a precedes b at the end of epoch e (see the host restatement in swap_host.cpp)
__device__ __forceinline__ bool ord_before(const OrdKey &a, const OrdKey &b, int e, const EpochTab &et) {
    bool flip = false;
    for (; e >= 0; e--) {
        const uint64_t ba = mod_p(a.h, et.P[e], et.M[e]), bb = mod_p(b.h, et.P[e], et.M[e]);
        if (ba != bb) return (ba < bb) != flip;
        const bool ia = a.ins == (uint32_t)e, ib = b.ins == (uint32_t)e;
        if (ia && ib) return (a.rank > b.rank) != flip;
        if (ia != ib) return ia != flip;
        flip = !flip;
    }
    return false;
}

// Bucket sort of a table's keys in two steps: bins by the top bits of the
// bucket (a count pass writing one histogram column per workgroup, a scan of
// the columns, a scatter pass whose slots come from LDS counters), then a
// segmented radix sort of every bin by the remaining bits.  Keys of one bucket
// come out in no set order: k_ord_ties orders every run of equal buckets.  (A
// whole-array radix sort spends a pass per 8 bits of the 23-27-bit buckets,
// each with its own state clearing, on the same data.)
constexpr int kBinThreads = 1024;
// 2^10 bins at most: with 2^12, each workgroup's keys of a bin were runs of
// ~4 on the 4.6M-key table, one partial line per store (scatter 0.12 ms); with
// 2^8, the segmented sort of 18K-key bins took longer than the scatter saved
constexpr int kMaxBinBits = 10;

__global__ void __launch_bounds__(kBinThreads) k_bin_count(const uint32_t *__restrict__ bucket, uint64_t n,
                                                           uint64_t per, int shift, uint32_t nbins,
                                                           uint32_t *__restrict__ cnt) {
    __shared__ uint32_t h[1u << kMaxBinBits];
    for (uint32_t i = threadIdx.x; i < nbins; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const uint64_t b = (uint64_t)blockIdx.x * per, e = min(n, b + per);
    for (uint64_t i = b + threadIdx.x; i < e; i += blockDim.x) atomicAdd(&h[bucket[i] >> shift], 1u);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nbins; i += blockDim.x) cnt[(uint64_t)i * gridDim.x + blockIdx.x] = h[i];
}

// off[bin * gridDim.x + block]: first slot of the block's keys of that bin.
// Workgroup 0 also writes the bin bounds for the segmented sort.
__global__ void __launch_bounds__(kBinThreads) k_bin_scatter(const uint32_t *__restrict__ bucket,
                                                             const uint32_t *__restrict__ idx, uint64_t n,
                                                             uint64_t per, int shift, uint32_t nbins,
                                                             const uint64_t *__restrict__ off,
                                                             uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                             uint32_t *__restrict__ bin_start) {
    __shared__ uint32_t base[1u << kMaxBinBits];
    for (uint32_t i = threadIdx.x; i < nbins; i += blockDim.x) {
        base[i] = (uint32_t)off[(uint64_t)i * gridDim.x + blockIdx.x];
        if (blockIdx.x == 0) bin_start[i] = base[i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) bin_start[nbins] = (uint32_t)n;
    __syncthreads();
    const uint64_t b = (uint64_t)blockIdx.x * per, e = min(n, b + per);
    for (uint64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
        const uint32_t k = bucket[i];
        const uint32_t slot = atomicAdd(&base[k >> shift], 1u);
        kout[slot] = k;
        vout[slot] = idx[i];
    }
}

// One thread per bucket run of the sorted array: insertion sort of its keys.
__global__ void k_ord_ties(const uint32_t *__restrict__ bucket, uint32_t *__restrict__ idx, uint64_t n,
                           const EpochTab *__restrict__ etp, const OrdKey *__restrict__ ok) {
    const uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const EpochTab &et = *etp;
    const uint32_t b = bucket[q];
    if (q > 0 && bucket[q - 1] == b) return;
    uint64_t e = q + 1;
    while (e < n && bucket[e] == b) e++;
    if (e - q < 2) return;
    const int f = et.n - 1;
    for (uint64_t x = q + 1; x < e; x++) {
        const uint32_t v = idx[x];
        const OrdKey kv = ok[v];
        uint64_t y = x;
        while (y > q && ord_before(kv, ok[idx[y - 1]], f, et)) {
            idx[y] = idx[y - 1];
            y--;
        }
        idx[y] = v;
    }
}

// Insertion epochs for new ranks (hash values and buckets unchanged).
__global__ void k_ord_ins(const uint64_t *__restrict__ rank, uint64_t n, const EpochTab *__restrict__ etp,
                          OrdKey *__restrict__ ok) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t r = rank[i];
    ok[i].rank = (uint32_t)r;
    ok[i].ins = (uint32_t)insert_epoch(*etp, r);
}

__global__ void k_ord_positions(const uint32_t *__restrict__ order, uint64_t n, uint64_t *__restrict__ pos) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pos[order[i]] = i;
}

__global__ void k_u32_to_u64_o(const uint32_t *__restrict__ in, uint64_t n, uint64_t *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}

#define OCHECK(x)                                                            \
    do {                                                                     \
        hipError_t e_ = (x);                                                 \
        if (e_ != hipSuccess) return fail(-5, hipGetErrorString(e_));        \
    } while (0)

static inline unsigned oblk(uint64_t n) { return (unsigned)((n + 255) / 256); }

// Iteration order of one table: sorted survivor indices into d_order.  The
// epoch table lives in device memory (d_et); tab is its host copy (the bucket
// counts are needed for the sort width; the starts may differ).
int Pipeline::table_order_device(const uint64_t *d_rank, const EpochTabHost &tab, const void *d_et, uint32_t *d_order) {
    const uint64_t S = n_surv;
    if (!S) return 0;
    const EpochTab *et = static_cast<const EpochTab *>(d_et);
    OCHECK(ob_[O_KEY].reserve(S * sizeof(OrdKey)));
    OCHECK(ob_[O_BKT].reserve(S * 4));
    OCHECK(ob_[O_BKT2].reserve(S * 4));
    OCHECK(ob_[O_IDX].reserve(S * 4));
    hipLaunchKernelGGL(k_ord_prep, dim3(oblk(S)), dim3(256), 0, st_, surv_key_.as<uint64_t>(), d_rank, S, et,
                       ob_[O_KEY].as<OrdKey>(), ob_[O_BKT].as<uint32_t>(),
                       ob_[O_IDX].as<uint32_t>());
    int bits = 1;
    while (bits < 32 && (1ull << bits) < tab.P[tab.n - 1]) bits++;
    // bins of at least ~1024 keys; the segmented sort orders each by the low bits
    int binb = 0;
    while (binb < kMaxBinBits && (S >> (binb + 1)) >= 1024) binb++;
    binb = std::min(binb, bits);
    const int shift = bits - binb;
    const uint32_t nbins = 1u << binb;
    int ncu = 0;
    OCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev_));
    const uint64_t nblk = std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)ncu, (S + 8191) / 8192));
    const uint64_t per = (S + nblk - 1) / nblk;
    OCHECK(ob_[O_BCNT].reserve((uint64_t)nbins * nblk * 4 + 4));
    OCHECK(ob_[O_BOFF].reserve((uint64_t)nbins * nblk * 8 + 8));
    OCHECK(ob_[O_BSTART].reserve((uint64_t)(nbins + 1) * 4));
    OCHECK(scan_scratch_.reserve((scan_scratch_words((uint64_t)nbins * nblk) + 16) * 8));
    const bool direct = shift == 0;  // one bucket per bin: the scatter is the sort
    if (!direct) {
        OCHECK(ob_[O_SK].reserve(S * 4));
        OCHECK(ob_[O_SV].reserve(S * 4));
    }
    uint32_t *kout = direct ? ob_[O_BKT2].as<uint32_t>() : ob_[O_SK].as<uint32_t>();
    uint32_t *vout = direct ? d_order : ob_[O_SV].as<uint32_t>();
    hipLaunchKernelGGL(k_bin_count, dim3((unsigned)nblk), dim3(kBinThreads), 0, st_, ob_[O_BKT].as<uint32_t>(), S, per,
                       shift, nbins, ob_[O_BCNT].as<uint32_t>());
    OCHECK(scan_excl<uint32_t>(ob_[O_BCNT].as<uint32_t>(), ob_[O_BOFF].as<uint64_t>(), (uint64_t)nbins * nblk,
                               scan_scratch_.as<uint64_t>(), ob_[O_BOFF].as<uint64_t>() + (uint64_t)nbins * nblk, st_));
    hipLaunchKernelGGL(k_bin_scatter, dim3((unsigned)nblk), dim3(kBinThreads), 0, st_, ob_[O_BKT].as<uint32_t>(),
                       ob_[O_IDX].as<uint32_t>(), S, per, shift, nbins, ob_[O_BOFF].as<uint64_t>(), kout, vout,
                       ob_[O_BSTART].as<uint32_t>());
    if (!direct) {
        const uint32_t *bs = ob_[O_BSTART].as<uint32_t>();
        size_t tb = 0;
        OCHECK(rocprim::segmented_radix_sort_pairs(nullptr, tb, kout, ob_[O_BKT2].as<uint32_t>(), vout, d_order,
                                                   (unsigned)S, nbins, bs, bs + 1, 0u, (unsigned)shift, st_));
        OCHECK(ob_[O_TMP].reserve(tb + 16));
        OCHECK(rocprim::segmented_radix_sort_pairs(ob_[O_TMP].p, tb, kout, ob_[O_BKT2].as<uint32_t>(), vout, d_order,
                                                   (unsigned)S, nbins, bs, bs + 1, 0u, (unsigned)shift, st_));
    }
    hipLaunchKernelGGL(k_ord_ties, dim3(oblk(S)), dim3(256), 0, st_, ob_[O_BKT2].as<uint32_t>(), d_order, S, et,
                       ob_[O_KEY].as<OrdKey>());
    return 0;
}

// Same table filled again in a new insertion order (the nodes table after the
// cutoff copy: equal bucket count, so the bucket sort of the previous call
// stands and only the order inside each bucket is redone).
int Pipeline::table_reorder_device(const uint64_t *d_rank, const void *d_et, uint32_t *d_order) {
    const uint64_t S = n_surv;
    const EpochTab *et = static_cast<const EpochTab *>(d_et);
    hipLaunchKernelGGL(k_ord_ins, dim3(oblk(S)), dim3(256), 0, st_, d_rank, S, et, ob_[O_KEY].as<OrdKey>());
    hipLaunchKernelGGL(k_ord_ties, dim3(oblk(S)), dim3(256), 0, st_, ob_[O_BKT2].as<uint32_t>(), d_order, S, et,
                       ob_[O_KEY].as<OrdKey>());
    return 0;
}

EpochTabHost epoch_table(uint64_t n_total, bool extra, const std::vector<uint64_t> *starts_override) {
    EpochTabHost t;
    std::vector<uint64_t> s = table_epoch_starts(n_total, extra);
    t.n = (int)std::min<size_t>(s.size(), 32);
    uint64_t P = table_buckets(0, false);
    for (int e = 0; e < t.n; e++) {
        if (e > 0) P = table_buckets(s[e] + 1, false);
        t.P[e] = P;
        t.M[e] = ~0ull / P;
        t.start[e] = s[e];
        if (starts_override) t.start[e] = e < (int)starts_override->size() ? (*starts_override)[e] : ~0ull;
    }
    return t;
}

// Visit priorities (position in the nodes table order) of all survivors.
int Pipeline::visit_priorities(uint64_t *d_prio) {
    const uint64_t S = n_surv;
    OCHECK(ob_[O_RANK].reserve(S * 8));
    OCHECK(ob_[O_ORD].reserve(S * 4));
    uint64_t *rank = ob_[O_RANK].as<uint64_t>();
    uint32_t *ord = ob_[O_ORD].as<uint32_t>();
    // kmolecules: ranks are first occurrences; the epoch table's starts are the
    // first occurrences of the growth ranks, selected on the device by the build
    hipLaunchKernelGGL(k_u32_to_u64_o, dim3(oblk(S)), dim3(256), 0, st_, surv_first_.as<uint32_t>(), S, rank);
    int rc = table_order_device(rank, epoch1_, etab_.p, ord);
    if (rc) return rc;
    // cutoff copy and nodes table: ranks are positions in the previous order
    *h_etabS_ = epoch_table(S, false, nullptr);
    OCHECK(etabS_.reserve(sizeof(EpochTabHost) + 64));
    OCHECK(hipMemcpyAsync(etabS_.p, h_etabS_, sizeof(EpochTabHost), hipMemcpyHostToDevice, st_));
    hipLaunchKernelGGL(k_ord_positions, dim3(oblk(S)), dim3(256), 0, st_, ord, S, rank);
    rc = table_order_device(rank, *h_etabS_, etabS_.p, ord);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ord_positions, dim3(oblk(S)), dim3(256), 0, st_, ord, S, rank);
    rc = table_reorder_device(rank, etabS_.p, ord);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ord_positions, dim3(oblk(S)), dim3(256), 0, st_, ord, S, d_prio);
    return 0;
}

}  // namespace swapk
