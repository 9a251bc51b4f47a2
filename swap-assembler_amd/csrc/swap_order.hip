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

// What the order of one key in a table depends on: its hash, its insertion
// rank (< 2^32: an input k-mer index or a position in another table) and the
// growth epoch it was inserted in.
struct OrdKey {
    uint64_t h;
    uint32_t rank;
    uint32_t ins;
};

// the last epoch whose start is at most r (0 if none): every start tested, so
// the loads issue together
__device__ __forceinline__ int insert_epoch(const EpochTab &et, uint64_t r) {
    int e = 0;
#pragma unroll
    for (int j = 1; j < 32; j++)
        if (j < et.n && et.start[j] <= r) e = j;
    return e;
}

// a precedes b in their common final bucket (epoch et.n - 1): back through
// the epochs, the later-inserted key of a bucket comes first and each rehash
// reverses the order (see the host restatement in swap_host.cpp)
__device__ __forceinline__ bool ord_before_tied(const OrdKey &a, const OrdKey &b, const EpochTab &et) {
    bool flip = false;
    for (int e = et.n - 1;; e--) {
        const bool ia = a.ins == (uint32_t)e, ib = b.ins == (uint32_t)e;
        if (ia && ib) return (a.rank > b.rank) != flip;
        if (ia != ib) return ia != flip;
        flip = !flip;
        if (e == 0) return false;
        const uint64_t ba = mod_p(a.h, et.P[e - 1], et.M[e - 1]), bb = mod_p(b.h, et.P[e - 1], et.M[e - 1]);
        if (ba != bb) return (ba < bb) != flip;
    }
}

// Each table's keys go to bins by the top bits of their final bucket, both
// tables in one pass over the survivors: a count pass writing one histogram
// column per workgroup and bin, one scan of the columns (table 1's bins, then
// table 2's), a scatter pass whose slots come from LDS cursors.  Every key
// moves as one 16-byte entry holding what its order depends on, so that the
// per-bin pass (k_bin_order) reads only its bin's contiguous entries:
//   table 1 (k-molecules): hash, first occurrence (its insertion rank), and
//     the key's slot among table 2's keys;
//   table 2 (cutoff copy): hash, survivor number.
// Table 1's pass writes each key's position in table 1's order at that slot
// (r2), which table 2's pass reads as the insertion rank, in slot order.
struct OrdEntry {
    uint64_t h;
    uint32_t a;  // table 1: first occurrence; table 2: survivor
    uint32_t b;  // table 1: slot among table 2's keys (0..S-1)
};

constexpr int kBinThreads = 1024;
constexpr int kBinBatch = 4;  // keys a thread loads at once in the count and scatter passes
// bins per table: with more, a workgroup's keys of a bin become runs too short
// for whole-line stores in the scatter.  The bins aim at order_bin_keys keys
// (2048 by default: 2^11 bins for the bacterial workload's 4.6M survivors);
// 2^15 keeps the bins of ~100M survivors within the kSortCap keys an LDS
// ordering holds.  The count and scatter passes hold one 32-bit cursor per bin
// in LDS (kBinCursors of them, 128 KB: they run one workgroup per CU anyway):
// both tables in one pass while their bins fit, else one pass per table.
constexpr int kMaxBinBits = 15;
constexpr uint32_t kBinCursors = 1u << 15;

struct BinShape {
    uint64_t P1, M1, P2, M2;  // final bucket counts of table 1 and 2 (M: see mod_p)
    int shift1, shift2;       // bucket bits below the bin
    uint32_t nb1, nb2;        // bins of each table
};

// The range of keys a count / scatter workgroup takes: workgroups are dealt
// to the 8 XCDs in turn, so the ranges are numbered XCD by XCD -- the runs two
// workgroups write next to each other (and share cache lines with) then come
// from one XCD's L2.
__device__ __forceinline__ uint32_t xcd_range(uint32_t wg, uint32_t nwg) {
    return (nwg % 8 == 0) ? (wg % 8) * (nwg / 8) + wg / 8 : wg;
}

__device__ __forceinline__ void bins_of(uint64_t key, const BinShape &bs, uint64_t &h, uint32_t &b1, uint32_t &b2) {
    h = dev_ref_hash(key);
    b1 = (uint32_t)(mod_p(h, bs.P1, bs.M1) >> bs.shift1);
    b2 = bs.nb1 + (uint32_t)(mod_p(h, bs.P2, bs.M2) >> bs.shift2);
}

// which: 0 both tables (bins 0 .. nb1 + nb2 - 1), 1 table 1 only, 2 table 2
// only (its bins' columns after table 1's)
__global__ void __launch_bounds__(kBinThreads) k_bin_count(const uint64_t *__restrict__ key, uint64_t n,
                                                           uint64_t per, BinShape bs, uint32_t *__restrict__ cnt,
                                                           int which) {
    __shared__ uint32_t hist[kBinCursors];
    const uint32_t nb = which == 0 ? bs.nb1 + bs.nb2 : which == 1 ? bs.nb1 : bs.nb2;
    const uint32_t col0 = which == 2 ? bs.nb1 : 0;
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const uint32_t w = xcd_range(blockIdx.x, gridDim.x);
    const uint64_t b = (uint64_t)w * per, e = min(n, b + per);
    for (uint64_t i0 = b + threadIdx.x; i0 < e; i0 += kBinBatch * blockDim.x) {
        uint64_t k[kBinBatch];
#pragma unroll
        for (int u = 0; u < kBinBatch; u++) {
            const uint64_t i = i0 + (uint64_t)u * blockDim.x;
            if (i < e) k[u] = key[i];
        }
#pragma unroll
        for (int u = 0; u < kBinBatch; u++) {
            if (i0 + (uint64_t)u * blockDim.x >= e) continue;
            uint64_t h;
            uint32_t b1, b2;
            bins_of(k[u], bs, h, b1, b2);
            if (which != 2) atomicAdd(&hist[b1], 1u);
            if (which != 1) atomicAdd(&hist[b2 - col0], 1u);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) cnt[(uint64_t)(col0 + i) * gridDim.x + w] = hist[i];
}

// off[bin * gridDim.x + block]: first slot of the block's keys of that bin
// (table 1's slots 0..S-1, table 2's S..2S-1).  Workgroup 0 also writes every
// bin's first slot and the end (2n).  which as in k_bin_count; with one pass
// per table, table 2's runs first and leaves every key's table-2 slot in s2
// (by survivor) for table 1's entries.
__global__ void __launch_bounds__(kBinThreads) k_bin_scatter(const uint64_t *__restrict__ key,
                                                             const uint32_t *__restrict__ first, uint64_t n,
                                                             uint64_t per, BinShape bs,
                                                             const uint64_t *__restrict__ off,
                                                             OrdEntry *__restrict__ ent,
                                                             uint32_t *__restrict__ bin_start, int which,
                                                             uint32_t *__restrict__ s2) {
    __shared__ uint32_t cur[kBinCursors];
    const uint32_t nb = which == 0 ? bs.nb1 + bs.nb2 : which == 1 ? bs.nb1 : bs.nb2;
    const uint32_t col0 = which == 2 ? bs.nb1 : 0;
    const uint32_t w = xcd_range(blockIdx.x, gridDim.x);
    for (uint32_t i = threadIdx.x; i < nb; i += blockDim.x) {
        cur[i] = (uint32_t)off[(uint64_t)(col0 + i) * gridDim.x + w];
        if (w == 0) bin_start[col0 + i] = cur[i];
    }
    if (w == 0 && threadIdx.x == 0 && which != 1) bin_start[bs.nb1 + bs.nb2] = (uint32_t)(2 * n);
    __syncthreads();
    const uint64_t b = (uint64_t)w * per, e = min(n, b + per);
    for (uint64_t i0 = b + threadIdx.x; i0 < e; i0 += kBinBatch * blockDim.x) {
        uint64_t k[kBinBatch];
        uint32_t fo[kBinBatch];
#pragma unroll
        for (int u = 0; u < kBinBatch; u++) {
            const uint64_t i = i0 + (uint64_t)u * blockDim.x;
            if (i < e) {
                k[u] = key[i];
                fo[u] = first[i];
            }
        }
#pragma unroll
        for (int u = 0; u < kBinBatch; u++) {
            const uint64_t i = i0 + (uint64_t)u * blockDim.x;
            if (i >= e) continue;
            uint64_t h;
            uint32_t b1, b2;
            bins_of(k[u], bs, h, b1, b2);
            if (which == 0) {
                const uint32_t s1 = atomicAdd(&cur[b1], 1u), t2 = atomicAdd(&cur[b2], 1u);
                ent[s1] = OrdEntry{h, fo[u], t2 - (uint32_t)n};
                ent[t2] = OrdEntry{h, (uint32_t)i, 0u};
            } else if (which == 2) {
                const uint32_t t2 = atomicAdd(&cur[b2 - col0], 1u);
                ent[t2] = OrdEntry{h, (uint32_t)i, 0u};
                s2[i] = t2 - (uint32_t)n;
            } else {
                const uint32_t s1 = atomicAdd(&cur[b1], 1u);
                ent[s1] = OrdEntry{h, fo[u], s2[i]};
            }
        }
    }
}

// One workgroup orders one bin: a counting sort of its keys into 2^12 groups
// by the top bucket bits below the bin (a group holds about one key), then one
// thread per group sorts it by bucket and, within a bucket, by the table's
// order (ord_before_tied) -- the hashes and ranks it compares held in LDS.  Bins of
// up to kSortCap keys are ordered in LDS; a larger one in device memory, by
// the same steps.  Table 1 leaves every key's position at the key's slot among
// table 2's keys; table 2 is ordered with those positions as insertion ranks,
// and then every bucket of it again with its own positions as ranks, which is
// table 3's (the nodes table's) order: both hold S keys, hence have the same
// buckets and growth epochs.  Table 3's positions are the visit priorities.
constexpr int kSortThreads = 1024;
constexpr uint32_t kSortCap = 6144;
constexpr int kSortItems = kSortCap / kSortThreads;
constexpr int kGroupBits = 12;

struct BinOrderArgs {
    const OrdEntry *ent;        // both tables' entries, grouped by bin
    const uint32_t *bin_start;  // this table's bins: first slot in ent, then one past the last bin
    uint32_t base;              // slot of this table's first key
    uint64_t P, M;              // final bucket count of this table
    int shift;                  // bucket bits below the bin
    const EpochTab *et;         // growth epochs of this table
    uint32_t *r2;               // positions in table 1's order, by table 2 slot (table 1: out, table 2: in)
    uint64_t *prio;             // table 3's positions, by survivor (table 2: out)
    uint64_t *gsort;            // S words of device memory for the bins above kSortCap keys
};

// Where a bin's sort words (bucket << 32 | key index in the bin) and its keys'
// hashes and ranks are: LDS, or device memory for a large bin.
struct BinLds {
    uint64_t *srt, *h;
    uint32_t *r;
    uint8_t *ins;  // insertion epochs, computed with the ranks
    __device__ OrdKey key(uint32_t i, const EpochTab &) const {
        OrdKey o;
        o.h = h[i];
        o.rank = r[i];
        o.ins = ins[i];
        return o;
    }
};
struct BinMem {
    uint64_t *srt;
    const OrdEntry *ent;  // the bin's entries
    const uint32_t *r;    // table 2: the bin's ranks; table 1: null (ranks in the entries)
    __device__ OrdKey key(uint32_t i, const EpochTab &et) const {
        OrdKey o;
        o.h = ent[i].h;
        o.rank = r ? r[i] : ent[i].a;
        o.ins = (uint32_t)insert_epoch(et, o.rank);
        return o;
    }
};

// insertion sort of srt[s, e): by bucket (the high word), then by table order
template <class Bin>
__device__ __forceinline__ void sort_group(const Bin &bin, const EpochTab &et, uint32_t s, uint32_t e) {
    uint64_t *srt = bin.srt;
    for (uint32_t x = s + 1; x < e; x++) {
        const uint64_t v = srt[x];
        const uint32_t bv = (uint32_t)(v >> 32), iv = (uint32_t)v;
        bool have = false;
        OrdKey kv;
        uint32_t y = x;
        while (y > s) {
            const uint64_t u = srt[y - 1];
            const uint32_t bu = (uint32_t)(u >> 32), iu = (uint32_t)u;
            bool before = bv < bu;
            if (bv == bu) {
                if (!have) {
                    kv = bin.key(iv, et);
                    have = true;
                }
                before = ord_before_tied(kv, bin.key(iu, et), et);
            }
            if (!before) break;
            srt[y] = u;
            y--;
        }
        srt[y] = v;
    }
}

// table 3's order key of a sort word holding its table-2 position as the
// rank; ins: the epoch of every position of the bin, or -1 if they differ
template <class Bin>
__device__ __forceinline__ OrdKey rank_key(const Bin &bin, const EpochTab &et, uint64_t w, int ins) {
    OrdKey o = bin.key((uint32_t)w, et);
    o.rank = (uint32_t)(w >> 32);
    o.ins = (uint32_t)(ins >= 0 ? ins : insert_epoch(et, o.rank));
    return o;
}

// table 3: every run of one bucket in srt[s, e) (in table 2's order, whose
// positions start at p0 + s) ordered again with those positions as ranks
template <class Bin>
__device__ __forceinline__ void reorder_runs(const Bin &bin, const EpochTab &et, uint32_t s, uint32_t e, uint32_t p0,
                                             int ins) {
    uint64_t *srt = bin.srt;
    uint32_t rs = s;
    while (rs < e) {
        const uint32_t b = (uint32_t)(srt[rs] >> 32);
        uint32_t re = rs + 1;
        while (re < e && (uint32_t)(srt[re] >> 32) == b) re++;
        if (re - rs >= 2) {
            for (uint32_t j = rs; j < re; j++) srt[j] = ((uint64_t)(p0 + j) << 32) | (uint32_t)srt[j];
            for (uint32_t x = rs + 1; x < re; x++) {
                const uint64_t v = srt[x];
                const OrdKey kv = rank_key(bin, et, v, ins);
                uint32_t y = x;
                while (y > rs) {
                    const uint64_t u = srt[y - 1];
                    if (!ord_before_tied(kv, rank_key(bin, et, u, ins), et)) break;
                    srt[y] = u;
                    y--;
                }
                srt[y] = v;
            }
        }
        rs = re;
    }
}

// groups: counted (grp holds the counts), then scanned; srt filled by the caller
__device__ __forceinline__ uint32_t group_of(uint64_t bucket, uint32_t sub_mask, int gs) {
    return ((uint32_t)bucket & sub_mask) >> gs;
}

// sorts every group of srt, then writes the bin's positions out.  The groups
// of two or more keys are listed first (mlist, *mcount zero on entry), so
// that each thread sorts about one of them instead of walking four.
template <bool kNodes, class Bin>
__device__ __forceinline__ void finish_bin(const BinOrderArgs &a, const Bin &bin, const EpochTab &et,
                                           const uint32_t *grp, uint32_t ng, uint32_t lo, uint32_t n,
                                           uint16_t *mlist, uint32_t *mcount) {
    const uint32_t p0 = lo - a.base;
    // table 3: the bin's positions p0 .. p0 + n - 1 nearly always fall in one
    // growth epoch (the starts ascend)
    int ins = -1;
    if (kNodes) {
        const int e0 = insert_epoch(et, p0), e1 = insert_epoch(et, p0 + n - 1);
        if (e0 == e1) ins = e0;
    }
    const uint32_t lane = threadIdx.x & 63;
    for (uint32_t g = threadIdx.x; g < ng; g += kSortThreads) {  // whole waves, or one wave's first lanes
        const bool multi = grp[g] - (g ? grp[g - 1] : 0) >= 2;
        const uint64_t m = __ballot(multi);
        uint32_t at = 0;
        if (lane == 0) at = atomicAdd(mcount, (uint32_t)__popcll(m));
        at = __shfl(at, 0);
        if (multi) mlist[at + __popcll(m & ((1ull << lane) - 1))] = (uint16_t)g;
    }
    __syncthreads();
    const uint32_t nm = *mcount;
    for (uint32_t q = threadIdx.x; q < nm; q += kSortThreads) {
        const uint32_t g = mlist[q];
        const uint32_t s = g ? grp[g - 1] : 0, e = grp[g];
        sort_group(bin, et, s, e);
        if (kNodes) reorder_runs(bin, et, s, e, p0, ins);
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < n; j += kSortThreads) {
        const OrdEntry &en = a.ent[lo + (uint32_t)bin.srt[j]];
        if (kNodes)
            a.prio[en.a] = (uint64_t)(p0 + j);
        else
            a.r2[en.b] = p0 + j;
    }
}

template <bool kNodes>
__global__ void __launch_bounds__(kSortThreads) k_bin_order(BinOrderArgs a) {
    __shared__ uint64_t lsrt[kSortCap], lh[kSortCap];
    __shared__ uint32_t lr[kSortCap];
    __shared__ uint8_t lins[kSortCap];
    __shared__ uint32_t grp[1u << kGroupBits];
    __shared__ uint16_t mlist[1u << kGroupBits];
    __shared__ uint32_t mcount;
    __shared__ EpochTab et;
    const uint32_t lo = a.bin_start[blockIdx.x], n = a.bin_start[blockIdx.x + 1] - lo;
    if (n == 0) return;
    {
        const uint32_t *src = reinterpret_cast<const uint32_t *>(a.et);
        uint32_t *dst = reinterpret_cast<uint32_t *>(&et);
        for (uint32_t i = threadIdx.x; i < sizeof(EpochTab) / 4; i += kSortThreads) dst[i] = src[i];
    }
    const int gb = min(a.shift, kGroupBits), gs = a.shift - gb;
    const uint32_t ng = 1u << gb;
    const uint32_t sub_mask = a.shift >= 32 ? ~0u : (1u << a.shift) - 1u;
    const uint32_t *rin = kNodes ? a.r2 + (lo - a.base) : nullptr;
    for (uint32_t i = threadIdx.x; i < ng; i += kSortThreads) grp[i] = 0;
    if (threadIdx.x == 0) mcount = 0;
    __syncthreads();  // also: the epoch table copied
    if (n <= kSortCap) {
        // keys in registers between the count and the placement
        uint64_t bk[kSortItems];
#pragma unroll
        for (int u = 0; u < kSortItems; u++) {
            const uint32_t i = u * kSortThreads + threadIdx.x;
            if (i < n) {
                const OrdEntry en = a.ent[lo + i];
                bk[u] = mod_p(en.h, a.P, a.M);
                lh[i] = en.h;
                const uint32_t r = kNodes ? rin[i] : en.a;
                lr[i] = r;
                lins[i] = (uint8_t)insert_epoch(et, r);
                atomicAdd(&grp[group_of(bk[u], sub_mask, gs)], 1u);
            }
        }
        __syncthreads();
        block_scan_excl<(1 << kGroupBits) / kSortThreads>(grp, grp, ng);
        __syncthreads();
#pragma unroll
        for (int u = 0; u < kSortItems; u++) {
            const uint32_t i = u * kSortThreads + threadIdx.x;
            if (i < n) lsrt[atomicAdd(&grp[group_of(bk[u], sub_mask, gs)], 1u)] = (bk[u] << 32) | i;
        }
        __syncthreads();  // grp[g]: one past group g
        finish_bin<kNodes>(a, BinLds{lsrt, lh, lr, lins}, et, grp, ng, lo, n, mlist, &mcount);
    } else {
        const BinMem bin{a.gsort + (lo - a.base), a.ent + lo, rin};
        for (uint32_t i = threadIdx.x; i < n; i += kSortThreads)
            atomicAdd(&grp[group_of(mod_p(a.ent[lo + i].h, a.P, a.M), sub_mask, gs)], 1u);
        __syncthreads();
        block_scan_excl<(1 << kGroupBits) / kSortThreads>(grp, grp, ng);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < n; i += kSortThreads) {
            const uint64_t bucket = mod_p(a.ent[lo + i].h, a.P, a.M);
            bin.srt[atomicAdd(&grp[group_of(bucket, sub_mask, gs)], 1u)] = (bucket << 32) | i;
        }
        __syncthreads();
        finish_bin<kNodes>(a, bin, et, grp, ng, lo, n, mlist, &mcount);
    }
}

#define OCHECK(x)                                                            \
    do {                                                                     \
        hipError_t e_ = (x);                                                 \
        if (e_ != hipSuccess) return fail(-5, hipGetErrorString(e_));        \
    } while (0)

static inline unsigned oblk(uint64_t n) { return (unsigned)((n + 255) / 256); }

// bins of a table with P buckets: the top bits of the bucket, as many as give
// bins of at least `keys` keys on average (at most 2^kMaxBinBits bins)
static int bin_bits(uint64_t P, uint64_t S, uint64_t keys, int *shift) {
    int bits = 0;
    while (bits < 32 && (1ull << bits) < P) bits++;
    int binb = 0;
    while (binb < kMaxBinBits && binb < bits && (S >> (binb + 1)) >= keys) binb++;
    *shift = bits - binb;
    return binb;
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
int Pipeline::visit_priorities(uint64_t *d_prio, hipStream_t s) {
    const uint64_t S = n_surv;
    if (!S) return 0;
    // the cutoff copy and the nodes table: S inserts each
    *h_etabS_ = epoch_table(S, false, nullptr);
    OCHECK(etabS_.reserve(sizeof(EpochTabHost) + 64));
    OCHECK(hipMemcpyAsync(etabS_.p, h_etabS_, sizeof(EpochTabHost), hipMemcpyHostToDevice, s));
    OCHECK(ob_[O_ENT].reserve(2 * S * sizeof(OrdEntry)));
    OCHECK(ob_[O_R2].reserve(S * 4));
    const uint64_t keys = std::max<uint32_t>(tune_.order_bin_keys, 1);
    BinShape bs;
    bs.P1 = epoch1_.P[epoch1_.n - 1];
    bs.M1 = epoch1_.M[epoch1_.n - 1];
    bs.P2 = h_etabS_->P[h_etabS_->n - 1];
    bs.M2 = h_etabS_->M[h_etabS_->n - 1];
    bs.nb1 = 1u << bin_bits(bs.P1, S, keys, &bs.shift1);
    bs.nb2 = 1u << bin_bits(bs.P2, S, keys, &bs.shift2);
    const uint32_t nb = bs.nb1 + bs.nb2;
    int ncu = 0;
    OCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev_));
    const uint64_t nblk = std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)ncu, (S + 8191) / 8192));
    const uint64_t per = (S + nblk - 1) / nblk;
    const uint64_t ncnt = (uint64_t)nb * nblk;
    OCHECK(ob_[O_BCNT].reserve(ncnt * 4 + 4));
    OCHECK(ob_[O_BOFF].reserve(ncnt * 8 + 8));
    OCHECK(ob_[O_BSTART].reserve((uint64_t)(nb + 1) * 4));
    OCHECK(scan_scratch2_.reserve((scan_scratch_words(ncnt) + 16) * 8));
    // room to order any bin in device memory (a bin above kSortCap keys: far
    // above the average, which stays below 2 x order_bin_keys by default)
    OCHECK(ob_[O_GSORT].reserve(S * 8));
    const uint64_t *key = surv_key_.as<uint64_t>();
    // both tables' bins in one count and scatter pass while their cursors fit
    // LDS, else a pass per table (table 2's scatter first: table 1's entries
    // carry table-2 slots, kept in r2 until table 1's ordering rewrites it)
    // (the wide-partition testing aid takes the per-table passes on any input)
    const bool split = nb > kBinCursors || wide_parts_;
    if (split) {
        hipLaunchKernelGGL(k_bin_count, dim3((unsigned)nblk), dim3(kBinThreads), 0, s, key, S, per, bs,
                           ob_[O_BCNT].as<uint32_t>(), 1);
        hipLaunchKernelGGL(k_bin_count, dim3((unsigned)nblk), dim3(kBinThreads), 0, s, key, S, per, bs,
                           ob_[O_BCNT].as<uint32_t>(), 2);
    } else {
        hipLaunchKernelGGL(k_bin_count, dim3((unsigned)nblk), dim3(kBinThreads), 0, s, key, S, per, bs,
                           ob_[O_BCNT].as<uint32_t>(), 0);
    }
    OCHECK(scan_excl<uint32_t>(ob_[O_BCNT].as<uint32_t>(), ob_[O_BOFF].as<uint64_t>(), ncnt,
                               scan_scratch2_.as<uint64_t>(), ob_[O_BOFF].as<uint64_t>() + ncnt, s));
    // table 1's ranks are first occurrences; its epoch table's starts are the
    // first occurrences of the growth ranks, selected on the device by the build
    const int passes[2] = {split ? 2 : 0, 1};
    for (int q = 0; q < (split ? 2 : 1); q++)
        hipLaunchKernelGGL(k_bin_scatter, dim3((unsigned)nblk), dim3(kBinThreads), 0, s, key,
                           surv_first_.as<uint32_t>(), S, per, bs, ob_[O_BOFF].as<uint64_t>(),
                           ob_[O_ENT].as<OrdEntry>(), ob_[O_BSTART].as<uint32_t>(), passes[q],
                           ob_[O_R2].as<uint32_t>());
    BinOrderArgs a;
    a.ent = ob_[O_ENT].as<OrdEntry>();
    a.r2 = ob_[O_R2].as<uint32_t>();
    a.prio = d_prio;
    a.gsort = ob_[O_GSORT].as<uint64_t>();
    a.bin_start = ob_[O_BSTART].as<uint32_t>();
    a.base = 0;
    a.P = bs.P1;
    a.M = bs.M1;
    a.shift = bs.shift1;
    a.et = static_cast<const EpochTab *>(etab_.p);
    hipLaunchKernelGGL(k_bin_order<false>, dim3(bs.nb1), dim3(kSortThreads), 0, s, a);
    a.bin_start += bs.nb1;
    a.base = (uint32_t)S;
    a.P = bs.P2;
    a.M = bs.M2;
    a.shift = bs.shift2;
    a.et = static_cast<const EpochTab *>(etabS_.p);
    hipLaunchKernelGGL(k_bin_order<true>, dim3(bs.nb2), dim3(kSortThreads), 0, s, a);
    return 0;
}

}  // namespace swapk
