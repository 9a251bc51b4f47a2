// k-mer graph construction on gfx950: FASTA record scan, 2-bit packing,
// canonical k-mer + edge-mask extraction, device-local hash partitioning and
// per-partition LDS hash tables with saturating edge counters, cutoff.
//
// Replaces (single rank):
//   kmerGraph::constructKmerGraph   kmerGraph.cpp:95-310  (records, encode, edges)
//   kmerGraph::distributeKmerGraph  kmerGraph.cpp:337-432 (owner-hash exchange)
//   distNodeGraph::constructDistNodeGraph graph.cpp:1049-1090 (table, counts)
//   distNodeGraph::cutoffGraph      graph.cpp:1092-1114
#include "swap_device.h"
#include "swap_host.h"
#include "swap_scan.h"

#include <algorithm>
#include <cstdio>
#include <cstring>

namespace swapk {

// ---------------------------------------------------------------------------
// small utilities

__device__ __forceinline__ uint32_t zero_bytes(uint32_t x) {
    // exact per-byte "is zero" mask: 0x80 in each zero byte
    return ~(((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x | 0x7F7F7F7Fu);
}

__device__ __forceinline__ uint32_t base_code(uint8_t c) {
    // sanitising map of kmerGraph.cpp:228-233 followed by sequence.cpp:164-168:
    // ACGT/acgt keep their base, every other byte reads as 'A'.
    switch (c) {
        case 'C': case 'c': return 1;
        case 'G': case 'g': return 2;
        case 'T': case 't': return 3;
        default: return 0;
    }
}

// The same map on four bytes at once: the 2-bit codes of bytes 0..3 (byte 0
// first, in the top bits of the result).  c|0x20 folds case; ((u>>1)^(u>>2))&3
// is 0,1,2,3 for a,c,g,t; any other byte gives 0.
__device__ __forceinline__ uint32_t base_codes4(uint32_t t) {
    const uint32_t u = t | 0x20202020u;
    const uint32_t code = ((u >> 1) ^ (u >> 2)) & 0x03030303u;
    const uint32_t valid = zero_bytes(u ^ 0x61616161u) | zero_bytes(u ^ 0x63636363u) |
                           zero_bytes(u ^ 0x67676767u) | zero_bytes(u ^ 0x74747474u);
    const uint32_t c = code & ((valid >> 7) * 3u);
    return ((c & 3u) << 6) | (((c >> 8) & 3u) << 4) | (((c >> 16) & 3u) << 2) | ((c >> 24) & 3u);
}

// ---------------------------------------------------------------------------
// FASTA record scan

constexpr int kTileThreads = 256;
constexpr int kTileBytes = kTileThreads * 16;

// Per tile: number of '\n' bytes; global: position of the first '>'.
__global__ void k_tile_count(const uint8_t *__restrict__ buf, uint64_t n, uint32_t *__restrict__ tile_nl,
                             unsigned long long *__restrict__ first_gt) {
    const uint64_t base = (uint64_t)blockIdx.x * kTileBytes + (uint64_t)threadIdx.x * 16;
    uint32_t cnt = 0;
    unsigned long long gt = ~0ull;
    if (base + 16 <= n) {
        const uint4 v = *reinterpret_cast<const uint4 *>(buf + base);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int i = 0; i < 4; i++) {
            cnt += __popc(zero_bytes(w[i] ^ 0x0A0A0A0Au));
            const uint32_t g = zero_bytes(w[i] ^ 0x3E3E3E3Eu);
            if (g && gt == ~0ull) gt = base + 4 * i + (__ffs(g) - 1) / 8;
        }
    } else {
        for (uint64_t i = base; i < n && i < base + 16; i++) {
            const uint8_t c = buf[i];
            cnt += (c == '\n');
            if (c == '>' && gt == ~0ull) gt = i;
        }
    }
    // block reduce
    __shared__ uint32_t scnt[kTileThreads / 64];
    __shared__ unsigned long long sgt[kTileThreads / 64];
    for (int d = 32; d > 0; d >>= 1) {
        cnt += __shfl_xor(cnt, d, 64);
        unsigned long long o = __shfl_xor(gt, d, 64);
        gt = o < gt ? o : gt;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) { scnt[w] = cnt; sgt[w] = gt; }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t c = 0;
        unsigned long long g = ~0ull;
        for (int i = 0; i < kTileThreads / 64; i++) { c += scnt[i]; g = sgt[i] < g ? sgt[i] : g; }
        tile_nl[blockIdx.x] = c;
        if (g != ~0ull && g < *(volatile unsigned long long *)first_gt) atomicMin(first_gt, g);
    }
}

// The build's small counters: [0] (first '>' position, taken by atomicMin)
// starts at ~0, the other 1023 words at 0.
__global__ void k_small_init(uint64_t *__restrict__ small) { small[threadIdx.x] = threadIdx.x ? 0 : ~0ull; }

// Writes the position of every '\n' in file order.  A thread's 16 bytes give
// a 16-bit mask of newline bytes; it stores one position per set bit.
__global__ void k_tile_emit(const uint8_t *__restrict__ buf, uint64_t n, const uint64_t *__restrict__ tile_off,
                            uint64_t *__restrict__ nl_pos) {
    __shared__ uint32_t wsum[kTileThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kTileBytes + (uint64_t)threadIdx.x * 16;
    uint32_t m = 0;
    if (base + 16 <= n) {
        const uint4 v = *reinterpret_cast<const uint4 *>(buf + base);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const uint32_t z = zero_bytes(w[i] ^ 0x0A0A0A0Au);  // 0x80 in each newline byte
            m |= (((z >> 7) & 1u) | ((z >> 14) & 2u) | ((z >> 21) & 4u) | ((z >> 28) & 8u)) << (4 * i);
        }
    } else {
        for (int i = 0; i < 16; i++)
            if (base + i < n && buf[base + i] == '\n') m |= 1u << i;
    }
    const uint32_t cnt = __popc(m);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = wave_incl_sum(cnt);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t off = inc - cnt;
    for (int i = 0; i < w; i++) off += wsum[i];
    uint64_t o = tile_off[blockIdx.x] + off;
    while (m) {
        nl_pos[o++] = base + (uint32_t)(__ffs(m) - 1);
        m &= m - 1;
    }
}

// Records are line pairs starting at the first '>' (kmerGraph.cpp:173-181): the
// header is skipped, the next line is the sequence.
__global__ void k_records(const uint64_t *__restrict__ nl_pos, const uint64_t *__restrict__ d_j0, uint64_t R, int k,
                          uint64_t *__restrict__ rec_start, uint32_t *__restrict__ rec_len,
                          uint32_t *__restrict__ occ_count, uint32_t *__restrict__ word_count,
                          unsigned long long *__restrict__ n_valid) {
    const uint64_t r0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = r0 < R;
    const uint64_t r = live ? r0 : R - 1;
    const uint64_t j0 = *d_j0;
    const uint64_t s0 = nl_pos[j0 + 2 * r] + 1;
    const uint64_t s1 = nl_pos[j0 + 2 * r + 1];
    const uint64_t len64 = s1 - s0;
    const uint32_t len = len64 > 0xFFFFFFF0ull ? 0xFFFFFFF0u : (uint32_t)len64;
    const bool ok = live && len >= (uint32_t)k;
    // valid-record count: one atomic per workgroup, spread over 32 counters
    __shared__ uint32_t wcnt;
    if (threadIdx.x == 0) wcnt = 0;
    __syncthreads();
    const uint64_t bal = __ballot(ok);
    if ((threadIdx.x & 63) == 0 && bal) atomicAdd(&wcnt, (uint32_t)__popcll(bal));
    __syncthreads();
    if (threadIdx.x == 0 && wcnt) atomicAdd(&n_valid[(blockIdx.x & 31) * 16], (unsigned long long)wcnt);
    if (!live) return;
    rec_start[r] = s0;
    rec_len[r] = len;
    occ_count[r] = ok ? len - k + 1 : 0;
    word_count[r] = ok ? (len + 31) / 32 : 0;
}

// j0 = number of newlines before the first '>'; R = complete records after it.
__global__ void k_locate(const uint64_t *__restrict__ nl_pos, const uint64_t *__restrict__ d_nnl,
                         const unsigned long long *__restrict__ first_gt, uint64_t *__restrict__ out /*j0,R*/) {
    const uint64_t nnl = *d_nnl;
    const unsigned long long gt = *first_gt;
    uint64_t lo = 0, hi = nnl;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) / 2;
        if (nl_pos[mid] < gt) lo = mid + 1; else hi = mid;
    }
    out[0] = lo;
    out[1] = (gt == ~0ull) ? 0 : (nnl - lo) / 2;
}

// ---------------------------------------------------------------------------
// 2-bit packing: one thread per 32-base word, first base in the top bits.  Each
// workgroup takes 256 consecutive words; the record of each word is found by a
// binary search in an LDS copy of the word-start table around the workgroup's
// first word.  Next to the packed word goes its occurrence descriptor:
//   bits 0-31  input index of the first k-mer starting in the word
//   bits 32-37 number of k-mers starting in the word (0..32)
//   bit 40     the word is not the first of its read (a predecessor base exists)
//   bit 41     the word holds its read's last k-mer (which has no successor)

constexpr int kPackThreads = 256;

__device__ __forceinline__ uint64_t upper_index(const uint64_t *__restrict__ a, uint64_t n, uint64_t v) {
    // last i with a[i] <= v (a non-decreasing, a[0] = 0 <= v)
    uint64_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) / 2;
        if (a[mid] <= v) lo = mid; else hi = mid;
    }
    return lo;
}

// block_rec[b] = the record holding word b * kPackThreads (the record r with
// word_base[r] <= b * kPackThreads < word_base[r + 1]).
__global__ void k_block_rec(const uint64_t *__restrict__ word_base, uint64_t R, uint64_t W,
                            uint64_t *__restrict__ block_rec) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const uint64_t wb = word_base[r], we = r + 1 < R ? word_base[r + 1] : W;
    for (uint64_t b = (wb + kPackThreads - 1) / kPackThreads; b * kPackThreads < we; b++) block_rec[b] = r;
}

__global__ void __launch_bounds__(kPackThreads) k_pack(const uint8_t *__restrict__ buf, uint64_t n,
                                                       const uint64_t *__restrict__ rec_start,
                                                       const uint32_t *__restrict__ rec_len,
                                                       const uint64_t *__restrict__ word_base,
                                                       const uint64_t *__restrict__ occ_base, uint64_t R, int k,
                                                       const uint64_t *__restrict__ block_rec, uint64_t W,
                                                       uint64_t *__restrict__ packed, uint64_t *__restrict__ wmeta) {
    __shared__ uint64_t swb[kPackThreads + 1];
    for (uint64_t w0 = (uint64_t)blockIdx.x * kPackThreads; w0 < W; w0 += (uint64_t)gridDim.x * kPackThreads) {
        const uint64_t r0 = block_rec[w0 / kPackThreads];
        const uint32_t nwin = (uint32_t)min<uint64_t>(kPackThreads + 1, R - r0);
        for (uint32_t i = threadIdx.x; i < nwin; i += blockDim.x) swb[i] = word_base[r0 + i];
        __syncthreads();
        const uint64_t w = w0 + threadIdx.x;
        if (w < W) {
            uint64_t r;
            if (swb[nwin - 1] <= w) {
                // many records without k-mers between: search past the window
                const uint64_t s = r0 + nwin - 1;
                r = s + upper_index(word_base + s, R - s, w);
            } else {
                uint32_t lo = 0, hi = nwin;
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) / 2;
                    if (swb[mid] <= w) lo = mid; else hi = mid;
                }
                r = r0 + lo;
            }
            const uint64_t j = w - word_base[r];
            const uint32_t len = rec_len[r];
            const uint64_t st = rec_start[r] + 32 * j;
            const int nb = (int)min<uint64_t>(32, len - 32 * j);
            uint64_t word = 0;
            // nine aligned dwords cover the 32 bytes from any offset
            const uint32_t sh = (uint32_t)(((uintptr_t)buf + st) & 3);
            const uint64_t a0 = st - sh;  // dword-aligned byte offset
            if (a0 + 36 <= n && sh <= st) {
                const uint32_t *d = reinterpret_cast<const uint32_t *>(buf + a0);
                uint32_t v[9];
#pragma unroll
                for (int q = 0; q < 9; q++) v[q] = d[q];
#pragma unroll
                for (int q = 0; q < 8; q++)
                    word |= (uint64_t)base_codes4(__builtin_amdgcn_alignbyte(v[q + 1], v[q], sh)) << (56 - 8 * q);
                if (nb < 32) word &= ~0ull << (64 - 2 * nb);
            } else {
                const uint8_t *s = buf + st;
                for (int i = 0; i < nb; i++) word |= (uint64_t)base_code(s[i]) << (62 - 2 * i);
            }
            packed[w] = word;
            const uint32_t nk = len - k + 1;
            const uint32_t p0 = (uint32_t)(32 * j);
            const uint32_t nh = p0 < nk ? min<uint32_t>(nk - p0, 32) : 0;
            wmeta[w] = (uint64_t)(uint32_t)(occ_base[r] + p0) | ((uint64_t)nh << 32) | ((uint64_t)(j > 0) << 40) |
                       ((uint64_t)(p0 + nh == nk) << 41);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// k-mer occurrences.  One thread per packed word rolls over the (up to) 32
// k-mers that start in it.  For each occurrence it forms the canonical k-mer
// X = max(x, rc x) and the 8-bit edge mask of kmerGraph.cpp:249-299 in closed
// form:  "x followed by b" sets slot b if x is X else 4+b, and "rc x followed by
// comp(prev)" sets slot 4+comp if x is X else comp.

struct OccCtx {
    const uint64_t *packed;
    const uint64_t *wmeta;
    uint64_t W;
    int k;
    Mixer mx;
};

// Calls f(h, mask, occ) for each occurrence starting in word w: h is the mixed
// canonical k-mer (2k bits), mask its 8 edge bits, occ its input k-mer index.
// Every base read lies inside the word's read: a k-mer's bases and its
// successor base (when the read has one) end before the read does.
// One word's inputs, loaded by the caller one word ahead of their use so that
// the loads of word w + blockDim.x overlap the k-mers of word w.
struct OccWord {
    uint64_t meta, w0, w1;
    uint32_t prev_base;  // last base of the previous word (used when the read continues)
};

__device__ __forceinline__ OccWord load_word(const OccCtx &c, uint64_t w) {
    OccWord o;
    o.meta = c.wmeta[w];
    o.w0 = c.packed[w];
    o.w1 = (w + 1 < c.W) ? c.packed[w + 1] : 0;
    o.prev_base = w > 0 ? (uint32_t)(c.packed[w - 1] & 3) : 0;
    return o;
}

template <typename F>
__device__ __forceinline__ void for_each_occ_word(const OccCtx &c, const OccWord &o, F &&f) {
    const uint64_t meta = o.meta;
    const uint32_t nh = (uint32_t)(meta >> 32) & 63;
    if (nh == 0) return;
    const uint32_t occ0 = (uint32_t)meta;
    const bool has_prev = (meta >> 40) & 1;
    const bool ends = (meta >> 41) & 1;
    const int k = c.k;
    const uint64_t w0 = o.w0, w1 = o.w1;
    const uint32_t prev_base = has_prev ? o.prev_base : 0;
    const uint64_t km = kmask(k);
    uint64_t x = w0 >> (64 - 2 * k);  // bases [0, k)
    uint64_t rcx = revcomp(x, k);
    // bases k, k+1, ... at the top of a shift register: the base after k-mer i
    // is taken from its top, the base before it is the top of k-mer i-1
    uint64_t st = (w0 << (2 * k)) | (k < 32 ? w1 >> (64 - 2 * k) : 0);
    uint32_t before = prev_base;
    for (uint32_t i = 0; i < nh; i++) {
        const uint32_t next = (uint32_t)(st >> 62);
        st <<= 2;
        const bool fwd = x > rcx;
        const uint64_t X = fwd ? x : rcx;
        uint32_t m = 0;
        if (!(ends && i + 1 == nh)) m |= 1u << (fwd ? next : 4 + next);
        if (i > 0 || has_prev) {
            const uint32_t a = 3 - before;
            m |= 1u << (fwd ? 4 + a : a);
        }
        f(mix_fwd(c.mx, X), m, (uint64_t)(occ0 + i));
        before = (uint32_t)(x >> (2 * k - 2));
        x = ((x << 2) | next) & km;
        rcx = (rcx >> 2) | ((uint64_t)(3 - next) << (2 * k - 2));
    }
}

// A workgroup's words taken in groups of blockDim.x, each group reordered by
// k-mer count (descending) before its words go to the threads: a read's words
// start 32, ..., 32 k-mers, then fewer, then none, so consecutive words mix
// full and short ones in every wave, whose loop over a word's k-mers then runs
// as long as its fullest lane; grouped by count, the full words fill whole
// waves and the short ones finish early.  The next group is sorted and its
// words loaded while the current one is processed.
struct WordSort {
    uint32_t ccnt[33];  // words of each class (32 - k-mers) of the group being sorted
    uint32_t coff[33];
    uint16_t perm[1024];  // group position of the word the thread takes
};

__device__ __forceinline__ uint32_t word_nh(uint64_t meta) { return (uint32_t)(meta >> 32) & 63; }

// This thread's word (offset in the group) after sorting the group by the
// k-mer counts nh of its words (nh of thread t's unsorted word).  Three
// barriers; counters cleared for the next group.
__device__ __forceinline__ uint32_t sort_group(WordSort &ws, uint32_t nh) {
    const uint32_t cls = 32 - min(nh, 32u);
    const uint32_t rank = atomicAdd(&ws.ccnt[cls], 1u);
    __syncthreads();
    if (threadIdx.x < 64) {
        const uint32_t v = threadIdx.x < 33 ? ws.ccnt[threadIdx.x] : 0;
        const uint32_t inc = wave_incl_sum(v);
        if (threadIdx.x < 33) {
            ws.coff[threadIdx.x] = inc - v;
            ws.ccnt[threadIdx.x] = 0;
        }
    }
    __syncthreads();
    ws.perm[ws.coff[cls] + rank] = (uint16_t)threadIdx.x;
    __syncthreads();
    return ws.perm[threadIdx.x];
}

// f(word, valid) once per group for every thread (uniform: f may hold
// barriers), words of the range [wb, we) in sorted groups; invalid words
// (past we) come last in their group.  ws.ccnt must be zero on entry.
template <typename F>
__device__ __forceinline__ void for_each_word_sorted(const OccCtx &c, WordSort &ws, uint64_t wb, uint64_t we, F &&f) {
    const uint32_t T = blockDim.x, t = threadIdx.x;
    const uint64_t ngroups = (we - wb + T - 1) / T;
    if (ngroups == 0) return;
    uint64_t m = wb + t;
    uint64_t w = wb + sort_group(ws, m < we ? word_nh(c.wmeta[m]) : 0);
    OccWord cur{};
    if (w < we) cur = load_word(c, w);
    m = wb + T + t;
    uint64_t mnext = m < we ? c.wmeta[m] : 0;
    for (uint64_t gi = 0; gi < ngroups; gi++) {
        OccWord nxt{};
        uint64_t wn = we;
        if (gi + 1 < ngroups) {
            wn = wb + (gi + 1) * T + sort_group(ws, word_nh(mnext));
            if (wn < we) nxt = load_word(c, wn);
            m = wb + (gi + 2) * T + t;
            mnext = m < we ? c.wmeta[m] : 0;
        }
        f(cur, w < we);
        cur = nxt;
        w = wn;
    }
}

// Two-level radix partitioning of the occurrence records by the top bits of the
// mixed key: 2^kL1Bits buckets first, then 2^l2bits sub-buckets of each
// (l2bits = pbits - kL1Bits, 7 for inputs up to ~2^28 occurrences).
// Level-1 records keep the key bits below the level-1 bits:
//   rec = (h mod 2^(2k-8)) << 8 | mask        (2k <= 62, so 2k bits in all)
// and the level-2 bucket is read back from the record.  Both levels write each
// record straight to its place: a workgroup owns a contiguous run of every
// bucket and hands out its slots with LDS counters, so the open end of each
// run (one partial cache line per bucket per workgroup) stays in L2 while it
// fills.  Record order inside a partition is not used by the counting pass.
constexpr int kL1Bits = 8;
constexpr uint32_t kL1 = 1u << kL1Bits;
// the histogram pass bins by the top kPartBitsMin bits (level-1 bucket and the
// next kHistL2Bits); finer partitions are counted after level 1
constexpr uint32_t kHistBins = 1u << kPartBitsMin;
constexpr int kHistL2Bits = kPartBitsMin - kL1Bits;
constexpr uint32_t kHistL2 = 1u << kHistL2Bits;
constexpr int kL2MaxBits = kPartBitsMax - kL1Bits;
constexpr int kOccThreads = 1024;
constexpr int kHistRowGroups = 8;     // row groups of the column sum
constexpr int kChunkRecs = 4096;      // level-2 records per workgroup
constexpr int kL2Threads = 512;
// with 2^10 or more sub-buckets per level-1 bucket (inputs of more than
// ~2^28 k-mers) a 4096-record chunk leaves runs of 2-4 records per
// sub-bucket: chunks of 8192 records (one 1024-thread workgroup per CU) double them
constexpr int kChunkRecsWide = 8192;
constexpr int kL2ThreadsWide = 1024;
constexpr int kL2WideBits = 10;

// Occurrence records between the passes.  Reference order carries the input
// index with each record: 12 bytes {rec lo, rec hi, occ}, moved as one
// dwordx3 access; hash order moves the bare 8-byte record.
struct __attribute__((aligned(4))) Rec12 {
    uint32_t lo, hi, occ;
};

template <bool WITH_OCC>
__device__ __forceinline__ void rec_store(uint32_t *base, uint64_t g, uint64_t rec, uint32_t occ) {
    if (WITH_OCC) {
        Rec12 v;
        v.lo = (uint32_t)rec;
        v.hi = (uint32_t)(rec >> 32);
        v.occ = occ;
        reinterpret_cast<Rec12 *>(base)[g] = v;
    } else {
        reinterpret_cast<uint64_t *>(base)[g] = rec;
    }
}

template <bool WITH_OCC>
__device__ __forceinline__ uint64_t rec_load(const uint32_t *base, uint64_t g, uint32_t &occ) {
    if (WITH_OCC) {
        const Rec12 v = reinterpret_cast<const Rec12 *>(base)[g];
        occ = v.occ;
        return (uint64_t)v.lo | ((uint64_t)v.hi << 32);
    }
    occ = 0;
    return reinterpret_cast<const uint64_t *>(base)[g];
}

constexpr int rec_words(bool with_occ) { return with_occ ? 3 : 2; }

// Per workgroup word range (same split in the histogram and level-1 passes).
__device__ __forceinline__ void word_range(uint64_t W, uint64_t &wb, uint64_t &we) {
    const uint64_t per = (W + gridDim.x - 1) / gridDim.x;
    wb = per * blockIdx.x;
    we = min<uint64_t>(W, wb + per);
    if (wb > W) wb = W;
}

// Final partition sizes of the workgroup (one row per workgroup, summed by
// k_col_sum) and its level-1 counts.
__global__ void __launch_bounds__(kOccThreads) k_part_hist(OccCtx c, uint32_t *__restrict__ rows,
                                                            uint32_t *__restrict__ hist1) {
    __shared__ uint32_t lh[kHistBins];
    __shared__ WordSort ws;
    for (uint32_t i = threadIdx.x; i < kHistBins; i += blockDim.x) lh[i] = 0;
    if (threadIdx.x < 33) ws.ccnt[threadIdx.x] = 0;
    __syncthreads();
    uint64_t wb, we;
    word_range(c.W, wb, we);
    const int sh = 2 * c.k - kPartBitsMin;
    for_each_word_sorted(c, ws, wb, we, [&](const OccWord &o, bool valid) {
        if (valid) for_each_occ_word(c, o, [&](uint64_t h, uint32_t, uint64_t) { atomicAdd(&lh[h >> sh], 1u); });
    });
    __syncthreads();
    uint32_t *row = rows + (uint64_t)blockIdx.x * kHistBins;
    for (uint32_t i = threadIdx.x; i < kHistBins; i += blockDim.x) row[i] = lh[i];
    for (uint32_t b = threadIdx.x; b < kL1; b += blockDim.x) {
        uint32_t s1 = 0;
        for (uint32_t j = 0; j < kHistL2; j++) s1 += lh[b * kHistL2 + j];
        hist1[(uint64_t)b * gridDim.x + blockIdx.x] = s1;
    }
}

// out[col] += sum of rows[r][col] over one group of rows (out zeroed before).
__global__ void k_col_sum(const uint32_t *__restrict__ rows, uint32_t nrows, uint32_t ncols,
                          uint32_t *__restrict__ out) {
    const uint32_t col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= ncols) return;
    const uint32_t per = (nrows + gridDim.y - 1) / gridDim.y;
    const uint32_t r0 = blockIdx.y * per, r1 = min(nrows, r0 + per);
    uint32_t s = 0;
    for (uint32_t r = r0; r < r1; r++) s += rows[(uint64_t)r * ncols + col];
    if (s) atomicAdd(&out[col], s);
}

// The occurrences of one word as a resumable generator (the loop of
// for_each_occ_word with its state in registers), so that a pass can take a
// word's k-mers a few at a time.
struct OccGen {
    uint64_t x, rcx, st, km;
    uint32_t before, i, nh, occ0, k;
    bool has_prev, ends;
};

__device__ __forceinline__ void occgen_init(const OccCtx &c, const OccWord &o, OccGen &g) {
    const uint64_t meta = o.meta;
    const int k = c.k;
    g.nh = (uint32_t)(meta >> 32) & 63;
    g.occ0 = (uint32_t)meta;
    g.has_prev = (meta >> 40) & 1;
    g.ends = (meta >> 41) & 1;
    g.k = (uint32_t)k;
    g.km = kmask(k);
    g.x = o.w0 >> (64 - 2 * k);
    g.rcx = revcomp(g.x, k);
    g.st = (o.w0 << (2 * k)) | (k < 32 ? o.w1 >> (64 - 2 * k) : 0);
    g.before = g.has_prev ? o.prev_base : 0;
    g.i = 0;
}

// Next occurrence (g.i < g.nh): mixed canonical k-mer, edge mask, input index.
__device__ __forceinline__ void occgen_next(const OccCtx &c, OccGen &g, uint64_t &h, uint32_t &m, uint32_t &occ) {
    const uint32_t k = g.k;
    const uint32_t next = (uint32_t)(g.st >> 62);
    g.st <<= 2;
    const bool fwd = g.x > g.rcx;
    const uint64_t X = fwd ? g.x : g.rcx;
    m = 0;
    if (!(g.ends && g.i + 1 == g.nh)) m |= 1u << (fwd ? next : 4 + next);
    if (g.i > 0 || g.has_prev) {
        const uint32_t a = 3 - g.before;
        m |= 1u << (fwd ? 4 + a : a);
    }
    h = mix_fwd(c.mx, X);
    occ = g.occ0 + g.i;
    g.before = (uint32_t)(g.x >> (2 * k - 2));
    g.x = ((g.x << 2) | next) & g.km;
    g.rcx = (g.rcx >> 2) | ((uint64_t)(3 - next) << (2 * k - 2));
    g.i++;
}

// A record as staged in LDS: the level-1 record with its input index
// (reference order) or alone.
template <bool WITH_OCC>
struct Staged {
    typedef Rec12 type;
};
template <>
struct Staged<false> {
    typedef uint64_t type;
};

// Level 1: every occurrence of the workgroup's words goes to the workgroup's
// run of its level-1 bucket (run starts from the histogram pass).  The
// occurrences are taken in rounds of up to R per thread (a word's k-mers over
// ceil(32 / R) rounds); each round is counted by bucket in LDS, ordered by
// bucket in LDS and stored in that order, so that consecutive lanes write
// consecutive records of one run (scattered 12-byte stores into 256 runs per
// workgroup left partial lines and cost 1.25x the time).  Three barriers per
// round: counts complete, offsets known, round staged.  Counters, offsets and
// staging of a round alternate with the next one's where they would race.
template <bool WITH_OCC, int R>
__global__ void __launch_bounds__(kOccThreads) k_part1(OccCtx c, const uint64_t *__restrict__ base1,
                                                        uint32_t *__restrict__ rec_out) {
    typedef typename Staged<WITH_OCC>::type S;
    constexpr uint32_t kStage = (uint32_t)R * kOccThreads;
    constexpr int kRounds = (32 + R - 1) / R;
    static_assert(kStage < (1u << 16), "round ranks are 16-bit");
    __shared__ S stage[kStage];
    __shared__ uint8_t sb[kStage];
    __shared__ __align__(16) uint32_t cnt[2][kL1];
    __shared__ __align__(16) uint32_t start[kL1];
    __shared__ __align__(16) uint32_t delta[kL1];
    __shared__ __align__(16) uint32_t fill[kL1];
    __shared__ uint32_t ntot;
    const uint32_t t = threadIdx.x;
    __shared__ WordSort ws;
    for (uint32_t b = t; b < kL1; b += blockDim.x) {
        fill[b] = (uint32_t)base1[(uint64_t)b * gridDim.x + blockIdx.x];
        cnt[0][b] = 0;
        cnt[1][b] = 0;
    }
    if (t < 33) ws.ccnt[t] = 0;
    __syncthreads();
    uint64_t wb, we;
    word_range(c.W, wb, we);
    const int sh1 = 2 * c.k - kL1Bits;
    const uint64_t low1 = (1ull << sh1) - 1;
    int p = 0;
    for_each_word_sorted(c, ws, wb, we, [&](const OccWord &o, bool valid) {
        OccGen g;
        g.nh = 0;
        g.i = 0;
        if (valid) occgen_init(c, o, g);
        for (int r = 0; r < kRounds; r++) {
            S rec[R];
            uint32_t pos[R];
#pragma unroll
            for (int j = 0; j < R; j++) {
                pos[j] = ~0u;
                if (g.i < g.nh) {
                    uint64_t h;
                    uint32_t m, o;
                    occgen_next(c, g, h, m, o);
                    const uint32_t b = (uint32_t)(h >> sh1);
                    const uint64_t v = ((h & low1) << 8) | m;
                    if constexpr (WITH_OCC) {
                        rec[j].lo = (uint32_t)v;
                        rec[j].hi = (uint32_t)(v >> 32);
                        rec[j].occ = o;
                    } else {
                        rec[j] = v;
                    }
                    pos[j] = (b << 16) | atomicAdd(&cnt[p][b], 1u);
                }
            }
            __syncthreads();  // the round's counts are complete
            if (t < 64) {
                // wave 0: bucket offsets of the round (4 buckets per lane), the
                // runs' next slots, the other parity's counters cleared
                const uint4 v = *reinterpret_cast<const uint4 *>(&cnt[p][4 * t]);
                const uint32_t sum = v.x + v.y + v.z + v.w;
                const uint32_t inc = wave_incl_sum(sum);
                const uint32_t e0 = inc - sum, e1 = e0 + v.x, e2 = e1 + v.y, e3 = e2 + v.z;
                *reinterpret_cast<uint4 *>(&start[4 * t]) = make_uint4(e0, e1, e2, e3);
                const uint4 f = *reinterpret_cast<const uint4 *>(&fill[4 * t]);
                *reinterpret_cast<uint4 *>(&delta[4 * t]) = make_uint4(f.x - e0, f.y - e1, f.z - e2, f.w - e3);
                *reinterpret_cast<uint4 *>(&fill[4 * t]) = make_uint4(f.x + v.x, f.y + v.y, f.z + v.z, f.w + v.w);
                *reinterpret_cast<uint4 *>(&cnt[p ^ 1][4 * t]) = make_uint4(0, 0, 0, 0);
                if (t == 63) ntot = inc;
            }
            __syncthreads();  // offsets known
#pragma unroll
            for (int j = 0; j < R; j++) {
                if (pos[j] != ~0u) {
                    const uint32_t b = pos[j] >> 16;
                    const uint32_t s = start[b] + (pos[j] & 0xFFFFu);
                    stage[s] = rec[j];
                    sb[s] = (uint8_t)b;
                }
            }
            __syncthreads();  // the round is staged in bucket order
            const uint32_t n = ntot;
            for (uint32_t i = t; i < n; i += blockDim.x) {
                const uint32_t gi = delta[sb[i]] + i;  // < 2^32: a build holds under 2^32 records
                reinterpret_cast<S *>(rec_out)[gi] = stage[i];
            }
            p ^= 1;
        }
    });
}

// Level-1 bucket b and its first record of the chunk `chunk` (chunk_start[b] =
// first chunk index of bucket b, kL1 + 1 entries; chunks of `recs` records).
__device__ __forceinline__ bool chunk_range(uint32_t chunk, const uint32_t *__restrict__ l1_start,
                                            const uint32_t *__restrict__ chunk_start, uint32_t recs, uint32_t &b1,
                                            uint64_t &a, uint64_t &e) {
    if (chunk >= chunk_start[kL1]) return false;
    uint32_t lo = 0, hi = kL1;  // last b with chunk_start[b] <= chunk
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (chunk_start[mid] <= chunk) lo = mid; else hi = mid;
    }
    b1 = lo;
    a = (uint64_t)l1_start[lo] + (uint64_t)(chunk - chunk_start[lo]) * recs;
    e = min<uint64_t>((uint64_t)l1_start[lo + 1], a + recs);
    return true;
}

// Level 2: each workgroup takes one chunk of one level-1 bucket, counts it by
// level-2 bucket in LDS, claims a run of each final partition with one atomic
// per sub-bucket, orders the chunk by sub-bucket in LDS and writes it out in
// that order, so consecutive lanes store consecutive records of one run.
// L2MAX bounds 2^l2bits (the LDS counters are sized by it).
//
// BATCH (many samples in one build): the output record's key bits above the
// partition's (implied by the partition from here on) are replaced by the
// record's sample, found from its input index among the samples' first input
// indices sbase[0..nsamp] (sbase[0] = 0, sbase[nsamp] = all k-mers): bits
// [8, 8 + 2k - pbits) keep the key's low bits, bits 55-63 hold the sample.
constexpr int kBatchSampleShift = 55;
constexpr uint32_t kMaxBatch = (uint32_t)kMaxBatchSamples;  // 9 sample bits

__device__ __forceinline__ uint32_t sample_of(const uint32_t *sb, uint32_t nsamp, uint32_t occ) {
    uint32_t lo = 0, hi = nsamp;  // last s with sb[s] <= occ (empty samples share their start)
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (sb[mid] <= occ) lo = mid; else hi = mid;
    }
    return lo;
}

template <bool WITH_OCC, int L2MAX, bool BATCH, int CHUNK, int THREADS>
__global__ void __launch_bounds__(THREADS) k_part2(const uint32_t *__restrict__ rec_in,
                                                       const uint32_t *__restrict__ l1_start,
                                                       const uint32_t *__restrict__ chunk_start, int k, int l2bits,
                                                       uint32_t *__restrict__ cursor, uint32_t *__restrict__ rec_out,
                                                       const uint32_t *__restrict__ sbase, uint32_t nsamp) {
    static_assert(!BATCH || WITH_OCC, "a batch build keeps the input index of every record");
    constexpr int kPer = CHUNK / THREADS;
    __shared__ uint32_t lcnt[L2MAX];
    __shared__ uint32_t lstart[L2MAX];
    __shared__ uint32_t gbase[L2MAX];
    __shared__ uint32_t slo[CHUNK], shi[CHUNK];  // the chunk in sub-bucket order
    __shared__ uint32_t socc[WITH_OCC ? CHUNK : 1];
    __shared__ uint32_t lsb[BATCH ? kMaxBatch + 1 : 1];
    const uint32_t nl2 = 1u << l2bits, m2 = nl2 - 1;
    uint32_t b1;
    uint64_t a, e;
    if (!chunk_range(blockIdx.x, l1_start, chunk_start, CHUNK, b1, a, e)) return;
    for (uint32_t b = threadIdx.x; b < nl2; b += blockDim.x) lcnt[b] = 0;
    if (BATCH)
        for (uint32_t s = threadIdx.x; s <= nsamp; s += blockDim.x) lsb[s] = sbase[s];
    __syncthreads();
    const uint32_t n = (uint32_t)(e - a);
    const int sh2 = 8 + 2 * k - (kL1Bits + l2bits);
    uint64_t r[kPer];
    uint32_t o[kPer], rank[kPer];
#pragma unroll
    for (int j = 0; j < kPer; j++) {
        const uint32_t i = j * THREADS + threadIdx.x;
        if (i < n) {
            r[j] = rec_load<WITH_OCC>(rec_in, a + i, o[j]);
            rank[j] = atomicAdd(&lcnt[(r[j] >> sh2) & m2], 1u);
        }
    }
    __syncthreads();
    block_scan_excl<(L2MAX + THREADS - 1) / THREADS>(lcnt, lstart, nl2);
    for (uint32_t b = threadIdx.x; b < nl2; b += blockDim.x) {
        const uint32_t c = lcnt[b];
        gbase[b] = c ? atomicAdd(&cursor[((uint64_t)b1 << l2bits) + b], c) : 0;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kPer; j++) {
        const uint32_t i = j * THREADS + threadIdx.x;
        if (i < n) {
            const uint32_t pos = lstart[(r[j] >> sh2) & m2] + rank[j];
            slo[pos] = (uint32_t)r[j];
            shi[pos] = (uint32_t)(r[j] >> 32);
            if (WITH_OCC) socc[pos] = o[j];
        }
    }
    __syncthreads();
    const uint64_t low = (1ull << sh2) - 1;  // mask and key bits below the partition's
    for (uint32_t i = threadIdx.x; i < n; i += THREADS) {
        uint64_t rec = (uint64_t)slo[i] | ((uint64_t)shi[i] << 32);
        const uint32_t b = (uint32_t)(rec >> sh2) & m2;
        const uint32_t oc = WITH_OCC ? socc[i] : 0u;
        if (BATCH) rec = (rec & low) | ((uint64_t)sample_of(lsb, nsamp, oc) << kBatchSampleShift);
        rec_store<WITH_OCC>(rec_out, (uint64_t)gbase[b] + (i - lstart[b]), rec, oc);
    }
}

// Level-1 bucket starts l1_start[b] = pstart[b << l2bits] (pstart: partition
// starts at l2bits sub-buckets per bucket), and the chunk table of chunks of
// `recs` records: one thread per bucket (blockDim.x == kL1), chunk counts
// scanned across the block.
__global__ void __launch_bounds__(kL1) k_l1_chunks(const uint32_t *__restrict__ pstart, int l2bits, uint32_t recs,
                                                   uint32_t *__restrict__ l1_start,
                                                   uint32_t *__restrict__ chunk_start) {
    __shared__ uint32_t wtot[kL1 / 64];
    const uint32_t b = threadIdx.x;
    const uint32_t s0 = pstart[(uint64_t)b << l2bits], s1 = pstart[(uint64_t)(b + 1) << l2bits];
    const uint32_t n = (s1 - s0 + recs - 1) / recs;
    const uint32_t inc = wave_incl_sum(n);
    if ((b & 63) == 63) wtot[b >> 6] = inc;
    __syncthreads();
    uint32_t off = inc - n;
    for (uint32_t w = 0; w < (b >> 6); w++) off += wtot[w];
    l1_start[b] = s0;
    chunk_start[b] = off;
    if (b == kL1 - 1) {
        l1_start[kL1] = s1;
        chunk_start[kL1] = off + n;
    }
}

// Final partition sizes when they are finer than the histogram pass's bins:
// each workgroup counts one chunk of one level-1 bucket by sub-bucket in LDS
// and adds its non-zero counts (large chunks keep the device atomics few).
constexpr uint32_t kFineChunkRecs = 1u << 16;

template <bool WITH_OCC>
__global__ void __launch_bounds__(1024) k_fine_hist(const uint32_t *__restrict__ rec_in,
                                                    const uint32_t *__restrict__ l1_start,
                                                    const uint32_t *__restrict__ chunk_start, int k, int l2bits,
                                                    uint32_t *__restrict__ sizes) {
    __shared__ uint32_t lh[1u << kL2MaxBits];
    const uint32_t nl2 = 1u << l2bits, m2 = nl2 - 1;
    uint32_t b1;
    uint64_t a, e;
    if (!chunk_range(blockIdx.x, l1_start, chunk_start, kFineChunkRecs, b1, a, e)) return;
    for (uint32_t b = threadIdx.x; b < nl2; b += blockDim.x) lh[b] = 0;
    __syncthreads();
    const int sh2 = 8 + 2 * k - (kL1Bits + l2bits);
    for (uint64_t i = a + threadIdx.x; i < e; i += blockDim.x) {
        uint32_t o;
        const uint64_t r = rec_load<WITH_OCC>(rec_in, i, o);
        atomicAdd(&lh[(r >> sh2) & m2], 1u);
    }
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < nl2; b += blockDim.x)
        if (lh[b]) atomicAdd(&sizes[((uint64_t)b1 << l2bits) + b], lh[b]);
}

// ---------------------------------------------------------------------------
// Per-partition counting.  A workgroup owns one partition at a time: its
// records go into an open-addressing table in LDS (key inserted by CAS), the 8
// edge counters of a slot are 16-bit LDS adds with 2-slot buckets probed in
// order (partitions under 2^16 records) or saturating bytes in one 64-bit word
// with linear probing, min(count, 255) of graph.cpp:1074-1078 either way, and
// the first-occurrence index is an atomicMin.  A partition whose distinct keys overflow the LDS table is redone
// by the global-memory variant of the same kernel.

constexpr int kTabBits = 12;
constexpr uint32_t kTab = 1u << kTabBits;
constexpr int kCountThreads = 1024;
constexpr uint64_t kEmpty = ~0ull;

struct CountArgs {
    const uint32_t *rec;      // final partitions, rec_words() words per record
    const uint32_t *bucket;   // start of each final partition, nparts + 1 entries
    uint32_t nparts;          // 2^pbits
    int l2bits;               // pbits - kL1Bits
    int k;
    int c;
    Mixer mx;
    uint64_t *surv_key;
    uint64_t *surv_mult;
    uint32_t *surv_first;
    unsigned long long *counters;  // [0]=survivors [1]=distinct [2]=overflow partitions
    uint32_t *part_surv;           // survivors of each partition (the survivor index sizes its regions by them)
    uint32_t *part_first;          // first survivor of each partition's run, ~0u if not one run
    uint32_t *overflow_list;
    int spill;                     // send every partition to the device-memory table
    uint32_t *dfo;                 // first-occurrence index of every distinct key (reference order), or null
    int gbits;                     // 2^gbits consecutive partitions share one fill of the LDS table
    // batch builds: record keys carry their sample above the partition's key
    // bits; per sample, the distinct keys and the largest first occurrence
    int batch;
    int pbits;
    uint32_t nsamp;
    uint32_t *samp_distinct;
    uint32_t *samp_last;
};

// The mixed key h of a counted key of partition `part`.  Ordinary builds keep
// the 2k - 8 bits below the level-1 bucket; batch builds the 2k - pbits bits
// below the partition, with the sample above them.
__device__ __forceinline__ uint64_t counted_key(const CountArgs &a, bool batch, uint32_t part, uint64_t key) {
    if (batch) {
        const int lb = 2 * a.k - a.pbits;
        return ((uint64_t)part << lb) | (key & ((1ull << lb) - 1));
    }
    return ((uint64_t)(part >> a.l2bits) << (2 * a.k - kL1Bits)) | key;
}
__device__ __forceinline__ uint32_t counted_sample(uint64_t key) {
    return (uint32_t)(key >> (kBatchSampleShift - 8));
}

__device__ __forceinline__ uint32_t slot_hash(uint64_t key, uint32_t mask) {
    return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 40) & mask;
}

// Inserts / updates one record; returns false if the table is full.
template <bool WITH_OCC>
__device__ __forceinline__ bool table_add(uint64_t *tkey, uint64_t *tcnt, uint32_t *tfo, uint32_t mask,
                                          uint64_t rec, uint32_t occ) {
    const uint64_t key = rec >> 8;
    const uint32_t m = (uint32_t)(rec & 0xFF);
    uint32_t s = slot_hash(key, mask);
    for (uint32_t probe = 0; probe <= mask; probe++) {
        uint64_t cur = tkey[s];
        if (cur == kEmpty) {
            const uint64_t prev = atomicCAS((unsigned long long *)&tkey[s], (unsigned long long)kEmpty,
                                            (unsigned long long)key);
            cur = (prev == kEmpty) ? key : prev;
        }
        if (cur == key) {
            if (m) {
                const uint64_t inc = mask_to_bytes(m);
                uint64_t old = tcnt[s];
                while (true) {
                    const uint64_t nv = sat_add_bytes(old, inc);
                    if (nv == old) break;
                    const uint64_t got = atomicCAS((unsigned long long *)&tcnt[s], (unsigned long long)old,
                                                   (unsigned long long)nv);
                    if (got == old) break;
                    old = got;
                }
            }
            if (WITH_OCC) atomicMin(&tfo[s], occ);
            return true;
        }
        s = (s + 1) & mask;
    }
    return false;
}

// Edge counters of one occurrence on slot s (16-bit counters, non-returning
// LDS adds).  An occurrence has at most two edge bits (one per side).
__device__ __forceinline__ void add_edges16(uint32_t *tc, uint32_t s, uint32_t m) {
    if (!m) return;
    const int t0 = __ffs(m) - 1;
    atomicAdd(&tc[4 * s + (t0 >> 1)], 1u << (16 * (t0 & 1)));
    m &= m - 1;
    if (!m) return;
    const int t1 = __ffs(m) - 1;
    atomicAdd(&tc[4 * s + (t1 >> 1)], 1u << (16 * (t1 & 1)));
    m &= m - 1;
    while (m) {  // not produced by the occurrence pass; kept for completeness
        const int t = __ffs(m) - 1;
        m &= m - 1;
        atomicAdd(&tc[4 * s + (t >> 1)], 1u << (16 * (t & 1)));
    }
}

// Bucketed probing for the 16-bit table: a key's home is a 2-slot bucket (16
// bytes of keys, one LDS read), buckets are probed in order.  Keys enter the
// first empty slot by CAS, so the filled slots of the table form a prefix of
// every key's probe sequence and a key is never stored twice.  The table runs
// under half full on the partitions the partition count aims at, where most
// lookups end in the home bucket: 2-slot buckets (one read) measured 5%
// faster than 4-slot ones (two reads).
constexpr uint32_t kBucketSlots = 2;
constexpr uint32_t kBuckets = kTab / kBucketSlots;

// The keys of the bucket at slot s0, in 16-byte LDS reads.
__device__ __forceinline__ void load_bucket(const uint64_t *tkey, uint32_t s0, uint64_t (&v)[kBucketSlots]) {
#pragma unroll
    for (uint32_t h = 0; h < kBucketSlots; h += 2) {
        const ulonglong2 x = *reinterpret_cast<const ulonglong2 *>(&tkey[s0 + h]);
        v[h] = x.x;
        v[h + 1] = x.y;
    }
}

__device__ __forceinline__ uint32_t home_bucket(uint64_t key) {
    return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> 40) & (kBuckets - 1);
}

// When 2^g partitions share the table, partition q of them owns sub-table q:
// the buckets whose slots belong to threads [q, q + 1) * kCountThreads / 2^g
// in every round of the emit (bucket b holds slots kBucketSlots * b..., owned
// by threads kBucketSlots * b mod kCountThreads...), so each partition's
// output is a run of whole waves.  Bucket of local index l (0 <= l < kBuckets / 2^g) of sub-table q:
constexpr int kRoundBucketBits = __builtin_ctz(kCountThreads / kBucketSlots);
__device__ __forceinline__ uint32_t sub_bucket(uint32_t l, uint32_t q, int g) {
    const int rb = kRoundBucketBits - g;
    return ((l >> rb) << kRoundBucketBits) | (q << rb) | (l & ((1u << rb) - 1));
}

// Finds or inserts key starting at local bucket l0 of sub-table q (fresh
// reads, one bucket at a time); returns its slot, or ~0u when the sub-table
// is full.
__device__ uint32_t bucket_probe(uint64_t *tkey, uint64_t key, uint32_t l0, uint32_t q, int g) {
    const uint32_t nsub = kBuckets >> g;
    for (uint32_t n = 0; n < nsub; n++) {
        const uint32_t s0 = sub_bucket((l0 + n) & (nsub - 1), q, g) * kBucketSlots;
        uint64_t v[kBucketSlots];
        load_bucket(tkey, s0, v);
#pragma unroll
        for (uint32_t j = 0; j < kBucketSlots; j++) {
            if (v[j] == key) return s0 + j;
            if (v[j] == kEmpty) {
                const uint64_t prev = atomicCAS((unsigned long long *)&tkey[s0 + j], (unsigned long long)kEmpty,
                                                (unsigned long long)key);
                if (prev == kEmpty || prev == key) return s0 + j;
            }
        }
    }
    return ~0u;
}

// Inserts / updates one record in the 16-bit table: one read of its home
// bucket, a CAS on the bucket's first empty slot when the key is not there,
// and bucket_probe only when the bucket is full of other keys or the CAS lost
// to another key.  Returns false when the table is full.
template <bool WITH_OCC>
__device__ __forceinline__ bool table_add16(uint64_t *tkey, uint32_t *tc, uint32_t *tfo, uint64_t rec,
                                            uint32_t occ, uint32_t q, int g) {
    const uint64_t key = rec >> 8;
    const uint32_t l0 = home_bucket(key) & ((kBuckets >> g) - 1);
    const uint32_t s0 = sub_bucket(l0, q, g) * kBucketSlots;
    uint64_t v[kBucketSlots];
    load_bucket(tkey, s0, v);
    // first slot of the bucket that holds the key or is empty
    uint32_t slot = ~0u;
    int cq = -1;
#pragma unroll
    for (int j = kBucketSlots - 1; j >= 0; j--) {
        if (v[j] == key) {
            slot = s0 + j;
            cq = -1;
        } else if (v[j] == kEmpty) {
            slot = ~0u;
            cq = j;
        }
    }
    if (cq >= 0) {
        const uint64_t prev = atomicCAS((unsigned long long *)&tkey[s0 + cq], (unsigned long long)kEmpty,
                                        (unsigned long long)key);
        if (prev == kEmpty || prev == key) slot = s0 + cq;
    }
    if (slot == ~0u) slot = bucket_probe(tkey, key, l0, q, g);
    if (slot == ~0u) return false;
    if (WITH_OCC) atomicMin(&tfo[slot], occ);
    add_edges16(tc, slot, (uint32_t)(rec & 0xFF));
    return true;
}

// Records whose key is not in its home bucket and cannot enter it (the
// bucket is full of other keys, or the CAS lost to another key) are queued in
// LDS during a fill's inserts and probed after them by all threads together:
// probing inline made nearly every wave run the probe loop for its few such
// lanes.  Counts and first-occurrence minima commute, so the order of the
// updates does not matter; a key enters the first empty slot of its probe
// sequence whichever path inserts it (buckets never lose keys during a fill).
constexpr uint32_t kQueueCap = 2048;

// The home-bucket part of table_add16: true when the record was counted,
// false when it needs the probe.
template <bool WITH_OCC>
__device__ __forceinline__ bool table_add16_home(uint64_t *tkey, uint32_t *tc, uint32_t *tfo, uint64_t rec,
                                                 uint32_t occ, uint32_t q, int g) {
    const uint64_t key = rec >> 8;
    const uint32_t l0 = home_bucket(key) & ((kBuckets >> g) - 1);
    const uint32_t s0 = sub_bucket(l0, q, g) * kBucketSlots;
    uint64_t v[kBucketSlots];
    load_bucket(tkey, s0, v);
    uint32_t slot = ~0u;
    int cq = -1;
#pragma unroll
    for (int j = kBucketSlots - 1; j >= 0; j--) {
        if (v[j] == key) {
            slot = s0 + j;
            cq = -1;
        } else if (v[j] == kEmpty) {
            slot = ~0u;
            cq = j;
        }
    }
    if (cq >= 0) {
        const uint64_t prev = atomicCAS((unsigned long long *)&tkey[s0 + cq], (unsigned long long)kEmpty,
                                        (unsigned long long)key);
        if (prev == kEmpty || prev == key) slot = s0 + cq;
    }
    if (slot == ~0u) return false;
    if (WITH_OCC) atomicMin(&tfo[slot], occ);
    add_edges16(tc, slot, (uint32_t)(rec & 0xFF));
    return true;
}

// The probe part: returns false when the sub-table is full.
template <bool WITH_OCC>
__device__ __forceinline__ bool table_add16_probe(uint64_t *tkey, uint32_t *tc, uint32_t *tfo, uint64_t rec,
                                                  uint32_t occ, uint32_t q, int g) {
    const uint64_t key = rec >> 8;
    const uint32_t slot = bucket_probe(tkey, key, home_bucket(key) & ((kBuckets >> g) - 1), q, g);
    if (slot == ~0u) return false;
    if (WITH_OCC) atomicMin(&tfo[slot], occ);
    add_edges16(tc, slot, (uint32_t)(rec & 0xFF));
    return true;
}

// table_add16 with the probe deferred to the queue (probed inline when the
// queue is full).  Returns false when the sub-table is full.
template <bool WITH_OCC>
__device__ __forceinline__ bool table_add16_q(uint64_t *tkey, uint32_t *tc, uint32_t *tfo, uint64_t rec, uint32_t occ,
                                              uint32_t q, int g, uint64_t *qrec, uint32_t *qocc, uint32_t *qn) {
    if (table_add16_home<WITH_OCC>(tkey, tc, tfo, rec, occ, q, g)) return true;
    const uint32_t qi = atomicAdd(qn, 1u);
    if (qi < kQueueCap) {
        qrec[qi] = rec;
        if (WITH_OCC) qocc[qi] = occ;
        return true;
    }
    return table_add16_probe<WITH_OCC>(tkey, tc, tfo, rec, occ, q, g);
}

// min(count, 255) and the cutoff of graph.cpp:1100-1112 on the 16-bit counters
// of one slot, two counters per packed 16-bit operation: a counter survives
// when min(count, 255) + (0x8000 - c) has its top bit set, i.e. when it is >= c.
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
typedef short i16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t bias16(int c) {
    const uint32_t b = 0x8000u - (uint32_t)min(max(c, 0), 256);
    return b | (b << 16);
}

__device__ __forceinline__ uint64_t cut16_to_bytes(const uint32_t *tc, uint32_t s, uint32_t bias) {
    uint32_t w[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        u16x2 x = __builtin_bit_cast(u16x2, tc[4 * s + q]);
        x = __builtin_elementwise_min(x, (u16x2)(255));
        const u16x2 t = x + __builtin_bit_cast(u16x2, bias);
        const i16x2 keep = __builtin_bit_cast(i16x2, t) >> 15;
        w[q] = __builtin_bit_cast(uint32_t, x & __builtin_bit_cast(u16x2, keep));
    }
    // counter 2q is byte 0 of w[q], counter 2q + 1 its byte 2
    const uint32_t lo = __builtin_amdgcn_perm(w[1], w[0], 0x06040200u);
    const uint32_t hi = __builtin_amdgcn_perm(w[3], w[2], 0x06040200u);
    return (uint64_t)lo | ((uint64_t)hi << 32);
}

// Cutoff of graph.cpp:1100-1112 on one slot: multiplicities below c become 0.
__device__ __forceinline__ uint64_t apply_cutoff(uint64_t cnt, int c) {
    uint64_t out = 0;
#pragma unroll
    for (int t = 0; t < 8; t++) {
        const uint32_t v = (uint32_t)((cnt >> (8 * t)) & 0xFF);
        if ((int)v >= c) out |= (uint64_t)v << (8 * t);
    }
    return out;
}

// Emits the survivors of one partition table and (reference order) the
// first-occurrence index of every distinct key.  Slots are taken 4 per thread
// per round; one atomic per workgroup and round reserves the output ranges.
template <bool WITH_OCC, typename CNT>
__device__ void table_emit(const uint64_t *tkey, CNT cnt_of, const uint32_t *tfo, uint32_t n_slots,
                           uint32_t part, const CountArgs &a) {
    __shared__ uint64_t wsum[kCountThreads / 64];
    __shared__ unsigned long long base_s, base_d;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    constexpr int kPer = 4;  // one round over the 4096-slot table
    for (uint32_t r0 = 0; r0 < n_slots; r0 += blockDim.x * kPer) {
        uint64_t keys[kPer], mults[kPer];
        uint32_t fos[kPer];
        uint32_t ns = 0, nd = 0;
#pragma unroll
        for (int j = 0; j < kPer; j++) {
            const uint32_t s = r0 + threadIdx.x * kPer + j;
            keys[j] = kEmpty;
            mults[j] = 0;
            fos[j] = 0;
            if (s < n_slots) {
                keys[j] = tkey[s];
                if (keys[j] != kEmpty) {
                    mults[j] = apply_cutoff(cnt_of(s), a.c);
                    if (WITH_OCC) fos[j] = tfo[s];
                    nd++;
                    ns += mults[j] != 0;
                }
            }
        }
        // block exclusive scan of (survivors, distinct) packed in one u64
        const uint64_t v = ((uint64_t)nd << 32) | ns;
        const uint64_t inc = wave_incl_sum(v);
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t run = 0;
            for (int i = 0; i < nw; i++) { const uint64_t t = wsum[i]; wsum[i] = run; run += t; }
            const uint32_t ts = (uint32_t)run, td = (uint32_t)(run >> 32);
            base_s = ts ? atomicAdd(&a.counters[0], (unsigned long long)ts) : 0;
            base_d = td ? atomicAdd(&a.counters[1], (unsigned long long)td) : 0;
            if (ts) atomicAdd(&a.part_surv[part], ts);
        }
        __syncthreads();
        const uint64_t ex = inc - v + wsum[w];
        uint64_t os = base_s + (uint32_t)ex, od = base_d + (ex >> 32);
#pragma unroll
        for (int j = 0; j < kPer; j++) {
            if (keys[j] == kEmpty) continue;
            if (WITH_OCC && a.dfo) a.dfo[od++] = fos[j];
            if (a.batch) {
                const uint32_t sm = counted_sample(keys[j]);
                atomicAdd(&a.samp_distinct[sm], 1u);
                atomicMax(&a.samp_last[sm], fos[j]);
            }
            if (mults[j]) {
                a.surv_key[os] = counted_key(a, a.batch, part, keys[j]);  // mixed; k_unmix inverts
                a.surv_mult[os] = mults[j];
                if (WITH_OCC) a.surv_first[os] = fos[j];
                os++;
            }
        }
        __syncthreads();
    }
}

// Records of the next partition each thread loads one partition ahead.
// Partitions hold N / 2^15 records on average (about 4900 for a 50x bacterial
// sample: under 5 per thread); the rest load inside the insert loop.
constexpr int kPre = 6;
// Table fills per workgroup whose bounds are staged in LDS (grid >= fills / this).
constexpr uint32_t kMaxFills = 512;
// Partitions that may share one table fill (2^kMaxFillBits).  The records a
// shared fill aims at are tuned (Tuning::fill_recs, 4096 by default: one record
// per sub-table slot on average; reads at 20x coverage or more carry one
// distinct key per 3-5 records, so sub-tables run under a third full); an
// input of nearly all distinct keys overflows shared fills to the
// device-memory table instead.
constexpr int kMaxFillBits = 3;

// Persistent and software-pipelined: one workgroup per CU walks the table
// fills blockIdx.x, blockIdx.x + gridDim.x, ...; fill f counts partitions
// f * 2^gbits ... (f + 1) * 2^gbits - 1 (consecutive records), each in its own
// sub-table.  Iteration i inserts fill i's records (loaded during iteration
// i-1) into the LDS table; then each thread stores its share of fill i-1's
// output (the output range reserved by one device atomic during iteration
// i-1), takes its 4 slots of fill i into registers, clears them, and the
// workgroup's (survivors, distinct) totals are scanned; thread 0 then reserves
// fill i's output range while fill i+1's records load.  The global-memory
// latencies of a fill (its record loads, the reservation, the output stores)
// thus overlap the LDS work of its neighbours.  Slot j*kCountThreads + t
// belongs to thread t, so the emit and the clear touch consecutive LDS words
// across a wave.  Sharing a fill among the small partitions of a small input
// divides the per-fill costs (barriers, the reservation on one device-wide
// counter, which every workgroup's fills serialise on) by 2^gbits.
//
// BATCH: the keys are (sample, key bits below the partition) and each
// workgroup tallies every sample's distinct keys and largest first occurrence
// in LDS, added to the device totals once at its end.
template <bool WITH_OCC, bool SHARED, bool BATCH>
__global__ void __launch_bounds__(kCountThreads) k_count_lds(CountArgs a) {
    static_assert(kTab == 4 * kCountThreads, "the emit takes 4 slots per thread");
    static_assert(!BATCH || (WITH_OCC && !SHARED), "batch builds count one partition per fill");
    __shared__ __align__(16) uint64_t tkey[kTab];
    __shared__ uint32_t tc[4 * kTab];  // 8 x u16 counters per slot, or 8 x u8 in the first 2*kTab words
    __shared__ uint32_t tfo[WITH_OCC ? kTab : 1];
    __shared__ uint64_t wsum[kCountThreads / 64];
    __shared__ unsigned long long base_s, base_d;
    __shared__ int overflow[2];  // set during the insert of iterations of each parity
    __shared__ uint64_t qrec[kQueueCap];  // records left for the probe (16-bit tables)
    __shared__ uint32_t qocc[WITH_OCC ? kQueueCap : 1];
    __shared__ uint32_t qn;
    __shared__ uint32_t bnd[kMaxFills + 1][2];  // [i] = record bounds of this workgroup's i-th fill
    __shared__ uint32_t sdist[BATCH ? kMaxBatch : 1], slast[BATCH ? kMaxBatch : 1];
    const uint32_t t = threadIdx.x;
    const int lane = t & 63, w = t >> 6;
    const uint32_t nparts = a.nparts;
    const int g = SHARED ? a.gbits : 0;  // single fills: the sub-table arithmetic folds away
    const uint32_t nfills = nparts >> g, qmask = (1u << g) - 1;
    // record bits that select the level-2 sub-bucket, i.e. the partition
    const int qsh = 8 + 2 * a.k - (kL1Bits + a.l2bits);
    constexpr int kWaves = kCountThreads / 64;
    const int wpp = kWaves >> g;  // waves per partition of a fill
    const uint32_t total = a.bucket[nparts];  // >= 1: the host skips empty builds
    const uint32_t bias = bias16(a.c);
    uint64_t *tcnt = reinterpret_cast<uint64_t *>(tc);
    for (uint32_t i = t; i < kTab; i += kCountThreads) {
        tkey[i] = kEmpty;
        if (WITH_OCC) tfo[i] = ~0u;
    }
    for (uint32_t i = t; i < 4 * kTab; i += kCountThreads) tc[i] = 0;
    if (t < 2) overflow[t] = 0;
    if (t == 0) qn = 0;
    if (BATCH)
        for (uint32_t i = t; i < a.nsamp; i += kCountThreads) sdist[i] = slast[i] = 0;
    for (uint32_t i = t; i <= kMaxFills; i += kCountThreads) {
        const uint64_t f = blockIdx.x + (uint64_t)i * gridDim.x;
        bnd[i][0] = f < nfills ? a.bucket[f << g] : 0;
        bnd[i][1] = f < nfills ? a.bucket[(f + 1) << g] : 0;
    }
    __syncthreads();

    // next fill's records: loaded from indices clamped to the fill's last
    // record, so that every load is issued and none reaches into the next fill
    // (entries past the fill's end are never used)
    uint64_t pr[kPre];
    uint32_t po[kPre];
    uint32_t b = 0, e = 0;
    auto prefetch = [&](uint32_t it) {
        b = bnd[it][0];
        e = bnd[it][1];
        const uint64_t last = e > b ? (uint64_t)e - 1 : min((uint64_t)b, (uint64_t)total - 1);
#pragma unroll
        for (int j = 0; j < kPre; j++) {
            const uint64_t i = min((uint64_t)b + t + (uint64_t)j * kCountThreads, last);
            pr[j] = rec_load<WITH_OCC>(a.rec, i, po[j]);
        }
    };
    // previous fill, held until its output range is known
    uint64_t keys[4], mults[4];
    uint32_t fos[4];
    uint64_t pv = 0, pinc = 0;
    uint32_t prev_p0 = 0;   // the previous fill's first partition
    uint32_t prev_off = 0;  // wave 0, lane q: survivors of the previous fill before its partition q
    bool have_prev = false, prev_ovf = false;
    unsigned long long pend = 0;  // thread 0: the output base reserved for the previous fill
#pragma unroll
    for (int j = 0; j < 4; j++) keys[j] = kEmpty;

    prefetch(0);
    for (uint32_t fill = blockIdx.x, it = 0;; fill += gridDim.x, it++) {
        const bool have = fill < nfills;  // uniform over the workgroup
        const uint32_t p0 = fill << g;
        const bool small = (e - b) < 0xFFFFu;
        bool ok = true;
        if (have) {
            if (small) {
                // 16-bit counters cannot wrap below 2^16 records
#pragma unroll
                for (int j = 0; j < kPre; j++) {
                    const uint64_t i = (uint64_t)b + t + (uint64_t)j * kCountThreads;
                    if (i < e && ok)
                        ok = table_add16_q<WITH_OCC>(tkey, tc, tfo, pr[j], po[j], (uint32_t)(pr[j] >> qsh) & qmask, g,
                                                      qrec, qocc, &qn);
                }
                for (uint64_t i = (uint64_t)b + t + (uint64_t)kPre * kCountThreads; i < e && ok; i += kCountThreads) {
                    uint32_t o;
                    const uint64_t r = rec_load<WITH_OCC>(a.rec, i, o);
                    ok = table_add16_q<WITH_OCC>(tkey, tc, tfo, r, o, (uint32_t)(r >> qsh) & qmask, g, qrec, qocc, &qn);
                }
            } else if (g > 0) {
                ok = false;  // a shared fill of 2^16 records: its partitions go to the device-memory table
            } else {
#pragma unroll
                for (int j = 0; j < kPre; j++) {
                    const uint64_t i = (uint64_t)b + t + (uint64_t)j * kCountThreads;
                    if (i < e && ok) ok = table_add<WITH_OCC>(tkey, tcnt, tfo, kTab - 1, pr[j], po[j]);
                }
                for (uint64_t i = (uint64_t)b + t + (uint64_t)kPre * kCountThreads; i < e && ok; i += kCountThreads) {
                    uint32_t o;
                    const uint64_t r = rec_load<WITH_OCC>(a.rec, i, o);
                    ok = table_add<WITH_OCC>(tkey, tcnt, tfo, kTab - 1, r, o);
                }
            }
            if (!ok) overflow[it & 1] = 1;
        }
        // the next fill's records are requested as soon as this fill's are
        // inserted: their loads are in flight through the output stores of the
        // previous fill, this fill's table read-out and both barriers
        if (fill + gridDim.x < nfills) prefetch(it + 1);
        if (w == 0 && have_prev && !prev_ovf) {
            // the reservation returns (distinct << 32) | survivors before this fill
            const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)pend);
            const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(pend >> 32));
            if (lane == 0) {
                base_s = lo;
                base_d = hi;
            }
            if (lane <= qmask) a.part_first[prev_p0 + lane] = lo + prev_off;
        }
        __syncthreads();  // inserts done, queue complete; previous output bases and wave offsets visible
        if (have && small) {
            // the queued records, probed by all threads
            const uint32_t nq = min(qn, kQueueCap);
            bool qok = true;
            for (uint32_t i = t; i < nq && qok; i += kCountThreads) {
                const uint64_t r = qrec[i];
                qok = table_add16_probe<WITH_OCC>(tkey, tc, tfo, r, WITH_OCC ? qocc[i] : 0u,
                                                  (uint32_t)(r >> qsh) & qmask, g);
            }
            if (!qok) overflow[it & 1] = 1;
        }
        if (have_prev && !prev_ovf) {
            const uint64_t ex = pinc - pv + wsum[w];
            uint64_t os = base_s + (uint32_t)ex, od = base_d + (ex >> 32);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if (keys[j] == kEmpty) continue;
                if (WITH_OCC && a.dfo) a.dfo[od++] = fos[j];
                if (BATCH) {
                    const uint32_t sm = counted_sample(keys[j]);
                    atomicAdd(&sdist[sm], 1u);
                    atomicMax(&slast[sm], fos[j]);
                }
                if (mults[j]) {
                    a.surv_key[os] = counted_key(a, BATCH, prev_p0, keys[j]);  // mixed; k_unmix inverts
                    a.surv_mult[os] = mults[j];
                    if (WITH_OCC) a.surv_first[os] = fos[j];
                    os++;
                }
            }
        }
        if (!have) break;
        __syncthreads();  // table complete
        if (t == 0) qn = 0;  // read by every thread before the barrier above
        const bool ovf = overflow[it & 1] != 0 || a.spill;
        uint32_t ns = 0, nd = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t s = (uint32_t)j * kCountThreads + t;
            const uint64_t kk = tkey[s];
            keys[j] = ovf ? kEmpty : kk;
            mults[j] = 0;
            fos[j] = 0;
            if (keys[j] != kEmpty) {
                mults[j] = small ? cut16_to_bytes(tc, s, bias) : apply_cutoff(tcnt[s], a.c);
                if (WITH_OCC) fos[j] = tfo[s];
                nd++;
                ns += mults[j] != 0;
            }
            // a slot's key, counters and first occurrence are read only by its
            // owner; counters and first occurrence change only in slots that
            // hold a key
            if (kk != kEmpty) {
                tkey[s] = kEmpty;
                if (WITH_OCC) tfo[s] = ~0u;
                if (small) {
#pragma unroll
                    for (int q = 0; q < 4; q++) tc[4 * s + q] = 0;
                } else {
                    tcnt[s] = 0;
                }
            }
        }
        // workgroup exclusive scan of (survivors, distinct) packed in one u64
        pv = ((uint64_t)nd << 32) | ns;
        pinc = wave_incl_sum(pv);
        if (lane == 63) wsum[w] = pinc;
        __syncthreads();  // wave sums complete; the table is empty
        if (w == 0) {
            // wave offsets: exclusive scan of the per-wave totals by wave 0
            const uint64_t x = (!ovf && lane < kWaves) ? wsum[lane] : 0;
            const uint64_t incw = wave_incl_sum(x);
            if (!ovf && lane < kWaves) wsum[lane] = incw - x;
            const uint64_t run = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(incw >> 32), kWaves - 1) << 32) |
                                 (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)incw, kWaves - 1);
            // survivors of the fill's partitions: partition q holds waves
            // [q, q + 1) * wpp, lane q takes the inclusive sums before and at its end
            const uint32_t inc_s = (uint32_t)incw;
            const int q = lane & (kWaves - 1);
            const uint32_t before = q ? (uint32_t)__shfl((int)inc_s, (q * wpp - 1) & (kWaves - 1)) : 0u;
            const uint32_t at_end = (uint32_t)__shfl((int)inc_s, ((q + 1) * wpp - 1) & (kWaves - 1));
            if (lane <= qmask) {
                if (ovf) {
                    const unsigned long long j = atomicAdd(&a.counters[2], 1ull);
                    a.overflow_list[j] = p0 + lane;
                    a.part_surv[p0 + lane] = 0;  // counted by the global-memory pass
                    a.part_first[p0 + lane] = ~0u;
                } else {
                    a.part_surv[p0 + lane] = at_end - before;
                    prev_off = before;  // part_first follows with the fill's output base
                }
            }
            if (lane == 0) {
                overflow[it & 1] = 0;
                if (!ovf) {
                    // one reservation of (distinct << 32) | survivors; a per-lane
                    // (zero) index keeps the compiler from turning it into a
                    // wave-aggregated atomic, whose result it would wait for here;
                    // the result is first used after the next insert
                    uint32_t zero;
                    asm volatile("v_mov_b32 %0, 0" : "=v"(zero));
                    pend = run ? atomicAdd(&a.counters[3 + zero], (unsigned long long)run) : 0;
                }
            }
        }
        have_prev = true;
        prev_ovf = ovf;
        prev_p0 = p0;
    }
    if (BATCH) {
        __syncthreads();  // every thread left the loop after the same fill
        for (uint32_t i = t; i < a.nsamp; i += kCountThreads)
            if (sdist[i]) {
                atomicAdd(&a.samp_distinct[i], sdist[i]);
                atomicMax(&a.samp_last[i], slast[i]);
            }
    }
    // the last workgroup to finish publishes the totals of the reservations as
    // the survivor and distinct counters (the device-memory pass appends to them)
    if (t == 0) {
        __threadfence();
        if (atomicAdd(&a.counters[5], 1ull) == gridDim.x - 1) {
            const unsigned long long tot = atomicAdd(&a.counters[3], 0ull);
            a.counters[0] = (uint32_t)tot;
            a.counters[1] = tot >> 32;
        }
    }
}

// The count kernels store survivors by their mixed key h; one pass turns them
// into k-mers (kept out of the count kernel, where it would run once per
// slot on every wave).
__global__ void k_unmix(uint64_t *__restrict__ key, uint64_t n, Mixer mx) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) key[i] = mix_inv(mx, key[i]);
}

// Global-memory fallback for partitions with more distinct keys than the LDS
// table holds.  Workgroup j takes overflow partition overflow_list[j] and a table
// of tab_size[j] slots at tab_off[j].
template <bool WITH_OCC>
__global__ void __launch_bounds__(kCountThreads) k_count_global(CountArgs a, const uint64_t *__restrict__ tab_off,
                                                                 const uint32_t *__restrict__ tab_bits,
                                                                 uint64_t *gkey, uint64_t *gcnt, uint32_t *gfo) {
    const uint32_t part = a.overflow_list[blockIdx.x];
    const uint64_t off = tab_off[blockIdx.x];
    const uint32_t n_slots = 1u << tab_bits[blockIdx.x];
    uint64_t *tkey = gkey + off;
    uint64_t *tcnt = gcnt + off;
    uint32_t *tfo = WITH_OCC ? gfo + off : nullptr;
    for (uint32_t i = threadIdx.x; i < n_slots; i += blockDim.x) {
        tkey[i] = kEmpty;
        tcnt[i] = 0;
        if (WITH_OCC) tfo[i] = ~0u;
    }
    __syncthreads();
    const uint32_t b = a.bucket[part], e = a.bucket[part + 1];
    for (uint32_t i = b + threadIdx.x; i < e; i += blockDim.x) {
        uint32_t o;
        const uint64_t r = rec_load<WITH_OCC>(a.rec, i, o);
        table_add<WITH_OCC>(tkey, tcnt, tfo, n_slots - 1, r, o);
    }
    __threadfence_block();
    __syncthreads();
    table_emit<WITH_OCC>(tkey, [&](uint32_t s) { return tcnt[s]; }, tfo, n_slots, part, a);
}

// ---------------------------------------------------------------------------
// Reference order support: the reference's k-mer table grows (rehashes) when
// its (P+1)-th distinct key arrives, P running through a fixed prime list.
// Which survivors were inserted before each growth follows from the first-
// occurrence index of the distinct key of each growth rank, selected from the
// first-occurrence indices of all distinct keys by a histogram pass and a
// collection pass over the few buckets that hold the wanted ranks.

constexpr int kSelBits = 14;
constexpr uint32_t kSelBins = 1u << kSelBits;

// One histogram row per workgroup (summed by k_col_sum).
__global__ void k_sel_hist(const uint32_t *__restrict__ dfo, const unsigned long long *__restrict__ d_n, int shift,
                           uint32_t *__restrict__ rows, unsigned long long *__restrict__ maxfo) {
    __shared__ uint32_t lh[kSelBins];
    __shared__ uint32_t lmax;
    for (uint32_t i = threadIdx.x; i < kSelBins; i += blockDim.x) lh[i] = 0;
    if (threadIdx.x == 0) lmax = 0;
    __syncthreads();
    const uint64_t n = *d_n;
    uint32_t mx = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t f = dfo[i];
        atomicAdd(&lh[f >> shift], 1u);
        mx = f > mx ? f : mx;
    }
    atomicMax(&lmax, mx);
    __syncthreads();
    uint32_t *row = rows + (uint64_t)blockIdx.x * kSelBins;
    for (uint32_t i = threadIdx.x; i < kSelBins; i += blockDim.x) row[i] = lh[i];
    if (threadIdx.x == 0) atomicMax(maxfo, (unsigned long long)lmax);
}

// Device-side selection bookkeeping (no host round trip): bin and in-bin
// position of each wanted rank, and the compact list of bins to mark.
struct SelInfo {
    int n_starts;
    int n_targets;
    int tgt[32];        // target index of each start
    uint32_t pos[32];   // position of the start's rank inside its bin
    uint32_t tbin[32];  // bin of each target
};

__global__ void __launch_bounds__(1024) k_sel_targets(const uint32_t *__restrict__ hist, const uint64_t *__restrict__ starts,
                                                      int n_starts, int *__restrict__ target, SelInfo *__restrict__ si,
                                                      uint32_t *__restrict__ sel_out, uint32_t sel_out_words) {
    // the marking pass's bitmaps start empty
    for (uint32_t i = threadIdx.x; i < sel_out_words; i += 1024) sel_out[i] = 0;
    __shared__ uint32_t cum[kSelBins + 1];
    __shared__ uint32_t wsum[16];
    constexpr int per = kSelBins / 1024;
    const int t = threadIdx.x;
    uint32_t loc[per];
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < per; j++) { loc[j] = hist[t * per + j]; sum += loc[j]; }
    const uint32_t inc = wave_incl_sum(sum);
    if ((t & 63) == 63) wsum[t >> 6] = inc;
    __syncthreads();
    uint32_t off = inc - sum;
    for (int w = 0; w < (t >> 6); w++) off += wsum[w];
#pragma unroll
    for (int j = 0; j < per; j++) { cum[t * per + j] = off; off += loc[j]; }
    if (t == 1023) cum[kSelBins] = off;
    for (uint32_t b = t; b < kSelBins; b += 1024) target[b] = -1;
    __syncthreads();
    if (t == 0) {
        int nt = 0;
        for (int j = 0; j < n_starts && j < 32; j++) {
            const uint64_t T = starts[j];
            uint32_t lo = 0, hi = kSelBins;  // last bin with cum <= T
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) / 2;
                if (cum[mid] <= T) lo = mid; else hi = mid;
            }
            int ti = -1;
            for (int q = 0; q < nt; q++) if (si->tbin[q] == lo) ti = q;
            if (ti < 0) {
                ti = nt++;
                si->tbin[ti] = lo;
                target[lo] = ti;
            }
            si->tgt[j] = ti;
            si->pos[j] = (uint32_t)(T - cum[lo]);
        }
        si->n_starts = n_starts;
        si->n_targets = nt;
    }
}

// Marks the first occurrences that fall in a target bin in that bin's bitmap
// (first occurrences of distinct keys are distinct, so the bitmap is exact).
// Only a few bins are targets: the atomics are few and spread over the bitmap.
__global__ void k_sel_mark(const uint32_t *__restrict__ dfo, const unsigned long long *__restrict__ d_n, int shift,
                           const int *__restrict__ target, uint32_t *__restrict__ bitmap) {
    // the few target bins as a bit per bin in LDS: the target table is read
    // only for the first occurrences that fall in one
    __shared__ uint32_t is_target[kSelBins / 32];
    for (uint32_t i = threadIdx.x; i < kSelBins / 32; i += blockDim.x) {
        uint32_t w = 0;
        for (uint32_t j = 0; j < 32; j++) w |= (target[32 * i + j] >= 0 ? 1u : 0u) << j;
        is_target[i] = w;
    }
    __syncthreads();
    const uint64_t n = *d_n;
    const uint32_t nw = (1u << shift) >= 32 ? (1u << shift) / 32 : 1;
    const uint32_t in_bin = (1u << shift) - 1;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t f = dfo[i];
        const uint32_t bin = f >> shift;
        if (!((is_target[bin >> 5] >> (bin & 31)) & 1)) continue;
        const int t = target[bin];
        if (t >= 0) {
            const uint32_t bit = f & in_bin;
            atomicOr(&bitmap[(uint32_t)t * nw + (bit >> 5)], 1u << (bit & 31));
        }
    }
}

// Workgroup per target bin: the wanted ranks by prefix popcount over the bin's
// bitmap.  Writes the first occurrence of each growth rank into the epoch
// table of the k-mer table.
__global__ void __launch_bounds__(1024) k_sel_pick(const SelInfo *__restrict__ si, const uint32_t *__restrict__ bitmap,
                                                   int shift, uint64_t *__restrict__ epoch_start) {
    __shared__ uint32_t part[1024];
    __shared__ uint32_t found;
    const int tg = blockIdx.x;
    if (tg >= si->n_targets) return;
    const uint32_t nw = (1u << shift) >= 32 ? (1u << shift) / 32 : 1;
    const uint32_t *bm = bitmap + (uint64_t)tg * nw;
    const uint32_t per = (nw + blockDim.x - 1) / blockDim.x;
    uint32_t c = 0;
    for (uint32_t j = 0; j < per; j++) {
        const uint32_t wi = threadIdx.x * per + j;
        if (wi < nw) c += __popc(bm[wi]);
    }
    part[threadIdx.x] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int i = 0; i < (int)blockDim.x; i++) { const uint32_t v = part[i]; part[i] = run; run += v; }
    }
    __syncthreads();
    const uint32_t base = si->tbin[tg] << shift;
    for (int j = 0; j < si->n_starts; j++) {
        if (si->tgt[j] != tg) continue;
        const uint32_t want = si->pos[j];  // 0-based rank inside the bin
        if (threadIdx.x == 0) found = 0xFFFFFFFFu;
        __syncthreads();
        const uint32_t lo = part[threadIdx.x];
        if (want >= lo && want < lo + c) {
            uint32_t r = want - lo;
            for (uint32_t q = 0; q < per; q++) {
                const uint32_t wi = threadIdx.x * per + q;
                if (wi >= nw) break;
                uint32_t w = bm[wi];
                const uint32_t pc = __popc(w);
                if (r < pc) {
                    for (uint32_t z = 0; z < r; z++) w &= w - 1;  // drop the r lowest set bits
                    found = wi * 32 + (__ffs(w) - 1);
                    break;
                }
                r -= pc;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) epoch_start[j] = (uint64_t)base + found;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Batch builds: many samples' FASTA bytes concatenated and built together.
// A sample's records are its own when every sample starts with '>' and ends
// with a newline and an even number of lines precedes every sample (records
// are line pairs from the first '>'): then no record spans two samples, and
// the records of sample s are those whose sequence starts in its bytes.

// One thread per sample boundary s = 0..nsamp: rfirst[s] = first record of
// sample s (R at the end), kbase[s] = its first input k-mer index (all k-mers
// at the end).  bad[0] is set when a sample breaks the rule above.
__global__ void k_batch_bounds(const uint8_t *__restrict__ buf, const uint64_t *__restrict__ off, uint32_t nsamp,
                               const uint64_t *__restrict__ nl_pos, const uint64_t *__restrict__ d_small,
                               const uint64_t *__restrict__ rec_start, const uint64_t *__restrict__ occ_base,
                               uint64_t *__restrict__ rfirst, uint32_t *__restrict__ kbase, uint32_t *__restrict__ bad) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > nsamp) return;
    const uint64_t nnl = d_small[1], R = d_small[3];
    const uint64_t b = off[s];
    if (s < nsamp) {
        const uint64_t e = off[s + 1];
        uint64_t lo = 0, hi = nnl;  // newlines before byte b
        while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            if (nl_pos[mid] < b) lo = mid + 1; else hi = mid;
        }
        if (e <= b || buf[b] != '>' || buf[e - 1] != '\n' || (lo & 1)) atomicOr(bad, 1u);
    }
    uint64_t lo = 0, hi = R;  // records whose sequence starts before byte b
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (rec_start[mid] < b) lo = mid + 1; else hi = mid;
    }
    rfirst[s] = lo;
    kbase[s] = (uint32_t)(lo < R ? occ_base[lo] : d_small[4]);  // the scan leaves its total in d_small[4]
}

// Reads of each sample that hold at least one k-mer (the count a build of the
// sample alone reports).
__global__ void __launch_bounds__(256) k_batch_reads(const uint32_t *__restrict__ occ_count, uint64_t R,
                                                     const uint64_t *__restrict__ rfirst, uint32_t nsamp,
                                                     unsigned long long *__restrict__ reads) {
    __shared__ uint64_t lrf[kMaxBatch + 1];
    __shared__ uint32_t lcnt[kMaxBatch];
    for (uint32_t i = threadIdx.x; i <= nsamp; i += blockDim.x) lrf[i] = rfirst[i];
    for (uint32_t i = threadIdx.x; i < nsamp; i += blockDim.x) lcnt[i] = 0;
    __syncthreads();
    for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < R; r += (uint64_t)gridDim.x * blockDim.x) {
        if (!occ_count[r]) continue;
        uint32_t lo = 0, hi = nsamp;  // last s with rfirst[s] <= r
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (lrf[mid] <= r) lo = mid; else hi = mid;
        }
        atomicAdd(&lcnt[lo], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nsamp; i += blockDim.x)
        if (lcnt[i]) atomicAdd(&reads[i], (unsigned long long)lcnt[i]);
}

// Growth ranks of many samples' k-mer tables at once.  Sample s's distinct
// keys hold the global first-occurrence ranks [D_<s, D_<s + D_s) (its input
// indices all lie above the earlier samples'), so its growth rank r is the
// global order statistic D_<s + r.  One workgroup: bin of each wanted rank
// (ranks ascending) and its position in the bin, the target bins numbered in
// bin order (target[bin] = its bitmap, -1 elsewhere).
__global__ void __launch_bounds__(1024) k_bsel_targets(const uint32_t *__restrict__ hist,
                                                       const uint64_t *__restrict__ rank, uint32_t T,
                                                       int *__restrict__ target, uint32_t *__restrict__ t_slot,
                                                       uint32_t *__restrict__ t_pos, uint32_t *__restrict__ slot_bin) {
    __shared__ uint32_t cum[kSelBins + 1];
    __shared__ uint8_t flag[kSelBins];
    __shared__ uint32_t wsum[16];
    constexpr int per = kSelBins / 1024;
    const int t = threadIdx.x;
    uint32_t loc[per];
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < per; j++) { loc[j] = hist[t * per + j]; sum += loc[j]; flag[t * per + j] = 0; }
    uint32_t inc = wave_incl_sum(sum);
    if ((t & 63) == 63) wsum[t >> 6] = inc;
    __syncthreads();
    uint32_t off = inc - sum;
    for (int w = 0; w < (t >> 6); w++) off += wsum[w];
#pragma unroll
    for (int j = 0; j < per; j++) { cum[t * per + j] = off; off += loc[j]; }
    if (t == 1023) cum[kSelBins] = off;
    __syncthreads();
    for (uint32_t i = t; i < T; i += 1024) {
        const uint64_t r = rank[i];
        uint32_t lo = 0, hi = kSelBins;  // last bin with cum <= r
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (cum[mid] <= r) lo = mid; else hi = mid;
        }
        t_pos[i] = (uint32_t)(r - cum[lo]);
        t_slot[i] = lo;  // the bin for now
        flag[lo] = 1;
    }
    __syncthreads();
    // number the target bins: exclusive scan of the flags
    sum = 0;
#pragma unroll
    for (int j = 0; j < per; j++) sum += flag[t * per + j];
    inc = wave_incl_sum(sum);
    __syncthreads();  // wsum reused
    if ((t & 63) == 63) wsum[t >> 6] = inc;
    __syncthreads();
    off = inc - sum;
    for (int w = 0; w < (t >> 6); w++) off += wsum[w];
#pragma unroll
    for (int j = 0; j < per; j++) {
        const uint32_t b = t * per + j;
        if (flag[b]) {
            target[b] = (int)off;
            slot_bin[off] = b;
            cum[b] = off;  // the bin's slot (the cumulative counts are no longer read)
            off++;
        } else {
            target[b] = -1;
        }
    }
    __syncthreads();
    for (uint32_t i = t; i < T; i += 1024) t_slot[i] = cum[t_slot[i]];
}

// Workgroup per wanted rank: its value by prefix popcount over its bin's
// bitmap (marked by k_sel_mark).
__global__ void __launch_bounds__(1024) k_bsel_pick(const uint32_t *__restrict__ t_slot,
                                                    const uint32_t *__restrict__ t_pos,
                                                    const uint32_t *__restrict__ slot_bin,
                                                    const uint32_t *__restrict__ bitmap, int shift,
                                                    uint64_t *__restrict__ value) {
    __shared__ uint32_t found;
    const uint32_t tg = blockIdx.x;
    const uint32_t slot = t_slot[tg], want = t_pos[tg];
    const uint32_t nw = (1u << shift) >= 32 ? (1u << shift) / 32 : 1;
    const uint32_t *bm = bitmap + (uint64_t)slot * nw;
    const uint32_t per = (nw + blockDim.x - 1) / blockDim.x;
    uint32_t c = 0;
    for (uint32_t j = 0; j < per; j++) {
        const uint32_t wi = threadIdx.x * per + j;
        if (wi < nw) c += __popc(bm[wi]);
    }
    const uint32_t inc = wave_incl_sum(c);
    __shared__ uint32_t wsum[16];
    if ((threadIdx.x & 63) == 63) wsum[threadIdx.x >> 6] = inc;
    if (threadIdx.x == 0) found = 0xFFFFFFFFu;
    __syncthreads();
    uint32_t lo = inc - c;  // set bits of the bitmap before this thread's words
    for (uint32_t w = 0; w < (threadIdx.x >> 6); w++) lo += wsum[w];
    if (want >= lo && want < lo + c) {
        uint32_t r = want - lo;
        for (uint32_t q = 0; q < per; q++) {
            const uint32_t wi = threadIdx.x * per + q;
            if (wi >= nw) break;
            uint32_t w = bm[wi];
            const uint32_t pc = __popc(w);
            if (r < pc) {
                for (uint32_t z = 0; z < r; z++) w &= w - 1;  // drop the r lowest set bits
                found = wi * 32 + (__ffs(w) - 1);
                break;
            }
            r -= pc;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) value[tg] = ((uint64_t)slot_bin[slot] << shift) + found;
}

// First survivor of each sample once the survivors are in first-occurrence
// order: surv0[s] = survivors whose first occurrence lies below kbase[s].
__global__ void k_batch_slices(const uint32_t *__restrict__ first, uint64_t S, const uint32_t *__restrict__ kbase,
                               uint32_t nsamp, uint64_t *__restrict__ surv0) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s > nsamp) return;
    const uint32_t b = kbase[s];
    uint64_t lo = 0, hi = S;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (first[mid] < b) lo = mid + 1; else hi = mid;
    }
    surv0[s] = s == nsamp ? S : lo;
}

// A sample's survivors taken from a batch build: first occurrences relative to
// its first input k-mer, and its survivors per count partition of its own.
__global__ void k_rebase_first(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst, uint64_t n,
                               uint32_t base) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i] - base;
}

__global__ void k_part_surv_count(const uint64_t *__restrict__ key, uint64_t n, Mixer mx, int shift,
                                  uint32_t *__restrict__ part_surv) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(&part_surv[mix_fwd(mx, key[i]) >> shift], 1u);
}

// ---------------------------------------------------------------------------
// Pipeline::build

static thread_local Arena *t_arena = nullptr;

Arena *current_arena() { return t_arena; }
ArenaScope::ArenaScope(Arena *a) : prev(t_arena) { t_arena = a; }
ArenaScope::~ArenaScope() { t_arena = prev; }

DevBuf::~DevBuf() {
    if (heap) (void)hipFree(heap);
}

hipError_t DevBuf::reserve(size_t bytes) {
    if (Arena *a = t_arena) {
        if (arena_epoch == a->epoch && p && bytes <= cap) return hipSuccess;
        const size_t want = (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
        a->wanted += want;
        if (a->used + want > a->size) return hipErrorOutOfMemory;
        p = a->base + a->used;
        cap = want;
        a->used += want;
        arena_epoch = a->epoch;
        return hipSuccess;
    }
    if (bytes <= cap && p) return hipSuccess;  // own memory, or a static call's space still being read
    arena_epoch = 0;
    if (heap && bytes <= heap_cap) {
        p = heap;
        cap = heap_cap;
        return hipSuccess;
    }
    p = nullptr;
    cap = 0;
    if (heap) {
        const hipError_t e = hipFree(heap);
        heap = nullptr;
        heap_cap = 0;
        if (e != hipSuccess) return e;
    }
    size_t want = bytes + bytes / 8 + 256;
    hipError_t e = hipMalloc(&heap, want);
    if (e != hipSuccess) { heap = nullptr; return e; }
    p = heap;
    cap = heap_cap = want;
    return hipSuccess;
}

void Pipeline::leave_arena() {
    DevBuf *bufs[] = {&fasta_, &tile_nl_, &tile_off_, &nl_pos_, &scan_scratch_, &small_, &rec_start_, &rec_len_,
                      &occ_count_, &word_count_, &occ_base_, &word_base_, &block_rec_, &packed_, &wmeta_, &rows_,
                      &hist_, &bucket_, &cursor_, &hist1_, &base1_, &l1info_, &rec1_, &rec_, &surv_key_, &surv_mult_,
                      &surv_first_, &dfo_, &sel_, &sel_target_, &sel_out_, &overflow_, &part_surv_, &part_first_,
                      &gtab_key_, &gtab_cnt_, &gtab_fo_, &etab_, &etabS_, &scan_scratch2_, &bdev_, &bsel_, &bbits_};
    for (DevBuf *b : bufs) b->leave_arena();
    for (DevBuf &b : mb_) b.leave_arena();
    for (DevBuf &b : ob_) b.leave_arena();
    for (DevBuf &b : rn_) b.leave_arena();
}

static uint64_t inv_odd(uint64_t a) {
    uint64_t x = a;  // Newton iteration for the inverse mod 2^64
    for (int i = 0; i < 6; i++) x *= 2 - a * x;
    return x;
}

Pipeline::Pipeline(int device, int k, int c, bool reference_order) : dev_(device), k_(k), c_(c), ref_order_(reference_order) {
    // the first failure is kept in init_err_ (checked by the C API at create)
    auto keep = [&](hipError_t e) {
        if (e != hipSuccess && init_err_ == hipSuccess) init_err_ = e;
    };
    keep(hipSetDevice(dev_));
    keep(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));
    keep(hipStreamCreateWithFlags(&st2_, hipStreamNonBlocking));
    keep(hipEventCreateWithFlags(&sel_done_, hipEventDisableTiming));
    keep(hipEventCreateWithFlags(&prio_done_, hipEventDisableTiming));
    keep(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
    keep(hipHostMalloc((void **)&h_small_, 1024 * sizeof(uint64_t), hipHostMallocDefault));
    keep(hipHostMalloc((void **)&h_starts_, 64 * sizeof(uint64_t), hipHostMallocDefault));
    keep(hipHostMalloc((void **)&h_etab_, sizeof(EpochTabHost), hipHostMallocDefault));
    keep(hipHostMalloc((void **)&h_etabS_, sizeof(EpochTabHost), hipHostMallocDefault));
    reconfigure(k, c, reference_order);
    for (auto &e : ev_) keep(hipEventCreate(&e));
}

void Pipeline::reconfigure(int k, int c, bool reference_order) {
    k_ = k;
    c_ = c;
    ref_order_ = reference_order;
    mx_.bits = 2 * k;
    mx_.m = kmask(k);
    // Fibonacci hashing on 2k bits: 2^2k / golden ratio, odd
    mx_.a = (0x9E3779B97F4A7C15ull >> (64 - 2 * k)) | 1;
    mx_.ia = inv_odd(mx_.a);
    mx_.s = k;  // half of the 2k bits
    n_reads = n_occ = n_distinct = n_surv = n_kept = n_overflow = n_special = 0;
}

// Phase timing is diagnostic: a failed event call leaves that phase at 0 ms.
void Pipeline::phase_begin(int i, hipStream_t s) {
    ev_used_[i] = hipEventRecord(ev_[2 * i], s ? s : st_) == hipSuccess;
}
void Pipeline::phase_end(int i, hipStream_t s) {
    if (hipEventRecord(ev_[2 * i + 1], s ? s : st_) != hipSuccess) ev_used_[i] = false;
}
// (a phase recorded on the side stream and not yet ended reads 0 ms)
void Pipeline::collect_phases(bool side) {
    const bool synced = hipStreamSynchronize(st_) == hipSuccess && (!side || hipStreamSynchronize(st2_) == hipSuccess);
    for (int i = 0; i < kPhases; i++) {
        phase_ms[i] = 0;
        if (synced && ev_used_[i] && hipEventElapsedTime(&phase_ms[i], ev_[2 * i], ev_[2 * i + 1]) != hipSuccess)
            phase_ms[i] = 0;
    }
}

Pipeline::~Pipeline() {
    // teardown: nothing useful can be done about a failing release here
    (void)hipSetDevice(dev_);
    if (st2_) (void)hipStreamSynchronize(st2_);
    for (auto &e : ev_)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : {sel_done_, prio_done_, fork_})
        if (e) (void)hipEventDestroy(e);
    if (st2_) (void)hipStreamDestroy(st2_);
    if (h_small_) (void)hipHostFree(h_small_);
    if (h_starts_) (void)hipHostFree(h_starts_);
    if (h_etab_) (void)hipHostFree(h_etab_);
    if (h_etabS_) (void)hipHostFree(h_etabS_);
    if (st_) (void)hipStreamDestroy(st_);
}

int Pipeline::fail(int code, const char *msg) {
    err_ = msg;
    return code;
}

#define SWAP_CHECK(x)                                                        \
    do {                                                                     \
        hipError_t e_ = (x);                                                 \
        if (e_ != hipSuccess) return fail(-5, hipGetErrorString(e_));        \
    } while (0)

// Layout of bsel_ in a batch build: per-sample distinct counts and largest
// first occurrences (u32), the samples' first survivors (u64), and per wanted
// growth rank (at most 32 per sample): rank (u64), slot, position in its bin,
// bin of each slot (u32), value (u64).
static uint64_t batch_sel_bytes(uint64_t ns) { return ns * 8 + (ns + 1) * 8 + 32 * ns * (8 + 4 + 4 + 4 + 8) + 256; }

int Pipeline::build(const uint8_t *d_in, uint64_t n) {
    SWAP_CHECK(hipSetDevice(dev_));
    // the previous build's selection or merge's replay may still read buffers
    // this build rewrites or grows
    SWAP_CHECK(hipStreamSynchronize(st2_));
    if (!current_arena()) leave_arena();  // an ordinary build never writes a static call's space
    n_reads = n_occ = n_distinct = n_surv = n_kept = n_overflow = n_special = 0;
    surv_renumbered_ = false;
    if (n == 0) return fail(-2, "empty input");
    const uint8_t *buf = d_in;
    if ((reinterpret_cast<uintptr_t>(d_in) & 15) != 0) {
        SWAP_CHECK(fasta_.reserve(n + 16));
        SWAP_CHECK(hipMemcpyAsync(fasta_.p, d_in, n, hipMemcpyDeviceToDevice, st_));
        buf = fasta_.as<uint8_t>();
    }
    uint64_t *hs = h_small_;
    // --- newline scan
    const uint64_t tiles = (n + kTileBytes - 1) / kTileBytes;
    SWAP_CHECK(tile_nl_.reserve(tiles * 4));
    SWAP_CHECK(tile_off_.reserve(tiles * 8));
    const uint64_t scan_words = scan_scratch_words(std::max<uint64_t>(tiles, 1)) + 16;
    SWAP_CHECK(small_.reserve(1024 * 8));
    uint64_t *d_small = small_.as<uint64_t>();  // [0]=first_gt [1]=nnl [2]=j0 [3]=R [4]=N [5]=W [8..]=counters
    hipLaunchKernelGGL(k_small_init, dim3(1), dim3(1024), 0, st_, d_small);
    for (int i = 0; i < kPhases; i++) ev_used_[i] = false;
    phase_begin(0);
    hipLaunchKernelGGL(k_tile_count, dim3((unsigned)tiles), dim3(kTileThreads), 0, st_, buf, n,
                       tile_nl_.as<uint32_t>(), (unsigned long long *)d_small);
    SWAP_CHECK(scan_scratch_.reserve(scan_words * 8));
    SWAP_CHECK(scan_excl<uint32_t>(tile_nl_.as<uint32_t>(), tile_off_.as<uint64_t>(), tiles, scan_scratch_.as<uint64_t>(),
                        d_small + 1, st_));
    SWAP_CHECK(hipMemcpyAsync(hs, d_small, 2 * 8, hipMemcpyDeviceToHost, st_));
    SWAP_CHECK(hipStreamSynchronize(st_));
    if (hs[0] == ~0ull) return fail(-3, "no '>' in the input");
    const uint64_t nnl = hs[1];
    SWAP_CHECK(nl_pos_.reserve(std::max<uint64_t>(nnl, 1) * 8));
    hipLaunchKernelGGL(k_tile_emit, dim3((unsigned)tiles), dim3(kTileThreads), 0, st_, buf, n,
                       tile_off_.as<uint64_t>(), nl_pos_.as<uint64_t>());
    hipLaunchKernelGGL(k_locate, dim3(1), dim3(1), 0, st_, nl_pos_.as<uint64_t>(), d_small + 1,
                       (const unsigned long long *)d_small, d_small + 2);
    SWAP_CHECK(hipMemcpyAsync(hs + 2, d_small + 2, 2 * 8, hipMemcpyDeviceToHost, st_));
    SWAP_CHECK(hipStreamSynchronize(st_));
    const uint64_t R = hs[3];
    if (R == 0) {
        return 0;  // no complete record: empty graph
    }
    // --- records
    SWAP_CHECK(rec_start_.reserve(R * 8));
    SWAP_CHECK(rec_len_.reserve(R * 4));
    SWAP_CHECK(occ_count_.reserve(R * 4));
    SWAP_CHECK(word_count_.reserve(R * 4));
    SWAP_CHECK(occ_base_.reserve((R + 1) * 8));
    SWAP_CHECK(word_base_.reserve((R + 1) * 8));
    SWAP_CHECK(scan_scratch_.reserve((scan_scratch_words(R) + 16) * 8));
    hipLaunchKernelGGL(k_records, dim3((unsigned)((R + 255) / 256)), dim3(256), 0, st_, nl_pos_.as<uint64_t>(),
                       d_small + 2, R, k_, rec_start_.as<uint64_t>(), rec_len_.as<uint32_t>(),
                       occ_count_.as<uint32_t>(), word_count_.as<uint32_t>(), (unsigned long long *)(d_small + 512));
    SWAP_CHECK(scan_excl<uint32_t>(occ_count_.as<uint32_t>(), occ_base_.as<uint64_t>(), R, scan_scratch_.as<uint64_t>(),
                        d_small + 4, st_));
    SWAP_CHECK(scan_excl<uint32_t>(word_count_.as<uint32_t>(), word_base_.as<uint64_t>(), R, scan_scratch_.as<uint64_t>(),
                        d_small + 5, st_));
    phase_end(0);
    SWAP_CHECK(hipMemcpyAsync(hs + 4, d_small + 4, 2 * 8, hipMemcpyDeviceToHost, st_));
    SWAP_CHECK(hipMemcpyAsync(hs + 512, d_small + 512, 512 * 8, hipMemcpyDeviceToHost, st_));
    SWAP_CHECK(hipStreamSynchronize(st_));
    const uint64_t N = hs[4], W = hs[5];
    n_reads = 0;
    for (int i = 0; i < 32; i++) n_reads += hs[512 + 16 * i];
    n_occ = N;
    if (N >= 0xFFFFFFF0ull) return fail(-4, "more than 2^32 k-mers in one build");
    // batch build: every sample's first record and first input k-mer, its reads
    const uint32_t nsamp = batch_off_ ? (uint32_t)(batch_off_->size() - 1) : 0;
    uint32_t *d_kbase = nullptr;
    if (batch_off_) {
        SWAP_CHECK(bdev_.reserve((uint64_t)(nsamp + 1) * 40 + 64));
        uint64_t *d_off = bdev_.as<uint64_t>();
        uint64_t *d_rfirst = d_off + (nsamp + 1);
        unsigned long long *d_reads = (unsigned long long *)(d_rfirst + (nsamp + 1));
        d_kbase = reinterpret_cast<uint32_t *>(d_reads + (nsamp + 1));
        uint32_t *d_bad = d_kbase + (nsamp + 1);
        SWAP_CHECK(hipMemcpyAsync(d_off, batch_off_->data(), (nsamp + 1) * 8, hipMemcpyHostToDevice, st_));
        SWAP_CHECK(hipMemsetAsync(d_reads, 0, (nsamp + 1) * 8, st_));
        SWAP_CHECK(hipMemsetAsync(d_bad, 0, 4, st_));
        hipLaunchKernelGGL(k_batch_bounds, dim3((nsamp + 256) / 256), dim3(256), 0, st_, buf, d_off, nsamp,
                           nl_pos_.as<uint64_t>(), d_small, rec_start_.as<uint64_t>(), occ_base_.as<uint64_t>(),
                           d_rfirst, d_kbase, d_bad);
        hipLaunchKernelGGL(k_batch_reads, dim3((unsigned)std::min<uint64_t>(1024, (R + 255) / 256)), dim3(256), 0, st_,
                           occ_count_.as<uint32_t>(), R, d_rfirst, nsamp, d_reads);
        std::vector<uint64_t> reads(nsamp);
        std::vector<uint32_t> kb(nsamp + 2);
        SWAP_CHECK(hipMemcpyAsync(reads.data(), d_reads, nsamp * 8, hipMemcpyDeviceToHost, st_));
        SWAP_CHECK(hipMemcpyAsync(kb.data(), d_kbase, (nsamp + 2) * 4, hipMemcpyDeviceToHost, st_));
        SWAP_CHECK(hipStreamSynchronize(st_));
        if (kb[nsamp + 1]) return fail(-1, "a sample does not start with '>', end with a newline and hold line pairs");
        batch_.assign(nsamp, BatchSample());
        for (uint32_t i = 0; i < nsamp; i++) {
            batch_[i].kbase = kb[i];
            batch_[i].kmers = (uint64_t)kb[i + 1] - kb[i];
            batch_[i].reads = reads[i];
        }
    }
    if (N == 0) return 0;
    // --- pack
    SWAP_CHECK(packed_.reserve((W + 1) * 8));
    SWAP_CHECK(wmeta_.reserve((W + 1) * 8));
    {
        const uint64_t nblocks = (W + kPackThreads - 1) / kPackThreads;
        const unsigned grid = (unsigned)std::min<uint64_t>(nblocks, 65535);
        SWAP_CHECK(block_rec_.reserve(nblocks * 8 + 8));
        phase_begin(1);
        hipLaunchKernelGGL(k_block_rec, dim3((unsigned)((R + 255) / 256)), dim3(256), 0, st_, word_base_.as<uint64_t>(),
                           R, W, block_rec_.as<uint64_t>());
        hipLaunchKernelGGL(k_pack, dim3(grid), dim3(kPackThreads), 0, st_, buf, n, rec_start_.as<uint64_t>(),
                           rec_len_.as<uint32_t>(), word_base_.as<uint64_t>(), occ_base_.as<uint64_t>(), R, k_,
                           block_rec_.as<uint64_t>(), W, packed_.as<uint64_t>(), wmeta_.as<uint64_t>());
        phase_end(1);
    }
    // --- partition
    OccCtx oc;
    oc.packed = packed_.as<uint64_t>();
    oc.wmeta = wmeta_.as<uint64_t>();
    oc.W = W;
    oc.k = k_;
    oc.mx = mx_;
    int ncu = 0;
    SWAP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev_));
    const unsigned occ_grid = (unsigned)std::max(8, ncu * (int)tune_.occ_groups);
    const uint64_t n1 = (uint64_t)kL1 * occ_grid;
    // partition bits: about part_recs records per partition (8192 by default),
    // so that its distinct keys fit the 4096-slot LDS table; at most 2k (the key
    // bits)
    int pbits = kPartBitsMin;
    while (pbits < kPartBitsMax && pbits < 2 * k_ && (N >> pbits) > tune_.part_recs) pbits++;
    if (wide_parts_) pbits = std::max(pbits, std::min(kPartBitsMax - 2, 2 * k_));
    pbits_ = pbits;
    const int l2bits = pbits - kL1Bits;
    const uint64_t nparts = 1ull << pbits;
    SWAP_CHECK(hist_.reserve((nparts + 1) * 4));   // final partition sizes, then starts (u32)
    SWAP_CHECK(bucket_.reserve((nparts + 1) * 8));
    SWAP_CHECK(cursor_.reserve((nparts + 1) * 4));
    SWAP_CHECK(hist1_.reserve(n1 * 4));
    SWAP_CHECK(base1_.reserve((n1 + 1) * 8));
    SWAP_CHECK(l1info_.reserve((3 * kL1 + 8) * 4));
    SWAP_CHECK(rows_.reserve((uint64_t)occ_grid * kHistBins * 4));
    SWAP_CHECK(hipMemsetAsync(hist_.p, 0, (kHistBins + 1) * 4, st_));
    phase_begin(2);
    hipLaunchKernelGGL(k_part_hist, dim3(occ_grid), dim3(kOccThreads), 0, st_, oc, rows_.as<uint32_t>(),
                       hist1_.as<uint32_t>());
    hipLaunchKernelGGL(k_col_sum, dim3(kHistBins / 256, kHistRowGroups), dim3(256), 0, st_, rows_.as<uint32_t>(),
                       occ_grid, kHistBins, hist_.as<uint32_t>());
    phase_end(2);
    SWAP_CHECK(scan_scratch_.reserve((scan_scratch_words(std::max<uint64_t>(n1, nparts + 1)) + 16) * 8));
    SWAP_CHECK(scan_excl<uint32_t>(hist_.as<uint32_t>(), bucket_.as<uint64_t>(), kHistBins + 1,
                                   scan_scratch_.as<uint64_t>(), d_small + 6, st_));
    SWAP_CHECK(scan_excl<uint32_t>(hist1_.as<uint32_t>(), base1_.as<uint64_t>(), n1, scan_scratch_.as<uint64_t>(),
                        d_small + 6, st_));
    // u32 starts of the histogram's bins: the final partition starts (and the
    // level-2 cursors) when the partitions are these bins
    hipLaunchKernelGGL(k_u64_to_u32, dim3((kHistBins + 1 + 255) / 256), dim3(256), 0, st_, bucket_.as<uint64_t>(),
                       (uint64_t)kHistBins + 1, hist_.as<uint32_t>(),
                       pbits == kPartBitsMin ? cursor_.as<uint32_t>() : (uint32_t *)nullptr);
    uint32_t *l1_start = l1info_.as<uint32_t>();
    uint32_t *chunk_start = l1_start + kL1 + 2;
    uint32_t *fchunk_start = chunk_start + kL1 + 2;
    const bool wide2 = l2bits >= kL2WideBits;
    const uint32_t chunk2 = wide2 ? kChunkRecsWide : kChunkRecs;
    hipLaunchKernelGGL(k_l1_chunks, dim3(1), dim3(kL1), 0, st_, hist_.as<uint32_t>(), kHistL2Bits, chunk2,
                       l1_start, chunk_start);
    if (pbits > kPartBitsMin)
        hipLaunchKernelGGL(k_l1_chunks, dim3(1), dim3(kL1), 0, st_, hist_.as<uint32_t>(), kHistL2Bits,
                           kFineChunkRecs, l1_start, fchunk_start);
    const uint64_t rec_bytes = N * 4 * (uint64_t)rec_words(ref_order_);
    SWAP_CHECK(rec1_.reserve(rec_bytes));
    SWAP_CHECK(rec_.reserve(rec_bytes));
    phase_begin(3);
    // rounds of 11 records per thread (12-byte records) or 16 (8-byte), the
    // most the 160 KB of LDS stages with one workgroup per CU
    if (ref_order_)
        hipLaunchKernelGGL((k_part1<true, 11>), dim3(occ_grid), dim3(kOccThreads), 0, st_, oc, base1_.as<uint64_t>(),
                           rec1_.as<uint32_t>());
    else
        hipLaunchKernelGGL((k_part1<false, 16>), dim3(occ_grid), dim3(kOccThreads), 0, st_, oc,
                           base1_.as<uint64_t>(), rec1_.as<uint32_t>());
    phase_end(3);
    if (pbits > kPartBitsMin) {
        // finer partitions: their sizes from the level-1 output
        SWAP_CHECK(hipMemsetAsync(hist_.p, 0, (nparts + 1) * 4, st_));
        const unsigned gridf = (unsigned)(N / kFineChunkRecs + kL1 + 1);
        if (ref_order_)
            hipLaunchKernelGGL(k_fine_hist<true>, dim3(gridf), dim3(1024), 0, st_, rec1_.as<uint32_t>(), l1_start,
                               fchunk_start, k_, l2bits, hist_.as<uint32_t>());
        else
            hipLaunchKernelGGL(k_fine_hist<false>, dim3(gridf), dim3(1024), 0, st_, rec1_.as<uint32_t>(), l1_start,
                               fchunk_start, k_, l2bits, hist_.as<uint32_t>());
        SWAP_CHECK(scan_excl<uint32_t>(hist_.as<uint32_t>(), bucket_.as<uint64_t>(), nparts + 1,
                                       scan_scratch_.as<uint64_t>(), d_small + 6, st_));
        hipLaunchKernelGGL(k_u64_to_u32, dim3((unsigned)((nparts + 1 + 255) / 256)), dim3(256), 0, st_,
                           bucket_.as<uint64_t>(), nparts + 1, hist_.as<uint32_t>(), cursor_.as<uint32_t>());
    }
    // (the level-2 cursors start as copies of the partition starts, written
    // by the conversion above)
    const unsigned grid2 = (unsigned)(N / chunk2 + kL1 + 1);
    phase_begin(7);
#define SWAP_L2(OCC, MAX, B, CH, TH)                                                                           \
    hipLaunchKernelGGL((k_part2<OCC, MAX, B, CH, TH>), dim3(grid2), dim3(TH), 0, st_, rec1_.as<uint32_t>(), l1_start,   \
                       chunk_start, k_, l2bits, cursor_.as<uint32_t>(), rec_.as<uint32_t>(), d_kbase, nsamp)
    if (pbits == kPartBitsMin) {
        if (batch_off_) SWAP_L2(true, kHistL2, true, kChunkRecs, kL2Threads);
        else if (ref_order_) SWAP_L2(true, kHistL2, false, kChunkRecs, kL2Threads);
        else SWAP_L2(false, kHistL2, false, kChunkRecs, kL2Threads);
    } else if (!wide2) {
        if (batch_off_) SWAP_L2(true, (1 << kL2MaxBits), true, kChunkRecs, kL2Threads);
        else if (ref_order_) SWAP_L2(true, (1 << kL2MaxBits), false, kChunkRecs, kL2Threads);
        else SWAP_L2(false, (1 << kL2MaxBits), false, kChunkRecs, kL2Threads);
    } else {
        if (batch_off_) SWAP_L2(true, (1 << kL2MaxBits), true, kChunkRecsWide, kL2ThreadsWide);
        else if (ref_order_) SWAP_L2(true, (1 << kL2MaxBits), false, kChunkRecsWide, kL2ThreadsWide);
        else SWAP_L2(false, (1 << kL2MaxBits), false, kChunkRecsWide, kL2ThreadsWide);
    }
#undef SWAP_L2
    phase_end(7);
    // --- count + cutoff
    // survivors <= distinct <= N; reserve N to avoid a sync (grow-only buffers)
    SWAP_CHECK(surv_key_.reserve(N * 8));
    SWAP_CHECK(surv_mult_.reserve(N * 8));
    if (ref_order_) {
        SWAP_CHECK(surv_first_.reserve(N * 4));
        SWAP_CHECK(dfo_.reserve(N * 4));
    }
    SWAP_CHECK(overflow_.reserve(nparts * 4));
    SWAP_CHECK(part_surv_.reserve(nparts * 4));
    SWAP_CHECK(part_first_.reserve(nparts * 4));
    CountArgs ca;
    ca.rec = rec_.as<uint32_t>();
    ca.bucket = hist_.as<uint32_t>();
    ca.nparts = (uint32_t)nparts;
    ca.l2bits = l2bits;
    ca.k = k_;
    ca.c = c_;
    ca.mx = mx_;
    ca.surv_key = surv_key_.as<uint64_t>();
    ca.surv_mult = surv_mult_.as<uint64_t>();
    ca.surv_first = ref_order_ ? surv_first_.as<uint32_t>() : nullptr;
    ca.counters = (unsigned long long *)(d_small + 8);
    ca.overflow_list = overflow_.as<uint32_t>();
    ca.part_surv = part_surv_.as<uint32_t>();
    ca.part_first = part_first_.as<uint32_t>();
    ca.spill = table_spill_ ? 1 : 0;
    ca.dfo = ref_order_ ? dfo_.as<uint32_t>() : nullptr;
    ca.batch = batch_off_ ? 1 : 0;
    ca.pbits = pbits;
    ca.nsamp = nsamp;
    ca.samp_distinct = nullptr;
    ca.samp_last = nullptr;
    if (batch_off_) {
        SWAP_CHECK(bsel_.reserve(batch_sel_bytes(nsamp)));
        ca.samp_distinct = bsel_.as<uint32_t>();
        ca.samp_last = ca.samp_distinct + nsamp;
        SWAP_CHECK(hipMemsetAsync(ca.samp_distinct, 0, (uint64_t)nsamp * 8, st_));
    }
    // small partitions share table fills: 2^gbits of them per fill, while a
    // fill of average partitions stays within fill_recs records (4096 by default)
    int gbits = 0;
    while (!single_fill_ && !batch_off_ && gbits < kMaxFillBits && gbits < l2bits && ((N >> pbits) << (gbits + 1)) <= tune_.fill_recs)
        gbits++;
    ca.gbits = gbits;
    const uint64_t nfills = nparts >> gbits;
    phase_begin(4);
    // one resident workgroup per CU (the LDS table takes most of a CU's LDS),
    // and at least the number whose fill bounds fit the kernel's LDS list
    const unsigned count_grid = (unsigned)std::min<int>(
        (int)nfills, std::max(ncu, (int)((nfills + kMaxFills - 1) / kMaxFills)));
    if (batch_off_)
        hipLaunchKernelGGL((k_count_lds<true, false, true>), dim3(count_grid), dim3(kCountThreads), 0, st_, ca);
    else if (ref_order_ && gbits)
        hipLaunchKernelGGL((k_count_lds<true, true, false>), dim3(count_grid), dim3(kCountThreads), 0, st_, ca);
    else if (ref_order_)
        hipLaunchKernelGGL((k_count_lds<true, false, false>), dim3(count_grid), dim3(kCountThreads), 0, st_, ca);
    else if (gbits)
        hipLaunchKernelGGL((k_count_lds<false, true, false>), dim3(count_grid), dim3(kCountThreads), 0, st_, ca);
    else
        hipLaunchKernelGGL((k_count_lds<false, false, false>), dim3(count_grid), dim3(kCountThreads), 0, st_, ca);
    phase_end(4);
    SWAP_CHECK(hipGetLastError());
    SWAP_CHECK(hipMemcpyAsync(hs + 8, d_small + 8, 3 * 8, hipMemcpyDeviceToHost, st_));
    SWAP_CHECK(hipStreamSynchronize(st_));
    n_overflow = hs[10];
    if (n_overflow) {
        // size a global table per overflowing partition from its record count
        std::vector<uint32_t> list(n_overflow), starts(nparts + 1);
        SWAP_CHECK(hipMemcpy(list.data(), overflow_.p, n_overflow * 4, hipMemcpyDeviceToHost));
        SWAP_CHECK(hipMemcpy(starts.data(), hist_.p, (nparts + 1) * 4, hipMemcpyDeviceToHost));
        std::vector<uint64_t> off(n_overflow);
        std::vector<uint32_t> bits(n_overflow);
        uint64_t tot = 0;
        for (uint64_t j = 0; j < n_overflow; j++) {
            const uint32_t p = list[j];
            const uint64_t cnt = (uint64_t)starts[p + 1] - starts[p];
            uint32_t b = 6;  // at least 2x the records (> 2x the distinct keys)
            while ((1ull << b) < 2 * cnt) b++;
            bits[j] = b;
            off[j] = tot;
            tot += 1ull << b;
        }
        SWAP_CHECK(gtab_key_.reserve(tot * 8));
        SWAP_CHECK(gtab_cnt_.reserve(tot * 8));
        if (ref_order_) SWAP_CHECK(gtab_fo_.reserve(tot * 4));
        DevBuf d_off, d_bits;
        SWAP_CHECK(d_off.reserve(n_overflow * 8));
        SWAP_CHECK(d_bits.reserve(n_overflow * 4));
        SWAP_CHECK(hipMemcpy(d_off.p, off.data(), n_overflow * 8, hipMemcpyHostToDevice));
        SWAP_CHECK(hipMemcpy(d_bits.p, bits.data(), n_overflow * 4, hipMemcpyHostToDevice));
        phase_begin(5);
        if (ref_order_)
            hipLaunchKernelGGL(k_count_global<true>, dim3((unsigned)n_overflow), dim3(kCountThreads), 0, st_, ca,
                               d_off.as<uint64_t>(), d_bits.as<uint32_t>(), gtab_key_.as<uint64_t>(),
                               gtab_cnt_.as<uint64_t>(), gtab_fo_.as<uint32_t>());
        else
            hipLaunchKernelGGL(k_count_global<false>, dim3((unsigned)n_overflow), dim3(kCountThreads), 0, st_, ca,
                               d_off.as<uint64_t>(), d_bits.as<uint32_t>(), gtab_key_.as<uint64_t>(),
                               gtab_cnt_.as<uint64_t>(), (uint32_t *)nullptr);
        phase_end(5);
        SWAP_CHECK(hipGetLastError());
        SWAP_CHECK(hipMemcpyAsync(hs + 8, d_small + 8, 3 * 8, hipMemcpyDeviceToHost, st_));
        SWAP_CHECK(hipStreamSynchronize(st_));
    }
    n_surv = hs[8];
    n_distinct = hs[9];
    if (n_surv) {
        phase_begin(15);
        hipLaunchKernelGGL(k_unmix, dim3((unsigned)((n_surv + 255) / 256)), dim3(256), 0, st_, surv_key_.as<uint64_t>(),
                           n_surv, mx_);
        phase_end(15);
        SWAP_CHECK(hipGetLastError());
    }
    if (ref_order_ && n_distinct && !batch_off_) {
        // growth ranks of the reference's k-mer table: the first occurrence of
        // the key of each growth rank, selected on the device
        std::vector<uint64_t> starts = table_epoch_starts(n_distinct, false);
        if (starts.size() > 32) starts.resize(32);
        int shift = 0;
        while ((N >> shift) >= kSelBins) shift++;
        SWAP_CHECK(sel_.reserve((kSelBins + 64) * 4 + 32 * 8 + sizeof(SelInfo) + sizeof(EpochTabHost) + 256));
        uint32_t *d_hist = sel_.as<uint32_t>();
        uint64_t *d_starts = reinterpret_cast<uint64_t *>(d_hist + kSelBins + 64);
        SelInfo *d_si = reinterpret_cast<SelInfo *>(d_starts + 32);
        // on the side stream: the merge's renumbering and index (main stream)
        // overlap it; the visit-order replay, queued behind it on the side
        // stream, and the k-mer graph export (sel_done_) read its epoch table
        hipStream_t ss = st2_;
        std::memcpy(h_starts_, starts.data(), starts.size() * 8);
        SWAP_CHECK(hipMemcpyAsync(d_starts, h_starts_, starts.size() * 8, hipMemcpyHostToDevice, ss));
        SWAP_CHECK(hipMemsetAsync(d_hist, 0, kSelBins * 4, ss));
        // (d_small[12], the last first occurrence, was zeroed at the build's start)
        SWAP_CHECK(sel_target_.reserve(kSelBins * 4 + 64));
        const uint32_t sel_words = (1u << shift) >= 32 ? (1u << shift) / 32 : 1;  // bitmap words per bin
        SWAP_CHECK(sel_out_.reserve((uint64_t)32 * sel_words * 4 + 16));
        // two histogram workgroups per CU (64 KB of LDS each): twice the waves
        // reading the first occurrences
        const unsigned sel_rows = 2 * (unsigned)ncu;
        SWAP_CHECK(rows_.reserve((uint64_t)sel_rows * kSelBins * 4));
        // the k-mer table's epoch table: bucket counts from D, starts filled below
        epoch1_ = epoch_table(n_distinct, false, nullptr);
        SWAP_CHECK(etab_.reserve(sizeof(EpochTabHost) + 64));
        std::memcpy(h_etab_, &epoch1_, sizeof(EpochTabHost));
        SWAP_CHECK(hipMemcpyAsync(etab_.p, h_etab_, sizeof(EpochTabHost), hipMemcpyHostToDevice, ss));
        EpochTabHost *d_et = etab_.as<EpochTabHost>();
        phase_begin(6, ss);
        hipLaunchKernelGGL(k_sel_hist, dim3(sel_rows), dim3(1024), 0, ss, dfo_.as<uint32_t>(),
                           (const unsigned long long *)(d_small + 9), shift, rows_.as<uint32_t>(),
                           (unsigned long long *)(d_small + 12));
        hipLaunchKernelGGL(k_col_sum, dim3(kSelBins / 256, kHistRowGroups), dim3(256), 0, ss, rows_.as<uint32_t>(),
                           sel_rows, kSelBins, d_hist);
        hipLaunchKernelGGL(k_sel_targets, dim3(1), dim3(1024), 0, ss, d_hist, d_starts, (int)starts.size(),
                           sel_target_.as<int>(), d_si, sel_out_.as<uint32_t>(), 32 * sel_words);
        hipLaunchKernelGGL(k_sel_mark, dim3(ncu * 4), dim3(1024), 0, ss, dfo_.as<uint32_t>(),
                           (const unsigned long long *)(d_small + 9), shift, sel_target_.as<int>(),
                           sel_out_.as<uint32_t>());
        hipLaunchKernelGGL(k_sel_pick, dim3(32), dim3(1024), 0, ss, d_si, sel_out_.as<uint32_t>(), shift,
                           &d_et->start[0]);
        phase_end(6, ss);
        SWAP_CHECK(hipEventRecord(sel_done_, ss));
        epoch_first_.clear();
        epoch_first_valid_ = false;
    }
    collect_phases(false);  // the selection's phase is read after the merge
    return 0;
}

int Pipeline::build_batch(const uint8_t *d_in, const std::vector<uint64_t> &off) {
    if (!ref_order_) return fail(-1, "a batch build replays the reference's table order");
    const size_t ns = off.empty() ? 0 : off.size() - 1;
    if (ns == 0 || ns > kMaxBatch) return fail(-1, "a batch build holds 1 to 512 samples");
    for (size_t i = 0; i < ns; i++)
        if (off[i + 1] <= off[i]) return fail(-2, "an empty sample");
    if (off[0] != 0) return fail(-1, "sample offsets start at 0");
    batch_.clear();
    batch_off_ = &off;
    int rc = build(d_in, off[ns]);
    batch_off_ = nullptr;
    if (rc) {
        batch_.clear();
        return rc;
    }
    // no complete record anywhere: the samples were never checked
    if (batch_.size() != ns) return fail(-1, "a batch without a complete record");
    const uint32_t nsamp = (uint32_t)ns;
    const uint64_t S = n_surv, N = n_occ;
    uint32_t *d_dist = bsel_.as<uint32_t>();
    uint32_t *d_last = d_dist + nsamp;
    uint64_t *d_surv0 = reinterpret_cast<uint64_t *>(d_last + nsamp);
    uint64_t *d_rank = d_surv0 + (nsamp + 1);
    uint32_t *d_tslot = reinterpret_cast<uint32_t *>(d_rank + 32 * ns);
    uint32_t *d_tpos = d_tslot + 32 * ns;
    uint32_t *d_sbin = d_tpos + 32 * ns;
    uint64_t *d_val = reinterpret_cast<uint64_t *>(d_sbin + 32 * ns);
    const uint32_t *d_kbase = reinterpret_cast<const uint32_t *>(bdev_.as<uint64_t>() + 3 * (nsamp + 1));
    if (N == 0) return 0;  // no k-mer: every sample is empty (kmers = 0 from the bounds)
    // survivors in first-occurrence order: each sample's are one run
    if (S) {
        bool renumbered = false;
        rc = renumber_survivors(&renumbered, true);
        if (rc) return rc;
        if (!renumbered) return fail(-4, "too many input k-mers in one batch to order its survivors");
        SWAP_CHECK(hipMemcpyAsync(surv_key_.p, rn_[R_KEY].p, S * 8, hipMemcpyDeviceToDevice, st_));
        surv_renumbered_ = true;
    }
    hipLaunchKernelGGL(k_batch_slices, dim3((nsamp + 256) / 256), dim3(256), 0, st_, surv_first_.as<uint32_t>(), S,
                       d_kbase, nsamp, d_surv0);
    std::vector<uint32_t> dl(2 * ns);
    std::vector<uint64_t> s0(ns + 1);
    SWAP_CHECK(hipMemcpyAsync(dl.data(), d_dist, 2 * ns * 4, hipMemcpyDeviceToHost, st_));
    SWAP_CHECK(hipMemcpyAsync(s0.data(), d_surv0, (ns + 1) * 8, hipMemcpyDeviceToHost, st_));
    SWAP_CHECK(hipStreamSynchronize(st_));
    // growth ranks of every sample's k-mer table as global order statistics
    std::vector<uint64_t> rank;
    std::vector<std::pair<uint32_t, uint32_t>> owner;  // (sample, growth) of each wanted rank
    uint64_t dpre = 0;
    for (uint32_t i = 0; i < nsamp; i++) {
        BatchSample &b = batch_[i];
        b.distinct = dl[i];
        b.last_first = dl[i] ? dl[ns + i] - b.kbase : 0;
        b.surv0 = s0[i];
        b.survivors = s0[i + 1] - s0[i];
        b.epoch_first.clear();
        if (b.distinct) {
            std::vector<uint64_t> st = table_epoch_starts(b.distinct, false);
            if (st.size() > 32) st.resize(32);
            b.epoch_first.assign(st.size(), 0);
            for (size_t j = 0; j < st.size(); j++) {
                rank.push_back(dpre + st[j]);
                owner.emplace_back(i, (uint32_t)j);
            }
        }
        dpre += b.distinct;
    }
    if (dpre != n_distinct) return fail(-5, "per-sample distinct counts do not add up");
    const uint32_t T = (uint32_t)rank.size();
    if (T) {
        int ncu = 0;
        SWAP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev_));
        int shift = 0;
        while ((N >> shift) >= kSelBins) shift++;
        const unsigned sel_rows = 2 * (unsigned)ncu;
        SWAP_CHECK(sel_.reserve((kSelBins + 64) * 4));
        SWAP_CHECK(sel_target_.reserve(kSelBins * 4 + 64));
        SWAP_CHECK(rows_.reserve((uint64_t)sel_rows * kSelBins * 4));
        const uint64_t nw = (1u << shift) >= 32 ? (1u << shift) / 32 : 1;
        SWAP_CHECK(bbits_.reserve((uint64_t)T * nw * 4 + 64));
        uint32_t *d_hist = sel_.as<uint32_t>();
        SWAP_CHECK(hipMemsetAsync(d_hist, 0, kSelBins * 4, st_));
        SWAP_CHECK(hipMemsetAsync(bbits_.p, 0, (uint64_t)T * nw * 4, st_));
        SWAP_CHECK(hipMemcpyAsync(d_rank, rank.data(), (uint64_t)T * 8, hipMemcpyHostToDevice, st_));
        hipLaunchKernelGGL(k_sel_hist, dim3(sel_rows), dim3(1024), 0, st_, dfo_.as<uint32_t>(),
                           (const unsigned long long *)(small_.as<uint64_t>() + 9), shift, rows_.as<uint32_t>(),
                           (unsigned long long *)(small_.as<uint64_t>() + 12));
        hipLaunchKernelGGL(k_col_sum, dim3(kSelBins / 256, kHistRowGroups), dim3(256), 0, st_, rows_.as<uint32_t>(),
                           sel_rows, kSelBins, d_hist);
        hipLaunchKernelGGL(k_bsel_targets, dim3(1), dim3(1024), 0, st_, d_hist, d_rank, T, sel_target_.as<int>(),
                           d_tslot, d_tpos, d_sbin);
        hipLaunchKernelGGL(k_sel_mark, dim3(ncu * 4), dim3(1024), 0, st_, dfo_.as<uint32_t>(),
                           (const unsigned long long *)(small_.as<uint64_t>() + 9), shift, sel_target_.as<int>(),
                           bbits_.as<uint32_t>());
        hipLaunchKernelGGL(k_bsel_pick, dim3(T), dim3(1024), 0, st_, d_tslot, d_tpos, d_sbin, bbits_.as<uint32_t>(),
                           shift, d_val);
        SWAP_CHECK(hipGetLastError());
        std::vector<uint64_t> val(T);
        SWAP_CHECK(hipMemcpyAsync(val.data(), d_val, (uint64_t)T * 8, hipMemcpyDeviceToHost, st_));
        SWAP_CHECK(hipStreamSynchronize(st_));
        for (uint32_t t = 0; t < T; t++) {
            BatchSample &b = batch_[owner[t].first];
            b.epoch_first[owner[t].second] = val[t] - b.kbase;
        }
    }
    return 0;
}

int Pipeline::take_sample(const Pipeline &src, size_t i) {
    SWAP_CHECK(hipSetDevice(dev_));
    SWAP_CHECK(hipStreamSynchronize(st2_));  // a selection or replay of an earlier build may still run
    if (!current_arena()) leave_arena();
    if (i >= src.batch_.size()) return fail(-1, "no such sample in the last batch build");
    if (src.k_ != k_ || src.c_ != c_ || !ref_order_ || !src.ref_order_)
        return fail(-1, "a sample is taken by a context of the batch's k, c and order");
    const BatchSample &b = src.batch_[i];
    n_reads = b.reads;
    n_occ = b.kmers;
    n_distinct = b.distinct;
    n_surv = b.survivors;
    n_kept = n_overflow = n_special = 0;
    surv_renumbered_ = true;  // the survivors come in first-occurrence order
    epoch_first_.clear();
    epoch_first_valid_ = false;
    for (int p = 0; p < kPhases; p++) {
        ev_used_[p] = false;
        phase_ms[p] = 0;
    }
    int pbits = kPartBitsMin;
    while (pbits < kPartBitsMax && pbits < 2 * k_ && (n_occ >> pbits) > tune_.part_recs) pbits++;
    pbits_ = pbits;
    const uint64_t nparts = 1ull << pbits, S = n_surv;
    SWAP_CHECK(small_.reserve(1024 * 8));
    SWAP_CHECK(part_surv_.reserve(nparts * 4));
    SWAP_CHECK(part_first_.reserve(nparts * 4));
    SWAP_CHECK(hipMemsetAsync(part_surv_.p, 0, nparts * 4, st_));
    if (S) {
        SWAP_CHECK(surv_key_.reserve(S * 8));
        SWAP_CHECK(surv_mult_.reserve(S * 8));
        SWAP_CHECK(surv_first_.reserve(S * 4));
        SWAP_CHECK(hipMemcpyAsync(surv_key_.p, src.surv_key_.as<uint64_t>() + b.surv0, S * 8, hipMemcpyDeviceToDevice,
                                  st_));
        SWAP_CHECK(hipMemcpyAsync(surv_mult_.p, src.surv_mult_.as<uint64_t>() + b.surv0, S * 8,
                                  hipMemcpyDeviceToDevice, st_));
        hipLaunchKernelGGL(k_rebase_first, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, st_,
                           src.surv_first_.as<uint32_t>() + b.surv0, surv_first_.as<uint32_t>(), S, (uint32_t)b.kbase);
        hipLaunchKernelGGL(k_part_surv_count, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, st_,
                           surv_key_.as<uint64_t>(), S, mx_, 2 * k_ - pbits, part_surv_.as<uint32_t>());
    }
    if (n_distinct) {
        epoch1_ = epoch_table(n_distinct, false, &b.epoch_first);
        SWAP_CHECK(etab_.reserve(sizeof(EpochTabHost) + 64));
        std::memcpy(h_etab_, &epoch1_, sizeof(EpochTabHost));
        SWAP_CHECK(hipMemcpyAsync(etab_.p, h_etab_, sizeof(EpochTabHost), hipMemcpyHostToDevice, st_));
    }
    // the last first occurrence, where a build leaves it for the exports
    h_small_[41] = b.last_first;
    SWAP_CHECK(hipMemcpyAsync(small_.as<uint64_t>() + 12, h_small_ + 41, 8, hipMemcpyHostToDevice, st_));
    SWAP_CHECK(hipGetLastError());
    SWAP_CHECK(hipStreamSynchronize(st_));
    return 0;
}

int Pipeline::export_kmer_graph(HostKmerGraph &out) {
    SWAP_CHECK(hipSetDevice(dev_));
    out.key.resize(n_surv);
    out.mult.resize(n_surv);
    out.first.clear();
    if (ref_order_ && n_distinct && !epoch_first_valid_) {
        SWAP_CHECK(hipStreamWaitEvent(st_, sel_done_, 0));
        EpochTabHost et;
        SWAP_CHECK(hipMemcpyAsync(&et, etab_.p, sizeof et, hipMemcpyDeviceToHost, st_));
        SWAP_CHECK(hipMemcpyAsync(hs_tail(), d_small_ptr() + 12, 8, hipMemcpyDeviceToHost, st_));
        SWAP_CHECK(hipStreamSynchronize(st_));
        epoch_first_.assign(et.start, et.start + et.n);
        last_first_ = *hs_tail();
        epoch_first_valid_ = true;
    }
    out.epoch_first = epoch_first_;
    out.distinct = n_distinct;
    // the table sees one more operator[] after the last new key when a later
    // input k-mer has an edge; taken as "any k-mer after the last first occurrence"
    out.tail_lookup = ref_order_ && n_occ > last_first_ + 1;
    if (n_surv) {
        SWAP_CHECK(hipMemcpyAsync(out.key.data(), surv_key_.p, n_surv * 8, hipMemcpyDeviceToHost, st_));
        SWAP_CHECK(hipMemcpyAsync(out.mult.data(), surv_mult_.p, n_surv * 8, hipMemcpyDeviceToHost, st_));
        if (ref_order_) {
            out.first.resize(n_surv);
            SWAP_CHECK(hipMemcpyAsync(out.first.data(), surv_first_.p, n_surv * 4, hipMemcpyDeviceToHost, st_));
        }
    }
    SWAP_CHECK(hipStreamSynchronize(st_));
    return 0;
}

}  // namespace swapk
