// Survivors renumbered in first-occurrence order before the merge.
//
// The count leaves survivors grouped by hash partition, so the k-mers a chain
// of the merge visits one after the other sit anywhere in the survivor arrays.
// Numbered by first occurrence instead, the k-mers of the read that first
// contains them are numbered consecutively: the neighbour a node's edge leads
// to is nearly always the next or the previous survivor, and the walks,
// strings and chain arrays of the merge touch neighbouring lines.  The
// numbering changes no result (every output is ordered by keys and the
// reference's visit order), only the memory traffic of the merge.
//
// A sample sort of the survivors by first occurrence:
//   1. every 16th survivor counted into coarse first-occurrence buckets
//      (global atomics on a few hundred thousand samples);
//   2. a scan of the sampled counts gives every coarse bucket its bin: bins
//      of consecutive coarse buckets, a few thousand survivors each, and
//      never across a block of 2^kRnBlockBits coarse buckets (first
//      occurrences thin out along the input and come in runs of up to a
//      read's length);
//   3. a count and a scatter pass move every survivor, as 24 bytes (key,
//      multiplicities, first occurrence, old number), to its bin;
//   4. one workgroup per bin ranks its survivors by first occurrence with
//      an LDS bitmap and writes them out in that order, and each old
//      number's new one.
// First occurrences are distinct, so the order is total and the numbering
// deterministic.
#include <algorithm>

#include "swap_device.h"
#include "swap_scan.h"

namespace swapk {

constexpr int kRnThreads = 1024;
constexpr uint32_t kRnSample = 16;          // survivors per sample
constexpr uint32_t kRnCoarseMax = 1u << 18;  // coarse first-occurrence buckets
constexpr int kRnCoarseBitsMin = 10;        // coarse bucket width >= 1024 first occurrences
// A bin takes the coarse buckets whose sampled prefix falls in one stretch of
// `keys` survivors: about `keys` plus at most one coarse bucket (as many
// survivors as its width at most) plus the sampling error, which stays below
// the kRnCap survivors a bin is ordered with in LDS (a larger bin is ordered
// in device memory).  A bin's survivors and its output stay in L2 while the
// workgroup reorders them (two workgroups per CU).
constexpr uint32_t kRnCap = 6144;
constexpr uint32_t kRnSlack = 768;  // sampling error allowed for
constexpr int kRnMaxBinBits = 14;   // LDS cursors of the count and scatter passes
constexpr int kRnGroupBits = 11;
// A bin spans at most 2^kRnBlockBits coarse buckets: new k-mers come in
// runs of up to a read's length of consecutive first occurrences, and the
// bitmap of a bin's first occurrences (2^kRnBlockBits coarse widths) stays
// small.
constexpr int kRnBlockBits = 5;
constexpr int kRnCoarseBitsMax = 13;  // the bitmap: at most 2^13 words

constexpr int kRnBatch = 4;  // survivors a thread loads at once in the count and scatter passes

__device__ __forceinline__ uint32_t rn_range(uint32_t wg, uint32_t nwg) {
    return (nwg % 8 == 0) ? (wg % 8) * (nwg / 8) + wg / 8 : wg;  // ranges XCD by XCD
}

__global__ void k_rn_sample(const uint32_t *__restrict__ fo, uint64_t S, int cshift, uint32_t *__restrict__ chist) {
    const uint64_t i = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * kRnSample;
    if (i < S) atomicAdd(&chist[fo[i] >> cshift], 1u);
}

// Coarse bucket c starts a bin (head[c] = 1) where its stretch --
// floor(sampled survivors before it (pre, an exclusive scan of the sampled
// counts) x kRnSample / keys) -- or its block of 2^block_bits coarse buckets
// differs from c - 1's.
__global__ void k_rn_heads(const uint64_t *__restrict__ pre, uint32_t nc, uint32_t keys, int block_bits,
                           uint32_t *__restrict__ head) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    head[c] = c == 0 || (pre[c] * kRnSample) / keys != (pre[c - 1] * kRnSample) / keys ||
              (c >> block_bits) != ((c - 1) >> block_bits);
}

// Coarse numbering: the survivors in their bins' order as the scatter left
// them (first occurrences ordered between bins, not inside one).
__global__ void k_rn_unpack(const ulonglong2 *__restrict__ kv, const uint2 *__restrict__ fi, uint64_t S,
                            uint64_t *__restrict__ key, uint64_t *__restrict__ mult, uint32_t *__restrict__ first,
                            uint32_t *__restrict__ perm) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= S) return;
    const ulonglong2 x = kv[j];
    const uint2 f = fi[j];
    key[j] = x.x;
    mult[j] = x.y;
    first[j] = f.x;
    perm[f.y] = (uint32_t)j;
}

// bin of every coarse bucket: the heads up to it, less one (hpre: exclusive scan of head)
__global__ void k_rn_binmap(const uint64_t *__restrict__ hpre, const uint32_t *__restrict__ head, uint32_t nc,
                            uint16_t *__restrict__ binmap) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nc) binmap[c] = (uint16_t)(hpre[c] + head[c] - 1);
}

struct RnShape {
    int cshift;
    uint32_t nb;
    uint64_t per;  // survivors per count / scatter workgroup
};

__global__ void __launch_bounds__(kRnThreads) k_rn_count(const uint32_t *__restrict__ fo, uint64_t S, RnShape sh,
                                                         const uint16_t *__restrict__ binmap,
                                                         uint32_t *__restrict__ cnt) {
    __shared__ uint32_t hist[1u << kRnMaxBinBits];
    for (uint32_t i = threadIdx.x; i < sh.nb; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const uint32_t w = rn_range(blockIdx.x, gridDim.x);
    const uint64_t b = (uint64_t)w * sh.per, e = min(S, b + sh.per);
    for (uint64_t i0 = b + threadIdx.x; i0 < e; i0 += kRnBatch * blockDim.x) {
        uint32_t f[kRnBatch];
#pragma unroll
        for (int u = 0; u < kRnBatch; u++) {
            const uint64_t i = i0 + (uint64_t)u * blockDim.x;
            if (i < e) f[u] = fo[i];
        }
        uint32_t bn[kRnBatch];
#pragma unroll
        for (int u = 0; u < kRnBatch; u++)
            if (i0 + (uint64_t)u * blockDim.x < e) bn[u] = binmap[f[u] >> sh.cshift];
#pragma unroll
        for (int u = 0; u < kRnBatch; u++)
            if (i0 + (uint64_t)u * blockDim.x < e) atomicAdd(&hist[bn[u]], 1u);
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < sh.nb; i += blockDim.x) cnt[(uint64_t)i * gridDim.x + w] = hist[i];
}

// off[bin * gridDim.x + range]: first slot of the range's survivors of the bin;
// the range of workgroup 0 also writes every bin's first slot and the end (S)
__global__ void __launch_bounds__(kRnThreads) k_rn_scatter(const uint32_t *__restrict__ fo,
                                                           const uint64_t *__restrict__ key,
                                                           const uint64_t *__restrict__ mult, uint64_t S, RnShape sh,
                                                           const uint16_t *__restrict__ binmap,
                                                           const uint64_t *__restrict__ off,
                                                           ulonglong2 *__restrict__ kv, uint2 *__restrict__ fi,
                                                           uint32_t *__restrict__ bin_start) {
    __shared__ uint32_t cur[1u << kRnMaxBinBits];
    const uint32_t w = rn_range(blockIdx.x, gridDim.x);
    for (uint32_t i = threadIdx.x; i < sh.nb; i += blockDim.x) {
        cur[i] = (uint32_t)off[(uint64_t)i * gridDim.x + w];
        if (w == 0) bin_start[i] = cur[i];
    }
    if (w == 0 && threadIdx.x == 0) bin_start[sh.nb] = (uint32_t)S;
    __syncthreads();
    const uint64_t b = (uint64_t)w * sh.per, e = min(S, b + sh.per);
    for (uint64_t i0 = b + threadIdx.x; i0 < e; i0 += kRnBatch * blockDim.x) {
        uint32_t f[kRnBatch], bn[kRnBatch];
        uint64_t k[kRnBatch], m[kRnBatch];
#pragma unroll
        for (int u = 0; u < kRnBatch; u++) {
            const uint64_t i = i0 + (uint64_t)u * blockDim.x;
            if (i < e) {
                f[u] = fo[i];
                k[u] = key[i];
                m[u] = mult[i];
            }
        }
#pragma unroll
        for (int u = 0; u < kRnBatch; u++)
            if (i0 + (uint64_t)u * blockDim.x < e) bn[u] = binmap[f[u] >> sh.cshift];
#pragma unroll
        for (int u = 0; u < kRnBatch; u++) {
            const uint64_t i = i0 + (uint64_t)u * blockDim.x;
            if (i < e) {
                const uint32_t slot = atomicAdd(&cur[bn[u]], 1u);
                kv[slot] = make_ulonglong2(k[u], m[u]);
                fi[slot] = make_uint2(f[u], (uint32_t)i);
            }
        }
    }
}

struct RnOrderArgs {
    const ulonglong2 *kv;       // key, multiplicities of every survivor, grouped by bin
    const uint2 *fi;            // first occurrence, old number
    const uint32_t *bin_start;  // nb + 1
    uint64_t *gsrt;             // S words for the bins above kRnCap survivors
    uint64_t *key;              // out, in first-occurrence order
    uint64_t *mult;
    uint32_t *first;
    uint32_t *perm;             // out: new number of every old one
};

// A bin above kRnCap survivors (device memory): sort words first occurrence
// << 32 | index in the bin, counted into 2^11 groups by first occurrence, the
// groups of two or more listed and each sorted by one thread; then every
// survivor gathered in that order.
__device__ void rn_sort_big(const RnOrderArgs &a, uint32_t lo, uint32_t n, uint64_t *srt, uint32_t *grp,
                            uint16_t *mlist, uint32_t *scal) {
    // scal: [0] smallest first occurrence, [1] largest, [2] listed groups
    for (uint32_t i = threadIdx.x; i < (1u << kRnGroupBits); i += kRnThreads) grp[i] = 0;
    if (threadIdx.x == 0) {
        scal[0] = ~0u;
        scal[1] = 0;
        scal[2] = 0;
    }
    __syncthreads();
    uint32_t mn = ~0u, mx = 0;
    for (uint32_t i = threadIdx.x; i < n; i += kRnThreads) {
        const uint32_t f = a.fi[lo + i].x;
        mn = min(mn, f);
        mx = max(mx, f);
    }
    atomicMin(&scal[0], mn);
    atomicMax(&scal[1], mx);
    __syncthreads();
    const uint32_t fmin = scal[0], range = scal[1] - fmin;
    int gs = 0;
    while ((range >> gs) >= (1u << kRnGroupBits)) gs++;
    const uint32_t ng = (range >> gs) + 1;
    for (uint32_t i = threadIdx.x; i < n; i += kRnThreads) atomicAdd(&grp[(a.fi[lo + i].x - fmin) >> gs], 1u);
    __syncthreads();
    block_scan_excl<(1 << kRnGroupBits) / kRnThreads>(grp, grp, ng);
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += kRnThreads) {
        const uint32_t f = a.fi[lo + i].x;
        srt[atomicAdd(&grp[(f - fmin) >> gs], 1u)] = ((uint64_t)f << 32) | i;
    }
    __syncthreads();  // grp[g]: one past group g
    const uint32_t lane = threadIdx.x & 63;
    for (uint32_t g = threadIdx.x; g < ng; g += kRnThreads) {  // whole waves, or one wave's first lanes
        const bool multi = grp[g] - (g ? grp[g - 1] : 0) >= 2;
        const uint64_t m = __ballot(multi);
        uint32_t at = 0;
        if (lane == 0) at = atomicAdd(&scal[2], (uint32_t)__popcll(m));
        at = __shfl(at, 0);
        if (multi) mlist[at + __popcll(m & ((1ull << lane) - 1))] = (uint16_t)g;
    }
    __syncthreads();
    const uint32_t nm = scal[2];
    for (uint32_t q = threadIdx.x; q < nm; q += kRnThreads) {
        const uint32_t g = mlist[q];
        const uint32_t s = g ? grp[g - 1] : 0, e = grp[g];
        for (uint32_t x = s + 1; x < e; x++) {  // first occurrences are distinct
            const uint64_t v = srt[x];
            uint32_t y = x;
            while (y > s && srt[y - 1] > v) {
                srt[y] = srt[y - 1];
                y--;
            }
            srt[y] = v;
        }
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < n; j += kRnThreads) {
        const uint64_t v = srt[j];
        const uint32_t i = (uint32_t)v;
        const ulonglong2 x = a.kv[lo + i];
        a.key[lo + j] = x.x;
        a.mult[lo + j] = x.y;
        a.first[lo + j] = (uint32_t)(v >> 32);
        a.perm[a.fi[lo + i].y] = lo + j;
    }
}

// One workgroup per bin.  A bin of up to kRnCap survivors lies inside one
// block of 2^kRnBlockBits coarse buckets, so its first occurrences fit a
// bitmap of `words` 32-bit words in LDS: a survivor's rank in the bin is the
// set bits below its own (an exclusive scan of the words' bit counts plus a
// popcount), with no sort.  The survivors, held in registers, are then staged
// through LDS in rank order (keys, multiplicities, first occurrences in turn),
// so that every output but the new numbers is stored whole-line.
__global__ void __launch_bounds__(kRnThreads) k_rn_order(RnOrderArgs a, uint32_t words) {
    extern __shared__ uint32_t dyn[];  // bitmap[words], then its prefix counts[words]
    __shared__ uint64_t stage[kRnCap];
    __shared__ uint32_t grp[1u << kRnGroupBits];
    __shared__ uint16_t mlist[1u << kRnGroupBits];
    __shared__ uint32_t scal[3];
    const uint32_t lo = a.bin_start[blockIdx.x], n = a.bin_start[blockIdx.x + 1] - lo;
    if (n == 0) return;
    if (n > kRnCap) {
        rn_sort_big(a, lo, n, a.gsrt + lo, grp, mlist, scal);
        return;
    }
    uint32_t *bm = dyn, *pre = dyn + words;
    for (uint32_t w = threadIdx.x; w < words; w += kRnThreads) bm[w] = 0;
    if (threadIdx.x == 0) scal[0] = ~0u;
    constexpr int kItems = kRnCap / kRnThreads;
    uint32_t f[kItems], id[kItems], r[kItems];
    uint64_t kk[kItems], mv[kItems];
    uint32_t mn = ~0u;
#pragma unroll
    for (int u = 0; u < kItems; u++) {
        const uint32_t i = u * kRnThreads + threadIdx.x;
        if (i < n) {
            const uint2 x = a.fi[lo + i];
            const ulonglong2 y = a.kv[lo + i];
            f[u] = x.x;
            id[u] = x.y;
            kk[u] = y.x;
            mv[u] = y.y;
            mn = min(mn, x.x);
        }
    }
    for (int o = 32; o > 0; o >>= 1) mn = min(mn, (uint32_t)__shfl_xor((int)mn, o));
    __syncthreads();  // the bitmap cleared, scal[0] set
    if ((threadIdx.x & 63) == 0) atomicMin(&scal[0], mn);
    __syncthreads();
    const uint32_t fmin = scal[0];
#pragma unroll
    for (int u = 0; u < kItems; u++) {
        const uint32_t i = u * kRnThreads + threadIdx.x;
        if (i < n) {
            const uint32_t d = f[u] - fmin;
            atomicOr(&bm[d >> 5], 1u << (d & 31));
        }
    }
    __syncthreads();
    for (uint32_t w = threadIdx.x; w < words; w += kRnThreads) pre[w] = __popc(bm[w]);
    __syncthreads();
    block_scan_excl<(1 << (kRnCoarseBitsMax + kRnBlockBits - 5)) / kRnThreads>(pre, pre, words);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kItems; u++) {
        const uint32_t i = u * kRnThreads + threadIdx.x;
        if (i < n) {
            const uint32_t d = f[u] - fmin;
            r[u] = pre[d >> 5] + __popc(bm[d >> 5] & ((1u << (d & 31)) - 1u));
            stage[r[u]] = kk[u];
            a.perm[id[u]] = lo + r[u];
        }
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < n; j += kRnThreads) a.key[lo + j] = stage[j];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kItems; u++)
        if (u * kRnThreads + threadIdx.x < n) stage[r[u]] = mv[u];
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < n; j += kRnThreads) a.mult[lo + j] = stage[j];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kItems; u++)
        if (u * kRnThreads + threadIdx.x < n) stage[r[u]] = f[u];
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < n; j += kRnThreads) a.first[lo + j] = (uint32_t)stage[j];
}

#define RCHECK(x)                                                            \
    do {                                                                     \
        hipError_t e_ = (x);                                                 \
        if (e_ != hipSuccess) return fail(-5, hipGetErrorString(e_));        \
    } while (0)

// New order of the survivors: keys into rn_[R_KEY] (the survivor index is
// built from the old order next), multiplicities and first occurrences in
// place (their old order is dead once the survivors are scattered into bins),
// and the new number of every old one into rn_[R_PERM].  exact_only: the exact
// first-occurrence order or nothing (*renumbered false, nothing launched,
// above 2^kRnMaxBinBits bins or 2^(kRnCoarseBitsMax + 18) input k-mers).
int Pipeline::renumber_survivors(bool *renumbered, bool exact_only) {
    *renumbered = false;
    const uint64_t S = n_surv;
    if (S == 0 || n_occ == 0 || !ref_order_) return 0;
    int cshift = kRnCoarseBitsMin;
    while (((n_occ - 1) >> cshift) >= kRnCoarseMax) cshift++;
    const uint32_t nc = (uint32_t)((n_occ - 1) >> cshift) + 1;
    const uint64_t nsamp = (S + kRnSample - 1) / kRnSample;
    RnShape sh;
    sh.cshift = cshift;
    // survivors a bin aims at: a bin of them plus one coarse bucket plus the
    // sampling error fits the LDS capacity
    const uint32_t width = 1u << cshift;
    uint32_t keys = kRnCap > width + kRnSlack + 1024 ? kRnCap - width - kRnSlack : 1024;
    int block_bits = kRnBlockBits;
    // bins: stretches, each cut at most once more by every block boundary
    uint64_t nb = nsamp * kRnSample / keys + 1 + ((nc - 1) >> kRnBlockBits);
    // Beyond the exact ordering's range (its bitmaps and bin cursors), or
    // with the testing aid, the survivors are numbered by bins only: bins of
    // about S / 2^14 survivors with no block cuts, first occurrences ordered
    // between bins and left in scatter order inside one.  A read's new k-mers
    // still share a bin, which is what the merge's locality needs; no result
    // depends on the numbering.  A batch build needs the exact order.
    const bool coarse = (coarse_renumber_ && !exact_only) || cshift > kRnCoarseBitsMax || nb > (1u << kRnMaxBinBits);
    if (coarse) {
        if (exact_only) return 0;
        block_bits = 31;
        keys = (uint32_t)std::max<uint64_t>(1024, nsamp * kRnSample / ((1u << kRnMaxBinBits) - 2) + 1);
        nb = nsamp * kRnSample / keys + 1;
    }
    if (nb > (1u << kRnMaxBinBits)) return 0;
    sh.nb = (uint32_t)nb;
    int ncu = 0;
    RCHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev_));
    const uint64_t nblk = std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)ncu, (S + 8191) / 8192));
    sh.per = (S + nblk - 1) / nblk;
    const uint64_t ncnt = nb * nblk;
    RCHECK(rn_[R_CHIST].reserve((uint64_t)nc * 4));
    RCHECK(rn_[R_CPRE].reserve((uint64_t)nc * 8 + 8));
    RCHECK(rn_[R_HEAD].reserve((uint64_t)nc * 4));
    RCHECK(rn_[R_HPRE].reserve((uint64_t)nc * 8 + 8));
    RCHECK(rn_[R_BINMAP].reserve((uint64_t)nc * 2));
    RCHECK(rn_[R_CNT].reserve(ncnt * 4 + 4));
    RCHECK(rn_[R_OFF].reserve(ncnt * 8 + 8));
    RCHECK(rn_[R_BSTART].reserve((nb + 1) * 4));
    RCHECK(rn_[R_KV].reserve(S * 16));
    RCHECK(rn_[R_FI].reserve(S * 8));
    RCHECK(rn_[R_GSRT].reserve(S * 8));
    RCHECK(rn_[R_KEY].reserve(S * 8));
    RCHECK(rn_[R_PERM].reserve(S * 4));
    RCHECK(scan_scratch_.reserve((scan_scratch_words(std::max<uint64_t>(ncnt, nc)) + 16) * 8));
    const uint32_t *fo = surv_first_.as<uint32_t>();
    uint32_t *chist = rn_[R_CHIST].as<uint32_t>();
    uint16_t *binmap = rn_[R_BINMAP].as<uint16_t>();
    RCHECK(hipMemsetAsync(chist, 0, (uint64_t)nc * 4, st_));
    hipLaunchKernelGGL(k_rn_sample, dim3((unsigned)((nsamp + 255) / 256)), dim3(256), 0, st_, fo, S, cshift, chist);
    uint64_t *cpre = rn_[R_CPRE].as<uint64_t>();
    RCHECK(scan_excl<uint32_t>(chist, cpre, nc, scan_scratch_.as<uint64_t>(), cpre + nc, st_));
    uint32_t *head = rn_[R_HEAD].as<uint32_t>();
    uint64_t *hpre = rn_[R_HPRE].as<uint64_t>();
    hipLaunchKernelGGL(k_rn_heads, dim3((nc + 255) / 256), dim3(256), 0, st_, cpre, nc, keys, block_bits, head);
    RCHECK(scan_excl<uint32_t>(head, hpre, nc, scan_scratch_.as<uint64_t>(), hpre + nc, st_));
    hipLaunchKernelGGL(k_rn_binmap, dim3((nc + 255) / 256), dim3(256), 0, st_, hpre, head, nc, binmap);
    hipLaunchKernelGGL(k_rn_count, dim3((unsigned)nblk), dim3(kRnThreads), 0, st_, fo, S, sh, binmap,
                       rn_[R_CNT].as<uint32_t>());
    uint64_t *off = rn_[R_OFF].as<uint64_t>();
    RCHECK(scan_excl<uint32_t>(rn_[R_CNT].as<uint32_t>(), off, ncnt, scan_scratch_.as<uint64_t>(), off + ncnt, st_));
    hipLaunchKernelGGL(k_rn_scatter, dim3((unsigned)nblk), dim3(kRnThreads), 0, st_, fo, surv_key_.as<uint64_t>(),
                       surv_mult_.as<uint64_t>(), S, sh, binmap, off, rn_[R_KV].as<ulonglong2>(),
                       rn_[R_FI].as<uint2>(), rn_[R_BSTART].as<uint32_t>());
    if (coarse) {
        hipLaunchKernelGGL(k_rn_unpack, dim3((unsigned)((S + 255) / 256)), dim3(256), 0, st_, rn_[R_KV].as<ulonglong2>(),
                           rn_[R_FI].as<uint2>(), S, rn_[R_KEY].as<uint64_t>(), surv_mult_.as<uint64_t>(),
                           surv_first_.as<uint32_t>(), rn_[R_PERM].as<uint32_t>());
        *renumbered = true;
        return 0;
    }
    RnOrderArgs a;
    a.kv = rn_[R_KV].as<ulonglong2>();
    a.fi = rn_[R_FI].as<uint2>();
    a.bin_start = rn_[R_BSTART].as<uint32_t>();
    a.gsrt = rn_[R_GSRT].as<uint64_t>();
    a.key = rn_[R_KEY].as<uint64_t>();
    a.mult = surv_mult_.as<uint64_t>();  // read only by the scatter above
    a.first = surv_first_.as<uint32_t>();
    a.perm = rn_[R_PERM].as<uint32_t>();
    const uint32_t words = 1u << (cshift + kRnBlockBits - 5);
    hipLaunchKernelGGL(k_rn_order, dim3((unsigned)nb), dim3(kRnThreads), 2 * words * 4, st_, a, words);
    *renumbered = true;
    return 0;
}

}  // namespace swapk
