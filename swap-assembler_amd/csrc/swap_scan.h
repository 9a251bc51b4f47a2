// Device-wide exclusive scan and wave helpers shared by the build and merge.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace swapk {

// Value of lane (i - shift) within each row of 16 lanes (row_shr), or of the
// last lane of the previous row / of lane 31 (row_bcast15 / row_bcast31) for
// the rows in row_mask; 0 where there is no such lane or the row is masked.
// DPP moves stay in the VALU: no LDS round trip per step.
template <int Ctrl, int RowMask>
__device__ __forceinline__ uint32_t dpp_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, Ctrl, RowMask, 0xf, false);
}

template <int Ctrl, int RowMask>
__device__ __forceinline__ uint64_t dpp_u64(uint64_t v) {
    const uint32_t lo = dpp_u32<Ctrl, RowMask>((uint32_t)v), hi = dpp_u32<Ctrl, RowMask>((uint32_t)(v >> 32));
    return (uint64_t)lo | ((uint64_t)hi << 32);
}

template <int Ctrl, int RowMask>
__device__ __forceinline__ uint32_t dpp_t(uint32_t v) { return dpp_u32<Ctrl, RowMask>(v); }
template <int Ctrl, int RowMask>
__device__ __forceinline__ uint64_t dpp_t(uint64_t v) { return dpp_u64<Ctrl, RowMask>(v); }

// Inclusive prefix sum over the 64 lanes of a wave (all lanes active):
// Hillis-Steele inside each row of 16, then the row totals carried across
// rows by two broadcasts.
template <typename T>
__device__ __forceinline__ T wave_incl_sum(T v) {
    v += dpp_t<0x111, 0xf>(v);  // row_shr:1
    v += dpp_t<0x112, 0xf>(v);  // row_shr:2
    v += dpp_t<0x114, 0xf>(v);  // row_shr:4
    v += dpp_t<0x118, 0xf>(v);  // row_shr:8
    v += dpp_t<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_t<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

// ---------------------------------------------------------------------------
// exclusive scan (u64 output) over u32 or u64 input: block reduce, recursive
// scan of the block sums, add back.

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

// total (may be null): where a single-block scan also leaves the sum
template <typename TIn>
static __global__ void k_scan_local(const TIn *__restrict__ in, uint64_t *__restrict__ out, uint64_t n,
                             uint64_t *__restrict__ block_sums, uint64_t *__restrict__ total) {
    __shared__ uint64_t wsum[kScanThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint64_t v[kScanItems];
    uint64_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        v[i] = (base + i < n) ? (uint64_t)in[base + i] : 0;
        s += v[i];
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint64_t inc = wave_incl_sum(s);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint64_t woff = 0;
    for (int i = 0; i < w; i++) woff += wsum[i];
    uint64_t run = inc - s + woff;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
    if (threadIdx.x == kScanThreads - 1) {
        block_sums[blockIdx.x] = run;
        if (total) *total = run;
    }
}

static __global__ void k_scan_add(uint64_t *__restrict__ out, uint64_t n, const uint64_t *__restrict__ offs) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += offs[i / kScanTile];
}

// out (and out2 when given) <- in, narrowed
static __global__ void k_u64_to_u32(const uint64_t *__restrict__ in, uint64_t n, uint32_t *__restrict__ out,
                                    uint32_t *__restrict__ out2) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t v = (uint32_t)in[i];
    out[i] = v;
    if (out2) out2[i] = v;
}

// Arrays up to kScanOneMax elements: one workgroup scans them in chunks of
// kScanOneThreads x kScanItems, carrying the running total (one launch instead
// of three; many samples' arrays are this small, and per-sample launches
// bound the many-sample throughput).
constexpr int kScanOneThreads = 1024;
constexpr uint64_t kScanOneMax = 1u << 14;

template <typename TIn>
static __global__ void __launch_bounds__(kScanOneThreads) k_scan_one(const TIn *__restrict__ in,
                                                                    uint64_t *__restrict__ out, uint64_t n,
                                                                    uint64_t *__restrict__ total) {
    __shared__ uint64_t wsum[kScanOneThreads / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint64_t carry = 0;
    for (uint64_t c0 = 0; c0 < n; c0 += (uint64_t)kScanOneThreads * kScanItems) {
        const uint64_t base = c0 + (uint64_t)threadIdx.x * kScanItems;
        uint64_t v[kScanItems];
        uint64_t s = 0;
#pragma unroll
        for (int i = 0; i < kScanItems; i++) {
            v[i] = (base + i < n) ? (uint64_t)in[base + i] : 0;
            s += v[i];
        }
        const uint64_t inc = wave_incl_sum(s);
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        uint64_t woff = 0, ctot = 0;
        for (int i = 0; i < kScanOneThreads / 64; i++) {
            const uint64_t t = wsum[i];
            if (i < w) woff += t;
            ctot += t;
        }
        uint64_t run = carry + inc - s + woff;
#pragma unroll
        for (int i = 0; i < kScanItems; i++) {
            if (base + i < n) out[base + i] = run;
            run += v[i];
        }
        carry += ctot;
        __syncthreads();  // wsum is rewritten by the next chunk
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

// total <- sum, out <- exclusive prefix.  scratch must hold scan_scratch_words(n).
static inline uint64_t scan_scratch_words(uint64_t n) {
    // each level holds its block sums and their exclusive scan
    uint64_t w = 0;
    while (n > 1) {
        n = (n + kScanTile - 1) / kScanTile;
        w += 2 * n + 2;
    }
    return w + 4;
}

// Returns the first launch or copy error (hipSuccess otherwise).
template <typename TIn>
static hipError_t scan_excl(const TIn *in, uint64_t *out, uint64_t n, uint64_t *scratch, uint64_t *d_total,
                            hipStream_t st) {
    if (n == 0) return hipMemsetAsync(d_total, 0, 8, st);
    const uint64_t nb = (n + kScanTile - 1) / kScanTile;
    uint64_t *sums = scratch;
    if (nb > 1 && n <= kScanOneMax) {
        hipLaunchKernelGGL(k_scan_one<TIn>, dim3(1), dim3(kScanOneThreads), 0, st, in, out, n, d_total);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_scan_local<TIn>, dim3((unsigned)nb), dim3(kScanThreads), 0, st, in, out, n, sums,
                       nb == 1 ? d_total : (uint64_t *)nullptr);
    if (nb == 1) return hipGetLastError();
    uint64_t *sum_scan = scratch + nb;
    const hipError_t e = scan_excl<uint64_t>(sums, sum_scan, nb, scratch + 2 * nb, d_total, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, out, n, sum_scan);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Stream compaction of a predicate over [0, n): list[j] = the j-th index i
// (ascending) with pred(i), and pos_of[i] = j for the selected i (FILL_NONE:
// ~0u for the others).  Blocks of kCompactItems indices: a count pass, an
// exclusive scan of the block counts (compact_count_scan), then a scatter
// pass that re-evaluates the predicate (compact_scatter), so a caller can
// size the list from the total in between.

constexpr int kCompactThreads = 256;
constexpr int kCompactPer = 16;
constexpr int kCompactItems = kCompactThreads * kCompactPer;

static inline uint64_t compact_blocks(uint64_t n) { return (n + kCompactItems - 1) / kCompactItems; }

template <typename Pred>
static __global__ void __launch_bounds__(kCompactThreads) k_compact_count(uint64_t n, Pred pred,
                                                                         uint32_t *__restrict__ blk_cnt) {
    __shared__ uint32_t wtot[kCompactThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kCompactItems + (uint64_t)threadIdx.x * kCompactPer;
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < kCompactPer; j++) c += (base + j < n && pred(base + j)) ? 1u : 0u;
    const uint32_t inc = wave_incl_sum(c);
    if ((threadIdx.x & 63) == 63) wtot[threadIdx.x >> 6] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t sum = 0;
        for (int w = 0; w < kCompactThreads / 64; w++) sum += wtot[w];
        blk_cnt[blockIdx.x] = sum;
    }
}

template <bool FILL_NONE, typename Pred>
static __global__ void __launch_bounds__(kCompactThreads) k_compact_scatter(uint64_t n, Pred pred,
                                                                           const uint64_t *__restrict__ blk_off,
                                                                           uint32_t *__restrict__ pos_of,
                                                                           uint32_t *__restrict__ list) {
    __shared__ uint32_t wtot[kCompactThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kCompactItems + (uint64_t)threadIdx.x * kCompactPer;
    uint32_t f = 0;  // bit j: index base + j selected
#pragma unroll
    for (int j = 0; j < kCompactPer; j++)
        if (base + j < n && pred(base + j)) f |= 1u << j;
    const uint32_t c = __popc(f);
    const uint32_t inc = wave_incl_sum(c);
    if ((threadIdx.x & 63) == 63) wtot[threadIdx.x >> 6] = inc;
    __syncthreads();
    uint64_t off = blk_off[blockIdx.x] + inc - c;
    for (uint32_t w = 0; w < (threadIdx.x >> 6); w++) off += wtot[w];
#pragma unroll
    for (int j = 0; j < kCompactPer; j++) {
        if (base + j >= n) break;
        if ((f >> j) & 1) {
            list[off] = (uint32_t)(base + j);
            pos_of[base + j] = (uint32_t)off;
            off++;
        } else if (FILL_NONE) {
            pos_of[base + j] = ~0u;
        }
    }
}

// Count pass and block-offset scan; the number selected goes to d_total.
// blk_cnt / blk_off hold compact_blocks(n) entries; scratch as for scan_excl.
template <typename Pred>
static hipError_t compact_count_scan(uint64_t n, Pred pred, uint32_t *blk_cnt, uint64_t *blk_off, uint64_t *scratch,
                                     uint64_t *d_total, hipStream_t st) {
    const uint64_t nb = compact_blocks(n);
    if (nb == 0) return hipMemsetAsync(d_total, 0, 8, st);
    hipLaunchKernelGGL(k_compact_count<Pred>, dim3((unsigned)nb), dim3(kCompactThreads), 0, st, n, pred, blk_cnt);
    return scan_excl<uint32_t>(blk_cnt, blk_off, nb, scratch, d_total, st);
}

template <bool FILL_NONE, typename Pred>
static void compact_scatter(uint64_t n, Pred pred, const uint64_t *blk_off, uint32_t *pos_of, uint32_t *list,
                            hipStream_t st) {
    const uint64_t nb = compact_blocks(n);
    if (nb) hipLaunchKernelGGL((k_compact_scatter<FILL_NONE, Pred>), dim3((unsigned)nb), dim3(kCompactThreads), 0, st,
                               n, pred, blk_off, pos_of, list);
}

// Exclusive scan of cnt[0..n) into start[] by the whole workgroup (n <= items
// per thread x blockDim.x; blockDim.x a multiple of 64, at most 1024).  Each
// thread reads its items before any is written, so start may be cnt.
template <int kItems>
__device__ __forceinline__ void block_scan_excl(const uint32_t *cnt, uint32_t *start, uint32_t n) {
    __shared__ uint32_t wtot[16];
    const uint32_t t = threadIdx.x, base = t * kItems;
    uint32_t v[kItems], sum = 0;
#pragma unroll
    for (int j = 0; j < kItems; j++) {
        v[j] = base + j < n ? cnt[base + j] : 0;
        sum += v[j];
    }
    const uint32_t inc = wave_incl_sum(sum);
    if ((t & 63) == 63) wtot[t >> 6] = inc;
    __syncthreads();
    uint32_t off = inc - sum;
    for (uint32_t w = 0; w < (t >> 6); w++) off += wtot[w];
#pragma unroll
    for (int j = 0; j < kItems; j++) {
        if (base + j < n) start[base + j] = off;
        off += v[j];
    }
}

}  // namespace swapk
