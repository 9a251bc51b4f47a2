// Shared device/host helpers for the k-mer graph path (gfx950).
//
// K-mer encoding follows kmerGraph.cpp:16-24 (A=0 C=1 G=2 T=3, first base in the
// highest bits) and the canonical form is max(x, rc x) (kmerGraph.cpp:246-247).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#define SWAP_HD __host__ __device__ __forceinline__

namespace swapk {

// Device-local hash partitioning: 2^pbits partitions keyed by the top bits of
// a bijective mix of the canonical k-mer, pbits chosen per build from the input
// size within [kPartBitsMin, kPartBitsMax] so that a partition's distinct keys
// fit the on-chip counting table.
constexpr int kPartBitsMin = 15;
constexpr int kPartBitsMax = 20;

// Smallest k the merge accepts: the reference's command line allows odd k in
// 19..31 (sequence.cpp:44).  Shorter k-mer graphs are built (-s) but not merged.
constexpr int kMergeMinK = 19;
// Samples of one batch build (swap_build_batch): the sample of a record rides
// in 9 key bits of the records between the partition passes and the count.
constexpr size_t kMaxBatchSamples = 512;

// The static device scratch of swap_assemble_static (one per GPU, held until
// the process ends) and the largest FASTA input that entry point accepts: the
// buffers of a build and merge come to about 130 device bytes per input byte
// on the densest inputs (k = 19, 250-base reads at 2% errors, c = 1: most
// k-mers survive and are merged), plus ~40 MB that do not scale.
constexpr size_t kStaticScratchBytes = 512ull << 20;
constexpr size_t kStaticMaxInput = 3ull << 20;

SWAP_HD uint64_t kmask(int k) { return (k >= 32) ? ~0ull : ((1ull << (2 * k)) - 1); }

// Reverse complement of a 2k-bit k-mer in O(1): complement every 2-bit group,
// reverse the group order, shift down to the low 2k bits.
SWAP_HD uint64_t revcomp(uint64_t x, int k) {
    x = ~x;
    x = ((x >> 2) & 0x3333333333333333ull) | ((x & 0x3333333333333333ull) << 2);
    x = ((x >> 4) & 0x0F0F0F0F0F0F0F0Full) | ((x & 0x0F0F0F0F0F0F0F0Full) << 4);
    x = ((x >> 8) & 0x00FF00FF00FF00FFull) | ((x & 0x00FF00FF00FF00FFull) << 8);
    x = ((x >> 16) & 0x0000FFFF0000FFFFull) | ((x & 0x0000FFFF0000FFFFull) << 16);
    x = (x >> 32) | (x << 32);
    return x >> (64 - 2 * k);
}

// Bijective mixer on the 2k-bit k-mer space used for partitioning: odd
// multiply mod 2^2k by 2^2k / golden ratio (Fibonacci hashing: the top bits,
// which pick the partition, spread even runs of consecutive keys), then an
// xor-shift by half the width so the low bits depend on the high ones.  The inverse is
// applied to the bits kept in the partition records to recover the k-mer.
struct Mixer {
    int bits;     // 2k
    uint64_t m;   // 2k-bit mask
    uint64_t a;   // odd multiplier
    uint64_t ia;  // its inverse mod 2^64 (truncated to 2k bits on use)
    int s;        // xor-shift amount, >= bits / 2
};

SWAP_HD uint64_t mix_fwd(const Mixer &mx, uint64_t x) {
    x = (x * mx.a) & mx.m;
    return x ^ (x >> mx.s);
}

SWAP_HD uint64_t mix_inv(const Mixer &mx, uint64_t y) {
    // y = x ^ (x >> s) with 2s >= bits: x = y ^ (y >> s)
    const uint64_t x = (y ^ (y >> mx.s)) & mx.m;
    return (x * mx.ia) & mx.m;
}

// Saturating add of per-byte increments (each byte of inc is 0 or 1).
SWAP_HD uint64_t sat_add_bytes(uint64_t v, uint64_t inc) {
    // bytes already at 255 drop their increment
    const uint64_t full = v & (v >> 1) & (v >> 2) & (v >> 3) & (v >> 4) & (v >> 5) & (v >> 6) & (v >> 7) & 0x0101010101010101ull;
    return v + (inc & ~full);
}

// Expand an 8-bit edge mask to one byte per bit.
SWAP_HD uint64_t mask_to_bytes(uint32_t m) {
    uint64_t r = 0;
    for (int t = 0; t < 8; t++) r |= (uint64_t)((m >> t) & 1u) << (8 * t);
    return r;
}

}  // namespace swapk
