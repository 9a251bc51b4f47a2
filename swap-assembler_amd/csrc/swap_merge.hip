// k-molecule merge on gfx950: unitig compaction by list ranking, edge strings
// by scatter, merged multiplicities by Cartesian-tree evaluation.
//
// Replaces (single rank):
//   distNodeGraph::buildKmoleculeGraph    graph.cpp:1117-1137
//   distNodeGraph::simplifyDistNodeGraph  graph.cpp:1392-1566 (SWAP)
//   distNodeGraph::masterSimplifyNodeGraph graph.cpp:2089-2192
//
// The reference absorbs, in its table's visit order, every node with exactly
// one left and one right edge (graph.cpp:1429-1438): the right neighbour's
// back edge is extended by the node's left edge and vice versa, and each
// extended edge takes m = floor((m_a*len_a + m_b*len_b) / (len_a + len_b))
// (graph.cpp:1339-1348).  Absorbing along a maximal chain of such nodes between
// two kept nodes therefore yields one edge string per end -- the unitig read
// from that end -- and a multiplicity equal to evaluating the chain's
// Cartesian tree keyed by visit time (the last-visited node is the root,
// combining the merged left and right parts).  Here:
//   * nodes are looked up through a device hash index of the survivors;
//   * chains are ranked over oriented nodes (2 per node) with a ruling set:
//     sampled rulers walk to the next ruler, then pointer jumping over rulers;
//   * each oriented node writes its base into the string of the end it is
//     reached from; kept-to-kept edges keep their one base;
//   * nearest-larger-priority neighbours come from a max hierarchy, and the
//     tree is evaluated bottom-up with one thread climbing from each leaf, the
//     last-arriving child finishing its parent.
// Chains that close on themselves (cycles) or fold back onto a node's other
// strand (palindromic paths) depend on the exact sequence of lock outcomes in
// the reference; they are rare and are replayed on the host.
#include "swap_host.h"
#include "swap_scan.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <numeric>
#include <unordered_map>

namespace swapk {

constexpr uint32_t kNil = 0xFFFFFFFFu;
constexpr uint64_t kInf = ~0ull;
constexpr uint64_t kIdxEmpty = ~0ull;

// ---------------------------------------------------------------------------
// survivor index

// The survivors of count partition p (p = top pbits bits of the mixed key,
// the partition the count kernel took them from) own one open-addressing
// region of entries [base[p], base[p+1]); its size is the power of two that
// holds them at no more than the tuned load (67% by default; 0 for a partition
// without survivors).  An entry is 16 bytes (key, survivor id), read by one
// load.

struct SurvIndex {
    const uint64_t *ent;   // 2 words per entry: key (kIdxEmpty if free), survivor id
    const uint64_t *base;  // nparts + 1 region starts
    Mixer mx;
    int shift;             // 2k - pbits
};

// Region of ns survivors: the smallest power of two holding ns at the tuned
// load (at least ns + 1 entries: a probe always meets a free entry).
// mul256 = 256 / load.
__device__ __forceinline__ uint32_t region_size(uint32_t ns, uint32_t mul256) {
    if (!ns) return 0;
    const uint32_t want = (uint32_t)(((uint64_t)ns * mul256) >> 8) + 1;
    return 1u << (32 - __clz(want - 1));
}

__global__ void k_index_sizes(const uint32_t *__restrict__ part_surv, uint32_t nparts, uint32_t mul256,
                              uint32_t *__restrict__ size) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < nparts) size[p] = region_size(part_surv[p], mul256);
}

// Regions built in LDS, one workgroup per partition whose survivors form one
// run of the survivor array and whose region fits kIdxLds entries: the region
// is written whole (free entries included) with coalesced stores.  Other
// regions are only cleared here and filled by k_index_loose.
//
// key: the survivors in the count's order.  Entries hold their numbers after
// renumbering, perm[i] for survivor i (i itself when perm is null).  With runs
// false (the survivor arrays were renumbered by an earlier merge of the build,
// and key is in that order) every region is filled by k_index_loose.
constexpr uint32_t kIdxLds = 1024;
constexpr uint32_t kIdxThreads = 128;

__global__ void __launch_bounds__(kIdxThreads) k_index_part(const uint64_t *__restrict__ key,
                                                    const uint32_t *__restrict__ part_surv,
                                                    const uint32_t *__restrict__ part_first, SurvIndex ix,
                                                    uint64_t *__restrict__ ent, const uint32_t *__restrict__ perm,
                                                    bool runs) {
    __shared__ unsigned long long lkey[kIdxLds];
    __shared__ uint32_t lid[kIdxLds];
    const uint32_t p = blockIdx.x;
    const uint64_t b = ix.base[p];
    const uint32_t size = (uint32_t)(ix.base[p + 1] - b);
    if (!size) return;
    const uint32_t first = runs ? part_first[p] : ~0u;
    if (first == ~0u || size > kIdxLds) {
        for (uint32_t s = threadIdx.x; s < 2 * size; s += blockDim.x) ent[2 * b + s] = kIdxEmpty;
        return;
    }
    for (uint32_t s = threadIdx.x; s < size; s += blockDim.x) lkey[s] = kIdxEmpty;
    __syncthreads();
    const uint32_t ns = part_surv[p], mask = size - 1;
    for (uint32_t i = threadIdx.x; i < ns; i += blockDim.x) {
        const uint64_t k = key[first + i];
        uint32_t s = (uint32_t)mix_fwd(ix.mx, k) & mask;
        for (uint32_t probe = 0; probe <= mask; probe++) {  // the region is at most 2/3 full
            const unsigned long long prev = atomicCAS(&lkey[s], (unsigned long long)kIdxEmpty, (unsigned long long)k);
            if (prev == kIdxEmpty) {
                lid[s] = perm ? perm[first + i] : first + i;
                break;
            }
            s = (s + 1) & mask;
        }
    }
    __syncthreads();
    for (uint32_t s = threadIdx.x; s < size; s += blockDim.x) {
        const uint64_t k = lkey[s];
        reinterpret_cast<ulonglong2 *>(ent)[b + s] = make_ulonglong2(k, k == kIdxEmpty ? kIdxEmpty : (uint64_t)lid[s]);
    }
}

// Survivors of the regions k_index_part left empty: device-memory inserts.
__global__ void k_index_loose(const uint64_t *__restrict__ key, uint64_t n, const uint32_t *__restrict__ part_first,
                              SurvIndex ix, uint64_t *ent, const uint32_t *__restrict__ perm, bool runs) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t k = key[i];
    const uint64_t h = mix_fwd(ix.mx, k);
    const uint64_t p = h >> ix.shift;
    const uint64_t b = ix.base[p];
    const uint64_t size = ix.base[p + 1] - b;  // the region holds key's partition: size >= 2
    if (runs && part_first[p] != ~0u && size <= kIdxLds) return;
    const uint64_t mask = size - 1;
    uint64_t s = h & mask;
    for (uint64_t probe = 0; probe <= mask; probe++) {  // the region is at most 2/3 full
        const uint64_t prev = atomicCAS((unsigned long long *)&ent[2 * (b + s)], (unsigned long long)kIdxEmpty,
                                        (unsigned long long)k);
        if (prev == kIdxEmpty) {
            ent[2 * (b + s) + 1] = perm ? perm[i] : i;
            return;
        }
        s = (s + 1) & mask;
    }
}

__device__ __forceinline__ uint32_t index_find(const SurvIndex &ix, uint64_t k) {
    const uint64_t h = mix_fwd(ix.mx, k);
    const uint64_t p = h >> ix.shift;
    const uint64_t b = ix.base[p];
    const uint64_t size = ix.base[p + 1] - b;
    if (!size) return kNil;
    const uint64_t mask = size - 1;
    uint64_t s = h & mask;
    for (uint64_t probe = 0; probe <= mask; probe++) {
        const ulonglong2 e = reinterpret_cast<const ulonglong2 *>(ix.ent)[b + s];
        if (e.x == k) return (uint32_t)e.y;
        if (e.x == kIdxEmpty) return kNil;
        s = (s + 1) & mask;
    }
    return kNil;
}

// ---------------------------------------------------------------------------
// classification

struct MergeArgs {
    const uint64_t *key;
    const uint64_t *mult;
    uint64_t S;
    int k;
    SurvIndex ix;
    bool near;  // survivors numbered in first-occurrence order
};

__device__ __forceinline__ uint64_t orient_kmer(uint64_t X, int slot, int k) { return slot >= 4 ? revcomp(X, k) : X; }

// one step along slot s of X: returns target oriented id (2j+o), or kNil if missing
__device__ __forceinline__ uint32_t hop(const MergeArgs &a, uint64_t X, int s) {
    const uint64_t km = kmask(a.k);
    const uint64_t t = ((orient_kmer(X, s, a.k) << 2) | (uint64_t)(s & 3)) & km;
    const uint64_t r = revcomp(t, a.k);
    const uint64_t Y = t > r ? t : r;
    const uint32_t j = index_find(a.ix, Y);
    if (j == kNil) return kNil;
    return (j << 1) | (t == Y ? 0u : 1u);
}

// hop() from survivor i: with survivors numbered in first-occurrence order,
// the k-mer that follows a survivor in the read that first contains it is
// nearly always the next or the previous survivor, so those two keys are
// compared before the index is probed.
__device__ __forceinline__ uint32_t hop_near(const MergeArgs &a, uint64_t X, int s, uint64_t i) {
    const uint64_t km = kmask(a.k);
    const uint64_t t = ((orient_kmer(X, s, a.k) << 2) | (uint64_t)(s & 3)) & km;
    const uint64_t r = revcomp(t, a.k);
    const uint64_t Y = t > r ? t : r;
    uint32_t j = kNil;
    if (a.near) {
        if (i + 1 < a.S && a.key[i + 1] == Y)
            j = (uint32_t)(i + 1);
        else if (i > 0 && a.key[i - 1] == Y)
            j = (uint32_t)(i - 1);
    }
    if (j == kNil) j = index_find(a.ix, Y);
    if (j == kNil) return kNil;
    return (j << 1) | (t == Y ? 0u : 1u);
}

__device__ __forceinline__ int nz_count(uint64_t m, int lo) {
    int c = 0;
#pragma unroll
    for (int t = 0; t < 4; t++) c += ((m >> (8 * (lo + t))) & 0xFF) != 0;
    return c;
}

__device__ __forceinline__ int nz_slot(uint64_t m, int lo) {
    int s = -1;
#pragma unroll
    for (int t = 0; t < 4; t++)
        if ((m >> (8 * (lo + t))) & 0xFF) s = lo + t;
    return s;
}

// is11[i]; for (1,1) nodes: nxt[2i] (through the right slot), nxt[2i+1] (left
// slot), exit multiplicities ex[2i], ex[2i+1], self-adjacency flag.
__global__ void k_classify(MergeArgs a, uint8_t *__restrict__ is11, uint32_t *__restrict__ nxt,
                           uint8_t *__restrict__ ex, uint8_t *__restrict__ selfadj,
                           unsigned long long *__restrict__ counters /*[0]=n11 [1]=missing*/) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.S) return;
    const uint64_t m = a.mult[i];
    const bool one = nz_count(m, 0) == 1 && nz_count(m, 4) == 1;
    is11[i] = one;
    if (!one) return;
    const uint64_t X = a.key[i];
    const int r = nz_slot(m, 0), l = nz_slot(m, 4);
    const uint32_t nr = hop_near(a, X, r, i), nl = hop_near(a, X, l, i);
    nxt[2 * i] = nr;
    nxt[2 * i + 1] = nl;
    ex[2 * i] = (uint8_t)(m >> (8 * r));
    ex[2 * i + 1] = (uint8_t)(m >> (8 * l));
    selfadj[i] = ((nr >> 1) == (uint32_t)i) || ((nl >> 1) == (uint32_t)i);
    if (nr == kNil || nl == kNil) atomicAdd(&counters[1], 1ull);
}

// Pointer-jumping state per oriented (1,1) node v = 2i+o.
struct Walk {
    uint32_t succ;   // next oriented (1,1) node still to jump over, or kNil
    uint32_t dist;   // nodes from v to the terminal (v included)
    uint32_t term;   // arrival at the kept node: 2*K + orientation
    uint32_t last;   // last oriented (1,1) node before the terminal
};

// Ruling-set list ranking.  Rulers: the first oriented node of every walk
// that starts at a kept node, plus a hashed 1/stride sample of the rest (so
// cycles and long chains are cut into short pieces; the stride is tuned, 32 by
// default).  Each ruler walks to the next ruler or to a kept node, labelling
// the nodes it passes with (ruler, offset); pointer jumping then runs over the
// rulers only.

// props[v]: bit 0 (1,1) node, bit 1 ruler, bit 2 self-adjacent, so that
// k_rs_prep reads one byte of each successor (and the ruler numbering one byte
// per oriented node).
__global__ void k_rs_mark(const uint8_t *__restrict__ is11, const uint32_t *__restrict__ nxt,
                          const uint8_t *__restrict__ selfadj, uint64_t S, uint8_t *__restrict__ props,
                          uint32_t stride, uint2 *__restrict__ label) {
    const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= 2 * S) return;
    uint32_t f = 0;
    const bool one = is11[v >> 1];
    if (one) {
        const uint32_t back = nxt[v ^ 1];
        f = (back == kNil || !is11[back >> 1]) ? 1u : ((uint32_t)((v * 0x9E3779B97F4A7C15ull) >> 32) % stride == 0);
    }
    props[v] = (uint8_t)((one ? 1u : 0u) | (f << 1) | (one && selfadj[v >> 1] ? 4u : 0u));
    label[v] = make_uint2(kNil, kNil);  // not reached by a ruler's walk (yet)
}

// Rulers (numbered in oriented-node order) and kept nodes (in survivor order):
// the predicates of the two compactions.
struct RulerPred {
    const uint8_t *props;
    __device__ bool operator()(uint64_t v) const { return (props[v] & 2) != 0; }
};
struct KeptPred {
    const uint8_t *is11;
    __device__ bool operator()(uint64_t i) const { return !is11[i]; }
};

// Per oriented (1,1) node: its successor with what the walk does there, so a
// walk step is one 8-byte load:  bits 0-31 successor, bits 32-33 kind
// (0 pass through, 1 next ruler, 2 kept node or none), bit 34 the
// successor's self-adjacency flag.
__global__ void k_rs_prep(const uint8_t *__restrict__ props, const uint32_t *__restrict__ nxt, uint64_t S,
                          uint64_t *__restrict__ nx) {
    const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= 2 * S || !(props[v] & 1)) return;
    const uint32_t t = nxt[v];
    const uint32_t pt = t == kNil ? 0u : props[t];
    uint64_t kind, sa = 0;
    if (!(pt & 1)) {
        kind = 2;
    } else {
        kind = (pt >> 1) & 1;
        sa = (pt >> 2) & 1;
    }
    nx[v] = (uint64_t)t | (kind << 32) | (sa << 34);
}

// Runs: a run is a maximal sequence of oriented (1,1) nodes of one tile of
// kRunTile consecutive ids, each linked to the next by a pass-through step
// (the successor lies in the tile and is not a ruler).  Survivors numbered in
// first-occurrence order put chain neighbours next to each other, so runs are
// long and the rulers' walks take one run per step.  Per node, by pointer
// doubling in LDS: the run's last node (its exit), the steps from the node to
// it, and the OR of the self-adjacency flags of the nodes after it up to the
// exit.  seg = exit | steps << 32 | flag << 48 | kSegNoExit (a cycle inside
// the tile: no exit).
constexpr int kRunBits = 11;
constexpr uint32_t kRunTile = 1u << kRunBits;
constexpr int kRunThreads = 1024;
constexpr uint64_t kSegNoExit = 1ull << 49;

__global__ void __launch_bounds__(kRunThreads) k_rs_runs(const uint8_t *__restrict__ props,
                                                         const uint64_t *__restrict__ nx, uint64_t n2,
                                                         uint64_t *__restrict__ seg) {
    __shared__ uint16_t snx[2][kRunTile];
    __shared__ uint16_t scnt[2][kRunTile];
    __shared__ uint8_t sor[2][kRunTile];  // self-adjacency seen
    const uint64_t base = (uint64_t)blockIdx.x * kRunTile;
    for (uint32_t x = threadIdx.x; x < kRunTile; x += kRunThreads) {
        const uint64_t v = base + x;
        uint32_t nxt = x, c = 0, o = 0;
        if (v < n2 && (props[v] & 1)) {
            const uint64_t e = nx[v];
            const uint32_t t = (uint32_t)e, kind = (uint32_t)(e >> 32) & 3;
            // (a node that is its own successor ends its run: only its own walk,
            // if it is a ruler, reaches it)
            if (kind == 0 && t != (uint32_t)v && t >= base && t - base < kRunTile) {
                nxt = t - (uint32_t)base;
                c = 1;
                o = (uint32_t)(e >> 34) & 1;
            }
        }
        snx[0][x] = (uint16_t)nxt;
        scnt[0][x] = (uint16_t)c;
        sor[0][x] = (uint8_t)o;
    }
    __syncthreads();
    int cur = 0;
    // after round r every pointer is 2^(r+1) steps ahead or at its exit; the
    // rounds stop when no pointer moved (runs are mostly far shorter than a tile)
    for (int r = 0; r < kRunBits; r++) {
        int moved = 0;
        for (uint32_t x = threadIdx.x; x < kRunTile; x += kRunThreads) {
            const uint32_t y = snx[cur][x];
            const uint32_t z = snx[cur][y];
            moved |= z != y;
            snx[cur ^ 1][x] = (uint16_t)z;
            scnt[cur ^ 1][x] = (uint16_t)(scnt[cur][x] + scnt[cur][y]);  // at most 2^11 in a cycle
            sor[cur ^ 1][x] = (uint8_t)(sor[cur][x] | sor[cur][y]);
        }
        cur ^= 1;
        if (!__syncthreads_or(moved)) break;
    }
    for (uint32_t x = threadIdx.x; x < kRunTile; x += kRunThreads) {
        const uint64_t v = base + x;
        if (v >= n2) break;
        if (!(props[v] & 1)) continue;
        const uint32_t xe = snx[cur][x];
        const bool cyc = snx[cur][xe] != xe;
        seg[v] = (base + xe) | ((uint64_t)scnt[cur][x] << 32) | ((uint64_t)sor[cur][x] << 48) |
                 (cyc ? kSegNoExit : 0);
    }
}

// one thread per ruler: walk to the next ruler / kept node a run per step,
// labelling each run's exit with (ruler, offset); a node's offset is its
// run exit's minus its steps to the exit
__global__ void k_rs_walk(const uint32_t *__restrict__ rlist, uint64_t nr, const uint64_t *__restrict__ nx,
                          const uint64_t *__restrict__ seg, const uint8_t *__restrict__ selfadj,
                          const uint32_t *__restrict__ rid, uint64_t S, uint2 *__restrict__ label,
                          Walk *__restrict__ rw, uint8_t *__restrict__ rsp, unsigned int *__restrict__ err) {
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    uint32_t u = rlist[r];
    uint32_t sp = selfadj[u >> 1];
    Walk x;
    x.succ = kNil;
    x.term = kNil;
    x.last = u;
    uint32_t off = 0;  // nodes before the current one in this walk
    bool ended = false;
    for (uint64_t guard = 0; guard <= 2 * S; guard++) {
        const uint64_t sg = seg[u];
        if (sg & kSegNoExit) break;  // (a run entered from outside always has an exit)
        const uint32_t xe = (uint32_t)sg;
        off += (uint32_t)(sg >> 32) & 0xFFFFu;
        sp |= (uint32_t)(sg >> 48) & 1;
        label[xe] = make_uint2((uint32_t)r, off);
        const uint64_t e = nx[xe];
        const uint32_t t = (uint32_t)e;
        const uint32_t kind = (uint32_t)(e >> 32) & 3;
        if (kind == 2) {  // arrives at a kept node
            x.term = t;
            x.last = xe;
            ended = true;
            break;
        }
        if (kind == 1) {  // next ruler
            x.succ = rid[t];
            x.last = xe;
            ended = true;
            break;
        }
        sp |= (uint32_t)(e >> 34) & 1;  // the next run's first node
        off++;
        u = t;
    }
    // every run of a walk is a different one: a walk still going revisits one
    if (!ended) atomicOr(err, 2u);
    x.dist = off + 1;
    rw[r] = x;
    rsp[r] = (uint8_t)sp;
}

// pointer jumping over rulers (succ is a ruler index), two hops of the
// previous round's pointers per round: a ruler's span grows at least threefold
// per launch until its walk ends, so ceil(log3(rulers + 1)) launches finish
// every walk
__global__ void k_rs_jump(const Walk *__restrict__ win, const uint8_t *__restrict__ spin, uint64_t nr,
                          Walk *__restrict__ wout, uint8_t *__restrict__ spout) {
    const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nr) return;
    Walk x = win[v];
    uint8_t f = spin[v];
#pragma unroll
    for (int hop = 0; hop < 2; hop++) {
        if (x.succ == kNil) break;
        const uint32_t s = x.succ;
        const Walk y = win[s];
        f |= spin[s];
        x.dist += y.dist;
        x.succ = y.succ;
        x.term = y.term;
        x.last = y.last;
    }
    wout[v] = x;
    spout[v] = f;
}

// All jump rounds in one workgroup when the rulers are few (small samples):
// one launch instead of one per round.  Round r reads buffer r & 1 and writes
// the other; the barrier's workgroup-scope release/acquire makes a round's
// device-memory writes visible to every wave of the workgroup (all of them
// run on one CU; a device-scope fence would write back and invalidate the
// L2 twice a round for nothing).  A thread holds its
// kJumpOnePer rulers in registers and issues each hop's loads for all of them
// together (three waves of independent loads per round instead of a chain of
// dependent ones per ruler).
constexpr int kJumpOneThreads = 1024;
constexpr uint64_t kJumpOneMax = 8192;
constexpr int kJumpOnePer = (int)(kJumpOneMax / kJumpOneThreads);

__global__ void __launch_bounds__(kJumpOneThreads) k_rs_jump_one(Walk *__restrict__ wa, uint8_t *__restrict__ spa,
                                                                Walk *__restrict__ wb, uint8_t *__restrict__ spb,
                                                                uint64_t nr, int rounds) {
    for (int r = 0; r < rounds; r++) {
        const Walk *win = (r & 1) ? wb : wa;
        const uint8_t *spin = (r & 1) ? spb : spa;
        Walk *wout = (r & 1) ? wa : wb;
        uint8_t *spout = (r & 1) ? spa : spb;
        Walk x[kJumpOnePer];
        uint8_t f[kJumpOnePer];
#pragma unroll
        for (int j = 0; j < kJumpOnePer; j++) {
            const uint64_t v = threadIdx.x + (uint64_t)j * kJumpOneThreads;
            if (v < nr) {
                x[j] = win[v];
                f[j] = spin[v];
            } else {
                x[j].succ = kNil;
            }
        }
#pragma unroll
        for (int hop = 0; hop < 2; hop++) {
            Walk y[kJumpOnePer];
            uint8_t g[kJumpOnePer];
#pragma unroll
            for (int j = 0; j < kJumpOnePer; j++)
                if (x[j].succ != kNil) {
                    y[j] = win[x[j].succ];
                    g[j] = spin[x[j].succ];
                }
#pragma unroll
            for (int j = 0; j < kJumpOnePer; j++)
                if (x[j].succ != kNil) {
                    f[j] |= g[j];
                    x[j].dist += y[j].dist;
                    x[j].succ = y[j].succ;
                    x[j].term = y[j].term;
                    x[j].last = y[j].last;
                }
        }
#pragma unroll
        for (int j = 0; j < kJumpOnePer; j++) {
            const uint64_t v = threadIdx.x + (uint64_t)j * kJumpOneThreads;
            if (v < nr) {
                wout[v] = x[j];
                spout[v] = f[j];
            }
        }
        __syncthreads();
    }
}

// per oriented node: walk state from its ruler
__global__ void k_rs_final(const uint8_t *__restrict__ is11, const uint32_t *__restrict__ rid,
                           const uint64_t *__restrict__ seg, const uint2 *__restrict__ label,
                           const Walk *__restrict__ rw, const uint8_t *__restrict__ rsp, uint64_t S,
                           Walk *__restrict__ w, uint8_t *__restrict__ sp) {
    const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= 2 * S) return;
    Walk x;
    x.succ = kNil; x.dist = 0; x.term = kNil; x.last = kNil;
    uint8_t f = 0;
    if (is11[v >> 1]) {
        uint32_t r = rid[v], off = 0;
        if (r == kNil) {
            const uint64_t sg = seg[v];
            if (!(sg & kSegNoExit)) {  // (else a cycle inside one tile)
                const uint2 lb = label[(uint32_t)sg];
                r = lb.x;
                off = lb.y - ((uint32_t)(sg >> 32) & 0xFFFFu);
            }
        }
        if (r == kNil) {  // a cycle no ruler reached: never terminates
            x.succ = (uint32_t)v;
            f = 1;
        } else {
            x = rw[r];
            x.dist -= off;
            f = rsp[r];
            if (x.succ != kNil) x.succ = (uint32_t)v;  // ruler chain without a kept end
        }
    }
    w[v] = x;
    sp[v] = f;
}

// kept-edge id of the end a walk arrives at: slot of the kept node K whose own
// walk enters the chain at `last` (reversed).
__device__ __forceinline__ uint32_t end_slot(const uint64_t *__restrict__ key, uint32_t term, uint32_t last, int k) {
    const uint32_t a = term & 1u;
    const uint64_t L = key[last >> 1];
    const uint32_t first = (last & 1u) ? (uint32_t)(3 - (L & 3)) : (uint32_t)((L >> (2 * (k - 1))) & 3);
    return (a ? 0u : 4u) + (3u - first);
}

// kept index of survivors (exclusive scan of !is11 gives it); per kept edge:
// length, initial multiplicity, owner flag and chain size.
struct EdgeArgs {
    const uint64_t *key;
    const uint64_t *mult;
    const uint8_t *is11;
    const Walk *walk;
    const uint8_t *sp;
    const uint32_t *kept_of;   // survivor -> kept index (valid for kept survivors)
    const uint32_t *kept_src;  // kept index -> survivor
    uint64_t n_kept;
    MergeArgs ma;
    uint32_t *elen;     // 8 per kept node (0: empty or special)
    uint8_t *emult;     // 8 per kept node
    uint32_t *eother;   // other end kept-edge id (for chains), else kNil
    uint32_t *csize;    // owner: chain nodes + 2 sentinels, else 0
    uint8_t *especial;  // 1 if the edge enters a cycle / fold-back component
    unsigned long long *counters;  // [2] special edges
};

__global__ void k_kept_edges(EdgeArgs e) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= 8 * e.n_kept) return;
    const uint32_t kap = (uint32_t)(g >> 3);
    const int s = (int)(g & 7);
    const uint32_t i = e.kept_src[kap];
    const uint64_t m = e.mult[i];
    const uint8_t mm = (uint8_t)(m >> (8 * s));
    e.emult[g] = mm;
    e.elen[g] = 0;
    e.eother[g] = kNil;
    e.csize[g] = 0;
    e.especial[g] = 0;
    if (!mm) return;
    const uint32_t t = hop(e.ma, e.key[i], s);
    if (t == kNil) return;
    const uint32_t j = t >> 1;
    if (!e.is11[j]) {
        e.elen[g] = 1;
        return;
    }
    const Walk w = e.walk[t];
    if (w.succ != kNil || e.sp[t] || w.term == kNil) {
        e.especial[g] = 1;
        atomicAdd(&e.counters[2], 1ull);
        return;
    }
    e.elen[g] = w.dist + 1;
    const uint32_t K2 = w.term >> 1;
    const uint32_t other = 8 * e.kept_of[K2] + end_slot(e.key, w.term, w.last, e.ma.k);
    e.eother[g] = other;
    if ((uint32_t)g < other) e.csize[g] = w.dist + 2;
}

// ---------------------------------------------------------------------------
// strings and chain arrays

struct FillArgs {
    const uint64_t *key;
    const uint8_t *is11;
    const Walk *walk;
    const uint8_t *sp;
    const uint32_t *kept_of;
    const uint64_t *eoff;     // pool offset per kept edge
    const uint32_t *elen;
    const uint64_t *coff;     // chain segment offset per owner edge
    const uint32_t *eother;
    const uint8_t *ex;        // exit multiplicities per oriented node
    const uint64_t *prio;     // visit priority per survivor
    uint64_t S;
    int k;
    char *pool;
    uint64_t *cpri;           // chain priorities (sentinels = kInf)
    uint8_t *cmul;            // unit-edge multiplicity leaving each chain position
    uint32_t *cedge;          // at sentinel positions: the kept edge id
};

__device__ __forceinline__ char base_char(uint32_t b) { return "ACGT"[b & 3]; }

__device__ __forceinline__ uint32_t last_base(uint64_t X, uint32_t o, int k) {
    return o ? (uint32_t)(3 - ((X >> (2 * (k - 1))) & 3)) : (uint32_t)(X & 3);
}

// For every normal (1,1) node and both orientations: its base in the string of
// the edge it is reached from; and its slot in the owner's chain array.
__global__ void k_fill_nodes(FillArgs f) {
    const uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= 2 * f.S) return;
    const uint64_t j = v >> 1;
    if (!f.is11[j]) return;
    const uint32_t o = (uint32_t)(v & 1);
    const uint64_t vr = v ^ 1;  // the walk back to the end this orientation starts from
    const Walk b = f.walk[vr];
    if (b.succ != kNil || b.term == kNil || f.sp[v] || f.sp[vr]) return;
    const Walk a = f.walk[v];
    const uint32_t startEdge = 8 * f.kept_of[b.term >> 1] + end_slot(f.key, b.term, b.last, f.k);
    const uint64_t X = f.key[j];
    f.pool[f.eoff[startEdge] + b.dist - 1] = base_char(last_base(X, o, f.k));
    // chain array (written once, from the owner's direction)
    const uint32_t endEdge = 8 * f.kept_of[a.term >> 1] + end_slot(f.key, a.term, a.last, f.k);
    if (startEdge < endEdge) {
        const uint64_t base = f.coff[startEdge];
        const uint64_t pos = base + b.dist;  // sentinel at base, nodes at 1..n
        f.cpri[pos] = f.prio[j];
        f.cmul[pos] = f.ex[v];
    }
}

// kept edges: final base; chain sentinels and the first unit edge
__global__ void k_fill_edges(FillArgs f, const uint8_t *__restrict__ emult, uint64_t n_edges) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_edges) return;
    const uint32_t L = f.elen[g];
    if (!L) return;
    // the last base of the edge is the last base of the oriented k-mer reached
    // one hop (L == 1) or a chain (L > 1) away; computed by the caller-side walk
    // through the survivor arrays would need the terminal, which is recorded in
    // eother for chains -- for L == 1 it is the slot base itself.
    if (L == 1) {
        f.pool[f.eoff[g]] = base_char((uint32_t)(g & 7));
        return;
    }
    const uint32_t other = f.eother[g];
    if ((uint32_t)g < other) {
        const uint64_t base = f.coff[g];
        const uint64_t n = L - 1;
        f.cpri[base] = kInf;
        f.cmul[base] = emult[g];
        f.cedge[base] = (uint32_t)g;
        f.cpri[base + n + 1] = kInf;
        f.cmul[base + n + 1] = 0;
        f.cedge[base + n + 1] = other;
    }
}

// Last base of every chain edge: the other end's kept node, as reached.
__global__ void k_fill_last(const uint32_t *__restrict__ elen, const uint32_t *__restrict__ eother,
                            const uint32_t *__restrict__ kept_src, const uint64_t *__restrict__ key,
                            const uint64_t *__restrict__ eoff, uint64_t n_edges, int k, char *__restrict__ pool) {
    const uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_edges) return;
    const uint32_t L = elen[g];
    if (L < 2) return;
    const uint32_t other = eother[g];
    // arrival orientation at the other kept node: its slot half says from which
    // strand its own walk leaves, the arrival is on the opposite strand
    const uint64_t K2 = key[kept_src[other >> 3]];
    const uint32_t oslot = other & 7;
    const uint32_t arr = oslot < 4 ? 1u : 0u;  // other walks from K2 (slots 0-3) => we arrive at rc(K2)
    pool[eoff[g] + L - 1] = base_char(last_base(K2, arr, k));
}

// ---------------------------------------------------------------------------
// nearest larger priority (max hierarchy) and Cartesian-tree evaluation

// Max hierarchy: level l + 1 holds the max of each pair of level l (0 past
// the end).  One workgroup reduces kMaxTile elements of level l0 through up
// to kMaxSteps levels in LDS; src is level l0 (level 0 is copied into the
// hierarchy on the way).
constexpr int kMaxSteps = 11;
constexpr uint32_t kMaxTile = 1u << kMaxSteps;
constexpr int kMaxThreads = 256;

struct LevTab {
    uint64_t off[64];
    uint64_t len[64];
};

__global__ void __launch_bounds__(kMaxThreads) k_max_levels(const uint64_t *__restrict__ src, uint64_t *__restrict__ lev,
                                                            LevTab lt, int l0, int l1) {
    __shared__ uint64_t buf[2][kMaxTile];
    const uint64_t base = (uint64_t)blockIdx.x * kMaxTile;
    const uint64_t n0 = lt.len[l0];
    for (uint32_t j = threadIdx.x; j < kMaxTile; j += kMaxThreads) {
        const uint64_t i = base + j;
        const uint64_t v = i < n0 ? src[i] : 0;
        buf[0][j] = v;
        if (l0 == 0 && i < n0) lev[i] = v;
    }
    __syncthreads();
    int cur = 0;
    for (int l = l0 + 1; l <= l1; l++) {
        const uint32_t size = kMaxTile >> (l - l0);
        const uint64_t lbase = base >> (l - l0);
        for (uint32_t j = threadIdx.x; j < size; j += kMaxThreads) {
            const uint64_t a = buf[cur][2 * j], b = buf[cur][2 * j + 1];
            const uint64_t m = a > b ? a : b;
            buf[cur ^ 1][j] = m;
            if (lbase + j < lt.len[l]) lev[lt.off[l] + lbase + j] = m;
        }
        cur ^= 1;
        __syncthreads();
    }
}

struct TreeArgs {
    const uint64_t *lev;      // concatenated levels (level 0: the chain priorities)
    uint64_t lev_off[64];     // offset of each level (passed by value: no copy to the device)
    int nlev;
    uint64_t M;               // chain array length
    const uint8_t *cmul;
    const uint32_t *cedge;
    uint32_t *par;            // parent position, kNil for roots / sentinels
    uint8_t *need;            // children to wait for (0..2); 255 for sentinels
    uint32_t *nl;             // nearest larger position on each side
    uint32_t *nr;
    uint32_t *far;            // positions with a nearest larger outside their tile
    unsigned int *nfar;
    unsigned int *err;        // bit 0: a search left the array (a chain without its sentinel); bit 2: a
                              // climb passed more parents than positions
};

__device__ __forceinline__ uint64_t levv(const TreeArgs &t, int l, uint64_t i) { return t.lev[t.lev_off[l] + i]; }

// Nearest position left / right of x whose priority exceeds v, by the global
// max hierarchy: skip aligned blocks whose max is at most v, then descend into
// the first block that holds a larger one.  A chain array position always has
// a sentinel (kInf) on both sides, so leaving the array is an error.
__device__ uint64_t nearest_left(const TreeArgs &t, uint64_t x, uint64_t v) {
    int64_t p = (int64_t)x - 1;
    int l = 0;
    while (true) {
        if (p < 0) { atomicOr(t.err, 1u); return x - 1; }
        while (l + 1 < t.nlev && (((uint64_t)p + 1) & ((2ull << l) - 1)) == 0) l++;
        if (levv(t, l, (((uint64_t)p + 1) >> l) - 1) > v) break;
        p -= (int64_t)1 << l;
    }
    while (l > 0) {
        l--;
        if (levv(t, l, (((uint64_t)p + 1) >> l) - 1) <= v) p -= (int64_t)1 << l;
    }
    return (uint64_t)p;
}

__device__ uint64_t nearest_right(const TreeArgs &t, uint64_t x, uint64_t v) {
    uint64_t p = x + 1;
    int l = 0;
    while (true) {
        if (p >= t.M) { atomicOr(t.err, 1u); return x + 1; }
        while (l + 1 < t.nlev && (p & ((2ull << l) - 1)) == 0 && (p >> (l + 1)) < ((t.M + (2ull << l) - 1) >> (l + 1))) l++;
        if (levv(t, l, p >> l) > v) break;
        p += 1ull << l;
    }
    while (l > 0) {
        l--;
        if (levv(t, l, p >> l) <= v) p += 1ull << l;
    }
    return p;
}

// parent and children to wait for, from the nearest larger neighbours
__device__ __forceinline__ void tree_node(const TreeArgs &t, uint64_t q, uint64_t L, uint64_t R, uint64_t vl,
                                          uint64_t vr) {
    t.nl[q] = (uint32_t)L;
    t.nr[q] = (uint32_t)R;
    t.par[q] = (vl == kInf && vr == kInf) ? kNil : (uint32_t)(vl < vr ? L : R);
    t.need[q] = (uint8_t)((L != q - 1) + (R != q + 1));
}

// Links of the Cartesian trees, one tile of kTreeTile positions per
// workgroup: each position's nearest larger neighbours are searched in an LDS
// max hierarchy of its tile (12 levels).  A position with one outside the
// tile (a few per tile: the maxima seen from the tile's edges) goes to a list
// that k_tree_far finishes with the global hierarchy.
constexpr int kTreeBits = 11;
constexpr uint32_t kTreeTile = 1u << kTreeBits;
constexpr int kTreeLinkThreads = 1024;
constexpr int kTreeEvalThreads = 1024;
constexpr uint32_t kTreeLevTotal = 2 * kTreeTile - 1;  // entries of levels 0..kTreeBits

__device__ __forceinline__ uint32_t tree_loff(int l) { return 2 * kTreeTile - (2 * kTreeTile >> l); }

__global__ void __launch_bounds__(kTreeLinkThreads) k_tree_links(TreeArgs t) {
    __shared__ uint64_t lv[kTreeLevTotal];
    __shared__ uint32_t sfar[kTreeTile];
    __shared__ uint32_t nfar, far_base;
    const uint64_t base = (uint64_t)blockIdx.x * kTreeTile;
    if (threadIdx.x == 0) nfar = 0;
    for (uint32_t j = threadIdx.x; j < kTreeTile; j += kTreeLinkThreads) {
        const uint64_t q = base + j;
        lv[j] = q < t.M ? t.lev[q] : 0;  // 0 past the end: never larger than a priority
    }
    __syncthreads();
    for (int l = 1; l <= kTreeBits; l++) {
        const uint32_t n = kTreeTile >> l, o = tree_loff(l), po = tree_loff(l - 1);
        for (uint32_t j = threadIdx.x; j < n; j += kTreeLinkThreads) {
            const uint64_t a = lv[po + 2 * j], b = lv[po + 2 * j + 1];
            lv[o + j] = a > b ? a : b;
        }
        __syncthreads();
    }
    for (uint32_t x = threadIdx.x; x < kTreeTile; x += kTreeLinkThreads) {
        const uint64_t q = base + x;
        if (q >= t.M) break;
        const uint64_t v = lv[x];
        if (v == kInf) {  // sentinel
            t.par[q] = kNil;
            t.need[q] = 255;
            continue;
        }
        uint32_t L = kNil, R = kNil;
        // most nearest larger neighbours are a few positions away: four direct
        // probes per side (independent reads) before the hierarchy search
        for (int d = 4; d >= 1; d--) {
            if ((int)x - d >= 0 && lv[x - d] > v) L = x - d;
            if (x + d < kTreeTile && lv[x + d] > v) R = x + d;
        }
        if (L == kNil) {
            int p = (int)x - 5, l = 0;
            while (p >= 0) {
                while (l < kTreeBits && ((p + 1) & ((2 << l) - 1)) == 0) l++;
                if (lv[tree_loff(l) + (((uint32_t)p + 1) >> l) - 1] > v) break;
                p -= 1 << l;
            }
            if (p >= 0) {
                while (l > 0) {
                    l--;
                    if (lv[tree_loff(l) + (((uint32_t)p + 1) >> l) - 1] <= v) p -= 1 << l;
                }
                L = (uint32_t)p;
            }
        }
        if (R == kNil) {
            uint32_t p = x + 5;
            int l = 0;
            while (p < kTreeTile) {
                while (l < kTreeBits && (p & ((2u << l) - 1)) == 0) l++;
                if (lv[tree_loff(l) + (p >> l)] > v) break;
                p += 1u << l;
            }
            if (p < kTreeTile) {
                while (l > 0) {
                    l--;
                    if (lv[tree_loff(l) + (p >> l)] <= v) p += 1u << l;
                }
                R = p;
            }
        }
        if (L != kNil && R != kNil) {
            tree_node(t, q, base + L, base + R, lv[L], lv[R]);
        } else {
            // finished by k_tree_far; the side found here travels in nl / nr
            t.nl[q] = L == kNil ? kNil : (uint32_t)(base + L);
            t.nr[q] = R == kNil ? kNil : (uint32_t)(base + R);
            sfar[atomicAdd(&nfar, 1u)] = (uint32_t)q;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0 && nfar) far_base = atomicAdd(t.nfar, nfar);
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < nfar; j += kTreeLinkThreads) t.far[far_base + j] = sfar[j];
}

// A side k_tree_links left open has nothing larger between q and its tile's
// edge, so the search starts at the edge: the hierarchy is entered at the tile
// level at once instead of climbing out of the tile one level per load.
__global__ void k_tree_far(TreeArgs t) {
    const uint32_t n = *t.nfar;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t q = t.far[j];
        const uint64_t v = t.lev[q];
        const uint32_t l0 = t.nl[q], r0 = t.nr[q];
        const uint64_t tb = q & ~(uint64_t)(kTreeTile - 1);  // q's tile: [tb, tb + kTreeTile)
        const uint64_t L = l0 != kNil ? l0 : nearest_left(t, tb, v);
        const uint64_t R = r0 != kNil ? r0 : nearest_right(t, tb + kTreeTile - 1, v);
        if (L >= t.M || R >= t.M) {  // reported through err; keep the position out of the evaluation
            t.par[q] = kNil;
            t.need[q] = 255;
            continue;
        }
        tree_node(t, q, L, R, t.lev[L], t.lev[R]);
    }
}

// merged (multiplicity, length) of two spliced edges, with the reference's
// arithmetic: the sum is formed in size_t, stored into an int, divided by the
// size_t length and stored into an unsigned char (graph.cpp:1343-1345).
__device__ __forceinline__ uint64_t splice(uint64_t a, uint64_t b) {
    const uint64_t ma = a & 0xFF, la = a >> 8, mb = b & 0xFF, lb = b >> 8;
    const uint64_t s = ma * la + mb * lb;
    const int32_t pre = (int32_t)(uint32_t)s;
    // a non-negative 32-bit numerator over a 32-bit length divides in 32 bits
    const uint64_t q = (pre >= 0 && la + lb <= 0xFFFFFFFFull)
                           ? (uint64_t)((uint32_t)pre / (uint32_t)(la + lb))
                           : (uint64_t)(int64_t)pre / (la + lb);
    return ((la + lb) << 8) | (q & 0xFF);
}

// multiplicity of node p's merged edge from its left and right inputs: the
// lengths are fixed by positions (p spans the unit edges between its nearest
// larger neighbours L and R)
__device__ __forceinline__ uint32_t tree_eval(uint64_t p, uint64_t L, uint64_t R, uint32_t mL, uint32_t mR) {
    return (uint32_t)(splice(((p - L) << 8) | mL, ((R - p) << 8) | mR) & 0xFF);
}

// Bottom-up evaluation, one tile per workgroup: one thread climbs from each
// leaf, the last-arriving child of a parent finishing it.  A node whose span
// (L, R) lies within [tile - 1, tile end] has all its descendants in the tile:
// its children arrive through a 32-bit LDS word (left multiplicity bits 0-7,
// right 8-15, arrival flags 16, 17) and it is finished in LDS.  The others --
// those whose span crosses a tile edge, a few per tile, and all their
// ancestors -- take their arrivals in a 64-bit word in device memory (same
// layout; the word carries the value, so one atomicOr publishes it with no
// fence).
__global__ void __launch_bounds__(kTreeEvalThreads) k_tree_eval(TreeArgs t, unsigned long long *__restrict__ word,
                                                               uint8_t *__restrict__ emult_out) {
    __shared__ uint32_t snl[kTreeTile], snr[kTreeTile], spar[kTreeTile], sword[kTreeTile];
    __shared__ uint8_t sneed[kTreeTile];
    __shared__ uint8_t smul[kTreeTile + 1];  // cmul of positions base - 1 .. base + kTreeTile - 1
    const uint64_t base = (uint64_t)blockIdx.x * kTreeTile;
    for (uint32_t j = threadIdx.x; j < kTreeTile; j += kTreeEvalThreads) {
        const uint64_t q = base + j;
        const bool in = q < t.M;
        sneed[j] = in ? t.need[q] : 255;
        const bool node = in && sneed[j] != 255;
        snl[j] = node ? t.nl[q] : kNil;
        snr[j] = node ? t.nr[q] : kNil;
        spar[j] = node ? t.par[q] : kNil;
        smul[j + 1] = in ? t.cmul[q] : 0;
        sword[j] = 0;
    }
    if (threadIdx.x == 0) smul[0] = base ? t.cmul[base - 1] : 0;
    __syncthreads();
    for (uint32_t x = threadIdx.x; x < kTreeTile; x += kTreeEvalThreads) {
        if (sneed[x] != 0) continue;
        uint64_t q = base + x;
        uint64_t L = snl[x], R = snr[x], p = spar[x];
        uint32_t m = tree_eval(q, L, R, smul[x], smul[x + 1]);
        // in the tile while the parent's span lies within it
        bool local = true;
        while (p != kNil) {
            const uint64_t pl = p - base;
            if (p < base || pl >= kTreeTile || (uint64_t)snl[pl] + 1 < base || snr[pl] > base + kTreeTile) {
                local = false;
                break;
            }
            const uint32_t side = q < p ? 0u : 1u;
            const uint32_t mine = (m << (8 * side)) | (1u << (16 + side));
            const uint32_t w = atomicOr(&sword[pl], mine) | mine;
            const uint64_t PL = snl[pl], PR = snr[pl];
            const bool needL = PL != p - 1, needR = PR != p + 1;
            if ((needL && !(w & (1u << 16))) || (needR && !(w & (1u << 17)))) break;  // the sibling finishes p
            const uint32_t mL = needL ? (w & 0xFF) : smul[pl];
            const uint32_t mR = needR ? ((w >> 8) & 0xFF) : smul[pl + 1];
            m = tree_eval(p, PL, PR, mL, mR);
            q = p;
            L = PL;
            R = PR;
            p = spar[pl];
        }
        if (local && p != kNil) continue;  // stopped at a parent its sibling finishes
        // through device memory from here (or q is a root)
        bool done = false;
        for (uint64_t steps = 0; steps <= t.M; steps++) {
            if (p == kNil) {  // the root: its merged edge is both ends' kept edge
                if (L < t.M && R < t.M) {  // (else q was excluded by a reported error)
                    emult_out[t.cedge[L]] = (uint8_t)m;
                    emult_out[t.cedge[R]] = (uint8_t)m;
                }
                done = true;
                break;
            }
            // the parent's links load while the arrival is published
            const uint64_t PL = t.nl[p], PR = t.nr[p], PP = t.par[p];
            const uint32_t side = q < p ? 0u : 1u;
            const unsigned long long mine = ((unsigned long long)m << (8 * side)) | (1ull << (16 + side));
            const unsigned long long w = atomicOr(&word[p], mine) | mine;
            if (t.need[p] == 255) { done = true; break; }  // excluded by a reported error
            const bool needL = PL != p - 1, needR = PR != p + 1;
            if ((needL && !(w & (1ull << 16))) || (needR && !(w & (1ull << 17)))) { done = true; break; }
            const uint32_t mL = needL ? (uint32_t)(w & 0xFF) : t.cmul[p - 1];
            const uint32_t mR = needR ? (uint32_t)((w >> 8) & 0xFF) : t.cmul[p];
            m = tree_eval(p, PL, PR, mL, mR);
            q = p;
            L = PL;
            R = PR;
            p = PP;
        }
        if (!done) atomicOr(t.err, 4u);  // more than M parents: the parent links hold a cycle
    }
}

// ---------------------------------------------------------------------------
// helpers

__global__ void k_prio_hash(const uint64_t *__restrict__ key, uint64_t S, Mixer mx, uint64_t *__restrict__ prio) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < S) prio[i] = mix_fwd(mx, key[i]);
}

__global__ void k_u32_to_u64(const uint32_t *__restrict__ in, uint64_t n, uint64_t *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = in[i];
}

__global__ void k_gather_u64(const uint64_t *__restrict__ src, const uint32_t *__restrict__ idx, uint64_t n,
                             uint64_t *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = src[idx[i]];
}

template <typename T>
__global__ void k_fill(T *__restrict__ p, uint64_t n, T v) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

static inline unsigned nblk(uint64_t n, unsigned b = 256) { return (unsigned)((n + b - 1) / b); }

}  // namespace swapk

namespace swapk {

__global__ void k_special_flags(const uint8_t *__restrict__ is11, const Walk *__restrict__ w, const uint8_t *__restrict__ sp,
                                uint64_t S, uint32_t *__restrict__ list, unsigned long long *__restrict__ n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool f = false;
    if (i < S && is11[i]) {
        const Walk a = w[2 * i], b = w[2 * i + 1];
        f = a.succ != kNil || b.succ != kNil || a.term == kNil || b.term == kNil || sp[2 * i] || sp[2 * i + 1];
    }
    const uint64_t bal = __ballot(f);
    if (!bal) return;
    const int lane = threadIdx.x & 63;
    unsigned long long base = 0;
    if (lane == __ffsll((long long)bal) - 1) base = atomicAdd(n, (unsigned long long)__popcll(bal));
    base = __shfl(base, __ffsll((long long)bal) - 1, 64);
    if (f) list[base + __popcll(bal & ((1ull << lane) - 1))] = (uint32_t)i;
}

#define MCHECK(x)                                                            \
    do {                                                                     \
        hipError_t e_ = (x);                                                 \
        if (e_ != hipSuccess) return fail(-5, hipGetErrorString(e_));        \
    } while (0)

// Simple cycles on the device.  An isolated cycle of (1,1) nodes, none of
// them its own neighbour and the cycle not its own reverse complement, merges
// in the reference's visit order with every lock granted (each node's two
// neighbours are other nodes while two or more are left), so every node but the
// last-visited one is absorbed: the result is the chain from the cycle's
// highest-priority node around to itself.  Making that node a kept node turns
// the cycle into such a chain for the list ranking, the strings and the
// Cartesian trees.  Candidates: oriented (1,1) nodes whose walk never reaches a
// kept node; each learns its cycle's maximum (priority, oriented id) and
// whether a self-adjacent node lies on it by pointer doubling.
struct CyclePred {
    const uint8_t *is11;
    const Walk *walk;
    __device__ bool operator()(uint64_t v) const { return is11[v >> 1] && walk[v].succ != kNil; }
};

__global__ void k_cyc_init(const uint32_t *__restrict__ cl, uint64_t nc, const uint32_t *__restrict__ cpos,
                           const uint32_t *__restrict__ nxt, const uint64_t *__restrict__ prio,
                           const uint8_t *__restrict__ selfadj, uint64_t *__restrict__ key, uint32_t *__restrict__ nn,
                           uint8_t *__restrict__ bad) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nc) return;
    const uint32_t v = cl[j];
    key[j] = (prio[v >> 1] << 32) | v;
    const uint32_t t = nxt[v];
    const uint32_t q = t == kNil ? kNil : cpos[t];
    nn[j] = q == kNil ? (uint32_t)j : q;  // (a successor off the candidates: not a cycle, never converted)
    bad[j] = (uint8_t)((selfadj[v >> 1] ? 1 : 0) | (q == kNil ? 1 : 0));
}

__global__ void k_cyc_jump(const uint64_t *__restrict__ kin, const uint32_t *__restrict__ nin,
                           const uint8_t *__restrict__ bin, uint64_t nc, uint64_t *__restrict__ kout,
                           uint32_t *__restrict__ nout, uint8_t *__restrict__ bout) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nc) return;
    const uint32_t q = nin[j];
    const uint64_t a = kin[j], b = kin[q];
    kout[j] = a > b ? a : b;
    bout[j] = bin[j] | bin[q];
    nout[j] = nin[q];
}

// the cycle's maximum node becomes a kept node where the cycle and its reverse
// complement are two cycles (their maxima are the node's two orientations)
__global__ void k_cyc_mark(const uint32_t *__restrict__ cl, uint64_t nc, const uint32_t *__restrict__ cpos,
                           const uint64_t *__restrict__ key, const uint8_t *__restrict__ bad,
                           uint8_t *__restrict__ is11, unsigned long long *__restrict__ nconv) {
    const uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nc) return;
    const uint32_t v = cl[j];
    const uint64_t m = key[j];
    if ((uint32_t)m != v) return;  // one thread per cycle: its maximum's own candidate
    const uint32_t r = cpos[v ^ 1u];
    if (r == kNil || bad[j] || bad[r]) return;
    if (key[r] != (m ^ 1ull)) return;  // the cycle is its own reverse complement
    is11[v >> 1] = 0;  // (both orientations' cycles decide alike)
    if (!(v & 1u)) atomicAdd(nconv, 1ull);
}

int Pipeline::merge_device(bool reference_order, MergeDev &out) {
    MCHECK(hipSetDevice(dev_));
    out = MergeDev();
    const uint64_t S = n_surv;
    if (S == 0) return 0;
    if (S >= 0x7FFFFFF0ull) return fail(-4, "too many survivors for the merge");
    uint64_t *d_small = small_.as<uint64_t>();
    uint64_t *hs = h_small_;
    MCHECK(hipMemsetAsync(d_small + 16, 0, 16 * 8, st_));
    unsigned long long *cnt = (unsigned long long *)(d_small + 16);
    // --- survivor index: one region per count partition
    const uint64_t nparts = 1ull << pbits_;
    // 256 / load, rounded up so that the load is never exceeded
    const uint32_t mul256 = (25600u + tune_.index_load - 1) / tune_.index_load;
    const uint64_t ent_bound = 2 * ((S * mul256) >> 8) + 2 * nparts + 2;  // region_size(n) <= 2 (n mul + 1)
    MCHECK(mb_[M_IBASE].reserve((2 * nparts + 1) * 8));
    MCHECK(mb_[M_IENT].reserve(ent_bound * 16));
    MCHECK(scan_scratch_.reserve((scan_scratch_words(nparts) + 16) * 8));
    uint64_t *ibase = mb_[M_IBASE].as<uint64_t>();
    uint32_t *isize = reinterpret_cast<uint32_t *>(ibase + nparts + 1);
    SurvIndex ix;
    ix.ent = mb_[M_IENT].as<uint64_t>();
    ix.base = ibase;
    ix.mx = mx_;
    ix.shift = 2 * k_ - pbits_;
    phase_begin(8);
    // survivors renumbered in first-occurrence order (reference order: the
    // first occurrences are known), once per build; the index is built from
    // the survivors in the count's partition runs, with their new numbers
    const bool runs = !surv_renumbered_;
    bool renumbered = false;
    if (reference_order && !surv_renumbered_) {
        const int rc = renumber_survivors(&renumbered);
        if (rc) return rc;
    }
    const uint32_t *perm = renumbered ? rn_[R_PERM].as<uint32_t>() : nullptr;
    hipLaunchKernelGGL(k_index_sizes, dim3((unsigned)(nparts / 256)), dim3(256), 0, st_, part_surv_.as<uint32_t>(),
                       (uint32_t)nparts, mul256, isize);
    MCHECK(scan_excl<uint32_t>(isize, ibase, nparts, scan_scratch_.as<uint64_t>(), ibase + nparts, st_));
    hipLaunchKernelGGL(k_index_part, dim3((unsigned)nparts), dim3(kIdxThreads), 0, st_, surv_key_.as<uint64_t>(),
                       part_surv_.as<uint32_t>(), part_first_.as<uint32_t>(), ix, mb_[M_IENT].as<uint64_t>(), perm,
                       runs);
    hipLaunchKernelGGL(k_index_loose, dim3(nblk(S)), dim3(256), 0, st_, surv_key_.as<uint64_t>(), S,
                       part_first_.as<uint32_t>(), ix, mb_[M_IENT].as<uint64_t>(), perm, runs);
    if (renumbered) {  // the new key order in place of the count's (multiplicities
        // and first occurrences were written in place); copied, not swapped: the
        // survivor arrays keep the capacity the count reserves, the
        // renumbering's stay at S
        MCHECK(hipMemcpyAsync(surv_key_.p, rn_[R_KEY].p, S * 8, hipMemcpyDeviceToDevice, st_));
        surv_renumbered_ = true;
    }
    // --- priorities (visit time)
    MCHECK(mb_[M_PRIO].reserve(S * 8));
    phase_end(8);
    // Reference order: the visit-order replay runs on the side stream (behind
    // the build's growth-rank selection there), overlapping classification,
    // list ranking and the kept edges; the main stream waits for it (prio_done_)
    // before the first reader of the priorities.  Every return from here on
    // leaves the side stream idle.
    struct SideIdle {
        hipStream_t s;
        ~SideIdle() { (void)hipStreamSynchronize(s); }
    } side_idle{st2_};
    if (reference_order) {
        MCHECK(hipEventRecord(fork_, st_));
        MCHECK(hipStreamWaitEvent(st2_, fork_, 0));
        phase_begin(13, st2_);
        const int rc = visit_priorities(mb_[M_PRIO].as<uint64_t>(), st2_);
        if (rc) return rc;
        phase_end(13, st2_);
        MCHECK(hipEventRecord(prio_done_, st2_));
    } else {
        phase_begin(13);
        hipLaunchKernelGGL(k_prio_hash, dim3(nblk(S)), dim3(256), 0, st_, surv_key_.as<uint64_t>(), S, mx_,
                           mb_[M_PRIO].as<uint64_t>());
        phase_end(13);
        MCHECK(hipEventRecord(prio_done_, st_));
    }
    phase_begin(14);
    // --- classification
    MergeArgs ma;
    ma.key = surv_key_.as<uint64_t>();
    ma.mult = surv_mult_.as<uint64_t>();
    ma.S = S;
    ma.k = k_;
    ma.ix = ix;
    ma.near = surv_renumbered_;
    MCHECK(mb_[M_IS11].reserve(S));
    MCHECK(mb_[M_NXT].reserve(2 * S * 4));
    MCHECK(mb_[M_EX].reserve(2 * S));
    MCHECK(mb_[M_SELF].reserve(S));
    hipLaunchKernelGGL(k_classify, dim3(nblk(S)), dim3(256), 0, st_, ma, mb_[M_IS11].as<uint8_t>(),
                       mb_[M_NXT].as<uint32_t>(), mb_[M_EX].as<uint8_t>(), mb_[M_SELF].as<uint8_t>(), cnt);
    phase_end(14);
    // --- list ranking: ruling set, then pointer jumping over the rulers; a
    // second pass when simple cycles became chains (their maxima kept)
    uint64_t nconv = 0;
    Walk *wa = nullptr;
    uint8_t *spa = nullptr;
    uint64_t n11 = 0, missing = 0, nk = 0, nspec = 0;
    const KeptPred kept{mb_[M_IS11].as<uint8_t>()};
    for (int pass = 0;; pass++) {
    MCHECK(mb_[M_WALKA].reserve(2 * S * sizeof(Walk)));
    MCHECK(mb_[M_WALKB].reserve(2 * S * sizeof(Walk)));
    MCHECK(mb_[M_SPA].reserve(2 * S));
    MCHECK(mb_[M_SPB].reserve(2 * S));
    const uint64_t nbc = compact_blocks(2 * S);
    MCHECK(mb_[M_BCNT].reserve(nbc * 4 + 4));
    MCHECK(mb_[M_BOFF].reserve(nbc * 8 + 8));
    MCHECK(mb_[M_NL].reserve(2 * S * 4));      // ruler id per oriented node
    MCHECK(mb_[M_NR].reserve(2 * S * 8));      // label: (ruler, offset)
    MCHECK(mb_[M_PAR].reserve(2 * S * 8));     // packed successor
    MCHECK(scan_scratch_.reserve((scan_scratch_words(nbc + 1) + 16) * 8));
    wa = mb_[M_WALKA].as<Walk>();
    Walk *wb = mb_[M_WALKB].as<Walk>();
    spa = mb_[M_SPA].as<uint8_t>();
    uint8_t *spb = mb_[M_SPB].as<uint8_t>();
    uint32_t *rid = mb_[M_NL].as<uint32_t>();
    uint2 *label = mb_[M_NR].as<uint2>();
    uint64_t *nx = mb_[M_PAR].as<uint64_t>();
    phase_begin(9);
    MCHECK(mb_[M_PROPS].reserve(2 * S));
    hipLaunchKernelGGL(k_rs_mark, dim3(nblk(2 * S)), dim3(256), 0, st_, mb_[M_IS11].as<uint8_t>(),
                       mb_[M_NXT].as<uint32_t>(), mb_[M_SELF].as<uint8_t>(), S, mb_[M_PROPS].as<uint8_t>(),
                       tune_.ruler_stride, label);
    const RulerPred ruler{mb_[M_PROPS].as<uint8_t>()};
    MCHECK(compact_count_scan(2 * S, ruler, mb_[M_BCNT].as<uint32_t>(), mb_[M_BOFF].as<uint64_t>(),
                              scan_scratch_.as<uint64_t>(), d_small + 26, st_));
    MCHECK(hipMemcpyAsync(hs + 26, d_small + 26, 8, hipMemcpyDeviceToHost, st_));
    MCHECK(hipStreamSynchronize(st_));
    const uint64_t nr = hs[26];
    MCHECK(mb_[M_ELEN].reserve(nr * 4 + 4));   // ruler list (reused below for edge lengths)
    MCHECK(mb_[M_CPRI].reserve((2 * nr + 2) * sizeof(Walk) + 16));
    MCHECK(mb_[M_CMUL].reserve(2 * nr + 16));
    uint32_t *rlist = mb_[M_ELEN].as<uint32_t>();
    Walk *rwa = mb_[M_CPRI].as<Walk>(), *rwb = rwa + nr + 1;
    uint8_t *rspa = mb_[M_CMUL].as<uint8_t>(), *rspb = rspa + nr + 1;
    hipLaunchKernelGGL(k_rs_prep, dim3(nblk(2 * S)), dim3(256), 0, st_, mb_[M_PROPS].as<uint8_t>(),
                       mb_[M_NXT].as<uint32_t>(), S, nx);
    MCHECK(mb_[M_SEG].reserve(2 * S * 8));
    uint64_t *seg = mb_[M_SEG].as<uint64_t>();
    hipLaunchKernelGGL(k_rs_runs, dim3((unsigned)((2 * S + kRunTile - 1) / kRunTile)), dim3(kRunThreads), 0, st_,
                       mb_[M_PROPS].as<uint8_t>(), nx, 2 * S, seg);
    compact_scatter<true>(2 * S, ruler, mb_[M_BOFF].as<uint64_t>(), rid, rlist, st_);
    if (nr) {
        hipLaunchKernelGGL(k_rs_walk, dim3(nblk(nr)), dim3(256), 0, st_, rlist, nr, nx, seg,
                           mb_[M_SELF].as<uint8_t>(), rid, S, label, rwa, rspa, (unsigned int *)(d_small + 25));
        int rounds = 1;
        for (uint64_t span = 3; span < nr + 2; span *= 3) rounds++;
        rounds++;
        if (nr <= kJumpOneMax) {
            hipLaunchKernelGGL(k_rs_jump_one, dim3(1), dim3(kJumpOneThreads), 0, st_, rwa, rspa, rwb, rspb, nr,
                               rounds);
            if (rounds & 1) {  // the last round wrote the second buffer
                std::swap(rwa, rwb);
                std::swap(rspa, rspb);
            }
        } else {
            for (int r = 0; r < rounds; r++) {
                hipLaunchKernelGGL(k_rs_jump, dim3(nblk(nr)), dim3(256), 0, st_, rwa, rspa, nr, rwb, rspb);
                std::swap(rwa, rwb);
                std::swap(rspa, rspb);
            }
        }
    }
    hipLaunchKernelGGL(k_rs_final, dim3(nblk(2 * S)), dim3(256), 0, st_, mb_[M_IS11].as<uint8_t>(), rid, seg, label,
                       rwa, rspa, S, wa, spa);
    phase_end(9);
    // --- kept nodes
    phase_begin(10);
    MCHECK(compact_count_scan(S, kept, mb_[M_BCNT].as<uint32_t>(), mb_[M_BOFF].as<uint64_t>(),
                              scan_scratch_.as<uint64_t>(), d_small + 20, st_));
    MCHECK(mb_[M_SPEC].reserve(S * 4));
    hipLaunchKernelGGL(k_special_flags, dim3(nblk(S)), dim3(256), 0, st_, mb_[M_IS11].as<uint8_t>(), wa, spa, S,
                       mb_[M_SPEC].as<uint32_t>(), (unsigned long long *)(d_small + 21));
    MCHECK(hipMemcpyAsync(hs + 16, d_small + 16, 8 * 8, hipMemcpyDeviceToHost, st_));
    MCHECK(hipStreamSynchronize(st_));
    n11 = hs[16] - nconv;
    missing = hs[17];
    nk = hs[20];
    nspec = hs[21];
    if (missing) return fail(-5, "edge to a k-mer missing from the graph (edge symmetry broken)");
    if (pass > 0 || nspec == 0) break;
    // simple cycles among the special components become chains
    // (block counts of their own: M_BOFF holds the kept nodes' for the scatter below)
    const CyclePred cyc{mb_[M_IS11].as<uint8_t>(), wa};
    MCHECK(mb_[M_LEV].reserve((nbc + 1) * 16 + 64));
    uint64_t *cboff = mb_[M_LEV].as<uint64_t>();
    uint32_t *cbcnt = reinterpret_cast<uint32_t *>(cboff + nbc + 1);
    MCHECK(compact_count_scan(2 * S, cyc, cbcnt, cboff, scan_scratch_.as<uint64_t>(), d_small + 28, st_));
    MCHECK(hipMemcpyAsync(hs + 28, d_small + 28, 8, hipMemcpyDeviceToHost, st_));
    MCHECK(hipStreamSynchronize(st_));
    const uint64_t nc = hs[28];
    if (nc == 0) break;
    MCHECK(mb_[M_INL].reserve(2 * S * 4 + 16));             // candidate position of each oriented node
    MCHECK(mb_[M_INR].reserve(nc * (4 + 2 * 8 + 2 * 4 + 2) + 64));
    uint32_t *cpos = mb_[M_INL].as<uint32_t>();
    uint32_t *cl = mb_[M_INR].as<uint32_t>();
    uint64_t *ka = reinterpret_cast<uint64_t *>(mb_[M_INR].as<uint8_t>() + ((nc * 4 + 15) & ~15ull));
    uint64_t *kb = ka + nc;
    uint32_t *na = reinterpret_cast<uint32_t *>(kb + nc), *nb2 = na + nc;
    uint8_t *ba = reinterpret_cast<uint8_t *>(nb2 + nc), *bb = ba + nc;
    compact_scatter<true>(2 * S, cyc, cboff, cpos, cl, st_);
    MCHECK(hipStreamWaitEvent(st_, prio_done_, 0));  // the cycles' maxima read the priorities
    hipLaunchKernelGGL(k_cyc_init, dim3(nblk(nc)), dim3(256), 0, st_, cl, nc, cpos, mb_[M_NXT].as<uint32_t>(),
                       mb_[M_PRIO].as<uint64_t>(), mb_[M_SELF].as<uint8_t>(), ka, na, ba);
    int crounds = 1;
    while ((1ull << crounds) < nc) crounds++;
    for (int r = 0; r <= crounds; r++) {  // 2^(crounds+1) >= 2 nc successors: every cycle covered
        hipLaunchKernelGGL(k_cyc_jump, dim3(nblk(nc)), dim3(256), 0, st_, ka, na, ba, nc, kb, nb2, bb);
        std::swap(ka, kb);
        std::swap(na, nb2);
        std::swap(ba, bb);
    }
    MCHECK(hipMemsetAsync(d_small + 29, 0, 8, st_));
    hipLaunchKernelGGL(k_cyc_mark, dim3(nblk(nc)), dim3(256), 0, st_, cl, nc, cpos, ka, ba,
                       mb_[M_IS11].as<uint8_t>(), (unsigned long long *)(d_small + 29));
    MCHECK(hipMemcpyAsync(hs + 29, d_small + 29, 8, hipMemcpyDeviceToHost, st_));
    MCHECK(hipStreamSynchronize(st_));
    nconv = hs[29];
    if (nconv == 0) break;
    // the list ranking again, the counters it adds to cleared
    MCHECK(hipMemsetAsync(d_small + 20, 0, 2 * 8, st_));
    }
    out.n11 = n11;
    out.n_kept = nk;
    MCHECK(mb_[M_KEPTOF].reserve(S * 4));
    MCHECK(mb_[M_KEPTSRC].reserve(nk * 4 + 4));
    compact_scatter<false>(S, kept, mb_[M_BOFF].as<uint64_t>(), mb_[M_KEPTOF].as<uint32_t>(),
                           mb_[M_KEPTSRC].as<uint32_t>(), st_);
    // --- kept edges
    const uint64_t ne = 8 * nk;
    MCHECK(mb_[M_ELEN].reserve(ne * 4 + 4));
    MCHECK(mb_[M_EMULT].reserve(ne + 4));
    MCHECK(mb_[M_EOTHER].reserve(ne * 4 + 4));
    MCHECK(mb_[M_CSIZE].reserve(ne * 4 + 4));
    MCHECK(mb_[M_ESPEC].reserve(ne + 4));
    EdgeArgs ea;
    ea.key = ma.key;
    ea.mult = ma.mult;
    ea.is11 = mb_[M_IS11].as<uint8_t>();
    ea.walk = wa;
    ea.sp = spa;
    ea.kept_of = mb_[M_KEPTOF].as<uint32_t>();
    ea.kept_src = mb_[M_KEPTSRC].as<uint32_t>();
    ea.n_kept = nk;
    ea.ma = ma;
    ea.elen = mb_[M_ELEN].as<uint32_t>();
    ea.emult = mb_[M_EMULT].as<uint8_t>();
    ea.eother = mb_[M_EOTHER].as<uint32_t>();
    ea.csize = mb_[M_CSIZE].as<uint32_t>();
    ea.especial = mb_[M_ESPEC].as<uint8_t>();
    ea.counters = cnt;
    if (ne) hipLaunchKernelGGL(k_kept_edges, dim3(nblk(ne)), dim3(256), 0, st_, ea);
    MCHECK(mb_[M_EOFF].reserve((ne + 1) * 8));
    MCHECK(mb_[M_COFF].reserve((ne + 1) * 8));
    MCHECK(scan_scratch_.reserve((scan_scratch_words(ne + 1) + 16) * 8));
    MCHECK(scan_excl<uint32_t>(mb_[M_ELEN].as<uint32_t>(), mb_[M_EOFF].as<uint64_t>(), ne, scan_scratch_.as<uint64_t>(),
                        d_small + 22, st_));
    MCHECK(scan_excl<uint32_t>(mb_[M_CSIZE].as<uint32_t>(), mb_[M_COFF].as<uint64_t>(), ne, scan_scratch_.as<uint64_t>(),
                        d_small + 23, st_));
    MCHECK(hipMemcpyAsync(hs + 22, d_small + 22, 2 * 8, hipMemcpyDeviceToHost, st_));
    MCHECK(hipStreamSynchronize(st_));
    phase_end(10);
    const uint64_t pool_n = ne ? hs[22] : 0, M = ne ? hs[23] : 0;
    // --- strings and chain arrays
    MCHECK(mb_[M_POOL].reserve(pool_n + 16));
    MCHECK(mb_[M_CPRI].reserve(M * 8 + 16));
    MCHECK(mb_[M_CMUL].reserve(M + 16));
    MCHECK(mb_[M_CEDGE].reserve(M * 4 + 16));
    FillArgs fa;
    fa.key = ma.key;
    fa.is11 = mb_[M_IS11].as<uint8_t>();
    fa.walk = wa;
    fa.sp = spa;
    fa.kept_of = mb_[M_KEPTOF].as<uint32_t>();
    fa.eoff = mb_[M_EOFF].as<uint64_t>();
    fa.elen = mb_[M_ELEN].as<uint32_t>();
    fa.coff = mb_[M_COFF].as<uint64_t>();
    fa.eother = mb_[M_EOTHER].as<uint32_t>();
    fa.ex = mb_[M_EX].as<uint8_t>();
    fa.prio = mb_[M_PRIO].as<uint64_t>();
    fa.S = S;
    fa.k = k_;
    fa.pool = mb_[M_POOL].as<char>();
    fa.cpri = mb_[M_CPRI].as<uint64_t>();
    fa.cmul = mb_[M_CMUL].as<uint8_t>();
    fa.cedge = mb_[M_CEDGE].as<uint32_t>();
    MCHECK(hipStreamWaitEvent(st_, prio_done_, 0));  // the fills read the visit priorities
    phase_begin(11);
    if (ne) {
        hipLaunchKernelGGL(k_fill_edges, dim3(nblk(ne)), dim3(256), 0, st_, fa, mb_[M_EMULT].as<uint8_t>(), ne);
        hipLaunchKernelGGL(k_fill_nodes, dim3(nblk(2 * S)), dim3(256), 0, st_, fa);
        hipLaunchKernelGGL(k_fill_last, dim3(nblk(ne)), dim3(256), 0, st_, mb_[M_ELEN].as<uint32_t>(),
                           mb_[M_EOTHER].as<uint32_t>(), mb_[M_KEPTSRC].as<uint32_t>(), ma.key,
                           mb_[M_EOFF].as<uint64_t>(), ne, k_, mb_[M_POOL].as<char>());
    }
    phase_end(11);
    // --- merged multiplicities: Cartesian trees over the chain arrays
    phase_begin(12);
    if (M) {
        std::vector<uint64_t> loff;
        uint64_t tot = 0, len = M;
        int nlev = 0;
        while (true) {
            loff.push_back(tot);
            tot += len;
            nlev++;
            if (len == 1) break;
            len = (len + 1) / 2;
        }
        MCHECK(mb_[M_LEV].reserve(tot * 8 + 16));
        uint64_t *lev = mb_[M_LEV].as<uint64_t>();
        if (nlev > 64) return fail(-4, "chain array too long for the max hierarchy");
        LevTab lt;
        len = M;
        for (int l = 0; l < nlev; l++) {
            lt.off[l] = loff[l];
            lt.len[l] = len;
            len = (len + 1) / 2;
        }
        for (int l0 = 0; l0 == 0 || l0 < nlev - 1; l0 += kMaxSteps) {
            const int l1 = std::min(nlev - 1, l0 + kMaxSteps);
            const uint64_t grid = (lt.len[l0] + kMaxTile - 1) / kMaxTile;
            hipLaunchKernelGGL(k_max_levels, dim3((unsigned)grid), dim3(kMaxThreads), 0, st_,
                               l0 == 0 ? (const uint64_t *)fa.cpri : (const uint64_t *)(lev + loff[l0]), lev, lt, l0,
                               l1);
        }
        MCHECK(mb_[M_PAR].reserve(M * 4));
        MCHECK(mb_[M_NEED].reserve(M));
        MCHECK(mb_[M_NL].reserve(M * 4));
        MCHECK(mb_[M_NR].reserve(M * 4));
        MCHECK(mb_[M_INL].reserve(M * 4));
        MCHECK(mb_[M_INR].reserve(M * 8));
        MCHECK(hipMemsetAsync(mb_[M_INR].p, 0, M * 8, st_));
        TreeArgs ta;
        ta.lev = lev;
        for (int l = 0; l < 64; l++) ta.lev_off[l] = l < nlev ? loff[l] : 0;
        ta.nlev = nlev;
        ta.M = M;
        ta.cmul = fa.cmul;
        ta.cedge = fa.cedge;
        ta.par = mb_[M_PAR].as<uint32_t>();
        ta.need = mb_[M_NEED].as<uint8_t>();
        ta.nl = mb_[M_NL].as<uint32_t>();
        ta.nr = mb_[M_NR].as<uint32_t>();
        ta.far = mb_[M_INL].as<uint32_t>();
        ta.nfar = (unsigned int *)(d_small + 27);  // zeroed with d_small[16..31] at the merge's start
        ta.err = (unsigned int *)(d_small + 25);
        const unsigned tiles = (unsigned)((M + kTreeTile - 1) / kTreeTile);
        hipLaunchKernelGGL(k_tree_links, dim3(tiles), dim3(kTreeLinkThreads), 0, st_, ta);
        hipLaunchKernelGGL(k_tree_far, dim3(std::min(tiles, 1024u)), dim3(256), 0, st_, ta);
        hipLaunchKernelGGL(k_tree_eval, dim3(tiles), dim3(kTreeEvalThreads), 0, st_, ta,
                           (unsigned long long *)mb_[M_INR].p, mb_[M_EMULT].as<uint8_t>());
    }
    phase_end(12);
    MCHECK(hipGetLastError());
    MCHECK(hipMemcpyAsync(hs + 25, d_small + 25, 8, hipMemcpyDeviceToHost, st_));
    MCHECK(hipStreamSynchronize(st_));
    const uint32_t merr = (uint32_t)hs[25];
    if (merr & 1u) return fail(-5, "merge: a chain array lost its end marker");
    if (merr & 2u) return fail(-5, "merge: a ruler walk passed every oriented node without ending");
    if (merr & 4u) return fail(-5, "merge: a Cartesian-tree climb passed more parents than chain positions");
    collect_phases();
    n_kept = nk;
    n_special = nspec;
    merged_ref_ = reference_order;
    merged_ne_ = ne;
    merged_pool_ = pool_n;
    merged_M_ = M;
    return 0;
}

// Copies the merge results of the last merge_device to the host.
// The merge's per-edge and per-kept-node results gathered into one buffer, so
// the host reads them with one copy: segment y of the grid, 8-byte aligned
// offsets in a.off.
struct PackArgs {
    const uint64_t *eoff, *prio, *key;
    const uint32_t *kept_src, *elen;
    const uint8_t *emult, *espec;
    uint64_t ne, nk;
    uint8_t *dst;
    uint64_t off[7];
};

__global__ void k_pack_merge(PackArgs a) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint8_t *d = a.dst + a.off[blockIdx.y];
    switch (blockIdx.y) {
        case 0: if (i < a.ne) reinterpret_cast<uint64_t *>(d)[i] = a.eoff[i]; break;
        case 1: if (i < a.nk) reinterpret_cast<uint64_t *>(d)[i] = a.prio[a.kept_src[i]]; break;
        case 2: if (i < a.nk) reinterpret_cast<uint64_t *>(d)[i] = a.key[a.kept_src[i]]; break;
        case 3: if (i < a.nk) reinterpret_cast<uint32_t *>(d)[i] = a.kept_src[i]; break;
        case 4: if (i < a.ne) reinterpret_cast<uint32_t *>(d)[i] = a.elen[i]; break;
        case 5: if (i < a.ne) d[i] = a.emult[i]; break;
        case 6: if (i < a.ne) d[i] = a.espec[i]; break;
        default: break;
    }
}

// What the host replay of cycles / fold-backs needs beyond the kept nodes:
// the special nodes (key, multiplicities, priority), the kept nodes'
// multiplicities, and a bitmap of the (1,1) nodes by visit rank.
__global__ void k_special_gather(const uint32_t *__restrict__ special, uint64_t nspec,
                                 const uint32_t *__restrict__ kept_src, uint64_t nk,
                                 const uint64_t *__restrict__ key, const uint64_t *__restrict__ mult,
                                 const uint64_t *__restrict__ prio, uint64_t *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nspec) {
        const uint32_t j = special[i];
        out[i] = key[j];
        out[nspec + i] = mult[j];
        out[2 * nspec + i] = prio[j];
    }
    if (i < nk) out[3 * nspec + i] = mult[kept_src[i]];
}

__global__ void k_one_bits(const uint8_t *__restrict__ is11, const uint64_t *__restrict__ prio, uint64_t S,
                           uint32_t *__restrict__ bits) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < S && is11[i]) {
        const uint64_t r = prio[i];  // a rank < S
        atomicOr(&bits[r >> 5], 1u << (r & 31));
    }
}

int Pipeline::fetch_merge(MergeDev &out) {
    MCHECK(hipSetDevice(dev_));
    out = MergeDev();
    const uint64_t S = n_surv, nk = n_kept, ne = merged_ne_, pool_n = merged_pool_, nspec = n_special;
    out.n_kept = nk;
    out.kept_src.resize(nk);
    out.elen.resize(ne);
    out.eoff.resize(ne);
    out.emult.resize(ne);
    out.especial.resize(ne);
    out.pool.resize(pool_n);
    if (pool_n) MCHECK(hipMemcpyAsync(out.pool.data(), mb_[M_POOL].p, pool_n, hipMemcpyDeviceToHost, st_));
    out.kept_prio.resize(nk);
    out.kept_key.resize(nk);
    std::vector<uint8_t> packed;
    PackArgs pa;
    if (nk || ne) {
        const uint64_t sz[7] = {ne * 8, nk * 8, nk * 8, nk * 4, ne * 4, ne, ne};
        uint64_t tot = 0;
        for (int j = 0; j < 7; j++) {
            pa.off[j] = tot;
            tot += (sz[j] + 7) & ~7ull;
        }
        MCHECK(mb_[M_INL].reserve(tot + 16));
        pa.eoff = mb_[M_EOFF].as<uint64_t>();
        pa.prio = mb_[M_PRIO].as<uint64_t>();
        pa.key = surv_key_.as<uint64_t>();
        pa.kept_src = mb_[M_KEPTSRC].as<uint32_t>();
        pa.elen = mb_[M_ELEN].as<uint32_t>();
        pa.emult = mb_[M_EMULT].as<uint8_t>();
        pa.espec = mb_[M_ESPEC].as<uint8_t>();
        pa.ne = ne;
        pa.nk = nk;
        pa.dst = mb_[M_INL].as<uint8_t>();
        hipLaunchKernelGGL(k_pack_merge, dim3(nblk(std::max(ne, nk)), 7), dim3(256), 0, st_, pa);
        packed.resize(tot);
        MCHECK(hipMemcpyAsync(packed.data(), pa.dst, tot, hipMemcpyDeviceToHost, st_));
    }
    std::vector<uint64_t> sg;
    if (nspec) {
        out.special.resize(nspec);
        MCHECK(hipMemcpyAsync(out.special.data(), mb_[M_SPEC].p, nspec * 4, hipMemcpyDeviceToHost, st_));
        if (merged_ref_) {
            // O(special + kept) values and S bits instead of three S-entry arrays
            const uint64_t words = (S + 31) / 32;
            MCHECK(mb_[M_PAR].reserve((3 * nspec + nk) * 8 + 16));
            MCHECK(mb_[M_NEED].reserve(words * 4 + 16));
            uint64_t *g = mb_[M_PAR].as<uint64_t>();
            uint32_t *bits = mb_[M_NEED].as<uint32_t>();
            MCHECK(hipMemsetAsync(bits, 0, words * 4, st_));
            hipLaunchKernelGGL(k_special_gather, dim3(nblk(std::max(nspec, nk))), dim3(256), 0, st_,
                               mb_[M_SPEC].as<uint32_t>(), nspec, mb_[M_KEPTSRC].as<uint32_t>(), nk,
                               surv_key_.as<uint64_t>(), surv_mult_.as<uint64_t>(), mb_[M_PRIO].as<uint64_t>(), g);
            hipLaunchKernelGGL(k_one_bits, dim3(nblk(S)), dim3(256), 0, st_, mb_[M_IS11].as<uint8_t>(),
                               mb_[M_PRIO].as<uint64_t>(), S, bits);
            sg.resize(3 * nspec + nk);
            out.one_bits.resize(words);
            MCHECK(hipMemcpyAsync(sg.data(), g, sg.size() * 8, hipMemcpyDeviceToHost, st_));
            MCHECK(hipMemcpyAsync(out.one_bits.data(), bits, words * 4, hipMemcpyDeviceToHost, st_));
        } else {
            out.all_key.resize(S);
            out.all_mult.resize(S);
            out.all_prio.resize(S);
            MCHECK(hipMemcpyAsync(out.all_key.data(), surv_key_.p, S * 8, hipMemcpyDeviceToHost, st_));
            MCHECK(hipMemcpyAsync(out.all_mult.data(), surv_mult_.p, S * 8, hipMemcpyDeviceToHost, st_));
            MCHECK(hipMemcpyAsync(out.all_prio.data(), mb_[M_PRIO].p, S * 8, hipMemcpyDeviceToHost, st_));
        }
    }
    MCHECK(hipStreamSynchronize(st_));
    if (!sg.empty()) {
        out.spec_key.assign(sg.begin(), sg.begin() + nspec);
        out.spec_mult.assign(sg.begin() + nspec, sg.begin() + 2 * nspec);
        out.spec_prio.assign(sg.begin() + 2 * nspec, sg.begin() + 3 * nspec);
        out.kept_mult.assign(sg.begin() + 3 * nspec, sg.end());
    }
    if (!packed.empty()) {
        const uint8_t *b = packed.data();
        if (ne) {
            std::memcpy(out.eoff.data(), b + pa.off[0], ne * 8);
            std::memcpy(out.elen.data(), b + pa.off[4], ne * 4);
            std::memcpy(out.emult.data(), b + pa.off[5], ne);
            std::memcpy(out.especial.data(), b + pa.off[6], ne);
        }
        if (nk) {
            std::memcpy(out.kept_prio.data(), b + pa.off[1], nk * 8);
            std::memcpy(out.kept_key.data(), b + pa.off[2], nk * 8);
            std::memcpy(out.kept_src.data(), b + pa.off[3], nk * 4);
        }
    }
    return 0;
}

}  // namespace swapk
