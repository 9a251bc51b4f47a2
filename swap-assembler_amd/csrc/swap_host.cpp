// Host side: the reference's hash-table orders (replayed, not rebuilt), the
// byte-exact writers, and the merge driver around the device kernels.
#include "swap_host.h"

#include <algorithm>
#include <cstring>
#include <functional>
#include <numeric>

namespace swapk {

namespace {

// Bucket counts of libstdc++'s backward hashtable (hashtable.h:211-219): a
// table starts at the first listed prime >= 100 and grows to the first listed
// prime >= n+1 when the (n+1)-th element arrives at a full table.
const uint64_t kPrimes[] = {5ul,         53ul,        97ul,         193ul,        389ul,       769ul,
                            1543ul,      3079ul,      6151ul,       12289ul,      24593ul,     49157ul,
                            98317ul,     196613ul,    393241ul,     786433ul,     1572869ul,   3145739ul,
                            6291469ul,   12582917ul,  25165843ul,   50331653ul,   100663319ul, 201326611ul,
                            402653189ul, 805306457ul, 1610612741ul, 3221225473ul, 4294967291ul};
const int kNumPrimes = sizeof(kPrimes) / sizeof(kPrimes[0]);

uint64_t next_prime(uint64_t n) {
    const uint64_t *p = std::lower_bound(kPrimes, kPrimes + kNumPrimes, n);
    return p == kPrimes + kNumPrimes ? kPrimes[kNumPrimes - 1] : *p;
}

struct Epochs {
    std::vector<uint64_t> start;  // first insert rank of each epoch
    std::vector<uint64_t> P;      // bucket count during the epoch
    int of(uint64_t r) const { return (int)(std::upper_bound(start.begin(), start.end(), r) - start.begin()) - 1; }
};

Epochs make_epochs(uint64_t n_total, bool extra_resize) {
    Epochs e;
    uint64_t P = next_prime(100);
    e.start.push_back(0);
    e.P.push_back(P);
    // insert of rank r first calls resize(r + 1): growth when r + 1 > P, i.e. at r == P
    while (P < n_total) {
        const uint64_t np = next_prime(P + 1);
        if (np <= P) break;
        e.start.push_back(P);
        e.P.push_back(np);
        P = np;
    }
    if (extra_resize && n_total == P) {
        const uint64_t np = next_prime(P + 1);
        if (np > P) {
            e.start.push_back(n_total);
            e.P.push_back(np);
        }
    }
    return e;
}

// LSD radix sort of (bucket, index) by bucket, stable.
void sort_by_bucket(std::vector<uint32_t> &bucket, std::vector<uint32_t> &idx) {
    const size_t n = idx.size();
    std::vector<uint32_t> b2(n), i2(n);
    for (int pass = 0; pass < 2; pass++) {
        const int sh = 16 * pass;
        std::vector<size_t> cnt(65537, 0);
        for (size_t i = 0; i < n; i++) cnt[((bucket[i] >> sh) & 0xFFFF) + 1]++;
        for (int d = 0; d < 65536; d++) cnt[d + 1] += cnt[d];
        for (size_t i = 0; i < n; i++) {
            const size_t pos = cnt[(bucket[i] >> sh) & 0xFFFF]++;
            b2[pos] = bucket[i];
            i2[pos] = idx[i];
        }
        bucket.swap(b2);
        idx.swap(i2);
    }
}

}  // namespace

uint64_t table_buckets(uint64_t n_total, bool extra_resize) { return make_epochs(n_total, extra_resize).P.back(); }

std::vector<uint32_t> table_order(const std::vector<uint64_t> &key, const std::vector<uint64_t> &rank,
                                  uint64_t n_total, bool extra_resize, uint64_t *bucket_count) {
    const Epochs ep = make_epochs(n_total, extra_resize);
    const int f = (int)ep.P.size() - 1;
    const size_t n = key.size();
    std::vector<uint64_t> h(n);
    for (size_t i = 0; i < n; i++) h[i] = ref_hash(key[i]);
    // a precedes b in the iteration order at the end of epoch e
    std::function<bool(uint32_t, uint32_t, int)> before = [&](uint32_t a, uint32_t b, int e) -> bool {
        const uint64_t ba = h[a] % ep.P[e], bb = h[b] % ep.P[e];
        if (ba != bb) return ba < bb;
        const bool ia = ep.of(rank[a]) == e, ib = ep.of(rank[b]) == e;
        if (ia && ib) return rank[a] > rank[b];  // head insertion: newest first
        if (ia != ib) return ia;                 // inserted after the rehash: in front
        return before(b, a, e - 1);              // rehash reverses each chain
    };
    std::vector<uint32_t> bucket(n), idx(n);
    for (size_t i = 0; i < n; i++) {
        bucket[i] = (uint32_t)(h[i] % ep.P[f]);
        idx[i] = (uint32_t)i;
    }
    sort_by_bucket(bucket, idx);
    for (size_t i = 0; i < n;) {
        size_t j = i + 1;
        while (j < n && bucket[j] == bucket[i]) j++;
        if (j - i > 1) std::sort(idx.begin() + i, idx.begin() + j, [&](uint32_t a, uint32_t b) { return before(a, b, f); });
        i = j;
    }
    if (bucket_count) *bucket_count = ep.P[f];
    return idx;
}

std::vector<uint32_t> visit_order(const std::vector<uint64_t> &key, const std::vector<uint64_t> &rank,
                                  uint64_t distinct, bool tail_lookup, uint64_t *node_buckets) {
    const size_t S = key.size();
    // kmolecules: every distinct k-mer inserted at its first occurrence
    // (graph.cpp:1058-1070)
    std::vector<uint32_t> o1 = table_order(key, rank, distinct, tail_lookup, nullptr);
    // cutoffGraph: survivors inserted into a fresh table in kmolecules order,
    // then copy-assigned back with the same buckets (graph.cpp:1096-1112)
    std::vector<uint64_t> r2(S);
    for (size_t i = 0; i < S; i++) r2[o1[i]] = i;
    std::vector<uint32_t> o2 = table_order(key, r2, S, false, nullptr);
    // buildKmoleculeGraph: nodes inserted in kmolecules order (graph.cpp:1123-1133)
    std::vector<uint64_t> r3(S);
    for (size_t i = 0; i < S; i++) r3[o2[i]] = i;
    return table_order(key, r3, S, false, node_buckets);
}

std::vector<uint32_t> gather_order(const std::vector<uint64_t> &key, const std::vector<uint32_t> &visit_pos,
                                   uint64_t node_buckets) {
    // cleared table keeps its buckets; re-insertion in visit order puts the
    // later-inserted node first within a bucket
    const size_t n = key.size();
    std::vector<uint32_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0u);
    std::vector<uint64_t> b(n);
    for (size_t i = 0; i < n; i++) b[i] = ref_hash(key[i]) % node_buckets;
    std::sort(idx.begin(), idx.end(), [&](uint32_t x, uint32_t y) {
        if (b[x] != b[y]) return b[x] < b[y];
        return visit_pos[x] > visit_pos[y];
    });
    return idx;
}

// ---------------------------------------------------------------------------
// writers

namespace {

inline void put_u64(std::string &out, uint64_t v) {
    char t[24];
    int n = 0;
    do { t[n++] = (char)('0' + v % 10); v /= 10; } while (v);
    while (n) out.push_back(t[--n]);
}

const char kACGT[4] = {'A', 'C', 'G', 'T'};

inline void put_kmer(std::string &out, uint64_t x, int k) {
    for (int i = k - 1; i >= 0; i--) out.push_back(kACGT[(x >> (2 * i)) & 3]);
}

inline uint32_t code_of(char ch) {
    switch (ch) { case 'C': return 1; case 'G': return 2; case 'T': return 3; default: return 0; }
}

// canonical terminal of orient(X, slot) + s: last k bases, max with its rc.
// Returns +1 if the walk ends on the forward strand of the result.
inline int terminal(uint64_t X, int slot, const char *s, size_t L, int k, uint64_t *id) {
    const uint64_t km = kmask(k);
    uint64_t t;
    if (L >= (size_t)k) {
        t = 0;
        for (size_t j = L - k; j < L; j++) t = (t << 2) | code_of(s[j]);
    } else {
        t = slot >= 4 ? revcomp(X, k) : X;
        for (size_t j = 0; j < L; j++) t = ((t << 2) | code_of(s[j])) & km;
    }
    const uint64_t r = revcomp(t, k);
    if (t < r) { *id = r; return -1; }
    *id = t;
    return 1;
}

}  // namespace

// printKmerGraph (graph.cpp:3352-3440).  With one-base edges the neighbour of
// slot i is the canonical form Y of (orient(X) << 2 | i%4); when the walk ends
// on Y's reverse strand the reference writes the value of reverse(str(Y)) with
// only character i complemented (its inner loop indexes with the slot, 3392-3393).
void format_kmer_graph(int k, const HostKmerGraph &g, const std::vector<uint32_t> &order, std::string &out) {
    const uint64_t km = kmask(k);
    const size_t n = g.key.size();
    out.reserve(out.size() + n * 64);
    for (size_t oi = 0; oi < n; oi++) {
        const size_t i = order.empty() ? oi : order[oi];
        const uint64_t X = g.key[i];
        const uint64_t m = g.mult[i];
        for (int half = 0; half < 2; half++) {
            const uint64_t o = half ? revcomp(X, k) : X;
            bool any = false;
            for (int s = 4 * half; s < 4 * half + 4; s++) {
                if (!((m >> (8 * s)) & 0xFF)) continue;
                if (!any) { put_u64(out, o); any = true; }
                const uint64_t t = ((o << 2) | (uint64_t)(s & 3)) & km;
                const uint64_t r = revcomp(t, k);
                uint64_t v;
                if (t >= r) v = t;
                else v = (revcomp(r, k) ^ km) ^ (3ull << (2 * (k - 1 - s)));
                out.push_back(' ');
                put_u64(out, v);
            }
            if (any) out.push_back('\n');
        }
    }
}

// gatherDistNodeGraph line format (graph.cpp:592-594, 641-644) and the final
// fprintf("%s\n") of graph.cpp:3617-3622.
void format_contig_graph(const HostCompacted &c, std::string &out) {
    const size_t n = c.key.size();
    for (size_t i = 0; i < n; i++) {
        const size_t start = out.size();
        size_t total = 61;
        put_u64(out, c.key[i]);
        out.push_back(' ');
        for (int s = 0; s < 8; s++) {
            const uint32_t L = c.len[8 * i + s];
            out.append(c.pool.data() + c.off[8 * i + s], L);
            out.push_back('#');
            if (s < 7) out.push_back(' ');
            total += L + 1;
        }
        for (int s = 0; s < 8; s++) {
            out.push_back(' ');
            put_u64(out, c.mult[8 * i + s]);
        }
        const size_t len = out.size() - start;
        if (len < total) out.append(total - len, ' ');
        out.resize(start + total);
        out[start + total - 1] = '\n';
    }
    out.push_back('\n');
}

// printContigs (graph.cpp:1139-1202): one header-only line per edge of at least
// Contig_Length - k bases whose start node is not smaller than its end.
void format_noce(int k, const HostCompacted &c, std::string &out) {
    const size_t n = c.key.size();
    for (size_t i = 0; i < n; i++) {
        const uint64_t X = c.key[i];
        for (int s = 0; s < 8; s++) {
            const uint32_t L = c.len[8 * i + s];
            if (L == 0) continue;
            if ((long)L < 100 - k) continue;
            const char *str = c.pool.data() + c.off[8 * i + s];
            uint64_t id;
            terminal(X, s, str, L, k, &id);
            if (X < id) continue;
            out += ">proc0: ";
            put_kmer(out, s >= 4 ? revcomp(X, k) : X, k);
            out.append(str, L);
            out.push_back('\n');
        }
    }
}

}  // namespace swapk
