// Host side: the reference's hash-table orders (replayed, not rebuilt), the
// byte-exact writers, and the merge driver around the device kernels.
#include "swap_host.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
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

// LSD radix sort of (bucket, index) by bucket, stable: four 8-bit passes over
// packed (bucket << 32 | index) words.
void sort_by_bucket(std::vector<uint32_t> &bucket, std::vector<uint32_t> &idx) {
    const size_t n = idx.size();
    std::vector<uint64_t> a(n), b(n);
    for (size_t i = 0; i < n; i++) a[i] = ((uint64_t)bucket[i] << 32) | idx[i];
    for (int pass = 0; pass < 4; pass++) {
        const int sh = 32 + 8 * pass;
        size_t cnt[257] = {0};
        for (size_t i = 0; i < n; i++) cnt[((a[i] >> sh) & 0xFF) + 1]++;
        for (int d = 0; d < 256; d++) cnt[d + 1] += cnt[d];
        for (size_t i = 0; i < n; i++) b[cnt[(a[i] >> sh) & 0xFF]++] = a[i];
        a.swap(b);
    }
    for (size_t i = 0; i < n; i++) {
        bucket[i] = (uint32_t)(a[i] >> 32);
        idx[i] = (uint32_t)a[i];
    }
}

}  // namespace

uint64_t table_buckets(uint64_t n_total, bool extra_resize) { return make_epochs(n_total, extra_resize).P.back(); }

std::vector<uint64_t> table_epoch_starts(uint64_t n_total, bool extra_resize) {
    return make_epochs(n_total, extra_resize).start;
}

namespace {

std::vector<uint32_t> order_with_epochs(const std::vector<uint64_t> &key, const std::vector<uint64_t> &rank,
                                        const Epochs &ep) {
    const int f = (int)ep.P.size() - 1;
    const size_t n = key.size();
    std::vector<uint64_t> h(n);
    std::vector<int16_t> ins(n);  // insertion epoch of each element
    for (size_t i = 0; i < n; i++) {
        h[i] = ref_hash(key[i]);
        ins[i] = (int16_t)ep.of(rank[i]);
    }
    // a precedes b in the iteration order at the end of epoch e.  Within a
    // bucket: elements inserted during e (newest first, head insertion), then
    // the ones carried over by the rehash that opened e, in reverse of their
    // order at the end of e-1.
    auto before = [&](uint32_t a, uint32_t b, int e) -> bool {
        bool flip = false;
        for (;; e--) {
            const uint64_t ba = h[a] % ep.P[e], bb = h[b] % ep.P[e];
            if (ba != bb) return (ba < bb) != flip;
            const bool ia = ins[a] == e, ib = ins[b] == e;
            if (ia && ib) return (rank[a] > rank[b]) != flip;
            if (ia != ib) return ia != flip;
            flip = !flip;
        }
    };
    std::vector<uint32_t> bucket(n), idx(n);
#ifdef SWAP_TIME_ORDER
    auto T0 = std::chrono::steady_clock::now();
#endif
    for (size_t i = 0; i < n; i++) {
        bucket[i] = (uint32_t)(h[i] % ep.P[f]);
        idx[i] = (uint32_t)i;
    }
    sort_by_bucket(bucket, idx);
#ifdef SWAP_TIME_ORDER
    auto T1 = std::chrono::steady_clock::now();
    fprintf(stderr, "sort %.3f\n", std::chrono::duration<double>(T1 - T0).count());
#endif
    for (size_t i = 0; i < n;) {
        size_t j = i + 1;
        while (j < n && bucket[j] == bucket[i]) j++;
        // insertion sort: buckets hold a handful of elements
        for (size_t x = i + 1; x < j; x++) {
            const uint32_t v = idx[x];
            size_t y = x;
            while (y > i && before(v, idx[y - 1], f)) { idx[y] = idx[y - 1]; y--; }
            idx[y] = v;
        }
        i = j;
    }
#ifdef SWAP_TIME_ORDER
    fprintf(stderr, "ties %.3f\n", std::chrono::duration<double>(std::chrono::steady_clock::now() - T1).count());
#endif
    return idx;
}

}  // namespace

std::vector<uint32_t> table_order(const std::vector<uint64_t> &key, const std::vector<uint64_t> &rank,
                                  uint64_t n_total, bool extra_resize, uint64_t *bucket_count) {
    const Epochs ep = make_epochs(n_total, extra_resize);
    if (bucket_count) *bucket_count = ep.P.back();
    return order_with_epochs(key, rank, ep);
}

std::vector<uint32_t> visit_order(const std::vector<uint64_t> &key, const std::vector<uint64_t> &first,
                                  const std::vector<uint64_t> &epoch_first, uint64_t distinct, bool tail_lookup,
                                  uint64_t *node_buckets) {
    const size_t S = key.size();
    // kmolecules: every distinct k-mer inserted at its first occurrence
    // (graph.cpp:1058-1070).  Insertion order among survivors is the order of
    // their first occurrences; epoch e starts at the first occurrence of the
    // distinct key whose insert grew the table.
    Epochs e1 = make_epochs(distinct, tail_lookup);
    for (size_t e = 0; e < e1.start.size() && e < epoch_first.size(); e++) e1.start[e] = epoch_first[e];
    if (e1.start.size() > epoch_first.size())  // growth on a lookup after the last insert
        for (size_t e = epoch_first.size(); e < e1.start.size(); e++) e1.start[e] = ~0ull;
    std::vector<uint32_t> o1 = order_with_epochs(key, first, e1);
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

namespace {
inline uint32_t code_of_base(char ch) {
    switch (ch) { case 'C': return 1; case 'G': return 2; case 'T': return 3; default: return 0; }
}
}  // namespace

// canonical terminal of orient(X, slot) + s: last k bases, max with its rc.
// Returns +1 if the walk ends on the forward strand of the result.
int terminal(uint64_t X, int slot, const char *s, size_t L, int k, uint64_t *id) {
    const uint64_t km = kmask(k);
    uint64_t t;
    if (L >= (size_t)k) {
        t = 0;
        for (size_t j = L - k; j < L; j++) t = (t << 2) | code_of_base(s[j]);
    } else {
        t = slot >= 4 ? revcomp(X, k) : X;
        for (size_t j = 0; j < L; j++) t = ((t << 2) | code_of_base(s[j])) & km;
    }
    const uint64_t r = revcomp(t, k);
    if (t < r) { *id = r; return -1; }
    *id = t;
    return 1;
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
    for (size_t oi = 0; oi < n; oi++) {
        const size_t i = c.noce_order.empty() ? oi : c.noce_order[oi];
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

// The -p contigs file (stringContigs, graph.cpp:3166-3238): the same contig
// selection as printContigs, walked in the order of the merge's own table (the
// contigGraph.txt order), written as ">contig_NodeID<X>_Index<slot>_Len<len>
// _Mul<merged multiplicity>_<running count>" and the sequence on its own line.
// Contig extension does not run in that mode, so no removed-edge records follow.
void format_perf_contigs(int k, const HostCompacted &c, std::string &out) {
    const size_t n = c.key.size();
    int count = 0;
    char head[160];
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
            count++;
            std::snprintf(head, sizeof head, ">contig_NodeID%llu_Index%d_Len%u_Mul%d_%d\n", (unsigned long long)X, s,
                          (unsigned)(k + L), (int)c.mult[8 * i + s], count);
            out += head;
            put_kmer(out, s >= 4 ? revcomp(X, k) : X, k);
            out.append(str, L);
            out.push_back('\n');
        }
    }
}

}  // namespace swapk

// ---------------------------------------------------------------------------
// merge driver

namespace swapk {

namespace {

// A node of a cycle / fold-back component (or a kept node next to one), with
// the reference's string edges (graph.cpp:42-49).
constexpr uint32_t kNoNode = 0xFFFFFFFFu;

// A replayed node.  Each slot's edge string is a rope (the splices only ever
// append one string to another), with its length and its last 31 bases
// packed two bits each: all the SWAP pass reads of a string are its length
// (the splice arithmetic) and its terminal k-mer (node::getNodeID), so the
// characters are only laid out for the output.
struct HNode {
    uint64_t id = 0;
    uint32_t rope[8];
    uint64_t len[8] = {0};
    uint64_t tail[8] = {0};
    uint8_t m[8] = {0};
    bool live = true;
    bool one = false;  // exactly one left and one right edge
    uint64_t prio = 0;
    int64_t kept = -1;  // kept index, or -1 for a (1,1) node of the component
    // the node at the end of each slot's string and the direction it is
    // reached in (node::getNodeID), kept current as the strings grow
    uint64_t nb[8] = {0};
    int8_t nbs[8] = {0};
    uint32_t nbi[8];  // index of that node among the replay's nodes (kNoNode: none)
};

struct Rope {
    uint32_t left, right;  // concatenation of two ropes, or a leaf (left == kNoNode)
    char base;
};

struct Replay {
    int k;
    uint64_t km;
    std::vector<HNode> nodes;
    std::unordered_map<uint64_t, uint32_t> index;  // key -> position in nodes
    std::vector<Rope> ropes;

    uint64_t orient(uint64_t X, int s) const { return s >= 4 ? revcomp(X, k) : X; }
    // node::getNodeID (graph.cpp:164-196) of slot s, after its string changed:
    // the last k bases of orient(X) followed by the string
    void refresh(HNode &n, int s) const {
        const uint64_t L = n.len[s];
        uint64_t t;
        if (L >= (uint64_t)k) t = n.tail[s] & km;
        else t = ((orient(n.id, s) << (2 * L)) | n.tail[s]) & km;
        const uint64_t r = revcomp(t, k);
        if (t < r) {
            n.nb[s] = r;
            n.nbs[s] = -1;
        } else {
            n.nb[s] = t;
            n.nbs[s] = 1;
        }
    }
    void refresh_all(HNode &n) const {
        for (int s = 0; s < 8; s++) refresh(n, s);
    }
    // the position of slot s's node (after refresh)
    void link(HNode &n, int s) const {
        auto it = index.find(n.nb[s]);
        n.nbi[s] = it == index.end() ? kNoNode : it->second;
    }
    void link_all(HNode &n, uint32_t self) const {
        for (int s = 0; s < 8; s++) {
            if (n.len[s]) link(n, s);
            else n.nbi[s] = self;  // an empty slot leads back to the node itself
        }
    }
    int neighbour(const HNode &n, int s, uint64_t *id) const {
        *id = n.nb[s];
        return n.nbs[s];
    }
    // node::getSrcNodeEdge, direction 2 (graph.cpp:263-324): every slot is
    // tested, an empty one included (it leads back to the node itself)
    int src_slot(const HNode &d, uint64_t src, int side) const {
        for (int i = 0; i < 8; i++)
            if (d.nb[i] == src && (d.nbs[i] == 1 ? 1 : 0) == side) return i;
        return -1;
    }
    uint32_t leaf(char base) {
        ropes.push_back(Rope{kNoNode, kNoNode, base});
        return (uint32_t)(ropes.size() - 1);
    }
    // string s of d += the string (rope, length, tail) of another slot
    void append(HNode &d, int s, uint32_t rope, uint64_t len, uint64_t tail) {
        if (!len) return;
        if (!d.len[s]) {
            d.rope[s] = rope;
        } else {
            ropes.push_back(Rope{d.rope[s], rope, 0});
            d.rope[s] = (uint32_t)(ropes.size() - 1);
        }
        d.tail[s] = len >= 31 ? tail : ((d.tail[s] << (2 * len)) | tail);
        d.tail[s] &= (1ull << 62) - 1;
        d.len[s] += len;
    }
    // the characters of a rope
    std::string text(uint32_t rope, uint64_t len) const {
        std::string out;
        out.reserve(len);
        if (!len) return out;
        std::vector<uint32_t> stack{rope};
        while (!stack.empty()) {
            const uint32_t r = stack.back();
            stack.pop_back();
            const Rope &x = ropes[r];
            if (x.left == kNoNode) {
                out.push_back(x.base);
            } else {
                stack.push_back(x.right);
                stack.push_back(x.left);
            }
        }
        return out;
    }
};

inline unsigned char splice_service(unsigned char ma, size_t la, unsigned char mb, size_t lb) {
    const size_t s = (size_t)ma * la + (size_t)mb * lb;
    const int pre = (int)(unsigned)s;
    return (unsigned char)((size_t)(long)pre / (la + lb));
}

inline unsigned char splice_master(unsigned char ma, size_t la, unsigned char mb, size_t lb) {
    return (unsigned char)(((size_t)ma * la + (size_t)mb * lb) / (la + lb));
}

}  // namespace

int host_merge(Pipeline &p, const swap_cfg &cfg, bool *materialized, HostCompacted &out) {
    const bool ref = !(cfg.flags & SWAP_FLAG_HASH_ORDER);
    MergeDev md;
    int rc = p.merge_device(ref, md);
    if (rc) return rc;
    *materialized = false;
    if (p.n_special == 0) return 0;  // results stay on the device until exported
    // cycles / fold-backs are replayed on the host now, as part of the merge
    rc = materialize_merge(p, cfg, out);
    if (rc) return rc;
    *materialized = true;
    return 0;
}

int materialize_merge(Pipeline &p, const swap_cfg &cfg, HostCompacted &out) {
    MergeDev md;
    int rc = p.fetch_merge(md);
    if (rc) return rc;
    uint64_t kept = 0;
    rc = replay_merge(md, p.n_surv, cfg, out, &kept);
    if (rc) return rc;
    p.n_kept = kept;
    return 0;
}

int replay_merge(const MergeDev &md, size_t S, const swap_cfg &cfg, HostCompacted &out, uint64_t *n_kept) {
    const bool ref = !(cfg.flags & SWAP_FLAG_HASH_ORDER);
    const int k = cfg.k;
    out = HostCompacted();
    const uint64_t P3 = ref ? table_buckets(S, false) : 0;
    const size_t nk = md.n_kept;
    // per kept node: 8 (string, multiplicity) taken from the device results
    struct Out {
        uint64_t id;
        uint64_t prio;
        std::string arcs[8];
        uint8_t m[8];
    };
    std::vector<Out> live, finished_live;
    live.reserve(nk + md.special.size());
    std::unordered_map<uint64_t, size_t> live_of;  // key -> index in live (only kept nodes next to specials)
    for (size_t i = 0; i < nk; i++) {
        Out o;
        o.id = md.kept_key[i];
        o.prio = md.kept_prio[i];
        bool any = false;
        for (int s = 0; s < 8; s++) {
            const uint32_t L = md.elen[8 * i + s];
            o.m[s] = L ? md.emult[8 * i + s] : 0;
            if (L) { o.arcs[s].assign(md.pool.data() + md.eoff[8 * i + s], L); any = true; }
        }
        (void)any;
        live.push_back(std::move(o));
    }
    // ---- cycles and fold-backs: the reference's sequential semantics, replayed
    if (!md.special.empty()) {
        Replay R;
        R.k = k;
        R.km = kmask(k);
        // reference order: the device sent the special nodes, the kept nodes'
        // multiplicities and a bitmap of the (1,1) visit ranks; hash order: all
        // survivors (its priorities are key hashes, not ranks)
        const bool ranks = md.all_key.empty();
        const std::vector<uint64_t> &gkey = md.all_key, &gmult = md.all_mult, &prio = md.all_prio;
        std::unordered_map<uint64_t, uint32_t> sidx_of;
        if (!ranks) {
            sidx_of.reserve(S * 2);
            for (size_t i = 0; i < S; i++) sidx_of[gkey[i]] = (uint32_t)i;
        }
        // the component's nodes: (1,1) nodes first, sorted by visit priority so
        // the SWAP pass walks them in memory order, then their kept neighbours;
        // a key's node is found through idx_of
        std::vector<HNode> &nodes = R.nodes;
        std::unordered_map<uint64_t, uint32_t> &idx_of = R.index;
        auto make = [&](uint64_t key, uint64_t mult, uint64_t pr, int64_t keptIdx) {
            HNode n;
            n.id = key;
            n.prio = pr;
            n.kept = keptIdx;
            int nr = 0, nl = 0;
            for (int s = 0; s < 8; s++) {
                n.m[s] = (uint8_t)(mult >> (8 * s));
                n.rope[s] = kNoNode;
                if (n.m[s]) {
                    n.rope[s] = R.leaf("ACGT"[s & 3]);
                    n.len[s] = 1;
                    n.tail[s] = (uint64_t)(s & 3);
                    (s < 4 ? nr : nl)++;
                }
            }
            n.one = nr == 1 && nl == 1;
            R.refresh_all(n);
            return n;
        };
        const size_t nspec = md.special.size();
        std::vector<std::pair<uint64_t, uint32_t>> pj(nspec);  // (priority, special index)
        for (size_t j = 0; j < nspec; j++) pj[j] = {ranks ? md.spec_prio[j] : prio[md.special[j]], (uint32_t)j};
        std::sort(pj.begin(), pj.end());
        std::vector<uint32_t> by_prio(nspec);
        for (size_t j = 0; j < nspec; j++) by_prio[j] = pj[j].second;
        nodes.reserve(nspec + 64);  // (kept neighbours: a few per component)
        idx_of.reserve(2 * nspec + 128);
        R.ropes.reserve(6 * nspec + 64);
        for (uint32_t j : by_prio) {
            HNode n = ranks ? make(md.spec_key[j], md.spec_mult[j], md.spec_prio[j], -1)
                            : make(gkey[md.special[j]], gmult[md.special[j]], prio[md.special[j]], -1);
            if (idx_of.emplace(n.id, (uint32_t)nodes.size()).second) nodes.push_back(std::move(n));
        }
        std::unordered_map<uint64_t, int64_t> kept_index;
        for (size_t i = 0; i < nk; i++) kept_index[md.kept_key[i]] = (int64_t)i;
        // kept neighbours of the component
        const size_t n_one = nodes.size();
        for (size_t q = 0; q < n_one; q++) {
            for (int s = 0; s < 8; s++) {
                if (!nodes[q].len[s]) continue;
                uint64_t id;
                R.neighbour(nodes[q], s, &id);
                if (idx_of.count(id)) continue;
                auto it = kept_index.find(id);
                if (it == kept_index.end()) continue;  // cannot happen: neighbours are survivors
                const size_t ki = (size_t)it->second;
                HNode n;
                if (ranks) {
                    n = make(id, md.kept_mult[ki], md.kept_prio[ki], it->second);
                } else {
                    const uint32_t sidx = sidx_of[id];
                    n = make(id, gmult[sidx], prio[sidx], it->second);
                }
                idx_of.emplace(id, (uint32_t)nodes.size());
                nodes.push_back(std::move(n));
            }
        }
        for (size_t q = 0; q < nodes.size(); q++) R.link_all(nodes[q], (uint32_t)q);
        // all (1,1) priorities, for the deferred unlock of a half-locked neighbour
        std::vector<uint64_t> one_prio;
        auto next_one_after = [&](uint64_t after, uint64_t excluding_prio) -> uint64_t {
            if (ranks) {  // the next set bit above `after` (ranks are unique: one may be excluded)
                const std::vector<uint32_t> &b = md.one_bits;
                for (uint64_t r = after + 1; (r >> 5) < b.size();) {
                    const uint32_t w = b[r >> 5] >> (r & 31);
                    if (!w) {
                        r = (r | 31) + 1;
                        continue;
                    }
                    r += (uint64_t)__builtin_ctz(w);
                    if (r != excluding_prio) return r;
                    r++;
                }
                return ~0ull;
            }
            if (one_prio.empty()) {
                for (size_t i = 0; i < S; i++) {
                    const uint64_t m = gmult[i];
                    int nr = 0, nl = 0;
                    for (int s = 0; s < 4; s++) nr += ((m >> (8 * s)) & 0xFF) != 0;
                    for (int s = 4; s < 8; s++) nl += ((m >> (8 * s)) & 0xFF) != 0;
                    if (nr == 1 && nl == 1) one_prio.push_back(prio[i]);
                }
                std::sort(one_prio.begin(), one_prio.end());
            }
            auto it = std::upper_bound(one_prio.begin(), one_prio.end(), after);
            while (it != one_prio.end() && *it == excluding_prio) ++it;
            return it == one_prio.end() ? ~0ull : *it;
        };
        // SWAP pass (graph.cpp:1392-1566) over the component's (1,1) nodes in
        // visit order: nodes[0 .. n_one)
        bool pending = false;
        uint64_t pendingId = 0, pendingAt = 0, pendingPrio = 0;
        for (size_t q = 0; q < n_one; q++) {
            HNode &M = nodes[q];
            if (!M.live) continue;
            if (pending && next_one_after(pendingAt, pendingPrio) < M.prio) pending = false;
            if (pending && pendingId == M.id) continue;  // still locked: its own trylock fails
            int li = -1, ri = -1, nl = 0, nr = 0;
            for (int s = 4; s < 8; s++) if (M.len[s]) { nl++; li = s; }
            for (int s = 0; s < 4; s++) if (M.len[s]) { nr++; ri = s; }
            if (nl != 1 || nr != 1) continue;
            uint64_t L, Rr;
            R.neighbour(M, li, &L);
            R.neighbour(M, ri, &Rr);
            pending = false;
            const bool lockL = L != M.id;
            const bool lockR = (Rr != L) ? (Rr != M.id) : lockL;
            if (!(lockL && lockR)) {
                if (lockL || lockR) {
                    pending = true;
                    pendingId = lockL ? L : Rr;
                    pendingAt = M.prio;
                    const uint32_t pi = lockL ? M.nbi[li] : M.nbi[ri];
                    pendingPrio = (pi != kNoNode && nodes[pi].kept < 0) ? nodes[pi].prio : ~0ull;
                }
                continue;
            }
            // the right neighbour's back edge takes M's left edge and the left
            // neighbour's M's right edge (M's two strings, taken before either
            // splice)
            const uint32_t srope[2] = {M.rope[li], M.rope[ri]};
            const uint64_t slen[2] = {M.len[li], M.len[ri]}, stail[2] = {M.tail[li], M.tail[ri]};
            const uint8_t ms[2] = {M.m[li], M.m[ri]};
            const uint32_t dsts[2] = {M.nbi[ri], M.nbi[li]};
            const int sides[2] = {0, 1};
            for (int q2 = 0; q2 < 2; q2++) {
                if (dsts[q2] == kNoNode) return SWAP_E_DEVICE;  // a neighbour outside the component
                HNode &D = nodes[dsts[q2]];
                const int idx = R.src_slot(D, M.id, sides[q2]);
                if (idx < 0) return SWAP_E_DEVICE;
                D.m[idx] = splice_service(D.m[idx], D.len[idx], ms[q2], slen[q2]);
                R.append(D, idx, srope[q2], slen[q2], stail[q2]);
                R.refresh(D, idx);
                R.link(D, idx);
            }
            M.live = false;
        }
        // The table as the merge leaves it: contigGraph.txt, the -p file and the
        // contig extension's input are written from this state (graph.cpp:3591-3622).
        auto collect = [&](std::vector<Out> &dst) {
            for (HNode &n : nodes) {
                if (n.kept >= 0) {
                    Out &o = dst[(size_t)n.kept];
                    for (int s = 0; s < 8; s++)
                        if (md.especial[8 * (size_t)n.kept + s]) {
                            o.arcs[s] = R.text(n.rope[s], n.len[s]);
                            o.m[s] = n.m[s];
                        }
                } else if (n.live) {
                    Out o;
                    o.id = n.id;
                    o.prio = n.prio;
                    for (int s = 0; s < 8; s++) {
                        o.arcs[s] = R.text(n.rope[s], n.len[s]);
                        o.m[s] = n.len[s] ? n.m[s] : 0;
                    }
                    dst.push_back(std::move(o));
                }
            }
        };
        std::vector<Out> after_swap = live;
        collect(after_swap);
        // masterSimplifyNodeGraph (graph.cpp:2089-2192) over the live (1,1) nodes
        // left, in the order of the table the gather re-fills (graph.cpp:3614-3615)
        std::vector<uint32_t> rest;
        for (size_t q = 0; q < n_one; q++)
            if (nodes[q].live) rest.push_back((uint32_t)q);
        if (ref && P3) {
            std::sort(rest.begin(), rest.end(), [&](uint32_t a, uint32_t b) {
                const uint64_t ba = ref_hash(nodes[a].id) % P3, bb = ref_hash(nodes[b].id) % P3;
                if (ba != bb) return ba < bb;
                return nodes[a].prio > nodes[b].prio;
            });
        } else {
            std::sort(rest.begin(), rest.end(), [&](uint32_t a, uint32_t b) { return nodes[a].prio < nodes[b].prio; });
        }
        bool finisher_merged = false;
        for (uint32_t q : rest) {
            HNode &M = nodes[q];
            if (!M.live) continue;
            int left = -1, right = -1, nl = 0, nr = 0;
            for (int s = 0; s < 4; s++) if (M.len[s]) { right = s; nr++; }
            for (int s = 4; s < 8; s++) if (M.len[s]) { left = s; nl++; }
            if (nr != 1 || nl != 1) continue;
            uint64_t rightNode, leftNode;
            R.neighbour(M, right, &rightNode);
            R.neighbour(M, left, &leftNode);
            const uint32_t ir = M.nbi[right], il = M.nbi[left];
            if (ir == kNoNode || il == kNoNode) continue;
            const int rr = R.src_slot(nodes[ir], M.id, 0);
            const int lr = R.src_slot(nodes[il], M.id, 1);
            if (rr == -1 || lr == -1) continue;
            if (leftNode == M.id || rightNode == M.id) continue;
            HNode &Rn = nodes[ir];
            Rn.m[rr] = splice_master(Rn.m[rr], Rn.len[rr], M.m[left], M.len[left]);
            R.append(Rn, rr, M.rope[left], M.len[left], M.tail[left]);
            R.refresh(Rn, rr);
            R.link(Rn, rr);
            HNode &Ln = nodes[il];
            Ln.m[lr] = splice_master(Ln.m[lr], Ln.len[lr], M.m[right], M.len[right]);
            R.append(Ln, lr, M.rope[right], M.len[right], M.tail[right]);
            R.refresh(Ln, lr);
            R.link(Ln, lr);
            M.live = false;
            finisher_merged = true;
        }
        if (finisher_merged) {
            finished_live = live;
            collect(finished_live);
        }
        live = std::move(after_swap);
    }
    // drop nodes without edges (graph.cpp:587-590)
    {
        size_t w = 0;
        for (size_t i = 0; i < live.size(); i++) {
            bool any = false;
            for (int s = 0; s < 8; s++) if (!live[i].arcs[s].empty()) any = true;
            if (any) { if (w != i) live[w] = std::move(live[i]); w++; }
        }
        live.resize(w);
    }
    // contigGraph.txt order: the gather walks the nodes table, i.e. visit order
    // (reference order); by key otherwise
    std::vector<uint32_t> ord(live.size());
    std::iota(ord.begin(), ord.end(), 0u);
    if (ref && P3)
        std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return live[a].prio < live[b].prio; });
    else
        std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return live[a].id < live[b].id; });
    // noCEcontigs.fasta order: the re-filled table (graph.cpp:3615, 891)
    out.node_buckets = table_buckets(S, false);
    out.noce_order.resize(live.size());
    if (ref && P3) {
        std::vector<uint64_t> keys(live.size());
        std::vector<uint32_t> pos(live.size());
        for (size_t r = 0; r < ord.size(); r++) { keys[r] = live[ord[r]].id; pos[r] = (uint32_t)r; }
        out.noce_order = gather_order(keys, pos, P3);
    } else {
        std::iota(out.noce_order.begin(), out.noce_order.end(), 0u);
    }
    // rows of a state in contigGraph.txt order
    auto fill = [&](HostCompacted &dst, const std::vector<const Out *> &rows) {
        size_t pool_n = 0;
        for (const Out *o : rows)
            if (o) for (int s = 0; s < 8; s++) pool_n += o->arcs[s].size();
        dst.key.resize(rows.size());
        dst.off.resize(8 * rows.size());
        dst.len.resize(8 * rows.size());
        dst.mult.resize(8 * rows.size());
        dst.pool.resize(pool_n);
        size_t at = 0;
        for (size_t r = 0; r < rows.size(); r++) {
            dst.key[r] = live[ord[r]].id;
            for (int s = 0; s < 8; s++) {
                const std::string empty;
                const std::string &a = rows[r] ? rows[r]->arcs[s] : empty;
                dst.off[8 * r + s] = at;
                dst.len[8 * r + s] = (uint32_t)a.size();
                dst.mult[8 * r + s] = a.empty() ? 0 : rows[r]->m[s];
                if (!a.empty()) { memcpy(dst.pool.data() + at, a.data(), a.size()); at += a.size(); }
            }
        }
    };
    std::vector<const Out *> rows(ord.size());
    for (size_t r = 0; r < ord.size(); r++) rows[r] = &live[ord[r]];
    fill(out, rows);
    if (!finished_live.empty()) {
        // the finisher's table: same rows; absorbed nodes keep no edge
        std::unordered_map<uint64_t, const Out *> fin_of;
        for (const Out &o : finished_live) fin_of[o.id] = &o;
        for (size_t r = 0; r < ord.size(); r++) {
            auto it = fin_of.find(live[ord[r]].id);
            rows[r] = it == fin_of.end() ? nullptr : it->second;
        }
        auto fin = std::make_shared<HostCompacted>();
        fill(*fin, rows);
        fin->noce_order = out.noce_order;
        fin->node_buckets = out.node_buckets;
        out.finished = std::move(fin);
    }
    *n_kept = live.size();
    return 0;
}

}  // namespace swapk
