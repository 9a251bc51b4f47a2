// CPU oracle: single-rank restatement of SWAP-Assembler 2's k-mer graph build
// and k-molecule merge.  TEST INFRASTRUCTURE ONLY (see swap_oracle.h for what
// pins each part).  Every function cites the reference code it restates.
//
// The reference keeps its tables in libstdc++'s __gnu_cxx::hash_map with a
// custom hash (graph.cpp:20-34, 353-354).  The iteration order of that
// container decides the merge visit order, and therefore the merged edge
// multiplicities and the output line order.  The oracle uses the same
// container with the same hash and issues the same sequence of find /
// operator[] calls, so its orders are the reference's np=1 orders by
// construction.
#include "swap_oracle.h"

#include <ext/hash_map>
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

// graph.cpp:29-32 -- (x >> 32) ^ hash<int>(x & 0xFFFFFFFF); hash<int> returns the
// int sign-extended to size_t.
struct RefHash {
    size_t operator()(const unsigned long long &x) const {
        return (size_t)(x >> 32) ^ (size_t)(long)(int)(unsigned)(x & 0xFFFFFFFFull);
    }
};

struct Arc {  // graph.cpp:36-40
    unsigned char m[8];
};

struct Node {  // graph.cpp:42-49 (mutex replaced by nothing: one rank, one thread)
    std::string arcs[8];
    unsigned char m[8];
    unsigned long long id;
    char live;  // the reference's deleteFlag: 1 = live, 0 = absorbed
    Node() : id(0), live(0) { memset(m, 0, 8); }
};

typedef __gnu_cxx::hash_map<unsigned long long, Arc, RefHash> ArcTable;
typedef __gnu_cxx::hash_map<unsigned long long, Node, RefHash> NodeTable;

const char kACGT[4] = {'A', 'C', 'G', 'T'};

inline int code(char ch) {  // sequence.cpp:164-168 (only ever applied to sanitised bases)
    switch (ch) { case 'A': return 0; case 'C': return 1; case 'G': return 2; default: return 3; }
}

struct Ctx {
    int k;
    unsigned long long kmask;  // 2k low bits set

    // kmerGraph.cpp:5-14 -- reverse complement, restated as a loop over bases.
    unsigned long long rc(unsigned long long x) const {
        unsigned long long r = 0;
        for (int i = 0; i < k; i++) { r = (r << 2) | (3 - (x & 3)); x >>= 2; }
        return r;
    }
    // kmerGraph.cpp:26-37
    std::string str(unsigned long long x) const {
        std::string s(k, 'A');
        for (int i = k - 1; i >= 0; i--) { s[i] = kACGT[x & 3]; x >>= 2; }
        return s;
    }
    // kmerGraph.cpp:16-24
    unsigned long long val(const char *s, size_t n) const {
        unsigned long long v = 0;
        for (size_t i = 0; i < n; i++) v = (v << 2) | (unsigned long long)code(s[i]);
        return v;
    }
    // Oriented start k-mer of edge slot i of node X: X for slots 0-3, rc(X) for 4-7
    // (graph.cpp:168-175).
    unsigned long long orient(unsigned long long X, int i) const { return i >= 4 ? rc(X) : X; }
    // Last k bases of orient(X, i) + arc (graph.cpp:177-181).
    unsigned long long terminal(unsigned long long X, int i, const std::string &arc) const {
        const size_t L = arc.size();
        if (L >= (size_t)k) return val(arc.data() + L - k, k);
        unsigned long long t = orient(X, i);
        for (size_t j = 0; j < L; j++) t = ((t << 2) | (unsigned long long)code(arc[j])) & kmask;
        return t;
    }
    // node::getNodeID (graph.cpp:164-196): canonical neighbour through slot i;
    // returns +1 if the walk ends on the neighbour's forward strand, -1 otherwise.
    int neighbour(const Node &n, int i, unsigned long long &id) const {
        const unsigned long long t = terminal(n.id, i, n.arcs[i]);
        const unsigned long long r = rc(t);
        if (t < r) { id = r; return -1; }
        id = t;
        return 1;
    }
    // node::getSrcNodeEdge with srcEdge="" srcMulti=0 direction=2 (graph.cpp:263-324):
    // first slot of dst whose walk ends on src in orientation side (1 = forward).
    int src_slot(const Node &dst, unsigned long long src, int side) const {
        for (int i = 0; i < 8; i++) {
            const unsigned long long t = terminal(dst.id, i, dst.arcs[i]);
            const unsigned long long r = rc(t);
            const int ret = (t < r) ? 0 : 1;
            const unsigned long long id = (t < r) ? r : t;
            if (id == src && ret == side) return i;
        }
        return -1;
    }
};

// Integer average of a splice in the service loop (graph.cpp:1343-1345): the sum
// is formed in size_t, stored into an int, and divided by the size_t length.
inline unsigned char splice_avg_service(unsigned char ma, size_t la, unsigned char mb, size_t lb) {
    const size_t s = (size_t)ma * la + (size_t)mb * lb;
    const int pre = (int)(unsigned)s;
    return (unsigned char)((size_t)(long)pre / (la + lb));
}
// The same average in masterSimplifyNodeGraph (graph.cpp:2162-2168), all size_t.
inline unsigned char splice_avg_master(unsigned char ma, size_t la, unsigned char mb, size_t lb) {
    return (unsigned char)(((size_t)ma * la + (size_t)mb * lb) / (la + lb));
}

}  // namespace

struct oracle_graph {
    Ctx cx;
    int c;
    int flags;
    uint64_t occurrences = 0, reads = 0, distinct = 0, survivors = 0, kept = 0;
    std::vector<unsigned long long> table_keys;   // sorted pre-cutoff keys
    std::vector<Arc> table_arcs;
    std::vector<unsigned long long> surv_keys;    // sorted survivor keys
    std::vector<Arc> surv_arcs;
    std::vector<uint64_t> surv_first;
    std::vector<uint64_t> surv_rank;              // index of the key among first occurrences
    std::vector<uint64_t> first_in_order;         // first occurrence of the i-th distinct key
    std::vector<unsigned long long> visit;        // survivors in merge visit order
    std::string kmer_graph_txt, contig_graph_txt, noce_txt, perf_txt;
};

namespace {

// kmerGraph.cpp:95-310 + graph.cpp:1049-1090 at np=1: records, sanitising, rolling
// encode, max-canonical, per-occurrence edge bits, table insert with saturating
// counts.  The reference processes the buffer in windows of 16 KiB; at np=1 that
// changes neither which records are read nor their order, so one pass stands in.
bool build_table(oracle_graph *g, const char *fasta, uint64_t n, ArcTable &kmol,
                 std::vector<uint64_t> *first_of, std::vector<unsigned long long> *first_key) {
    const Ctx &cx = g->cx;
    const int k = cx.k;
    std::string buf(fasta, n);
    // kmerGraph.cpp:112-114: start at the first '>'.
    size_t pos = buf.find('>');
    if (pos == std::string::npos) return false;
    const unsigned long long MASK = ~(3ull << (2 * k));  // sequence.cpp:180
    std::vector<unsigned long long> kmers;
    std::vector<unsigned char> arcs;
    while (pos < n) {
        // header line, then exactly one sequence line (kmerGraph.cpp:173-181)
        const size_t hdr_end = buf.find('\n', pos);
        if (hdr_end == std::string::npos) break;
        const size_t s0 = hdr_end + 1;
        const size_t s1 = buf.find('\n', s0);
        if (s1 == std::string::npos) break;  // unterminated last record: never consumed
        const size_t len = s1 - s0;
        pos = s1 + 1;
        if (len < (size_t)k) continue;  // kmerGraph.cpp:182-187
        char *p = &buf[s0];
        for (size_t j = 0; j < len; j++) {  // kmerGraph.cpp:228-233
            char ch = p[j];
            if (ch == 'A' || ch == 'C' || ch == 'G' || ch == 'T') continue;
            if (ch == 'a' || ch == 'c' || ch == 'g' || ch == 't') p[j] = ch + 'A' - 'a';
            else p[j] = 'A';
        }
        g->reads++;
        // kmerGraph.cpp:240-300: rolling encode and the 4-case edge table.
        const size_t nk = len - k + 1;
        kmers.assign(nk, 0);
        arcs.assign(nk, 0);
        unsigned long long cur = cx.val(p, k - 1), pre = 0, prePrime = 0;
        for (size_t j = 0; j < nk; j++) {
            cur = ((cur << 2) | (unsigned long long)code(p[j + k - 1])) & MASK;
            const unsigned long long twin = cx.rc(cur);
            const unsigned long long prime = cur > twin ? cur : twin;
            kmers[j] = prime;
            if (j > 0) {
                // arcPos (kmerGraph.cpp:39-47) folded: the previous k-mer gains
                // "followed by p[j+k-1]" on the strand it was read on; the current one
                // gains "rc followed by complement of p[j-1]".
                const bool preFwd = (pre == prePrime), curFwd = (cur == prime);
                const int b = code(p[j + k - 1]);
                const int a = 3 - code(p[j - 1]);
                arcs[j - 1] |= (unsigned char)(1u << (preFwd ? b : 4 + b));
                arcs[j] |= (unsigned char)(1u << (curFwd ? 4 + a : a));
            }
            pre = cur;
            prePrime = prime;
        }
        // graph.cpp:1058-1080, in occurrence order, with the same find/operator[]
        // sequence as the reference so the container evolves identically.
        for (size_t j = 0; j < nk; j++) {
            const unsigned long long key = kmers[j];
            const unsigned char a = arcs[j];
            if (kmol.find(key) == kmol.end()) {
                Arc na;
                for (int t = 0; t < 8; t++) na.m[t] = (a >> t) & 1;
                kmol[key] = na;
                if (first_of) { first_of->push_back(g->occurrences + j); first_key->push_back(key); }
            } else {
                for (int t = 0; t < 8; t++)
                    if ((a & (1 << t)) && kmol[key].m[t] < 255) kmol[key].m[t]++;
            }
        }
        g->occurrences += nk;
    }
    return true;
}

// graph.cpp:1092-1114
void cutoff(ArcTable &kmol, int threshold) {
    ArcTable ret;
    for (ArcTable::iterator it = kmol.begin(); it != kmol.end(); ++it) {
        int cnt = 0;
        for (int i = 0; i < 8; i++) {
            if ((int)it->second.m[i] < threshold) it->second.m[i] = 0;
            if (it->second.m[i] != 0) cnt++;
        }
        if (cnt > 0) ret[it->first] = it->second;
    }
    kmol.clear();
    kmol = ret;
}

// node::init + buildKmoleculeGraph (graph.cpp:81-102, 1117-1137)
void build_nodes(ArcTable &kmol, NodeTable &nodes) {
    for (ArcTable::iterator it = kmol.begin(); it != kmol.end(); ++it) {
        Node nn;
        nn.id = it->first;
        nn.live = 1;
        for (int i = 0; i < 8; i++) {
            if (it->second.m[i]) { nn.arcs[i] = std::string(1, kACGT[i % 4]); nn.m[i] = it->second.m[i]; }
        }
        nodes[it->first] = nn;
    }
    kmol.clear();
}

// simplifyDistNodeGraph + recvProc at np=1 (graph.cpp:1204-1566).  With one rank
// every lock request is answered by the same process, so a lock fails only when
// the requested node is the visited node itself, and the two merge messages are
// served in the order they were sent (right neighbour first).
//
// One message is served late: after a half-successful lock, the unlock of the
// locked neighbour (graph.cpp:1520-1548) is sent with MPI_Isend into the already
// posted service receive, so recvProc(...,1,...) returns as soon as the send
// completes and the unlock is only applied when the next lock request enters
// the service loop.  A node visited while it is still locked that way fails its
// own trylock (graph.cpp:1422-1427) and is skipped by this pass.
void swap_merge(const Ctx &cx, NodeTable &nodes, int flags) {
    const bool plant_eager_unlock = flags & 2, plant_double_lock = flags & 4;
    bool pending = false;
    unsigned long long pendingId = 0;
    for (NodeTable::iterator it = nodes.begin(); it != nodes.end(); ++it) {
        Node &M = it->second;
        if (!M.live) continue;
        if (pending && pendingId == M.id && !plant_eager_unlock) continue;  // still locked: trylock fails
        int nl = 0, li = -1, nr = 0, ri = -1;
        for (int i = 4; i < 8; i++) if (!M.arcs[i].empty()) { nl++; li = i; }
        for (int i = 0; i < 4; i++) if (!M.arcs[i].empty()) { nr++; ri = i; }
        if (nl != 1 || nr != 1) continue;  // graph.cpp:1429-1438
        unsigned long long L, R;
        cx.neighbour(M, li, L);
        cx.neighbour(M, ri, R);
        pending = false;  // the first lock request of this visit serves the late unlock
        const bool lockL = (L != M.id);
        const bool lockR = (R != L) ? (R != M.id) : (plant_double_lock ? false : lockL);  // graph.cpp:1464-1475
        if (!(lockL && lockR)) {
            if (lockL || lockR) { pending = true; pendingId = lockL ? L : R; }
            continue;
        }
        // message to R: M's left edge, side 0 (graph.cpp:1480-1495); then to L: M's
        // right edge, side 1 (graph.cpp:1506-1511).
        const std::string leftEdge = M.arcs[li], rightEdge = M.arcs[ri];
        const unsigned char leftM = M.m[li], rightM = M.m[ri];
        struct Msg { unsigned long long dst; int side; const std::string *s; unsigned char m; };
        const Msg msgs[2] = {{R, 0, &leftEdge, leftM}, {L, 1, &rightEdge, rightM}};
        for (const Msg &msg : msgs) {
            NodeTable::iterator d = nodes.find(msg.dst);
            if (d == nodes.end()) { fprintf(stderr, "oracle: neighbour %llu missing\n", msg.dst); abort(); }
            Node &D = d->second;
            const int idx = cx.src_slot(D, M.id, msg.side);
            if (idx < 0) { fprintf(stderr, "oracle: no back edge %llu -> %llu\n", msg.dst, M.id); abort(); }
            const size_t la = D.arcs[idx].size(), lb = msg.s->size();
            D.m[idx] = splice_avg_service(D.m[idx], la, msg.m, lb);
            D.arcs[idx] += *msg.s;
        }
        M.live = 0;
    }
}

// One contigGraph.txt line (graph.cpp:592-594, 641-644): the sprintf'd record
// padded with spaces to sum(len_i + 1) + 61 bytes, the last byte '\n'.
void graph_line(const Node &n, std::string &out) {
    char head[32];
    snprintf(head, sizeof head, "%llu ", n.id);
    std::string s = head;
    size_t total = 0;
    for (int i = 0; i < 8; i++) { s += n.arcs[i]; s += "# "; total += n.arcs[i].size() + 1; }
    s.pop_back();  // "%s# %d ..." : the space after the last '#' is the separator below
    for (int i = 0; i < 8; i++) { char t[8]; snprintf(t, sizeof t, " %d", n.m[i]); s += t; }
    total += 20 + 8 * 3 + 16 + 1;
    if (s.size() < total) s.append(total - s.size(), ' ');
    s.resize(total);
    s[total - 1] = '\n';
    out += s;
}

// gatherDistNodeGraph at np=1 (graph.cpp:576-696): live nodes with at least one
// edge, in table order.  recvBufDistNodeGraph (graph.cpp:845-896) then re-inserts
// them into the cleared table, which keeps its bucket count.
void gather(NodeTable &nodes, std::string &text) {
    std::vector<Node> live;
    for (NodeTable::iterator it = nodes.begin(); it != nodes.end(); ++it) {
        const Node &n = it->second;
        if (!n.live) continue;
        bool any = false;
        for (int i = 0; i < 8; i++) if (!n.arcs[i].empty()) any = true;
        if (!any) continue;
        graph_line(n, text);
        live.push_back(n);
    }
    nodes.clear();
    for (Node &n : live) { n.live = 1; nodes[n.id] = n; }
}

// masterSimplifyNodeGraph (graph.cpp:2089-2192)
void master_simplify(const Ctx &cx, NodeTable &nodes) {
    for (NodeTable::iterator it = nodes.begin(); it != nodes.end(); ++it) {
        if (!it->second.live) continue;
        int left = -1, right = -1, nl = 0, nr = 0;
        for (int i = 0; i < 4; i++) if (!it->second.arcs[i].empty()) { right = i; nr++; }
        for (int i = 4; i < 8; i++) if (!it->second.arcs[i].empty()) { left = i; nl++; }
        if (nr != 1 || nl != 1) continue;
        unsigned long long rightNode, leftNode;
        cx.neighbour(it->second, right, rightNode);
        cx.neighbour(it->second, left, leftNode);
        int rr = cx.src_slot(nodes[rightNode], it->first, 0);
        int lr = cx.src_slot(nodes[leftNode], it->first, 1);
        if (rr == -1 || lr == -1) continue;
        if (leftNode == it->first || rightNode == it->first) continue;
        Node &Rn = nodes[rightNode];
        Rn.m[rr] = splice_avg_master(Rn.m[rr], Rn.arcs[rr].size(), it->second.m[left], it->second.arcs[left].size());
        Rn.arcs[rr] += it->second.arcs[left];
        Node &Ln = nodes[leftNode];
        Ln.m[lr] = splice_avg_master(Ln.m[lr], Ln.arcs[lr].size(), it->second.m[right], it->second.arcs[right].size());
        Ln.arcs[lr] += it->second.arcs[right];
        it->second.live = 0;
    }
}

// printContigs (graph.cpp:1139-1202) with Contig_Length = 100 (sequence.h:15).
void print_contigs(const Ctx &cx, NodeTable &nodes, std::string &out) {
    for (NodeTable::iterator it = nodes.begin(); it != nodes.end(); ++it) {
        const Node &n = it->second;
        if (!n.live) continue;
        for (int i = 0; i < 8; i++) {
            if (n.arcs[i].empty()) continue;
            if ((long)n.arcs[i].size() < 100 - cx.k) continue;
            unsigned long long id;
            cx.neighbour(n, i, id);
            if (n.id < id) continue;
            out += ">proc0: ";
            out += cx.str(cx.orient(n.id, i));
            out += n.arcs[i];
            out += "\n";
        }
    }
}

// stringContigs (graph.cpp:3166-3238), the -p output: printContigs' selection
// over the table the merge leaves (before the gather), each contig as
// ">contig_NodeID<id>_Index<slot>_Len<len>_Mul<m>_<running count>\n<seq>\n".
void print_perf_contigs(const Ctx &cx, NodeTable &nodes, std::string &out) {
    int count = 0;
    for (NodeTable::iterator it = nodes.begin(); it != nodes.end(); ++it) {
        const Node &n = it->second;
        if (!n.live) continue;
        for (int i = 0; i < 8; i++) {
            if (n.arcs[i].empty()) continue;
            if ((long)n.arcs[i].size() < 100 - cx.k) continue;
            unsigned long long id;
            cx.neighbour(n, i, id);
            if (n.id < id) continue;
            count++;
            const std::string seq = cx.str(cx.orient(n.id, i)) + n.arcs[i];
            out += ">contig_NodeID" + std::to_string(it->first) + "_Index" + std::to_string(i) + "_Len" +
                   std::to_string(seq.size()) + "_Mul" + std::to_string((int)n.m[i]) + "_" + std::to_string(count) +
                   "\n" + seq + "\n";
        }
    }
}

// printKmerGraph (graph.cpp:3352-3440), including its neighbour encoding: when the
// walk ends on the reverse strand the value written is that of reverse(str(Y))
// with only the character at the slot index complemented (the loop at
// graph.cpp:3392-3393 indexes with the slot variable).
void print_kmer_graph(const Ctx &cx, NodeTable &nodes, std::string &out) {
    const char comp[4] = {'T', 'G', 'C', 'A'};
    for (NodeTable::iterator it = nodes.begin(); it != nodes.end(); ++it) {
        const Node &n = it->second;
        if (!n.live) continue;
        for (int half = 0; half < 2; half++) {
            unsigned long long next[4];
            int kc = 0;
            for (int i = 4 * half; i < 4 * half + 4; i++) {
                if (n.arcs[i].empty()) continue;
                unsigned long long id;
                if (cx.neighbour(n, i, id) == -1) {
                    std::string s = cx.str(id);
                    std::reverse(s.begin(), s.end());
                    if (i < (int)s.size()) s[i] = comp[code(s[i])];
                    id = cx.val(s.data(), s.size());
                }
                next[kc++] = id;
            }
            if (kc == 0) continue;
            char t[32];
            snprintf(t, sizeof t, "%llu", half == 0 ? n.id : cx.rc(n.id));
            out += t;
            for (int j = 0; j < kc; j++) { snprintf(t, sizeof t, " %llu", next[j]); out += t; }
            out += "\n";
        }
    }
}

}  // namespace

extern "C" {

oracle_graph *oracle_run(const char *fasta, uint64_t n, int k, int c, int flags) {
    if (k < 1 || k > 31 || (k % 2) == 0) return nullptr;
    oracle_graph *g = new oracle_graph();
    g->cx.k = k;
    g->cx.kmask = (k == 32) ? ~0ull : ((1ull << (2 * k)) - 1);
    g->c = c;
    g->flags = flags;
    ArcTable kmol;
    std::vector<uint64_t> first_of;
    std::vector<unsigned long long> first_key;
    if (!build_table(g, fasta, n, kmol, &first_of, &first_key)) { delete g; return nullptr; }
    g->distinct = kmol.size();
    g->first_in_order = first_of;
    {
        std::vector<std::pair<unsigned long long, Arc>> v(kmol.begin(), kmol.end());
        std::sort(v.begin(), v.end(), [](const std::pair<unsigned long long, Arc> &a, const std::pair<unsigned long long, Arc> &b) { return a.first < b.first; });
        g->table_keys.reserve(v.size());
        g->table_arcs.reserve(v.size());
        for (auto &e : v) { g->table_keys.push_back(e.first); g->table_arcs.push_back(e.second); }
    }
    cutoff(kmol, c);
    g->survivors = kmol.size();
    {
        std::vector<std::pair<unsigned long long, Arc>> v(kmol.begin(), kmol.end());
        std::sort(v.begin(), v.end(), [](const std::pair<unsigned long long, Arc> &a, const std::pair<unsigned long long, Arc> &b) { return a.first < b.first; });
        std::vector<std::pair<unsigned long long, uint64_t>> fk(first_key.size());
        for (size_t i = 0; i < first_key.size(); i++) fk[i] = {first_key[i], i};
        std::sort(fk.begin(), fk.end());
        for (auto &e : v) {
            g->surv_keys.push_back(e.first);
            g->surv_arcs.push_back(e.second);
            auto f = std::lower_bound(fk.begin(), fk.end(), std::make_pair(e.first, (uint64_t)0));
            g->surv_first.push_back(first_of[f->second]);
            g->surv_rank.push_back(f->second);
        }
    }
    NodeTable nodes;
    build_nodes(kmol, nodes);
    for (NodeTable::iterator it = nodes.begin(); it != nodes.end(); ++it) g->visit.push_back(it->first);
    if (!(flags & 1)) {
        swap_merge(g->cx, nodes, flags);  // graph.cpp:3576
        print_perf_contigs(g->cx, nodes, g->perf_txt);  // what -p writes at this point (graph.cpp:3591-3611)
    }
    gather(nodes, g->contig_graph_txt);
    g->contig_graph_txt += "\n";  // graph.cpp:3617-3622: fprintf("%s\n")
    if (flags & 1) {
        print_kmer_graph(g->cx, nodes, g->kmer_graph_txt);
    } else {
        master_simplify(g->cx, nodes);
        print_contigs(g->cx, nodes, g->noce_txt);
    }
    for (NodeTable::iterator it = nodes.begin(); it != nodes.end(); ++it) if (it->second.live) g->kept++;
    return g;
}

void oracle_free(oracle_graph *g) { delete g; }

uint64_t oracle_num_occurrences(const oracle_graph *g) { return g->occurrences; }
uint64_t oracle_num_reads(const oracle_graph *g) { return g->reads; }
uint64_t oracle_num_distinct(const oracle_graph *g) { return g->distinct; }
uint64_t oracle_num_survivors(const oracle_graph *g) { return g->survivors; }
uint64_t oracle_num_kept(const oracle_graph *g) { return g->kept; }

uint64_t oracle_table(const oracle_graph *g, uint64_t *keys, uint8_t *mult, uint64_t cap) {
    const uint64_t n = g->table_keys.size();
    for (uint64_t i = 0; i < n && i < cap; i++) {
        if (keys) keys[i] = g->table_keys[i];
        if (mult) memcpy(mult + 8 * i, g->table_arcs[i].m, 8);
    }
    return n;
}

uint64_t oracle_survivors(const oracle_graph *g, uint64_t *keys, uint8_t *mult, uint64_t cap) {
    const uint64_t n = g->surv_keys.size();
    for (uint64_t i = 0; i < n && i < cap; i++) {
        if (keys) keys[i] = g->surv_keys[i];
        if (mult) memcpy(mult + 8 * i, g->surv_arcs[i].m, 8);
    }
    return n;
}

uint64_t oracle_survivor_first(const oracle_graph *g, uint64_t *first, uint64_t cap) {
    const uint64_t n = g->surv_first.size();
    for (uint64_t i = 0; i < n && i < cap; i++) first[i] = g->surv_first[i];
    return n;
}

uint64_t oracle_survivor_rank(const oracle_graph *g, uint64_t *rank, uint64_t cap) {
    const uint64_t n = g->surv_rank.size();
    for (uint64_t i = 0; i < n && i < cap; i++) rank[i] = g->surv_rank[i];
    return n;
}

uint64_t oracle_first_of_rank(const oracle_graph *g, uint64_t rank) {
    return rank < g->first_in_order.size() ? g->first_in_order[rank] : ~0ull;
}

uint64_t oracle_visit_order(const oracle_graph *g, uint64_t *keys, uint64_t cap) {
    const uint64_t n = g->visit.size();
    for (uint64_t i = 0; i < n && i < cap; i++) keys[i] = g->visit[i];
    return n;
}

uint64_t oracle_file(const oracle_graph *g, int which, char *buf, uint64_t cap) {
    const std::string *s = which == 0   ? &g->kmer_graph_txt
                           : which == 1 ? &g->contig_graph_txt
                           : which == 2 ? &g->noce_txt
                                        : &g->perf_txt;
    if (buf && cap >= s->size()) memcpy(buf, s->data(), s->size());
    return s->size();
}

}  // extern "C"
