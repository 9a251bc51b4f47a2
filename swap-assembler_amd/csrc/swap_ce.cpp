// Contig extension on the host: the reference's rank-0 finisher and graph
// clean-up over the merged graph, and the files written from it.
//
// The merged graph comes in as contigGraph.txt holds it (kept nodes in the
// gather order, 8 edge strings and multiplicities each).  The reference refills
// its cleared nodes table from that text (recvBufDistNodeGraph,
// graph.cpp:845-896) and then runs, on rank 0 (graph.cpp:3636-3682):
//   masterSimplifyNodeGraph; printContigs (noCEcontigs.fasta);
//   4 x { masterTipsRemoval, masterSimplifyNodeGraph,
//         masterLowCoverageEdgeRemoval, masterSimplifyNodeGraph,
//         masterLoopBubbleRemoval, masterMultipleEdgeBubbleRemoval,
//         masterSimplifyNodeGraph, masterGraphStatistic };
//   masterRemoveCrossEdge; masterPrintContigs (CEContig.fasta);
//   printJungGraph with -j.
// Every pass walks the table in its iteration order and edits it in place, so
// the table here is the same libstdc++ __gnu_cxx::hash_map, with the same hash
// and the same bucket count as the reference's (the merge's node table, sized
// by its S inserts, then cleared), filled in the same order; a lookup of an
// absent key inserts an empty node exactly where the reference's operator[]
// does.  (Further lookups of a key already present only ask the table to hold
// size+1 elements, which changes nothing here: the table was sized for all S
// survivors and holds the kept ones.)  The passes keep the reference's
// arithmetic -- integer averages over size_t lengths, the low-coverage
// threshold compared as a double -- and its quirks, noted where they occur.
#define _GLIBCXX_PERMIT_BACKWARD_HASH  // the reference's own container, on purpose
#include <ext/hash_map>

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "swap_host.h"

namespace swapk {

namespace {

struct CeNode {
    unsigned long long id = 0;  // the node's k-mer (0 in a node created by a lookup)
    std::string arcs[8];
    unsigned char m[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // deleteFlag.  A node created by a lookup gets an indeterminate flag in the
    // reference (node() leaves it unset); such a node never carries an edge, so
    // no pass and no writer can tell, and it is dead here.
    bool live = false;
};

struct CeHash {
    size_t operator()(unsigned long long x) const { return (size_t)ref_hash(x); }
};

typedef __gnu_cxx::hash_map<unsigned long long, CeNode, CeHash> CeTable;

constexpr int kContigLength = 100;  // sequence.h:15

// (a*la + b*lb) / (la + lb) in size_t, 0 when both lengths are 0 (where the
// reference would divide by zero)
inline unsigned char avg2(size_t a, size_t la, size_t b, size_t lb) {
    const size_t d = la + lb;
    return d ? (unsigned char)((a * la + b * lb) / d) : 0;
}

class Ce {
public:
    Ce(int k, CeTable &t) : k_(k), t_(t) {}

    std::vector<std::string> removed;  // RemovedEdges: bubble paths taken out
    bool divide_by_zero = false;       // an average coverage over no edge base

    std::string kmer_str(unsigned long long x) const {
        std::string s((size_t)k_, 'A');
        for (int i = 0; i < k_; i++) s[(size_t)i] = "ACGT"[(x >> (2 * (k_ - 1 - i))) & 3];
        return s;
    }

    static std::string revcomp_str(const std::string &s) {
        std::string r(s.rbegin(), s.rend());
        for (char &c : r) c = c == 'A' ? 'T' : c == 'C' ? 'G' : c == 'G' ? 'C' : c == 'T' ? 'A' : '\0';
        return r;
    }

    static unsigned long long value(const std::string &s) {
        unsigned long long v = 0;
        for (char c : s) v = (v << 2) | (unsigned long long)(c == 'C' ? 1 : c == 'G' ? 2 : c == 'T' ? 3 : 0);
        return v;
    }

    // the node's k-mer read on the side of slot i (itself for 0-3, its reverse
    // complement for 4-7) followed by the slot's edge string
    std::string descriptor(const CeNode &n, int i) const {
        std::string d = kmer_str(n.id);
        if (i > 3) d = revcomp_str(d);
        return d + n.arcs[i];
    }

    // node::getNodeID (graph.cpp:164-196): canonical id of the last k bases of
    // the descriptor; 1 if they read forward, -1 otherwise
    int next_id(const CeNode &n, int i, unsigned long long &id) const {
        const std::string d = descriptor(n, i);
        const std::string tail = d.substr(d.size() - (size_t)k_);
        const std::string rc = revcomp_str(tail);
        if (tail < rc) {
            id = value(rc);
            return -1;
        }
        id = value(tail);
        return 1;
    }

    // node::getSrcNodeEdge (graph.cpp:263-324): the first slot of sides 4-7,
    // 0-3 or all (direction 0, 1, 2) whose walk ends at `src` read with
    // orientation src_dir (1 forward, 0 reverse), with multiplicity src_multi
    // and, reversed, the string src_str when those are given; side + (slot<<2),
    // or -1.  Empty slots are examined too, as in the reference.
    int src_edge(const CeNode &n, unsigned long long src, int src_dir, const std::string &src_str, int src_multi,
                 int direction) const {
        int lo = 0, hi = 8;
        if (direction == 0) lo = 4;
        if (direction == 1) hi = 4;
        for (int i = lo; i < hi; i++) {
            const std::string d = descriptor(n, i);
            const std::string tail = d.substr(d.size() - (size_t)k_);
            const std::string rc = revcomp_str(tail);
            int ret = 1;
            unsigned long long id = value(tail);
            if (tail < rc) {
                ret = 0;
                id = value(rc);
            }
            if (id != src || src_dir != ret) continue;
            if (src_multi != 0 && n.m[i] != src_multi) continue;
            if (!src_str.empty() && src_str != revcomp_str(d).substr((size_t)k_)) continue;
            return ret + (i << 2);
        }
        return -1;
    }

    int src_edge(const CeNode &n, unsigned long long src, int src_dir) const {
        static const std::string none;
        return src_edge(n, src, src_dir, none, 0, 2);
    }

    CeNode &at(unsigned long long key) { return t_[key]; }

    static void clear_slot(CeNode &n, int i) {
        n.arcs[i].clear();
        n.m[i] = 0;
    }

    // masterSimplifyNodeGraph (graph.cpp:2089-2192): absorbs every node with
    // one left and one right edge, in table order
    void simplify() {
        for (CeTable::iterator it = t_.begin(); it != t_.end(); ++it) {
            CeNode &n = it->second;
            if (!n.live) continue;
            int left = -1, right = -1, nl = 0, nr = 0;
            for (int i = 0; i < 4; i++) if (!n.arcs[i].empty()) right = i, nr++;
            for (int i = 4; i < 8; i++) if (!n.arcs[i].empty()) left = i, nl++;
            if (nr != 1 || nl != 1) continue;
            unsigned long long rn, ln;
            next_id(n, right, rn);
            next_id(n, left, ln);
            int rr = src_edge(at(rn), it->first, 0);
            int lr = src_edge(at(ln), it->first, 1);
            if (rr == -1 || lr == -1) continue;
            rr >>= 2;
            lr >>= 2;
            if (ln == it->first || rn == it->first) continue;
            CeNode &R = at(rn);
            R.m[rr] = avg2(R.m[rr], R.arcs[rr].size(), n.m[left], n.arcs[left].size());
            R.arcs[rr] += n.arcs[left];
            CeNode &L = at(ln);
            L.m[lr] = avg2(L.m[lr], L.arcs[lr].size(), n.m[right], n.arcs[right].size());
            L.arcs[lr] += n.arcs[right];
            n.live = false;
        }
    }

    // masterTipsRemoval (graph.cpp:2194-2265): a node with a single edge shorter
    // than k is cut off together with its neighbour's back edge
    void tips() {
        for (CeTable::iterator it = t_.begin(); it != t_.end(); ++it) {
            CeNode &n = it->second;
            if (!n.live) continue;
            int arcs = 0, index = -1;
            for (int i = 0; i < 8; i++) if (!n.arcs[i].empty()) arcs++, index = i;
            if (arcs != 1 || n.arcs[index].size() >= (size_t)k_) continue;
            unsigned long long nx;
            next_id(n, index, nx);
            int back = src_edge(at(nx), it->first, index < 4 ? 0 : 1);
            if (back == -1) {
                clear_slot(n, index);
                continue;
            }
            back >>= 2;
            if (nx == it->first) continue;
            clear_slot(at(nx), back);
            clear_slot(n, index);
            n.live = false;
        }
    }

    // masterLowCoverageEdgeRemoval (graph.cpp:2267-2300): edges shorter than 2k
    // below a quarter of the length-weighted mean multiplicity are dropped, on
    // this side only (the neighbour keeps its back edge)
    void low_coverage() {
        unsigned long long cov = 0, bases = 0;
        for (CeTable::iterator it = t_.begin(); it != t_.end(); ++it) {
            const CeNode &n = it->second;
            if (!n.live) continue;
            for (int i = 0; i < 8; i++) {
                if (n.arcs[i].empty()) continue;
                bases += n.arcs[i].size();
                cov += n.arcs[i].size() * n.m[i];
            }
        }
        if (bases == 0) {
            divide_by_zero = true;
            return;
        }
        const unsigned long long avg = cov / bases;
        for (CeTable::iterator it = t_.begin(); it != t_.end(); ++it) {
            CeNode &n = it->second;
            if (!n.live) continue;
            for (int i = 0; i < 8; i++)
                if (n.arcs[i].size() < 2 * (size_t)k_ && n.m[i] < 0.25 * (double)avg) clear_slot(n, i);
        }
    }

    // masterLoopBubbleRemoval (graph.cpp:2302-2477): nodes with an edge back to
    // themselves
    void loop_bubbles() {
        for (CeTable::iterator it = t_.begin(); it != t_.end(); ++it) {
            CeNode &n = it->second;
            if (!n.live) continue;
            unsigned long long nx = 0, nx2 = 0;
            int degree = 0, loop_degree = 0;
            std::vector<int> loop_slot, loop_dir, out;
            for (int i = 0; i < 8; i++) {
                if (n.arcs[i].empty()) continue;
                degree++;
                const int r = next_id(n, i, nx);
                if (nx == it->first) loop_degree++, loop_slot.push_back(i), loop_dir.push_back(r);
                else out.push_back(i);
            }
            if (out.empty()) continue;
            if (loop_degree == 2 && loop_dir[0] != loop_dir[1]) {
                if (degree == 4) {
                    next_id(n, out[0], nx);
                    int b1 = src_edge(at(nx), it->first, out[0] < 4 ? 0 : 1);
                    if (b1 == -1) continue;
                    next_id(n, out[1], nx2);
                    int b2 = src_edge(at(nx2), it->first, out[1] < 4 ? 0 : 1);
                    if (b2 == -1) continue;
                    if (b1 % 2 == b2 % 2) continue;  // a dead sink
                }
                next_id(n, out[0], nx);
                int back = src_edge(at(nx), it->first, out[0] < 4 ? 0 : 1);
                if (back == -1) continue;
                const int d = back % 2;
                back >>= 2;
                if (degree == 3 || degree == 4) {
                    // the loop arc on the exit side takes the out edge; the
                    // neighbour's back edge takes the other loop arc (k = 1)
                    const int a = loop_slot[(size_t)d], o = out[0], b = loop_slot[(size_t)(1 - d)];
                    n.m[a] = avg2(n.m[a], n.arcs[a].size(), n.m[o], n.arcs[o].size());
                    n.arcs[a] += n.arcs[o];
                    clear_slot(n, o);
                    CeNode &N = at(nx);
                    N.m[back] = avg2(N.m[back], N.arcs[back].size(), n.m[b], n.arcs[b].size());
                    N.arcs[back] += n.arcs[b];
                    clear_slot(n, b);
                }
            }
            // two loop arcs in the same direction: left alone (graph.cpp:2376-2379)
            if (loop_degree == 1) {
                const int L = loop_slot[0];
                if (degree == 2) {
                    next_id(n, out[0], nx);
                    int back = src_edge(at(nx), it->first, out[0] < 4 ? 0 : 1);
                    if (back == -1) continue;
                    back >>= 2;
                    const int O = out[0];
                    if ((L < 4 && loop_dir[0] == -1 && O > 3) || (L > 3 && loop_dir[0] == 1 && O < 4)) {
                        // half the loop's weight goes each way (m*len/2 in size_t)
                        CeNode &N = at(nx);
                        const size_t ln = N.arcs[back].size(), ll = n.arcs[L].size(), lo = n.arcs[O].size();
                        const size_t dn = ln + ll, dl = ll + lo;
                        N.m[back] = dn ? (unsigned char)((N.m[back] * ln + n.m[L] * ll / 2) / dn) : 0;
                        N.arcs[back] += n.arcs[L];
                        n.m[L] = dl ? (unsigned char)((n.m[L] * ll / 2 + n.m[O] * lo) / dl) : 0;
                        n.arcs[L] += n.arcs[O];
                        clear_slot(n, O);
                    }
                } else if (degree == 3) {
                    next_id(n, out[0], nx);
                    int b1 = src_edge(at(nx), it->first, out[0] < 4 ? 0 : 1);
                    if (b1 == -1) continue;
                    const int d1 = b1 % 2;
                    b1 >>= 2;
                    next_id(n, out[1], nx2);
                    int b2 = src_edge(at(nx2), it->first, out[1] < 4 ? 0 : 1);
                    if (b2 == -1) continue;
                    const int d2 = b2 % 2;
                    // quirk kept: the neighbours' multiplicities are read at this
                    // node's out-slot indices, not at their back slots
                    if ((at(nx2).m[out[1]] * 0.5 > n.m[L] || at(nx).m[out[0]] * 0.5 > n.m[L]) && d2 != d1 &&
                        n.arcs[L].size() < 5)
                        clear_slot(n, L);  // a short bubble loop
                    if (d2 == d1 && d1 == (L > 3 ? 0 : 1)) {
                        const int O = out[0];
                        CeNode &N = at(nx);
                        N.m[b1] = avg2(N.m[b1], N.arcs[b1].size(), n.m[L], n.arcs[L].size());
                        N.arcs[b1] += n.arcs[L];
                        n.m[L] = avg2(n.m[L], n.arcs[L].size(), n.m[O], n.arcs[O].size());
                        n.arcs[L] += n.arcs[O];
                        clear_slot(n, O);
                    }
                }
            }
            // the rest of the reference's loop body only prints (graph.cpp:2452-2474)
        }
    }

    // masterMultipleEdgeBubbleRemoval (graph.cpp:2480-2729): two edges of a
    // degree-3 node that reach the same neighbour
    void multi_edge_bubbles() {
        for (CeTable::iterator it = t_.begin(); it != t_.end(); ++it) {
            CeNode &n = it->second;
            if (!n.live) continue;
            for (int i = 0; i < 8; i++) {
                if (n.arcs[i].empty()) continue;
                unsigned long long ni, nj;
                const int Di = next_id(n, i, ni) == -1 ? 0 : 1;
                for (int j = i + 1; j < 8; j++) {
                    if (n.arcs[j].empty()) continue;
                    const int Dj = next_id(n, j, nj) == -1 ? 0 : 1;
                    if (ni != nj) continue;
                    int degree = 0;
                    for (int q = 0; q < 8; q++) degree += !n.arcs[q].empty();
                    if (degree != 3) continue;
                    // the third edge of this node and the last edge of the
                    // neighbour that does not come back here
                    unsigned long long oni = 0, onj = 0, tmp;
                    int oi = -1, oj = -1;
                    for (int q = 0; q < 8; q++) {
                        if (n.arcs[q].empty()) continue;
                        next_id(n, q, tmp);
                        if (tmp != ni) oni = tmp, oi = q;
                    }
                    {
                        CeNode &N = at(ni);
                        for (int q = 0; q < 8; q++) {
                            if (N.arcs[q].empty()) continue;
                            next_id(N, q, tmp);
                            if (tmp != it->first) onj = tmp, oj = q;
                        }
                    }
                    int ri = src_edge(at(ni), it->first, i < 4 ? 0 : 1, n.arcs[i], n.m[i], 1 - Di);
                    int rj = src_edge(at(nj), it->first, j < 4 ? 0 : 1, n.arcs[j], n.m[j], 1 - Dj);
                    if (ri == -1 || rj == -1) continue;
                    ri >>= 2;
                    rj >>= 2;
                    int ro = src_edge(at(oni), it->first, oi < 4 ? 0 : 1);
                    if (ro == -1) continue;
                    ro >>= 2;
                    if (Di != Dj) {
                        if (!((i < 4 && j > 3) || (i > 3 && j < 4))) continue;
                        // both branches of graph.cpp:2585-2641 do the same: the
                        // two parallel edges are chained into one path
                        if (oi < 4) {
                            CeNode &O = at(oni);
                            O.arcs[ro] = O.arcs[ro] + n.arcs[j] + at(ni).arcs[ri];
                            n.arcs[oi] = n.arcs[i] + at(nj).arcs[rj] + n.arcs[oi];
                            clear_slot(n, i);
                            clear_slot(at(ni), ri);
                        } else {
                            n.arcs[oi] = n.arcs[j] + at(ni).arcs[ri] + n.arcs[oi];
                            CeNode &O = at(oni);
                            O.arcs[ro] = O.arcs[ro] + n.arcs[i] + at(nj).arcs[rj];
                            clear_slot(n, j);
                            clear_slot(at(nj), rj);
                        }
                    } else if ((i < 4 && j < 4 && oi > 3) || (i > 3 && j > 3 && oi < 4)) {
                        // a bubble: the weaker edge goes, its path is kept aside
                        std::string path;
                        if (n.m[i] > n.m[j]) {
                            path = j < 4 ? kmer_str(n.id) + n.arcs[j] : kmer_str(nj) + at(nj).arcs[rj];
                            clear_slot(n, j);
                            clear_slot(at(nj), rj);
                        } else {
                            path = i < 4 ? kmer_str(n.id) + n.arcs[i] : kmer_str(ni) + at(ni).arcs[ri];
                            clear_slot(n, i);
                            clear_slot(at(ni), ri);
                        }
                        removed.push_back(path);
                    }
                    (void)oj;
                    (void)onj;
                }
            }
        }
    }

    // masterGraphStatistic (graph.cpp:2730-2799): counts only, but it looks up
    // the neighbour reached twice, which can add an empty node, and it divides
    // the coverage by the edge bases
    void statistic() {
        unsigned long long bases = 0;
        for (CeTable::iterator it = t_.begin(); it != t_.end(); ++it) {
            CeNode &n = it->second;
            if (!n.live) continue;
            for (int i = 0; i < 8; i++) bases += n.arcs[i].size();
            unsigned long long nx[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int i = 0; i < 8; i++)
                if (!n.arcs[i].empty()) next_id(n, i, nx[i]);
            for (int i = 0; i < 8; i++)
                for (int j = i + 1; j < 8; j++)
                    if (nx[i] == nx[j] && nx[i] != 0) at(nx[i]);
        }
        if (bases == 0) divide_by_zero = true;
    }

    // masterRemoveCrossEdge (graph.cpp:2892-3163): an edge between two nodes of
    // degree 3 whose other two edges leave on opposite sides is split into two
    // paths, pairing the weaker edges and the stronger edges
    void cross_edges() {
        for (CeTable::iterator it = t_.begin(); it != t_.end(); ++it) {
            CeNode &n = it->second;
            if (!n.live) continue;
            int pos = 0, neg = 0;
            for (int i = 0; i < 8; i++) {
                if (n.arcs[i].empty()) continue;
                if (i < 4) pos++;
                else neg++;
            }
            if (pos + neg != 3 || pos == 0 || neg == 0) continue;
            // the single edge on the lone side; quirk kept: nx and next keep the
            // values of the last slot examined even when that slot's back edge
            // was not found and an earlier slot set index
            unsigned long long nx = 0, id;
            int index = -1, next = -1;
            for (int i = 0; i < 8; i++) {
                if (pos == 1 && i > 3) continue;
                if (pos == 2 && i < 4) continue;
                if (n.arcs[i].empty()) continue;
                next_id(n, i, nx);
                next = src_edge(at(nx), it->first, i < 4 ? 0 : 1);
                if (next == -1) continue;
                next >>= 2;
                index = i;
            }
            if (index == -1) continue;
            if (it->first == nx) continue;
            CeNode &N = at(nx);
            int total = 0, outer = 0;
            for (int i = 0; i < 8; i++) {
                if (!N.arcs[i].empty()) total++;
                if (next < 4 && i < 4) continue;
                if (next > 3 && i > 3) continue;
                if (!N.arcs[i].empty()) outer++;
            }
            if (total != 3 || outer != 2) continue;
            int li[2], ri[2], tk = 0;
            for (int i = 0; i < 8 && tk < 2; i++) {
                if (pos == 1 && i < 4) continue;
                if (pos == 2 && i > 3) continue;
                if (n.arcs[i].empty()) continue;
                next_id(n, i, id);
                if (src_edge(at(id), it->first, i < 4 ? 0 : 1) == -1) continue;
                li[tk++] = i;
            }
            if (tk != 2) continue;
            tk = 0;
            for (int i = 0; i < 8 && tk < 2; i++) {
                if (next < 4 && i < 4) continue;
                if (next > 3 && i > 3) continue;
                if (N.arcs[i].empty()) continue;
                next_id(N, i, id);
                if (src_edge(at(id), nx, i < 4 ? 0 : 1) == -1) continue;
                ri[tk++] = i;
            }
            if (tk != 2) continue;
            if (n.m[li[0]] > n.m[li[1]]) std::swap(li[0], li[1]);
            if (N.m[ri[0]] > N.m[ri[1]]) std::swap(ri[0], ri[1]);
            // four back edges extended through the cross edge; the new
            // multiplicity is (m + m1 + m2) / new length, not a weighted mean
            bool ok = true;
            for (int s = 0; s < 4 && ok; s++) {
                const bool left = (s % 2) == 0;
                const int pair = s / 2;
                if (!left && next < 0) {  // the reference reads the slot before arcs[0] here
                    ok = false;
                    break;
                }
                const CeNode &from = left ? n : N;
                const unsigned long long from_id = left ? it->first : nx;
                const int slot = left ? li[pair] : ri[pair];
                next_id(from, slot, id);
                int back = src_edge(at(id), from_id, slot < 4 ? 0 : 1);
                if (back == -1) {
                    ok = false;
                    break;
                }
                back >>= 2;
                CeNode &D = at(id);
                const std::string add = left ? n.arcs[index] + N.arcs[ri[pair]] : N.arcs[next] + n.arcs[li[pair]];
                const int m1 = left ? n.m[index] : N.m[next];
                const int m2 = left ? N.m[ri[pair]] : n.m[li[pair]];
                D.arcs[back] += add;
                D.m[back] = (unsigned char)((size_t)((int)D.m[back] + m1 + m2) / D.arcs[back].size());
            }
            if (!ok) continue;
            n.live = false;
            N.live = false;
        }
    }

    // printContigs / masterPrintContigs selection (graph.cpp:1139-1202,
    // 3240-3318): every edge of at least Contig_Length - k bases, from the end
    // whose k-mer is not below the one it reaches
    template <typename F>
    void contigs(F &&emit) const {
        for (CeTable::const_iterator it = t_.begin(); it != t_.end(); ++it) {
            const CeNode &n = it->second;
            if (!n.live) continue;
            for (int i = 0; i < 8; i++) {
                if (n.arcs[i].empty()) continue;
                if ((long)n.arcs[i].size() < kContigLength - k_) continue;
                unsigned long long id;
                next_id(n, i, id);
                if (n.id < id) continue;
                emit(it->first, i, n, descriptor(n, i));
            }
        }
    }

    int k_;
    CeTable &t_;
};

}  // namespace

int contig_extension(int k, const HostCompacted &g, CeFiles &out) {
    out = CeFiles();
    CeTable table(g.node_buckets ? g.node_buckets : 1);
    // recvBufDistNodeGraph: contigGraph.txt order, operator[] per line
    for (size_t r = 0; r < g.key.size(); r++) {
        CeNode nd;
        nd.id = g.key[r];
        nd.live = true;
        for (int s = 0; s < 8; s++) {
            const uint32_t L = g.len[8 * r + s];
            if (L) nd.arcs[s].assign(g.pool.data() + g.off[8 * r + s], L);
            nd.m[s] = g.mult[8 * r + s];
        }
        table[nd.id] = nd;
    }
    Ce ce(k, table);
    ce.simplify();
    ce.contigs([&](unsigned long long, int, const CeNode &, const std::string &seq) {
        out.noce += ">proc0: ";
        out.noce += seq;
        out.noce += '\n';
    });
    for (int round = 0; round < 4; round++) {
        ce.tips();
        ce.simplify();
        ce.low_coverage();
        if (ce.divide_by_zero) return SWAP_E_EMPTY;
        ce.simplify();
        ce.loop_bubbles();
        ce.multi_edge_bubbles();
        ce.simplify();
        ce.statistic();
        if (ce.divide_by_zero) return SWAP_E_EMPTY;
    }
    ce.cross_edges();
    char head[200];
    int count = 0;
    ce.contigs([&](unsigned long long key, int slot, const CeNode &n, const std::string &seq) {
        count++;
        std::snprintf(head, sizeof head, ">contig_NodeID%llu_Index%d_Len%d_Mul%d_%d\n", key, slot, (int)seq.size(),
                      (int)n.m[slot], count);
        out.ce += head;
        out.ce += seq;
        out.ce += '\n';
    });
    int removed = 0;
    for (const std::string &p : ce.removed) {
        if ((long)p.size() <= kContigLength) continue;
        std::snprintf(head, sizeof head, ">contig_RemovedNo%d\n", removed++);
        out.ce += head;
        out.ce += p;
        out.ce += '\n';
    }
    // printJungGraph (graph.cpp:3320-3350)
    for (CeTable::const_iterator it = table.begin(); it != table.end(); ++it) {
        const CeNode &n = it->second;
        if (!n.live) continue;
        for (int i = 0; i < 8; i++) {
            if (n.arcs[i].empty()) continue;
            unsigned long long id;
            ce.next_id(n, i, id);
            std::snprintf(head, sizeof head, "%llu\t%llu\t%llu\n", it->first, id, (unsigned long long)n.arcs[i].size());
            out.jung_arc += head;
            std::snprintf(head, sizeof head, "%llu\t%llu\t%llu\n", it->first, id, (unsigned long long)n.m[i]);
            out.jung_mul += head;
        }
    }
    return 0;
}

// contigGraph.txt bytes -> compacted graph, read the way recvBufDistNodeGraph
// (graph.cpp:845-896) reads them: a decimal id, spaces, 8 fields each ended by
// '#' and one more character, 8 decimal multiplicities separated by spaces,
// then any number of newlines.
int parse_contig_graph(const char *text, size_t n, uint64_t survivors, HostCompacted &g) {
    g = HostCompacted();
    size_t p = 0;
    auto at = [&](size_t q) -> char { return q < n ? text[q] : '\0'; };
    while (p < n && at(p) != '\0') {
        unsigned long long id = 0;
        while (at(p) >= '0' && at(p) <= '9') id = id * 10 + (unsigned long long)(at(p++) - '0');
        while (at(p) == ' ') p++;
        g.key.push_back(id);
        for (int s = 0; s < 8; s++) {
            size_t e = p;
            while (e < n && text[e] != '#') e++;
            if (e >= n) return SWAP_E_FORMAT;
            g.off.push_back(g.pool.size());
            g.len.push_back((uint32_t)(e - p));
            g.pool.insert(g.pool.end(), text + p, text + e);
            p = e + 2;
        }
        for (int s = 0; s < 8; s++) {
            int m = 0;
            while (at(p) >= '0' && at(p) <= '9') m = m * 10 + (at(p++) - '0');
            g.mult.push_back((uint8_t)m);
            while (at(p) == ' ') p++;
        }
        while (at(p) == '\n') p++;
    }
    g.noce_order.resize(g.key.size());
    for (size_t i = 0; i < g.key.size(); i++) g.noce_order[i] = (uint32_t)i;
    g.node_buckets = table_buckets(survivors, false);
    return 0;
}

}  // namespace swapk

extern "C" int64_t swap_contig_extension(int k, uint64_t survivors, const char *contig_graph, size_t n, int which,
                                         char *buf, size_t cap) {
    if (k < 9 || k > 31 || (k % 2) == 0 || (!contig_graph && n)) return SWAP_E_INVAL;
    if (which != SWAP_FILE_NOCE && which != SWAP_FILE_CE_CONTIGS && which != SWAP_FILE_JUNG_ARC &&
        which != SWAP_FILE_JUNG_MUL)
        return SWAP_E_INVAL;
    swapk::HostCompacted g;
    int rc = swapk::parse_contig_graph(contig_graph, n, survivors, g);
    if (rc) return rc;
    swapk::CeFiles f;
    rc = swapk::contig_extension(k, g, f);
    if (rc) return rc;
    const std::string &out = which == SWAP_FILE_NOCE        ? f.noce
                             : which == SWAP_FILE_CE_CONTIGS ? f.ce
                             : which == SWAP_FILE_JUNG_ARC   ? f.jung_arc
                                                             : f.jung_mul;
    if (buf) {
        if (cap < out.size()) return SWAP_E_RANGE;
        std::memcpy(buf, out.data(), out.size());
    }
    return (int64_t)out.size();
}
