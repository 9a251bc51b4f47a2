// Host side of the path: the reference's table orders, the output writers and
// the merge driver.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "swap_device.h"
#include "swap_hip.h"

namespace swapk {

// Hash of graph.cpp:29-32: (x >> 32) ^ (size_t)(int)(x & 0xFFFFFFFF).
inline uint64_t ref_hash(uint64_t x) {
    return (x >> 32) ^ (uint64_t)(int64_t)(int32_t)(uint32_t)(x & 0xFFFFFFFFull);
}

// Iteration order of a __gnu_cxx::hash_map (libstdc++ backward/hashtable.h)
// after inserting n_total keys one by one into an empty table of 193 buckets,
// restricted to the subset of keys given here.  rank[i] is the insertion index
// (0-based among all n_total inserts) of key[i]; extra_resize asks for one more
// resize check after the last insert (an operator[] lookup that saw
// num_elements == bucket_count).  Returns the subset's indices in iteration
// order and the final bucket count.
std::vector<uint32_t> table_order(const std::vector<uint64_t> &key, const std::vector<uint64_t> &rank,
                                  uint64_t n_total, bool extra_resize, uint64_t *bucket_count);

// Bucket count of the table after n inserts.
uint64_t table_buckets(uint64_t n_total, bool extra_resize);
// Insert ranks at which the table starts a new bucket count (first entry 0).
std::vector<uint64_t> table_epoch_starts(uint64_t n_total, bool extra_resize);

// The reference's visit order of the merge (the nodes table, graph.cpp:1123-1133)
// -- three table builds chained (kmolecules, cutoff copy, nodes).  first[i] is
// the first-occurrence index of survivor i; epoch_first[e] the first-occurrence
// index of the distinct key whose insert started growth epoch e of the k-mer
// table (D distinct keys in all).  Returns survivor indices in visit order and
// the nodes table bucket count.
std::vector<uint32_t> visit_order(const std::vector<uint64_t> &key, const std::vector<uint64_t> &first,
                                  const std::vector<uint64_t> &epoch_first, uint64_t distinct, bool tail_lookup,
                                  uint64_t *node_buckets);

// Order after the gather (graph.cpp:3614-3615, 891): live nodes re-inserted, in
// visit order, into the cleared nodes table (bucket count kept).
std::vector<uint32_t> gather_order(const std::vector<uint64_t> &key, const std::vector<uint32_t> &visit_pos_sorted,
                                   uint64_t node_buckets);

void format_kmer_graph(int k, const HostKmerGraph &g, const std::vector<uint32_t> &order, std::string &out);
void format_contig_graph(const HostCompacted &c, std::string &out);
void format_noce(int k, const HostCompacted &c, std::string &out);
void format_perf_contigs(int k, const HostCompacted &c, std::string &out);
// Contig extension of the merged graph (the reference's rank-0 finisher,
// clean-up passes and writers, graph.cpp:2089-3350); 0, or SWAP_E_EMPTY where
// the reference divides by zero (no edge base left).
int contig_extension(int k, const HostCompacted &g, CeFiles &out);

// Runs the merge (device kernels; cycles and fold-backs replayed on the host).
// When nothing needs the host replay the results stay on the device and
// *materialized is false; materialize_merge() copies and assembles them.
int host_merge(Pipeline &p, const swap_cfg &cfg, bool *materialized, HostCompacted &out);
int materialize_merge(Pipeline &p, const swap_cfg &cfg, HostCompacted &out);
// The host half of materialize_merge: the replay of cycles and fold-backs over
// the merge results fetched from the device (S survivors), and the compacted
// graph in file order; *n_kept receives its node count.
int replay_merge(const MergeDev &md, size_t S, const swap_cfg &cfg, HostCompacted &out, uint64_t *n_kept);

}  // namespace swapk
