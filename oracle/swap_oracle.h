// CPU oracle for the k-mer graph build + k-molecule merge path.
//
// TEST INFRASTRUCTURE ONLY.  This is a from-scratch restatement of the
// single-rank (np=1) semantics of SWAP-Assembler 2's path, used by tests/, the
// smoke check and the CPU-baseline leg of the benchmark as the checker.  The
// product (libswaphip.so and the swap driver) never links or calls it.
//
// Pinning (what checks each part):
//   every output file of the single-rank path -- kmerGraph.txt, contigGraph.txt
//       (merged multiplicities, padding and line order included),
//       noCEcontigs.fasta and the -p file -- byte for byte against the
//       unmodified reference run as one rank (oracle/ref_swap1.cpp) on 23
//       committed inputs and the 1 Mbp / 4.6 Mbp workloads
//       (tests/golden/make_golden_np1.py, tests/test_oracle.py).  The inputs
//       include isolated cycles, self-neighbours, two-node cycles, fold-backs
//       and the late unlock of a half-locked neighbour; the planted-mistake
//       flags below must each break at least one of them.
//   parse / encode / canonical / edge bits / counts / cutoff
//       also closed forms (max-canonical, (k+1)-mer occurrence counts
//       saturating at 255) and brute force on tiny inputs.
#pragma once
#include <cstdint>
#include <cstddef>

extern "C" {

typedef struct oracle_graph oracle_graph;

// Runs the whole single-rank path on an in-memory FASTA buffer.
//   flags: bit0 = stop after the k-mer graph (like -s, no merge)
//          bit1 = PLANTED MISTAKE for tests: the unlock after a half-successful
//                 lock takes effect at once instead of at the next lock request
//          bit2 = PLANTED MISTAKE for tests: a neighbour that is both left and
//                 right is locked twice (the second lock fails)
// Returns NULL on malformed input (no '>' in the buffer) or bad k.
oracle_graph *oracle_run(const char *fasta, uint64_t n, int k, int c, int flags);
void oracle_free(oracle_graph *g);

// Counters of the run.
uint64_t oracle_num_occurrences(const oracle_graph *g);   // input k-mers (N)
uint64_t oracle_num_reads(const oracle_graph *g);         // records with >= k bases
uint64_t oracle_num_distinct(const oracle_graph *g);      // D (table size before cutoff)
uint64_t oracle_num_survivors(const oracle_graph *g);     // S (after cutoff)
uint64_t oracle_num_kept(const oracle_graph *g);          // live nodes after the merge

// Pre-cutoff table, sorted by key: keys[i], mult[8*i..8*i+8).  Returns D; writes
// at most cap records.  Only sensible for small inputs.
uint64_t oracle_table(const oracle_graph *g, uint64_t *keys, uint8_t *mult, uint64_t cap);
// Same for the survivors of the cutoff (before the merge), sorted by key.
uint64_t oracle_survivors(const oracle_graph *g, uint64_t *keys, uint8_t *mult, uint64_t cap);
// First-occurrence index (0-based input k-mer index) of each survivor, in the
// order of oracle_survivors.
uint64_t oracle_survivor_first(const oracle_graph *g, uint64_t *first, uint64_t cap);
// Index of each survivor's key among all distinct keys in first-occurrence
// order (its insertion index into the reference's k-mer table).
uint64_t oracle_survivor_rank(const oracle_graph *g, uint64_t *rank, uint64_t cap);
// First-occurrence index of the distinct key inserted rank-th into the table.
uint64_t oracle_first_of_rank(const oracle_graph *g, uint64_t rank);
// Survivor keys in the merge visit order (the reference's hash-table order).
uint64_t oracle_visit_order(const oracle_graph *g, uint64_t *keys, uint64_t cap);

// Output files, byte for byte, into caller buffers.  Each returns the byte
// count; pass buf=NULL to query the size.
//   which: 0 = kmerGraph.txt (run with flags bit0), 1 = contigGraph.txt,
//          2 = noCEcontigs.fasta, 3 = noCEcontigs.fasta as the -p mode writes it
uint64_t oracle_file(const oracle_graph *g, int which, char *buf, uint64_t cap);

}  // extern "C"
