// CPU oracle for the k-mer graph build + k-molecule merge path.
//
// TEST INFRASTRUCTURE ONLY.  This is a from-scratch restatement of the
// single-rank (np=1) semantics of SWAP-Assembler 2's path, used by tests/, the
// smoke check and the CPU-baseline leg of the benchmark as the checker.  The
// product (libswaphip.so and the swap driver) never links or calls it.
//
// Pinning (what checks each part):
//   parse / encode / canonical / edge bits / counts / cutoff
//       closed forms (max-canonical, (k+1)-mer occurrence counts saturating at
//       255), brute force on tiny inputs, and the unmodified reference run at
//       np>=3 (kmerGraph.txt, compared as sorted lines) on committed fixtures.
//   compaction topology + edge strings + noCEcontigs.fasta
//       the reference run at np>=3 (noCEcontigs set), textbook unitig
//       compaction on random genomes (the genome itself comes back).
//   merged multiplicities (order-dependent integer averages) and the
//       reference's output line order: parity unpinned by a reference run --
//       the reference deadlocks at np=1 in this MPI build (blocking self-send,
//       kmerGraph.cpp:121-122) and its multi-rank merge is order-racy.  Checked
//       by invariants only (ends of one unitig agree, min<=m<=max of the unit
//       edges, exact on hand-worked chains).
#pragma once
#include <cstdint>
#include <cstddef>

extern "C" {

typedef struct oracle_graph oracle_graph;

// Runs the whole single-rank path on an in-memory FASTA buffer.
//   flags: bit0 = stop after the k-mer graph (like -s, no merge)
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
