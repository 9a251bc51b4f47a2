/*
 * swap_hip.h -- C ABI of the MI355X (gfx950) k-mer graph construction and
 * k-molecule merging path of SWAP-Assembler 2.
 *
 * Plain C types only: callers bind it from C, C++ (the swap driver), Python
 * (ctypes) or any FFI.  Errors are negative return codes (never exit()); the
 * text of the last error of a context is available from swap_last_error().
 * A context owns its device memory; host buffers belong to the caller.  One
 * context per host thread: the API is thread-compatible, not thread-safe.
 *
 * Reference interfaces replaced (file:line in SWAP-Assembler 2):
 *   swap_create / swap_destroy   parameter::getParameters (sequence.cpp:121-183)
 *                                for k, c and the flags; new distNodeGraph /
 *                                new kmerGraph in main (graph.cpp:3474-3477)
 *   swap_build                   distNodeGraph::constructDistNodeGraph
 *                                (graph.cpp:1049-1090) with the
 *                                kmerGraph::constructKmerGraph /
 *                                distributeKmerGraph loop (kmerGraph.cpp:95-310,
 *                                337-432), then distNodeGraph::cutoffGraph
 *                                (graph.cpp:1092-1114)
 *   swap_merge                   buildKmoleculeGraph (graph.cpp:1117-1137) +
 *                                simplifyDistNodeGraph (graph.cpp:1392-1566) +
 *                                masterSimplifyNodeGraph (graph.cpp:2089-2192)
 *   swap_export_kmer_graph       the kmolecules table after cutoffGraph
 *                                (graph.cpp:353, 1112)
 *   swap_export_compacted        the nodes table after the merge (graph.cpp:354)
 *   swap_build_batch /           constructDistNodeGraph + cutoffGraph of many
 *   swap_batch_take              samples in one build (the reference builds one
 *                                sample per run); a taken sample is as after
 *                                swap_build_device
 *   swap_reconfigure, swap_server_*  no counterpart (many samples per process)
 *   swap_format_file /           printKmerGraph (graph.cpp:3352-3440),
 *   swap_write_file              gatherDistNodeGraph + the contigGraph.txt write
 *                                (graph.cpp:576-696, 3617-3622), printContigs
 *                                (graph.cpp:1139-1202), stringContigs (-p,
 *                                graph.cpp:3166-3238), and the contig extension
 *                                with its writers (graph.cpp:2089-3350)
 */
#ifndef SWAP_HIP_H
#define SWAP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* return codes */
#define SWAP_OK 0
#define SWAP_E_INVAL (-1)     /* bad argument (k not odd in [9,31], a merge with k < 19, NULL pointer) */
#define SWAP_E_EMPTY (-2)     /* empty input buffer */
#define SWAP_E_FORMAT (-3)    /* no '>' in the input (reference: assert at kmerGraph.cpp:114) */
#define SWAP_E_TOOBIG (-4)    /* more than 2^32-16 k-mers in one build; for swap_assemble_static, an input
                                 beyond its static scratch */
#define SWAP_E_DEVICE (-5)    /* HIP runtime error (see swap_last_error) */
#define SWAP_E_NOGPU (-6)     /* no usable gfx950 device */
#define SWAP_E_IO (-7)        /* file could not be written */
#define SWAP_E_STATE (-8)     /* call out of order (e.g. merge before build) */
#define SWAP_E_RANGE (-9)     /* caller buffer too small */

/* swap_cfg.flags */
#define SWAP_FLAG_KMER_GRAPH 1u  /* -s: stop after the k-mer graph, no merge */
#define SWAP_FLAG_HASH_ORDER 2u  /* merge in a GPU-native (hash) visit order instead of
                                    replaying the reference's table order; topology,
                                    edge strings and contigs are the same, merged
                                    multiplicities follow the different order */
#define SWAP_FLAG_TABLE_SPILL 4u /* testing aid: every count partition is recounted in the
                                    device-memory table that partitions too large for the
                                    on-chip table use; same results, slower */
#define SWAP_FLAG_WIDE_PARTITIONS 8u /* testing aid: at least 2^18 count partitions and
                                        one visit-order binning pass per table, the
                                        paths inputs of billions of k-mers take; same results */
#define SWAP_FLAG_SINGLE_FILL 16u /* testing aid: one count partition per fill of the on-chip
                                     table, as on large inputs, also where small inputs'
                                     partitions would share fills; same results */
#define SWAP_FLAG_COARSE_RENUMBER 32u /* testing aid: the merge numbers the survivors by coarse
                                         first-occurrence bins only, the order inputs of more
                                         than 2^31 k-mers get; same results */

typedef struct swap_cfg {
    int k;          /* odd k-mer length, 9..31; swap_merge needs 19..31 (the reference CLI's range) */
    int c;          /* cutoff: edge multiplicities below c are removed */
    int device;     /* HIP device ordinal */
    unsigned flags; /* SWAP_FLAG_* */
} swap_cfg;

/* Performance parameters of a build and merge (no counterpart in the reference,
 * whose container and hash are fixed, graph.cpp:20-34, 353-354).  None of them
 * changes a result, only the time it takes:
 *   part_recs     records per count partition the partition count aims at,
 *                 which sets the load of the on-chip counting table (1024..65536)
 *   fill_recs     records a shared fill of the on-chip table aims at on small
 *                 inputs (512..8192)
 *   index_load    largest load factor of a survivor-index region, percent (50..85)
 *   ruler_stride  one ruler per this many oriented nodes in the list ranking (4..256)
 *   occ_groups    workgroups per compute unit of the k-mer passes (1..4)
 *   order_bin_keys  keys per bin the visit-order sort aims at on average; a
 *                 bin is ordered in LDS, or in device memory above 6144 keys
 *                 (256..2^30)
 * Values outside the ranges are clamped. */
typedef struct swap_tuning {
    uint32_t part_recs;
    uint32_t fill_recs;
    uint32_t index_load;
    uint32_t ruler_stride;
    uint32_t occ_groups;
    uint32_t order_bin_keys;
} swap_tuning;

typedef struct swap_stats {
    uint64_t reads;        /* records with at least k bases */
    uint64_t kmers;        /* input k-mer occurrences (N) */
    uint64_t distinct;     /* canonical k-mers before cutoff (D) */
    uint64_t survivors;    /* k-molecules after cutoff (S) */
    uint64_t kept;         /* nodes after the merge */
    uint64_t overflow;     /* partitions that fell back to the global-memory table */
    uint64_t special;      /* (1,1) nodes merged on the host (cycles, fold-backs) */
    double build_ms;       /* wall time of the last swap_build (device + syncs) */
    double merge_ms;       /* wall time of the last swap_merge */
    /* device time of each phase (events on the context's stream): 0 record scan,
     * 1 2-bit pack, 2 partition histogram, 3 level-1 partition, 4 LDS count +
     * cutoff, 5 global-table count (overflow), 6 first-occurrence ranks,
     * 7 level-2 partition, 8 survivor renumbering and index, 9 list ranking,
     * 10 kept edges,
     * 11 edge strings, 12 Cartesian trees, 13 visit order, 14 classification,
     * 15 survivor keys (un-mixing) */
    double phase_ms[16];
} swap_stats;

typedef struct swap_ctx swap_ctx;

int swap_create(const swap_cfg *cfg, swap_ctx **out);
void swap_destroy(swap_ctx *ctx);

/* Build from FASTA bytes in host memory (copied to the device). */
int swap_build(swap_ctx *ctx, const char *fasta, size_t n);
/* Build from FASTA bytes already in device memory on the context's device. */
int swap_build_device(swap_ctx *ctx, const void *d_fasta, size_t n);

/* Read-ahead of a later swap_build's input, the device-side counterpart of the
 * reference's windowed reads (kmerGraph.cpp:100-156 reads the next window while
 * the previous one is processed): starts copying the n host bytes at fasta to
 * a device buffer of the context on a copy stream of its own and returns; the
 * next swap_build with the same pointer and size takes the copied bytes
 * instead of copying them, so the transfer of input i+1 overlaps the build and
 * merge of input i.  Builds consume prefetches in the order they were made; a
 * swap_build of any other input drops them.  At most two inputs may be
 * prefetched and not yet built (SWAP_E_STATE beyond).  The bytes must stay
 * valid and unchanged until the swap_build that consumes them returns (or the
 * context is destroyed). */
int swap_prefetch(swap_ctx *ctx, const char *fasta, size_t n);

/* Many samples in one build: the k-mer graphs of nsamples (1..512) samples
 * whose FASTA bytes are in device memory (d_fasta[i], n[i] bytes) are built by
 * one pass of the build kernels over all of them: one record scan, one packing,
 * two partition passes and one persistent count whose table keys carry the
 * sample, then the survivors of all samples are ordered by first occurrence
 * and every sample's growth ranks selected in one pass.  The reference builds
 * one sample per run (graph.cpp:3443-3698, constructDistNodeGraph at
 * graph.cpp:1049-1090); a sample's results equal those of its own build.
 * Needs reference order and no SWAP_FLAG_KMER_GRAPH on ctx; every sample must
 * start with '>', end with a newline and, except the last, hold whole line
 * pairs (SWAP_E_INVAL otherwise: build those samples alone); at most 2^32-16
 * k-mers in all.  The samples are then taken out one by one with
 * swap_batch_take. */
int swap_build_batch(swap_ctx *ctx, const void *const *d_fasta, const size_t *n, size_t nsamples);
/* Makes dst's state that of a build of sample i of batch's last
 * swap_build_batch alone (dst: same k, c and device, reference order): merge,
 * exports and output files follow as after swap_build_device.  Copies the
 * sample's survivors; batch keeps its results until its next build. */
int swap_batch_take(swap_ctx *batch, size_t i, swap_ctx *dst);

/* Merge the k-mer graph of the last build into unitig edges (k >= 19, else
 * SWAP_E_INVAL). */
int swap_merge(swap_ctx *ctx);

/* swap_build_device followed by swap_merge (build only for a SWAP_FLAG_KMER_GRAPH
 * context) in one call that allocates nothing: every device buffer of the call
 * is carved from one fixed scratch array per GPU (512 MiB, held by the library
 * from the first context created on that GPU until the process ends), and no
 * hipMalloc / hipFree or other allocating routine is called.  No counterpart in
 * the reference: for callers that cannot allocate device memory (work issued
 * under HIP graph capture, tight serving loops, harnesses that forbid it).
 * Inputs up to 3 MiB of FASTA; larger ones, or an input whose buffers would
 * not fit, return SWAP_E_TOOBIG.  Results are read with the usual export and
 * file calls; they live in the scratch and stay valid until the next
 * swap_assemble_static call on the same GPU from any context.  Calls that
 * overlap on one GPU race on the scratch: the caller serialises them.
 * swap_merge after this call returns SWAP_E_STATE. */
int swap_assemble_static(swap_ctx *ctx, const void *d_fasta, size_t n);

int swap_get_stats(const swap_ctx *ctx, swap_stats *out);

/* The tuning a build of input_bytes of FASTA takes from the autotuned table
 * (swap-assembler_amd/tune/swap_tune.txt beside the library's directory, or the
 * file named by the SWAP_TUNE_FILE environment variable; built-in defaults
 * without one).  Needs no device.  No counterpart in the reference. */
int swap_tuning_for(uint64_t input_bytes, swap_tuning *out);
/* Fixes the tuning of the context's next builds (NULL: back to the table).
 * Used by the autotuner; no counterpart in the reference. */
int swap_set_tuning(swap_ctx *ctx, const swap_tuning *t);
/* The tuning the context's last build used. */
int swap_get_tuning(const swap_ctx *ctx, swap_tuning *out);

/* Survivors of the cutoff sorted by key: keys[i] and 8 multiplicities
 * mult[8*i .. 8*i+7] (slot t = "X followed by base t" for t<4, "rc(X) followed
 * by base t-4" for t>=4).  Returns the number of records (pass NULL/0 to query),
 * or a negative code. */
int64_t swap_export_kmer_graph(swap_ctx *ctx, uint64_t *keys, uint8_t *mult, size_t cap);

/* Output files (byte formats of the reference):
 *   SWAP_FILE_KMER_GRAPH   kmerGraph.txt      (after a SWAP_FLAG_KMER_GRAPH build)
 *   SWAP_FILE_CONTIG_GRAPH contigGraph.txt    (after swap_merge)
 *   SWAP_FILE_NOCE         noCEcontigs.fasta  (after swap_merge)
 *   SWAP_FILE_PERF_CONTIGS noCEcontigs.fasta as the -p mode writes it (after swap_merge)
 *   SWAP_FILE_CE_CONTIGS, SWAP_FILE_JUNG_ARC, SWAP_FILE_JUNG_MUL: see below */
#define SWAP_FILE_KMER_GRAPH 0
#define SWAP_FILE_CONTIG_GRAPH 1
#define SWAP_FILE_NOCE 2
#define SWAP_FILE_PERF_CONTIGS 3 /* the -p contigs file (stringContigs, graph.cpp:3166-3238), after swap_merge */
/* after swap_merge, from the contig extension run on the host (the reference's
 * rank-0 finisher and clean-up passes, graph.cpp:2089-3350; computed once per merge) */
#define SWAP_FILE_CE_CONTIGS 4   /* CEContig.fasta (masterPrintContigs, graph.cpp:3240-3318) */
#define SWAP_FILE_JUNG_ARC 5     /* JungGraph_arc.txt (-j, printJungGraph, graph.cpp:3320-3350) */
#define SWAP_FILE_JUNG_MUL 6     /* JungGraph_mul.txt (-j) */
/* Writes the file bytes into buf; returns the byte count (buf NULL: size query). */
int64_t swap_format_file(swap_ctx *ctx, int which, char *buf, size_t cap);
/* Writes the file to path; returns bytes written or a negative code. */
int64_t swap_write_file(swap_ctx *ctx, int which, const char *path);

/* Host utility: iteration order of the reference's __gnu_cxx::hash_map (hash of
 * graph.cpp:29-32, libstdc++ bucket growth) after n_total inserts, restricted to
 * the n keys given; ranks[i] is the insertion index of keys[i].  extra_resize
 * adds the growth check of one more operator[] lookup after the last insert.
 * Writes the indices of the keys in iteration order to order_out; returns n. */
int64_t swap_table_order(const uint64_t *keys, const uint64_t *ranks, size_t n, uint64_t n_total,
                         int extra_resize, uint32_t *order_out);

/* Host utility: the reference's merge visit order (its nodes table order,
 * graph.cpp:1123-1133) for n survivors with keys[i] and first-occurrence
 * indices first[i]; epoch_first[e] is the first-occurrence index of the
 * distinct k-mer whose insert started growth epoch e of the k-mer table
 * (n_epochs entries, the first 0), distinct the number of distinct k-mers.
 * Writes survivor indices in visit order; returns n. */
int64_t swap_visit_order(const uint64_t *keys, const uint64_t *first, size_t n, const uint64_t *epoch_first,
                         size_t n_epochs, uint64_t distinct, int tail_lookup, uint32_t *order_out);
/* Insert ranks at which a table of n_total inserts grows (first entry 0);
 * returns the count written (at most cap). */
int64_t swap_table_epochs(uint64_t n_total, int extra_resize, uint64_t *starts, size_t cap);

const char *swap_strerror(int code);
const char *swap_last_error(const swap_ctx *ctx);

/* Host utility, no device: the contig extension on a contigGraph.txt (the
 * reference reads that file back the same way, recvBufDistNodeGraph,
 * graph.cpp:845-896), as the rank-0 part of its main runs it
 * (graph.cpp:3636-3682).  survivors = the k-molecule count after the cutoff
 * (it sizes the table the graph is read into, which fixes the walk order).
 * which: SWAP_FILE_NOCE, SWAP_FILE_CE_CONTIGS, SWAP_FILE_JUNG_ARC or
 * SWAP_FILE_JUNG_MUL.  Returns the byte count (buf NULL: size query). */
int64_t swap_contig_extension(int k, uint64_t survivors, const char *contig_graph, size_t n, int which, char *buf,
                              size_t cap);

/* Re-targets a context to a new k, c and flags for its next build, keeping its
 * device buffers (same device).  No counterpart in the reference (one run =
 * one parameter set); used by callers that assemble many samples. */
int swap_reconfigure(swap_ctx *ctx, const swap_cfg *cfg);

/* The compacted graph after swap_merge (the reference's nodes table after the
 * merge, graph.cpp:354): n kept nodes, in contigGraph.txt order.  Node i has key
 * keys[i]; its edge slot t (0..7, as in swap_export_kmer_graph) is the string
 * pool[edge_off[8*i+t] .. + edge_len[8*i+t]) of 'A','C','G','T' (length 0: no
 * edge) with merged multiplicity edge_mult[8*i+t].  *pool_len receives the pool
 * size.  Returns n; NULL buffers query the sizes. */
int64_t swap_export_compacted(swap_ctx *ctx, uint64_t *keys, uint64_t *edge_off, uint32_t *edge_len,
                              uint8_t *edge_mult, size_t cap_nodes, char *pool, size_t cap_pool,
                              uint64_t *pool_len);

/* ---------------------------------------------------------------------------
 * Many assemblies on one GPU (no counterpart in the reference, which assembles
 * one sample per run).  A server runs jobs on `slots` concurrent contexts (one
 * host thread and one HIP stream each; their kernels overlap on the device).
 * A job starts, in submission order, when a slot is free and the estimated
 * device footprint of the running jobs plus its own fits hbm_budget (bytes;
 * 0 = 80% of the free device memory); a job alone always starts.  Every job's
 * outputs are byte-identical to running it alone on a fresh context. */
#define SWAP_JOB_QUEUED 1
#define SWAP_JOB_RUNNING 2
#define SWAP_JOB_DEVICE_INPUT 1u /* swap_job.job_flags: fasta is device memory on the server's device */

typedef struct swap_job {
    const void *fasta;  /* FASTA bytes; must stay valid until the job ends */
    size_t n;
    int k, c;
    unsigned flags;     /* SWAP_FLAG_* */
    unsigned outputs;   /* bit (1u << SWAP_FILE_*) per file to keep: kmerGraph.txt only with
                           SWAP_FLAG_KMER_GRAPH, the others only without it */
    unsigned job_flags; /* SWAP_JOB_* */
} swap_job;

typedef struct swap_job_info {
    int status;        /* SWAP_JOB_QUEUED, SWAP_JOB_RUNNING, 0 done, or a negative code */
    double submit_ms;  /* times since the server was created */
    double start_ms;
    double end_ms;
    swap_stats stats;  /* of the job's build and merge */
} swap_job_info;

typedef struct swap_server swap_server;

int swap_server_create(int device, int slots, size_t hbm_budget, swap_server **out);
/* Queues a job; returns its id (>= 0) or a negative code. */
int64_t swap_server_submit(swap_server *s, const swap_job *job);
/* Queues n jobs, largest input first; ids[i] receives the id of jobs[i]. */
int swap_server_submit_batch(swap_server *s, const swap_job *jobs, size_t n, int64_t *ids);
/* Queues n merge jobs of one k, c and flags (reference order, no
 * SWAP_FLAG_KMER_GRAPH) whose FASTA is in device memory
 * (SWAP_JOB_DEVICE_INPUT), built together: largest input first, the jobs are
 * cut into groups of at most group_bytes of FASTA (0: 256 MiB; at most 512 MiB) and 512
 * samples; a builder thread of the server builds each group with one
 * swap_build_batch (the next group while the slots merge the previous one),
 * and the slot that runs a job takes its sample (swap_batch_take) instead of
 * building it.  A group whose batch build fails runs as ordinary jobs.
 * Outputs, statuses and the wait calls are those of ordinary jobs; each
 * job's files equal those of its own build.  No counterpart in the reference
 * (one sample per run, graph.cpp:3443-3698). */
int swap_server_submit_fused(swap_server *s, const swap_job *jobs, size_t n, size_t group_bytes, int64_t *ids);
/* Blocks until the job has ended; returns its status (0 or a negative code). */
int swap_server_wait(swap_server *s, int64_t id);
/* Blocks until every queued job has ended; returns 0 or the first job error. */
int swap_server_wait_all(swap_server *s);
int swap_server_job_info(swap_server *s, int64_t id, swap_job_info *info);
const char *swap_server_job_error(swap_server *s, int64_t id);
/* Bytes of output file `which` of an ended job (buf NULL: size query). */
int64_t swap_server_output(swap_server *s, int64_t id, int which, char *buf, size_t cap);
/* Frees the outputs kept for an ended job. */
int swap_server_release(swap_server *s, int64_t id);
/* Finishes the queued and running jobs, then frees the server. */
void swap_server_destroy(swap_server *s);

#ifdef __cplusplus
}
#endif

#endif /* SWAP_HIP_H */
