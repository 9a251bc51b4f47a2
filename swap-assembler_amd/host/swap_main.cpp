// swap -- drop-in driver for the k-mer graph build + k-molecule merge path.
//
//   swap -k <kmerlength> -c <cutoff> -i <reads.fa> -o <outdir> [-s] [-d] [-p] [-h] [-v]
//        [--device N] [--hash-order]
//
// Same flags, messages and output files as the reference's main
// (graph.cpp:3443-3698, sequence.cpp:4-183), in its branch order: -p writes
// noCEcontigs.fasta in the performance-mode format (stringContigs,
// graph.cpp:3591-3611) from the table the merge left -- with -s, an unmerged
// table, so the file is empty -- and nothing else; otherwise -s writes
// kmerGraph.txt and stops after the k-mer graph; otherwise the graph is merged,
// contigGraph.txt is written with -d, then noCEcontigs.fasta, the contig
// extension on the host and CEContig.fasta (plus JungGraph_arc.txt /
// JungGraph_mul.txt with -j).  logtime.txt has the reference's lines for each
// mode.  Unlike the reference, errors exit with status 1.
#include <sys/resource.h>
#include <sys/stat.h>
#include <sys/types.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "swap_hip.h"

static void usage(const char *argv0) {
    printf("Usage: %s -k <kmerlength> -c <cutoff Threshod> -i <input Fasta data> -o <output data directory>\n", argv0);
}

int main(int argc, char **argv) {
    int k = 23, c = 5, device = 0;  // sequence.cpp:8-9 defaults
    std::string in, out;
    bool kmer_graph = false, dist_graph = false, hash_order = false, perf = false, jung = false;
    for (int i = 1; i < argc; i++) {
        const char *a = argv[i];
        if (a[0] != '-') { usage(argv[0]); return 1; }
        if (!strcmp(a, "--device") && i + 1 < argc) { device = atoi(argv[++i]); continue; }
        if (!strcmp(a, "--hash-order")) { hash_order = true; continue; }
        switch (a[1]) {
            case 'h': usage(argv[0]); return 0;
            case 'v': printf("SWAP-Assembler version 0.2 (gfx950 k-mer graph path)\n"); return 0;
            case 'k':
                if (i + 1 == argc || sscanf(argv[i + 1], "%d", &k) != 1 || k < 19 || k > 31 || k % 2 != 1) {
                    printf("Error: -k needs a value (odd number between 18 and 32) for the kmer size\n");
                    usage(argv[0]);
                    return 1;
                }
                i++;
                break;
            case 'c':
                if (i + 1 == argc || sscanf(argv[i + 1], "%d", &c) != 1) {
                    printf("Error: -c needs a value for the cutoff threshold\n");
                    usage(argv[0]);
                    return 1;
                }
                i++;
                break;
            case 'i':
                if (i + 1 == argc) { printf("Error: -i needs a path for the file in fasta format\n"); usage(argv[0]); return 1; }
                in = argv[++i];
                break;
            case 'o':
                if (i + 1 == argc) { printf("Error: -o needs a directory name used for the output data\n"); usage(argv[0]); return 1; }
                out = argv[++i];
                break;
            case 's': kmer_graph = true; break;
            case 'd': dist_graph = true; break;
            case 'p': perf = true; break;
            case 'j': jung = true; break;
            default: break;
        }
    }
    if (in.empty() || out.empty()) { usage(argv[0]); return 1; }
    if (mkdir(out.c_str(), S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH) == -1) {
        printf("Error: Creating directory %s\n", out.c_str());
        return 1;
    }
    printf("Runing command: %s -k %d -c %d -i %s -o %s\n", argv[0], k, c, in.c_str(), out.c_str());
    // logtime.txt keeps the reference's lines (graph.cpp:3458-3694): the ones it
    // writes through MPIEnviroment::print carry "Proc:0 [Mem=<ru_maxrss> MB] ->"
    // and end in a blank line (mympi.cpp:30-40); times here are wall-clock
    // seconds of this driver's phases.
    FILE *log = fopen((out + "/logtime.txt").c_str(), "w");
    auto t0 = std::chrono::steady_clock::now();
    auto since = [&](std::chrono::steady_clock::time_point t) {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
    };
    auto logp = [&](const std::string &msg) {
        if (!log) return;
        struct rusage ru;
        getrusage(RUSAGE_SELF, &ru);
        fprintf(log, "Proc:0 [Mem=%ld MB] ->%s\n\n", ru.ru_maxrss, msg.c_str());
    };
    auto fmt = [](const char *f, double a, double b = 0, double c2 = 0, double d = 0, double e = 0) {
        char t[256];
        snprintf(t, sizeof t, f, a, b, c2, d, e);
        return std::string(t);
    };
    auto fail = [&](int rc, swap_ctx *ctx) {
        printf("Error: %s (%s)\n", swap_strerror(rc), ctx ? swap_last_error(ctx) : "");
        if (log) fclose(log);
        if (ctx) swap_destroy(ctx);
        return 1;
    };
    logp(fmt("Time spend in Geting Reads %.5f", since(t0)));

    auto tr = std::chrono::steady_clock::now();
    FILE *f = fopen(in.c_str(), "rb");
    if (!f) { printf("Proc:0 Unable to open file \"%s\"\n", in.c_str()); return 1; }
    fseek(f, 0, SEEK_END);
    const long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    std::vector<char> buf((size_t)(n > 0 ? n : 0));
    if (n > 0 && fread(buf.data(), 1, (size_t)n, f) != (size_t)n) { printf("read error on %s\n", in.c_str()); return 1; }
    fclose(f);
    const double t_read = since(tr);

    swap_cfg cfg;
    cfg.k = k;
    cfg.c = c;
    cfg.device = device;
    // -s builds the k-mer graph only (the reference skips the merge, graph.cpp:3576)
    cfg.flags = (kmer_graph ? SWAP_FLAG_KMER_GRAPH : 0u) | (hash_order ? SWAP_FLAG_HASH_ORDER : 0u);
    swap_ctx *ctx = nullptr;
    int rc = swap_create(&cfg, &ctx);
    if (rc) return fail(rc, nullptr);
    auto t1 = std::chrono::steady_clock::now();
    rc = swap_build(ctx, buf.data(), buf.size());
    if (rc) return fail(rc, ctx);
    const double t_build = since(t1);
    swap_stats st;
    swap_get_stats(ctx, &st);
    printf("Proc 0: tot_kmer_num=%llu reads=%llu distinct=%llu survivors=%llu\n", (unsigned long long)st.kmers,
           (unsigned long long)st.reads, (unsigned long long)st.distinct, (unsigned long long)st.survivors);
    logp(fmt("Time spend in I/O part in MPIcontrol %.5f", t_read));
    logp(fmt("Time spend in locate function %.5f", 0.0));
    logp(fmt("Time spend in read function %.5f", t_read));
    // the count kernel applies the cutoff as it empties each table: its device
    // time (and the survivor keys pass) is reported as cutoffGraph
    const double t_cut = std::min(t_build, (st.phase_ms[4] + st.phase_ms[5] + st.phase_ms[15]) / 1e3);
    logp(fmt("Time spend in Constructing Graph %.5f (cutTimetot=%.5f, communication time=%.5f, "
             "communicationtimetot=%.5f)",
             t_build - t_cut, t_cut, 0.0, 0.0));
    logp(fmt("Time spend in cutoffGraph %.5f", t_cut));

    double t_nodes = 0, t_simplify = 0;
    if (!kmer_graph) {
        auto t2 = std::chrono::steady_clock::now();
        rc = swap_merge(ctx);
        if (rc) return fail(rc, ctx);
        const double t_merge = since(t2);
        swap_get_stats(ctx, &st);
        // the node table of the reference (buildKmoleculeGraph) is the survivor index here
        t_nodes = std::min(t_merge, st.phase_ms[8] / 1e3);
        t_simplify = t_merge - t_nodes;
    }
    logp(fmt("Time spend in Constructing DistGraph %.5f", t_nodes));
    logp(fmt("Time spend in Simplify DistGraph %.5f (workTime=%.5f(%.5f), totalTime=%.5f(%.5f))", t_simplify,
             t_simplify, t_simplify, t_simplify, t_simplify));

    auto t3 = std::chrono::steady_clock::now();
    if (perf) {
        // stringContigs on the table the merge left (graph.cpp:3591-3611); with -s
        // that table was never merged and holds no edge long enough for a contig
        if (kmer_graph) {
            FILE *e = fopen((out + "/noCEcontigs.fasta").c_str(), "wb");
            if (!e) return fail(SWAP_E_IO, ctx);
            fclose(e);
        } else {
            int64_t w = swap_write_file(ctx, SWAP_FILE_PERF_CONTIGS, (out + "/noCEcontigs.fasta").c_str());
            if (w < 0) return fail((int)w, ctx);
        }
        if (log) {
            fprintf(log, "Time spend in printDistGraph %.5f\n", since(t3));
            fprintf(log, "Time spend in All step %.5f\n", since(t0));
            fclose(log);
        }
        swap_destroy(ctx);
        return 0;
    }
    if (dist_graph && !kmer_graph) {
        int64_t w = swap_write_file(ctx, SWAP_FILE_CONTIG_GRAPH, (out + "/contigGraph.txt").c_str());
        if (w < 0) return fail((int)w, ctx);
    }
    if (log) fprintf(log, "Time spend in printDistGraph %.5f\n", since(t3));
    auto t4 = std::chrono::steady_clock::now();
    if (kmer_graph) {
        // printKmerGraph, then the reference returns without the last log line
        int64_t w = swap_write_file(ctx, SWAP_FILE_KMER_GRAPH, (out + "/kmerGraph.txt").c_str());
        if (w < 0) return fail((int)w, ctx);
        if (log) fclose(log);
        swap_destroy(ctx);
        return 0;
    }
    if (log) fprintf(log, "Time spend in reading DistGraph  %.5f\n", since(t4));
    int64_t w = swap_write_file(ctx, SWAP_FILE_NOCE, (out + "/noCEcontigs.fasta").c_str());
    if (w < 0) return fail((int)w, ctx);
    printf("|proc:0 %llu nodes in distGraph\n", (unsigned long long)st.kept);
    // the contig extension runs when its first file is formatted
    w = swap_write_file(ctx, SWAP_FILE_CE_CONTIGS, (out + "/CEContig.fasta").c_str());
    if (w < 0) return fail((int)w, ctx);
    if (log) fprintf(log, "Time spend in Contig Extension %.5f\n", since(t4));
    if (jung) {
        w = swap_write_file(ctx, SWAP_FILE_JUNG_ARC, (out + "/JungGraph_arc.txt").c_str());
        if (w >= 0) w = swap_write_file(ctx, SWAP_FILE_JUNG_MUL, (out + "/JungGraph_mul.txt").c_str());
        if (w < 0) return fail((int)w, ctx);
    }
    if (log) {
        fprintf(log, "Time spend in All step %.5f\n", since(t0));
        fclose(log);
    }
    swap_destroy(ctx);
    return 0;
}
