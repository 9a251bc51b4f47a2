// swap -- drop-in driver for the k-mer graph build + k-molecule merge path.
//
//   swap -k <kmerlength> -c <cutoff> -i <reads.fa> -o <outdir> [-s] [-d] [-p] [-h] [-v]
//        [--device N] [--hash-order]
//
// Same flags, messages and output files as the reference's main
// (graph.cpp:3443-3698, sequence.cpp:4-183): -s writes kmerGraph.txt and stops
// after the k-mer graph; otherwise the graph is merged and noCEcontigs.fasta is
// written, then the contig extension runs on the host and CEContig.fasta is
// written (plus JungGraph_arc.txt / JungGraph_mul.txt with -j); contigGraph.txt
// with -d; with -p the merged graph's contigs are written in the
// performance-mode format (stringContigs, graph.cpp:3591-3611) instead and
// nothing else.  logtime.txt holds the phase timings.  Unlike the reference,
// errors exit with status 1.
#include <sys/stat.h>
#include <sys/types.h>

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
    FILE *log = fopen((out + "/logtime.txt").c_str(), "w");
    auto t0 = std::chrono::steady_clock::now();
    auto since = [&](std::chrono::steady_clock::time_point t) {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
    };

    FILE *f = fopen(in.c_str(), "rb");
    if (!f) { printf("Proc:0 Unable to open file \"%s\"\n", in.c_str()); return 1; }
    fseek(f, 0, SEEK_END);
    const long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    std::vector<char> buf((size_t)(n > 0 ? n : 0));
    if (n > 0 && fread(buf.data(), 1, (size_t)n, f) != (size_t)n) { printf("read error on %s\n", in.c_str()); return 1; }
    fclose(f);
    if (log) fprintf(log, "Time spend in Geting Reads %.5f\n", since(t0));

    swap_cfg cfg;
    cfg.k = k;
    cfg.c = c;
    cfg.device = device;
    cfg.flags = (kmer_graph ? SWAP_FLAG_KMER_GRAPH : 0u) | (hash_order ? SWAP_FLAG_HASH_ORDER : 0u);
    swap_ctx *ctx = nullptr;
    int rc = swap_create(&cfg, &ctx);
    if (rc) { printf("Error: %s\n", swap_strerror(rc)); return 1; }
    auto t1 = std::chrono::steady_clock::now();
    rc = swap_build(ctx, buf.data(), buf.size());
    if (rc) { printf("Error: %s (%s)\n", swap_strerror(rc), swap_last_error(ctx)); return 1; }
    swap_stats st;
    swap_get_stats(ctx, &st);
    printf("Proc 0: tot_kmer_num=%llu reads=%llu distinct=%llu survivors=%llu\n", (unsigned long long)st.kmers,
           (unsigned long long)st.reads, (unsigned long long)st.distinct, (unsigned long long)st.survivors);
    if (log) fprintf(log, "Time spend in Constructing Graph %.5f\n", since(t1));

    if (kmer_graph) {
        auto t2 = std::chrono::steady_clock::now();
        int64_t w = swap_write_file(ctx, SWAP_FILE_KMER_GRAPH, (out + "/kmerGraph.txt").c_str());
        if (w < 0) { printf("Error: %s (%s)\n", swap_strerror((int)w), swap_last_error(ctx)); return 1; }
        if (log) fprintf(log, "Time spend in printDistGraph %.5f\n", since(t2));
    } else {
        auto t2 = std::chrono::steady_clock::now();
        rc = swap_merge(ctx);
        if (rc) { printf("Error: %s (%s)\n", swap_strerror(rc), swap_last_error(ctx)); return 1; }
        swap_get_stats(ctx, &st);
        if (log) fprintf(log, "Time spend in Simplify DistGraph %.5f\n", since(t2));
        auto t3 = std::chrono::steady_clock::now();
        if (perf) {
            int64_t w = swap_write_file(ctx, SWAP_FILE_PERF_CONTIGS, (out + "/noCEcontigs.fasta").c_str());
            if (w < 0) { printf("Error: %s (%s)\n", swap_strerror((int)w), swap_last_error(ctx)); return 1; }
            if (log) {
                fprintf(log, "Time spend in printDistGraph %.5f\n", since(t3));
                fprintf(log, "Time spend in All step %.5f\n", since(t0));
                fclose(log);
            }
            swap_destroy(ctx);
            return 0;
        }
        if (dist_graph) {
            int64_t w = swap_write_file(ctx, SWAP_FILE_CONTIG_GRAPH, (out + "/contigGraph.txt").c_str());
            if (w < 0) { printf("Error: %s (%s)\n", swap_strerror((int)w), swap_last_error(ctx)); return 1; }
        }
        int64_t w = swap_write_file(ctx, SWAP_FILE_NOCE, (out + "/noCEcontigs.fasta").c_str());
        if (w < 0) { printf("Error: %s (%s)\n", swap_strerror((int)w), swap_last_error(ctx)); return 1; }
        if (log) fprintf(log, "Time spend in printDistGraph %.5f\n", since(t3));
        printf("|proc:0 %llu nodes in distGraph\n", (unsigned long long)st.kept);
        auto t4 = std::chrono::steady_clock::now();
        w = swap_write_file(ctx, SWAP_FILE_CE_CONTIGS, (out + "/CEContig.fasta").c_str());
        if (w < 0) { printf("Error: %s (%s)\n", swap_strerror((int)w), swap_last_error(ctx)); return 1; }
        if (log) fprintf(log, "Time spend in Contig Extension %.5f\n", since(t4));
        if (jung) {
            w = swap_write_file(ctx, SWAP_FILE_JUNG_ARC, (out + "/JungGraph_arc.txt").c_str());
            if (w >= 0) w = swap_write_file(ctx, SWAP_FILE_JUNG_MUL, (out + "/JungGraph_mul.txt").c_str());
            if (w < 0) { printf("Error: %s (%s)\n", swap_strerror((int)w), swap_last_error(ctx)); return 1; }
        }
    }
    if (log) {
        fprintf(log, "Time spend in All step %.5f\n", since(t0));
        fclose(log);
    }
    swap_destroy(ctx);
    return 0;
}
