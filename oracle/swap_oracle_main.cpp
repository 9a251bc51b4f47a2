// Command-line front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).
//
//   swap_oracle -k K -c C -i reads.fa -o outdir [-s] [-d]
//
// Writes the files the reference writes at np=1 for the path: kmerGraph.txt
// (-s), contigGraph.txt (-d without -s) and noCEcontigs.fasta (without -s), plus
// a stats line on stdout.  Used to time the CPU baseline and to diff against
// reference runs.
#include "swap_oracle.h"

#include <sys/stat.h>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

static bool write_file(const std::string &path, const std::vector<char> &b) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = fwrite(b.data(), 1, b.size(), f) == b.size();
    fclose(f);
    return ok;
}

int main(int argc, char **argv) {
    int k = 23, c = 5;
    const char *in = nullptr, *out = nullptr;
    bool s = false, d = false;
    for (int i = 1; i < argc; i++) {
        if (!strcmp(argv[i], "-k") && i + 1 < argc) k = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-c") && i + 1 < argc) c = atoi(argv[++i]);
        else if (!strcmp(argv[i], "-i") && i + 1 < argc) in = argv[++i];
        else if (!strcmp(argv[i], "-o") && i + 1 < argc) out = argv[++i];
        else if (!strcmp(argv[i], "-s")) s = true;
        else if (!strcmp(argv[i], "-d")) d = true;
        else { fprintf(stderr, "usage: swap_oracle -k K -c C -i reads.fa -o outdir [-s] [-d]\n"); return 2; }
    }
    if (!in || !out) { fprintf(stderr, "usage: swap_oracle -k K -c C -i reads.fa -o outdir [-s] [-d]\n"); return 2; }
    FILE *f = fopen(in, "rb");
    if (!f) { perror(in); return 1; }
    std::vector<char> buf;
    {
        fseek(f, 0, SEEK_END);
        long n = ftell(f);
        fseek(f, 0, SEEK_SET);
        buf.resize((size_t)n);
        if (n > 0 && fread(buf.data(), 1, (size_t)n, f) != (size_t)n) { perror(in); fclose(f); return 1; }
        fclose(f);
    }
    if (mkdir(out, 0755) != 0) { fprintf(stderr, "Error: Creating directory %s\n", out); return 1; }
    auto t0 = std::chrono::steady_clock::now();
    oracle_graph *g = oracle_run(buf.data(), buf.size(), k, c, s ? 1 : 0);
    auto t1 = std::chrono::steady_clock::now();
    if (!g) { fprintf(stderr, "oracle: malformed input or bad k\n"); return 1; }
    auto dump = [&](int which, const char *name) {
        std::vector<char> b(oracle_file(g, which, nullptr, 0));
        oracle_file(g, which, b.data(), b.size());
        write_file(std::string(out) + "/" + name, b);
    };
    if (s) dump(0, "kmerGraph.txt");
    else {
        if (d) dump(1, "contigGraph.txt");
        dump(2, "noCEcontigs.fasta");
    }
    printf("{\"reads\": %llu, \"kmers\": %llu, \"distinct\": %llu, \"survivors\": %llu, \"kept\": %llu, \"seconds\": %.6f}\n",
           (unsigned long long)oracle_num_reads(g), (unsigned long long)oracle_num_occurrences(g),
           (unsigned long long)oracle_num_distinct(g), (unsigned long long)oracle_num_survivors(g),
           (unsigned long long)oracle_num_kept(g), std::chrono::duration<double>(t1 - t0).count());
    oracle_free(g);
    return 0;
}
