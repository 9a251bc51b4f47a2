// swap_stats: contig-set statistics with the command line and output of the
// reference's stats tool (stats.cpp:18-86):
//
//   swap_stats <contigs.fasta> <reference genome length>
//
// A contig is the concatenation of the non-header lines between two header
// lines ('>' first).  Quirks kept from stats.cpp:
//   * a contig shorter than 200 bases is not recorded and its bases are NOT
//     discarded: they stay in front of the next contig (stats.cpp:32-37);
//   * N50_self is the length at which the running sum, longest first, first
//     exceeds tot/2 (integer division); N50_abs the same against half the given
//     genome length (stats.cpp:55-78);
//   * "mean" prints the median element vs[n/2] (stats.cpp:85);
//   * with the header-only noCEcontigs.fasta format every line is a header, so
//     no contig is recorded.
// Difference: with no recorded contig the reference reads vs[0] of an empty
// vector (undefined behaviour); this tool prints "Max = 0, mean = 0".
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: swap_stats <contigs.fasta> <reference genome length>\n");
        return 1;
    }
    const long long ref_len = std::atoll(argv[2]);
    std::ifstream in(argv[1]);
    std::vector<size_t> len;  // contig lengths (only lengths are reported)
    std::string line;
    size_t cur = 0;
    while (std::getline(in, line)) {
        if (!line.empty() && line[0] == '>') {
            if (cur < 200) continue;  // short contig: kept, joins the next one
            len.push_back(cur);
            cur = 0;
            continue;
        }
        cur += line.size();
    }
    if (cur >= 200) len.push_back(cur);
    std::printf("Input %d contigs\n", (int)len.size());
    std::sort(len.begin(), len.end(), [](size_t a, size_t b) { return a > b; });
    std::printf("Sort end\n");
    long long tot = 0;
    for (size_t v : len) tot += (long long)v;
    long long n50_self = 0, n50_abs = 0, sum = 0;
    for (size_t v : len) {
        sum += (long long)v;
        if (sum > tot / 2) {
            n50_self = (long long)v;
            break;
        }
    }
    sum = 0;
    for (size_t v : len) {
        sum += (long long)v;
        if (sum > ref_len / 2) {
            n50_abs = (long long)v;
            break;
        }
    }
    std::printf("ContigNum=%d, tot = %lld, N50_self=%lld, N50_abs=%lld\n", (int)len.size(), tot, n50_self, n50_abs);
    std::printf("Max = %d, mean = %d\n", len.empty() ? 0 : (int)len[0], len.empty() ? 0 : (int)len[len.size() / 2]);
    return 0;
}
