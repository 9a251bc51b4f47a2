// Seeded synthetic read generator for the k-mer graph benchmarks.
//
// Produces FASTA in the only grammar the assembler accepts: one header line
// ">r<i>" and exactly one upper-case sequence line per read, every line ending
// in '\n' (the record grammar read by kmerGraph.cpp:160-233).
//
// RNG: xoshiro256** seeded through splitmix64.  Genome bases are i.i.d.
// uniform ACGT.  A read starts uniformly in [0, G-L], comes from either strand
// with probability 1/2 (reverse-complemented), and carries i.i.d. substitution
// errors at rate e, the new base uniform over the other three.  R = round(cov*G/L).
//
// Workloads (sizes of the benchmark configurations):
//   small     G=1,000,000  L=100 cov=30 e=0.005 seed=1     (k=31 c=3)
//   bacterial G=4,600,000  L=100 cov=50 e=0.01  seed=2     (k=31 c=5)
//   batch i   G=round(exp U[ln 2e4, ln 2e6]) L=150 cov~U[20,80] e=0.005 seed=3+i (k=31 c=4)
//   repeat    G=100,000,000, 5% in 200 repeat families (unit U[1,5] kbp,
//             copies with 2% point divergence), L=150 cov=40 e=0.005 seed=4
//   request i G=round(exp U[ln 5e4, ln 5e6]) L in {100,150,250} cov~U[20,60]
//             e=0.005 k in {21,25,31} c=3 seed=5+i
//
// Usage:
//   swap_gen --workload small|bacterial|repeat -o reads.fa
//   swap_gen --workload batch   --index I -o reads.fa
//   swap_gen --workload request --index I -o reads.fa   (prints "k c" on stdout)
//   swap_gen -G len -L readlen -x cov -e err -s seed [--repeat-frac f --families n] -o reads.fa
//
// The C entry point swapgen_fasta() generates into a caller buffer so that
// benchmarks can avoid the file round trip.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct Rng {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t &x) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (int i = 0; i < 4; i++) s[i] = splitmix(x);
    }
    static uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    // uniform integer in [0, n)
    uint64_t below(uint64_t n) { return (uint64_t)(((unsigned __int128)next() * n) >> 64); }
    // uniform double in [0, 1)
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

const char kBases[4] = {'A', 'C', 'G', 'T'};

inline int code_of(char c) {
    switch (c) { case 'A': return 0; case 'C': return 1; case 'G': return 2; default: return 3; }
}

}  // namespace

struct GenParams {
    uint64_t genome_len;
    uint32_t read_len;
    double coverage;
    double error_rate;
    uint64_t seed;
    double repeat_frac;   // fraction of the genome covered by repeat copies
    uint32_t families;    // number of repeat families
};

static void make_genome(const GenParams &p, Rng &rng, std::vector<char> &g) {
    g.resize(p.genome_len);
    for (uint64_t i = 0; i < p.genome_len; i++) g[i] = kBases[rng.next() >> 62];
    if (p.repeat_frac <= 0.0 || p.families == 0) return;
    const uint64_t per_family = (uint64_t)(p.repeat_frac * (double)p.genome_len / p.families);
    std::vector<char> unit;
    for (uint32_t f = 0; f < p.families; f++) {
        const uint64_t ulen = 1000 + rng.below(4001);  // U[1,5] kbp
        unit.resize(ulen);
        for (uint64_t i = 0; i < ulen; i++) unit[i] = kBases[rng.next() >> 62];
        uint64_t placed = 0;
        while (placed < per_family && ulen < p.genome_len) {
            const uint64_t pos = rng.below(p.genome_len - ulen + 1);
            for (uint64_t i = 0; i < ulen; i++) {
                char b = unit[i];
                if (rng.unit() < 0.02) b = kBases[(code_of(b) + 1 + rng.below(3)) & 3];
                g[pos + i] = b;
            }
            placed += ulen;
        }
    }
}

// Number of reads for a parameter set.
static uint64_t read_count(const GenParams &p) {
    return (uint64_t)std::llround(p.coverage * (double)p.genome_len / (double)p.read_len);
}

// Writes the FASTA text into out (cap bytes).  Returns the byte count, or the
// required size when out is NULL / too small (nothing useful written then).
static uint64_t emit_fasta(const GenParams &p, char *out, uint64_t cap) {
    const uint64_t R = read_count(p);
    const uint32_t L = p.read_len;
    uint64_t need = 0;
    {
        char hdr[32];
        for (uint64_t i = 0; i < R; i++) need += (uint64_t)snprintf(hdr, sizeof hdr, ">r%llu\n", (unsigned long long)i) + L + 1;
    }
    if (out == nullptr || cap < need || L > p.genome_len) return need;
    Rng rng(p.seed);
    std::vector<char> g;
    make_genome(p, rng, g);
    uint64_t o = 0;
    std::vector<char> read(L);
    for (uint64_t i = 0; i < R; i++) {
        const uint64_t start = rng.below(p.genome_len - L + 1);
        const bool reverse = (rng.next() >> 63) != 0;
        for (uint32_t j = 0; j < L; j++) {
            char b = reverse ? kBases[3 - code_of(g[start + L - 1 - j])] : g[start + j];
            if (p.error_rate > 0.0 && rng.unit() < p.error_rate) b = kBases[(code_of(b) + 1 + rng.below(3)) & 3];
            read[j] = b;
        }
        o += (uint64_t)sprintf(out + o, ">r%llu\n", (unsigned long long)i);
        memcpy(out + o, read.data(), L);
        o += L;
        out[o++] = '\n';
    }
    return o;
}

struct Workload {
    GenParams p;
    int k;
    int c;
};

static bool named_workload(const char *name, uint64_t index, Workload &w) {
    memset(&w, 0, sizeof w);
    if (!strcmp(name, "small")) {
        w.p = {1000000, 100, 30.0, 0.005, 1, 0.0, 0}; w.k = 31; w.c = 3; return true;
    }
    if (!strcmp(name, "bacterial")) {
        w.p = {4600000, 100, 50.0, 0.01, 2, 0.0, 0}; w.k = 31; w.c = 5; return true;
    }
    if (!strcmp(name, "repeat")) {
        w.p = {100000000, 150, 40.0, 0.005, 4, 0.05, 200}; w.k = 31; w.c = 4; return true;
    }
    if (!strcmp(name, "batch")) {
        Rng r(3 + index);
        const double lo = std::log(2e4), hi = std::log(2e6);
        const uint64_t G = (uint64_t)std::llround(std::exp(lo + (hi - lo) * r.unit()));
        const double cov = 20.0 + 60.0 * r.unit();
        w.p = {G, 150, cov, 0.005, 0x5A5A0000ull + 3 + index, 0.0, 0}; w.k = 31; w.c = 4; return true;
    }
    if (!strcmp(name, "request")) {
        Rng r(5 + index);
        const double lo = std::log(5e4), hi = std::log(5e6);
        const uint64_t G = (uint64_t)std::llround(std::exp(lo + (hi - lo) * r.unit()));
        const uint32_t lens[3] = {100, 150, 250};
        const int ks[3] = {21, 25, 31};
        const uint32_t L = lens[r.below(3)];
        const double cov = 20.0 + 40.0 * r.unit();
        const int k = ks[r.below(3)];
        w.p = {G, L, cov, 0.005, 0xA5A50000ull + 5 + index, 0.0, 0}; w.k = k; w.c = 3; return true;
    }
    return false;
}

extern "C" {

// Size in bytes of the FASTA for a named workload (small, bacterial, repeat,
// batch, request; index selects the sample / request).  k and c receive the
// assembly parameters of the workload.  Returns 0 for an unknown name.
uint64_t swapgen_workload_size(const char *name, uint64_t index, int *k, int *c) {
    Workload w;
    if (!named_workload(name, index, w)) return 0;
    if (k) *k = w.k;
    if (c) *c = w.c;
    return emit_fasta(w.p, nullptr, 0);
}

// Generates the FASTA of a named workload into out.  Returns bytes written, or
// 0 if the name is unknown or cap is too small.
uint64_t swapgen_workload(const char *name, uint64_t index, char *out, uint64_t cap) {
    Workload w;
    if (!named_workload(name, index, w)) return 0;
    const uint64_t need = emit_fasta(w.p, nullptr, 0);
    if (cap < need) return 0;
    return emit_fasta(w.p, out, cap);
}

// Fully parameterised generator.  Same return convention as emit_fasta.
uint64_t swapgen_fasta(uint64_t genome_len, uint32_t read_len, double coverage, double error_rate,
                       uint64_t seed, double repeat_frac, uint32_t families, char *out, uint64_t cap) {
    GenParams p = {genome_len, read_len, coverage, error_rate, seed, repeat_frac, families};
    return emit_fasta(p, out, cap);
}

}  // extern "C"

#ifndef SWAPGEN_NO_MAIN
static void usage() {
    fprintf(stderr,
            "usage: swap_gen --workload NAME [--index I] -o FILE\n"
            "       swap_gen -G len -L readlen -x cov -e err -s seed [--repeat-frac f --families n] -o FILE\n");
}

int main(int argc, char **argv) {
    const char *wname = nullptr, *outpath = nullptr;
    uint64_t index = 0;
    GenParams p = {0, 100, 30.0, 0.005, 1, 0.0, 0};
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto val = [&](void) -> const char * {
            if (i + 1 >= argc) { usage(); exit(2); }
            return argv[++i];
        };
        if (a == "--workload") wname = val();
        else if (a == "--index") index = strtoull(val(), nullptr, 10);
        else if (a == "-o") outpath = val();
        else if (a == "-G") p.genome_len = strtoull(val(), nullptr, 10);
        else if (a == "-L") p.read_len = (uint32_t)strtoul(val(), nullptr, 10);
        else if (a == "-x") p.coverage = atof(val());
        else if (a == "-e") p.error_rate = atof(val());
        else if (a == "-s") p.seed = strtoull(val(), nullptr, 10);
        else if (a == "--repeat-frac") p.repeat_frac = atof(val());
        else if (a == "--families") p.families = (uint32_t)strtoul(val(), nullptr, 10);
        else { usage(); return 2; }
    }
    if (!outpath) { usage(); return 2; }
    int k = 0, c = 0;
    if (wname) {
        Workload w;
        if (!named_workload(wname, index, w)) { fprintf(stderr, "unknown workload %s\n", wname); return 2; }
        p = w.p; k = w.k; c = w.c;
    }
    if (p.genome_len == 0 || p.read_len == 0 || p.read_len > p.genome_len) { usage(); return 2; }
    const uint64_t need = emit_fasta(p, nullptr, 0);
    std::vector<char> buf(need);
    const uint64_t n = emit_fasta(p, buf.data(), need);
    FILE *f = fopen(outpath, "wb");
    if (!f) { perror(outpath); return 1; }
    if (fwrite(buf.data(), 1, n, f) != n) { perror(outpath); fclose(f); return 1; }
    fclose(f);
    if (wname) printf("%d %d %llu\n", k, c, (unsigned long long)n);
    return 0;
}
#endif
