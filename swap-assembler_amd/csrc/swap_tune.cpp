// Tuning table: the performance parameters of a build and merge (Tuning in
// swap_device.h) chosen per input-size class by the autotuner
// (swap-assembler_amd/tune/swap_autotune.py), read from tune/swap_tune.txt
// beside the library's directory, or from the file SWAP_TUNE_FILE names.
//
// One row per size class, ascending:
//   <max input bytes> <part_recs> <fill_recs> <index_load> <ruler_stride> <occ_groups> [<order_bin_keys>]
// '#' starts a comment; a row without the last column takes its default.  An input takes the first row whose bound it does not
// exceed (the last row above every bound); without a table, the built-in
// defaults.  No parameter changes a result, only the time it takes.
#include <dlfcn.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "swap_device.h"

namespace swapk {

Tuning clamp_tuning(const Tuning &t) {
    Tuning c = t;
    c.part_recs = std::min<uint32_t>(std::max<uint32_t>(c.part_recs, 1024), 65536);
    c.fill_recs = std::min<uint32_t>(std::max<uint32_t>(c.fill_recs, 512), 8192);
    c.index_load = std::min<uint32_t>(std::max<uint32_t>(c.index_load, 50), 85);
    c.ruler_stride = std::min<uint32_t>(std::max<uint32_t>(c.ruler_stride, 4), 256);
    c.occ_groups = std::min<uint32_t>(std::max<uint32_t>(c.occ_groups, 1), 4);
    c.order_bin_keys = std::min<uint32_t>(std::max<uint32_t>(c.order_bin_keys, 256), 1u << 30);
    return c;
}

namespace {

struct Row {
    uint64_t max_bytes;
    Tuning t;
};

std::string table_path() {
    if (const char *e = std::getenv("SWAP_TUNE_FILE")) return e;
    Dl_info info;
    if (!dladdr(reinterpret_cast<void *>(&clamp_tuning), &info) || !info.dli_fname) return "";
    std::string lib = info.dli_fname;  // .../swap-assembler_amd/lib/libswaphip.so
    const size_t slash = lib.rfind('/');
    const std::string dir = slash == std::string::npos ? "." : lib.substr(0, slash);
    return dir + "/../tune/swap_tune.txt";
}

std::vector<Row> read_table(const std::string &path) {
    std::vector<Row> rows;
    FILE *f = path.empty() ? nullptr : std::fopen(path.c_str(), "r");
    if (!f) return rows;
    char line[512];
    while (std::fgets(line, sizeof line, f)) {
        if (char *h = std::strchr(line, '#')) *h = 0;
        unsigned long long mb;
        unsigned pr, fr, il, rs, og, ob = Tuning().order_bin_keys;
        const int got = std::sscanf(line, "%llu %u %u %u %u %u %u", &mb, &pr, &fr, &il, &rs, &og, &ob);
        if (got != 6 && got != 7) continue;
        Row r;
        r.max_bytes = mb;
        r.t.part_recs = pr;
        r.t.fill_recs = fr;
        r.t.index_load = il;
        r.t.ruler_stride = rs;
        r.t.occ_groups = og;
        r.t.order_bin_keys = ob;
        r.t = clamp_tuning(r.t);
        rows.push_back(r);
    }
    std::fclose(f);
    std::sort(rows.begin(), rows.end(), [](const Row &a, const Row &b) { return a.max_bytes < b.max_bytes; });
    return rows;
}

std::once_flag g_once;
std::vector<Row> g_rows;

}  // namespace

Tuning tuning_for(uint64_t n_bytes) {
    std::call_once(g_once, [] { g_rows = read_table(table_path()); });
    if (g_rows.empty()) return Tuning();
    for (const Row &r : g_rows)
        if (n_bytes <= r.max_bytes) return r.t;
    return g_rows.back().t;
}

}  // namespace swapk
