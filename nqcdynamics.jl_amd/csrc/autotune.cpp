// Process-wide launch-shape table of the persistent kernels (see autotune.h).
#include "autotune.h"

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "../../include/nqcd_mi355x.h"
#include "handle.h"

namespace nqcd {
namespace {

using Key = std::tuple<int32_t, int32_t, int32_t, int32_t>;

std::mutex g_mu;
std::map<Key, TuneShape> g_table;
bool g_loaded = false;
std::string g_path;  // cache file the table was last read from

bool valid(const TuneShape& s) { return s.block >= 64 && s.block <= 256 && s.block % 64 == 0 && s.wpc >= 0; }

// Reads the cache file on first use and again whenever NQCD_MI355X_TUNE_CACHE names another file
// (its entries are merged over the table); malformed lines are skipped.  Caller holds g_mu.
void load_locked() {
    const char* path = std::getenv("NQCD_MI355X_TUNE_CACHE");
    const std::string p = path ? path : "";
    if (g_loaded && p == g_path) return;
    g_loaded = true;
    g_path = p;
    if (p.empty()) return;
    FILE* f = std::fopen(path, "r");
    if (!f) return;
    char line[256];
    while (std::fgets(line, sizeof line, f)) {
        int fam, ns, prec, cu, block, wpc;
        if (std::sscanf(line, "%d %d %d %d %d %d", &fam, &ns, &prec, &cu, &block, &wpc) != 6) continue;
        TuneShape s;
        s.block = block;
        s.wpc = wpc;
        if (valid(s)) g_table[Key{fam, ns, prec, cu}] = s;
    }
    std::fclose(f);
}

}  // namespace

bool tune_lookup(int32_t family, int32_t n_states, int32_t precision, int32_t cu_count, TuneShape* out) {
    std::lock_guard<std::mutex> lock(g_mu);
    load_locked();
    auto it = g_table.find(Key{family, n_states, precision, cu_count});
    if (it == g_table.end()) it = g_table.find(Key{family, n_states, precision, 0});  // any device size
    if (it == g_table.end()) return false;
    *out = it->second;
    return true;
}

void tune_store(int32_t family, int32_t n_states, int32_t precision, int32_t cu_count, TuneShape s) {
    if (!valid(s)) return;
    std::lock_guard<std::mutex> lock(g_mu);
    load_locked();
    g_table[Key{family, n_states, precision, cu_count}] = s;
    const char* path = std::getenv("NQCD_MI355X_TUNE_CACHE");
    if (!path || !*path) return;
    if (FILE* f = std::fopen(path, "a")) {
        std::fprintf(f, "%d %d %d %d %d %d\n", family, n_states, precision, cu_count, s.block, s.wpc);
        std::fclose(f);
    }
}

TuneShape launch_shape(const nqcd_handle* h, int32_t family) {
    TuneShape s;
    if (h->force_block > 0) {
        s.block = h->force_block;
        s.wpc = h->force_wpc;
        return s;
    }
    TuneShape t;
    if (tune_lookup(family, h->prob.n_states, h->prob.precision, h->cu_count, &t)) return t;
    return s;
}

int64_t persistent_grid(int32_t occupancy, const TuneShape& s, int32_t cu_count, int64_t ntraj) {
    int64_t bpc = occupancy < 1 ? 1 : occupancy;
    if (s.wpc > 0 && s.wpc < bpc) bpc = s.wpc;
    int64_t grid = bpc * cu_count;
    const int64_t need = (ntraj + s.block - 1) / s.block;
    if (grid > need) grid = need;
    return grid < 1 ? 1 : grid;
}

}  // namespace nqcd
