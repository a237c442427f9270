// Launch shapes of the persistent work-queue kernels and the table the autotuner fills.
//
// A persistent kernel's results do not depend on its launch shape: every lane claims
// trajectory ids from the same queue and each trajectory's random numbers are keyed by its
// id.  The shape (threads per workgroup, resident workgroups per compute unit) only sets how
// many lanes share a compute unit, so it is chosen by timing, per kernel family, state count,
// precision and device size.
#pragma once

#include <cstdint>

struct nqcd_handle;

namespace nqcd {

enum TuneFamily : int32_t {
    TUNE_FSSH2 = 1,  // two-state one-dof persistent kernel (fssh2_persistent_*)
    TUNE_ND = 2,     // multi-dof analytic-model persistent kernel (fssh_nd_*)
};

struct TuneShape {
    int32_t block = 256;  // threads per workgroup (a multiple of 64, at most 256)
    int32_t wpc = 0;      // resident workgroups per compute unit; 0 = the occupancy limit
};

// Shape for a run: the handle's forced shape while the autotuner sweeps, else the table's
// entry for (family, n_states, precision, cu_count), else the default.
TuneShape launch_shape(const nqcd_handle* h, int32_t family);

// Process-wide table, guarded by a mutex.  When NQCD_MI355X_TUNE_CACHE names a file, the
// table is read from it on first use (and again when the variable names another file) and
// every new entry is appended to it (one line per entry: family n_states precision cu_count
// block wpc).  An entry with cu_count 0 applies to devices of any size that have no entry of
// their own.
bool tune_lookup(int32_t family, int32_t n_states, int32_t precision, int32_t cu_count, TuneShape* out);
void tune_store(int32_t family, int32_t n_states, int32_t precision, int32_t cu_count, TuneShape s);

// Grid for a persistent launch: min(wpc, occupancy) workgroups on each compute unit, but no more
// workgroups than the ensemble fills.
int64_t persistent_grid(int32_t occupancy, const TuneShape& s, int32_t cu_count, int64_t ntraj);

}  // namespace nqcd
