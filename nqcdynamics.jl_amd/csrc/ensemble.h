// ensemble.h -- sampling and reduction launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nqcd_device.h"

namespace nqcd {

enum { OUTK_FINAL_ADIABATIC = 4, OUTK_SCATTERING_1D = 2, OUTK_POP_CORRELATION = 3, OUTK_SCATTERING_1D_DIABATIC = 8 };

hipError_t launch_sample_tully_grid(int64_t ntraj, int32_t nm, double x0, double p_lo, int32_t n_p, double mass,
                                    uint64_t key, double* r0, double* v0, int32_t* model_id, hipStream_t s);
hipError_t launch_sample_gaussian(int64_t ntraj, int32_t D, const double* prm, uint64_t key, double* r0, double* v0,
                                  hipStream_t s);
hipError_t launch_init_density2(int64_t ntraj, int32_t state, int32_t diabatic, const ModelSlot* models,
                                const int32_t* model_id, const double* r0, double* sig0, hipStream_t s);
// Source of an output computed per trajectory from the final state (final adiabatic populations,
// state-resolved scattering) or from the diabatic save slots (population correlation).
struct OutputSrc {
    int32_t kind = 0;             // OUTK_*
    int32_t method = 0;           // 0 FSSH (one-hot active state), 1 Ehrenfest (diagonal of sigma)
    int32_t nsave = 0;
    int32_t xdof = 0;             // row of r classified by sign for scattering
    const double* r = nullptr;    // [D][ntraj]
    const double* sig = nullptr;  // [4][ntraj]
    const int32_t* state = nullptr;
    const double* save_pop_di = nullptr;  // [nsave][nst][ntraj]
    const double* final_pop_di = nullptr; // [2][ntraj] diabatic populations of the final state
    // population correlation: states, the initial estimator [nst][ntraj] (null: save slot 0) and the
    // normalisation (TimeCorrelationFunctions.jl:26-40)
    int32_t nst = 2, pad = 0;
    const double* init_pop = nullptr;
    double norm = 1.0;
};
hipError_t launch_fill_output2(const OutputSrc& S, int64_t ntraj, double* out, hipStream_t s);
int32_t fill_output2_rows(int32_t kind, int32_t nsave);
// Clears `bytes` (a multiple of 4) at p with a fill kernel: save slots a trajectory never reaches
// (it terminated first) read as zero, whatever the buffer held before.
hipError_t launch_zero(void* p, size_t bytes, hipStream_t s);
// out[0] = x[0] + ... + x[n - 1] (int64)
hipError_t launch_sum_i32(const int32_t* x, int64_t n, long long* out, hipStream_t s);
int reduce_nchunk(int64_t ntraj);
hipError_t launch_reduce_rows(const double* data, int32_t rows, int64_t ntraj, double* partial, double scale, double* out,
                              hipStream_t s);

// Allocation-free reduction: a per-device scratch array of kReduceScratchDoubles doubles (2 MiB).
constexpr size_t kReduceScratchDoubles = (size_t)1 << 18;
// Resolves the scratch array's address on `device` (which must be current) once per process.
hipError_t reduce_scratch_init(int device);
// Address of the scratch array on `device`, or null before reduce_scratch_init(device).
double* reduce_scratch(int device);
// Doubles of scratch a reduction of `rows` rows over ntraj trajectories uses.
size_t reduce_noalloc_doubles(int32_t rows, int64_t ntraj, bool two_sources);
// Sums `rows` rows (data, or S computed on the fly when data is null; plus data2 when given, summed
// separately and added after scaling) into scratch_ptr[reduce_noalloc_doubles - rows ..] (the last
// `rows` doubles of the region used).
hipError_t launch_reduce_rows_noalloc(const double* data, const double* data2, const OutputSrc& S, int32_t rows,
                                      int64_t ntraj, double scale, double* scratch_ptr, hipStream_t s);

}  // namespace nqcd
