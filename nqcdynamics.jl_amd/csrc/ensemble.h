// ensemble.h -- sampling and reduction launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nqcd_device.h"

namespace nqcd {

enum { OUTK_FINAL_ADIABATIC = 4, OUTK_SCATTERING_1D = 2, OUTK_POP_CORRELATION = 3 };

hipError_t launch_sample_tully_grid(int64_t ntraj, int32_t nm, double x0, double p_lo, int32_t n_p, double mass,
                                    uint64_t key, double* r0, double* v0, int32_t* model_id, hipStream_t s);
hipError_t launch_sample_gaussian(int64_t ntraj, int32_t D, const double* prm, uint64_t key, double* r0, double* v0,
                                  hipStream_t s);
hipError_t launch_init_density2(int64_t ntraj, int32_t state, int32_t diabatic, const ModelSlot* models,
                                const int32_t* model_id, const double* r0, double* sig0, hipStream_t s);
hipError_t launch_fill_output2(int32_t kind, int32_t method, int64_t ntraj, int32_t nsave, int32_t xdof, const double* r,
                               const double* sig, const int32_t* state, const double* save_pop_di, double* out,
                               hipStream_t s);
// Clears `bytes` (a multiple of 4) at p with a fill kernel: save slots a trajectory never reaches
// (it terminated first) read as zero, whatever the buffer held before.
hipError_t launch_zero(void* p, size_t bytes, hipStream_t s);
// out[0] = x[0] + ... + x[n - 1] (int64)
hipError_t launch_sum_i32(const int32_t* x, int64_t n, long long* out, hipStream_t s);
int reduce_nchunk(int64_t ntraj);
hipError_t launch_reduce_rows(const double* data, int32_t rows, int64_t ntraj, double* partial, double scale, double* out,
                              hipStream_t s);

}  // namespace nqcd
