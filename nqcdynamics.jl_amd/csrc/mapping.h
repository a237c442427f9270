// mapping.h -- spin-mapping dynamics (SpinMappingW) integrated by MInt, entry points used by abi.cpp.
#pragma once
#include "handle.h"

namespace nqcd {

constexpr int kMapMaxStates = 4;

bool map_is_problem(const nqcd_handle_t* h);
int32_t map_check_problem(nqcd_handle_t* h);
int32_t map_load_initial(nqcd_handle_t* h, int64_t ntraj, const double* r, const double* v, const int32_t* state,
                         const int32_t* model_id);
int32_t map_sample_initial(nqcd_handle_t* h, const nqcd_sampler_desc* S, uint64_t seed);
int32_t map_load_variables(nqcd_handle_t* h, int64_t ntraj, const double* X, const double* P);
int32_t map_run(nqcd_handle_t* h, double t0, int64_t nsteps, uint64_t key);
int32_t map_download(nqcd_handle_t* h, int32_t field, void* out, int64_t bytes);

}  // namespace nqcd
