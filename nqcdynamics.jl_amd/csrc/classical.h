// classical.h -- classical dynamics (Simulation{Classical}, VelocityVerlet) entry points used by abi.cpp.
#pragma once
#include "handle.h"

namespace nqcd {

bool cls_is_problem(const nqcd_handle_t* h);
int32_t cls_check_problem(nqcd_handle_t* h);
int32_t cls_load_initial(nqcd_handle_t* h, int64_t ntraj, const double* r, const double* v);
int32_t cls_sample_initial(nqcd_handle_t* h, const nqcd_sampler_desc* S, uint64_t seed);
int32_t cls_run(nqcd_handle_t* h, double t0, int64_t nsteps, const nqcd_term_desc* term);

}  // namespace nqcd
