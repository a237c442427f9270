// ring_polymer.h -- ring-polymer (RPMD / RPSH / RP-Ehrenfest) entry points used by abi.cpp.
#pragma once
#include "handle.h"

namespace nqcd {

int32_t rp_check_problem(nqcd_handle_t* h);
int32_t rp_sample_initial(nqcd_handle_t* h, const nqcd_sampler_desc* S, uint64_t seed);
int32_t rp_load_initial(nqcd_handle_t* h, int64_t ntraj, const double* r, const double* v, const double* sig_re,
                        const double* sig_im, const int32_t* state);
int32_t rp_run(nqcd_handle_t* h, double t0, int64_t nsteps, const nqcd_term_desc* term, uint64_t key);

}  // namespace nqcd
