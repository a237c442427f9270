// spin_boson.h -- spin-boson (system-bath) FSSH / Ehrenfest entry points used by abi.cpp.
#pragma once
#include "handle.h"

namespace nqcd {

int32_t sb_sample_initial(nqcd_handle_t* h, const nqcd_sampler_desc* S, uint64_t seed);
int32_t sb_run(nqcd_handle_t* h, double t0, int64_t nsteps, const nqcd_term_desc* term, uint64_t key);

}  // namespace nqcd
