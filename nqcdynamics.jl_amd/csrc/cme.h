// cme.h -- classical master equation surface hopping (CME / BCME) entry points used by abi.cpp.
#pragma once
#include "handle.h"

namespace nqcd {

bool cme_is_problem(const nqcd_handle_t* h);
int32_t cme_check_problem(nqcd_handle_t* h);
int32_t cme_load_initial(nqcd_handle_t* h, int64_t ntraj, const double* r, const double* v, const int32_t* state);
int32_t cme_sample_initial(nqcd_handle_t* h, const nqcd_sampler_desc* S, uint64_t seed);
int32_t cme_run(nqcd_handle_t* h, double t0, int64_t nsteps, const nqcd_term_desc* term, uint64_t key);
int32_t cme_test_broadening(nqcd_handle_t* h, const double* hs, int64_t count, double mu, double kT, double Gam, double W,
                            double* n, double* n2);

}  // namespace nqcd
