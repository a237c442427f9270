// iesh.h -- IESH / EhrenfestNA (many-electron wavefunction, Newns-Anderson model) entry points used by abi.cpp.
#pragma once
#include "handle.h"

namespace nqcd {

int32_t ie_check_problem(nqcd_handle_t* h);
int32_t ie_set_model(nqcd_handle_t* h, int32_t slot, const double* params, int64_t n);
int32_t ie_load_initial(nqcd_handle_t* h, int64_t ntraj, const double* r, const double* v, const double* psi_re,
                        const double* psi_im, const int32_t* occ);
int32_t ie_sample_initial(nqcd_handle_t* h, const nqcd_sampler_desc* S, uint64_t seed);
// Occupations from the Fermi-Dirac (Boltzmann-factor) chain at the loaded positions; avail = NULL
// makes every state available; basis 0 = adiabatic energies, 1 = diabatic (orbitals built from U).
int32_t ie_sample_fermi_dirac(nqcd_handle_t* h, double beta, const int32_t* avail, int32_t navail, int32_t basis,
                              uint64_t seed);
int32_t ie_run(nqcd_handle_t* h, double t0, int64_t nsteps, const nqcd_term_desc* term, uint64_t key);
int32_t ie_download(nqcd_handle_t* h, int32_t field, void* out, int64_t bytes);
int32_t ie_test_det(nqcd_handle_t* h, const double* re, const double* im, int32_t Ne, int32_t nmat, double* out);

}  // namespace nqcd
