// fssh_n.h -- FSSH / Ehrenfest for up to eight electronic states on one-coordinate models, with
// mixed state counts in one persistent launch; entry points used by abi.cpp.
#pragma once
#include "fssh2.h"
#include "handle.h"

namespace nqcd {

constexpr int kNsMaxStates = 8;

// Kernel arguments (read by the kernel through a device copy).  Arrays are structure-of-arrays with
// the trajectory index fastest; density matrices are NM x NM, element (i, j) in row i + NM j, where NM
// is the kernel's state count.
struct FsshNArgs {
    int64_t ntraj;
    int32_t method, rescaling, nsub, save_every;
    int32_t nsave, term_kind, init_diabatic, pad;
    int32_t slot_n[kMaxModels];     // states of each model slot
    int32_t slot_tab[kMaxModels];   // offset of a MORSE_N slot's table in mtab, -1 for closed-form models
    double term_value, dt, t0, inv_mass;
    int64_t nsteps;
    uint64_t key;
    const ModelSlot* models;
    const double* mtab;
    const double* r0;
    const double* v0;
    const double* sig0_re;
    const double* sig0_im;
    const int32_t* state0;
    const int32_t* model_id;
    const int32_t* order;
    double* r;
    double* v;
    double* sig_re;
    double* sig_im;
    int32_t* state;
    int32_t* steps;
    int32_t* nhops;
    double* save_pop_ad;   // [nsave][NM][ntraj]
    double* save_pop_di;
    int32_t* save_state;   // [nsave][ntraj]
    unsigned long long* queue;
    CellDesc cell;         // cell callback (one coordinate); mode 0: none
};

// Kernel state count used for a problem of n states.
int ns_kernel_states(int n);
bool ns_is_problem(const nqcd_handle_t* h);
int32_t ns_set_model(nqcd_handle_t* h, int32_t slot, const double* params, int64_t n);
int32_t ns_load_initial(nqcd_handle_t* h, int64_t ntraj, const double* r, const double* v, const double* sig_re,
                        const double* sig_im, const int32_t* state, const int32_t* model_id);
int32_t ns_sample_initial(nqcd_handle_t* h, const nqcd_sampler_desc* S, uint64_t seed);
int32_t ns_run(nqcd_handle_t* h, double t0, int64_t nsteps, const nqcd_term_desc* term, uint64_t key);
int32_t ns_download(nqcd_handle_t* h, int32_t field, void* out, int64_t bytes);
// Launch and occupancy of the five- to eight-state kernels (fssh_n8.hip).
hipError_t ns_launch_hi(int NM, int precision, int grid, int block, hipStream_t s, const FsshNArgs* Adev);
hipError_t ns_occupancy_hi(int NM, int precision, int block, int* bpc);

}  // namespace nqcd
