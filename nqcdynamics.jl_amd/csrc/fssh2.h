// fssh2.h -- launch interface of the two-state FSSH / Ehrenfest kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nqcd_device.h"

namespace nqcd {

enum { METHOD_FSSH = 0, METHOD_EHRENFEST = 1 };
constexpr int kMaxModels = 4;
constexpr int kMaxDof1 = 8;

// Kernel arguments (passed by value).  Arrays are structure-of-arrays with the
// trajectory index fastest: element (row, traj) at row * ntraj + traj.
struct Fssh2Args {
    int64_t ntraj;
    int32_t method, rescaling, nsub, save_every;
    int32_t nsave, term_kind, term_dof, nmodels;
    double term_value, dt, t0;
    int64_t nsteps;
    uint64_t key;
    const double* inv_mass;     // [D] per flattened dof
    const ModelSlot* models;    // [kMaxModels]
    // inputs
    const double* r0;      // [D][ntraj]
    const double* v0;      // [D][ntraj]
    const double* sig0;    // [4][ntraj]: sigma00, sigma11, Re sigma01, Im sigma01
    const int32_t* state0; // [ntraj] or null
    const int32_t* model_id;
    const int32_t* order;  // queue order (permutation of trajectory ids) or null
    // outputs
    double* r;
    double* v;
    double* sig;
    int32_t* state;
    int32_t* steps;
    int32_t* nhops;
    double* save_pop_ad;   // [nsave][2][ntraj] or null
    double* save_pop_di;   // [nsave][2][ntraj] or null
    int32_t* save_state;   // [nsave][ntraj] or null
    double* save_ekin;     // [nsave][ntraj] classical kinetic energy, or null
    double* save_epot;     // [nsave][ntraj] classical potential energy (fssh.jl:150-154, ehrenfest.jl:85-95), or null
    unsigned long long* queue;
};

// Adev: device copy of the arguments (the kernel reads them through this pointer).
hipError_t fssh2_launch_persistent(const Fssh2Args& A, Fssh2Args* Adev, int precision, int grid, int block,
                                   hipStream_t stream);
hipError_t fssh2_persistent_occupancy(int precision, int block, int* blocks_per_cu);
hipError_t fssh2_launch_plain_init(const Fssh2Args& A, double* S, int32_t* Si, hipStream_t stream);
hipError_t fssh2_launch_plain_step(const Fssh2Args& A, double* S, int32_t* Si, hipStream_t stream);
hipError_t fssh2_launch_plain_finish(const Fssh2Args& A, double* S, int32_t* Si, hipStream_t stream);
constexpr int kPlainRows = 20;
constexpr int kPlainIntRows = 4;

}  // namespace nqcd
