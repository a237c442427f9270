// fssh2.h -- launch interface of the two-state FSSH / Ehrenfest kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "nqcd_device.h"

namespace nqcd {

enum { METHOD_FSSH = 0, METHOD_EHRENFEST = 1 };
constexpr int kMaxModels = 4;
constexpr int kMaxDof1 = 8;

// Trajectories a persistent launch sets aside once its queue has drained and a wave has gone sparse,
// for the next launch to take up densely (structure-of-arrays, slot fastest).  Results do not depend on
// it: the random numbers are keyed by trajectory id and step.
struct SuspSet {
    double* f;                  // [kSuspRows][cap]: r, v, k, q, x, y, tr, U[4], dEc, wc, tc, t
    int64_t* tid;               // [cap]
    int32_t* i;                 // [3][cap]: active state, hops, steps
    unsigned long long* count;  // trajectories set aside
    unsigned long long* claim;  // claim counter of the launch that takes them up
    int64_t cap;
};
constexpr int kSuspRows = 15;
// A persistent run is up to kSuspRounds launches (default kSuspDefaultRounds): the first over the queue,
// then each over the trajectories the previous one set aside; a drained wave with at most the threshold
// (default kSuspDefaultThreshold) active lanes sets them aside.  NQCD_MI355X_SUSPEND="rounds,threshold"
// overrides both (rounds 1: a single launch).
constexpr int kSuspRounds = 8;
constexpr int kSuspDefaultRounds = 1;
constexpr int kSuspDefaultThreshold = 40;
void fssh2_suspend_policy(int* rounds, int* threshold);

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
    double* final_pop_di;  // [2][ntraj] diabatic populations of the final state, or null
    unsigned long long* queue;
    CellDesc cell;         // cell callback (mode 0: none)
    SuspSet sin;           // trajectories to take up (sin.f null: the queue over ntraj)
    SuspSet sout;          // where sparse waves set theirs aside (sout.f null: run to the end)
    int32_t susp_thr;
    int32_t pad_s;
    // Wave priority by claim order (0: off): a trajectory claimed as queue entry q has class
    // 3 - min(3, q / prio_span), and a wave runs at the highest class among its running lanes, so the
    // waves holding the earliest (with a longest-first queue order, the longest) trajectories get the
    // issue slots first.
    int64_t prio_span;
};
// prio_span of a run on `lanes` resident lanes: kPrioDefaultSpan lanes' worth of claims per class, or
// NQCD_MI355X_PRIO (a multiple of the lane count, 0 = off)
constexpr double kPrioDefaultSpan = 1.0;
int64_t fssh2_prio_span(int64_t lanes);

// Adev: room for kSuspRounds device copies of the arguments (each launch reads its own through this
// pointer); sets: two buffers of grid * block slots each, used in turn (null: a single launch).
hipError_t fssh2_launch_persistent(const Fssh2Args& A, Fssh2Args* Adev, int precision, int grid, int block,
                                   hipStream_t stream, const SuspSet* sets = nullptr);
hipError_t fssh2_persistent_occupancy(int precision, int block, int* blocks_per_cu);
hipError_t fssh2_launch_plain_init(const Fssh2Args& A, double* S, int32_t* Si, hipStream_t stream);
hipError_t fssh2_launch_plain_step(const Fssh2Args& A, double* S, int32_t* Si, hipStream_t stream);
hipError_t fssh2_launch_plain_finish(const Fssh2Args& A, double* S, int32_t* Si, hipStream_t stream);
constexpr int kPlainRows = 20;
constexpr int kPlainIntRows = 4;

}  // namespace nqcd
