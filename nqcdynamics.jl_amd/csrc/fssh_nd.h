// fssh_nd.h -- two-state FSSH / Ehrenfest on the six-coordinate diatomic-on-surface model,
// evaluated analytically or through the MLP surrogate; entry points used by abi.cpp.
#pragma once
#include "handle.h"

namespace nqcd {

constexpr int kNdDof = 6;
constexpr int kDiatomParams = 16;

int32_t nd_set_model(nqcd_handle_t* h, int32_t kind, const double* params, int64_t n);
bool nd_is_problem(const nqcd_handle_t* h);
// Resident workgroups per compute unit of the analytic-model persistent kernel at `block` threads.
hipError_t nd_persistent_occupancy(int precision, int block, int* blocks_per_cu);
int32_t nd_run(nqcd_handle_t* h, double t0, int64_t nsteps, const nqcd_term_desc* term, uint64_t key);

}  // namespace nqcd
