// classical.hip -- classical dynamics of one coordinate on a single-state model (Simulation{Classical},
// ClassicalMethods/classical.jl), integrated as the reference integrates it: OrdinaryDiffEq's
// VelocityVerlet (the algorithm classical.jl selects), one trajectory per lane, FP64.
//
// A VelocityVerlet step evaluates the acceleration at the step start and at its end:
//   x' = x + dt v + dt^2 (a / 2),   a' = a(x'),   v' = v + dt (a / 2 + a' / 2).
// The start acceleration equals the previous step's end acceleration unless a callback moved the
// coordinate (CellBoundaryCallback), in which case it is evaluated again at the wrapped position.
// Saves take the state before the step's callbacks; the cell callback and the termination predicate
// follow the step (DiscreteCallbacks, never at t0).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "classical.h"
#include "ensemble.h"

namespace nqcd {

struct ClassicalArgs {
    int64_t ntraj, nsteps;
    int32_t save_every, nsave, term_kind, pad;
    double term_value, dt, mass;
    ModelSlot model;
    CellDesc cell;
    const double* r0;
    const double* v0;
    double* r;
    double* v;
    int32_t* steps;
    double* save_r;       // [slot][traj]
    double* save_v;
    double* save_ekin;
    double* save_epot;
};

// V and dV/dx of the single-state model (element [1, 1] of the model matrix)
__device__ __forceinline__ void cls_model(const ModelRegs& m, double x, double& V, double& dV) {
    double v22, v12, d22, d12;
    model1d_eval<double>(m, x, V, v22, v12, dV, d22, d12);
}

__global__ __launch_bounds__(256) void classical_kernel(const ClassicalArgs A) {
    const ModelRegs m = model_regs(A.model);
    const int64_t nt = A.ntraj;
    for (int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; tid < nt; tid += (int64_t)gridDim.x * blockDim.x) {
        double x = A.r0[tid], v = A.v0[tid];
        double V, dV;
        cls_model(m, x, V, dV);
        double a = -dV / A.mass;
        auto save = [&](int slot) {
            if (slot >= A.nsave) return;
            A.save_r[(int64_t)slot * nt + tid] = x;
            A.save_v[(int64_t)slot * nt + tid] = v;
            A.save_ekin[(int64_t)slot * nt + tid] = 0.5 * A.mass * v * v;   // classical_kinetic_energy
            A.save_epot[(int64_t)slot * nt + tid] = V;                      // classical_potential_energy
        };
        save(0);
        const double dt = A.dt, dtsq = dt * dt;
        int64_t step = 0;
        while (step < A.nsteps) {
            const double xn = x + dt * v + dtsq * (0.5 * a);
            cls_model(m, xn, V, dV);
            const double an = -dV / A.mass;
            v = v + dt * (0.5 * a + 0.5 * an);
            x = xn;
            a = an;
            if (A.save_every > 0 && (step + 1) % A.save_every == 0) save((int)((step + 1) / A.save_every));
            double xx[1] = {x};
            const bool cell_end = cell_callback<1, 1, double>(A.cell, xx);
            if (xx[0] != x) {  // wrapped: VelocityVerlet evaluates the start acceleration at the new position
                x = xx[0];
                cls_model(m, x, V, dV);
                a = -dV / A.mass;
            }
            ++step;
            if (cell_end || term_hit(A.term_kind, x, A.term_value)) break;
        }
        A.r[tid] = x;
        A.v[tid] = v;
        A.steps[tid] = (int32_t)step;
    }
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
bool cls_is_problem(const nqcd_handle_t* h) { return h->prob.method == NQCD_METHOD_CLASSICAL; }

int32_t cls_check_problem(nqcd_handle_t* h) {
    const nqcd_problem_desc& d = h->prob;
    if (d.algorithm != NQCD_ALG_VELOCITY_VERLET)
        return fail(h, NQCD_ERR_UNSUPPORTED, "Simulation{Classical} integrates with VelocityVerlet");
    if (d.n_states != 1 || d.ndofs * d.natoms != 1 || d.nbeads != 1)
        return fail(h, NQCD_ERR_UNSUPPORTED, "classical dynamics: a single-state model of one coordinate");
    if (d.precision != NQCD_PREC_FP64) return fail(h, NQCD_ERR_UNSUPPORTED, "classical dynamics run in FP64");
    return NQCD_OK;
}

int32_t cls_load_initial(nqcd_handle_t* h, int64_t ntraj, const double* r, const double* v) {
    h->ntraj = ntraj;
    HIPCHK(h, h->r0.ensure(sizeof(double) * ntraj));
    HIPCHK(h, h->v0.ensure(sizeof(double) * ntraj));
    HIPCHK(h, hipMemcpy(h->r0.p, r, sizeof(double) * ntraj, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->v0.p, v, sizeof(double) * ntraj, hipMemcpyHostToDevice));
    return NQCD_OK;
}

int32_t cls_sample_initial(nqcd_handle_t* h, const nqcd_sampler_desc* S, uint64_t seed) {
    if (S->kind != NQCD_SAMPLER_GAUSSIAN || S->nparams != 4 || !S->params)
        return fail(h, NQCD_ERR_UNSUPPORTED, "classical dynamics sample with GAUSSIAN (params mu_r, sd_r, mu_v, sd_v)");
    const int64_t n = S->ntraj;
    h->ntraj = n;
    HIPCHK(h, h->r0.ensure(sizeof(double) * n));
    HIPCHK(h, h->v0.ensure(sizeof(double) * n));
    HIPCHK(h, h->small.ensure(sizeof(double) * 4));
    HIPCHK(h, hipMemcpy(h->small.p, S->params, sizeof(double) * 4, hipMemcpyHostToDevice));
    HIPCHK(h, launch_sample_gaussian(n, 1, h->small.as<double>(), seed, h->r0.as<double>(), h->v0.as<double>(), h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return NQCD_OK;
}

int32_t cls_run(nqcd_handle_t* h, double t0, int64_t nsteps, const nqcd_term_desc* term) {
    (void)t0;
    if (!h->model_set[0]) return fail(h, NQCD_ERR_STATE, "model slot 0 is not set");
    if (h->models[0].kind != NQCD_MODEL_HARMONIC)
        return fail(h, NQCD_ERR_UNSUPPORTED, "classical dynamics models: HARMONIC");
    const int64_t n = h->ntraj;
    const int ns = h->nsave;
    HIPCHK(h, h->r.ensure(sizeof(double) * n));
    HIPCHK(h, h->v.ensure(sizeof(double) * n));
    HIPCHK(h, h->steps.ensure(sizeof(int32_t) * n));
    HIPCHK(h, h->save_centroid.ensure(sizeof(double) * ns * n));
    HIPCHK(h, h->save_v.ensure(sizeof(double) * ns * n));
    HIPCHK(h, h->save_ekin.ensure(sizeof(double) * ns * n));
    HIPCHK(h, h->save_epot.ensure(sizeof(double) * ns * n));
    HIPCHK(h, launch_zero(h->save_centroid.p, sizeof(double) * ns * n, h->stream));
    HIPCHK(h, launch_zero(h->save_v.p, sizeof(double) * ns * n, h->stream));
    HIPCHK(h, launch_zero(h->save_ekin.p, sizeof(double) * ns * n, h->stream));
    HIPCHK(h, launch_zero(h->save_epot.p, sizeof(double) * ns * n, h->stream));
    ClassicalArgs A{};
    A.ntraj = n; A.nsteps = nsteps;
    A.save_every = h->save_every; A.nsave = ns;
    A.term_kind = term ? term->kind : 0; A.term_value = term ? term->value : 0.0;
    A.dt = h->prob.dt; A.mass = h->masses[0];
    A.model = h->models[0];
    A.cell = h->cell;
    A.r0 = h->r0.as<double>(); A.v0 = h->v0.as<double>();
    A.r = h->r.as<double>(); A.v = h->v.as<double>(); A.steps = h->steps.as<int32_t>();
    A.save_r = h->save_centroid.as<double>(); A.save_v = h->save_v.as<double>();
    A.save_ekin = h->save_ekin.as<double>(); A.save_epot = h->save_epot.as<double>();
    const int block = 256;
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((n + block - 1) / block, (int64_t)h->cu_count * 8));
    HIPCHK(h, hipEventRecord(h->ev0, h->stream));
    hipLaunchKernelGGL(classical_kernel, dim3((unsigned)grid), dim3(block), 0, h->stream, A);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipEventRecord(h->ev1, h->stream));
    h->stats.block = block;
    h->stats.grid = (int)grid;
    h->stats.launches = 1;
    return NQCD_OK;
}

}  // namespace nqcd
