// mapping.hip -- spin-mapping dynamics (Simulation{SpinMappingW}, spin_mapping.jl:1-131) integrated by
// MInt (mint.jl:42-181), one coordinate, two to four electronic states, FP64, one trajectory per lane.
//
// With gamma = 2 (sqrt(n + 1) - 1) / n the mapping Hamiltonian of positions X and momenta P is
//   H = m v^2 / 2 + sum_i V_ii (X_i^2 + P_i^2 - gamma) / 2 + sum_{i<j} V_ij (X_i X_j + P_i P_j)
// (classical_potential_energy, spin_mapping.jl:102-118; the models here have no state-independent
// part).  An MInt step:
//   r' = r + dt/2 v;  (w, Z) the eigenpairs of V(r');  a = Z^T X, b = Z^T P;
//   a_k <- cos(w_k dt) a_k + sin(w_k dt) b_k,  b_k <- cos(w_k dt) b_k - sin(w_k dt) a_k
//     (X <- C X - D P, P <- C P + D X with C = Z cos(w dt) Z^T, D = Z sin(-w dt) Z^T, mint.jl:88-124,
//     applied in the eigenbasis);
//   W = Z^T dV Z;  Gamma_ii = W_ii dt,  Gamma_ij = sin(D_ij dt) W_ij / D_ij,
//   Xi_ij = (1 - cos(D_ij dt)) W_ij / D_ij with D_ij = w_j - w_i (i != j)  (mint.jl:132-162);
//   force = (X^T E X + P^T E P) / 2 - X^T F P with E = Z Gamma Z^T, F = Z Xi Z^T  (mint.jl:164-181),
//     evaluated in the eigenbasis as (a^T Gamma a + b^T Gamma b) / 2 - a^T Xi b;
//   v <- v - force / m + tr(dV) gamma dt / (2 m);  r <- r' + dt/2 v.
// Xi's off-diagonal part is antisymmetric (D_ji = -D_ij, W symmetric).  The reference's set_xi! leaves
// the diagonal of its cache matrix at the previous step's diagonal of F; F's diagonal is the quadratic
// form of Xi in the rows of Z, to which the antisymmetric part contributes nothing, so starting from a
// zero matrix that diagonal stays zero and the carried term is dropped here.
//
// Initial mapping variables (spin_mapping.jl:15-33): theta_i = 2 pi u with u from Philox stream 4 at
// counter (trajectory, i), radius sqrt(2 + gamma) on the occupied diabatic state and sqrt(gamma) on
// the others, X_i = radius sin(theta_i), P_i = radius cos(theta_i) (the reference stores the sines as
// the mapping positions); or uploaded by nqcd_mi355x_load_mapping.
//
// eCMM (Simulation{eCMM}, cmm.jl:1-120) runs the same kernel: its Hamiltonian
// (q^T V q + p^T V p - 2 gamma tr V) / 2 is the form above with 2 gamma in place of spin mapping's gamma on
// the diagonal; MInt's trace term is tr(dV) gamma dt / 2m for either method (mint.jl:85), which for eCMM
// is half of its Hamiltonian's and is kept as the reference computes it; its initial variables are a point
// uniform on the 2n-sphere of radius sqrt(2 + 2 n gamma) (2n Philox normals, pair k from stream 4 at
// counter (trajectory, k), normalised; the first n coordinates are the positions); its diabatic
// population is the inverse mapping kernel (1 + n) / (1 + n gamma)^2 (X^2 + P^2) / 2 - (1 - gamma) /
// (1 + n gamma) (cmm.jl:96-112).
//
// Device layout: X, P [state][traj]; diabatic populations (spin mapping (X_i^2 + P_i^2 - gamma) / 2,
// spin_mapping.jl:120-131) [slot][state][traj]; kinetic and potential energy [slot][traj].
#include "mapping.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "ensemble.h"
#include "fssh_n_kern.h"

namespace nqcd {

constexpr uint32_t kMapStream = 4;

struct MapArgs {
    int64_t ntraj, nsteps;
    int32_t save_every, nsave;
    int32_t slot_tab[kMaxModels];   // offset of a MORSE_N slot's table in mtab, -1 for closed-form models
    int32_t ecmm, pad;              // 1: eCMM's initial variables (sphere) instead of spin mapping's
    double dt, inv_mass, mass;
    double gamma;                   // constant of the diagonal terms V_ii (X_i^2 + P_i^2 - gamma) / 2 of H
    double gamma_force;             // gamma of MInt's trace term tr(dV) gamma dt / 2m
    double pop_a, pop_b;            // population estimator pop_a (X_i^2 + P_i^2) / 2 - pop_b
    double radius;                  // eCMM: radius of the initial sphere
    double occ_radius, free_radius; // spin mapping: initial radii of the occupied and the other states
    uint64_t key;
    const ModelSlot* models;
    const double* mtab;
    const double* r0;
    const double* v0;
    const double* X0;        // [state][traj] uploaded mapping variables, or null (drawn)
    const double* P0;
    const int32_t* state0;   // occupied diabatic state per trajectory
    const int32_t* model_id; // or null (slot 0)
    double* r;
    double* v;
    double* X;
    double* P;
    int32_t* steps;          // nuclear steps taken (every trajectory runs to t1)
    double* save_pop;        // [slot][state][traj]
    double* save_ekin;       // [slot][traj]
    double* save_epot;
};

// Potential part of H at coordinate x for mapping variables X, P.
template <int NM>
__device__ __forceinline__ double map_potential(const LaneN<NM, double>& L, double x, const double X[NM],
                                                const double P[NM], double gamma) {
    constexpr int NP = NM * (NM - 1) / 2;
    double vd[NM], vo[NP], dd[NM], dof[NP];
    ns_model<NM, double>(L, x, vd, vo, dd, dof);
    double e = 0.0;
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        e += vd[i] * (X[i] * X[i] + P[i] * P[i] - gamma) * 0.5;
#pragma unroll
        for (int j = i + 1; j < NM; ++j) e += vo[pidx<NM>(i, j)] * (X[i] * X[j] + P[i] * P[j]);
    }
    return e;
}

template <int NM>
__global__ __launch_bounds__(256) void mapping_kernel(const MapArgs A) {
    constexpr int NP = NM * (NM - 1) / 2;
    const int64_t nt = A.ntraj;
    const double dt = A.dt, hdt = 0.5 * A.dt, g = A.gamma, gf = A.gamma_force, im = A.inv_mass;
    for (int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; tid < nt; tid += (int64_t)gridDim.x * blockDim.x) {
        const int mid = A.model_id ? A.model_id[tid] : 0;
        LaneN<NM, double> L;
        L.n = NM;
        L.tab = A.slot_tab[mid] >= 0 ? A.mtab + A.slot_tab[mid] : nullptr;
        L.m = model_regs(A.models, mid);
        double x = A.r0[tid], v = A.v0[tid];
        double X[NM], P[NM];
        if (A.X0) {
#pragma unroll
            for (int i = 0; i < NM; ++i) { X[i] = A.X0[(int64_t)i * nt + tid]; P[i] = A.P0[(int64_t)i * nt + tid]; }
        } else if (A.ecmm) {
            double z[2 * NM], nrm = 0.0;
#pragma unroll
            for (int k = 0; k < NM; ++k) {
                philox_normal2((uint64_t)tid, (uint32_t)k, kMapStream, A.key, z[2 * k], z[2 * k + 1]);
                nrm += z[2 * k] * z[2 * k] + z[2 * k + 1] * z[2 * k + 1];
            }
            const double sc = A.radius / sqrt(nrm);
#pragma unroll
            for (int i = 0; i < NM; ++i) { X[i] = z[i] * sc; P[i] = z[NM + i] * sc; }
        } else {
            const int st = A.state0[tid];
#pragma unroll
            for (int i = 0; i < NM; ++i) {
                const double th = 6.283185307179586 * philox_uniform((uint64_t)tid, (uint32_t)i, kMapStream, A.key);
                const double rad = i == st ? A.occ_radius : A.free_radius;
                double s, c;
                sincos(th, &s, &c);
                X[i] = rad * s;
                P[i] = rad * c;
            }
        }
        auto save = [&](int slot) {
#pragma unroll
            for (int i = 0; i < NM; ++i)
                A.save_pop[((int64_t)slot * NM + i) * nt + tid] = A.pop_a * ((X[i] * X[i] + P[i] * P[i]) * 0.5) - A.pop_b;
            A.save_ekin[(int64_t)slot * nt + tid] = 0.5 * A.mass * v * v;
            A.save_epot[(int64_t)slot * nt + tid] = map_potential<NM>(L, x, X, P, g);
        };
        save(0);
        int slot = 1, until = A.save_every;
        for (int64_t step = 0; step < A.nsteps; ++step) {
            const double rt = x + hdt * v;
            double vd[NM], vo[NP > 0 ? NP : 1], dd[NM], dof[NP > 0 ? NP : 1];
            ns_model<NM, double>(L, rt, vd, vo, dd, dof);
            double w[NM], Z[NM * NM];
            if constexpr (NM == 2) eigen2(vd[0], vo[0], vd[1], w[0], w[1], Z, false);
            else ns_eigen<NM>(NM, vd, vo, w, Z, false);
            // mapping variables in the eigenbasis, rotated by exp(-i w dt)
            double a[NM], b[NM], sk[NM], ck[NM];
#pragma unroll
            for (int k = 0; k < NM; ++k) {
                double sa = 0.0, sb = 0.0;
#pragma unroll
                for (int i = 0; i < NM; ++i) { sa += Z[i * NM + k] * X[i]; sb += Z[i * NM + k] * P[i]; }
                sincos(w[k] * dt, &sk[k], &ck[k]);
                a[k] = ck[k] * sa + sk[k] * sb;
                b[k] = ck[k] * sb - sk[k] * sa;
            }
#pragma unroll
            for (int i = 0; i < NM; ++i) {
                double sx = 0.0, sp = 0.0;
#pragma unroll
                for (int k = 0; k < NM; ++k) { sx += Z[i * NM + k] * a[k]; sp += Z[i * NM + k] * b[k]; }
                X[i] = sx;
                P[i] = sp;
            }
            // nuclear force from the rotated variables
            double Wd[NM], Wo[NP > 0 ? NP : 1];
            ns_rotate<NM>(Z, dd, dof, Wd, Wo);
            double f = 0.0, tr = 0.0;
#pragma unroll
            for (int i = 0; i < NM; ++i) {
                f += 0.5 * Wd[i] * dt * (a[i] * a[i] + b[i] * b[i]);
                tr += dd[i];
#pragma unroll
                for (int j = i + 1; j < NM; ++j) {
                    // sin and cos of (w_j - w_i) dt from those of the rotation angles
                    const double dl = w[j] - w[i];
                    const double s = sk[j] * ck[i] - ck[j] * sk[i], c = ck[j] * ck[i] + sk[j] * sk[i];
                    const double wij = Wo[pidx<NM>(i, j)] / dl;
                    f += s * wij * (a[i] * a[j] + b[i] * b[j]) - (1.0 - c) * wij * (a[i] * b[j] - a[j] * b[i]);
                }
            }
            v = v - f * im + tr * gf * dt * 0.5 * im;
            x = rt + hdt * v;
            if (A.save_every > 0 && step + 1 == until) {
                if (slot < A.nsave) save(slot);
                ++slot;
                until += A.save_every;
            }
        }
        A.r[tid] = x;
        A.v[tid] = v;
        A.steps[tid] = (int32_t)A.nsteps;
#pragma unroll
        for (int i = 0; i < NM; ++i) { A.X[(int64_t)i * nt + tid] = X[i]; A.P[(int64_t)i * nt + tid] = P[i]; }
    }
}

template __global__ void mapping_kernel<2>(const MapArgs);
template __global__ void mapping_kernel<3>(const MapArgs);
template __global__ void mapping_kernel<4>(const MapArgs);

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
bool map_is_problem(const nqcd_handle_t* h) {
    return h->prob.method == NQCD_METHOD_SPIN_MAPPING || h->prob.method == NQCD_METHOD_ECMM;
}

static double map_gamma(int n) { return 2.0 / n * (std::sqrt((double)n + 1.0) - 1.0); }

static int map_slot_states(const nqcd_handle_t* h, int s) {
    return h->models[s].kind == NQCD_MODEL_MORSE_N ? (int)h->models[s].p[0] : 2;
}

int32_t map_check_problem(nqcd_handle_t* h) {
    const nqcd_problem_desc& d = h->prob;
    if (d.algorithm != NQCD_ALG_MINT)
        return fail(h, NQCD_ERR_UNSUPPORTED, "Simulation{SpinMappingW | eCMM} integrates with MInt");
    if (d.n_states < 2 || d.n_states > kMapMaxStates || d.ndofs * d.natoms != 1 || d.nbeads != 1)
        return fail(h, NQCD_ERR_UNSUPPORTED, "spin mapping / eCMM: 2 to 4 states, one nuclear coordinate");
    if (d.precision != NQCD_PREC_FP64) return fail(h, NQCD_ERR_UNSUPPORTED, "spin mapping / eCMM run in FP64");
    if (d.method == NQCD_METHOD_ECMM && (!std::isfinite(d.mapping_gamma) || 2.0 + 2.0 * d.n_states * d.mapping_gamma <= 0.0))
        return fail(h, NQCD_ERR_ARG, "eCMM: mapping_gamma must be finite with 2 + 2 n gamma > 0");
    return NQCD_OK;
}

static int32_t map_check_models(nqcd_handle_t* h) {
    for (int s = 0; s < kMaxModels; ++s) {
        if (!h->model_set[s]) continue;
        const int k = h->models[s].kind;
        if (k != NQCD_MODEL_MORSE_N && k != NQCD_MODEL_TULLY1 && k != NQCD_MODEL_TULLY2 && k != NQCD_MODEL_TULLY3 &&
            k != NQCD_MODEL_DW_GAUSSIAN)
            return fail(h, NQCD_ERR_UNSUPPORTED, "spin mapping models: TULLY1/2/3, DW_GAUSSIAN, MORSE_N");
        if (map_slot_states(h, s) != h->prob.n_states)
            return fail(h, NQCD_ERR_ARG, "spin mapping: every model slot must have n_states states");
    }
    return NQCD_OK;
}

int32_t map_load_initial(nqcd_handle_t* h, int64_t ntraj, const double* r, const double* v, const int32_t* state,
                         const int32_t* model_id) {
    if (int32_t rc = map_check_models(h)) return rc;
    std::vector<int32_t> zeros;
    if (!state && h->prob.method == NQCD_METHOD_ECMM) {   // eCMM's draw does not depend on a state (cmm.jl:26)
        zeros.assign(ntraj, 0);
        state = zeros.data();
    }
    if (!state) return fail(h, NQCD_ERR_ARG, "spin mapping needs the occupied diabatic state of each trajectory");
    h->ntraj = ntraj;
    HIPCHK(h, h->r0.ensure(sizeof(double) * ntraj));
    HIPCHK(h, h->v0.ensure(sizeof(double) * ntraj));
    HIPCHK(h, h->state0.ensure(sizeof(int32_t) * ntraj));
    HIPCHK(h, hipMemcpyAsync(h->r0.p, r, sizeof(double) * ntraj, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->v0.p, v, sizeof(double) * ntraj, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->state0.p, state, sizeof(int32_t) * ntraj, hipMemcpyHostToDevice, h->stream));
    h->model_id_given = model_id != nullptr;
    if (model_id) {
        HIPCHK(h, h->model_id.ensure(sizeof(int32_t) * ntraj));
        HIPCHK(h, hipMemcpyAsync(h->model_id.p, model_id, sizeof(int32_t) * ntraj, hipMemcpyHostToDevice, h->stream));
    }
    h->map_given = false;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return NQCD_OK;
}

int32_t map_sample_initial(nqcd_handle_t* h, const nqcd_sampler_desc* S, uint64_t seed) {
    if (int32_t rc = map_check_models(h)) return rc;
    if (S->kind != NQCD_SAMPLER_GAUSSIAN || S->nparams != 4 || !S->params)
        return fail(h, NQCD_ERR_UNSUPPORTED, "spin mapping samples with GAUSSIAN (params mu_r, sd_r, mu_v, sd_v)");
    if (S->state_basis != 1 && h->prob.method == NQCD_METHOD_SPIN_MAPPING)
        return fail(h, NQCD_ERR_UNSUPPORTED, "spin mapping starts from a diabatic pure state (spin_mapping.jl:15)");
    if (!h->model_set[0]) return fail(h, NQCD_ERR_STATE, "model slot 0 is not set");
    if (S->initial_state < 0 || S->initial_state >= h->prob.n_states) return fail(h, NQCD_ERR_ARG, "initial_state out of range");
    const int64_t ntraj = S->ntraj;
    h->ntraj = ntraj;
    HIPCHK(h, h->r0.ensure(sizeof(double) * ntraj));
    HIPCHK(h, h->v0.ensure(sizeof(double) * ntraj));
    HIPCHK(h, h->small.ensure(sizeof(double) * 4));
    HIPCHK(h, hipMemcpyAsync(h->small.p, S->params, sizeof(double) * 4, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, launch_sample_gaussian(ntraj, 1, h->small.as<double>(), seed, h->r0.as<double>(), h->v0.as<double>(),
                                     h->stream));
    std::vector<int32_t> st(ntraj, S->initial_state);
    HIPCHK(h, h->state0.ensure(sizeof(int32_t) * ntraj));
    HIPCHK(h, hipMemcpyAsync(h->state0.p, st.data(), sizeof(int32_t) * ntraj, hipMemcpyHostToDevice, h->stream));
    h->model_id_given = false;
    h->map_given = false;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return NQCD_OK;
}

int32_t map_load_variables(nqcd_handle_t* h, int64_t ntraj, const double* X, const double* P) {
    if (!X && !P) { h->map_given = false; return NQCD_OK; }
    if (!X || !P) return fail(h, NQCD_ERR_ARG, "mapping positions and momenta are given together");
    if (ntraj != h->ntraj) return fail(h, NQCD_ERR_ARG, "ntraj differs from the loaded initial conditions");
    const int n = h->prob.n_states;
    std::vector<double> t((size_t)2 * n * ntraj);
    for (int64_t k = 0; k < ntraj; ++k)
        for (int i = 0; i < n; ++i) {
            t[(size_t)i * ntraj + k] = X[(size_t)k * n + i];
            t[(size_t)(n + i) * ntraj + k] = P[(size_t)k * n + i];
        }
    HIPCHK(h, h->sig0.ensure(sizeof(double) * t.size()));
    HIPCHK(h, hipMemcpy(h->sig0.p, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice));
    h->map_given = true;
    return NQCD_OK;
}

int32_t map_run(nqcd_handle_t* h, double t0, int64_t nsteps, uint64_t key) {
    (void)t0;
    if (int32_t rc = map_check_models(h)) return rc;
    if (!h->model_set[0] && !h->model_id_given) return fail(h, NQCD_ERR_STATE, "model slot 0 is not set");
    const int n = h->prob.n_states;
    const int64_t nt = h->ntraj;
    const int ns = h->nsave;
    HIPCHK(h, h->r.ensure(sizeof(double) * nt));
    HIPCHK(h, h->v.ensure(sizeof(double) * nt));
    HIPCHK(h, h->sig.ensure(sizeof(double) * 2 * n * nt));
    HIPCHK(h, h->steps.ensure(sizeof(int32_t) * nt));
    HIPCHK(h, h->save_di.ensure(sizeof(double) * n * ns * nt));
    HIPCHK(h, h->save_ekin.ensure(sizeof(double) * ns * nt));
    HIPCHK(h, h->save_epot.ensure(sizeof(double) * ns * nt));
    HIPCHK(h, launch_zero(h->save_di.p, sizeof(double) * n * ns * nt, h->stream));
    HIPCHK(h, launch_zero(h->save_ekin.p, sizeof(double) * ns * nt, h->stream));
    HIPCHK(h, launch_zero(h->save_epot.p, sizeof(double) * ns * nt, h->stream));
    MapArgs A{};
    std::vector<double> tab;
    for (int s = 0; s < kMaxModels; ++s) {
        A.slot_tab[s] = -1;
        if (h->model_set[s] && h->models[s].kind == NQCD_MODEL_MORSE_N) {
            A.slot_tab[s] = (int32_t)tab.size();
            tab.insert(tab.end(), h->ns_tab[s].begin(), h->ns_tab[s].end());
        }
    }
    if (tab.empty()) tab.push_back(0.0);
    HIPCHK(h, h->ns_tab_dev.ensure(sizeof(double) * tab.size()));
    HIPCHK(h, hipMemcpyAsync(h->ns_tab_dev.p, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, h->models_dev.ensure(sizeof(ModelSlot) * kMaxModels));
    HIPCHK(h, hipMemcpyAsync(h->models_dev.p, h->models, sizeof(ModelSlot) * kMaxModels, hipMemcpyHostToDevice, h->stream));
    A.ntraj = nt;
    A.nsteps = nsteps;
    A.save_every = h->save_every;
    A.nsave = ns;
    A.dt = h->prob.dt;
    A.mass = h->masses[0];
    A.inv_mass = h->inv_mass[0];
    const bool ecmm = h->prob.method == NQCD_METHOD_ECMM;
    const double ge = h->prob.mapping_gamma, gs = map_gamma(n);
    A.ecmm = ecmm ? 1 : 0;
    A.gamma = ecmm ? 2.0 * ge : gs;
    A.gamma_force = ecmm ? ge : gs;
    A.pop_a = ecmm ? (1.0 + n) / ((1.0 + n * ge) * (1.0 + n * ge)) : 1.0;
    A.pop_b = ecmm ? (1.0 - ge) / (1.0 + n * ge) : 0.5 * gs;
    A.radius = std::sqrt(2.0 + 2.0 * n * ge);
    A.occ_radius = std::sqrt(2.0 + gs);
    A.free_radius = std::sqrt(gs);
    A.key = key;
    A.models = h->models_dev.as<ModelSlot>();
    A.mtab = h->ns_tab_dev.as<double>();
    A.r0 = h->r0.as<double>();
    A.v0 = h->v0.as<double>();
    A.X0 = h->map_given ? h->sig0.as<double>() : nullptr;
    A.P0 = h->map_given ? h->sig0.as<double>() + (size_t)n * nt : nullptr;
    A.state0 = h->state0.as<int32_t>();
    A.model_id = h->model_id_given ? h->model_id.as<int32_t>() : nullptr;
    A.r = h->r.as<double>();
    A.v = h->v.as<double>();
    A.X = h->sig.as<double>();
    A.P = h->sig.as<double>() + (size_t)n * nt;
    A.steps = h->steps.as<int32_t>();
    A.save_pop = h->save_di.as<double>();
    A.save_ekin = h->save_ekin.as<double>();
    A.save_epot = h->save_epot.as<double>();
    const int block = 256;
    int bpc = 0;
    switch (n) {
    case 2: HIPCHK(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, mapping_kernel<2>, block, 0)); break;
    case 3: HIPCHK(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, mapping_kernel<3>, block, 0)); break;
    default: HIPCHK(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, mapping_kernel<4>, block, 0)); break;
    }
    if (bpc < 1) bpc = 1;
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((nt + block - 1) / block, (int64_t)bpc * h->cu_count));
    HIPCHK(h, hipEventRecord(h->ev0, h->stream));
    switch (n) {
    case 2: hipLaunchKernelGGL(mapping_kernel<2>, dim3((unsigned)grid), dim3(block), 0, h->stream, A); break;
    case 3: hipLaunchKernelGGL(mapping_kernel<3>, dim3((unsigned)grid), dim3(block), 0, h->stream, A); break;
    default: hipLaunchKernelGGL(mapping_kernel<4>, dim3((unsigned)grid), dim3(block), 0, h->stream, A); break;
    }
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipEventRecord(h->ev1, h->stream));
    h->stats.block = block;
    h->stats.grid = (int)grid;
    h->stats.launches = 1;
    return NQCD_OK;
}

int32_t map_download(nqcd_handle_t* h, int32_t field, void* out, int64_t bytes) {
    const int n = h->prob.n_states;
    const int64_t nt = h->ntraj;
    const int ns = h->nsave;
    auto need = [&](int64_t b) -> int32_t {
        return b == bytes ? NQCD_OK : fail(h, NQCD_ERR_ARG, "bytes does not match the field size (" + std::to_string(b) + ")");
    };
    std::vector<double> tmp;
    auto fetch = [&](const void* src, size_t count) -> hipError_t {
        tmp.resize(count);
        return hipMemcpy(tmp.data(), src, sizeof(double) * count, hipMemcpyDeviceToHost);
    };
    double* o = static_cast<double*>(out);
    switch (field) {
    case NQCD_FIELD_R:
    case NQCD_FIELD_V: {
        if (int32_t rc = need(8 * nt)) return rc;
        HIPCHK(h, hipMemcpy(out, field == NQCD_FIELD_R ? h->r.p : h->v.p, 8 * nt, hipMemcpyDeviceToHost));
        return NQCD_OK;
    }
    case NQCD_FIELD_STEPS: {
        if (int32_t rc = need(4 * nt)) return rc;
        HIPCHK(h, hipMemcpy(out, h->steps.p, 4 * nt, hipMemcpyDeviceToHost));
        return NQCD_OK;
    }
    case NQCD_FIELD_MAP_X:
    case NQCD_FIELD_MAP_P: {
        if (int32_t rc = need(8 * (int64_t)n * nt)) return rc;
        HIPCHK(h, fetch(h->sig.as<double>() + (field == NQCD_FIELD_MAP_P ? (size_t)n * nt : 0), (size_t)n * nt));
        for (int64_t t = 0; t < nt; ++t)
            for (int i = 0; i < n; ++i) o[(size_t)t * n + i] = tmp[(size_t)i * nt + t];
        return NQCD_OK;
    }
    case NQCD_FIELD_SAVE_POP_DI: {
        if (int32_t rc = need(8 * (int64_t)n * ns * nt)) return rc;
        HIPCHK(h, fetch(h->save_di.p, (size_t)n * ns * nt));
        for (int64_t t = 0; t < nt; ++t)
            for (int s = 0; s < ns; ++s)
                for (int i = 0; i < n; ++i) o[((size_t)t * ns + s) * n + i] = tmp[((size_t)s * n + i) * nt + t];
        return NQCD_OK;
    }
    case NQCD_FIELD_SAVE_EKIN:
    case NQCD_FIELD_SAVE_EPOT: {
        if (int32_t rc = need(8 * (int64_t)ns * nt)) return rc;
        HIPCHK(h, fetch(field == NQCD_FIELD_SAVE_EKIN ? h->save_ekin.p : h->save_epot.p, (size_t)ns * nt));
        for (int64_t t = 0; t < nt; ++t)
            for (int s = 0; s < ns; ++s) o[(size_t)t * ns + s] = tmp[(size_t)s * nt + t];
        return NQCD_OK;
    }
    default:
        return fail(h, NQCD_ERR_UNSUPPORTED, "spin-mapping fields: R, V, STEPS, MAP_X, MAP_P, SAVE_POP_DI, SAVE_EKIN, SAVE_EPOT");
    }
}

}  // namespace nqcd
