// fssh_n.hip -- FSSH / Ehrenfest with BABwithTsit5 for any number of electronic states (up to 8) on
// one-coordinate models, and ensembles that mix state counts in one persistent launch.
//
// Restates bab_electronics.jl:48-91 for n states: velocity Verlet with the active-surface force
// (fssh.jl:67-74) or the mean-field force with the full U^T dV U (ehrenfest.jl:50-68), the n x n
// density propagated by Tsit5 under V(tau) = diag(E) - i (v.d) interpolated linearly between the
// step ends (electronic_dynamics.jl:15-79,104-140), then the HoppingCallback (fssh.jl:96-121,
// surface_hopping.jl:64-168).
//
// One trajectory per lane, all of its state in registers, lanes refilled from a work queue as
// trajectories terminate (as the two-state kernel).  The kernel is instantiated for a maximum state
// count NM (2 to 8); a trajectory whose model has n < NM states carries its n x n block in the
// leading rows and columns and zeros elsewhere, and every loop over states is bounded by its own n,
// so a launch mixes state counts (for instance two-state Tully and three-state Morse trajectories)
// and each trajectory follows exactly its own n-state dynamics.
//
// The density is carried as the real diagonal a_i and the real / imaginary parts x_ij, y_ij of the
// upper triangle (n^2 reals instead of the 2 n^2 of the complex matrix).  With W = v.d real
// antisymmetric, X = Re sigma and Y = Im sigma, -i[V, sigma] splits into
//   dx_ij = (E_i - E_j) y_ij - (M_ij + M_ji),  dy_ij = -(E_i - E_j) x_ij - (N_ij - N_ji),
//   da_i = -2 M_ii,   M = W X, N = W Y,
// an exact restatement of the reference's dense complex commutator (two real n x n products per
// right-hand side instead of two complex ones).
// The eigen-decomposition is cyclic Jacobi on the n x n model matrix with ascending eigenvalues, the
// largest-magnitude component of each fresh eigenvector positive and tracked columns kept at positive
// overlap with the previous step (NQCCalculators' eigenvector tracking).
#include "fssh_n_kern.h"

#include <algorithm>
#include <string>
#include <vector>

#include "ensemble.h"

namespace nqcd {

NQCD_NS_KERNELS(2)
NQCD_NS_KERNELS(3)
NQCD_NS_KERNELS(4)

__global__ void ns_queue_reset(unsigned long long* q) { *q = 0ull; }

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
int ns_kernel_states(int n) { return n < 2 ? 2 : n; }

static hipError_t ns_launch(int NM, int precision, int grid, int block, hipStream_t s, const FsshNArgs* Adev) {
    if (NM > 4) return ns_launch_hi(NM, precision, grid, block, s, Adev);
#define NQCD_CASE(X)                                                                                         \
    case X:                                                                                                  \
        if (precision == 0) hipLaunchKernelGGL(fssh_n_f64_##X, dim3(grid), dim3(block), 0, s, Adev);         \
        else hipLaunchKernelGGL(fssh_n_f32_##X, dim3(grid), dim3(block), 0, s, Adev);                        \
        break;
    switch (NM) {
        NQCD_CASE(2) NQCD_CASE(3) NQCD_CASE(4)
    default: return hipErrorInvalidValue;
    }
#undef NQCD_CASE
    return hipGetLastError();
}

static hipError_t ns_occupancy(int NM, int precision, int block, int* bpc) {
    if (NM > 4) return ns_occupancy_hi(NM, precision, block, bpc);
#define NQCD_CASE(X)                                                                                         \
    case X:                                                                                                  \
        return precision == 0 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(bpc, fssh_n_f64_##X, block, 0)  \
                              : hipOccupancyMaxActiveBlocksPerMultiprocessor(bpc, fssh_n_f32_##X, block, 0);
    switch (NM) {
        NQCD_CASE(2) NQCD_CASE(3) NQCD_CASE(4)
    default: return hipErrorInvalidValue;
    }
#undef NQCD_CASE
}

bool ns_is_problem(const nqcd_handle_t* h) {
    const nqcd_problem_desc& d = h->prob;
    if (!(d.method == NQCD_METHOD_FSSH || d.method == NQCD_METHOD_EHRENFEST) || d.algorithm != NQCD_ALG_BAB_TSIT5 ||
        h->D != 1 || d.nbeads != 1)
        return false;
    if (d.n_states > 2) return true;
    for (int s = 0; s < kMaxModels; ++s)
        if (h->model_set[s] && h->models[s].kind == NQCD_MODEL_MORSE_N) return true;
    return false;
}

int32_t ns_set_model(nqcd_handle_t* h, int32_t slot, const double* params, int64_t n) {
    if (n < 1 || !params) return fail(h, NQCD_ERR_ARG, "MORSE_N params: n, then the table");
    const int ns = (int)params[0];
    if (ns < 2 || ns > kNsMaxStates || (double)ns != params[0])
        return fail(h, NQCD_ERR_ARG, "MORSE_N: the state count must be an integer in 2..8");
    const int64_t need = 1 + 4 * ns + 3 * ns * (ns - 1) / 2;
    if (n != need)
        return fail(h, NQCD_ERR_ARG, "MORSE_N params: n, [D, beta, r, c] per state, [A, a, r] per pair i < j (" +
                                         std::to_string(need) + " values for n = " + std::to_string(ns) + ")");
    h->ns_tab[slot].assign(params + 1, params + n);
    ModelSlot m{};
    m.kind = NQCD_MODEL_MORSE_N;
    m.p[0] = (double)ns;
    h->models[slot] = m;
    h->model_set[slot] = true;
    return NQCD_OK;
}

static int slot_states(const nqcd_handle_t* h, int s) {
    return h->models[s].kind == NQCD_MODEL_MORSE_N ? (int)h->models[s].p[0] : 2;
}

int32_t ns_check_models(nqcd_handle_t* h) {
    for (int s = 0; s < kMaxModels; ++s) {
        if (!h->model_set[s]) continue;
        const int k = h->models[s].kind;
        if (k != NQCD_MODEL_MORSE_N && k != NQCD_MODEL_TULLY1 && k != NQCD_MODEL_TULLY2 && k != NQCD_MODEL_TULLY3 &&
            k != NQCD_MODEL_DW_GAUSSIAN)
            return fail(h, NQCD_ERR_UNSUPPORTED, "n-state FSSH / Ehrenfest models: TULLY1/2/3, DW_GAUSSIAN, MORSE_N");
        if (slot_states(h, s) > h->prob.n_states)
            return fail(h, NQCD_ERR_ARG, "a model slot has more states than the problem's n_states");
    }
    return NQCD_OK;
}

int32_t ns_load_initial(nqcd_handle_t* h, int64_t ntraj, const double* r, const double* v, const double* sig_re,
                        const double* sig_im, const int32_t* state, const int32_t* model_id) {
    if (int32_t rc = ns_check_models(h)) return rc;
    const int n = h->prob.n_states, NM = ns_kernel_states(n);
    h->ntraj = ntraj;
    // sigma (n, n, ntraj) column-major -> rows (i + NM j) of an NM x NM matrix per trajectory
    std::vector<double> sr((size_t)NM * NM * ntraj, 0.0), si((size_t)NM * NM * ntraj, 0.0);
    for (int64_t t = 0; t < ntraj; ++t) {
        const int mid = model_id ? model_id[t] : 0;
        const int nt = slot_states(h, mid);
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const double re = sig_re[(size_t)t * n * n + i + (size_t)n * j];
                const double im = sig_im[(size_t)t * n * n + i + (size_t)n * j];
                if ((i >= nt || j >= nt) && (re != 0.0 || im != 0.0))
                    return fail(h, NQCD_ERR_ARG, "density has population or coherence beyond the trajectory model's states");
                sr[(size_t)(i + NM * j) * ntraj + t] = re;
                si[(size_t)(i + NM * j) * ntraj + t] = im;
            }
        if (state && state[t] >= nt) return fail(h, NQCD_ERR_ARG, "state beyond the trajectory model's states");
    }
    HIPCHK(h, h->r0.ensure(sizeof(double) * ntraj));
    HIPCHK(h, h->v0.ensure(sizeof(double) * ntraj));
    HIPCHK(h, h->sig0.ensure(sizeof(double) * 2 * sr.size()));
    HIPCHK(h, hipMemcpyAsync(h->r0.p, r, sizeof(double) * ntraj, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->v0.p, v, sizeof(double) * ntraj, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->sig0.p, sr.data(), sizeof(double) * sr.size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->sig0.as<double>() + sr.size(), si.data(), sizeof(double) * si.size(),
                             hipMemcpyHostToDevice, h->stream));
    h->state0_given = state != nullptr;
    if (state) {
        HIPCHK(h, h->state0.ensure(sizeof(int32_t) * ntraj));
        HIPCHK(h, hipMemcpyAsync(h->state0.p, state, sizeof(int32_t) * ntraj, hipMemcpyHostToDevice, h->stream));
    }
    h->model_id_given = model_id != nullptr;
    std::vector<int32_t> ord(ntraj);
    {   // queue order grouped by model slot, so a wavefront refills with one state count
        int64_t pos = 0;
        for (int m = 0; m < kMaxModels; ++m)
            for (int64_t t = 0; t < ntraj; ++t)
                if ((model_id ? model_id[t] : 0) == m) ord[pos++] = (int32_t)t;
    }
    if (model_id) {
        HIPCHK(h, h->model_id.ensure(sizeof(int32_t) * ntraj));
        HIPCHK(h, hipMemcpyAsync(h->model_id.p, model_id, sizeof(int32_t) * ntraj, hipMemcpyHostToDevice, h->stream));
    }
    HIPCHK(h, h->order.ensure(sizeof(int32_t) * ntraj));
    HIPCHK(h, hipMemcpyAsync(h->order.p, ord.data(), sizeof(int32_t) * ntraj, hipMemcpyHostToDevice, h->stream));
    h->order_given = true;
    h->ns_init_diabatic = -1;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return NQCD_OK;
}

int32_t ns_sample_initial(nqcd_handle_t* h, const nqcd_sampler_desc* S, uint64_t seed) {
    if (int32_t rc = ns_check_models(h)) return rc;
    if (S->kind != NQCD_SAMPLER_GAUSSIAN || S->nparams != 4 || !S->params)
        return fail(h, NQCD_ERR_UNSUPPORTED, "n-state problems sample with GAUSSIAN (params mu_r, sd_r, mu_v, sd_v)");
    if (!h->model_set[0]) return fail(h, NQCD_ERR_STATE, "model slot 0 is not set");
    if (S->initial_state >= slot_states(h, 0)) return fail(h, NQCD_ERR_ARG, "initial_state beyond the model's states");
    const int64_t ntraj = S->ntraj;
    const int n = h->prob.n_states, NM = ns_kernel_states(n);
    h->ntraj = ntraj;
    HIPCHK(h, h->r0.ensure(sizeof(double) * ntraj));
    HIPCHK(h, h->v0.ensure(sizeof(double) * ntraj));
    HIPCHK(h, h->small.ensure(sizeof(double) * 4));
    HIPCHK(h, hipMemcpyAsync(h->small.p, S->params, sizeof(double) * 4, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, launch_sample_gaussian(ntraj, 1, h->small.as<double>(), seed, h->r0.as<double>(), h->v0.as<double>(),
                                     h->stream));
    std::vector<double> s0((size_t)2 * NM * NM * ntraj, 0.0);
    std::vector<int32_t> st(ntraj, S->initial_state);
    if (S->state_basis == 0)
        for (int64_t t = 0; t < ntraj; ++t) s0[(size_t)(S->initial_state * (NM + 1)) * ntraj + t] = 1.0;
    HIPCHK(h, h->sig0.ensure(sizeof(double) * s0.size()));
    HIPCHK(h, hipMemcpyAsync(h->sig0.p, s0.data(), sizeof(double) * s0.size(), hipMemcpyHostToDevice, h->stream));
    h->ns_init_diabatic = S->state_basis == 1 ? S->initial_state : -1;
    h->state0_given = S->state_basis == 0;  // the diabatic start draws the active state (fssh.jl:54)
    if (h->state0_given) {
        HIPCHK(h, h->state0.ensure(sizeof(int32_t) * ntraj));
        HIPCHK(h, hipMemcpyAsync(h->state0.p, st.data(), sizeof(int32_t) * ntraj, hipMemcpyHostToDevice, h->stream));
    }
    h->model_id_given = false;
    h->order_given = false;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return NQCD_OK;
}

int32_t ns_run(nqcd_handle_t* h, double t0, int64_t nsteps, const nqcd_term_desc* term, uint64_t key) {
    if (int32_t rc = ns_check_models(h)) return rc;
    const int n = h->prob.n_states, NM = ns_kernel_states(n);
    const int64_t nt = h->ntraj;
    HIPCHK(h, h->r.ensure(sizeof(double) * nt));
    HIPCHK(h, h->v.ensure(sizeof(double) * nt));
    HIPCHK(h, h->sig.ensure(sizeof(double) * 2 * NM * NM * nt));
    HIPCHK(h, h->state.ensure(sizeof(int32_t) * nt));
    HIPCHK(h, h->steps.ensure(sizeof(int32_t) * nt));
    HIPCHK(h, h->nhops.ensure(sizeof(int32_t) * nt));
    HIPCHK(h, h->save_ad.ensure(sizeof(double) * NM * h->nsave * nt));
    HIPCHK(h, h->save_di.ensure(sizeof(double) * NM * h->nsave * nt));
    HIPCHK(h, h->save_state.ensure(sizeof(int32_t) * h->nsave * nt));
    HIPCHK(h, h->queue.ensure(sizeof(unsigned long long)));
    HIPCHK(h, launch_zero(h->save_ad.p, sizeof(double) * NM * h->nsave * nt, h->stream));
    HIPCHK(h, launch_zero(h->save_di.p, sizeof(double) * NM * h->nsave * nt, h->stream));
    HIPCHK(h, launch_zero(h->save_state.p, sizeof(int32_t) * h->nsave * nt, h->stream));
    // model tables: MORSE_N slots in one buffer, offsets per slot
    std::vector<double> tab;
    FsshNArgs A{};
    for (int s = 0; s < kMaxModels; ++s) {
        A.slot_n[s] = h->model_set[s] ? slot_states(h, s) : 2;
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
    A.method = h->prob.method == NQCD_METHOD_FSSH ? 0 : 1;
    A.rescaling = h->prob.rescaling;
    A.nsub = h->prob.electronic_substeps > 0 ? h->prob.electronic_substeps : 5;
    A.save_every = h->save_every;
    A.nsave = h->nsave;
    A.term_kind = term ? term->kind : 0;
    A.term_value = term ? term->value : 0.0;
    A.cell = h->cell;
    A.init_diabatic = h->ns_init_diabatic;
    A.dt = h->prob.dt;
    A.t0 = t0;
    A.nsteps = nsteps;
    A.key = key;
    A.inv_mass = h->inv_mass[0];
    A.models = h->models_dev.as<ModelSlot>();
    A.mtab = h->ns_tab_dev.as<double>();
    A.r0 = h->r0.as<double>(); A.v0 = h->v0.as<double>();
    A.sig0_re = h->sig0.as<double>(); A.sig0_im = h->sig0.as<double>() + (size_t)NM * NM * nt;
    A.state0 = h->state0_given ? h->state0.as<int32_t>() : nullptr;
    A.model_id = h->model_id_given ? h->model_id.as<int32_t>() : nullptr;
    A.order = h->order_given ? h->order.as<int32_t>() : nullptr;
    A.r = h->r.as<double>(); A.v = h->v.as<double>();
    A.sig_re = h->sig.as<double>(); A.sig_im = h->sig.as<double>() + (size_t)NM * NM * nt;
    A.state = h->state.as<int32_t>(); A.steps = h->steps.as<int32_t>(); A.nhops = h->nhops.as<int32_t>();
    A.save_pop_ad = h->save_ad.as<double>(); A.save_pop_di = h->save_di.as<double>();
    A.save_state = h->save_state.as<int32_t>();
    A.queue = h->queue.as<unsigned long long>();
    const int block = 256;
    int bpc = 0;
    HIPCHK(h, ns_occupancy(NM, h->prob.precision, block, &bpc));
    if (bpc < 1) bpc = 1;
    const int64_t lanes = (int64_t)bpc * h->cu_count * block;
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((nt + block - 1) / block, lanes / block));
    HIPCHK(h, h->args_dev.ensure(sizeof(FsshNArgs)));
    HIPCHK(h, hipMemcpyAsync(h->args_dev.p, &A, sizeof(FsshNArgs), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipEventRecord(h->ev0, h->stream));
    hipLaunchKernelGGL(ns_queue_reset, dim3(1), dim3(1), 0, h->stream, A.queue);
    HIPCHK(h, ns_launch(NM, h->prob.precision, (int)grid, block, h->stream, h->args_dev.as<const FsshNArgs>()));
    HIPCHK(h, hipEventRecord(h->ev1, h->stream));
    h->stats.block = block;
    h->stats.grid = (int)grid;
    h->stats.launches = 2;
    return NQCD_OK;
}

int32_t ns_download(nqcd_handle_t* h, int32_t field, void* out, int64_t bytes) {
    const int n = h->prob.n_states, NM = ns_kernel_states(n);
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
    switch (field) {
    case NQCD_FIELD_R:
    case NQCD_FIELD_V: {
        if (int32_t rc = need(8 * nt)) return rc;
        HIPCHK(h, hipMemcpy(out, field == NQCD_FIELD_R ? h->r.p : h->v.p, 8 * nt, hipMemcpyDeviceToHost));
        return NQCD_OK;
    }
    case NQCD_FIELD_SIGMA_RE:
    case NQCD_FIELD_SIGMA_IM: {
        if (int32_t rc = need(8 * (int64_t)n * n * nt)) return rc;
        const double* src = h->sig.as<double>() + (field == NQCD_FIELD_SIGMA_IM ? (size_t)NM * NM * nt : 0);
        HIPCHK(h, fetch(src, (size_t)NM * NM * nt));
        double* o = static_cast<double*>(out);
        for (int64_t t = 0; t < nt; ++t)
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < n; ++i) o[(size_t)t * n * n + i + (size_t)n * j] = tmp[(size_t)(i + NM * j) * nt + t];
        return NQCD_OK;
    }
    case NQCD_FIELD_STATE:
    case NQCD_FIELD_STEPS:
    case NQCD_FIELD_NHOPS: {
        if (int32_t rc = need(4 * nt)) return rc;
        const DevBuf& b = field == NQCD_FIELD_STATE ? h->state : field == NQCD_FIELD_STEPS ? h->steps : h->nhops;
        HIPCHK(h, hipMemcpy(out, b.p, 4 * nt, hipMemcpyDeviceToHost));
        return NQCD_OK;
    }
    case NQCD_FIELD_SAVE_POP_AD:
    case NQCD_FIELD_SAVE_POP_DI: {
        if (int32_t rc = need(8 * (int64_t)n * ns * nt)) return rc;
        HIPCHK(h, fetch(field == NQCD_FIELD_SAVE_POP_AD ? h->save_ad.p : h->save_di.p, (size_t)NM * ns * nt));
        double* o = static_cast<double*>(out);
        for (int64_t t = 0; t < nt; ++t)
            for (int s = 0; s < ns; ++s)
                for (int i = 0; i < n; ++i) o[((size_t)t * ns + s) * n + i] = tmp[((size_t)s * NM + i) * nt + t];
        return NQCD_OK;
    }
    case NQCD_FIELD_SAVE_STATE: {
        if (int32_t rc = need(4 * (int64_t)ns * nt)) return rc;
        std::vector<int32_t> ti((size_t)ns * nt);
        HIPCHK(h, hipMemcpy(ti.data(), h->save_state.p, 4 * (size_t)ns * nt, hipMemcpyDeviceToHost));
        int32_t* o = static_cast<int32_t*>(out);
        for (int64_t t = 0; t < nt; ++t)
            for (int s = 0; s < ns; ++s) o[(size_t)t * ns + s] = ti[(size_t)s * nt + t];
        return NQCD_OK;
    }
    default: return fail(h, NQCD_ERR_UNSUPPORTED, "field not available for n-state problems");
    }
}

}  // namespace nqcd
