// abi.cpp -- implementation of the C ABI declared in include/nqcd_mi355x.h.
//
// Host side of EnsembleMI355X: owns the device buffers of one ensemble, converts
// between the reference's column-major (dofs, atoms, traj) host layout and the
// device's structure-of-arrays layout, dispatches the trajectory kernels and
// the deterministic reductions, and times the trajectory phase with HIP events.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/nqcd_mi355x.h"
#include "ensemble.h"
#include "fssh2.h"
#include "handle.h"
#include "spin_boson.h"
#include "ring_polymer.h"
#include "iesh.h"
#include "fssh_nd.h"
#include "fssh_n.h"
#include "classical.h"
#include "mapping.h"
#include "cme.h"
#include "mlp_host.h"
#include "autotune.h"

using namespace nqcd;

namespace nqcd {

thread_local std::string g_tls_error;

int32_t fail(nqcd_handle_t* h, int32_t code, const std::string& msg) {
    if (h) h->err = msg; else g_tls_error = msg;
    return code;
}

}  // namespace nqcd

namespace {

bool is_two_state_1d(const nqcd_handle* h) {
    return (h->prob.method == NQCD_METHOD_FSSH || h->prob.method == NQCD_METHOD_EHRENFEST) &&
           h->prob.algorithm == NQCD_ALG_BAB_TSIT5 && h->prob.n_states == 2 && h->D == 1 && !ns_is_problem(h);
}

bool is_spin_boson(const nqcd_handle* h) {
    return (h->prob.method == NQCD_METHOD_FSSH || h->prob.method == NQCD_METHOD_EHRENFEST) &&
           h->prob.algorithm == NQCD_ALG_BAB_TSIT5 && h->prob.n_states == 2 && h->model_set[0] &&
           h->models[0].kind == MODEL_SPIN_BOSON;
}

bool is_rp(const nqcd_handle* h) {
    return h->prob.method == NQCD_METHOD_RPMD || h->prob.method == NQCD_METHOD_RPSH ||
           h->prob.method == NQCD_METHOD_RP_EHRENFEST;
}

bool is_ie(const nqcd_handle* h) {
    return h->prob.method == NQCD_METHOD_IESH || h->prob.method == NQCD_METHOD_EHRENFEST_NA;
}

// rows of the per-trajectory position / velocity arrays (dofs x atoms x beads)
int64_t nqcd_rows(const nqcd_handle* h) { return (int64_t)h->D * (is_rp(h) ? h->prob.nbeads : 1); }

int nsub_of(const nqcd_handle* h) { return h->prob.electronic_substeps > 0 ? h->prob.electronic_substeps : 5; }

// Queue order: stable grouping by model so a wavefront refills with one model.
std::vector<int32_t> order_by_model(const std::vector<int32_t>& mid, int64_t ntraj) {
    std::vector<int32_t> ord(ntraj);
    int64_t pos = 0;
    for (int m = 0; m < kMaxModels; ++m)
        for (int64_t i = 0; i < ntraj; ++i)
            if ((mid.empty() ? 0 : mid[i]) == m) ord[pos++] = (int32_t)i;
    return ord;
}

int32_t alloc_results(nqcd_handle* h) {
    const int64_t n = h->ntraj;
    const int D = h->D;
    HIPCHK(h, h->r.ensure(sizeof(double) * D * n));
    HIPCHK(h, h->v.ensure(sizeof(double) * D * n));
    HIPCHK(h, h->sig.ensure(sizeof(double) * 4 * n));
    HIPCHK(h, h->state.ensure(sizeof(int32_t) * n));
    HIPCHK(h, h->steps.ensure(sizeof(int32_t) * n));
    HIPCHK(h, h->nhops.ensure(sizeof(int32_t) * n));
    HIPCHK(h, h->save_ad.ensure(sizeof(double) * 2 * h->nsave * n));
    HIPCHK(h, h->save_di.ensure(sizeof(double) * 2 * h->nsave * n));
    HIPCHK(h, h->save_state.ensure(sizeof(int32_t) * h->nsave * n));
    HIPCHK(h, h->queue.ensure(sizeof(unsigned long long)));
    HIPCHK(h, launch_zero(h->save_ad.p, sizeof(double) * 2 * h->nsave * n, h->stream));
    HIPCHK(h, launch_zero(h->save_di.p, sizeof(double) * 2 * h->nsave * n, h->stream));
    HIPCHK(h, launch_zero(h->save_state.p, sizeof(int32_t) * h->nsave * n, h->stream));
    HIPCHK(h, h->save_ekin.ensure(sizeof(double) * h->nsave * n));
    HIPCHK(h, h->save_epot.ensure(sizeof(double) * h->nsave * n));
    HIPCHK(h, launch_zero(h->save_ekin.p, sizeof(double) * h->nsave * n, h->stream));
    HIPCHK(h, launch_zero(h->save_epot.p, sizeof(double) * h->nsave * n, h->stream));
    HIPCHK(h, h->final_di.ensure(sizeof(double) * 2 * n));
    // constant tables read by the kernels (models per slot, 1/m per flattened dof)
    HIPCHK(h, h->models_dev.ensure(sizeof(ModelSlot) * kMaxModels));
    HIPCHK(h, h->inv_mass_dev.ensure(sizeof(double) * D));
    HIPCHK(h, hipMemcpyAsync(h->models_dev.p, h->models, sizeof(ModelSlot) * kMaxModels, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->inv_mass_dev.p, h->inv_mass.data(), sizeof(double) * D, hipMemcpyHostToDevice, h->stream));
    return NQCD_OK;
}

Fssh2Args make_args(nqcd_handle* h, double t0, int64_t nsteps, const nqcd_term_desc* term, uint64_t key) {
    Fssh2Args A{};
    A.ntraj = h->ntraj;
    A.method = (h->prob.method == NQCD_METHOD_FSSH) ? METHOD_FSSH : METHOD_EHRENFEST;
    A.rescaling = h->prob.rescaling;
    A.nsub = nsub_of(h);
    A.save_every = h->save_every;
    A.nsave = h->nsave;
    A.term_kind = term ? term->kind : 0;
    A.term_dof = term ? term->dof : 0;
    A.term_value = term ? term->value : 0.0;
    A.cell = h->cell;
    A.dt = h->prob.dt;
    A.t0 = t0;
    A.nsteps = nsteps;
    A.key = key;
    int nm = 0;
    for (int m = 0; m < kMaxModels; ++m) if (h->model_set[m]) nm = m + 1;
    A.nmodels = nm;
    A.inv_mass = h->inv_mass_dev.as<double>();
    A.models = h->models_dev.as<ModelSlot>();
    A.r0 = h->r0.as<double>(); A.v0 = h->v0.as<double>(); A.sig0 = h->sig0.as<double>();
    A.state0 = h->state0_given ? h->state0.as<int32_t>() : nullptr;
    A.model_id = h->model_id_given ? h->model_id.as<int32_t>() : nullptr;
    A.order = h->order_given ? h->order.as<int32_t>() : nullptr;
    A.r = h->r.as<double>(); A.v = h->v.as<double>(); A.sig = h->sig.as<double>();
    A.state = h->state.as<int32_t>(); A.steps = h->steps.as<int32_t>(); A.nhops = h->nhops.as<int32_t>();
    A.save_pop_ad = h->save_ad.as<double>(); A.save_pop_di = h->save_di.as<double>();
    A.save_state = h->save_state.as<int32_t>();
    A.save_ekin = h->save_ekin.as<double>(); A.save_epot = h->save_epot.as<double>();
    A.final_pop_di = h->final_di.as<double>();
    A.queue = h->queue.as<unsigned long long>();
    return A;
}

}  // namespace

extern "C" {

int32_t nqcd_mi355x_abi_version(void) { return NQCD_MI355X_ABI_VERSION; }

int32_t nqcd_mi355x_test_broadening(nqcd_handle_t* h, const double* hs, int64_t count, double mu, double kT, double Gamma,
                                    double W, double* n, double* n2) {
    if (!h || !hs || !n || !n2 || count < 1) return fail(h, NQCD_ERR_ARG, "null argument or empty list");
    if (!(kT > 0.0) || !(Gamma > 0.0) || !(W > 0.0)) return fail(h, NQCD_ERR_ARG, "kT, Gamma and W must be positive");
    HIPCHK(h, hipSetDevice(h->device));
    return cme_test_broadening(h, hs, count, mu, kT, Gamma, W, n, n2);
}

const char* nqcd_mi355x_last_error(nqcd_handle_t* h) { return h ? h->err.c_str() : g_tls_error.c_str(); }

int32_t nqcd_mi355x_open(int32_t device, nqcd_handle_t** out) {
    if (!out) return fail(nullptr, NQCD_ERR_ARG, "out is null");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        return fail(nullptr, NQCD_ERR_DEVICE, std::string("no HIP device: ") + hipGetErrorString(e));
    if (device < 0 || device >= count) return fail(nullptr, NQCD_ERR_ARG, "device index out of range");
    e = hipSetDevice(device);
    if (e != hipSuccess) return fail(nullptr, NQCD_ERR_DEVICE, hipGetErrorString(e));
    // address of this device's scratch array for nqcd_mi355x_reduce_noalloc (resolved once per device)
    e = reduce_scratch_init(device);
    if (e != hipSuccess)
        return fail(nullptr, NQCD_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    nqcd_handle* h = new nqcd_handle();
    h->device = device;
    if (hipStreamCreate(&h->stream) != hipSuccess || hipEventCreate(&h->ev0) != hipSuccess ||
        hipEventCreate(&h->ev1) != hipSuccess) {
        delete h;
        return fail(nullptr, NQCD_ERR_DEVICE, "stream/event creation failed");
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) h->cu_count = prop.multiProcessorCount;
    if (h->cu_count <= 0) h->cu_count = 256;
    *out = h;
    return NQCD_OK;
}

void nqcd_mi355x_close(nqcd_handle_t* h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
    DevBuf* bufs[] = {&h->mlp_master, &h->mlp_packed, &h->mlp_adam_m, &h->mlp_adam_v, &h->mlp_x, &h->mlp_t, &h->mlp_idx,
                      &h->mlp_partial, &h->mlp_losspart, &h->mlp_lossbuf, &h->mlp_grad, &h->fallbacks,
                      &h->r0, &h->v0, &h->sig0, &h->state0, &h->model_id, &h->order, &h->models_dev, &h->bath, &h->inv_mass_dev, &h->args_dev,
                      &h->r, &h->v, &h->sig, &h->state, &h->steps, &h->nhops, &h->save_ad, &h->save_di,
                      &h->save_state, &h->queue, &h->plainS, &h->plainSi, &h->tmp, &h->partial, &h->small, &h->save_centroid, &h->sum64,
                      &h->save_ekin, &h->save_epot, &h->ns_tab_dev, &h->final_di, &h->save_v, &h->susp_f[0],
                      &h->susp_f[1], &h->susp_tid[0], &h->susp_tid[1], &h->susp_i[0], &h->susp_i[1], &h->susp_ctr};
    for (DevBuf* b : bufs) b->release();
    (void)hipEventDestroy(h->ev0);
    (void)hipEventDestroy(h->ev1);
    (void)hipStreamDestroy(h->stream);
    delete h;
}

int32_t nqcd_mi355x_set_problem(nqcd_handle_t* h, const nqcd_problem_desc* d) {
    if (!h || !d) return fail(h, NQCD_ERR_ARG, "null argument");
    if (d->n_states < 1 || d->ndofs < 1 || d->natoms < 1 || d->nbeads < 1)
        return fail(h, NQCD_ERR_ARG, "n_states, ndofs, natoms and nbeads must be >= 1");
    if (!(d->dt > 0.0) || !std::isfinite(d->dt)) return fail(h, NQCD_ERR_ARG, "dt must be positive and finite");
    if (!d->masses) return fail(h, NQCD_ERR_ARG, "masses is null");
    if (d->rescaling < 0 || d->rescaling > 2) return fail(h, NQCD_ERR_ARG, "This mode of rescaling is not implemented");
    if (d->precision < 0 || d->precision > 1) return fail(h, NQCD_ERR_ARG, "unknown precision");
    if (d->schedule < 0 || d->schedule > 1) return fail(h, NQCD_ERR_ARG, "unknown schedule");
    const bool sh_or_eh = d->method == NQCD_METHOD_FSSH || d->method == NQCD_METHOD_EHRENFEST;
    const bool rp = d->method == NQCD_METHOD_RPMD || d->method == NQCD_METHOD_RPSH || d->method == NQCD_METHOD_RP_EHRENFEST;
    const bool ie = d->method == NQCD_METHOD_IESH || d->method == NQCD_METHOD_EHRENFEST_NA;
    const bool cme = d->method == NQCD_METHOD_CME || d->method == NQCD_METHOD_BCME;
    const bool cls = d->method == NQCD_METHOD_CLASSICAL;
    const bool map = d->method == NQCD_METHOD_SPIN_MAPPING || d->method == NQCD_METHOD_ECMM || d->method == NQCD_METHOD_NRPMD;
    if (!rp && !ie && !cme && !cls && !map && !(sh_or_eh && d->algorithm == NQCD_ALG_BAB_TSIT5 && d->nbeads == 1 && d->n_states >= 2 &&
                        d->n_states <= kNsMaxStates))
        return fail(h, NQCD_ERR_UNSUPPORTED,
                    "EnsembleMI355X supports Simulation{FSSH|Ehrenfest} (2 to 8 states, BABwithTsit5), "
                    "Simulation{AdiabaticIESH|EhrenfestNA} (VerletwithElectronics), Simulation{CME|BCME}, "
                    "Simulation{Classical} (VelocityVerlet), Simulation{SpinMappingW|eCMM} (MInt), RingPolymerSimulation{NRPMD|eCMM} (RingPolymerMInt) and "
                    "RingPolymerSimulation{Classical|FSSH|Ehrenfest} (BCB / BCBwithTsit5)");
    if (sh_or_eh && d->n_states > 2 && d->ndofs * d->natoms != 1)
        return fail(h, NQCD_ERR_UNSUPPORTED, "FSSH / Ehrenfest with more than two states: one nuclear coordinate");
    h->prob = *d;
    h->masses.assign(d->masses, d->masses + d->natoms);
    for (double m : h->masses)
        if (!(m > 0.0)) return fail(h, NQCD_ERR_ARG, "masses must be positive");
    h->D = d->ndofs * d->natoms;
    h->inv_mass.resize(h->D);
    for (int I = 0; I < h->D; ++I) h->inv_mass[I] = 1.0 / h->masses[I / d->ndofs];
    h->prob.masses = nullptr;
    if (rp) {
        int32_t rc = rp_check_problem(h);
        if (rc != NQCD_OK) return rc;
    }
    if (ie) {
        int32_t rc = ie_check_problem(h);
        if (rc != NQCD_OK) return rc;
    }
    if (cme) {
        int32_t rc = cme_check_problem(h);
        if (rc != NQCD_OK) return rc;
    }
    if (cls) {
        int32_t rc = cls_check_problem(h);
        if (rc != NQCD_OK) return rc;
    }
    if (map) {
        int32_t rc = map_check_problem(h);
        if (rc != NQCD_OK) return rc;
    }
    h->map_given = false;
    h->cell = CellDesc{};
    h->have_problem = true;
    h->have_initial = h->have_results = false;
    return NQCD_OK;
}

int32_t nqcd_mi355x_set_cell(nqcd_handle_t* h, int32_t mode, int32_t ndofs, const double* vectors, const int32_t* periodic) {
    if (!h) return fail(h, NQCD_ERR_ARG, "null handle");
    if (!h->have_problem) return fail(h, NQCD_ERR_STATE, "set_problem must be called first");
    if (mode == NQCD_CELL_NONE) { h->cell = CellDesc{}; return NQCD_OK; }
    if (mode != NQCD_CELL_BOUNDARY && mode != NQCD_CELL_TERMINATE) return fail(h, NQCD_ERR_ARG, "unknown cell mode");
    if (ndofs != h->prob.ndofs || ndofs < 1 || ndofs > 3)
        return fail(h, NQCD_ERR_ARG, "the cell has ndofs x ndofs vectors (ndofs of the problem, at most 3)");
    if (!vectors) return fail(h, NQCD_ERR_ARG, "vectors is null");
    CellDesc c{};
    c.mode = mode;
    c.nd = ndofs;
    for (int j = 0; j < ndofs; ++j) c.per[j] = periodic ? (periodic[j] != 0) : 1;
    for (int j = 0; j < ndofs; ++j)
        for (int k = 0; k < ndofs; ++k) c.a[j + 3 * k] = vectors[j + ndofs * k];
    // inverse by cofactors (the cell matrix is at most 3 x 3)
    double det;
    if (ndofs == 1) {
        det = c.a[0];
        if (det != 0.0) c.inv[0] = 1.0 / det;
    } else if (ndofs == 2) {
        det = c.a[0] * c.a[4] - c.a[3] * c.a[1];
        if (det != 0.0) { c.inv[0] = c.a[4] / det; c.inv[3] = -c.a[3] / det; c.inv[1] = -c.a[1] / det; c.inv[4] = c.a[0] / det; }
    } else {
        auto A = [&](int i, int j) { return c.a[i + 3 * j]; };
        const double c00 = A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1), c01 = A(1, 2) * A(2, 0) - A(1, 0) * A(2, 2),
                     c02 = A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0);
        det = A(0, 0) * c00 + A(0, 1) * c01 + A(0, 2) * c02;
        if (det != 0.0) {
            // inv(i, j) = cofactor(j, i) / det
            c.inv[0 + 3 * 0] = c00 / det; c.inv[1 + 3 * 0] = c01 / det; c.inv[2 + 3 * 0] = c02 / det;
            c.inv[0 + 3 * 1] = (A(0, 2) * A(2, 1) - A(0, 1) * A(2, 2)) / det;
            c.inv[1 + 3 * 1] = (A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0)) / det;
            c.inv[2 + 3 * 1] = (A(0, 1) * A(2, 0) - A(0, 0) * A(2, 1)) / det;
            c.inv[0 + 3 * 2] = (A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1)) / det;
            c.inv[1 + 3 * 2] = (A(0, 2) * A(1, 0) - A(0, 0) * A(1, 2)) / det;
            c.inv[2 + 3 * 2] = (A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0)) / det;
        }
    }
    if (!(std::fabs(det) > 0.0) || !std::isfinite(det)) return fail(h, NQCD_ERR_ARG, "the cell vectors are singular");
    h->cell = c;
    return NQCD_OK;
}

int32_t nqcd_mi355x_set_model(nqcd_handle_t* h, int32_t slot, int32_t kind, const double* params, int64_t n) {
    if (!h) return fail(h, NQCD_ERR_ARG, "null handle");
    if (slot < 0 || slot >= kMaxModels) return fail(h, NQCD_ERR_ARG, "model slot must be in 0..3");
    if (n < 0 || (n > 0 && !params)) return fail(h, NQCD_ERR_ARG, "params");
    ModelSlot m{};
    m.kind = kind;
    int need = -1;
    switch (kind) {
    case NQCD_MODEL_TULLY1: need = 4; break;
    case NQCD_MODEL_TULLY2: need = 5; break;
    case NQCD_MODEL_TULLY3: need = 3; break;
    case NQCD_MODEL_DW_GAUSSIAN: need = 5; break;
    case NQCD_MODEL_HARMONIC: need = 3; break;
    case NQCD_MODEL_ANDERSON_HOLSTEIN_1: need = 4; break;
    case NQCD_MODEL_NEWNS_ANDERSON: return ie_set_model(h, slot, params, n);
    case NQCD_MODEL_MORSE_N: return ns_set_model(h, slot, params, n);
    case NQCD_MODEL_DIATOM_SURFACE:
    case NQCD_MODEL_MLP_SURROGATE:
        if (slot != 0) return fail(h, NQCD_ERR_UNSUPPORTED, "the diatom-surface / surrogate model must use slot 0");
        return nd_set_model(h, kind, params, n);
    case NQCD_MODEL_SPIN_BOSON: {
        if (slot != 0) return fail(h, NQCD_ERR_UNSUPPORTED, "spin-boson model must use slot 0");
        if (n < 4 || (n - 2) % 2 != 0) return fail(h, NQCD_ERR_ARG, "spin-boson params: eps, Delta, omega[nb], c[nb]");
        const int64_t nb = (n - 2) / 2;
        h->sb_eps = params[0]; h->sb_delta = params[1];
        h->sb_omega.assign(params + 2, params + 2 + nb);
        h->sb_c.assign(params + 2 + nb, params + 2 + 2 * nb);
        m.p[0] = params[0]; m.p[1] = params[1]; m.p[2] = (double)nb;
        h->models[slot] = m;
        h->model_set[slot] = true;
        return NQCD_OK;
    }
    default: return fail(h, NQCD_ERR_UNSUPPORTED, "unknown model kind");
    }
    if (n != need) return fail(h, NQCD_ERR_ARG, "wrong number of model parameters");
    for (int i = 0; i < need; ++i) m.p[i] = params[i];
    h->models[slot] = m;
    h->model_set[slot] = true;
    return NQCD_OK;
}

int32_t nqcd_mi355x_load_initial(nqcd_handle_t* h, int64_t ntraj, const double* r, const double* v,
                                 const double* sig_re, const double* sig_im, const int32_t* state,
                                 const int32_t* model_id) {
    if (!h) return fail(h, NQCD_ERR_ARG, "null handle");
    if (!h->have_problem) return fail(h, NQCD_ERR_STATE, "set_problem must be called first");
    if (ntraj < 1 || ntraj > (int64_t)1 << 31) return fail(h, NQCD_ERR_ARG, "ntraj out of range");
    if (is_ie(h)) {
        if (!r || !v) return fail(h, NQCD_ERR_ARG, "null input array");
        HIPCHK(h, hipSetDevice(h->device));
        int32_t rc = ie_load_initial(h, ntraj, r, v, sig_re, sig_im, state);
        if (rc == NQCD_OK) { h->have_initial = true; h->sb_initial = false; h->have_results = false; }
        return rc;
    }
    if (!r || !v || ((!sig_re || !sig_im) && h->prob.method != NQCD_METHOD_RPMD && !cme_is_problem(h) &&
                     !cls_is_problem(h) && !map_is_problem(h)))
        return fail(h, NQCD_ERR_ARG, "null input array");
    if (cls_is_problem(h)) {
        if (!h->model_set[0]) return fail(h, NQCD_ERR_STATE, "model slot 0 is not set");
        HIPCHK(h, hipSetDevice(h->device));
        int32_t rc = cls_load_initial(h, ntraj, r, v);
        if (rc == NQCD_OK) { h->have_initial = true; h->sb_initial = false; h->have_results = false; }
        return rc;
    }
    const int D = h->D, n = h->prob.n_states;
    if (model_id)
        for (int64_t i = 0; i < ntraj; ++i)
            if (model_id[i] < 0 || model_id[i] >= kMaxModels || !h->model_set[model_id[i]])
                return fail(h, NQCD_ERR_ARG, "model_id refers to an unset model slot");
    if (!h->model_set[0] && !model_id) return fail(h, NQCD_ERR_STATE, "model slot 0 is not set");
    if (state)
        for (int64_t i = 0; i < ntraj; ++i)
            if (state[i] < -1 || state[i] >= n) return fail(h, NQCD_ERR_ARG, "state out of range");
    if (map_is_problem(h)) {
        HIPCHK(h, hipSetDevice(h->device));
        int32_t rc = map_load_initial(h, ntraj, r, v, state, model_id);
        if (rc == NQCD_OK) { h->have_initial = true; h->sb_initial = false; h->have_results = false; }
        return rc;
    }
    if (cme_is_problem(h)) {
        if (!r || !v) return fail(h, NQCD_ERR_ARG, "null input array");
        HIPCHK(h, hipSetDevice(h->device));
        int32_t rc = cme_load_initial(h, ntraj, r, v, state);
        if (rc == NQCD_OK) { h->have_initial = true; h->sb_initial = false; h->have_results = false; }
        return rc;
    }
    if (ns_is_problem(h)) {
        HIPCHK(h, hipSetDevice(h->device));
        int32_t rc = ns_load_initial(h, ntraj, r, v, sig_re, sig_im, state, model_id);
        if (rc == NQCD_OK) { h->have_initial = true; h->sb_initial = false; h->have_results = false; }
        return rc;
    }
    if (is_rp(h)) {
        HIPCHK(h, hipSetDevice(h->device));
        int32_t rc = rp_load_initial(h, ntraj, r, v, sig_re, sig_im, state);
        if (rc == NQCD_OK) { h->have_initial = true; h->sb_initial = false; h->have_results = false; }
        return rc;
    }
    h->ntraj = ntraj;
    std::vector<double> tr((size_t)D * ntraj), tv((size_t)D * ntraj), ts((size_t)4 * ntraj);
    for (int64_t i = 0; i < ntraj; ++i)
        for (int I = 0; I < D; ++I) {
            tr[(size_t)I * ntraj + i] = r[(size_t)i * D + I];
            tv[(size_t)I * ntraj + i] = v[(size_t)i * D + I];
        }
    for (int64_t i = 0; i < ntraj; ++i) {  // (n, n, ntraj) column-major -> compact Hermitian rows
        const size_t o = (size_t)i * n * n;
        ts[i] = sig_re[o + 0];
        ts[ntraj + i] = sig_re[o + 3];
        ts[2 * ntraj + i] = sig_re[o + 2];  // element (1,2) in 1-based = row 0, col 1
        ts[3 * ntraj + i] = sig_im[o + 2];
    }
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, h->r0.ensure(sizeof(double) * tr.size()));
    HIPCHK(h, h->v0.ensure(sizeof(double) * tv.size()));
    HIPCHK(h, h->sig0.ensure(sizeof(double) * ts.size()));
    HIPCHK(h, hipMemcpyAsync(h->r0.p, tr.data(), sizeof(double) * tr.size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->v0.p, tv.data(), sizeof(double) * tv.size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->sig0.p, ts.data(), sizeof(double) * ts.size(), hipMemcpyHostToDevice, h->stream));
    h->state0_given = state != nullptr;
    if (state) {
        HIPCHK(h, h->state0.ensure(sizeof(int32_t) * ntraj));
        HIPCHK(h, hipMemcpyAsync(h->state0.p, state, sizeof(int32_t) * ntraj, hipMemcpyHostToDevice, h->stream));
    }
    std::vector<int32_t> mid;
    h->model_id_given = model_id != nullptr;
    if (model_id) {
        mid.assign(model_id, model_id + ntraj);
        HIPCHK(h, h->model_id.ensure(sizeof(int32_t) * ntraj));
        HIPCHK(h, hipMemcpyAsync(h->model_id.p, model_id, sizeof(int32_t) * ntraj, hipMemcpyHostToDevice, h->stream));
    }
    std::vector<int32_t> ord = order_by_model(mid, ntraj);
    HIPCHK(h, h->order.ensure(sizeof(int32_t) * ntraj));
    HIPCHK(h, hipMemcpyAsync(h->order.p, ord.data(), sizeof(int32_t) * ntraj, hipMemcpyHostToDevice, h->stream));
    h->order_given = true;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->have_initial = true;
    h->sb_initial = false;
    h->have_results = false;
    return NQCD_OK;
}

int32_t nqcd_mi355x_load_mapping(nqcd_handle_t* h, int64_t ntraj, const double* X, const double* P) {
    if (!h) return fail(h, NQCD_ERR_ARG, "null handle");
    if (!h->have_problem || !map_is_problem(h)) return fail(h, NQCD_ERR_STATE, "set a spin-mapping problem first");
    if (!h->have_initial) return fail(h, NQCD_ERR_STATE, "load or sample the initial conditions first");
    HIPCHK(h, hipSetDevice(h->device));
    const int32_t rc = map_load_variables(h, ntraj, X, P);
    if (rc == NQCD_OK) h->have_results = false;
    return rc;
}

int32_t nqcd_mi355x_sample_fermi_dirac(nqcd_handle_t* h, double beta, const int32_t* available_states,
                                       int32_t n_available, int32_t basis, uint64_t seed) {
    if (!h) return fail(h, NQCD_ERR_ARG, "null argument");
    if (!h->have_problem) return fail(h, NQCD_ERR_STATE, "set_problem must be called first");
    if (!is_ie(h)) return fail(h, NQCD_ERR_UNSUPPORTED, "Fermi-Dirac occupations are for IESH / EhrenfestNA");
    if (available_states && n_available < 1) return fail(h, NQCD_ERR_ARG, "n_available must be positive");
    HIPCHK(h, hipSetDevice(h->device));
    const int32_t rc = ie_sample_fermi_dirac(h, beta, available_states, n_available, basis, seed);
    if (rc == NQCD_OK) h->have_results = false;
    return rc;
}

int32_t nqcd_mi355x_sample_initial(nqcd_handle_t* h, const nqcd_sampler_desc* S, uint64_t seed) {
    if (!h || !S) return fail(h, NQCD_ERR_ARG, "null argument");
    if (!h->have_problem) return fail(h, NQCD_ERR_STATE, "set_problem must be called first");
    if (S->ntraj < 1 || S->ntraj > (int64_t)1 << 31) return fail(h, NQCD_ERR_ARG, "ntraj out of range");
    const int n = h->prob.n_states;
    if (S->initial_state < 0 || S->initial_state >= n) return fail(h, NQCD_ERR_ARG, "initial_state out of range");
    const int64_t ntraj = S->ntraj;
    const int D = h->D;
    HIPCHK(h, hipSetDevice(h->device));
    if (is_ie(h)) {
        int32_t rc = ie_sample_initial(h, S, seed);
        if (rc == NQCD_OK) { h->have_initial = true; h->sb_initial = false; h->have_results = false; }
        return rc;
    }
    if (is_rp(h)) {
        int32_t rc = rp_sample_initial(h, S, seed);
        if (rc == NQCD_OK) { h->have_initial = true; h->sb_initial = false; h->have_results = false; }
        return rc;
    }
    if (S->kind == NQCD_SAMPLER_HARMONIC_WIGNER) {
        if (!is_spin_boson(h)) return fail(h, NQCD_ERR_UNSUPPORTED, "harmonic Wigner sampling needs the spin-boson model");
        int32_t rc = sb_sample_initial(h, S, seed);
        if (rc == NQCD_OK) { h->have_initial = true; h->sb_initial = true; h->have_results = false; }
        return rc;
    }
    if (cls_is_problem(h)) {
        int32_t rc = cls_sample_initial(h, S, seed);
        if (rc == NQCD_OK) { h->have_initial = true; h->sb_initial = false; h->have_results = false; }
        return rc;
    }
    if (cme_is_problem(h)) {
        int32_t rc = cme_sample_initial(h, S, seed);
        if (rc == NQCD_OK) { h->have_initial = true; h->sb_initial = false; h->have_results = false; }
        return rc;
    }
    if (map_is_problem(h)) {
        int32_t rc = map_sample_initial(h, S, seed);
        if (rc == NQCD_OK) { h->have_initial = true; h->sb_initial = false; h->have_results = false; }
        return rc;
    }
    if (ns_is_problem(h)) {
        int32_t rc = ns_sample_initial(h, S, seed);
        if (rc == NQCD_OK) { h->have_initial = true; h->sb_initial = false; h->have_results = false; }
        return rc;
    }
    if (!is_two_state_1d(h) && !nd_is_problem(h))
        return fail(h, NQCD_ERR_UNSUPPORTED, "sampler needs a two-state one-dof or diatom-surface problem");
    if (nd_is_problem(h) && (S->kind != NQCD_SAMPLER_GAUSSIAN || S->state_basis != 0))
        return fail(h, NQCD_ERR_UNSUPPORTED, "diatom-surface sampling: GAUSSIAN with an adiabatic initial state");
    h->ntraj = ntraj;
    HIPCHK(h, h->r0.ensure(sizeof(double) * D * ntraj));
    HIPCHK(h, h->v0.ensure(sizeof(double) * D * ntraj));
    HIPCHK(h, h->sig0.ensure(sizeof(double) * 4 * ntraj));
    HIPCHK(h, h->model_id.ensure(sizeof(int32_t) * ntraj));
    HIPCHK(h, h->order.ensure(sizeof(int32_t) * ntraj));
    HIPCHK(h, h->models_dev.ensure(sizeof(ModelSlot) * kMaxModels));
    HIPCHK(h, hipMemcpyAsync(h->models_dev.p, h->models, sizeof(ModelSlot) * kMaxModels, hipMemcpyHostToDevice, h->stream));
    std::vector<int32_t> ord(ntraj);
    if (S->kind == NQCD_SAMPLER_TULLY_GRID) {
        if (S->nparams != 3 || !S->params) return fail(h, NQCD_ERR_ARG, "TULLY_GRID params: x0, p_lo, n_p");
        const int nm = S->n_models;
        if (nm < 1 || nm > kMaxModels) return fail(h, NQCD_ERR_ARG, "n_models must be 1..4");
        for (int m = 0; m < nm; ++m)
            if (!h->model_set[m]) return fail(h, NQCD_ERR_STATE, "model slot not set");
        const int n_p = (int)S->params[2];
        if (n_p < 1) return fail(h, NQCD_ERR_ARG, "n_p must be >= 1");
        HIPCHK(h, launch_sample_tully_grid(ntraj, nm, S->params[0], S->params[1], n_p, h->masses[0], seed,
                                           h->r0.as<double>(), h->v0.as<double>(), h->model_id.as<int32_t>(), h->stream));
        // Queue order: slowest initial momentum first (longest trajectories), grouped by model.
        int64_t pos = 0;
        for (int pi = 0; pi < n_p; ++pi)
            for (int m = 0; m < nm; ++m)
                for (int64_t i = (int64_t)m + (int64_t)nm * pi; i < ntraj; i += (int64_t)nm * n_p) ord[pos++] = (int32_t)i;
        h->model_id_given = true;
    } else if (S->kind == NQCD_SAMPLER_GAUSSIAN) {
        if (S->nparams != 4 * D || !S->params) return fail(h, NQCD_ERR_ARG, "GAUSSIAN params: 4 per dof");
        if (!h->model_set[0]) return fail(h, NQCD_ERR_STATE, "model slot 0 not set");
        HIPCHK(h, h->small.ensure(sizeof(double) * 4 * D));
        HIPCHK(h, hipMemcpyAsync(h->small.p, S->params, sizeof(double) * 4 * D, hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, launch_sample_gaussian(ntraj, D, h->small.as<double>(), seed, h->r0.as<double>(), h->v0.as<double>(),
                                         h->stream));
        for (int64_t i = 0; i < ntraj; ++i) ord[i] = (int32_t)i;
        h->model_id_given = false;
    } else {
        return fail(h, NQCD_ERR_UNSUPPORTED, "unknown sampler kind");
    }
    HIPCHK(h, launch_init_density2(ntraj, S->initial_state, S->state_basis == 1, h->models_dev.as<ModelSlot>(),
                                   h->model_id_given ? h->model_id.as<int32_t>() : nullptr, h->r0.as<double>(),
                                   h->sig0.as<double>(), h->stream));
    HIPCHK(h, hipMemcpyAsync(h->order.p, ord.data(), sizeof(int32_t) * ntraj, hipMemcpyHostToDevice, h->stream));
    h->order_given = true;
    if (S->state_basis == 1) {
        h->state0_given = false;  // active state drawn from diag(sigma) (fssh.jl:54)
    } else {
        std::vector<int32_t> st(ntraj, S->initial_state);
        HIPCHK(h, h->state0.ensure(sizeof(int32_t) * ntraj));
        HIPCHK(h, hipMemcpyAsync(h->state0.p, st.data(), sizeof(int32_t) * ntraj, hipMemcpyHostToDevice, h->stream));
        h->state0_given = true;
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->have_initial = true;
    h->sb_initial = false;
    h->have_results = false;
    return NQCD_OK;
}

int32_t nqcd_mi355x_run(nqcd_handle_t* h, double t0, double t1, int32_t save_every, const nqcd_term_desc* term,
                        uint64_t key) {
    if (!h) return fail(h, NQCD_ERR_ARG, "null handle");
    if (!h->have_initial) return fail(h, NQCD_ERR_STATE, "no initial conditions loaded");
    if (!(t1 >= t0)) return fail(h, NQCD_ERR_ARG, "t1 must be >= t0");
    if (save_every < 0) return fail(h, NQCD_ERR_ARG, "save_every must be >= 0");
    if (term && term->kind != NQCD_TERM_NONE) {
        if (term->kind < NQCD_TERM_ABS_GT || term->kind > NQCD_TERM_LT)
            return fail(h, NQCD_ERR_UNSUPPORTED, "unknown termination predicate");
        if (term->dof < 0 || term->dof >= h->D) return fail(h, NQCD_ERR_ARG, "termination dof out of range");
    }
    const double nsteps_d = std::round((t1 - t0) / h->prob.dt);
    if (nsteps_d > 2.0e9) return fail(h, NQCD_ERR_ARG, "too many steps");
    const int64_t nsteps = (int64_t)nsteps_d;
    h->save_every = save_every;
    h->nsave = (save_every > 0) ? (int32_t)(nsteps / save_every + 1) : 1;
    HIPCHK(h, hipSetDevice(h->device));
    if (h->cell.mode != NQCD_CELL_NONE) {
        const bool one = h->cell.nd == 1 && h->D == 1 &&
                         (cls_is_problem(h) || cme_is_problem(h) || ns_is_problem(h) ||
                          (is_two_state_1d(h) && !is_rp(h) && !is_ie(h) && !is_spin_boson(h)));
        const bool diatom = nd_is_problem(h) && h->cell.nd == 3;
        if (!one && !diatom)
            return fail(h, NQCD_ERR_UNSUPPORTED, "cell callbacks: one-coordinate FSSH / Ehrenfest, CME / BCME and classical "
                                                 "problems, and the diatom-surface problem (3 x 3 cell)");
    }
    const auto wall0 = std::chrono::steady_clock::now();
    h->stats = nqcd_run_stats{};
    if (cls_is_problem(h)) {
        if (term && term->kind != NQCD_TERM_NONE && term->dof != 0)
            return fail(h, NQCD_ERR_ARG, "termination dof out of range");
        int32_t rc = cls_run(h, t0, nsteps, term);
        if (rc != NQCD_OK) return rc;
    } else if (is_rp(h)) {
        int32_t rc = rp_run(h, t0, nsteps, term, key);
        if (rc != NQCD_OK) return rc;
    } else if (is_ie(h)) {
        int32_t rc = ie_run(h, t0, nsteps, term, key);
        if (rc != NQCD_OK) return rc;
    } else if (nd_is_problem(h)) {
        int32_t rc = nd_run(h, t0, nsteps, term, key);
        if (rc != NQCD_OK) return rc;
    } else if (cme_is_problem(h)) {
        if (term && term->kind != NQCD_TERM_NONE && term->dof != 0)
            return fail(h, NQCD_ERR_ARG, "termination dof out of range");
        int32_t rc = cme_run(h, t0, nsteps, term, key);
        if (rc != NQCD_OK) return rc;
    } else if (map_is_problem(h)) {
        if (term && term->kind != NQCD_TERM_NONE)
            return fail(h, NQCD_ERR_UNSUPPORTED, "spin mapping runs to t1 (no termination predicate)");
        int32_t rc = map_run(h, t0, nsteps, key);
        if (rc != NQCD_OK) return rc;
    } else if (ns_is_problem(h)) {
        if (h->prob.schedule != NQCD_SCHED_PERSISTENT)
            return fail(h, NQCD_ERR_UNSUPPORTED, "n-state FSSH / Ehrenfest run on the persistent schedule");
        if (term && term->kind != NQCD_TERM_NONE && term->dof != 0)
            return fail(h, NQCD_ERR_ARG, "termination dof out of range");
        int32_t rc = ns_run(h, t0, nsteps, term, key);
        if (rc != NQCD_OK) return rc;
    } else if (is_spin_boson(h)) {
        if (!h->sb_initial) return fail(h, NQCD_ERR_STATE, "spin-boson runs need nqcd_mi355x_sample_initial");
        int32_t rc = sb_run(h, t0, nsteps, term, key);
        if (rc != NQCD_OK) return rc;
    } else if (is_two_state_1d(h)) {
        if (int32_t rc = alloc_results(h)) return rc;
        Fssh2Args A = make_args(h, t0, nsteps, term, key);
        if (h->prob.schedule == NQCD_SCHED_PERSISTENT) {
            const TuneShape sh = launch_shape(h, TUNE_FSSH2);
            const int block = sh.block;
            int bpc = 0;
            HIPCHK(h, fssh2_persistent_occupancy(h->prob.precision, block, &bpc));
            const int grid = (int)persistent_grid(bpc, sh, h->cu_count, h->ntraj);
            HIPCHK(h, hipEventRecord(h->ev0, h->stream));
            HIPCHK(h, h->args_dev.ensure(sizeof(Fssh2Args) * kSuspRounds));
            const int64_t cap = (int64_t)grid * block;
            SuspSet sets[2];
            HIPCHK(h, h->susp_ctr.ensure(sizeof(unsigned long long) * 4));
            for (int b = 0; b < 2; ++b) {
                HIPCHK(h, h->susp_f[b].ensure(sizeof(double) * kSuspRows * cap));
                HIPCHK(h, h->susp_tid[b].ensure(sizeof(int64_t) * cap));
                HIPCHK(h, h->susp_i[b].ensure(sizeof(int32_t) * 3 * cap));
                sets[b].f = h->susp_f[b].as<double>();
                sets[b].tid = h->susp_tid[b].as<int64_t>();
                sets[b].i = h->susp_i[b].as<int32_t>();
                sets[b].count = h->susp_ctr.as<unsigned long long>() + 2 * b;
                sets[b].claim = h->susp_ctr.as<unsigned long long>() + 2 * b + 1;
                sets[b].cap = cap;
            }
            A.prio_span = fssh2_prio_span(cap);
            HIPCHK(h, fssh2_launch_persistent(A, h->args_dev.as<Fssh2Args>(), h->prob.precision, grid, block, h->stream,
                                              sets));
            HIPCHK(h, hipEventRecord(h->ev1, h->stream));
            int rounds = 1, thr = 0;
            fssh2_suspend_policy(&rounds, &thr);
            h->stats.block = block; h->stats.grid = grid; h->stats.launches = 2 * rounds;
        } else {
            if (h->prob.precision != NQCD_PREC_FP64)
                return fail(h, NQCD_ERR_UNSUPPORTED, "the step-per-launch schedule runs in FP64 only");
            const int64_t n = h->ntraj;
            HIPCHK(h, h->plainS.ensure(sizeof(double) * kPlainRows * n));
            HIPCHK(h, h->plainSi.ensure(sizeof(int32_t) * kPlainIntRows * n));
            double* S = h->plainS.as<double>();
            int32_t* Si = h->plainSi.as<int32_t>();
            HIPCHK(h, hipEventRecord(h->ev0, h->stream));
            HIPCHK(h, fssh2_launch_plain_init(A, S, Si, h->stream));
            // Identical step launches captured once into a graph and replayed.
            const int64_t chunk = 64;
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            if (nsteps >= chunk) {
                HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
                for (int64_t j = 0; j < chunk; ++j) (void)fssh2_launch_plain_step(A, S, Si, h->stream);
                HIPCHK(h, hipStreamEndCapture(h->stream, &graph));
                HIPCHK(h, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            }
            int64_t done = 0, launches = 1;
            for (; done + chunk <= nsteps && exec; done += chunk) {
                HIPCHK(h, hipGraphLaunch(exec, h->stream));
                launches += chunk;
            }
            for (; done < nsteps; ++done, ++launches) HIPCHK(h, fssh2_launch_plain_step(A, S, Si, h->stream));
            HIPCHK(h, fssh2_launch_plain_finish(A, S, Si, h->stream));
            HIPCHK(h, hipEventRecord(h->ev1, h->stream));
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
            h->stats.block = 256; h->stats.grid = (int)((n + 255) / 256); h->stats.launches = launches + 1;
        }
    } else {
        return fail(h, NQCD_ERR_UNSUPPORTED, "unsupported problem for run");
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    float ms = 0.0f;
    HIPCHK(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->stats.kernel_ms = ms;
    // every run path writes the steps taken per trajectory; a path that did not is a library error, not
    // something to read
    if (!h->steps.p || h->steps.bytes < sizeof(int32_t) * (size_t)h->ntraj)
        return fail(h, NQCD_ERR_DEVICE, "internal: the run did not record the steps per trajectory");
    HIPCHK(h, h->sum64.ensure(sizeof(long long)));
    HIPCHK(h, launch_sum_i32(h->steps.as<int32_t>(), h->ntraj, h->sum64.as<long long>(), h->stream));
    long long tot = 0;
    HIPCHK(h, hipMemcpy(&tot, h->sum64.p, sizeof(long long), hipMemcpyDeviceToHost));
    h->stats.traj_steps = tot;
    h->stats.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
    h->have_results = true;
    return NQCD_OK;
}

}  // extern "C"

namespace {

// What a Mean/Sum reduction of an output sums: materialised rows (src, and for the total energy the
// potential-energy rows src2, summed separately and added after scaling), or rows computed per
// trajectory from the final state / the diabatic save slots (fill).
struct ReducePlan {
    int32_t rows = 0;
    const double* src = nullptr;
    const double* src2 = nullptr;
    bool fill = false;
    OutputSrc S{};
    double scale = 1.0;
};

int32_t reduce_plan(nqcd_handle_t* h, int32_t kind, int32_t red, const double* out, int64_t len, ReducePlan& P) {
    if (!h || !out) return fail(h, NQCD_ERR_ARG, "null argument");
    if (!h->have_results) return fail(h, NQCD_ERR_STATE, "no results: call nqcd_mi355x_run first");
    if (red != NQCD_RED_MEAN && red != NQCD_RED_SUM) return fail(h, NQCD_ERR_ARG, "unknown reduction");
    const int n = h->prob.n_states;
    const int64_t nt = h->ntraj;
    int32_t rows = 0;
    const double* src = nullptr;
    if (is_ie(h)) {
        if (kind == NQCD_OUT_ADIABATIC_POPULATION) { rows = n * h->nsave; src = h->save_ad.as<double>(); }
        else if (kind == NQCD_OUT_DIABATIC_POPULATION) { rows = n * h->nsave; src = h->save_di.as<double>(); }
        else return fail(h, NQCD_ERR_UNSUPPORTED, "IESH / EhrenfestNA outputs: ADIABATIC_POPULATION, DIABATIC_POPULATION");
    }
    if (map_is_problem(h)) {
        if (kind == NQCD_OUT_DIABATIC_POPULATION) { rows = n * h->nsave; src = h->save_di.as<double>(); }
        else if (kind == NQCD_OUT_KINETIC_ENERGY || kind == NQCD_OUT_TOTAL_ENERGY) { rows = h->nsave; src = h->save_ekin.as<double>(); }
        else if (kind == NQCD_OUT_CENTROID_POSITION && h->prob.algorithm == NQCD_ALG_RP_MINT) {
            rows = h->nsave; src = h->save_centroid.as<double>();
        }
        else if (kind == NQCD_OUT_POPULATION_CORRELATION) rows = n * n * h->nsave;
        else return fail(h, NQCD_ERR_UNSUPPORTED, "mapping outputs: DIABATIC_POPULATION, POPULATION_CORRELATION, "
                                                  "KINETIC_ENERGY, TOTAL_ENERGY (CENTROID_POSITION for ring polymers)");
        if (len != rows) return fail(h, NQCD_ERR_ARG, "len does not match the output size");
        P = ReducePlan{};
        P.rows = rows;
        P.src = src;
        if (kind == NQCD_OUT_POPULATION_CORRELATION) {   // normalisation F for eCMM (cmm.jl:114-119), 1 otherwise
            P.fill = true;
            P.S.kind = kind;
            P.S.nsave = h->nsave;
            P.S.nst = n;
            P.S.save_pop_di = h->save_di.as<double>();
            P.S.init_pop = h->final_di.as<double>();
            P.S.norm = h->prob.method == NQCD_METHOD_ECMM ? (double)n : 1.0;
        }
        P.src2 = kind == NQCD_OUT_TOTAL_ENERGY ? h->save_epot.as<double>() : nullptr;
        P.scale = (red == NQCD_RED_MEAN) ? 1.0 / (double)nt : 1.0;
        return NQCD_OK;
    }
    if (cme_is_problem(h) || cls_is_problem(h)) {
        const bool cls = cls_is_problem(h);
        if (kind == NQCD_OUT_KINETIC_ENERGY) { rows = h->nsave; src = h->save_ekin.as<double>(); }
        else if (kind == NQCD_OUT_TOTAL_ENERGY && h->prob.method != NQCD_METHOD_BCME) { rows = h->nsave; src = h->save_ekin.as<double>(); }
        else if (kind == NQCD_OUT_POSITION) { rows = h->nsave; src = h->save_centroid.as<double>(); }
        else if (kind == NQCD_OUT_VELOCITY && cls) { rows = h->nsave; src = h->save_v.as<double>(); }
        else return fail(h, NQCD_ERR_UNSUPPORTED, cls ? "classical outputs: KINETIC_ENERGY, TOTAL_ENERGY, POSITION, VELOCITY"
                                                      : "CME outputs: KINETIC_ENERGY, TOTAL_ENERGY (CME), POSITION");
        if (len != rows) return fail(h, NQCD_ERR_ARG, "len does not match the output size");
        P = ReducePlan{};
        P.rows = rows;
        P.src = src;
        P.src2 = kind == NQCD_OUT_TOTAL_ENERGY ? h->save_epot.as<double>() : nullptr;
        P.scale = (red == NQCD_RED_MEAN) ? 1.0 / (double)nt : 1.0;
        return NQCD_OK;
    }
    if (kind == NQCD_OUT_POSITION || kind == NQCD_OUT_VELOCITY)
        return fail(h, NQCD_ERR_UNSUPPORTED, "position / velocity outputs: CME / BCME and classical problems");
    if (ns_is_problem(h) && kind != NQCD_OUT_ADIABATIC_POPULATION && kind != NQCD_OUT_DIABATIC_POPULATION)
        return fail(h, NQCD_ERR_UNSUPPORTED, "n-state FSSH / Ehrenfest outputs: ADIABATIC_POPULATION, DIABATIC_POPULATION");
    const bool energy = kind == NQCD_OUT_KINETIC_ENERGY || kind == NQCD_OUT_TOTAL_ENERGY;
    if (energy && (is_rp(h) || nd_is_problem(h) || is_spin_boson(h) || !is_two_state_1d(h)))
        return fail(h, NQCD_ERR_UNSUPPORTED, "energy outputs exist for two-state one-dof FSSH / Ehrenfest");
    if (!src) switch (kind) {
    case NQCD_OUT_KINETIC_ENERGY:
    case NQCD_OUT_TOTAL_ENERGY: rows = h->nsave; src = h->save_ekin.as<double>(); break;
    case NQCD_OUT_ADIABATIC_POPULATION: rows = n * h->nsave; src = h->save_ad.as<double>(); break;
    case NQCD_OUT_DIABATIC_POPULATION: rows = n * h->nsave; src = h->save_di.as<double>(); break;
    case NQCD_OUT_FINAL_ADIABATIC_POPULATION: rows = n; break;
    case NQCD_OUT_SCATTERING_1D: rows = 2 * n; break;
    case NQCD_OUT_SCATTERING_1D_DIABATIC:
        if (!is_two_state_1d(h))
            return fail(h, NQCD_ERR_UNSUPPORTED, "diabatic final populations exist for two-state one-dof FSSH / Ehrenfest");
        rows = 2 * n; break;
    case NQCD_OUT_POPULATION_CORRELATION: rows = n * n * h->nsave; break;
    case NQCD_OUT_CENTROID_POSITION:
        if (!is_rp(h)) return fail(h, NQCD_ERR_UNSUPPORTED, "centroid output needs a ring-polymer problem");
        rows = h->nsave; src = h->save_centroid.as<double>(); break;
    default: return fail(h, NQCD_ERR_UNSUPPORTED, "unknown output kind");
    }
    if (len != rows) return fail(h, NQCD_ERR_ARG, "len does not match the output size");
    if (h->prob.method == NQCD_METHOD_RPMD && kind != NQCD_OUT_CENTROID_POSITION)
        return fail(h, NQCD_ERR_UNSUPPORTED, "RPMD has no electronic outputs");
    P = ReducePlan{};
    P.rows = rows;
    P.src = src;
    P.src2 = kind == NQCD_OUT_TOTAL_ENERGY ? h->save_epot.as<double>() : nullptr;
    if (!src) {
        P.fill = true;
        P.S.kind = kind;
        P.S.method = h->prob.method == NQCD_METHOD_FSSH ? 0 : 1;
        P.S.nsave = h->nsave;
        P.S.xdof = 0;
        P.S.r = h->r.as<double>();
        P.S.sig = h->sig.as<double>();
        P.S.state = h->state.as<int32_t>();
        P.S.save_pop_di = h->save_di.as<double>();
        P.S.final_pop_di = h->final_di.as<double>();
    }
    P.scale = (red == NQCD_RED_MEAN) ? 1.0 / (double)nt : 1.0;
    return NQCD_OK;
}

}  // namespace

extern "C" {

int32_t nqcd_mi355x_reduce(nqcd_handle_t* h, int32_t kind, int32_t red, double* out, int64_t len) {
    ReducePlan P;
    if (int32_t rc = reduce_plan(h, kind, red, out, len, P)) return rc;
    const int64_t nt = h->ntraj;
    const int32_t rows = P.rows;
    HIPCHK(h, hipSetDevice(h->device));
    const double* src = P.src;
    if (P.fill) {
        HIPCHK(h, h->tmp.ensure(sizeof(double) * rows * nt));
        HIPCHK(h, launch_fill_output2(P.S, nt, h->tmp.as<double>(), h->stream));
        src = h->tmp.as<double>();
    }
    HIPCHK(h, h->partial.ensure(sizeof(double) * rows * reduce_nchunk(nt)));
    HIPCHK(h, h->small.ensure(sizeof(double) * rows));
    HIPCHK(h, launch_reduce_rows(src, rows, nt, h->partial.as<double>(), P.scale, h->small.as<double>(), h->stream));
    HIPCHK(h, hipMemcpyAsync(out, h->small.p, sizeof(double) * rows, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (P.src2) {  // total energy: reduce the potential energy too; the reduction is linear
        std::vector<double> pot(rows);
        HIPCHK(h, launch_reduce_rows(P.src2, rows, nt, h->partial.as<double>(), P.scale, h->small.as<double>(), h->stream));
        HIPCHK(h, hipMemcpyAsync(pot.data(), h->small.p, sizeof(double) * rows, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        for (int32_t k = 0; k < rows; ++k) out[k] += pot[k];
    }
    return NQCD_OK;
}

int32_t nqcd_mi355x_reduce_noalloc(nqcd_handle_t* h, int32_t kind, int32_t red, double* out, int64_t len) {
    ReducePlan P;
    if (int32_t rc = reduce_plan(h, kind, red, out, len, P)) return rc;
    const int64_t nt = h->ntraj;
    const size_t need = reduce_noalloc_doubles(P.rows, nt, P.src2 != nullptr);
    if (need > kReduceScratchDoubles)
        return fail(h, NQCD_ERR_ARG, "reduction needs " + std::to_string(need) + " doubles of scratch (rows x chunks of " +
                                         "4096 trajectories + rows), more than the " + std::to_string(kReduceScratchDoubles) +
                                         " the allocation-free entry point holds; use nqcd_mi355x_reduce");
    HIPCHK(h, hipSetDevice(h->device));
    double* scr = reduce_scratch(h->device);
    if (!scr) return fail(h, NQCD_ERR_STATE, "scratch address of this device not resolved (open a handle on it first)");
    HIPCHK(h, launch_reduce_rows_noalloc(P.fill ? nullptr : P.src, P.src2, P.S, P.rows, nt, P.scale, scr, h->stream));
    HIPCHK(h, hipMemcpyAsync(out, scr + (need - (size_t)P.rows), sizeof(double) * P.rows, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return NQCD_OK;
}

int32_t nqcd_mi355x_download(nqcd_handle_t* h, int32_t field, void* out, int64_t bytes) {
    if (!h || !out) return fail(h, NQCD_ERR_ARG, "null argument");
    if (!h->have_results) return fail(h, NQCD_ERR_STATE, "no results: call nqcd_mi355x_run first");
    const int64_t nt = h->ntraj;
    const int D = (int)nqcd_rows(h), n = h->prob.n_states == 1 ? 2 : h->prob.n_states, ns = h->nsave;
    HIPCHK(h, hipSetDevice(h->device));
    if (is_ie(h)) return ie_download(h, field, out, bytes);
    if (ns_is_problem(h)) return ns_download(h, field, out, bytes);
    if (map_is_problem(h)) return map_download(h, field, out, bytes);
    if (cme_is_problem(h) && field != NQCD_FIELD_R && field != NQCD_FIELD_V && field != NQCD_FIELD_STATE &&
        field != NQCD_FIELD_STEPS && field != NQCD_FIELD_NHOPS && field != NQCD_FIELD_SAVE_R &&
        field != NQCD_FIELD_SAVE_EKIN && field != NQCD_FIELD_SAVE_EPOT && field != NQCD_FIELD_SAVE_STATE)
        return fail(h, NQCD_ERR_UNSUPPORTED, "CME / BCME fields: R, V, STATE, STEPS, NHOPS, SAVE_R, SAVE_EKIN, SAVE_EPOT, SAVE_STATE");
    if (cls_is_problem(h) && field != NQCD_FIELD_R && field != NQCD_FIELD_V && field != NQCD_FIELD_STEPS &&
        field != NQCD_FIELD_SAVE_R && field != NQCD_FIELD_SAVE_V && field != NQCD_FIELD_SAVE_EKIN &&
        field != NQCD_FIELD_SAVE_EPOT)
        return fail(h, NQCD_ERR_UNSUPPORTED, "classical fields: R, V, STEPS, SAVE_R, SAVE_V, SAVE_EKIN, SAVE_EPOT");
    auto fetch = [&](const DevBuf& b, size_t sz, std::vector<char>& host) -> hipError_t {
        host.resize(sz);
        return hipMemcpy(host.data(), b.p, sz, hipMemcpyDeviceToHost);
    };
    std::vector<char> raw;
    switch (field) {
    case NQCD_FIELD_R:
    case NQCD_FIELD_V: {
        if (bytes != (int64_t)sizeof(double) * D * nt) return fail(h, NQCD_ERR_ARG, "bytes mismatch");
        HIPCHK(h, fetch(field == NQCD_FIELD_R ? h->r : h->v, sizeof(double) * D * nt, raw));
        const double* s = reinterpret_cast<const double*>(raw.data());
        double* o = static_cast<double*>(out);
        for (int64_t i = 0; i < nt; ++i)
            for (int I = 0; I < D; ++I) o[(size_t)i * D + I] = s[(size_t)I * nt + i];
        return NQCD_OK;
    }
    case NQCD_FIELD_SIGMA_RE:
    case NQCD_FIELD_SIGMA_IM: {
        if (bytes != (int64_t)sizeof(double) * n * n * nt) return fail(h, NQCD_ERR_ARG, "bytes mismatch");
        HIPCHK(h, fetch(h->sig, sizeof(double) * 4 * nt, raw));
        const double* s = reinterpret_cast<const double*>(raw.data());
        double* o = static_cast<double*>(out);
        for (int64_t i = 0; i < nt; ++i) {
            double* m = o + (size_t)i * 4;
            if (field == NQCD_FIELD_SIGMA_RE) { m[0] = s[i]; m[3] = s[nt + i]; m[1] = m[2] = s[2 * nt + i]; }
            else { m[0] = m[3] = 0.0; m[2] = s[3 * nt + i]; m[1] = -s[3 * nt + i]; }
        }
        return NQCD_OK;
    }
    case NQCD_FIELD_STATE:
    case NQCD_FIELD_STEPS:
    case NQCD_FIELD_NHOPS: {
        if (bytes != (int64_t)sizeof(int32_t) * nt) return fail(h, NQCD_ERR_ARG, "bytes mismatch");
        const DevBuf& b = field == NQCD_FIELD_STATE ? h->state : (field == NQCD_FIELD_STEPS ? h->steps : h->nhops);
        HIPCHK(h, hipMemcpy(out, b.p, sizeof(int32_t) * nt, hipMemcpyDeviceToHost));
        return NQCD_OK;
    }
    case NQCD_FIELD_SAVE_POP_AD:
    case NQCD_FIELD_SAVE_POP_DI: {
        if (bytes != (int64_t)sizeof(double) * n * ns * nt) return fail(h, NQCD_ERR_ARG, "bytes mismatch");
        HIPCHK(h, fetch(field == NQCD_FIELD_SAVE_POP_AD ? h->save_ad : h->save_di, sizeof(double) * n * ns * nt, raw));
        const double* s = reinterpret_cast<const double*>(raw.data());
        double* o = static_cast<double*>(out);
        for (int64_t i = 0; i < nt; ++i)
            for (int sl = 0; sl < ns; ++sl)
                for (int c = 0; c < n; ++c) o[((size_t)i * ns + sl) * n + c] = s[((size_t)sl * n + c) * nt + i];
        return NQCD_OK;
    }
    case NQCD_FIELD_SAVE_CENTROID: {
        if (!is_rp(h)) return fail(h, NQCD_ERR_UNSUPPORTED, "centroids exist for ring-polymer problems only");
        if (bytes != (int64_t)sizeof(double) * ns * nt) return fail(h, NQCD_ERR_ARG, "bytes mismatch");
        HIPCHK(h, fetch(h->save_centroid, sizeof(double) * ns * nt, raw));
        const double* s = reinterpret_cast<const double*>(raw.data());
        double* o = static_cast<double*>(out);
        for (int64_t i = 0; i < nt; ++i)
            for (int sl = 0; sl < ns; ++sl) o[(size_t)i * ns + sl] = s[(size_t)sl * nt + i];
        return NQCD_OK;
    }
    case NQCD_FIELD_SAVE_V: {
        if (!cls_is_problem(h)) return fail(h, NQCD_ERR_UNSUPPORTED, "velocities per save exist for classical problems");
        if (bytes != (int64_t)sizeof(double) * ns * nt) return fail(h, NQCD_ERR_ARG, "bytes mismatch");
        HIPCHK(h, fetch(h->save_v, sizeof(double) * ns * nt, raw));
        const double* s = reinterpret_cast<const double*>(raw.data());
        double* o = static_cast<double*>(out);
        for (int64_t i = 0; i < nt; ++i)
            for (int sl = 0; sl < ns; ++sl) o[(size_t)i * ns + sl] = s[(size_t)sl * nt + i];
        return NQCD_OK;
    }
    case NQCD_FIELD_SAVE_R: {
        if (!nd_is_problem(h) && !cme_is_problem(h) && !cls_is_problem(h))
            return fail(h, NQCD_ERR_UNSUPPORTED, "configurations per save exist for the diatom-surface and CME models");
        if (bytes != (int64_t)sizeof(double) * D * ns * nt) return fail(h, NQCD_ERR_ARG, "bytes mismatch");
        HIPCHK(h, fetch(h->save_centroid, sizeof(double) * D * ns * nt, raw));
        const double* s = reinterpret_cast<const double*>(raw.data());
        double* o = static_cast<double*>(out);
        for (int64_t i = 0; i < nt; ++i)
            for (int sl = 0; sl < ns; ++sl)
                for (int I = 0; I < D; ++I) o[((size_t)i * ns + sl) * D + I] = s[((size_t)sl * D + I) * nt + i];
        return NQCD_OK;
    }
    case NQCD_FIELD_FALLBACKS: {
        if (!nd_is_problem(h) || h->models[0].kind != NQCD_MODEL_MLP_SURROGATE)
            return fail(h, NQCD_ERR_UNSUPPORTED, "fallback counts exist for surrogate runs only");
        if (bytes != (int64_t)sizeof(int32_t) * nt) return fail(h, NQCD_ERR_ARG, "bytes mismatch");
        HIPCHK(h, hipMemcpy(out, h->fallbacks.p, sizeof(int32_t) * nt, hipMemcpyDeviceToHost));
        return NQCD_OK;
    }
    case NQCD_FIELD_SAVE_EKIN:
    case NQCD_FIELD_SAVE_EPOT: {
        if (!cme_is_problem(h) && !cls_is_problem(h) &&
            (is_rp(h) || nd_is_problem(h) || is_spin_boson(h) || !is_two_state_1d(h)))
            return fail(h, NQCD_ERR_UNSUPPORTED, "energies per save exist for two-state one-dof FSSH / Ehrenfest");
        if (bytes != (int64_t)sizeof(double) * ns * nt) return fail(h, NQCD_ERR_ARG, "bytes mismatch");
        HIPCHK(h, fetch(field == NQCD_FIELD_SAVE_EKIN ? h->save_ekin : h->save_epot, sizeof(double) * ns * nt, raw));
        const double* s = reinterpret_cast<const double*>(raw.data());
        double* o = static_cast<double*>(out);
        for (int64_t i = 0; i < nt; ++i)
            for (int sl = 0; sl < ns; ++sl) o[(size_t)i * ns + sl] = s[(size_t)sl * nt + i];
        return NQCD_OK;
    }
    case NQCD_FIELD_FINAL_POP_DI: {
        if (!is_two_state_1d(h))
            return fail(h, NQCD_ERR_UNSUPPORTED, "diabatic final populations exist for two-state one-dof FSSH / Ehrenfest");
        if (bytes != (int64_t)sizeof(double) * 2 * nt) return fail(h, NQCD_ERR_ARG, "bytes mismatch");
        HIPCHK(h, fetch(h->final_di, sizeof(double) * 2 * nt, raw));
        const double* s = reinterpret_cast<const double*>(raw.data());
        double* o = static_cast<double*>(out);
        for (int64_t i = 0; i < nt; ++i) { o[2 * i] = s[i]; o[2 * i + 1] = s[nt + i]; }
        return NQCD_OK;
    }
    case NQCD_FIELD_SAVE_STATE: {
        if (bytes != (int64_t)sizeof(int32_t) * ns * nt) return fail(h, NQCD_ERR_ARG, "bytes mismatch");
        HIPCHK(h, fetch(h->save_state, sizeof(int32_t) * ns * nt, raw));
        const int32_t* s = reinterpret_cast<const int32_t*>(raw.data());
        int32_t* o = static_cast<int32_t*>(out);
        for (int64_t i = 0; i < nt; ++i)
            for (int sl = 0; sl < ns; ++sl) o[(size_t)i * ns + sl] = s[(size_t)sl * nt + i];
        return NQCD_OK;
    }
    default: return fail(h, NQCD_ERR_ARG, "unknown field");
    }
}

int32_t nqcd_mi355x_surrogate_train(nqcd_handle_t* h, const double* X, const double* Y, int64_t n,
                                    const nqcd_train_desc* desc, double* loss_history) {
    if (!h || !X || !Y || !desc) return fail(h, NQCD_ERR_ARG, "null argument");
    return mlp_train(h, X, Y, n, desc, loss_history);
}

int32_t nqcd_mi355x_surrogate_eval(nqcd_handle_t* h, int32_t member, const double* X, int64_t n, double* H, double* dH) {
    if (!h || !X || !H) return fail(h, NQCD_ERR_ARG, "null argument");
    return mlp_eval(h, member, X, n, H, dH);
}

int32_t nqcd_mi355x_surrogate_gradient(nqcd_handle_t* h, int32_t member, const double* X, const double* Y, int64_t n,
                                       const double* w, double* grad, double* loss) {
    if (!h || !X || !Y || !w || !grad || !loss) return fail(h, NQCD_ERR_ARG, "null argument");
    return mlp_gradient(h, member, X, Y, n, w, grad, loss);
}

int32_t nqcd_mi355x_surrogate_set_params(nqcd_handle_t* h, int32_t member, const float* params, const float* norm) {
    if (!h || !params) return fail(h, NQCD_ERR_ARG, "null argument");
    return mlp_set_params(h, member, params, norm);
}

int32_t nqcd_mi355x_surrogate_get_params(nqcd_handle_t* h, int32_t member, float* params, float* norm) {
    if (!h || !params) return fail(h, NQCD_ERR_ARG, "null argument");
    return mlp_get_params(h, member, params, norm);
}

int32_t nqcd_mi355x_test_determinant(nqcd_handle_t* h, const double* re, const double* im, int32_t n, int32_t nmat,
                                     double* out) {
    if (!h || !re || !im || !out) return fail(h, NQCD_ERR_ARG, "null argument");
    return ie_test_det(h, re, im, n, nmat, out);
}

int32_t nqcd_mi355x_last_run_stats(nqcd_handle_t* h, nqcd_run_stats* out) {
    if (!h || !out) return fail(h, NQCD_ERR_ARG, "null argument");
    if (!h->have_results) return fail(h, NQCD_ERR_STATE, "no run yet");
    *out = h->stats;
    return NQCD_OK;
}

int32_t nqcd_mi355x_autotune(nqcd_handle_t* h, double t0, double t1, const nqcd_term_desc* term, uint64_t philox_key,
                             nqcd_tune_result* out) {
    if (!h || !out) return fail(h, NQCD_ERR_ARG, "null argument");
    if (!h->have_problem || !h->have_initial) return fail(h, NQCD_ERR_STATE, "autotune needs a problem and initial conditions");
    int32_t family = 0;
    if (is_rp(h) || is_ie(h) || is_spin_boson(h)) family = 0;
    else if (nd_is_problem(h)) family = h->models[0].kind == NQCD_MODEL_DIATOM_SURFACE ? TUNE_ND : 0;
    else if (is_two_state_1d(h) && h->prob.schedule == NQCD_SCHED_PERSISTENT) family = TUNE_FSSH2;
    if (!family)
        return fail(h, NQCD_ERR_UNSUPPORTED,
                    "autotune covers the persistent two-state one-dof and diatom-surface (analytic model) kernels");
    HIPCHK(h, hipSetDevice(h->device));
    std::vector<TuneShape> cand;
    for (int block : {256, 128, 64}) {
        int occ = 0;
        HIPCHK(h, family == TUNE_FSSH2 ? fssh2_persistent_occupancy(h->prob.precision, block, &occ)
                                        : nd_persistent_occupancy(h->prob.precision, block, &occ));
        if (occ < 1) continue;
        for (int w : {occ, occ * 3 / 4, occ / 2, occ / 4}) {
            if (w < 1) continue;
            const TuneShape s{block, w == occ ? 0 : w};
            bool dup = false;
            for (const TuneShape& c : cand) dup = dup || (c.block == s.block && c.wpc == s.wpc);
            if (!dup) cand.push_back(s);
        }
    }
    if (cand.empty() || cand[0].block != 256 || cand[0].wpc != 0)
        return fail(h, NQCD_ERR_UNSUPPORTED, "the persistent kernel does not fit on a compute unit");
    double best = 0.0, def_ms = 0.0;
    TuneShape best_s = cand[0];
    for (size_t i = 0; i < cand.size(); ++i) {
        h->force_block = cand[i].block;
        h->force_wpc = cand[i].wpc;
        double ms = 0.0;
        for (int rep = 0; rep < 2; ++rep) {
            const int32_t rc = nqcd_mi355x_run(h, t0, t1, 0, term, philox_key);
            if (rc != NQCD_OK) {
                h->force_block = h->force_wpc = 0;
                return rc;
            }
            ms = rep == 0 ? h->stats.kernel_ms : std::min(ms, h->stats.kernel_ms);
        }
        if (i == 0) def_ms = best = ms;
        else if (ms < best) { best = ms; best_s = cand[i]; }
    }
    h->force_block = h->force_wpc = 0;
    if (best > 0.98 * def_ms) { best_s = cand[0]; best = def_ms; }
    tune_store(family, h->prob.n_states, h->prob.precision, h->cu_count, best_s);
    h->have_results = false;
    out->block = best_s.block;
    out->workgroups_per_cu = best_s.wpc;
    out->n_candidates = (int32_t)cand.size();
    out->family = family;
    out->best_ms = best;
    out->default_ms = def_ms;
    return NQCD_OK;
}

}  // extern "C"
