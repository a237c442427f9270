// cme.hip -- classical master equation surface hopping (cme.jl:26-182): CME (Dou, Nitzan, Subotnik,
// J. Chem. Phys. 142, 084110 (2015)) and BCME (broadened, Dou and Subotnik, J. Chem. Phys. 144, 024116
// (2016)) on a two-state diabatic model of one coordinate, one trajectory per lane, FP64.
//
// Each step integrates motion! (dr = v, dv = a(r, state), cme.jl:7-18) with one Tsit5 step of dt
// (select_algorithm, IntegrationAlgorithms.jl:68; the reference's solve adapts the step to abstol /
// reltol, here it is fixed), then the HoppingCallback (surface_hopping.jl:2-16): Gamma = 2 pi V12^2,
// f = fermi(V22 - V11, mu, beta), P = Gamma f dt from state 1 and Gamma (1 - f) dt from state 2
// (cme.jl:26-41), one uniform, hop when u < P (cme.jl:43-53), no velocity rescaling (cme.jl:55-57).
// CME force: -dV_ss/dx / m (cme.jl:96-105).  BCME force (cme.jl:137-160): -dV_ss + f dh - dh n - (dGamma /
// Gamma) n2 with n(h) = int A(e) f(e) de and n2(h) = int (e - h) A(e) f(e) de over the bandwidth around mu,
// A the Lorentzian of half width Gamma / 2 at h (cme.jl:167-182).  The reference integrates them with
// QuadGK; here by parts: n = [F f] - int F f' de, F(e) = atan((e - h)/(Gamma/2))/pi, and
// n2 = [G f] - int G f' de, G(e) = (Gamma / 4pi) ln((e - h)^2 + (Gamma/2)^2), where f' = -beta f (1 - f)
// vanishes beyond 40 kT of mu.  When the band holds that window the remaining integrals have a closed form
// through the complex digamma function (bcme_broadening); a narrower band takes 8-point Gauss-Legendre panels
// over the window (panel width min(2 kT, Gamma/2), with panel edges at h and h +- 4 Gamma/2).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "cme.h"
#include "ensemble.h"

namespace nqcd {

struct CmeArgs {
    int64_t ntraj, nsteps;
    int32_t method, save_every, nsave, term_kind;
    double term_value, dt, t0, mass, kT, mu, bandwidth;
    uint64_t key;
    ModelSlot model;
    const double* r0;
    const double* v0;
    const int32_t* state0;
    double* r;
    double* v;
    int32_t* state;
    int32_t* steps;
    int32_t* nhops;
    double* save_r;       // [slot][traj]
    double* save_ekin;
    double* save_epot;
    int32_t* save_state;
    CellDesc cell;        // cell callback (one coordinate); mode 0: none
};

// Two-state diabatic model: the one-mode Anderson-Holstein impurity (params k, Ed, c, V12: V11 = k x^2/2,
// V22 = V11 + Ed + c x, V12 constant) or the closed-form two-state models.
__device__ __forceinline__ void cme_model(const ModelRegs& m, double x, double& v11, double& v22, double& v12,
                                          double& d11, double& d22, double& d12) {
    if (m.kind == NQCD_MODEL_ANDERSON_HOLSTEIN_1) {
        const double k = m.p[0], Ed = m.p[1], c = m.p[2];
        v11 = 0.5 * k * x * x; v22 = v11 + Ed + c * x; v12 = m.p[3];
        d11 = k * x; d22 = k * x + c; d12 = 0.0;
        return;
    }
    model1d_eval<double>(m, x, v11, v22, v12, d11, d22, d12);
}

__device__ __forceinline__ double fermi_dev(double e, double mu, double beta) { return 1.0 / (1.0 + exp(beta * (e - mu))); }

// beta f (1 - f), without overflow
__device__ __forceinline__ double fermi_deriv(double e, double mu, double beta) {
    const double x = beta * (e - mu);
    const double q = exp(-fabs(x));
    return beta * q / ((1.0 + q) * (1.0 + q));
}

// Digamma of z = x + i y for x >= 1/2: psi(z) = psi(z + m) - sum_{k<m} 1/(z + k), with m the fewest shifts
// that give |z + m| >= 10, then Stirling's series through the B14 term (truncation below 5e-17 at |w| >= 10).
// The shift sum is accumulated as one fraction, so the shifts cost no division.
__device__ __forceinline__ void digamma_c(double x, double y, double& re, double& im) {
    constexpr double R = 10.0;
    const double y2 = y * y;
    const int m = (y2 >= R * R) ? 0 : (int)fmax(0.0, ceil(sqrt(R * R - y2) - x));
    double nr = 0.0, ni = 0.0, dr = 1.0, di = 0.0;
    for (int k = 0; k < m; ++k) {
        const double wr = x + k;
        const double tr = nr * wr - ni * y + dr, ti = nr * y + ni * wr + di;
        const double ur = dr * wr - di * y, ui = dr * y + di * wr;
        nr = tr; ni = ti; dr = ur; di = ui;
    }
    const double dd = 1.0 / (dr * dr + di * di);
    const double sr = (nr * dr + ni * di) * dd, si = (ni * dr - nr * di) * dd;
    const double wr = x + m;
    const double ww = wr * wr + y2;
    const double iw = 1.0 / ww;
    const double tr = wr * iw, ti = -y * iw;                 // t = 1 / w
    const double qr = tr * tr - ti * ti, qi = 2.0 * tr * ti;  // t^2
    // P = 1/12 - t2 (1/120 - t2 (1/252 - t2 (1/240 - t2 (1/132 - t2 (691/32760 - t2 / 12)))))
    double pr = 691.0 / 32760.0 - qr / 12.0, pi = -qi / 12.0;
    const double cs[5] = {1.0 / 132.0, 1.0 / 240.0, 1.0 / 252.0, 1.0 / 120.0, 1.0 / 12.0};
    for (int j = 0; j < 5; ++j) {
        const double ar = qr * pr - qi * pi, ai = qr * pi + qi * pr;
        pr = cs[j] - ar; pi = -ai;
    }
    const double er = qr * pr - qi * pi, ei = qr * pi + qi * pr;  // t^2 P
    re = 0.5 * log(ww) - 0.5 * tr - er - sr;
    im = atan2(y, wr) - 0.5 * ti - ei - si;
}

// n(h) = int_a^b A f and n2(h) = int_a^b (e - h) A f, a, b = mu -+ W/2, A the Lorentzian of half-width
// Gamma / 2 at h.  Integrated by parts: the boundary terms plus int (-f') G with G = atan((e - h) / (Gamma / 2))
// for n and (Gamma / 4 pi) log((e - h)^2 + Gamma^2 / 4) for n2.  -f' is the logistic density of scale kT about
// mu; when the band holds +-40 kT of it (the rest weighs below 1e-17), the remaining integral is taken over the
// whole line in closed form:  int (-f') log(Gamma / 2 + i (e - h)) de = log(2 pi kT) + psi(1/2 + (Gamma / 2 +
// i (mu - h)) / (2 pi kT)), whose real and imaginary parts give the two integrals.  A narrower band takes
// eight-point Gauss-Legendre panels.
__device__ void bcme_broadening(double h, double mu, double beta, double Gam, double W, double& n, double& n2) {
    const double hg = 0.5 * Gam;
    const double c4 = Gam / (4.0 * M_PI);
    const double a = mu - 0.5 * W, b = mu + 0.5 * W;
    if (0.5 * W * beta >= 40.0) {
        // f(a) = 1 / (1 + e^{-beta W/2}) rounds to 1 and f(b) < 4.3e-18, so the boundary terms at b (below
        // 2.2e-18 for n, and Gamma/4pi ln(...) f(b) for n2) are dropped and those at a taken with f(a) = 1
        n = -atan((a - h) / hg) / M_PI;
        n2 = -c4 * log((a - h) * (a - h) + hg * hg);
        double pr, pi;
        digamma_c(0.5 + beta * hg / (2.0 * M_PI), beta * (mu - h) / (2.0 * M_PI), pr, pi);
        n += pi / M_PI;
        n2 += 2.0 * c4 * (log(2.0 * M_PI / beta) + pr);
        return;
    }
    const double fa = fermi_dev(a, mu, beta), fb = fermi_dev(b, mu, beta);
    n = atan((b - h) / hg) / M_PI * fb - atan((a - h) / hg) / M_PI * fa;
    n2 = c4 * log((b - h) * (b - h) + hg * hg) * fb - c4 * log((a - h) * (a - h) + hg * hg) * fa;
    const double lo = fmax(a, mu - 40.0 / beta), hi = fmin(b, mu + 40.0 / beta);
    if (!(hi > lo)) return;
    // 8-point Gauss-Legendre
    const double gx[4] = {0.1834346424956498, 0.5255324099163290, 0.7966664774136267, 0.9602898564975363};
    const double gw[4] = {0.3626837833783620, 0.3137066458778873, 0.2223810344533745, 0.1012285362903763};
    const double pw = fmin(2.0 / beta, hg);
    const double pts[5] = {lo, fmin(fmax(h - 4.0 * hg, lo), hi), fmin(fmax(h, lo), hi), fmin(fmax(h + 4.0 * hg, lo), hi), hi};
    double sF = 0.0, sG = 0.0;
    for (int seg = 0; seg < 4; ++seg) {
        const double s0 = pts[seg], s1 = pts[seg + 1];
        if (!(s1 > s0)) continue;
        const int np = (int)fmin(256.0, ceil((s1 - s0) / pw));
        const double w = (s1 - s0) / np;
        for (int p = 0; p < np; ++p) {
            const double c = s0 + (p + 0.5) * w, hw = 0.5 * w;
            for (int j = 0; j < 8; ++j) {
                const double e = c + ((j & 1) ? hw : -hw) * gx[j >> 1];
                const double wt = hw * gw[j >> 1] * fermi_deriv(e, mu, beta);
                sF += wt * atan((e - h) / hg);
                sG += wt * log((e - h) * (e - h) + hg * hg);
            }
        }
    }
    n += sF / M_PI;
    n2 += c4 * sG;
}

// BCME: the broadened force (A.method 1); a kernel of its own, so the CME kernel carries none of its code
template <bool BCME>
__device__ __forceinline__ double cme_accel(const CmeArgs& A, const ModelRegs& m, double x, int s) {
    double v11, v22, v12, d11, d22, d12;
    cme_model(m, x, v11, v22, v12, d11, d22, d12);
    double f = -((s == 0) ? d11 : d22);
    if constexpr (BCME) {
        const double Gam = 2.0 * M_PI * v12 * v12;
        const double beta = 1.0 / A.kT;
        const double h = v22 - v11;
        double n, n2;
        bcme_broadening(h, A.mu, beta, Gam, A.bandwidth, n, n2);
        const double fe = fermi_dev(h, A.mu, beta);
        const double dh = d22 - d11;
        const double dGam = 4.0 * M_PI * v12 * d12;
        f += fe * dh - dh * n - dGam / Gam * n2;
    }
    return f / A.mass;
}

template <bool BCME>
__global__ __launch_bounds__(256) void cme_kernel(const CmeArgs A) {
    const ModelRegs m = model_regs(A.model);
    for (int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; tid < A.ntraj;
         tid += (int64_t)gridDim.x * blockDim.x) {
        double x = A.r0[tid], v = A.v0[tid];
        int s = A.state0[tid];
        int32_t nh = 0;
        const int64_t nt = A.ntraj;
        auto save = [&](int slot) {
            if (slot >= A.nsave) return;
            double v11, v22, v12, d11, d22, d12;
            cme_model(m, x, v11, v22, v12, d11, d22, d12);
            A.save_r[(int64_t)slot * nt + tid] = x;
            A.save_ekin[(int64_t)slot * nt + tid] = 0.5 * A.mass * v * v;
            A.save_epot[(int64_t)slot * nt + tid] = (s == 0) ? v11 : v22;   // classical_potential_energy (cme.jl:107-113)
            A.save_state[(int64_t)slot * nt + tid] = s;
        };
        save(0);
        int64_t step = 0;
        while (step < A.nsteps) {
            using namespace ts5;
            const double h = A.dt;
            const double k1x = v, k1v = cme_accel<BCME>(A, m, x, s);
            double tx = x + h * (a21 * k1x), tv = v + h * (a21 * k1v);
            const double k2x = tv, k2v = cme_accel<BCME>(A, m, tx, s);
            tx = x + h * (a31 * k1x + a32 * k2x); tv = v + h * (a31 * k1v + a32 * k2v);
            const double k3x = tv, k3v = cme_accel<BCME>(A, m, tx, s);
            tx = x + h * (a41 * k1x + a42 * k2x + a43 * k3x); tv = v + h * (a41 * k1v + a42 * k2v + a43 * k3v);
            const double k4x = tv, k4v = cme_accel<BCME>(A, m, tx, s);
            tx = x + h * (a51 * k1x + a52 * k2x + a53 * k3x + a54 * k4x);
            tv = v + h * (a51 * k1v + a52 * k2v + a53 * k3v + a54 * k4v);
            const double k5x = tv, k5v = cme_accel<BCME>(A, m, tx, s);
            tx = x + h * (a61 * k1x + a62 * k2x + a63 * k3x + a64 * k4x + a65 * k5x);
            tv = v + h * (a61 * k1v + a62 * k2v + a63 * k3v + a64 * k4v + a65 * k5v);
            const double k6x = tv, k6v = cme_accel<BCME>(A, m, tx, s);
            x = x + h * (a71 * k1x + a72 * k2x + a73 * k3x + a74 * k4x + a75 * k5x + a76 * k6x);
            v = v + h * (a71 * k1v + a72 * k2v + a73 * k3v + a74 * k4v + a75 * k5v + a76 * k6v);
            // the save precedes the hop callback (a save time holds the state before the step's callbacks)
            if (A.save_every > 0 && (step + 1) % A.save_every == 0) save((int)((step + 1) / A.save_every));
            {   // HoppingCallback with the CME probability
                double v11, v22, v12, d11, d22, d12;
                cme_model(m, x, v11, v22, v12, d11, d22, d12);
                const double Gam = 2.0 * M_PI * v12 * v12;
                const double fe = fermi_dev(v22 - v11, A.mu, 1.0 / A.kT);
                const double P = (s == 0) ? Gam * fe * A.dt : Gam * (1.0 - fe) * A.dt;
                const double u = philox_uniform((uint64_t)tid, (uint32_t)step, STREAM_HOP, A.key);
                if (u < P) { s = 1 - s; ++nh; }
            }
            double xx[1] = {x};
            const bool cell_end = cell_callback<1, 1, double>(A.cell, xx);   // after the hop callback
            x = xx[0];
            ++step;
            if (cell_end) break;
            if (term_hit(A.term_kind, x, A.term_value)) break;   // after each step, not at t0
        }
        A.r[tid] = x;
        A.v[tid] = v;
        A.state[tid] = s;
        A.steps[tid] = (int32_t)step;
        A.nhops[tid] = nh;
    }
}

// Test hook: the broadening integrals n(h), n2(h) for each h of a list.
__global__ void bcme_broadening_kernel(const double* h, int64_t count, double mu, double beta, double Gam, double W,
                                       double* n, double* n2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    bcme_broadening(h[i], mu, beta, Gam, W, n[i], n2[i]);
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
bool cme_is_problem(const nqcd_handle_t* h) {
    return h->prob.method == NQCD_METHOD_CME || h->prob.method == NQCD_METHOD_BCME;
}

int32_t cme_check_problem(nqcd_handle_t* h) {
    const nqcd_problem_desc& d = h->prob;
    if (d.algorithm != NQCD_ALG_TSIT5) return fail(h, NQCD_ERR_UNSUPPORTED, "CME / BCME integrate with Tsit5");
    if (d.n_states != 2 || d.ndofs * d.natoms != 1 || d.nbeads != 1)
        return fail(h, NQCD_ERR_UNSUPPORTED, "CME / BCME: two states, one coordinate");
    if (!(d.temperature > 0.0)) return fail(h, NQCD_ERR_ARG, "CME / BCME need temperature > 0");
    if (d.method == NQCD_METHOD_BCME && !(d.bandwidth > 0.0)) return fail(h, NQCD_ERR_ARG, "BCME needs bandwidth > 0");
    return NQCD_OK;
}

int32_t cme_load_initial(nqcd_handle_t* h, int64_t ntraj, const double* r, const double* v, const int32_t* state) {
    if (!state) return fail(h, NQCD_ERR_ARG, "CME / BCME start from a diabatic state (state array)");
    for (int64_t i = 0; i < ntraj; ++i)
        if (state[i] < 0 || state[i] > 1) return fail(h, NQCD_ERR_ARG, "CME / BCME states are 0 or 1");
    h->ntraj = ntraj;
    HIPCHK(h, h->r0.ensure(sizeof(double) * ntraj));
    HIPCHK(h, h->v0.ensure(sizeof(double) * ntraj));
    HIPCHK(h, h->state0.ensure(sizeof(int32_t) * ntraj));
    HIPCHK(h, hipMemcpy(h->r0.p, r, sizeof(double) * ntraj, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->v0.p, v, sizeof(double) * ntraj, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->state0.p, state, sizeof(int32_t) * ntraj, hipMemcpyHostToDevice));
    h->state0_given = true;
    return NQCD_OK;
}

int32_t cme_sample_initial(nqcd_handle_t* h, const nqcd_sampler_desc* S, uint64_t seed) {
    if (S->kind != NQCD_SAMPLER_GAUSSIAN || S->nparams != 4 || !S->params)
        return fail(h, NQCD_ERR_UNSUPPORTED, "CME / BCME sample with GAUSSIAN (params mu_r, sd_r, mu_v, sd_v)");
    if (S->initial_state > 1) return fail(h, NQCD_ERR_ARG, "initial_state must be 0 or 1");
    const int64_t n = S->ntraj;
    h->ntraj = n;
    HIPCHK(h, h->r0.ensure(sizeof(double) * n));
    HIPCHK(h, h->v0.ensure(sizeof(double) * n));
    HIPCHK(h, h->small.ensure(sizeof(double) * 4));
    HIPCHK(h, hipMemcpy(h->small.p, S->params, sizeof(double) * 4, hipMemcpyHostToDevice));
    HIPCHK(h, launch_sample_gaussian(n, 1, h->small.as<double>(), seed, h->r0.as<double>(), h->v0.as<double>(), h->stream));
    std::vector<int32_t> st(n, S->initial_state);
    HIPCHK(h, h->state0.ensure(sizeof(int32_t) * n));
    HIPCHK(h, hipMemcpyAsync(h->state0.p, st.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->state0_given = true;
    return NQCD_OK;
}

int32_t cme_run(nqcd_handle_t* h, double t0, int64_t nsteps, const nqcd_term_desc* term, uint64_t key) {
    if (!h->model_set[0]) return fail(h, NQCD_ERR_STATE, "model slot 0 is not set");
    const int k = h->models[0].kind;
    if (k != NQCD_MODEL_ANDERSON_HOLSTEIN_1 && k != NQCD_MODEL_TULLY1 && k != NQCD_MODEL_TULLY2 && k != NQCD_MODEL_TULLY3 &&
        k != NQCD_MODEL_DW_GAUSSIAN)
        return fail(h, NQCD_ERR_UNSUPPORTED, "CME / BCME models: ANDERSON_HOLSTEIN_1, TULLY1/2/3, DW_GAUSSIAN");
    if (h->prob.precision != NQCD_PREC_FP64) return fail(h, NQCD_ERR_UNSUPPORTED, "CME / BCME run in FP64");
    const int64_t n = h->ntraj;
    const int ns = h->nsave;
    HIPCHK(h, h->r.ensure(sizeof(double) * n));
    HIPCHK(h, h->v.ensure(sizeof(double) * n));
    HIPCHK(h, h->state.ensure(sizeof(int32_t) * n));
    HIPCHK(h, h->steps.ensure(sizeof(int32_t) * n));
    HIPCHK(h, h->nhops.ensure(sizeof(int32_t) * n));
    HIPCHK(h, h->save_centroid.ensure(sizeof(double) * ns * n));
    HIPCHK(h, h->save_ekin.ensure(sizeof(double) * ns * n));
    HIPCHK(h, h->save_epot.ensure(sizeof(double) * ns * n));
    HIPCHK(h, h->save_state.ensure(sizeof(int32_t) * ns * n));
    HIPCHK(h, launch_zero(h->save_centroid.p, sizeof(double) * ns * n, h->stream));
    HIPCHK(h, launch_zero(h->save_ekin.p, sizeof(double) * ns * n, h->stream));
    HIPCHK(h, launch_zero(h->save_epot.p, sizeof(double) * ns * n, h->stream));
    HIPCHK(h, launch_zero(h->save_state.p, sizeof(int32_t) * ns * n, h->stream));
    CmeArgs A{};
    A.ntraj = n; A.nsteps = nsteps;
    A.method = h->prob.method == NQCD_METHOD_CME ? 0 : 1;
    A.save_every = h->save_every; A.nsave = ns;
    A.term_kind = term ? term->kind : 0; A.term_value = term ? term->value : 0.0;
    A.cell = h->cell;
    A.dt = h->prob.dt; A.t0 = t0; A.mass = h->masses[0];
    A.kT = h->prob.temperature; A.mu = h->prob.fermi_level; A.bandwidth = h->prob.bandwidth;
    A.key = key;
    A.model = h->models[0];
    A.r0 = h->r0.as<double>(); A.v0 = h->v0.as<double>(); A.state0 = h->state0.as<int32_t>();
    A.r = h->r.as<double>(); A.v = h->v.as<double>(); A.state = h->state.as<int32_t>();
    A.steps = h->steps.as<int32_t>(); A.nhops = h->nhops.as<int32_t>();
    A.save_r = h->save_centroid.as<double>(); A.save_ekin = h->save_ekin.as<double>();
    A.save_epot = h->save_epot.as<double>(); A.save_state = h->save_state.as<int32_t>();
    const int block = 128;
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((n + block - 1) / block, (int64_t)h->cu_count * 16));
    HIPCHK(h, hipEventRecord(h->ev0, h->stream));
    if (A.method == 1) hipLaunchKernelGGL(cme_kernel<true>, dim3((unsigned)grid), dim3(block), 0, h->stream, A);
    else hipLaunchKernelGGL(cme_kernel<false>, dim3((unsigned)grid), dim3(block), 0, h->stream, A);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipEventRecord(h->ev1, h->stream));
    h->stats.block = block;
    h->stats.grid = (int)grid;
    h->stats.launches = 1;
    return NQCD_OK;
}

int32_t cme_test_broadening(nqcd_handle_t* h, const double* hs, int64_t count, double mu, double kT, double Gam, double W,
                            double* n, double* n2) {
    DevBuf b;
    HIPCHK(h, b.ensure(sizeof(double) * 3 * count));
    double* d = b.as<double>();
    HIPCHK(h, hipMemcpy(d, hs, sizeof(double) * count, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(bcme_broadening_kernel, dim3((unsigned)((count + 63) / 64)), dim3(64), 0, h->stream, d, count, mu,
                       1.0 / kT, Gam, W, d + count, d + 2 * count);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipStreamSynchronize(h->stream));
    HIPCHK(h, hipMemcpy(n, d + count, sizeof(double) * count, hipMemcpyDeviceToHost));
    HIPCHK(h, hipMemcpy(n2, d + 2 * count, sizeof(double) * count, hipMemcpyDeviceToHost));
    b.release();
    return NQCD_OK;
}

}  // namespace nqcd
