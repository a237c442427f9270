// spin_boson.hip -- FSSH / Ehrenfest for the spin-boson model (two states coupled to a
// harmonic bath), BABwithTsit5 (bab_electronics.jl:48-91) with the hop callback
// (surface_hopping.jl:2-16).
//
// V(R) = [[eps + c.R, Delta], [Delta, -eps - c.R]] + R^T W^2 R / 2
// (docs/src/NQCModels/systembathmodels.md:21-27).  The state-dependent part of dV/dR_j
// is c_j sigma_z, so every per-mode quantity the reference forms per degree of freedom
// (U^T dV_j U, the coupling d_j, v.d, the rescaling vectors) is c_j times a 2x2 quantity
// of the collective coordinate X = sum_j c_j R_j.  One trajectory per lane; the bath
// positions and velocities live in registers; all trajectories run the same number of
// steps, so the launch is a plain grid (no refill queue).
#include "spin_boson.h"

#include <cmath>

#include "ensemble.h"
#include <vector>

namespace nqcd {

constexpr int kBathModes = 100;  // instantiated bath size

struct SbArgs {
    int64_t ntraj;
    int32_t method, rescaling, nsub, save_every;
    int32_t nsave, initial_state, diabatic, pad;
    double dt, t0, eps, Delta, S2, cnorm;  // S2 = sum c^2/m, cnorm = |c|
    int64_t nsteps;
    uint64_t key;
    const double* w2;    // [NB] omega^2
    const double* c;     // [NB] coupling
    const double* invm;  // [NB] 1/m
    const double* kd;    // [3][NB] half-kick coefficients dt/2 w^2/m, dt/2 c/m, and c (FP64)
    const float* kf;     // [3][NB] the same in FP32
    const double* r0;    // [NB][ntraj] sampled initial positions
    const double* v0;    // [NB][ntraj] sampled initial velocities
    double* r;           // [NB][ntraj]
    double* v;           // [NB][ntraj]
    double* sig;         // [4][ntraj]
    int32_t* state;
    int32_t* steps;
    int32_t* nhops;
    double* save_pop_ad; // [nsave][2][ntraj]
    double* save_pop_di; // [nsave][2][ntraj]
    int32_t* save_state; // [nsave][ntraj]
};

template <class T>
__device__ __forceinline__ const T* sb_opaque(const T* p) {
    asm volatile("" : "+s"(p));
    return p;
}

// U^T sigma_z U for the 2x2 eigenvectors U (row-major).
__device__ __forceinline__ void rotate_sz(const double U[4], double& g00, double& g11, double& g01) {
    g00 = U[0] * U[0] - U[2] * U[2];
    g11 = U[1] * U[1] - U[3] * U[3];
    g01 = U[0] * U[1] - U[2] * U[3];
}

__device__ __forceinline__ void sb_save(const SbArgs* Ap, int64_t tid, int slot, int method, int st, double tr,
                                        const Sig2& s, const double U[4]) {
    const SbArgs& A = *sb_opaque(Ap);
    if (slot >= A.nsave) return;
    const int64_t n = A.ntraj;
    double pa0, pa1;
    if (method == 0) { pa0 = (st == 0) ? 1.0 : 0.0; pa1 = 1.0 - pa0; }
    else { pa0 = 0.5 * (tr + s.q); pa1 = 0.5 * (tr - s.q); }
    const double pd0 = U[0] * U[0] * pa0 + U[1] * U[1] * pa1 + 2.0 * U[0] * U[1] * s.x;
    const double pd1 = U[2] * U[2] * pa0 + U[3] * U[3] * pa1 + 2.0 * U[2] * U[3] * s.x;
    A.save_pop_ad[((int64_t)slot * 2) * n + tid] = pa0;
    A.save_pop_ad[((int64_t)slot * 2 + 1) * n + tid] = pa1;
    A.save_pop_di[((int64_t)slot * 2) * n + tid] = pd0;
    A.save_pop_di[((int64_t)slot * 2 + 1) * n + tid] = pd1;
    A.save_state[(int64_t)slot * n + tid] = st;
}

__device__ __forceinline__ const double* kick_table(const SbArgs& A, double) { return A.kd; }
__device__ __forceinline__ const float* kick_table(const SbArgs& A, float) { return A.kf; }

template <int NB, typename R>
__device__ __forceinline__ void sb_body(const SbArgs* __restrict__ Ap) {
    const SbArgs& A = *Ap;
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= A.ntraj) return;
    const int method = A.method;
    R x[NB], v[NB];
    {
        const SbArgs& B = *sb_opaque(Ap);
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            x[j] = (R)B.r0[(int64_t)j * B.ntraj + tid];
            v[j] = (R)B.v0[(int64_t)j * B.ntraj + tid];
        }
    }
    // collective coordinate and the initial cache
    R X = (R)0;
    {
        const R* K = kick_table(*sb_opaque(Ap), (R)0);
#pragma unroll
        for (int j = 0; j < NB; ++j) X += K[2 * NB + j] * x[j];
    }
    double E0, E1, U[4];
    eigen2(A.eps + (double)X, A.Delta, -A.eps - (double)X, E0, E1, U, false);
    // initial density: pure adiabatic state, or U^T rho U for a diabatic one (density_matrix_dynamics.jl:37-75)
    double a, b, sx;
    if (A.diabatic) {
        const double u0 = U[A.initial_state * 2], u1 = U[A.initial_state * 2 + 1];
        a = u0 * u0; b = u1 * u1; sx = u0 * u1;
    } else {
        a = (A.initial_state == 0) ? 1.0 : 0.0; b = 1.0 - a; sx = 0.0;
    }
    const double tr = a + b;
    Sig2 s; s.q = a - b; s.x = sx; s.y = 0.0;
    int st = A.initial_state;
    if (method == 0 && A.diabatic) {  // fssh.jl:54
        const double u = philox_uniform((uint64_t)tid, 0u, STREAM_STATE, A.key) * tr;
        st = (u < a) ? 0 : 1;
    }
    double g00, g11, g01;
    rotate_sz(U, g00, g11, g01);
    // system force factor F: k_j = -(w2_j x_j + c_j F) / m_j
    double F = (method == 0) ? ((st == 0) ? g00 : g11) : (g00 * a + g11 * b + 2.0 * g01 * sx);
    double dEc = 0.0, wc = 0.0, tc = 0.0, t = A.t0;
    int nh = 0;
    sb_save(Ap, tid, 0, method, st, tr, s, U);
    const R dtR = (R)A.dt;
    for (int step = 0; step < A.nsteps; ++step) {
        // per-mode coefficients re-read every step (scalar loads) instead of being held in registers
        const R* K = kick_table(*sb_opaque(Ap), (R)0);
        const R Fr = (R)F;
        R Xn = (R)0;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            v[j] = v[j] - (K[j] * x[j] + K[NB + j] * Fr);   // step_B!: v += dt/2 * k
            x[j] = x[j] + dtR * v[j];                        // step_A!
            Xn += K[2 * NB + j] * x[j];
        }
        eigen2(A.eps + (double)Xn, A.Delta, -A.eps - (double)Xn, E0, E1, U, true);  // update_cache!
        rotate_sz(U, g00, g11, g01);
        {
            const double aa = 0.5 * (tr + s.q), bb = 0.5 * (tr - s.q);
            F = (method == 0) ? ((st == 0) ? g00 : g11) : (g00 * aa + g11 * bb + 2.0 * g01 * s.x);
        }
        const R* K2 = kick_table(*sb_opaque(Ap), (R)0);
        const R Fn = (R)F;
        R Vc = (R)0;
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            v[j] = v[j] - (K2[j] * x[j] + K2[NB + j] * Fn);  // step_B!
            Vc += K2[2 * NB + j] * v[j];
        }
        const double gap = E1 - E0;
        const double d01c = g01 / gap;               // d_j[0][1] = c_j * d01c
        const double w = d01c * (double)Vc;
        const double tn = t + A.dt;
        const double dEn = E0 - E1;
        const double inv = 1.0 / (tn - tc);
        Interp2 P;
        P.dE0 = dEc; P.dE1 = (dEn - dEc) * inv; P.w0 = wc; P.w1 = (w - wc) * inv;
        propagate2(P, s, t, tc, A.dt, A.nsub);
        dEc = dEn; wc = w; tc = tn; t = tn;
        // the save precedes the hop callback (a save time holds the state before the step's callbacks)
        if (A.save_every > 0 && ((step + 1) % A.save_every) == 0) sb_save(Ap, tid, (step + 1) / A.save_every, method, st, tr, s, U);
        if (method == 0) {
            const double sss = 0.5 * (tr + ((st == 0) ? s.q : -s.q));
            const double ratio = s.x / sss;
            const double dsm = (st == 0) ? d01c : -d01c;
            double p = 2.0 * (double)Vc * ratio * dsm * A.dt;
            if (p > 0.0) {
                p = fmin(p, 1.0);
                const double rnd = philox_uniform((uint64_t)tid, (uint32_t)step, STREAM_HOP, A.key);
                if (p > rnd) {
                    // rescale along d_j = c_j * dms (coupling[new, old]); surface_hopping.jl:64-99
                    const double dms = -dsm;
                    const double dE = (st == 0) ? gap : -gap;
                    bool accept = true;
                    if (A.rescaling != 2) {
                        const double aq = 0.5 * dms * dms * A.S2;
                        const double bq = dms * (double)Vc;
                        const double disc = bq * bq - 4.0 * aq * dE;
                        if (disc < 0.0) {
                            accept = false;
                            if (A.rescaling == 1) {
                                const SbArgs& C = *sb_opaque(Ap);
                                const double nrm = fabs(dms) * A.cnorm;
                                const double gam = bq / nrm;
#pragma unroll
                                for (int j = 0; j < NB; ++j)
                                    v[j] = (R)((double)v[j] - 2.0 * gam * (dms * C.c[j] / nrm));
                            }
                        } else {
                            const SbArgs& C = *sb_opaque(Ap);
                            const double root = sqrt(disc);
                            const double gam = (bq < 0.0) ? (bq + root) / (2.0 * aq) : (bq - root) / (2.0 * aq);
#pragma unroll
                            for (int j = 0; j < NB; ++j) v[j] = (R)((double)v[j] - gam * dms * C.c[j] * C.invm[j]);
                        }
                    }
                    if (accept) { st = 1 - st; ++nh; }
                }
            }
        }
    }
    const SbArgs& B = *sb_opaque(Ap);
    const int64_t n = B.ntraj;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        B.r[(int64_t)j * n + tid] = (double)x[j];
        B.v[(int64_t)j * n + tid] = (double)v[j];
    }
    B.sig[tid] = 0.5 * (tr + s.q);
    B.sig[n + tid] = 0.5 * (tr - s.q);
    B.sig[2 * n + tid] = s.x;
    B.sig[3 * n + tid] = s.y;
    B.state[tid] = st;
    B.steps[tid] = (int32_t)A.nsteps;
    B.nhops[tid] = nh;
}

__global__ __launch_bounds__(64) void sb_kernel_f32(const SbArgs* __restrict__ Ap) { sb_body<kBathModes, float>(Ap); }
__global__ __launch_bounds__(64) void sb_kernel_f64(const SbArgs* __restrict__ Ap) { sb_body<kBathModes, double>(Ap); }

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
int32_t sb_sample_initial(nqcd_handle_t* h, const nqcd_sampler_desc* S, uint64_t seed) {
    const int nb = (int)h->sb_omega.size();
    if (nb != kBathModes)
        return fail(h, NQCD_ERR_UNSUPPORTED, "spin-boson kernels are built for 100 bath modes");
    if (h->D != nb) return fail(h, NQCD_ERR_ARG, "problem dofs must equal the number of bath modes");
    if (S->nparams != 4 * nb || !S->params)
        return fail(h, NQCD_ERR_ARG, "bath sampler params: mu_r[nb], sd_r[nb], mu_v[nb], sd_v[nb]");
    h->ntraj = S->ntraj;
    HIPCHK(h, h->bath.ensure(sizeof(double) * 10 * nb + sizeof(float) * 3 * nb));
    std::vector<double> tab(10 * nb);
    for (int j = 0; j < nb; ++j) {
        tab[j] = h->sb_omega[j] * h->sb_omega[j];
        tab[nb + j] = h->sb_c[j];
        tab[2 * nb + j] = h->inv_mass[j];
        for (int k = 0; k < 4; ++k) tab[3 * nb + 4 * j + k] = S->params[k * nb + j];  // per-mode interleave
    }
    const double hdt = 0.5 * h->prob.dt;
    std::vector<float> tabf(3 * nb);
    for (int j = 0; j < nb; ++j) {
        tab[7 * nb + j] = hdt * tab[j] * tab[2 * nb + j];
        tab[8 * nb + j] = hdt * tab[nb + j] * tab[2 * nb + j];
        tab[9 * nb + j] = tab[nb + j];
        for (int k = 0; k < 3; ++k) tabf[k * nb + j] = (float)tab[(7 + k) * nb + j];
    }
    HIPCHK(h, hipMemcpy(h->bath.p, tab.data(), sizeof(double) * 10 * nb, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->bath.as<char>() + sizeof(double) * 10 * nb, tabf.data(), sizeof(float) * 3 * nb,
                        hipMemcpyHostToDevice));
    // harmonic Wigner sampling: counter (traj, mode, stream 0) -> z0 for R, z1 for V
    HIPCHK(h, h->r0.ensure(sizeof(double) * nb * h->ntraj));
    HIPCHK(h, h->v0.ensure(sizeof(double) * nb * h->ntraj));
    HIPCHK(h, launch_sample_gaussian(h->ntraj, nb, h->bath.as<double>() + 3 * nb, seed, h->r0.as<double>(),
                                     h->v0.as<double>(), h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->sb_state = S->initial_state;
    h->sb_diabatic = S->state_basis == 1;
    h->sb_seed = seed;
    return NQCD_OK;
}

int32_t sb_run(nqcd_handle_t* h, double t0, int64_t nsteps, const nqcd_term_desc* term, uint64_t key) {
    if (term && term->kind != NQCD_TERM_NONE)
        return fail(h, NQCD_ERR_UNSUPPORTED, "termination predicates are not supported for system-bath models");
    const int nb = kBathModes;
    const int64_t n = h->ntraj;
    HIPCHK(h, h->r.ensure(sizeof(double) * nb * n));
    HIPCHK(h, h->v.ensure(sizeof(double) * nb * n));
    HIPCHK(h, h->sig.ensure(sizeof(double) * 4 * n));
    HIPCHK(h, h->state.ensure(sizeof(int32_t) * n));
    HIPCHK(h, h->steps.ensure(sizeof(int32_t) * n));
    HIPCHK(h, h->nhops.ensure(sizeof(int32_t) * n));
    HIPCHK(h, h->save_ad.ensure(sizeof(double) * 2 * h->nsave * n));
    HIPCHK(h, h->save_di.ensure(sizeof(double) * 2 * h->nsave * n));
    HIPCHK(h, h->save_state.ensure(sizeof(int32_t) * h->nsave * n));
    HIPCHK(h, launch_zero(h->save_ad.p, sizeof(double) * 2 * h->nsave * n, h->stream));
    HIPCHK(h, launch_zero(h->save_di.p, sizeof(double) * 2 * h->nsave * n, h->stream));
    HIPCHK(h, launch_zero(h->save_state.p, sizeof(int32_t) * h->nsave * n, h->stream));
    SbArgs A{};
    A.ntraj = n;
    A.method = (h->prob.method == NQCD_METHOD_FSSH) ? 0 : 1;
    A.rescaling = h->prob.rescaling;
    A.nsub = h->prob.electronic_substeps > 0 ? h->prob.electronic_substeps : 5;
    A.save_every = h->save_every;
    A.nsave = h->nsave;
    A.initial_state = h->sb_state;
    A.diabatic = h->sb_diabatic ? 1 : 0;
    A.dt = h->prob.dt;
    A.t0 = t0;
    A.eps = h->sb_eps;
    A.Delta = h->sb_delta;
    double S2 = 0.0, c2 = 0.0;
    for (int j = 0; j < nb; ++j) { S2 += h->sb_c[j] * h->sb_c[j] * h->inv_mass[j]; c2 += h->sb_c[j] * h->sb_c[j]; }
    A.S2 = S2;
    A.cnorm = std::sqrt(c2);
    A.nsteps = nsteps;
    A.key = key;
    const double* tab = h->bath.as<double>();
    A.w2 = tab; A.c = tab + nb; A.invm = tab + 2 * nb;
    A.kd = tab + 7 * nb;
    A.kf = reinterpret_cast<const float*>(h->bath.as<char>() + sizeof(double) * 10 * nb);
    A.r0 = h->r0.as<double>(); A.v0 = h->v0.as<double>();
    A.r = h->r.as<double>(); A.v = h->v.as<double>(); A.sig = h->sig.as<double>();
    A.state = h->state.as<int32_t>(); A.steps = h->steps.as<int32_t>(); A.nhops = h->nhops.as<int32_t>();
    A.save_pop_ad = h->save_ad.as<double>(); A.save_pop_di = h->save_di.as<double>();
    A.save_state = h->save_state.as<int32_t>();
    HIPCHK(h, h->args_dev.ensure(sizeof(SbArgs)));
    HIPCHK(h, hipMemcpyAsync(h->args_dev.p, &A, sizeof(SbArgs), hipMemcpyHostToDevice, h->stream));
    const int block = 64;
    const int grid = (int)((n + block - 1) / block);
    HIPCHK(h, hipEventRecord(h->ev0, h->stream));
    if (h->prob.precision == NQCD_PREC_FP64)
        hipLaunchKernelGGL(sb_kernel_f64, dim3(grid), dim3(block), 0, h->stream, h->args_dev.as<const SbArgs>());
    else
        hipLaunchKernelGGL(sb_kernel_f32, dim3(grid), dim3(block), 0, h->stream, h->args_dev.as<const SbArgs>());
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipEventRecord(h->ev1, h->stream));
    h->stats.block = block;
    h->stats.grid = grid;
    h->stats.launches = 1;
    return NQCD_OK;
}

}  // namespace nqcd
