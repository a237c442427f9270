// fssh2.hip -- two-state FSSH / Ehrenfest trajectory kernels (BABwithTsit5).
//
// Restates bab_electronics.jl:48-91 (B-A-B velocity Verlet, acceleration on the
// active surface or mean field, Tsit5 density propagation with linearly
// interpolated eigenvalues and v.d) followed by the HoppingCallback of
// surface_hopping.jl:2-16 (fssh.jl:96-121 probability and selection,
// surface_hopping.jl:64-99 rescaling).
//
// Two schedules:
//  * fssh2_persistent: one trajectory per lane, all trajectory state in VGPRs,
//    a grid sized to the device that refills lanes from a global work queue as
//    trajectories terminate (uneven lengths), the electronic 2x2 Hermitian
//    density carried as four reals.  Templated on the nuclear precision.
//  * fssh2_plain_*: the straightforward port -- one launch per nuclear step, the
//    state round-trips through HBM every step, the electronic step uses the
//    dense complex 2x2 formulation, FP64 throughout, terminated trajectories are
//    masked rather than replaced.
#include "fssh2.h"

#include <cstdio>
#include <algorithm>
#include <cstdlib>

namespace nqcd {

// Makes a pointer opaque to the optimiser at this point, so loads through it are
// issued here (scalar loads of the argument block) instead of being hoisted into
// SGPRs that stay live, and spill, across the whole persistent loop.
template <class T>
__device__ __forceinline__ const T* opaque(const T* p) {
    asm volatile("" : "+s"(p));
    return p;
}

// ---------------------------------------------------------------------------
// Per-lane trajectory state for the persistent kernel.
// ---------------------------------------------------------------------------
template <int D, typename R>
struct Lane2 {
    R r[D], v[D], k[D];
    Sig2 s;             // q = sigma00 - sigma11, x + iy = sigma01
    double tr;          // sigma00 + sigma11 (constant)
    double U[4];
    double dEc, wc, tc; // electronic buffer at the start of the step
    double t;
    ModelLane<R> m;     // the lane's model, its parameters in the nuclear precision
    int64_t tid;
    int32_t st, nh, step;
};

// The argument fields the step reads, copied once into scalar registers for the whole persistent
// loop.  Read per step through the argument pointer, each is a vector-memory round trip with a full
// wait (the loop's stores keep the compiler from using scalar loads there), on the critical path of
// the longest trajectories once their waves run alone.
template <int D>
struct Fssh2Hot {
    int32_t method, rescaling, nsub, save_every;
    int32_t cell_mode, term_kind, term_dof, pad;
    int64_t nsteps;
    double dt, term_value;
    uint64_t key;
    double inv_mass[D];
};

template <int D>
__device__ __forceinline__ Fssh2Hot<D> fssh2_hot(const Fssh2Args& A) {
    Fssh2Hot<D> H;
    H.method = uniform_sgpr(A.method); H.rescaling = uniform_sgpr(A.rescaling);
    H.nsub = uniform_sgpr(A.nsub); H.save_every = uniform_sgpr(A.save_every);
    H.cell_mode = uniform_sgpr(A.cell.mode); H.term_kind = uniform_sgpr(A.term_kind);
    H.term_dof = uniform_sgpr(A.term_dof); H.pad = 0;
    H.nsteps = uniform_sgpr(A.nsteps);
    H.dt = uniform_sgpr(A.dt); H.term_value = uniform_sgpr(A.term_value);
    H.key = uniform_sgpr(A.key);
#pragma unroll
    for (int I = 0; I < D; ++I) H.inv_mass[I] = uniform_sgpr(A.inv_mass[I]);
    return H;
}

template <int D, typename R>
__device__ __forceinline__ void eval_cache(Lane2<D, R>& L, bool track, double& E0, double& E1, double G00[D],
                                           double G11[D], double G01[D]) {
    static_assert(D == 1, "one-dimensional models");
    R v11, v22, v12, d11, d22, d12;
    model1d_eval_lane<R>(L.m, L.r[0], v11, v22, v12, d11, d22, d12);
    eigen2((double)v11, (double)v12, (double)v22, E0, E1, L.U, track);
    rotate_derivative(L.U, (double)d11, (double)d22, (double)d12, G00[0], G11[0], G01[0]);
}

template <int D, typename R>
__device__ __forceinline__ void set_accel(const Fssh2Hot<D>& H, Lane2<D, R>& L, const double G00[D],
                                          const double G11[D], const double G01[D]) {
    const double a = 0.5 * (L.tr + L.s.q), b = 0.5 * (L.tr - L.s.q);
#pragma unroll
    for (int I = 0; I < D; ++I) {
        double f;
        if (H.method == METHOD_FSSH) f = (L.st == 0) ? G00[I] : G11[I];
        else f = G00[I] * a + G11[I] * b + 2.0 * G01[I] * L.s.x;
        L.k[I] = (R)(-f * H.inv_mass[I]);
    }
}

template <int D, typename R>
__device__ __forceinline__ void save_slot(const Fssh2Args* Ap, const Lane2<D, R>& L, int slot) {
    const Fssh2Args& A = *opaque(Ap);
    if (slot >= A.nsave) return;
    const int64_t n = A.ntraj;
    double pa0, pa1;
    if (A.method == METHOD_FSSH) { pa0 = (L.st == 0) ? 1.0 : 0.0; pa1 = 1.0 - pa0; }
    else { pa0 = 0.5 * (L.tr + L.s.q); pa1 = 0.5 * (L.tr - L.s.q); }
    // diag(U S U^T) with S = [[pa0, Re sigma01],[Re sigma01, pa1]] (fssh.jl:132-142, ehrenfest.jl:75-83)
    const double S01 = L.s.x;
    const double pd0 = L.U[0] * L.U[0] * pa0 + L.U[1] * L.U[1] * pa1 + 2.0 * L.U[0] * L.U[1] * S01;
    const double pd1 = L.U[2] * L.U[2] * pa0 + L.U[3] * L.U[3] * pa1 + 2.0 * L.U[2] * L.U[3] * S01;
    if (A.save_pop_ad) {
        A.save_pop_ad[((int64_t)slot * 2 + 0) * n + L.tid] = pa0;
        A.save_pop_ad[((int64_t)slot * 2 + 1) * n + L.tid] = pa1;
    }
    if (A.save_pop_di) {
        A.save_pop_di[((int64_t)slot * 2 + 0) * n + L.tid] = pd0;
        A.save_pop_di[((int64_t)slot * 2 + 1) * n + L.tid] = pd1;
    }
    if (A.save_state) A.save_state[(int64_t)slot * n + L.tid] = L.st;
    if (A.save_ekin) {  // classical_kinetic_energy and classical_potential_energy at the slot
        R v11, v22, v12, d11, d22, d12;
        model1d_eval_lane<R>(L.m, L.r[0], v11, v22, v12, d11, d22, d12);
        double E0, E1, Ut[4] = {1.0, 0.0, 0.0, 1.0};
        eigen2((double)v11, (double)v12, (double)v22, E0, E1, Ut, false);
        double ek = 0.0;
#pragma unroll
        for (int I = 0; I < D; ++I) ek += 0.5 * (double)L.v[I] * (double)L.v[I] / A.inv_mass[I];
        A.save_ekin[(int64_t)slot * n + L.tid] = ek;
        A.save_epot[(int64_t)slot * n + L.tid] = (A.method == METHOD_FSSH) ? ((L.st == 0) ? E0 : E1) : pa0 * E0 + pa1 * E1;
    }
}

template <int D, typename R>
__device__ __forceinline__ void load_traj(const Fssh2Args* Ap, const Fssh2Hot<D>& H, Lane2<D, R>& L, int64_t tid) {
    const Fssh2Args& A = *opaque(Ap);
    const int64_t n = A.ntraj;
    L.tid = tid;
#pragma unroll
    for (int I = 0; I < D; ++I) {
        L.r[I] = (R)A.r0[(int64_t)I * n + tid];
        L.v[I] = (R)A.v0[(int64_t)I * n + tid];
    }
    const double a = A.sig0[0 * n + tid], b = A.sig0[1 * n + tid];
    L.tr = a + b;
    L.s.q = a - b;
    L.s.x = A.sig0[2 * n + tid];
    L.s.y = A.sig0[3 * n + tid];
    const int mid = A.model_id ? A.model_id[tid] : 0;
    L.m = model_lane<R>(model_regs(A.models, mid));
    int st = A.state0 ? A.state0[tid] : -1;
    if (A.method == METHOD_FSSH && st < 0) {  // fssh.jl:54: sample(Weights(diag(sigma)))
        const double u = philox_uniform((uint64_t)tid, 0u, STREAM_STATE, A.key) * (a + b);
        st = (u < a) ? 0 : 1;
    }
    L.st = st;
    double E0, E1, G00[D], G11[D], G01[D];
    eval_cache(L, false, E0, E1, G00, G11, G01);  // initialize! (bab_electronics.jl:48-59)
    set_accel(H, L, G00, G11, G01);
    L.dEc = 0.0; L.wc = 0.0; L.tc = 0.0;  // zero buffers stamped t = 0 (electronic_dynamics.jl:118-127)
    L.t = A.t0;
    L.nh = 0;
    L.step = 0;
    save_slot(Ap, L, 0);
}

// Set a running trajectory aside in slot `slot` of S, and take one up again (bit-identical state: FP32
// nuclei are stored widened and narrowed back exactly; the model registers come from the model id).
template <int D, typename R>
__device__ __forceinline__ void suspend_traj(const SuspSet& S, uint64_t slot, const Lane2<D, R>& L) {
    static_assert(D == 1, "one-dimensional models");
    const int64_t c = S.cap;
    double* f = S.f + slot;
    f[0 * c] = (double)L.r[0]; f[1 * c] = (double)L.v[0]; f[2 * c] = (double)L.k[0];
    f[3 * c] = L.s.q; f[4 * c] = L.s.x; f[5 * c] = L.s.y; f[6 * c] = L.tr;
    f[7 * c] = L.U[0]; f[8 * c] = L.U[1]; f[9 * c] = L.U[2]; f[10 * c] = L.U[3];
    f[11 * c] = L.dEc; f[12 * c] = L.wc; f[13 * c] = L.tc; f[14 * c] = L.t;
    S.tid[slot] = L.tid;
    S.i[0 * c + slot] = L.st; S.i[1 * c + slot] = L.nh; S.i[2 * c + slot] = L.step;
}

template <int D, typename R>
__device__ __forceinline__ void resume_traj(const Fssh2Args* Ap, const SuspSet& S, uint64_t slot, Lane2<D, R>& L) {
    static_assert(D == 1, "one-dimensional models");
    const Fssh2Args& A = *opaque(Ap);
    const int64_t c = S.cap;
    const double* f = S.f + slot;
    L.r[0] = (R)f[0 * c]; L.v[0] = (R)f[1 * c]; L.k[0] = (R)f[2 * c];
    L.s.q = f[3 * c]; L.s.x = f[4 * c]; L.s.y = f[5 * c]; L.tr = f[6 * c];
    L.U[0] = f[7 * c]; L.U[1] = f[8 * c]; L.U[2] = f[9 * c]; L.U[3] = f[10 * c];
    L.dEc = f[11 * c]; L.wc = f[12 * c]; L.tc = f[13 * c]; L.t = f[14 * c];
    L.tid = S.tid[slot];
    L.st = S.i[0 * c + slot]; L.nh = S.i[1 * c + slot]; L.step = S.i[2 * c + slot];
    const int mid = A.model_id ? A.model_id[L.tid] : 0;
    L.m = model_lane<R>(model_regs(A.models, mid));
}

template <int D, typename R>
__device__ __forceinline__ void finish_traj(const Fssh2Args* Ap, const Lane2<D, R>& L) {
    const Fssh2Args& A = *opaque(Ap);
    const int64_t n = A.ntraj;
#pragma unroll
    for (int I = 0; I < D; ++I) {
        A.r[(int64_t)I * n + L.tid] = (double)L.r[I];
        A.v[(int64_t)I * n + L.tid] = (double)L.v[I];
    }
    A.sig[0 * n + L.tid] = 0.5 * (L.tr + L.s.q);
    A.sig[1 * n + L.tid] = 0.5 * (L.tr - L.s.q);
    A.sig[2 * n + L.tid] = L.s.x;
    A.sig[3 * n + L.tid] = L.s.y;
    A.state[L.tid] = L.st;
    A.steps[L.tid] = L.step;
    A.nhops[L.tid] = L.nh;
    if (A.final_pop_di) {  // diag(U S U^T) of the final state (fssh.jl:132-142, ehrenfest.jl:75-83)
        double pa0, pa1;
        if (A.method == METHOD_FSSH) { pa0 = (L.st == 0) ? 1.0 : 0.0; pa1 = 1.0 - pa0; }
        else { pa0 = 0.5 * (L.tr + L.s.q); pa1 = 0.5 * (L.tr - L.s.q); }
        A.final_pop_di[L.tid] = L.U[0] * L.U[0] * pa0 + L.U[1] * L.U[1] * pa1 + 2.0 * L.U[0] * L.U[1] * L.s.x;
        A.final_pop_di[n + L.tid] = L.U[2] * L.U[2] * pa0 + L.U[3] * L.U[3] * pa1 + 2.0 * L.U[2] * L.U[3] * L.s.x;
    }
}

template <int D, typename R>
__device__ __forceinline__ bool is_done(const Fssh2Args& A, const Fssh2Hot<D>& H, const Lane2<D, R>& L) {
    if (L.step >= H.nsteps) return true;
    if (H.cell_mode == 2 && cell_outside<1, D, R>(A.cell, L.r)) return true;  // TerminateCellCallback
    if (H.term_kind != 0) {
        double x = 0.0;  // select, not a runtime index, so the lane state stays in registers
#pragma unroll
        for (int I = 0; I < D; ++I) x = (I == H.term_dof) ? (double)L.r[I] : x;
        if (term_hit(H.term_kind, x, H.term_value)) return true;
    }
    return false;
}

// One nuclear step + hopping callback.
template <int D, typename R>
__device__ __forceinline__ void step_traj(const Fssh2Args* Ap, const Fssh2Args& A, const Fssh2Hot<D>& H,
                                          Lane2<D, R>& L) {
    const R hdt = (R)(0.5 * H.dt), dtR = (R)H.dt;
    R vt[D];
#pragma unroll
    for (int I = 0; I < D; ++I) vt[I] = L.v[I] + hdt * L.k[I];           // step_B!
#pragma unroll
    for (int I = 0; I < D; ++I) L.r[I] = L.r[I] + dtR * vt[I];           // step_A!
    double E0, E1, G00[D], G11[D], G01[D];
    eval_cache(L, true, E0, E1, G00, G11, G01);                          // update_cache!
    set_accel(H, L, G00, G11, G01);                                      // sigma of the step start
#pragma unroll
    for (int I = 0; I < D; ++I) L.v[I] = vt[I] + hdt * L.k[I];           // step_B!
    const double gap = E1 - E0;
    const double inv_gap = 1.0 / gap;
    double d01[D];
    double w = 0.0;
#pragma unroll
    for (int I = 0; I < D; ++I) { d01[I] = G01[I] * inv_gap; w += d01[I] * (double)L.v[I]; }
    const double tn = L.t + H.dt;
    const double dEn = E0 - E1;
    const double inv = 1.0 / (tn - L.tc);
    Interp2 P;
    P.dE0 = L.dEc; P.dE1 = (dEn - L.dEc) * inv; P.w0 = L.wc; P.w1 = (w - L.wc) * inv;
    propagate2(P, L.s, L.t, L.tc, H.dt, H.nsub);
    L.dEc = dEn; L.wc = w; L.tc = tn;
    L.t = tn;
    // A save time's values are those before the step's discrete callbacks (DiffEqBase's
    // apply_discrete_callback! saves the due saveat points before affect!): the hop comes after the save.
    if (H.save_every > 0 && ((L.step + 1) % H.save_every) == 0) save_slot(Ap, L, (L.step + 1) / H.save_every);
    if (H.method == METHOD_FSSH) {
        const int s = L.st;
        const double sss = 0.5 * (L.tr + ((s == 0) ? L.s.q : -L.s.q));
        const double ratio = L.s.x / sss;  // Re(sigma_ms / sigma_ss)
        double p = 0.0;
#pragma unroll
        for (int I = 0; I < D; ++I) {
            const double dsm = (s == 0) ? d01[I] : -d01[I];
            p += 2.0 * (double)L.v[I] * ratio * dsm * H.dt;
        }
        if (p > 0.0) {  // a uniform in [0,1) can only fall below a positive probability
            p = fmin(p, 1.0);
            const double rnd = philox_uniform((uint64_t)L.tid, (uint32_t)L.step, STREAM_HOP, H.key);
            if (p > rnd) {
                double dms[D];
#pragma unroll
                for (int I = 0; I < D; ++I) dms[I] = (s == 0) ? -d01[I] : d01[I];  // coupling[new, old]
                const double dE = (s == 0) ? gap : -gap;
                if (rescale_velocity<D, R>(H.rescaling, L.v, H.inv_mass, dms, dE)) { L.st = 1 - s; ++L.nh; }
            }
        }
    }
    // CellBoundaryCallback after the hop callback (the problem's callbacks precede the user's); the
    // acceleration of the step end is kept, as BABwithTsit5's cache keeps it (bab_electronics.jl:61-74)
    if (H.cell_mode == 1 && cell_outside<1, D, R>(A.cell, L.r)) cell_wrap<1, D, R>(A.cell, L.r);
    ++L.step;
}

template <int D, typename R>
__device__ __forceinline__ void persistent_body(const Fssh2Args* __restrict__ Ap) {
    const Fssh2Args& A = *Ap;
    const Fssh2Hot<D> H = fssh2_hot<D>(A);
    const int lane = __lane_id();
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    Lane2<D, R> L;
    bool active = false, exhausted = false;
    int lvl = 0, wave_lvl = 0;   // priority class of the lane's trajectory; the wave's current priority
    for (;;) {
        const uint64_t want = __ballot(!active && !exhausted);
        if (want) {
            // the first launch claims trajectory ids from the queue, a resuming launch slots of sin
            const bool resuming = opaque(Ap)->sin.f != nullptr;
            const unsigned long long limit = resuming ? *opaque(Ap)->sin.count : (unsigned long long)A.ntraj;
            const int cnt = __popcll(want);
            const int leader = __ffsll((unsigned long long)want) - 1;
            unsigned long long base = 0;
            if (lane == leader)
                base = atomicAdd(resuming ? opaque(Ap)->sin.claim : opaque(Ap)->queue, (unsigned long long)cnt);
            base = __shfl(base, leader);
            if (!active && !exhausted) {
                const unsigned long long q = base + (unsigned long long)__popcll(want & below);
                if (q < limit) {
                    if (resuming) {
                        resume_traj(Ap, opaque(Ap)->sin, q, L);
                        active = true;
                        lvl = 3;
                    } else {
                        const int64_t span = opaque(Ap)->prio_span;
                        lvl = span > 0 ? 3 - (int)min((unsigned long long)3, q / (unsigned long long)span) : 0;
                        const int32_t* order = opaque(Ap)->order;
                        load_traj(Ap, H, L, order ? (int64_t)order[q] : (int64_t)q);
                        // the termination predicate is a DiscreteCallback, evaluated after each step and
                        // not at t0 (callbacks.jl:15-29): a fresh trajectory is done only for zero steps
                        if (H.nsteps <= 0) finish_traj(Ap, L); else active = true;
                    }
                } else {
                    exhausted = true;
                }
            }
        }
        if (__ballot(active) == 0) break;
        if (A.prio_span > 0) {   // s_setprio takes an immediate
            const int w = __ballot(active && lvl >= 3) ? 3 : __ballot(active && lvl >= 2) ? 2
                        : __ballot(active && lvl >= 1) ? 1 : 0;
            if (w != wave_lvl) {
                wave_lvl = w;
                if (w == 3) __builtin_amdgcn_s_setprio(3);
                else if (w == 2) __builtin_amdgcn_s_setprio(2);
                else if (w == 1) __builtin_amdgcn_s_setprio(1);
                else __builtin_amdgcn_s_setprio(0);
            }
        }
        if (active) {
            step_traj(Ap, A, H, L);
            if (is_done(A, H, L)) { finish_traj(Ap, L); active = false; }
        }
        // Once the queue has drained, a wave with few running lanes sets them aside for the next launch,
        // which packs them into full waves (instead of issuing for its idle lanes to the end).
        if (__ballot(exhausted) != 0 && opaque(Ap)->sout.f != nullptr) {
            const uint64_t act = __ballot(active);
            if (__popcll(act) <= opaque(Ap)->susp_thr) {
                if (act) {
                    const int leader = __ffsll((unsigned long long)act) - 1;
                    unsigned long long base = 0;
                    if (lane == leader) base = atomicAdd(opaque(Ap)->sout.count, (unsigned long long)__popcll(act));
                    base = __shfl(base, leader);
                    if (active) suspend_traj(opaque(Ap)->sout, base + (unsigned long long)__popcll(act & below), L);
                }
                break;
            }
        }
    }
}

// FP32 nuclei: fits 128 VGPRs, so four waves per SIMD.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void fssh2_persistent_f32(
    const Fssh2Args* __restrict__ Ap) {
    persistent_body<1, float>(Ap);
}

// FP64 (parity) path: three waves per SIMD without spilling.
__global__ __launch_bounds__(256) void fssh2_persistent_f64(const Fssh2Args* __restrict__ Ap) {
    persistent_body<1, double>(Ap);
}

// ---------------------------------------------------------------------------
// Plain schedule: HBM-resident state, one launch per nuclear step.
// State rows (double): r, v, k, sigma re[4], im[4], E0c, E1c, Wc, tc, U[4], t.
// Int rows: state, nhops, step, done.
// ---------------------------------------------------------------------------
enum { PL_R = 0, PL_V = 1, PL_K = 2, PL_SRE = 3, PL_SIM = 7, PL_E0C = 11, PL_E1C = 12, PL_WC = 13, PL_TC = 14,
       PL_U = 15, PL_T = 19, PL_NROWS = 20 };
enum { PI_ST = 0, PI_NH = 1, PI_STEP = 2, PI_DONE = 3, PI_NROWS = 4 };

__device__ __forceinline__ void plain_save(const Fssh2Args& A, int64_t tid, int slot, int st, const C2& s,
                                           const double U[4], double v, double E0, double E1) {
    if (slot >= A.nsave) return;
    const int64_t n = A.ntraj;
    double S00, S11, pa0, pa1;
    if (A.method == METHOD_FSSH) { pa0 = (st == 0) ? 1.0 : 0.0; pa1 = 1.0 - pa0; S00 = pa0; S11 = pa1; }
    else { pa0 = s.re[0]; pa1 = s.re[3]; S00 = pa0; S11 = pa1; }
    const double S01 = s.re[1];
    const double pd0 = U[0] * U[0] * S00 + U[1] * U[1] * S11 + 2.0 * U[0] * U[1] * S01;
    const double pd1 = U[2] * U[2] * S00 + U[3] * U[3] * S11 + 2.0 * U[2] * U[3] * S01;
    if (A.save_pop_ad) { A.save_pop_ad[((int64_t)slot * 2) * n + tid] = pa0; A.save_pop_ad[((int64_t)slot * 2 + 1) * n + tid] = pa1; }
    if (A.save_pop_di) { A.save_pop_di[((int64_t)slot * 2) * n + tid] = pd0; A.save_pop_di[((int64_t)slot * 2 + 1) * n + tid] = pd1; }
    if (A.save_state) A.save_state[(int64_t)slot * n + tid] = st;
    if (A.save_ekin) {  // classical_kinetic_energy and classical_potential_energy at the slot
        A.save_ekin[(int64_t)slot * n + tid] = 0.5 * v * v / A.inv_mass[0];
        A.save_epot[(int64_t)slot * n + tid] = (A.method == METHOD_FSSH) ? ((st == 0) ? E0 : E1) : pa0 * E0 + pa1 * E1;
    }
}

__global__ __launch_bounds__(256) void fssh2_plain_init(const Fssh2Args A, double* S, int32_t* Si) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= A.ntraj) return;
    const int64_t n = A.ntraj;
    const double r = A.r0[tid], v = A.v0[tid];
    C2 s;
    s.re[0] = A.sig0[tid]; s.re[3] = A.sig0[n + tid]; s.re[1] = s.re[2] = A.sig0[2 * n + tid];
    s.im[0] = s.im[3] = 0.0; s.im[1] = A.sig0[3 * n + tid]; s.im[2] = -A.sig0[3 * n + tid];
    const int mid = A.model_id ? A.model_id[tid] : 0;
    int st = A.state0 ? A.state0[tid] : -1;
    if (A.method == METHOD_FSSH && st < 0) {
        const double tot = s.re[0] + s.re[3];
        st = (philox_uniform((uint64_t)tid, 0u, STREAM_STATE, A.key) * tot < s.re[0]) ? 0 : 1;
    }
    double v11, v22, v12, d11, d22, d12, E0, E1, U[4], G00, G11, G01;
    model1d_eval<double>(model_regs(A.models, mid), r, v11, v22, v12, d11, d22, d12);
    eigen2(v11, v12, v22, E0, E1, U, false);
    rotate_derivative(U, d11, d22, d12, G00, G11, G01);
    const double f = (A.method == METHOD_FSSH) ? ((st == 0) ? G00 : G11) : (G00 * s.re[0] + G11 * s.re[3] + 2.0 * G01 * s.re[1]);
    S[PL_R * n + tid] = r; S[PL_V * n + tid] = v; S[PL_K * n + tid] = -f * A.inv_mass[0];
    for (int J = 0; J < 4; ++J) { S[(PL_SRE + J) * n + tid] = s.re[J]; S[(PL_SIM + J) * n + tid] = s.im[J]; }
    S[PL_E0C * n + tid] = 0.0; S[PL_E1C * n + tid] = 0.0; S[PL_WC * n + tid] = 0.0; S[PL_TC * n + tid] = 0.0;
    for (int J = 0; J < 4; ++J) S[(PL_U + J) * n + tid] = U[J];
    S[PL_T * n + tid] = A.t0;
    Si[PI_ST * n + tid] = st; Si[PI_NH * n + tid] = 0; Si[PI_STEP * n + tid] = 0;
    const bool done = A.nsteps <= 0;  // the termination predicate is first evaluated after step 1
    Si[PI_DONE * n + tid] = done ? 1 : 0;
    plain_save(A, tid, 0, st, s, U, v, E0, E1);
}

__global__ __launch_bounds__(256) void fssh2_plain_step(const Fssh2Args A, double* S, int32_t* Si) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= A.ntraj) return;
    const int64_t n = A.ntraj;
    if (Si[PI_DONE * n + tid]) return;
    const int mid = A.model_id ? A.model_id[tid] : 0;
    const ModelRegs M = model_regs(A.models, mid);
    double r = S[PL_R * n + tid], v = S[PL_V * n + tid], k = S[PL_K * n + tid];
    C2 s;
    for (int J = 0; J < 4; ++J) { s.re[J] = S[(PL_SRE + J) * n + tid]; s.im[J] = S[(PL_SIM + J) * n + tid]; }
    double U[4];
    for (int J = 0; J < 4; ++J) U[J] = S[(PL_U + J) * n + tid];
    const double t = S[PL_T * n + tid];
    int st = Si[PI_ST * n + tid], nh = Si[PI_NH * n + tid], step = Si[PI_STEP * n + tid];
    const double dt = A.dt;
    const double vt = v + 0.5 * dt * k;
    r = r + dt * vt;
    double v11, v22, v12, d11, d22, d12, E0, E1, G00, G11, G01;
    model1d_eval<double>(M, r, v11, v22, v12, d11, d22, d12);
    eigen2(v11, v12, v22, E0, E1, U, true);
    rotate_derivative(U, d11, d22, d12, G00, G11, G01);
    const double f = (A.method == METHOD_FSSH) ? ((st == 0) ? G00 : G11) : (G00 * s.re[0] + G11 * s.re[3] + 2.0 * G01 * s.re[1]);
    k = -f * A.inv_mass[0];
    v = vt + 0.5 * dt * k;
    const double d01 = G01 / (E1 - E0);
    const double w = d01 * v;
    const double tn = t + dt;
    InterpDense P;
    P.E0c = S[PL_E0C * n + tid]; P.E1c = S[PL_E1C * n + tid]; P.Wc = S[PL_WC * n + tid]; P.tc = S[PL_TC * n + tid];
    P.E0n = E0; P.E1n = E1; P.Wn = w; P.inv = 1.0 / (tn - P.tc);
    propagate_dense(P, s, t, dt, A.nsub);
    if (A.save_every > 0 && (step + 1) % A.save_every == 0)  // before the hop, as the persistent kernel
        plain_save(A, tid, (step + 1) / A.save_every, st, s, U, v, E0, E1);
    if (A.method == METHOD_FSSH) {
        const int m = 1 - st;
        // real(sigma[m,s] / sigma[s,s]) with full complex division
        const double nr = s.re[m * 2 + st], ni = s.im[m * 2 + st];
        const double dr = s.re[st * 3], di = s.im[st * 3];
        const double ratio = (nr * dr + ni * di) / (dr * dr + di * di);
        const double dsm = (st == 0) ? d01 : -d01;
        double p = 2.0 * v * ratio * dsm * dt;
        p = fmin(fmax(p, 0.0), 1.0);
        const double rnd = philox_uniform((uint64_t)tid, (uint32_t)step, STREAM_HOP, A.key);
        if (p > rnd) {
            double vv[1] = {v};
            const double dms[1] = {(st == 0) ? -d01 : d01};
            const double dE = (st == 0) ? (E1 - E0) : (E0 - E1);
            if (rescale_velocity<1, double>(A.rescaling, vv, A.inv_mass, dms, dE)) { st = m; ++nh; }
            v = vv[0];
        }
    }
    bool cell_end = false;
    if (A.cell.mode != 0) {  // cell callback after the hop callback, as the persistent kernel
        double rr[1] = {r};
        cell_end = cell_callback<1, 1, double>(A.cell, rr);
        r = rr[0];
    }
    ++step;
    S[PL_R * n + tid] = r; S[PL_V * n + tid] = v; S[PL_K * n + tid] = k;
    for (int J = 0; J < 4; ++J) { S[(PL_SRE + J) * n + tid] = s.re[J]; S[(PL_SIM + J) * n + tid] = s.im[J]; }
    S[PL_E0C * n + tid] = E0; S[PL_E1C * n + tid] = E1; S[PL_WC * n + tid] = w; S[PL_TC * n + tid] = tn;
    for (int J = 0; J < 4; ++J) S[(PL_U + J) * n + tid] = U[J];
    S[PL_T * n + tid] = tn;
    Si[PI_ST * n + tid] = st; Si[PI_NH * n + tid] = nh; Si[PI_STEP * n + tid] = step;
    const bool done = (step >= A.nsteps) || cell_end || term_hit(A.term_kind, r, A.term_value);
    if (done) Si[PI_DONE * n + tid] = 1;
}

__global__ __launch_bounds__(256) void fssh2_plain_finish(const Fssh2Args A, const double* S, const int32_t* Si) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= A.ntraj) return;
    const int64_t n = A.ntraj;
    A.r[tid] = S[PL_R * n + tid];
    A.v[tid] = S[PL_V * n + tid];
    A.sig[tid] = S[PL_SRE * n + tid];
    A.sig[n + tid] = S[(PL_SRE + 3) * n + tid];
    A.sig[2 * n + tid] = S[(PL_SRE + 1) * n + tid];
    A.sig[3 * n + tid] = S[(PL_SIM + 1) * n + tid];
    A.state[tid] = Si[PI_ST * n + tid];
    A.nhops[tid] = Si[PI_NH * n + tid];
    A.steps[tid] = Si[PI_STEP * n + tid];
    if (A.final_pop_di) {  // diag(U S U^T) of the final state (fssh.jl:132-142, ehrenfest.jl:75-83)
        double U[4];
        for (int J = 0; J < 4; ++J) U[J] = S[(PL_U + J) * n + tid];
        const int st = Si[PI_ST * n + tid];
        double S00, S11;
        if (A.method == METHOD_FSSH) { S00 = (st == 0) ? 1.0 : 0.0; S11 = 1.0 - S00; }
        else { S00 = S[PL_SRE * n + tid]; S11 = S[(PL_SRE + 3) * n + tid]; }
        const double S01 = S[(PL_SRE + 1) * n + tid];
        A.final_pop_di[tid] = U[0] * U[0] * S00 + U[1] * U[1] * S11 + 2.0 * U[0] * U[1] * S01;
        A.final_pop_di[n + tid] = U[2] * U[2] * S00 + U[3] * U[3] * S11 + 2.0 * U[2] * U[3] * S01;
    }
}

__global__ void queue_reset(unsigned long long* q) { *q = 0ull; }

// zeroes the counters of a set of set-aside trajectories before a launch fills it
__global__ void susp_reset(unsigned long long* count, unsigned long long* claim) {
    if (threadIdx.x == 0) { *count = 0ull; *claim = 0ull; }
}

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------
int64_t fssh2_prio_span(int64_t lanes) {  // read at every run
    double f = kPrioDefaultSpan;
    if (const char* e = std::getenv("NQCD_MI355X_PRIO")) f = std::atof(e);
    if (!(f > 0.0) || f > 1e6) return 0;
    return std::max<int64_t>(1, (int64_t)(f * (double)lanes));
}

void fssh2_suspend_policy(int* rounds, int* threshold) {  // read at every run
    int r = kSuspDefaultRounds, t = kSuspDefaultThreshold;
    if (const char* e = std::getenv("NQCD_MI355X_SUSPEND")) {
        int a = 0, b = 0;
        if (std::sscanf(e, "%d,%d", &a, &b) == 2 && a >= 1 && a <= kSuspRounds && b >= 0 && b <= 64) { r = a; t = b; }
    }
    *rounds = r;
    *threshold = t;
}

hipError_t fssh2_launch_persistent(const Fssh2Args& A, Fssh2Args* Adev, int precision, int grid, int block,
                                   hipStream_t stream, const SuspSet* sets) {
    // round r reads sets[(r - 1) % 2] and fills sets[r % 2]; the last round runs every trajectory it
    // holds to the end, so no trajectory is left set aside
    int rounds = 1, thr = 0;
    if (sets) fssh2_suspend_policy(&rounds, &thr);
    Fssh2Args R[kSuspRounds];
    for (int r = 0; r < rounds; ++r) {
        R[r] = A;
        R[r].sin = (r == 0) ? SuspSet{} : sets[(r - 1) % 2];
        R[r].sout = (r == rounds - 1) ? SuspSet{} : sets[r % 2];
        R[r].susp_thr = thr;
    }
    hipError_t e = hipMemcpyAsync(Adev, R, sizeof(Fssh2Args) * rounds, hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(queue_reset, dim3(1), dim3(1), 0, stream, A.queue);
    for (int r = 0; r < rounds; ++r) {
        if (R[r].sout.f) hipLaunchKernelGGL(susp_reset, dim3(1), dim3(64), 0, stream, R[r].sout.count, R[r].sout.claim);
        const Fssh2Args* Ar = Adev + r;
        if (precision == 0)
            hipLaunchKernelGGL(fssh2_persistent_f64, dim3(grid), dim3(block), 0, stream, Ar);
        else
            hipLaunchKernelGGL(fssh2_persistent_f32, dim3(grid), dim3(block), 0, stream, Ar);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t fssh2_persistent_occupancy(int precision, int block, int* blocks_per_cu) {
    if (precision == 0)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, fssh2_persistent_f64, block, 0);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, fssh2_persistent_f32, block, 0);
}

hipError_t fssh2_launch_plain_init(const Fssh2Args& A, double* S, int32_t* Si, hipStream_t stream) {
    const int block = 256;
    const int grid = (int)((A.ntraj + block - 1) / block);
    hipLaunchKernelGGL(fssh2_plain_init, dim3(grid), dim3(block), 0, stream, A, S, Si);
    return hipGetLastError();
}

hipError_t fssh2_launch_plain_step(const Fssh2Args& A, double* S, int32_t* Si, hipStream_t stream) {
    const int block = 256;
    const int grid = (int)((A.ntraj + block - 1) / block);
    hipLaunchKernelGGL(fssh2_plain_step, dim3(grid), dim3(block), 0, stream, A, S, Si);
    return hipGetLastError();
}

hipError_t fssh2_launch_plain_finish(const Fssh2Args& A, double* S, int32_t* Si, hipStream_t stream) {
    const int block = 256;
    const int grid = (int)((A.ntraj + block - 1) / block);
    hipLaunchKernelGGL(fssh2_plain_finish, dim3(grid), dim3(block), 0, stream, A, S, Si);
    return hipGetLastError();
}

}  // namespace nqcd
