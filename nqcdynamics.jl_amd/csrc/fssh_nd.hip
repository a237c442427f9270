// fssh_nd.hip -- two-state FSSH / Ehrenfest (BABwithTsit5) for a molecule above a surface: six
// Cartesian coordinates (ndofs = 3, natoms = 2), persistent work-queue kernel with one trajectory
// per lane.
//
// Same step as the one-dimensional kernel (bab_electronics.jl:61-91, surface_hopping.jl:2-16,
// fssh.jl:96-121, surface_hopping.jl:64-99); the loop is split in phases around the model
// evaluation so that a model evaluated cooperatively by the whole wavefront (the MLP surrogate)
// sees every lane at the same point: (1) refill lanes from the queue (raw inputs only), (2) half
// kick + drift of stepping lanes, (3) model evaluation at every lane's current coordinates,
// (4) initialisation of refilled lanes / rest of the step for stepping lanes.
//
// Model (diatomic molecule, bond b = |R1 - R2|, height Z = mu z1 + (1 - mu) z2):
//   V11 = Dn (1 - e^{-an (b - bn)})^2 + Aw e^{-bw Z}
//   V22 = Di (1 - e^{-ai (b - bi)})^2 + Ds (1 - e^{-as (Z - Zs)})^2 + Delta
//   V12 = C e^{-g (Z - Zc)^2}
#include "ensemble.h"
#include "fssh2.h"
#include "fssh_nd.h"
#include "mlp.h"
#include "autotune.h"

namespace nqcd {

struct FsshNdArgs {
    int64_t ntraj;
    int32_t method, rescaling, nsub, save_every;
    int32_t nsave, term_kind, term_dof, pad0;
    double term_value, dt, t0;
    int64_t nsteps;
    uint64_t key;
    double inv_mass[kNdDof];
    double p[kDiatomParams];
    // surrogate
    const unsigned char* mlp_packed;  // two packed members
    MlpNorm nrm;
    float threshold;                  // fallback when the members' H differ by more (Hartree)
    int32_t pad1;
    int32_t* fallbacks;               // per trajectory
    const double* r0; const double* v0; const double* sig0; const int32_t* state0; const int32_t* order;
    double* r; double* v; double* sig; int32_t* state; int32_t* steps; int32_t* nhops;
    double* save_pop_ad; double* save_pop_di; int32_t* save_state;
    double* save_r;                   // [slot][dof][traj] or null
    unsigned long long* queue;
    CellDesc cell;                    // cell callback per atom (3 coordinates, 2 atoms); mode 0: none
};

namespace {

template <class T>
__device__ __forceinline__ const T* nd_opaque(const T* p) {
    asm volatile("" : "+s"(p));
    return p;
}

// The launch arguments the step reads, in scalar registers for the whole persistent loop (read per step
// through the argument pointer, each is a vector-memory round trip with a full wait).
struct NdHot {
    int32_t method, rescaling, nsub, save_every;
    int32_t cell_mode, term_kind, term_dof, pad;
    int64_t nsteps;
    double dt, term_value;
    uint64_t key;
    double inv_mass[kNdDof];
};

__device__ __forceinline__ NdHot nd_hot(const FsshNdArgs& A) {
    NdHot H;
    H.method = uniform_sgpr(A.method); H.rescaling = uniform_sgpr(A.rescaling);
    H.nsub = uniform_sgpr(A.nsub); H.save_every = uniform_sgpr(A.save_every);
    H.cell_mode = uniform_sgpr(A.cell.mode); H.term_kind = uniform_sgpr(A.term_kind);
    H.term_dof = uniform_sgpr(A.term_dof); H.pad = 0;
    H.nsteps = uniform_sgpr(A.nsteps);
    H.dt = uniform_sgpr(A.dt); H.term_value = uniform_sgpr(A.term_value);
    H.key = uniform_sgpr(A.key);
#pragma unroll
    for (int I = 0; I < kNdDof; ++I) H.inv_mass[I] = uniform_sgpr(A.inv_mass[I]);
    return H;
}

// The model parameters through a constant-address-space pointer in scalar registers, made opaque where
// it is used so the sixteen parameters are scalar loads at the model evaluation (they do not change
// during a launch) instead of values held across the loop.
using NdParams = __attribute__((address_space(4))) const double*;

__device__ __forceinline__ NdParams nd_params(const FsshNdArgs* Ap) {
    const uint64_t a = (uint64_t)(&Ap->p[0]);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a), hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    return (NdParams)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ NdParams nd_params_here(NdParams p) {
    asm volatile("" : "+s"(p));
    return p;
}

template <typename R>
__device__ __forceinline__ void morse(R D, R a, R q0, R x, R& f, R& df) {
    const R e = fexp(-a * (x - q0));
    f = D * ((R)1 - e) * ((R)1 - e);
    df = (R)2 * D * a * e * ((R)1 - e);
}

// Diabatic matrix and its gradient with respect to the six coordinates (P: the 16 parameters, through a
// generic or a constant-address-space pointer).
template <typename R, class PT>
__device__ __forceinline__ void diatom_eval(PT P, const R r[kNdDof], R& v11, R& v22, R& v12,
                                            R d11[kNdDof], R d22[kNdDof], R d12[kNdDof]) {
    const R Dn = (R)P[0], an = (R)P[1], bn = (R)P[2], Aw = (R)P[3], bw = (R)P[4], Di = (R)P[5], ai = (R)P[6];
    const R bi = (R)P[7], Ds = (R)P[8], as = (R)P[9], Zs = (R)P[10], Delta = (R)P[11], C = (R)P[12], g = (R)P[13];
    const R Zc = (R)P[14], mu = (R)P[15];
    const R dx = r[0] - r[3], dy = r[1] - r[4], dz = r[2] - r[5];
    const R b = sqrt(dx * dx + dy * dy + dz * dz);
    const R Z = mu * r[2] + ((R)1 - mu) * r[5];
    R mn, dmn, mi, dmi, ms, dms;
    morse(Dn, an, bn, b, mn, dmn);
    morse(Di, ai, bi, b, mi, dmi);
    morse(Ds, as, Zs, Z, ms, dms);
    const R wall = Aw * fexp(-bw * Z);
    const R zc = Z - Zc;
    v12 = C * fexp(-g * zc * zc);
    v11 = mn + wall;
    v22 = mi + ms + Delta;
    const R ib = (R)1 / b;
    const R ub[kNdDof] = {dx * ib, dy * ib, dz * ib, -dx * ib, -dy * ib, -dz * ib};
    const R uZ[kNdDof] = {(R)0, (R)0, mu, (R)0, (R)0, (R)1 - mu};
    const R dZ11 = -bw * wall, dZ12 = (R)-2 * g * zc * v12;
#pragma unroll
    for (int I = 0; I < kNdDof; ++I) {
        d11[I] = dmn * ub[I] + dZ11 * uZ[I];
        d22[I] = dmi * ub[I] + dms * uZ[I];
        d12[I] = dZ12 * uZ[I];
    }
}

// Lane (0-based) of the k-th set bit of a 64-bit mask (k < popcount).
__device__ __forceinline__ int select_bit(uint64_t m, int k) {
    int base = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        const uint64_t low = m & ((1ull << w) - 1ull);
        const int c = __popcll(low);
        if (k >= c) { k -= c; m >>= w; base += w; }
    }
    return base;
}

// Both surrogate members at the coordinates of the wavefront's active lanes; the ensemble mean is
// returned in each active lane, and `far` says the members disagree by more than the threshold on
// some element of H.  Active lanes are packed into the first ceil(n_active / 16) sample tiles (slot
// s takes the lane of the s-th set bit of the active mask), so idle lanes cost no network work.
__device__ __forceinline__ void surrogate_eval(const FsshNdArgs& A, const unsigned char* lds, const float r[kNdDof],
                                               bool active, float& v11, float& v22, float& v12, float d11[kNdDof],
                                               float d22[kNdDof], float d12[kNdDof], bool& far) {
    const int l = __lane_id(), j = l & 15;
    // idle lanes (all coordinates zero) evaluate a harmless configuration instead
    const float bb = (r[0] - r[3]) * (r[0] - r[3]) + (r[1] - r[4]) * (r[1] - r[4]) + (r[2] - r[5]) * (r[2] - r[5]);
    float x[kMlpIn], xn[kMlpIn];
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) x[i] = (bb > 0.0f) ? r[i] : (i == 0 ? 1.0f : 0.0f);
    mlp_descriptor(x, xn);
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) xn[i] = (xn[i] - A.nrm.in_shift[i]) * A.nrm.in_scale[i];
    const uint64_t mask = __ballot(active);
    const int nact = __popcll(mask);
    const int ntile = (nact + 15) >> 4;
    const int slot = __popcll(mask & ((1ull << l) - 1ull));               // this lane's slot, if active
    const int src = (l < nact) ? select_bit(mask, l) : 0;               // lane feeding slot l
    const int mt = slot >> 4, mj = slot & 15;
    // members and sample tiles in one rolled loop: member it / ntile, tile it % ntile
    float H0[kMlpOut], H1[kMlpOut], Js[kMlpOut][kMlpIn];
#pragma unroll
    for (int r3 = 0; r3 < kMlpOut; ++r3) {
        H0[r3] = 0.0f;
        H1[r3] = 0.0f;
#pragma unroll
        for (int i = 0; i < kMlpIn; ++i) Js[r3][i] = 0.0f;
    }
#pragma unroll 1
    for (int it = 0; it < 2 * ntile; ++it) {
        const int m = it >= ntile ? 1 : 0, T = it - m * ntile;
        const int sl = __shfl(src, 16 * T + j);  // lane whose configuration is sample j of tile T
        float xs[kMlpIn];
#pragma unroll
        for (int i = 0; i < kMlpIn; ++i) xs[i] = __shfl(xn[i], sl);
        float o[kMlpOut], Jt[kMlpOut][kMlpIn];
        mlp_tile<true>(lds + m * kPkBytes, xs, o, Jt);
        // outputs sit in lanes 0..15 and the gradient in every lane of the sample: the owner of
        // slot 16 T + mj takes those of sample mj
        const bool mine = active && (mt == T);
#pragma unroll
        for (int r3 = 0; r3 < kMlpOut; ++r3) {
            const float ov = __shfl(o[r3], mj);
            if (mine && m == 0) H0[r3] = ov;
            if (mine && m == 1) H1[r3] = ov;
#pragma unroll
            for (int i = 0; i < kMlpIn; ++i) {
                const float gv = __shfl(Jt[r3][i], mj);
                if (mine) Js[r3][i] += gv;
            }
        }
    }
    float dmax = 0.0f;
    float Hm[kMlpOut];
#pragma unroll
    for (int r3 = 0; r3 < kMlpOut; ++r3) {
        const float a = H0[r3] * A.nrm.out_scale[r3], b = H1[r3] * A.nrm.out_scale[r3];
        dmax = fmaxf(dmax, fabsf(a - b));
        Hm[r3] = 0.5f * (a + b) + A.nrm.out_shift[r3];
    }
    far = !(dmax <= A.threshold);
    v11 = Hm[0]; v22 = Hm[1]; v12 = Hm[2];
    float g11[kMlpIn], g22[kMlpIn], g12[kMlpIn];
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) {
        g11[i] = 0.5f * Js[0][i] * A.nrm.out_scale[0] * A.nrm.in_scale[i];
        g22[i] = 0.5f * Js[1][i] * A.nrm.out_scale[1] * A.nrm.in_scale[i];
        g12[i] = 0.5f * Js[2][i] * A.nrm.out_scale[2] * A.nrm.in_scale[i];
    }
    mlp_descriptor_pullback(x, g11, d11);
    mlp_descriptor_pullback(x, g22, d22);
    mlp_descriptor_pullback(x, g12, d12);
}

template <typename R>
struct LaneN {
    R r[kNdDof], v[kNdDof], k[kNdDof], vt[kNdDof];
    Sig2 s;
    double tr;
    double U[4];
    double dEc, wc, tc, t;
    int64_t tid;
    int32_t st, nh, step, nfb;
};

template <typename R>
__device__ __forceinline__ void nd_save(const FsshNdArgs* Ap, const LaneN<R>& L, int slot) {
    const FsshNdArgs& A = *nd_opaque(Ap);
    if (slot >= A.nsave) return;
    const int64_t n = A.ntraj;
    double pa0, pa1;
    if (A.method == METHOD_FSSH) { pa0 = (L.st == 0) ? 1.0 : 0.0; pa1 = 1.0 - pa0; }
    else { pa0 = 0.5 * (L.tr + L.s.q); pa1 = 0.5 * (L.tr - L.s.q); }
    const double S01 = L.s.x;  // fssh.jl:132-142, ehrenfest.jl:75-83
    const double pd0 = L.U[0] * L.U[0] * pa0 + L.U[1] * L.U[1] * pa1 + 2.0 * L.U[0] * L.U[1] * S01;
    const double pd1 = L.U[2] * L.U[2] * pa0 + L.U[3] * L.U[3] * pa1 + 2.0 * L.U[2] * L.U[3] * S01;
    A.save_pop_ad[((int64_t)slot * 2 + 0) * n + L.tid] = pa0;
    A.save_pop_ad[((int64_t)slot * 2 + 1) * n + L.tid] = pa1;
    A.save_pop_di[((int64_t)slot * 2 + 0) * n + L.tid] = pd0;
    A.save_pop_di[((int64_t)slot * 2 + 1) * n + L.tid] = pd1;
    A.save_state[(int64_t)slot * n + L.tid] = L.st;
    if (A.save_r) {
#pragma unroll
        for (int I = 0; I < kNdDof; ++I) A.save_r[((int64_t)slot * kNdDof + I) * n + L.tid] = (double)L.r[I];
    }
}

template <typename R>
__device__ __forceinline__ void nd_load(const FsshNdArgs* Ap, LaneN<R>& L, int64_t tid) {
    const FsshNdArgs& A = *nd_opaque(Ap);
    const int64_t n = A.ntraj;
    L.tid = tid;
#pragma unroll
    for (int I = 0; I < kNdDof; ++I) {
        L.r[I] = (R)A.r0[(int64_t)I * n + tid];
        L.v[I] = (R)A.v0[(int64_t)I * n + tid];
    }
    const double a = A.sig0[0 * n + tid], b = A.sig0[1 * n + tid];
    L.tr = a + b;
    L.s.q = a - b;
    L.s.x = A.sig0[2 * n + tid];
    L.s.y = A.sig0[3 * n + tid];
    int st = A.state0 ? A.state0[tid] : -1;
    if (A.method == METHOD_FSSH && st < 0) {  // fssh.jl:54
        const double u = philox_uniform((uint64_t)tid, 0u, STREAM_STATE, A.key) * (a + b);
        st = (u < a) ? 0 : 1;
    }
    L.st = st;
}

template <typename R>
__device__ __forceinline__ void nd_finish(const FsshNdArgs* Ap, const LaneN<R>& L) {
    const FsshNdArgs& A = *nd_opaque(Ap);
    const int64_t n = A.ntraj;
#pragma unroll
    for (int I = 0; I < kNdDof; ++I) {
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
    if (A.fallbacks) A.fallbacks[L.tid] = L.nfb;
}

template <typename R>
__device__ __forceinline__ bool nd_done(const FsshNdArgs& A, const NdHot& H, const LaneN<R>& L) {
    if (L.step >= H.nsteps) return true;
    if (H.cell_mode == 2 && cell_outside<3, 2, R>(A.cell, L.r)) return true;  // TerminateCellCallback
    if (H.term_kind != 0) {
        double x = 0.0;
#pragma unroll
        for (int I = 0; I < kNdDof; ++I) x = (I == H.term_dof) ? (double)L.r[I] : x;
        if (term_hit(H.term_kind, x, H.term_value)) return true;
    }
    return false;
}

// Eigen-decomposition, U^T dV U and the acceleration on the active surface or mean field.
template <typename R>
__device__ __forceinline__ void nd_adiabatic(const NdHot& A, LaneN<R>& L, bool track, R v11, R v22, R v12,
                                             const R d11[kNdDof], const R d22[kNdDof], const R d12[kNdDof],
                                             double& E0, double& E1, double G01[kNdDof]) {
    eigen2((double)v11, (double)v12, (double)v22, E0, E1, L.U, track);
    const double a = 0.5 * (L.tr + L.s.q), b = 0.5 * (L.tr - L.s.q);
#pragma unroll
    for (int I = 0; I < kNdDof; ++I) {
        double G00, G11;
        rotate_derivative(L.U, (double)d11[I], (double)d22[I], (double)d12[I], G00, G11, G01[I]);
        double f;
        if (A.method == METHOD_FSSH) f = (L.st == 0) ? G00 : G11;
        else f = G00 * a + G11 * b + 2.0 * G01[I] * L.s.x;
        L.k[I] = (R)(-f * A.inv_mass[I]);
    }
}

// Rest of the step after the model at the drifted coordinates: kick, electronics, hop.
template <typename R>
__device__ __forceinline__ void nd_finish_step(const FsshNdArgs* Ap, const FsshNdArgs& Af, const NdHot& A, LaneN<R>& L,
                                               R v11, R v22, R v12, const R d11[kNdDof], const R d22[kNdDof],
                                               const R d12[kNdDof]) {
    const R hdt = (R)(0.5 * A.dt);
    double E0, E1, G01[kNdDof];
    nd_adiabatic(A, L, true, v11, v22, v12, d11, d22, d12, E0, E1, G01);  // sigma of the step start
#pragma unroll
    for (int I = 0; I < kNdDof; ++I) L.v[I] = L.vt[I] + hdt * L.k[I];
    const double gap = E1 - E0;
    const double inv_gap = 1.0 / gap;
    double d01[kNdDof];
    double w = 0.0;
#pragma unroll
    for (int I = 0; I < kNdDof; ++I) { d01[I] = G01[I] * inv_gap; w += d01[I] * (double)L.v[I]; }
    const double tn = L.t + A.dt;
    const double dEn = E0 - E1;
    const double inv = 1.0 / (tn - L.tc);
    Interp2 P;
    P.dE0 = L.dEc; P.dE1 = (dEn - L.dEc) * inv; P.w0 = L.wc; P.w1 = (w - L.wc) * inv;
    propagate2(P, L.s, L.t, L.tc, A.dt, A.nsub);
    L.dEc = dEn; L.wc = w; L.tc = tn;
    L.t = tn;
    // the save precedes the hop callback (a save time holds the state before the step's callbacks)
    if (A.save_every > 0 && ((L.step + 1) % A.save_every) == 0) nd_save(Ap, L, (L.step + 1) / A.save_every);
    if (A.method == METHOD_FSSH) {
        const int s = L.st;
        const double sss = 0.5 * (L.tr + ((s == 0) ? L.s.q : -L.s.q));
        const double ratio = L.s.x / sss;
        double p = 0.0;
#pragma unroll
        for (int I = 0; I < kNdDof; ++I) {
            const double dsm = (s == 0) ? d01[I] : -d01[I];
            p += 2.0 * (double)L.v[I] * ratio * dsm * A.dt;
        }
        if (p > 0.0) {
            p = fmin(p, 1.0);
            const double rnd = philox_uniform((uint64_t)L.tid, (uint32_t)L.step, STREAM_HOP, A.key);
            if (p > rnd) {
                double dms[kNdDof];
#pragma unroll
                for (int I = 0; I < kNdDof; ++I) dms[I] = (s == 0) ? -d01[I] : d01[I];
                const double dE = (s == 0) ? gap : -gap;
                if (rescale_velocity<kNdDof, R>(A.rescaling, L.v, A.inv_mass, dms, dE)) { L.st = 1 - s; ++L.nh; }
            }
        }
    }
    // CellBoundaryCallback after the hop callback; the step-end acceleration is kept (bab_electronics.jl:61-74)
    if (A.cell_mode == 1 && cell_outside<3, 2, R>(Af.cell, L.r)) cell_wrap<3, 2, R>(Af.cell, L.r);
    ++L.step;
}

template <typename R, bool kMlp>
__device__ __forceinline__ void nd_body(const FsshNdArgs* __restrict__ Ap, const unsigned char* lds) {
    const FsshNdArgs& A = *Ap;
    const NdHot H = nd_hot(A);
    const NdParams P0 = nd_params(Ap);
    const int lane = __lane_id();
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    LaneN<R> L;
#pragma unroll
    for (int I = 0; I < kNdDof; ++I) L.r[I] = (R)0;  // idle lanes still feed the wave-wide surrogate
    bool active = false, exhausted = false, fresh = false;
    for (;;) {
        // (1) refill idle lanes from the queue
        const uint64_t want = __ballot(!active && !exhausted);
        if (want) {
            const int cnt = __popcll(want);
            const int leader = __ffsll((unsigned long long)want) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(nd_opaque(Ap)->queue, (unsigned long long)cnt);
            base = __shfl(base, leader);
            if (!active && !exhausted) {
                const unsigned long long q = base + (unsigned long long)__popcll(want & below);
                if (q < (unsigned long long)A.ntraj) {
                    const int32_t* order = nd_opaque(Ap)->order;
                    nd_load(Ap, L, order ? (int64_t)order[q] : (int64_t)q);
                    L.nfb = 0;
                    active = true;
                    fresh = true;
                } else {
                    exhausted = true;
                }
            }
        }
        if (__ballot(active) == 0) break;
        // (2) half kick + drift (step_B!, step_A!)
        if (active && !fresh) {
            const R hdt = (R)(0.5 * H.dt), dtR = (R)H.dt;
#pragma unroll
            for (int I = 0; I < kNdDof; ++I) {
                L.vt[I] = L.v[I] + hdt * L.k[I];
                L.r[I] = L.r[I] + dtR * L.vt[I];
            }
        }
        // (3) model at the current coordinates (the surrogate packs the wavefront's active lanes)
        R v11, v22, v12, d11[kNdDof], d22[kNdDof], d12[kNdDof];
        if constexpr (kMlp) {
            bool far;
            surrogate_eval(A, lds, L.r, active, v11, v22, v12, d11, d22, d12, far);
            if (active && far) {  // members disagree: the analytic model for this lane and step
                diatom_eval<R>(nd_params_here(P0), L.r, v11, v22, v12, d11, d22, d12);
                ++L.nfb;
            }
        } else {
            if (active) diatom_eval<R>(nd_params_here(P0), L.r, v11, v22, v12, d11, d22, d12);
        }
        // (4) initialise refilled lanes (bab_electronics.jl:48-59), or finish the step
        if (active) {
            if (fresh) {
                double E0, E1, G01[kNdDof];
                nd_adiabatic(H, L, false, v11, v22, v12, d11, d22, d12, E0, E1, G01);
                L.dEc = 0.0; L.wc = 0.0; L.tc = 0.0;  // zero buffers stamped t = 0
                L.t = A.t0;
                L.nh = 0;
                L.step = 0;
                nd_save(Ap, L, 0);
                fresh = false;
                // the termination predicate is a DiscreteCallback, evaluated after each step and not
                // at t0 (callbacks.jl:15-29): a fresh trajectory is done only for zero steps
                if (H.nsteps <= 0) { nd_finish(Ap, L); active = false; }
            } else {
                nd_finish_step(Ap, A, H, L, v11, v22, v12, d11, d22, d12);
                if (nd_done(A, H, L)) { nd_finish(Ap, L); active = false; }
            }
        }
    }
}

__global__ __launch_bounds__(256) void fssh_nd_f64(const FsshNdArgs* __restrict__ Ap) { nd_body<double, false>(Ap, nullptr); }
__global__ __launch_bounds__(256) void fssh_nd_f32(const FsshNdArgs* __restrict__ Ap) { nd_body<float, false>(Ap, nullptr); }

// Surrogate: the two packed members staged once into LDS, read by all eight wavefronts.
constexpr int kMlpBlock = 512;
__global__ __launch_bounds__(kMlpBlock) void fssh_nd_mlp_f32(const FsshNdArgs* __restrict__ Ap) {
    extern __shared__ __align__(16) unsigned char lds[];
    const uint4* src = reinterpret_cast<const uint4*>(nd_opaque(Ap)->mlp_packed);
    uint4* dst = reinterpret_cast<uint4*>(lds);
    for (int i = threadIdx.x; i < 2 * kPkBytes / 16; i += blockDim.x) dst[i] = src[i];
    __syncthreads();
    nd_body<float, true>(Ap, lds);
}
__global__ void nd_queue_reset(unsigned long long* q) { *q = 0ull; }

// ---------------------------------------------------------------------------
// Plain schedule for the diatom-surface model: the straightforward port, one launch per nuclear step,
// the state round-trips through HBM every step, the electronic step uses the dense complex 2x2
// formulation (propagate_dense), FP64 throughout, terminated trajectories are masked.  Analytic
// model only.  State rows (double): r[6], v[6], k[6], sigma re[4], im[4], E0c, E1c, Wc, tc, U[4], t;
// int rows: state, nhops, step, done.
// ---------------------------------------------------------------------------
enum { NP_R = 0, NP_V = 6, NP_K = 12, NP_SRE = 18, NP_SIM = 22, NP_E0C = 26, NP_E1C = 27, NP_WC = 28, NP_TC = 29,
       NP_U = 30, NP_T = 34, NP_NROWS = 35 };
enum { NPI_ST = 0, NPI_NH = 1, NPI_STEP = 2, NPI_DONE = 3, NPI_NROWS = 4 };

__device__ __forceinline__ void nd_plain_save(const FsshNdArgs& A, int64_t tid, int slot, int st, const C2& s,
                                              const double U[4], const double r[kNdDof]) {
    if (slot >= A.nsave) return;
    const int64_t n = A.ntraj;
    double pa0, pa1;
    if (A.method == METHOD_FSSH) { pa0 = (st == 0) ? 1.0 : 0.0; pa1 = 1.0 - pa0; }
    else { pa0 = s.re[0]; pa1 = s.re[3]; }
    const double S01 = s.re[1];
    A.save_pop_ad[((int64_t)slot * 2 + 0) * n + tid] = pa0;
    A.save_pop_ad[((int64_t)slot * 2 + 1) * n + tid] = pa1;
    A.save_pop_di[((int64_t)slot * 2 + 0) * n + tid] = U[0] * U[0] * pa0 + U[1] * U[1] * pa1 + 2.0 * U[0] * U[1] * S01;
    A.save_pop_di[((int64_t)slot * 2 + 1) * n + tid] = U[2] * U[2] * pa0 + U[3] * U[3] * pa1 + 2.0 * U[2] * U[3] * S01;
    A.save_state[(int64_t)slot * n + tid] = st;
    if (A.save_r) {
        for (int I = 0; I < kNdDof; ++I) A.save_r[((int64_t)slot * kNdDof + I) * n + tid] = r[I];
    }
}

// Eigen-decomposition at r, U^T dV U per coordinate and the acceleration on the active surface or
// the mean field of the dense density.
__device__ __forceinline__ void nd_plain_forces(const FsshNdArgs& A, const double r[kNdDof], int st, const C2& s,
                                                bool track, double U[4], double& E0, double& E1, double k[kNdDof],
                                                double G01[kNdDof]) {
    double v11, v22, v12, d11[kNdDof], d22[kNdDof], d12[kNdDof];
    diatom_eval<double>(A.p, r, v11, v22, v12, d11, d22, d12);
    eigen2(v11, v12, v22, E0, E1, U, track);
    for (int I = 0; I < kNdDof; ++I) {
        double G00, G11;
        rotate_derivative(U, d11[I], d22[I], d12[I], G00, G11, G01[I]);
        const double f = (A.method == METHOD_FSSH) ? ((st == 0) ? G00 : G11)
                                                   : (G00 * s.re[0] + G11 * s.re[3] + 2.0 * G01[I] * s.re[1]);
        k[I] = -f * A.inv_mass[I];
    }
}

__global__ __launch_bounds__(256) void nd_plain_init(const FsshNdArgs A, double* S, int32_t* Si) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= A.ntraj) return;
    const int64_t n = A.ntraj;
    double r[kNdDof], v[kNdDof], k[kNdDof], G01[kNdDof], U[4] = {1.0, 0.0, 0.0, 1.0}, E0, E1;
    for (int I = 0; I < kNdDof; ++I) { r[I] = A.r0[(int64_t)I * n + tid]; v[I] = A.v0[(int64_t)I * n + tid]; }
    C2 s;
    s.re[0] = A.sig0[tid]; s.re[3] = A.sig0[n + tid]; s.re[1] = s.re[2] = A.sig0[2 * n + tid];
    s.im[0] = s.im[3] = 0.0; s.im[1] = A.sig0[3 * n + tid]; s.im[2] = -A.sig0[3 * n + tid];
    int st = A.state0 ? A.state0[tid] : -1;
    if (A.method == METHOD_FSSH && st < 0) {  // fssh.jl:54
        const double tot = s.re[0] + s.re[3];
        st = (philox_uniform((uint64_t)tid, 0u, STREAM_STATE, A.key) * tot < s.re[0]) ? 0 : 1;
    }
    nd_plain_forces(A, r, st, s, false, U, E0, E1, k, G01);
    for (int I = 0; I < kNdDof; ++I) {
        S[(NP_R + I) * n + tid] = r[I]; S[(NP_V + I) * n + tid] = v[I]; S[(NP_K + I) * n + tid] = k[I];
    }
    for (int J = 0; J < 4; ++J) { S[(NP_SRE + J) * n + tid] = s.re[J]; S[(NP_SIM + J) * n + tid] = s.im[J]; }
    S[NP_E0C * n + tid] = 0.0; S[NP_E1C * n + tid] = 0.0; S[NP_WC * n + tid] = 0.0; S[NP_TC * n + tid] = 0.0;
    for (int J = 0; J < 4; ++J) S[(NP_U + J) * n + tid] = U[J];
    S[NP_T * n + tid] = A.t0;
    Si[NPI_ST * n + tid] = st; Si[NPI_NH * n + tid] = 0; Si[NPI_STEP * n + tid] = 0;
    Si[NPI_DONE * n + tid] = A.nsteps <= 0 ? 1 : 0;  // the predicates are first evaluated after step 1
    nd_plain_save(A, tid, 0, st, s, U, r);
}

__global__ __launch_bounds__(256) void nd_plain_step(const FsshNdArgs A, double* S, int32_t* Si) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= A.ntraj) return;
    const int64_t n = A.ntraj;
    if (Si[NPI_DONE * n + tid]) return;
    double r[kNdDof], v[kNdDof], k[kNdDof], G01[kNdDof], U[4], E0, E1;
    for (int I = 0; I < kNdDof; ++I) {
        r[I] = S[(NP_R + I) * n + tid]; v[I] = S[(NP_V + I) * n + tid]; k[I] = S[(NP_K + I) * n + tid];
    }
    C2 s;
    for (int J = 0; J < 4; ++J) { s.re[J] = S[(NP_SRE + J) * n + tid]; s.im[J] = S[(NP_SIM + J) * n + tid]; }
    for (int J = 0; J < 4; ++J) U[J] = S[(NP_U + J) * n + tid];
    const double t = S[NP_T * n + tid];
    int st = Si[NPI_ST * n + tid], nh = Si[NPI_NH * n + tid], step = Si[NPI_STEP * n + tid];
    const double dt = A.dt;
    double vt[kNdDof];
    for (int I = 0; I < kNdDof; ++I) { vt[I] = v[I] + 0.5 * dt * k[I]; r[I] = r[I] + dt * vt[I]; }  // step_B!, step_A!
    nd_plain_forces(A, r, st, s, true, U, E0, E1, k, G01);
    for (int I = 0; I < kNdDof; ++I) v[I] = vt[I] + 0.5 * dt * k[I];                            // step_B!
    const double inv_gap = 1.0 / (E1 - E0);
    double d01[kNdDof], w = 0.0;
    for (int I = 0; I < kNdDof; ++I) { d01[I] = G01[I] * inv_gap; w += d01[I] * v[I]; }
    const double tn = t + dt;
    InterpDense P;
    P.E0c = S[NP_E0C * n + tid]; P.E1c = S[NP_E1C * n + tid]; P.Wc = S[NP_WC * n + tid]; P.tc = S[NP_TC * n + tid];
    P.E0n = E0; P.E1n = E1; P.Wn = w; P.inv = 1.0 / (tn - P.tc);
    propagate_dense(P, s, t, dt, A.nsub);
    if (A.save_every > 0 && (step + 1) % A.save_every == 0)  // before the hop, as the persistent kernel
        nd_plain_save(A, tid, (step + 1) / A.save_every, st, s, U, r);
    if (A.method == METHOD_FSSH) {
        const int m = 1 - st;
        // real(sigma[m,s] / sigma[s,s]) with full complex division (fssh.jl:96-121)
        const double nr = s.re[m * 2 + st], ni = s.im[m * 2 + st];
        const double dr = s.re[st * 3], di = s.im[st * 3];
        const double ratio = (nr * dr + ni * di) / (dr * dr + di * di);
        double p = 0.0;
        for (int I = 0; I < kNdDof; ++I) {
            const double dsm = (st == 0) ? d01[I] : -d01[I];
            p += 2.0 * v[I] * ratio * dsm * dt;
        }
        p = fmin(fmax(p, 0.0), 1.0);
        const double rnd = philox_uniform((uint64_t)tid, (uint32_t)step, STREAM_HOP, A.key);
        if (p > rnd) {
            double dms[kNdDof];
            for (int I = 0; I < kNdDof; ++I) dms[I] = (st == 0) ? -d01[I] : d01[I];
            const double dE = (st == 0) ? (E1 - E0) : (E0 - E1);
            if (rescale_velocity<kNdDof, double>(A.rescaling, v, A.inv_mass, dms, dE)) { st = m; ++nh; }
        }
    }
    // CellBoundaryCallback after the hop callback, the step-end acceleration kept
    if (A.cell.mode == 1 && cell_outside<3, 2, double>(A.cell, r)) cell_wrap<3, 2, double>(A.cell, r);
    ++step;
    for (int I = 0; I < kNdDof; ++I) {
        S[(NP_R + I) * n + tid] = r[I]; S[(NP_V + I) * n + tid] = v[I]; S[(NP_K + I) * n + tid] = k[I];
    }
    for (int J = 0; J < 4; ++J) { S[(NP_SRE + J) * n + tid] = s.re[J]; S[(NP_SIM + J) * n + tid] = s.im[J]; }
    S[NP_E0C * n + tid] = E0; S[NP_E1C * n + tid] = E1; S[NP_WC * n + tid] = w; S[NP_TC * n + tid] = tn;
    for (int J = 0; J < 4; ++J) S[(NP_U + J) * n + tid] = U[J];
    S[NP_T * n + tid] = tn;
    Si[NPI_ST * n + tid] = st; Si[NPI_NH * n + tid] = nh; Si[NPI_STEP * n + tid] = step;
    bool done = step >= A.nsteps;
    if (A.cell.mode == 2 && cell_outside<3, 2, double>(A.cell, r)) done = true;  // TerminateCellCallback
    if (A.term_kind != 0 && term_hit(A.term_kind, r[A.term_dof], A.term_value)) done = true;
    if (done) Si[NPI_DONE * n + tid] = 1;
}

__global__ __launch_bounds__(256) void nd_plain_finish(const FsshNdArgs A, const double* S, const int32_t* Si) {
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= A.ntraj) return;
    const int64_t n = A.ntraj;
    for (int I = 0; I < kNdDof; ++I) {
        A.r[(int64_t)I * n + tid] = S[(NP_R + I) * n + tid];
        A.v[(int64_t)I * n + tid] = S[(NP_V + I) * n + tid];
    }
    A.sig[tid] = S[NP_SRE * n + tid];
    A.sig[n + tid] = S[(NP_SRE + 3) * n + tid];
    A.sig[2 * n + tid] = S[(NP_SRE + 1) * n + tid];
    A.sig[3 * n + tid] = S[(NP_SIM + 1) * n + tid];
    A.state[tid] = Si[NPI_ST * n + tid];
    A.nhops[tid] = Si[NPI_NH * n + tid];
    A.steps[tid] = Si[NPI_STEP * n + tid];
}

}  // namespace

hipError_t nd_persistent_occupancy(int precision, int block, int* blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(
        blocks_per_cu, precision == NQCD_PREC_FP64 ? (const void*)fssh_nd_f64 : (const void*)fssh_nd_f32, block, 0);
}

bool nd_is_problem(const nqcd_handle_t* h) {
    return (h->prob.method == NQCD_METHOD_FSSH || h->prob.method == NQCD_METHOD_EHRENFEST) &&
           h->prob.algorithm == NQCD_ALG_BAB_TSIT5 && h->prob.n_states == 2 && h->model_set[0] &&
           (h->models[0].kind == NQCD_MODEL_DIATOM_SURFACE || h->models[0].kind == NQCD_MODEL_MLP_SURROGATE);
}

int32_t nd_set_model(nqcd_handle_t* h, int32_t kind, const double* params, int64_t n) {
    const bool mlp = kind == NQCD_MODEL_MLP_SURROGATE;
    if (kind != NQCD_MODEL_DIATOM_SURFACE && !mlp) return fail(h, NQCD_ERR_UNSUPPORTED, "unknown model kind");
    if (n != kDiatomParams + (mlp ? 1 : 0))
        return fail(h, NQCD_ERR_ARG, mlp ? "surrogate params: threshold, then the 16 diatom-surface parameters"
                                         : "diatom-surface params: Dn an bn Aw bw Di ai bi Ds as Zs Delta C g Zc mu");
    const double* p = params + (mlp ? 1 : 0);
    if (!(p[15] > 0.0 && p[15] < 1.0)) return fail(h, NQCD_ERR_ARG, "diatom-surface: mu must be in (0, 1)");
    if (mlp && !(params[0] >= 0.0)) return fail(h, NQCD_ERR_ARG, "surrogate: threshold must be >= 0");
    for (int i = 0; i < kDiatomParams; ++i) h->nd_params[i] = p[i];
    if (mlp) h->mlp_threshold = params[0];
    ModelSlot m{};
    m.kind = kind;
    h->models[0] = m;
    h->model_set[0] = true;
    return NQCD_OK;
}

int32_t nd_run(nqcd_handle_t* h, double t0, int64_t nsteps, const nqcd_term_desc* term, uint64_t key) {
    if (h->D != kNdDof) return fail(h, NQCD_ERR_UNSUPPORTED, "the diatom-surface model has ndofs = 3, natoms = 2");
    if (h->model_id_given) return fail(h, NQCD_ERR_UNSUPPORTED, "the diatom-surface model uses model slot 0 only");
    const int64_t n = h->ntraj;
    HIPCHK(h, h->r.ensure(sizeof(double) * kNdDof * n));
    HIPCHK(h, h->v.ensure(sizeof(double) * kNdDof * n));
    HIPCHK(h, h->sig.ensure(sizeof(double) * 4 * n));
    HIPCHK(h, h->state.ensure(sizeof(int32_t) * n));
    HIPCHK(h, h->steps.ensure(sizeof(int32_t) * n));
    HIPCHK(h, h->nhops.ensure(sizeof(int32_t) * n));
    HIPCHK(h, h->save_ad.ensure(sizeof(double) * 2 * h->nsave * n));
    HIPCHK(h, h->save_di.ensure(sizeof(double) * 2 * h->nsave * n));
    HIPCHK(h, h->save_state.ensure(sizeof(int32_t) * h->nsave * n));
    HIPCHK(h, h->queue.ensure(sizeof(unsigned long long)));
    HIPCHK(h, h->save_centroid.ensure(sizeof(double) * kNdDof * h->nsave * n));  // configuration per save slot
    HIPCHK(h, launch_zero(h->save_ad.p, sizeof(double) * 2 * h->nsave * n, h->stream));
    HIPCHK(h, launch_zero(h->save_di.p, sizeof(double) * 2 * h->nsave * n, h->stream));
    HIPCHK(h, launch_zero(h->save_state.p, sizeof(int32_t) * h->nsave * n, h->stream));
    HIPCHK(h, launch_zero(h->save_centroid.p, sizeof(double) * kNdDof * h->nsave * n, h->stream));
    FsshNdArgs A{};
    A.ntraj = n;
    A.method = (h->prob.method == NQCD_METHOD_FSSH) ? METHOD_FSSH : METHOD_EHRENFEST;
    A.rescaling = h->prob.rescaling;
    A.nsub = h->prob.electronic_substeps > 0 ? h->prob.electronic_substeps : 5;
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
    for (int I = 0; I < kNdDof; ++I) A.inv_mass[I] = h->inv_mass[I];
    for (int i = 0; i < kDiatomParams; ++i) A.p[i] = h->nd_params[i];
    A.r0 = h->r0.as<double>(); A.v0 = h->v0.as<double>(); A.sig0 = h->sig0.as<double>();
    A.state0 = h->state0_given ? h->state0.as<int32_t>() : nullptr;
    A.order = h->order_given ? h->order.as<int32_t>() : nullptr;
    A.r = h->r.as<double>(); A.v = h->v.as<double>(); A.sig = h->sig.as<double>();
    A.state = h->state.as<int32_t>(); A.steps = h->steps.as<int32_t>(); A.nhops = h->nhops.as<int32_t>();
    A.save_pop_ad = h->save_ad.as<double>(); A.save_pop_di = h->save_di.as<double>();
    A.save_state = h->save_state.as<int32_t>();
    A.save_r = h->save_centroid.as<double>();
    A.queue = h->queue.as<unsigned long long>();
    const bool mlp = h->models[0].kind == NQCD_MODEL_MLP_SURROGATE;
    if (mlp) {
        if (!h->mlp_ready) return fail(h, NQCD_ERR_STATE, "the surrogate has not been trained or loaded");
        if (h->prob.precision != NQCD_PREC_FP32_NUCLEI)
            return fail(h, NQCD_ERR_UNSUPPORTED, "the surrogate runs with FP32 nuclei / FP64 electronics");
        HIPCHK(h, h->fallbacks.ensure(sizeof(int32_t) * n));
        A.mlp_packed = h->mlp_packed.as<unsigned char>();
        A.nrm = h->mlp_norm;
        A.threshold = (float)h->mlp_threshold;
        A.fallbacks = h->fallbacks.as<int32_t>();
    }
    if (h->prob.schedule == NQCD_SCHED_STEP_PER_LAUNCH) {
        if (mlp) return fail(h, NQCD_ERR_UNSUPPORTED, "the step-per-launch schedule runs the analytic model");
        if (h->prob.precision != NQCD_PREC_FP64)
            return fail(h, NQCD_ERR_UNSUPPORTED, "the step-per-launch schedule runs in FP64 only");
        HIPCHK(h, h->plainS.ensure(sizeof(double) * NP_NROWS * n));
        HIPCHK(h, h->plainSi.ensure(sizeof(int32_t) * NPI_NROWS * n));
        double* S = h->plainS.as<double>();
        int32_t* Si = h->plainSi.as<int32_t>();
        const dim3 grid((unsigned)((n + 255) / 256)), block(256);
        HIPCHK(h, hipEventRecord(h->ev0, h->stream));
        hipLaunchKernelGGL(nd_plain_init, grid, block, 0, h->stream, A, S, Si);
        HIPCHK(h, hipGetLastError());
        // identical step launches captured once into a graph and replayed
        const int64_t chunk = 64;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        if (nsteps >= chunk) {
            HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            for (int64_t j = 0; j < chunk; ++j) hipLaunchKernelGGL(nd_plain_step, grid, block, 0, h->stream, A, S, Si);
            HIPCHK(h, hipStreamEndCapture(h->stream, &graph));
            HIPCHK(h, hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        }
        int64_t done = 0, launches = 1;
        for (; done + chunk <= nsteps && exec; done += chunk) {
            HIPCHK(h, hipGraphLaunch(exec, h->stream));
            launches += chunk;
        }
        for (; done < nsteps; ++done, ++launches) {
            hipLaunchKernelGGL(nd_plain_step, grid, block, 0, h->stream, A, S, Si);
            HIPCHK(h, hipGetLastError());
        }
        hipLaunchKernelGGL(nd_plain_finish, grid, block, 0, h->stream, A, (const double*)S, (const int32_t*)Si);
        HIPCHK(h, hipGetLastError());
        HIPCHK(h, hipEventRecord(h->ev1, h->stream));
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        h->stats.block = 256;
        h->stats.grid = (int)grid.x;
        h->stats.launches = launches + 1;
        return NQCD_OK;
    }
    const bool f64 = h->prob.precision == NQCD_PREC_FP64;
    const void* kern = mlp ? (const void*)fssh_nd_mlp_f32 : (f64 ? (const void*)fssh_nd_f64 : (const void*)fssh_nd_f32);
    // The surrogate kernel's workgroup is fixed (its eight waves share the packed members in LDS);
    // the analytic-model kernels take the tuned shape.
    TuneShape sh = launch_shape(h, TUNE_ND);
    if (mlp) sh = TuneShape{kMlpBlock, 0};
    const int block = sh.block;
    const size_t lds = mlp ? 2 * (size_t)kPkBytes : 0;
    if (mlp) HIPCHK(h, hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int bpc = 0;
    HIPCHK(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, kern, block, lds));
    const int64_t grid = persistent_grid(bpc, sh, h->cu_count, n);
    HIPCHK(h, h->args_dev.ensure(sizeof(FsshNdArgs)));
    HIPCHK(h, hipEventRecord(h->ev0, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->args_dev.p, &A, sizeof(FsshNdArgs), hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(nd_queue_reset, dim3(1), dim3(1), 0, h->stream, A.queue);
    if (mlp) hipLaunchKernelGGL(fssh_nd_mlp_f32, dim3((unsigned)grid), dim3(block), lds, h->stream, h->args_dev.as<const FsshNdArgs>());
    else if (f64) hipLaunchKernelGGL(fssh_nd_f64, dim3((unsigned)grid), dim3(block), 0, h->stream, h->args_dev.as<const FsshNdArgs>());
    else hipLaunchKernelGGL(fssh_nd_f32, dim3((unsigned)grid), dim3(block), 0, h->stream, h->args_dev.as<const FsshNdArgs>());
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipEventRecord(h->ev1, h->stream));
    h->stats.block = block;
    h->stats.grid = (int)grid;
    h->stats.launches = 2;
    return NQCD_OK;
}

}  // namespace nqcd
