// ring_polymer.hip -- ring-polymer dynamics: RPMD (BCB, bcb.jl:81-116) and ring-polymer
// surface hopping / Ehrenfest (BCBwithTsit5, bcb_electronics.jl:53-97) on one-dimensional
// models, with the centroid hop and bead-broadcast rescaling of rpsh.jl:30-50.
//
// Layout: a ring polymer of P beads occupies P consecutive lanes; lane l holds bead l and
// normal mode l.  The state is carried in normal-mode coordinates in the discrete Hartley
// basis H[j][k] = (cos 2pi jk/P + sin 2pi jk/P)/sqrt(P).  H is orthogonal and symmetric and
// its column k lies in the spring eigenspace of w_k = 2 w_n sin(k pi/P) (ring_polymer.jl:59-60),
// so the B-C-B step written in these modes equals the reference's bead-space step
// (transform, Cayley per mode, transform back) in exact arithmetic.  Each lane keeps its row
// of H in registers; the two transforms per step (modes -> bead positions, bead forces ->
// modes) are P-term dot products against values broadcast through LDS.
// The centroid electronics (eigen, coupling, Tsit5 density, hop and rescaling; once per ring
// polymer, bcb_electronics.jl:83-96) run on electronics lanes: lane e of the workgroup's first
// waves takes ring polymer e of the workgroup.  The mode-0 lanes hand the centroid position and
// velocity over through LDS, and the electronics lanes hand back the active state / density of
// the step start (for the bead forces) and the centroid velocity change of an accepted hop, so
// each ring polymer's electronics is evaluated once instead of on all P of its lanes.  A workgroup
// holds 64 ring polymers where the block size allows (one full electronics wavefront).
#include "ensemble.h"
#include "ring_polymer.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace nqcd {

// Workgroup size per bead count, precision and kernel family.  RPMD (E = false): 64 ring polymers up
// to 512 lanes in FP32 for up to 16 beads, else 256 (FP64 Hartley rows take twice the registers, and
// 512-lane workgroups at 32 and 64 beads spill).  With the centroid
// electronics (E = true) at most 256 lanes: the electronics lanes sit in one wavefront of the
// workgroup while the others wait for them, so several smaller workgroups per compute unit keep its
// SIMDs busy (their electronics wavefronts land on different SIMDs).
template <int P, typename R, bool E>
struct RpShape {
    static constexpr int kMax = (E || P > 16 || sizeof(R) != 4) ? 256 : 512;
    static constexpr int block = (64 * P < kMax) ? 64 * P : kMax;
    static constexpr int nr = block / P;   // ring polymers per workgroup
};
constexpr int kRpBlock = 256;              // the sampler's workgroup


struct RpArgs {
    int64_t ntraj;
    int32_t method, rescaling, nsub, save_every;
    int32_t nsave, initial_state, diabatic, nstates;
    double dt, t0, mass, wn;
    int64_t nsteps;
    uint64_t key;
    ModelSlot model;
    const double* r0;     // [P][ntraj] bead positions
    const double* v0;     // [P][ntraj]
    const double* sig0;   // [4][ntraj] or null
    const int32_t* state0;
    double* r;            // [P][ntraj]
    double* v;
    double* sig;          // [4][ntraj]
    int32_t* state;
    int32_t* steps;
    int32_t* nhops;
    double* save_centroid;  // [nsave][ntraj]
    double* save_pop_ad;    // [nsave][2][ntraj]
    double* save_pop_di;
    int32_t* save_state;
    // Cayley coefficients c11[P], c12[P], c21[P] in the bead precision, written on the device before
    // the one-ring-polymer-per-lane kernels of 32 and more beads (rp_cayley_fill)
    double cay[3 * 64];
};

template <class T>
__device__ __forceinline__ const T* rp_opaque(const T* p) {
    asm volatile("" : "+s"(p));
    return p;
}

// y_l = sum_k H[l][k] x_k over the lanes of one ring polymer (x broadcast through LDS).
template <int P, typename R>
__device__ __forceinline__ R hartley(R x, const R H[P], R* lds, int gbase) {
    lds[threadIdx.x] = x;
    __syncthreads();
    R acc = (R)0;
#pragma unroll
    for (int k = 0; k < P; ++k) acc += H[k] * lds[gbase + k];
    __syncthreads();
    __builtin_amdgcn_sched_barrier(0);   // keeps the compiler from overlapping transforms (register pressure)
    return acc;
}

template <int P, typename R>
__device__ __forceinline__ R group_value(R x, R* lds, int gbase, int lane_src) {
    lds[threadIdx.x] = x;
    __syncthreads();
    const R y = lds[gbase + lane_src];
    __syncthreads();
    return y;
}

// Gradient of the surface a bead moves on: the single-state potential (RPMD, sh = false) or the
// active adiabatic surface of two states (RPSH), without the eigenvectors: E = m -+ rad with
// m = (V11 + V22) / 2, h = (V11 - V22) / 2, rad = sqrt(h^2 + V12^2), so
// dE/dx = dm -+ (h dh + V12 dV12) / rad, in the bead precision.
template <typename R>
__device__ __forceinline__ R surface_gradient(const ModelRegs& m, R x, bool sh, int st) {
    R v11, v22, v12, d11, d22, d12;
    model1d_eval<R>(m, x, v11, v22, v12, d11, d22, d12);
    if (!sh) return d11;
    const R h = (R)0.5 * (v11 - v22), dh = (R)0.5 * (d11 - d22), dm = (R)0.5 * (d11 + d22);
    const R rad = sqrt(h * h + v12 * v12);
    const R g = (rad > (R)0) ? (h * dh + v12 * d12) / rad : (R)0;
    return (st == 0) ? dm - g : dm + g;
}

// Bead force (per lane): RPMD uses the single-state potential; RPSH the active adiabatic
// surface; RP-Ehrenfest the mean field with the centroid density (ehrenfest_rpmd.jl:23-43).
template <typename R, bool E>
__device__ __forceinline__ double bead_force(const RpArgs& A, const ModelRegs& m, R x, double Ub[4], bool track,
                                             int st, double a, double b, double sx) {
    if (!E || A.method != 2) return (double)surface_gradient<R>(m, x, E && A.method == 1, st);
    R v11, v22, v12, d11, d22, d12;
    model1d_eval<R>(m, x, v11, v22, v12, d11, d22, d12);
    double E0, E1, G00, G11, G01;
    eigen2((double)v11, (double)v12, (double)v22, E0, E1, Ub, track);
    rotate_derivative(Ub, (double)d11, (double)d22, (double)d12, G00, G11, G01);
    return G00 * a + G11 * b + 2.0 * G01 * sx;
}

// Centroid electronics of one ring polymer (bcb_electronics.jl:83-96): the density carried as its
// trace tr and (q, x, y), the active state and hop count, the tracked eigenvectors at the centroid,
// the interpolation buffer's last end (dE and w at time tc), and the step end's coupling and gap.
struct RpElec {
    double tr;
    Sig2 s;
    int est, nh;
    double Uc[4];
    double dEc, wc, tc;
    double d01, gap;
};

// Initial density and active state of ring polymer id from the centroid at rc (fssh.jl:54 draws a
// state when none is given); returns the density's diagonal and real coherence as given.
template <typename R>
__device__ __forceinline__ void rp_elec_init(const RpArgs* __restrict__ Ap, const ModelRegs& m, R rc, int64_t id,
                                             RpElec& E, double& ea, double& eb, double& ex) {
    const RpArgs& B = *rp_opaque(Ap);
    double ey = 0.0;
    ex = 0.0;
    if (B.sig0) {
        ea = B.sig0[id]; eb = B.sig0[B.ntraj + id]; ex = B.sig0[2 * B.ntraj + id]; ey = B.sig0[3 * B.ntraj + id];
    } else {
        ea = (B.initial_state == 0) ? 1.0 : 0.0; eb = 1.0 - ea;
    }
    E.est = B.state0 ? B.state0[id] : -1;
    R v11, v22, v12, d11, d22, d12;
    model1d_eval<R>(m, rc, v11, v22, v12, d11, d22, d12);
    double E0c, E1c;
    eigen2((double)v11, (double)v12, (double)v22, E0c, E1c, E.Uc, false);
    if (B.method == 1 && E.est < 0) {
        const double u = philox_uniform((uint64_t)id, 0u, STREAM_STATE, B.key) * (ea + eb);
        E.est = (u < ea) ? 0 : 1;
    }
    E.tr = ea + eb;
    E.s.q = ea - eb; E.s.x = ex; E.s.y = ey;
}

// The density over [t, tn] under the centroid at the step end (rc, vc): the interpolation buffer's
// new end is dE = E0 - E1 and w = d01 vc at the centroid (the tracked eigenvectors, bcb_electronics.jl:88-92).
// A: RpArgs, or RpHot (its fields in scalar registers)
template <typename R, class AT>
__device__ __forceinline__ void rp_elec_propagate(const AT& A, const ModelRegs& m, R rc, double vc, double t,
                                                  double tn, RpElec& E) {
    R v11, v22, v12, d11, d22, d12;
    model1d_eval<R>(m, rc, v11, v22, v12, d11, d22, d12);
    double E0c, E1c, G00c, G11c, G01c;
    eigen2((double)v11, (double)v12, (double)v22, E0c, E1c, E.Uc, true);
    rotate_derivative(E.Uc, (double)d11, (double)d22, (double)d12, G00c, G11c, G01c);
    E.gap = E1c - E0c;
    E.d01 = G01c / E.gap;
    const double w = E.d01 * vc;
    const double dEn = E0c - E1c;
    const double inv = 1.0 / (tn - E.tc);
    Interp2 Pi;
    Pi.dE0 = E.dEc; Pi.dE1 = (dEn - E.dEc) * inv; Pi.w0 = E.wc; Pi.w1 = (w - E.wc) * inv;
    propagate2(Pi, E.s, t, E.tc, A.dt, A.nsub);
    E.dEc = dEn; E.wc = w; E.tc = tn;
}

// The centroid hop at the end of step `step` (rpsh.jl:30-50): returns the centroid mode's momentum
// change (sqrt(P) times the centroid velocity change), zero unless a hop is accepted or reflected.
template <class AT>
__device__ __forceinline__ double rp_elec_hop(const AT& A, double vc, double sqP, int64_t id, int step,
                                              RpElec& E) {
    double dp = 0.0;
    const double sss = 0.5 * (E.tr + ((E.est == 0) ? E.s.q : -E.s.q));
    const double ratio = E.s.x / sss;
    const double dsm = (E.est == 0) ? E.d01 : -E.d01;
    double pr = 2.0 * vc * ratio * dsm * A.dt;
    if (pr > 0.0) {
        pr = fmin(pr, 1.0);
        const double rnd = philox_uniform((uint64_t)id, (uint32_t)step, STREAM_HOP, A.key);
        if (pr > rnd) {
            const double d = -dsm;  // coupling[new, old]
            const double dE = (E.est == 0) ? E.gap : -E.gap;
            bool ok = true;
            if (A.rescaling != 2) {
                const double aq = d * d / A.mass / 2.0;
                const double bq = d * vc;
                const double disc = bq * bq - 4.0 * aq * dE;
                if (disc < 0.0) {
                    ok = false;
                    if (A.rescaling == 1) {  // all beads: only the centroid mode moves
                        const double dn = d / fabs(d);
                        dp = -(sqP * 2.0 * (vc * dn) * dn);
                    }
                } else {
                    const double root = sqrt(disc);
                    const double gam = (bq < 0.0) ? (bq + root) / (2.0 * aq) : (bq - root) / (2.0 * aq);
                    dp = -(sqP * gam * d / A.mass);
                }
            }
            if (ok) { E.est = 1 - E.est; ++E.nh; }
        }
    }
    return dp;
}

// The launch arguments the centroid electronics read every step, in scalar registers.
struct RpHot {
    double dt, mass;
    uint64_t key;
    int32_t nsub, rescaling;
};

__device__ __forceinline__ RpHot rp_hot(const RpArgs& A) {
    RpHot H;
    H.dt = uniform_sgpr(A.dt); H.mass = uniform_sgpr(A.mass); H.key = uniform_sgpr(A.key);
    H.nsub = uniform_sgpr(A.nsub); H.rescaling = uniform_sgpr(A.rescaling);
    return H;
}

// Electronic save slot of ring polymer id: adiabatic populations (the active state's indicator for
// RPSH, the density's diagonal for RP-Ehrenfest), diabatic populations through the centroid's
// eigenvectors, and the active state.
__device__ __forceinline__ void rp_save_elec(const RpArgs& B, int slot, int64_t id, const RpElec& E) {
    const int64_t n = B.ntraj;
    double pa0, pa1;
    if (B.method == 1) { pa0 = (E.est == 0) ? 1.0 : 0.0; pa1 = 1.0 - pa0; }
    else { pa0 = 0.5 * (E.tr + E.s.q); pa1 = 0.5 * (E.tr - E.s.q); }
    const double* U = E.Uc;
    B.save_pop_ad[((int64_t)slot * 2) * n + id] = pa0;
    B.save_pop_ad[((int64_t)slot * 2 + 1) * n + id] = pa1;
    B.save_pop_di[((int64_t)slot * 2) * n + id] = U[0] * U[0] * pa0 + U[1] * U[1] * pa1 + 2.0 * U[0] * U[1] * E.s.x;
    B.save_pop_di[((int64_t)slot * 2 + 1) * n + id] = U[2] * U[2] * pa0 + U[3] * U[3] * pa1 + 2.0 * U[2] * U[3] * E.s.x;
    B.save_state[(int64_t)slot * n + id] = E.est;
}

// Final density, state, step and hop counts of ring polymer id.
__device__ __forceinline__ void rp_store_elec(const RpArgs& B, int64_t id, const RpElec& E) {
    const int64_t n = B.ntraj;
    B.sig[id] = 0.5 * (E.tr + E.s.q);
    B.sig[n + id] = 0.5 * (E.tr - E.s.q);
    B.sig[2 * n + id] = E.s.x;
    B.sig[3 * n + id] = E.s.y;
    B.state[id] = E.est;
    B.steps[id] = (int32_t)B.nsteps;
    B.nhops[id] = E.nh;
}

// Final state of an RPMD ring polymer (one electronic state).
__device__ __forceinline__ void rp_store_single(const RpArgs& B, int64_t id) {
    const int64_t n = B.ntraj;
    B.sig[id] = 1.0; B.sig[n + id] = 0.0; B.sig[2 * n + id] = 0.0; B.sig[3 * n + id] = 0.0;
    B.state[id] = 0; B.steps[id] = (int32_t)B.nsteps; B.nhops[id] = 0;
}

template <int P, typename R, bool E>
__device__ __forceinline__ void rp_body(const RpArgs* __restrict__ Ap) {
    constexpr int NB = RpShape<P, R, E>::block, NR = RpShape<P, R, E>::nr;
    __shared__ R lds[NB];
    // hand-over between the mode-0 lanes and the electronics lanes, one entry per ring polymer
    __shared__ R x_rc[NR];                          // centroid position (bead precision)
    __shared__ double x_vc[NR], x_a[NR], x_b[NR], x_sx[NR], x_dp[NR];
    __shared__ int x_st[NR];
    const RpArgs& A = *Ap;
    const int l = threadIdx.x % P;
    const int j = threadIdx.x / P;                  // ring polymer within the workgroup
    const int gbase = threadIdx.x - l;
    const int64_t rp = (int64_t)blockIdx.x * NR + j;
    const bool valid = rp < A.ntraj;
    const int64_t id = valid ? rp : 0;
    const int64_t n = A.ntraj;
    // row l of the Hartley matrix
    R H[P];
    {
        const double s = 1.0 / sqrt((double)P);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            double sn, cs;
            sincospi(2.0 * (double)((l * k) % P) / (double)P, &sn, &cs);
            H[k] = (R)((cs + sn) * s);
        }
    }
    const double sqP = sqrt((double)P);
    const ModelRegs m = model_regs(A.model);
    // Cayley map of mode l (ring_polymer.jl:71-82, half = false)
    double c11, c12, c21;
    {
        const double w = 2.0 * A.wn * sinpi((double)l / (double)P);
        const double u = A.dt * A.dt * w * w / 4.0;
        const double s = 1.0 / (1.0 + u);
        c11 = (1.0 - u) * s; c12 = A.dt * s; c21 = -A.dt * w * w * s;
    }
    const R C11 = (R)c11, C12 = (R)c12, C21 = (R)c21;
    // initial beads -> modes
    R rb, vb;
    {
        const RpArgs& B = *rp_opaque(Ap);
        rb = valid ? (R)B.r0[(int64_t)l * n + id] : (R)0;
        vb = valid ? (R)B.v0[(int64_t)l * n + id] : (R)0;
    }
    R q = hartley<P, R>(rb, H, lds, gbase);
    R p = hartley<P, R>(vb, H, lds, gbase);
    // the electronics kernel also serves RPMD above 16 beads (electronics off at run time): the
    // compiler allocates the RPMD-only kernel at those sizes badly (256 VGPRs and spills)
    const bool elec = E && A.method != 0;
    // electronics lanes: ring polymer e on lane e of wavefront blockIdx.x % (NB / 64) (rotating the
    // wavefront, and with it the SIMD, from one workgroup to the next)
    const int e = (int)threadIdx.x - 64 * (int)(blockIdx.x % (unsigned)(NB / 64));
    const bool eln = elec && e >= 0 && e < NR;
    const int64_t erp = (int64_t)blockIdx.x * NR + e;
    const bool evalid = eln && erp < n;
    const int64_t eid = evalid ? erp : 0;
    RpElec El;
    El.tr = 1.0; El.s = {1.0, 0.0, 0.0}; El.est = 0; El.nh = 0;
    El.Uc[0] = 1.0; El.Uc[1] = 0.0; El.Uc[2] = 0.0; El.Uc[3] = 1.0;
    El.dEc = 0.0; El.wc = 0.0; El.tc = 0.0; El.d01 = 0.0; El.gap = 0.0;
    double Ub[4] = {1.0, 0.0, 0.0, 1.0};
    double t = A.t0;
    // active state / density of the step start as the bead lanes see them
    int st = 0;
    double a = 1.0, b = 0.0, sx = 0.0;
    if (elec) {
        if (l == 0) x_rc[j] = q / (R)sqP;
        __syncthreads();
        if (eln) {
            double ea = 1.0, eb = 0.0, ex = 0.0;
            if (evalid) rp_elec_init<R>(Ap, m, x_rc[e], eid, El, ea, eb, ex);
            x_st[e] = El.est; x_a[e] = ea; x_b[e] = eb; x_sx[e] = ex; x_dp[e] = 0.0;
        }
        __syncthreads();
        st = x_st[j]; a = x_a[j]; b = x_b[j]; sx = x_sx[j];
    }
    // initial accelerations in modes
    R acc = (R)(-bead_force<R, E>(A, m, rb, Ub, false, st, a, b, sx) / A.mass);
    R an = hartley<P, R>(acc, H, lds, gbase);
    // centroid slots come from the mode-0 lanes, electronic slots from the electronics lanes
    auto save_centroid = [&](int slot) {
        const RpArgs& B = *rp_opaque(Ap);
        if (slot >= B.nsave) return;
        if (valid && l == 0) B.save_centroid[(int64_t)slot * n + id] = (double)q / sqP;
    };
    auto save_elec = [&](int slot) {
        const RpArgs& B = *rp_opaque(Ap);
        if (slot >= B.nsave || !evalid) return;
        rp_save_elec(B, slot, eid, El);
    };
    save_centroid(0);
    save_elec(0);
    const R hdt = (R)(0.5 * A.dt);
#pragma unroll 1
    for (int step = 0; step < A.nsteps; ++step) {
        if (elec) {
            if (step > 0) {  // hand-over from the previous step's electronics
                st = x_st[j];
                a = x_a[j]; b = x_b[j]; sx = x_sx[j];
                if (l == 0) {
                    const double dp = x_dp[j];
                    if (dp != 0.0) p = (R)((double)p + dp);
                }
            } else {         // the initial density in its carried form (trace and population difference)
                const double trr = a + b, qq = a - b;
                a = 0.5 * (trr + qq); b = 0.5 * (trr - qq);
            }
        }
        p = p + hdt * an;                                        // step_B! (in modes)
        {                                                        // step_C! (mode l)
            const R qn = C11 * q + C12 * p;
            const R pn = C21 * q + C11 * p;
            q = qn; p = pn;
        }
        rb = hartley<P, R>(q, H, lds, gbase);                    // bead positions
        acc = (R)(-bead_force<R, E>(A, m, rb, Ub, true, st, a, b, sx) / A.mass);
        an = hartley<P, R>(acc, H, lds, gbase);
        p = p + hdt * an;                                        // step_B!
        const double tn = t + A.dt;
        if (elec) {
            if (l == 0) { x_rc[j] = q / (R)sqP; x_vc[j] = (double)p / sqP; }
            __syncthreads();
            if (eln) {
                double dp = 0.0;
                if (evalid) {
                    const double vc = x_vc[e];
                    rp_elec_propagate<R>(A, m, x_rc[e], vc, t, tn, El);
                    // the save precedes the hop callback (a save time holds the state before the step's
                    // callbacks; the hop's velocity change reaches the beads at the next step)
                    if (A.save_every > 0 && ((step + 1) % A.save_every) == 0) save_elec((step + 1) / A.save_every);
                    if (A.method == 1) dp = rp_elec_hop(A, vc, sqP, eid, step, El);
                }
                x_st[e] = El.est;
                x_a[e] = 0.5 * (El.tr + El.s.q); x_b[e] = 0.5 * (El.tr - El.s.q); x_sx[e] = El.s.x;
                x_dp[e] = dp;
            }
        }
        t = tn;
        if (A.save_every > 0 && ((step + 1) % A.save_every) == 0) save_centroid((step + 1) / A.save_every);
        if (elec) __syncthreads();
    }
    if (elec && A.nsteps > 0 && l == 0) {  // the last step's hop
        const double dp = x_dp[j];
        if (dp != 0.0) p = (R)((double)p + dp);
    }
    rb = hartley<P, R>(q, H, lds, gbase);
    vb = hartley<P, R>(p, H, lds, gbase);
    const RpArgs& B = *rp_opaque(Ap);
    if (valid) {
        B.r[(int64_t)l * n + id] = (double)rb;
        B.v[(int64_t)l * n + id] = (double)vb;
        if (!elec && l == 0) rp_store_single(B, id);  // RPMD: no electronic state
    }
    if (evalid) rp_store_elec(B, eid, El);
}

// ---------------------------------------------------------------------------
// One ring polymer per lane (RPMD and RPSH; up to 64 beads in FP32, 32 in FP64).  The lane keeps all
// P normal modes in registers and applies the Hartley transform (orthogonal and its own inverse) in
// registers, so nothing passes through LDS in the step, no workgroup barrier is needed, and every
// lane runs its own ring polymer's centroid electronics: in the bead-per-lane layout P - 1 of every P
// lanes wait while the FP64 electronics run.  From 8 beads up the transform is the radix-2 fast
// Hartley transform (P log2 P butterflies instead of the P^2 products of the matrix form).
// ---------------------------------------------------------------------------
template <int P, typename R>
__device__ __forceinline__ void hartley_regs(const R c[P], const R x[P], R y[P]) {
#pragma unroll
    for (int j = 0; j < P; ++j) {
        R acc = (R)0;
#pragma unroll
        for (int k = 0; k < P; ++k) acc += c[(j * k) % P] * x[k];
        y[j] = acc;
    }
}

constexpr int bit_reverse(int i, int P) {
    int r = 0;
    for (int b = 1; b < P; b <<= 1) { r = (r << 1) | (i & 1); i >>= 1; }
    return r;
}

// x <- H x with H[j][k] = cas(2 pi jk / P) / sqrt(P), P a power of two: radix-2 decimation in time
// (Bracewell).  With E and O the transforms of the even and odd halves (length h), the length-2h
// transform is H_k = E_k + cos(pi k/h) O_k + sin(pi k/h) O_{h-k} and H_{k+h} = E_k - (the same),
// indices mod h; modes k and h - k are combined together so the butterfly runs in place.  The input
// permutation and the twiddles are compile-time constants.
template <int P, typename R>
__device__ __forceinline__ void fht_regs(R x[P]) {
    R t[P];
#pragma unroll
    for (int i = 0; i < P; ++i) t[i] = x[bit_reverse(i, P)];
#pragma unroll
    for (int h = 1; h < P; h <<= 1) {
#pragma unroll
        for (int b = 0; b < P; b += 2 * h) {
            {
                const R e = t[b], o = t[b + h];
                t[b] = e + o; t[b + h] = e - o;
            }
            if (h >= 2) {   // k = h/2: cos = 0, sin = 1, O_{h-k} = O_k
                const int k = h / 2;
                const R e = t[b + k], o = t[b + h + k];
                t[b + k] = e + o; t[b + h + k] = e - o;
            }
#pragma unroll
            for (int k = 1; k < h / 2; ++k) {
                const int kk = h - k;
                double sn, cs;
                sincospi((double)k / (double)h, &sn, &cs);
                const R c = (R)cs, s = (R)sn;
                const R ek = t[b + k], ekk = t[b + kk], ok = t[b + h + k], okk = t[b + h + kk];
                const R t1 = c * ok + s * okk, t2 = s * ok - c * okk;
                t[b + k] = ek + t1; t[b + k + h] = ek - t1;
                t[b + kk] = ekk + t2; t[b + kk + h] = ekk - t2;
            }
        }
    }
    const R sc = (R)(1.0 / sqrt((double)P));
#pragma unroll
    for (int i = 0; i < P; ++i) x[i] = t[i] * sc;
}

// x <- H x in place: the fast transform from 8 beads, the matrix form below that.
template <int P, typename R>
__device__ __forceinline__ void hartley_inplace(R x[P]) {
    if constexpr (P >= 8) {
        fht_regs<P, R>(x);
    } else {
        R c[P];
        const double s = 1.0 / sqrt((double)P);
#pragma unroll
        for (int k = 0; k < P; ++k) {
            double sn, cs;
            sincospi(2.0 * (double)k / (double)P, &sn, &cs);
            c[k] = (R)((cs + sn) * s);
        }
        R y[P];
        hartley_regs<P, R>(c, x, y);
#pragma unroll
        for (int k = 0; k < P; ++k) x[k] = y[k];
    }
}

constexpr int kRpLaneBlock = 64;   // one wavefront per workgroup: small ensembles spread over every CU

// Accelerations of P beads from the gradients of the surface each moves on (surface_gradient), under
// one model dispatch so the beads form straight-line code; in place (x: bead positions in,
// accelerations out).  FP32 beads take 1/rad from the hardware reciprocal square root and multiply by
// the FP32 inverse mass; FP64 beads keep the rounded square root and the divisions.
template <int P, typename R, bool SH>
__device__ __forceinline__ void bead_accelerations(const ModelRegs& m, R x[P], int st, double mass) {
    model1d_dispatch(m, [&](auto kind) {
        constexpr int K = decltype(kind)::value;
        const R inv_m = (R)(1.0 / mass);
#pragma unroll
        for (int l = 0; l < P; ++l) {
            R v11, v22, v12, d11, d22, d12;
            model1d_kind<K, R>(m, x[l], v11, v22, v12, d11, d22, d12);
            R g = d11;
            if constexpr (SH) {
                const R h = (R)0.5 * (v11 - v22), dh = (R)0.5 * (d11 - d22), dm = (R)0.5 * (d11 + d22);
                const R r2 = h * h + v12 * v12;
                const R num = h * dh + v12 * d12;
                R c;
                if constexpr (sizeof(R) == 4) {
                    c = (r2 > 0.0f) ? num * __builtin_amdgcn_rsqf(r2) : 0.0f;
                } else {
                    const R rad = sqrt(r2);
                    c = (rad > (R)0) ? num / rad : (R)0;
                }
                g = (st == 0) ? dm - c : dm + c;
            }
            if constexpr (sizeof(R) == 4) x[l] = -g * inv_m;
            else x[l] = -g / mass;
        }
    });
}

// Cayley map of mode k (ring_polymer.jl:71-82, half = false) in the bead precision.
template <typename R>
__device__ __forceinline__ void cayley_mode(const RpArgs& A, int k, int P, R& c11, R& c12, R& c21) {
    const double w = 2.0 * A.wn * sinpi((double)k / (double)P);
    const double u = A.dt * A.dt * w * w / 4.0;
    const double si = 1.0 / (1.0 + u);
    c11 = (R)((1.0 - u) * si); c12 = (R)(A.dt * si); c21 = (R)(-A.dt * w * w * si);
}

// The Cayley table of the argument block (the first 3 P entries in the bead precision).
template <int P, typename R>
__global__ void rp_cayley_fill(RpArgs* A) {
    const int k = threadIdx.x;
    if (k < P) {
        R* t = reinterpret_cast<R*>(A->cay);
        cayley_mode<R>(*A, k, P, t[k], t[P + k], t[2 * P + k]);
    }
}

template <typename R>
using RpConstTab = __attribute__((address_space(4))) const R*;

template <typename R>
__device__ __forceinline__ RpConstTab<R> rp_cay_table(const RpArgs* Ap) {
    const uint64_t a = (uint64_t)reinterpret_cast<const R*>(Ap->cay);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a), hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    return (RpConstTab<R>)(((uint64_t)hi << 32) | lo);
}

// The table pointer made opaque where it is used, so its entries are scalar-loaded there each step
// (operands straight from SGPRs) instead of being hoisted into registers for the whole loop.
template <typename R>
__device__ __forceinline__ RpConstTab<R> rp_cay_here(RpConstTab<R> p) {
    asm volatile("" : "+s"(p));
    return p;
}

// Whether the lane kernel of P beads reads its Cayley coefficients from the argument block's table.
constexpr bool rp_lane_table(int P, bool fp64) { return P >= 32 || (fp64 && P >= 16); }

template <int P, typename R, bool SH>
__device__ __forceinline__ void rp_lane_body(const RpArgs* __restrict__ Ap) {
    // From 32 beads (16 in FP64) the 3 P Cayley coefficients are scalar loads from the argument
    // block's table (written by rp_cayley_fill) instead of registers, which hold the 2 P modes and the
    // P-bead work array.
    constexpr bool kTab = rp_lane_table(P, sizeof(R) == 8);
    const RpArgs& A = *Ap;
    const int64_t n = A.ntraj;
    const int64_t id = (int64_t)blockIdx.x * kRpLaneBlock + threadIdx.x;
    if (id >= n) return;   // no barriers below
    constexpr bool sh = SH;
    R C11[kTab ? 1 : P], C12[kTab ? 1 : P], C21[kTab ? 1 : P];
    if constexpr (!kTab) {
#pragma unroll
        for (int k = 0; k < P; ++k) cayley_mode<R>(A, k, P, C11[k], C12[k], C21[k]);
    }
    const double sqP = sqrt((double)P);
    const ModelRegs m = model_regs(A.model);
    R q[P], p[P], w[P];   // modes, and the work array (bead positions -> accelerations -> their modes)
    RpElec El;
    El.tr = 1.0; El.s = {1.0, 0.0, 0.0}; El.est = 0; El.nh = 0;
    El.Uc[0] = 1.0; El.Uc[1] = 0.0; El.Uc[2] = 0.0; El.Uc[3] = 1.0;
    El.dEc = 0.0; El.wc = 0.0; El.tc = 0.0; El.d01 = 0.0; El.gap = 0.0;
    {
        const RpArgs& B = *rp_opaque(Ap);
#pragma unroll
        for (int l = 0; l < P; ++l) {
            q[l] = (R)B.r0[(int64_t)l * n + id];
            p[l] = (R)B.v0[(int64_t)l * n + id];
            w[l] = q[l];
        }
        hartley_inplace<P, R>(q);
        hartley_inplace<P, R>(p);
        if (sh) {
            double ea, eb, ex;
            rp_elec_init<R>(Ap, m, q[0] / (R)sqP, id, El, ea, eb, ex);
        }
        bead_accelerations<P, R, SH>(m, w, El.est, A.mass);   // at the initial beads
        hartley_inplace<P, R>(w);
    }
    auto save = [&](int slot) {   // centroid and electronics of save slot `slot`
        const RpArgs& B = *rp_opaque(Ap);
        if (slot >= B.nsave) return;
        B.save_centroid[(int64_t)slot * n + id] = (double)q[0] / sqP;
        if (sh) rp_save_elec(B, slot, id, El);
    };
    save(0);
    // the step loop's launch arguments in scalar registers (no memory access per step for them)
    const int64_t nsteps = uniform_sgpr(A.nsteps);
    const int32_t save_every = uniform_sgpr(A.save_every);
    const RpHot Hs = rp_hot(A);
    const R hdt = (R)(0.5 * Hs.dt);
    double t = A.t0;
    const RpConstTab<R> cay0 = kTab ? rp_cay_table<R>(Ap) : nullptr;
    // The accelerations of a step serve both its closing half kick and the next step's opening one,
    // applied together at the step's end (each rounded as before) so they are not held across steps.
    if (nsteps > 0) {
#pragma unroll
        for (int k = 0; k < P; ++k) p[k] = p[k] + hdt * w[k];  // step_B! of the first step
    }
#pragma unroll 1
    for (int step = 0; step < nsteps; ++step) {
        const RpConstTab<R> cay = kTab ? rp_cay_here<R>(cay0) : cay0;
#pragma unroll
        for (int k = 0; k < P; ++k) {                            // step_C! (mode k)
            R c11, c12, c21;
            if constexpr (kTab) { c11 = cay[k]; c12 = cay[P + k]; c21 = cay[2 * P + k]; }
            else { c11 = C11[k]; c12 = C12[k]; c21 = C21[k]; }
            const R qn = c11 * q[k] + c12 * p[k];
            p[k] = c21 * q[k] + c11 * p[k];
            q[k] = qn;
        }
#pragma unroll
        for (int k = 0; k < P; ++k) w[k] = q[k];
        hartley_inplace<P, R>(w);                                // bead positions
        bead_accelerations<P, R, SH>(m, w, El.est, A.mass);
        hartley_inplace<P, R>(w);                                // accelerations in modes
#pragma unroll
        for (int k = 0; k < P; ++k) p[k] = p[k] + hdt * w[k];   // step_B!
        // the next step's opening kick of modes 1.. now (the hop changes only the centroid mode), so
        // the accelerations are dead during the electronics
        const R w0 = w[0];
        const bool more = step + 1 < nsteps;
        if (more) {
#pragma unroll
            for (int k = 1; k < P; ++k) p[k] = p[k] + hdt * w[k];
        }
        const double tn = t + Hs.dt;
        const bool save_now = save_every > 0 && ((step + 1) % save_every) == 0;
        double dp = 0.0;
        if (sh) {
            const double vc = (double)p[0] / sqP;
            rp_elec_propagate<R>(Hs, m, q[0] / (R)sqP, vc, t, tn, El);
            // the save precedes the hop callback (a save time holds the state before the step's callbacks)
            if (save_now) save((step + 1) / save_every);
            dp = rp_elec_hop(Hs, vc, sqP, id, step, El);
        } else if (save_now) {
            save((step + 1) / save_every);
        }
        if (dp != 0.0) p[0] = (R)((double)p[0] + dp);
        t = tn;
        if (more) p[0] = p[0] + hdt * w0;                        // the next step's opening kick, centroid mode
    }
    const RpArgs& B = *rp_opaque(Ap);
    hartley_inplace<P, R>(q);
    hartley_inplace<P, R>(p);
#pragma unroll
    for (int l = 0; l < P; ++l) {
        B.r[(int64_t)l * n + id] = (double)q[l];
        B.v[(int64_t)l * n + id] = (double)p[l];
    }
    if (sh) rp_store_elec(B, id, El);
    else rp_store_single(B, id);
}

#define NQCD_RPMD_KERNELS(P)                                                                                \
    __global__ __launch_bounds__((RpShape<P, float, false>::block)) void rp_kernel_f32_##P(const RpArgs* __restrict__ Ap) { \
        rp_body<P, float, false>(Ap);                                                                       \
    }                                                                                                       \
    __global__ __launch_bounds__((RpShape<P, double, false>::block)) void rp_kernel_f64_##P(const RpArgs* __restrict__ Ap) { \
        rp_body<P, double, false>(Ap);                                                                      \
    }
#define NQCD_RP_KERNELS(P)                                                                                  \
    __global__ __launch_bounds__((RpShape<P, float, true>::block)) void rp_elec_f32_##P(const RpArgs* __restrict__ Ap) { \
        rp_body<P, float, true>(Ap);                                                                        \
    }                                                                                                       \
    __global__ __launch_bounds__((RpShape<P, double, true>::block)) void rp_elec_f64_##P(const RpArgs* __restrict__ Ap) { \
        rp_body<P, double, true>(Ap);                                                                       \
    }
NQCD_RPMD_KERNELS(1)
NQCD_RPMD_KERNELS(2)
NQCD_RPMD_KERNELS(4)
NQCD_RPMD_KERNELS(8)
NQCD_RPMD_KERNELS(16)
NQCD_RP_KERNELS(1)
NQCD_RP_KERNELS(2)
NQCD_RP_KERNELS(4)
NQCD_RP_KERNELS(8)
NQCD_RP_KERNELS(16)
NQCD_RP_KERNELS(32)
NQCD_RP_KERNELS(64)
#undef NQCD_RP_KERNELS
#undef NQCD_RPMD_KERNELS

// rp_lane_*: RPMD; rp_lanesh_*: RPSH
// WSH: the RPSH kernel's minimum waves per SIMD (2 where that fits without spilling much: 1e5 ring
// polymers are 1.5 wavefronts per SIMD, so a one-wave kernel runs the rest as a second round).
#define NQCD_RP_LANE_KERNELS(P, R, SUFFIX, WSH)                                                             \
    __global__ __launch_bounds__(kRpLaneBlock) void rp_lane_##SUFFIX(const RpArgs* __restrict__ Ap) {      \
        rp_lane_body<P, R, false>(Ap);                                                                      \
    }                                                                                                       \
    __global__ __launch_bounds__(kRpLaneBlock) __attribute__((amdgpu_waves_per_eu(WSH)))                    \
    void rp_lanesh_##SUFFIX(const RpArgs* __restrict__ Ap) {                                               \
        rp_lane_body<P, R, true>(Ap);                                                                       \
    }
NQCD_RP_LANE_KERNELS(1, float, f32_1, 1)
NQCD_RP_LANE_KERNELS(2, float, f32_2, 1)
NQCD_RP_LANE_KERNELS(4, float, f32_4, 1)
NQCD_RP_LANE_KERNELS(8, float, f32_8, 1)
NQCD_RP_LANE_KERNELS(16, float, f32_16, 1)
NQCD_RP_LANE_KERNELS(32, float, f32_32, 2)
NQCD_RP_LANE_KERNELS(64, float, f32_64, 1)
NQCD_RP_LANE_KERNELS(1, double, f64_1, 1)
NQCD_RP_LANE_KERNELS(2, double, f64_2, 1)
NQCD_RP_LANE_KERNELS(4, double, f64_4, 1)
NQCD_RP_LANE_KERNELS(8, double, f64_8, 1)
NQCD_RP_LANE_KERNELS(16, double, f64_16, 1)
NQCD_RP_LANE_KERNELS(32, double, f64_32, 1)
#undef NQCD_RP_LANE_KERNELS

// Thermal ring polymers in the harmonic approximation of a well at x0 with frequency w_well:
// mode k ~ N(sqrt(P) x0 delta_k0, 1/(beta_P m (w_k^2 + w_well^2))), velocity ~ N(0, 1/(beta_P m)),
// beta_P = beta/P; counter (rp, mode, stream 0) gives (z0 position, z1 velocity).
template <int P>
__global__ __launch_bounds__(kRpBlock) void rp_sample(int64_t ntraj, double x0, double wwell, double beta, double mass,
                                                      double wn, uint64_t key, double* r0, double* v0) {
    __shared__ double lds[kRpBlock];
    const int l = threadIdx.x % P;
    const int gbase = threadIdx.x - l;
    const int64_t rp = (int64_t)blockIdx.x * (kRpBlock / P) + threadIdx.x / P;
    const bool valid = rp < ntraj;
    double H[P];
    const double s = 1.0 / sqrt((double)P);
#pragma unroll
    for (int k = 0; k < P; ++k) {
        double sn, cs;
        sincospi(2.0 * (double)((l * k) % P) / (double)P, &sn, &cs);
        H[k] = (cs + sn) * s;
    }
    double z0 = 0.0, z1 = 0.0;
    if (valid) philox_normal2((uint64_t)rp, (uint32_t)l, STREAM_INIT, key, z0, z1);
    const double w = 2.0 * wn * sinpi((double)l / (double)P);
    const double betaP = beta / P;
    const double q = ((l == 0) ? sqrt((double)P) * x0 : 0.0) + z0 / sqrt(betaP * mass * (w * w + wwell * wwell));
    const double pv = z1 / sqrt(betaP * mass);
    const double rb = hartley<P, double>(q, H, lds, gbase);
    const double vb = hartley<P, double>(pv, H, lds, gbase);
    if (valid) { r0[(int64_t)l * ntraj + rp] = rb; v0[(int64_t)l * ntraj + rp] = vb; }
}

// Path-integral Monte Carlo (MetropolisHastings.jl:58-82,131-148,170-258): one Metropolis-Hastings
// chain per ring polymer, P lanes per chain (lane l holds bead l and Hartley mode l), FP64.  Each
// step: a centroid move (every bead shifted by (u - 1/2) Delta / sqrt(P)), pass_length = P / L mode
// passes (L draws of a mode k in 1..P-1 resampled from N(0, sigma_k), last draw of a mode winning),
// each followed by the Metropolis test min(1, exp(-(E' - E)/w_n)) > u on the bead-summed potential
// (element [1] of the model matrix), then the reference's second test of the last proposal.  The
// draws of a chain come from counters shared by its lanes (Philox stream 4), so every lane takes
// the same decisions; the potential sum is a butterfly over the chain's lanes, whose result carries
// the same bits in all of them.  Segment draw s is made by lane s and handed to the lane of its mode
// through LDS (the highest s wins, as the reference's sequential resampling does).
struct PimcArgs {
    int64_t ntraj, steps;
    int32_t L, width, reassess, pad;
    double x0, delta, mass, wn;
    uint64_t key;
    ModelSlot model;
};

template <int P>
__device__ __forceinline__ double chain_sum(double v) {
#pragma unroll
    for (int o = P / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ double pimc_u(uint64_t rp, int64_t it, uint32_t sub, uint64_t key, int half) {
    uint32_t c0 = (uint32_t)rp, c1 = (uint32_t)(rp >> 32), c2 = (uint32_t)it, c3 = 4u + 16u * sub;
    philox4x32_10(c0, c1, c2, c3, (uint32_t)key, (uint32_t)(key >> 32));
    return half ? u53(c2, c3) : u53(c0, c1);
}

template <int P>
__global__ __launch_bounds__(kRpBlock) void rp_pimc(const PimcArgs A, double* r0, double* v0) {
    __shared__ double lds[kRpBlock];
    __shared__ double zbuf[kRpBlock];
    __shared__ int last[kRpBlock];
    const int l = threadIdx.x % P;
    const int gbase = threadIdx.x - l;
    const int64_t rp = (int64_t)blockIdx.x * (kRpBlock / P) + threadIdx.x / P;
    const bool valid = rp < A.ntraj;
    const uint64_t cid = valid ? (uint64_t)rp : 0ull;
    double H[P];
    const double s = 1.0 / sqrt((double)P);
#pragma unroll
    for (int k = 0; k < P; ++k) {
        double sn, cs;
        sincospi(2.0 * (double)((l * k) % P) / (double)P, &sn, &cs);
        H[k] = (cs + sn) * s;
    }
    const ModelRegs m = model_regs(A.model);
    auto pot = [&](double x) {
        double v11, v22, v12, d11, d22, d12;
        model1d_eval<double>(m, x, v11, v22, v12, d11, d22, d12);
        return v11;
    };
    double sig = 0.0;
    if (l > 0) {
        const double wk = 2.0 * A.wn * sinpi((double)l / (double)P);
        sig = A.width == 0 ? sqrt(A.wn / A.mass) / (0.5 * wk * wk * A.mass) : sqrt(A.wn / A.mass) / wk;
    }
    const int L = A.L, J = P / A.L;
    double x = A.x0, xp = A.x0;
    double Ei = chain_sum<P>(pot(x)), Ep = Ei;
    last[threadIdx.x] = -1;
    auto assess = [&](double u) {
        Ep = chain_sum<P>(pot(xp));
        if (fmin(1.0, exp(-(Ep - Ei) / A.wn)) > u) { Ei = Ep; x = xp; }
    };
    for (int64_t it = 0; it < A.steps; ++it) {
        {   // centroid move (MetropolisHastings.jl:208-229)
            uint32_t c0 = (uint32_t)cid, c1 = (uint32_t)(cid >> 32), c2 = (uint32_t)it, c3 = 4u;
            philox4x32_10(c0, c1, c2, c3, (uint32_t)A.key, (uint32_t)(A.key >> 32));
            const double shift = (u53(c0, c1) - 0.5) * A.delta / sqrt(1.0);
            xp = x + shift * H[0];
            assess(u53(c2, c3));
        }
        if (P > 1) {
            for (int j = 0; j < J; ++j) {   // normal-mode moves (:232-248)
                double y = hartley<P, double>(x, H, lds, gbase);
                if (l < L) {
                    uint32_t c0 = (uint32_t)cid, c1 = (uint32_t)(cid >> 32), c2 = (uint32_t)it;
                    uint32_t c3 = 4u + 16u * (uint32_t)(1 + j * (L + 1) + l);
                    philox4x32_10(c0, c1, c2, c3, (uint32_t)A.key, (uint32_t)(A.key >> 32));
                    const int k = 1 + (int)(((uint64_t)c0 * (uint64_t)(P - 1)) >> 32);
                    zbuf[threadIdx.x] = sqrt(-2.0 * log(1.0 - u53(c1, c2))) * cos(6.283185307179586 * ((double)c3 * 0x1.0p-32));
                    atomicMax(&last[gbase + k], l);
                }
                __syncthreads();
                const int sl = last[threadIdx.x];
                if (sl >= 0) y = sig * zbuf[gbase + sl];
                __syncthreads();
                last[threadIdx.x] = -1;
                xp = hartley<P, double>(y, H, lds, gbase);
                assess(pimc_u(cid, it, (uint32_t)(1 + j * (L + 1) + L), A.key, 0));
            }
        }
        if (A.reassess) {   // the last proposal assessed once more (:194)
            const double u = pimc_u(cid, it, (uint32_t)(1 + J * (L + 1)), A.key, 0);
            if (fmin(1.0, exp(-(Ep - Ei) / A.wn)) > u) { Ei = Ep; x = xp; }
        }
    }
    double z0, z1;
    philox_normal2(cid, (uint32_t)l, STREAM_INIT, A.key, z0, z1);
    if (valid) {
        r0[(int64_t)l * A.ntraj + rp] = x;
        v0[(int64_t)l * A.ntraj + rp] = sqrt(A.wn / A.mass) * z1;
    }
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
typedef void (*RpKernel)(const RpArgs*);

static bool rp_supported(int P) { return P == 1 || P == 2 || P == 4 || P == 8 || P == 16 || P == 32 || P == 64; }

// Workgroup size and ring polymers per workgroup of the (P, precision, electronics) kernel.
static void rp_shape(int P, int precision, bool elec, int& block, int& nr) {
#define NQCD_CASE(PP)                                                                                         \
    case PP:                                                                                                  \
        if (elec) {                                                                                           \
            block = precision == 0 ? (RpShape<PP, double, true>::block) : (RpShape<PP, float, true>::block);  \
            nr = precision == 0 ? (RpShape<PP, double, true>::nr) : (RpShape<PP, float, true>::nr);           \
        } else {                                                                                              \
            block = precision == 0 ? (RpShape<PP, double, false>::block) : (RpShape<PP, float, false>::block); \
            nr = precision == 0 ? (RpShape<PP, double, false>::nr) : (RpShape<PP, float, false>::nr);         \
        }                                                                                                     \
        return;
    switch (P) {
        NQCD_CASE(1) NQCD_CASE(2) NQCD_CASE(4) NQCD_CASE(8) NQCD_CASE(16) NQCD_CASE(32) NQCD_CASE(64)
    default: block = 0; nr = 0;
    }
#undef NQCD_CASE
}

// elec selects the electronics kernel; the RPMD-only kernels exist up to 16 beads.
static bool rp_needs_elec(int method, int P) { return method != NQCD_METHOD_RPMD || P > 16; }

// The one-ring-polymer-per-lane kernels: RPMD and RPSH up to 64 beads in FP32 and 32 in FP64 (the
// largest hold one wave per SIMD; their few spills are outside the step loop).
static bool rp_lane_exists(int method, int P, int precision) {
    if (method != NQCD_METHOD_RPMD && method != NQCD_METHOD_RPSH) return false;
    return P <= (precision == 0 ? 32 : 64);
}

static hipError_t rp_fill_cayley(int P, int precision, hipStream_t s, RpArgs* Adev) {
    if (!rp_lane_table(P, precision == 0)) return hipSuccess;
    switch (P * (precision == 0 ? -1 : 1)) {
    case 32: hipLaunchKernelGGL((rp_cayley_fill<32, float>), dim3(1), dim3(64), 0, s, Adev); break;
    case 64: hipLaunchKernelGGL((rp_cayley_fill<64, float>), dim3(1), dim3(64), 0, s, Adev); break;
    case -16: hipLaunchKernelGGL((rp_cayley_fill<16, double>), dim3(1), dim3(64), 0, s, Adev); break;
    case -32: hipLaunchKernelGGL((rp_cayley_fill<32, double>), dim3(1), dim3(64), 0, s, Adev); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// Layout of a run: one ring polymer per lane where that kernel exists, P lanes per ring polymer
// otherwise.  NQCD_MI355X_RP_LAYOUT=lane or bead overrides the choice where
// both kernels exist.
static bool rp_use_lane(int method, int P, int precision) {
    if (!rp_lane_exists(method, P, precision)) return false;
    if (const char* env = getenv("NQCD_MI355X_RP_LAYOUT")) {
        if (!strcmp(env, "bead")) return false;
    }
    return true;
}

static hipError_t rp_launch_lane(int P, int precision, bool sh, int grid, hipStream_t s, const RpArgs* Adev) {
    const dim3 g(grid), b(kRpLaneBlock);
#define NQCD_LANE_CASE(PP, SUFFIX)                                                                            \
    case PP:                                                                                                  \
        if (sh) hipLaunchKernelGGL(rp_lanesh_##SUFFIX, g, b, 0, s, Adev);                                     \
        else hipLaunchKernelGGL(rp_lane_##SUFFIX, g, b, 0, s, Adev);                                          \
        break;
    if (precision == 0) {
        switch (P) {
            NQCD_LANE_CASE(1, f64_1) NQCD_LANE_CASE(2, f64_2) NQCD_LANE_CASE(4, f64_4) NQCD_LANE_CASE(8, f64_8)
            NQCD_LANE_CASE(16, f64_16) NQCD_LANE_CASE(32, f64_32)
        default: return hipErrorInvalidValue;
        }
    } else {
        switch (P) {
            NQCD_LANE_CASE(1, f32_1) NQCD_LANE_CASE(2, f32_2) NQCD_LANE_CASE(4, f32_4) NQCD_LANE_CASE(8, f32_8)
            NQCD_LANE_CASE(16, f32_16) NQCD_LANE_CASE(32, f32_32) NQCD_LANE_CASE(64, f32_64)
        default: return hipErrorInvalidValue;
        }
    }
#undef NQCD_LANE_CASE
    return hipGetLastError();
}

static hipError_t rp_launch(int P, int precision, bool elec, int grid, hipStream_t s, const RpArgs* Adev) {
#define NQCD_ELEC_CASE(PP)                                                                                    \
    case PP:                                                                                                  \
        if (precision == 0)                                                                                   \
            hipLaunchKernelGGL(rp_elec_f64_##PP, dim3(grid), dim3((RpShape<PP, double, true>::block)), 0, s, Adev); \
        else hipLaunchKernelGGL(rp_elec_f32_##PP, dim3(grid), dim3((RpShape<PP, float, true>::block)), 0, s, Adev); \
        break;
#define NQCD_RPMD_CASE(PP)                                                                                    \
    case PP:                                                                                                  \
        if (precision == 0)                                                                                   \
            hipLaunchKernelGGL(rp_kernel_f64_##PP, dim3(grid), dim3((RpShape<PP, double, false>::block)), 0, s, Adev); \
        else hipLaunchKernelGGL(rp_kernel_f32_##PP, dim3(grid), dim3((RpShape<PP, float, false>::block)), 0, s, Adev); \
        break;
    if (elec) {
        switch (P) {
            NQCD_ELEC_CASE(1) NQCD_ELEC_CASE(2) NQCD_ELEC_CASE(4) NQCD_ELEC_CASE(8) NQCD_ELEC_CASE(16)
            NQCD_ELEC_CASE(32) NQCD_ELEC_CASE(64)
        default: return hipErrorInvalidValue;
        }
    } else {
        switch (P) {
            NQCD_RPMD_CASE(1) NQCD_RPMD_CASE(2) NQCD_RPMD_CASE(4) NQCD_RPMD_CASE(8) NQCD_RPMD_CASE(16)
        default: return hipErrorInvalidValue;
        }
    }
#undef NQCD_ELEC_CASE
#undef NQCD_RPMD_CASE
    return hipGetLastError();
}

int32_t rp_check_problem(nqcd_handle_t* h) {
    const nqcd_problem_desc& d = h->prob;
    if (!rp_supported(d.nbeads)) return fail(h, NQCD_ERR_UNSUPPORTED, "ring polymers: nbeads must be 1, 2, 4, 8, 16, 32 or 64");
    if (d.ndofs * d.natoms != 1) return fail(h, NQCD_ERR_UNSUPPORTED, "ring polymers: one degree of freedom per bead");
    if (!(d.temperature > 0.0)) return fail(h, NQCD_ERR_ARG, "ring polymers need temperature > 0");
    if (d.method == NQCD_METHOD_RPMD && d.algorithm != NQCD_ALG_BCB)
        return fail(h, NQCD_ERR_UNSUPPORTED, "RingPolymerSimulation{Classical} integrates with BCB");
    if (d.method != NQCD_METHOD_RPMD && d.algorithm != NQCD_ALG_BCB_TSIT5)
        return fail(h, NQCD_ERR_UNSUPPORTED, "ring-polymer surface hopping / Ehrenfest integrate with BCBwithTsit5");
    if (d.method == NQCD_METHOD_RPMD && d.n_states != 1)
        return fail(h, NQCD_ERR_UNSUPPORTED, "RPMD needs a single-state model");
    if (d.method != NQCD_METHOD_RPMD && d.n_states != 2)
        return fail(h, NQCD_ERR_UNSUPPORTED, "ring-polymer surface hopping / Ehrenfest need two states");
    return NQCD_OK;
}

int32_t rp_sample_initial(nqcd_handle_t* h, const nqcd_sampler_desc* S, uint64_t seed) {
    const bool pimc = S->kind == NQCD_SAMPLER_RING_POLYMER_PIMC;
    if (!pimc && (S->kind != NQCD_SAMPLER_RING_POLYMER_THERMAL || S->nparams != 3 || !S->params))
        return fail(h, NQCD_ERR_ARG, "ring-polymer sampler params: x0, w_well, beta");
    if (pimc && (!S->params || S->nparams < 4 || S->nparams > 6))
        return fail(h, NQCD_ERR_ARG, "PIMC sampler params: x0, Delta, passes, segment[, width[, reassess]]");
    const int P = h->prob.nbeads;
    const int64_t n = S->ntraj;
    const int grid = (int)((n + (kRpBlock / P) - 1) / (kRpBlock / P));
    const double wn = P * h->prob.temperature;
    const double* p = S->params;
    PimcArgs Q{};
    if (pimc) {
        if (!h->model_set[0]) return fail(h, NQCD_ERR_STATE, "model slot 0 is not set");
        const double passes = p[2], segment = p[3];
        if (!(passes >= 0.0) || passes > 1e9) return fail(h, NQCD_ERR_ARG, "PIMC passes must be in 0..1e9");
        const int L = std::max(1, (int)std::lround(segment * P));   // MetropolisHastings.jl:77
        if (!(segment > 0.0) || L > P) return fail(h, NQCD_ERR_ARG, "PIMC segment must be in (0, 1]");
        Q.ntraj = n; Q.steps = (int64_t)passes; Q.L = L;
        Q.width = S->nparams > 4 ? (int32_t)p[4] : 0;
        Q.reassess = S->nparams > 5 ? (int32_t)p[5] : 1;
        if (Q.width < 0 || Q.width > 1 || Q.reassess < 0 || Q.reassess > 1)
            return fail(h, NQCD_ERR_ARG, "PIMC width and reassess flags are 0 or 1");
        Q.x0 = p[0]; Q.delta = p[1]; Q.mass = h->masses[0]; Q.wn = wn; Q.key = seed; Q.model = h->models[0];
    }
    h->ntraj = n;
    HIPCHK(h, h->r0.ensure(sizeof(double) * P * n));
    HIPCHK(h, h->v0.ensure(sizeof(double) * P * n));
#define NQCD_CASE(PP)                                                                                           \
    case PP:                                                                                                    \
        if (pimc)                                                                                               \
            hipLaunchKernelGGL(rp_pimc<PP>, dim3(grid), dim3(kRpBlock), 0, h->stream, Q, h->r0.as<double>(),   \
                               h->v0.as<double>());                                                             \
        else                                                                                                    \
            hipLaunchKernelGGL(rp_sample<PP>, dim3(grid), dim3(kRpBlock), 0, h->stream, n, p[0], p[1], p[2],  \
                               h->masses[0], wn, seed, h->r0.as<double>(), h->v0.as<double>());                 \
        break;
    switch (P) {
        NQCD_CASE(1) NQCD_CASE(2) NQCD_CASE(4) NQCD_CASE(8) NQCD_CASE(16) NQCD_CASE(32) NQCD_CASE(64)
    default: return fail(h, NQCD_ERR_UNSUPPORTED, "nbeads");
    }
#undef NQCD_CASE
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->rp_sig_given = false;
    h->rp_state = S->initial_state;
    h->state0_given = false;
    if (S->state_basis == 1) return fail(h, NQCD_ERR_UNSUPPORTED, "ring-polymer sampler: adiabatic initial state only");
    return NQCD_OK;
}

int32_t rp_load_initial(nqcd_handle_t* h, int64_t ntraj, const double* r, const double* v, const double* sig_re,
                        const double* sig_im, const int32_t* state) {
    const int P = h->prob.nbeads;
    h->ntraj = ntraj;
    std::vector<double> tr((size_t)P * ntraj), tv((size_t)P * ntraj), ts((size_t)4 * ntraj);
    for (int64_t i = 0; i < ntraj; ++i)
        for (int b = 0; b < P; ++b) {
            tr[(size_t)b * ntraj + i] = r[(size_t)i * P + b];
            tv[(size_t)b * ntraj + i] = v[(size_t)i * P + b];
        }
    HIPCHK(h, h->r0.ensure(sizeof(double) * tr.size()));
    HIPCHK(h, h->v0.ensure(sizeof(double) * tv.size()));
    HIPCHK(h, hipMemcpy(h->r0.p, tr.data(), sizeof(double) * tr.size(), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->v0.p, tv.data(), sizeof(double) * tv.size(), hipMemcpyHostToDevice));
    h->rp_sig_given = false;
    if (h->prob.n_states == 2) {
        if (!sig_re || !sig_im) return fail(h, NQCD_ERR_ARG, "null density");
        for (int64_t i = 0; i < ntraj; ++i) {
            const size_t o = (size_t)i * 4;
            ts[i] = sig_re[o]; ts[ntraj + i] = sig_re[o + 3]; ts[2 * ntraj + i] = sig_re[o + 2];
            ts[3 * ntraj + i] = sig_im[o + 2];
        }
        HIPCHK(h, h->sig0.ensure(sizeof(double) * ts.size()));
        HIPCHK(h, hipMemcpy(h->sig0.p, ts.data(), sizeof(double) * ts.size(), hipMemcpyHostToDevice));
        h->rp_sig_given = true;
    }
    h->state0_given = state != nullptr;
    if (state) {
        HIPCHK(h, h->state0.ensure(sizeof(int32_t) * ntraj));
        HIPCHK(h, hipMemcpy(h->state0.p, state, sizeof(int32_t) * ntraj, hipMemcpyHostToDevice));
    }
    return NQCD_OK;
}

int32_t rp_run(nqcd_handle_t* h, double t0, int64_t nsteps, const nqcd_term_desc* term, uint64_t key) {
    if (term && term->kind != NQCD_TERM_NONE)
        return fail(h, NQCD_ERR_UNSUPPORTED, "termination predicates are not supported for ring polymers");
    if (!h->model_set[0]) return fail(h, NQCD_ERR_STATE, "model slot 0 is not set");
    const int P = h->prob.nbeads;
    const int64_t n = h->ntraj;
    HIPCHK(h, h->r.ensure(sizeof(double) * P * n));
    HIPCHK(h, h->v.ensure(sizeof(double) * P * n));
    HIPCHK(h, h->sig.ensure(sizeof(double) * 4 * n));
    HIPCHK(h, h->state.ensure(sizeof(int32_t) * n));
    HIPCHK(h, h->steps.ensure(sizeof(int32_t) * n));
    HIPCHK(h, h->nhops.ensure(sizeof(int32_t) * n));
    HIPCHK(h, h->save_ad.ensure(sizeof(double) * 2 * h->nsave * n));
    HIPCHK(h, h->save_di.ensure(sizeof(double) * 2 * h->nsave * n));
    HIPCHK(h, h->save_state.ensure(sizeof(int32_t) * h->nsave * n));
    HIPCHK(h, h->save_centroid.ensure(sizeof(double) * h->nsave * n));
    HIPCHK(h, launch_zero(h->save_ad.p, sizeof(double) * 2 * h->nsave * n, h->stream));
    HIPCHK(h, launch_zero(h->save_di.p, sizeof(double) * 2 * h->nsave * n, h->stream));
    HIPCHK(h, launch_zero(h->save_state.p, sizeof(int32_t) * h->nsave * n, h->stream));
    HIPCHK(h, launch_zero(h->save_centroid.p, sizeof(double) * h->nsave * n, h->stream));
    RpArgs A{};
    A.ntraj = n;
    A.method = (h->prob.method == NQCD_METHOD_RPMD) ? 0 : ((h->prob.method == NQCD_METHOD_RPSH) ? 1 : 2);
    A.rescaling = h->prob.rescaling;
    A.nsub = h->prob.electronic_substeps > 0 ? h->prob.electronic_substeps : 5;
    A.save_every = h->save_every;
    A.nsave = h->nsave;
    A.initial_state = h->rp_state;
    A.nstates = h->prob.n_states;
    A.dt = h->prob.dt;
    A.t0 = t0;
    A.mass = h->masses[0];
    A.wn = P * h->prob.temperature;
    A.nsteps = nsteps;
    A.key = key;
    A.model = h->models[0];
    A.r0 = h->r0.as<double>(); A.v0 = h->v0.as<double>();
    A.sig0 = h->rp_sig_given ? h->sig0.as<double>() : nullptr;
    A.state0 = h->state0_given ? h->state0.as<int32_t>() : nullptr;
    A.r = h->r.as<double>(); A.v = h->v.as<double>(); A.sig = h->sig.as<double>();
    A.state = h->state.as<int32_t>(); A.steps = h->steps.as<int32_t>(); A.nhops = h->nhops.as<int32_t>();
    A.save_centroid = h->save_centroid.as<double>();
    A.save_pop_ad = h->save_ad.as<double>(); A.save_pop_di = h->save_di.as<double>();
    A.save_state = h->save_state.as<int32_t>();
    HIPCHK(h, h->args_dev.ensure(sizeof(RpArgs)));
    HIPCHK(h, hipMemcpyAsync(h->args_dev.p, &A, sizeof(RpArgs), hipMemcpyHostToDevice, h->stream));
    int block = 0, nr = 0;
    const bool lane = rp_use_lane(h->prob.method, P, h->prob.precision);
    const bool elec = rp_needs_elec(h->prob.method, P);
    if (lane) { block = kRpLaneBlock; nr = kRpLaneBlock; }
    else rp_shape(P, h->prob.precision, elec, block, nr);
    if (nr < 1) return fail(h, NQCD_ERR_UNSUPPORTED, "nbeads");
    const int grid = (int)((n + nr - 1) / nr);
    HIPCHK(h, hipEventRecord(h->ev0, h->stream));
    if (lane) HIPCHK(h, rp_fill_cayley(P, h->prob.precision, h->stream, h->args_dev.as<RpArgs>()));
    if (lane) HIPCHK(h, rp_launch_lane(P, h->prob.precision, h->prob.method == NQCD_METHOD_RPSH, grid, h->stream,
                                       h->args_dev.as<const RpArgs>()));
    else HIPCHK(h, rp_launch(P, h->prob.precision, elec, grid, h->stream, h->args_dev.as<const RpArgs>()));
    HIPCHK(h, hipEventRecord(h->ev1, h->stream));
    h->stats.block = block;
    h->stats.grid = grid;
    h->stats.launches = 1;
    return NQCD_OK;
}

}  // namespace nqcd
