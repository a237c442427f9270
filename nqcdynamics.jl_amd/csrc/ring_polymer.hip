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
#include <vector>

namespace nqcd {

// Workgroup size per bead count and precision: 64 ring polymers (one full wavefront of
// electronics lanes) up to 512 lanes in FP32 (two waves per SIMD, so up to 256 VGPRs without
// spilling: the kernel needs 150-250); at most 256 lanes in FP64, whose Hartley rows take twice the
// registers.
template <int P, typename R>
struct RpShape {
    static constexpr int kMax = sizeof(R) == 4 ? 512 : 256;
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

// Bead force (per lane): RPMD uses the single-state potential; RPSH the active adiabatic
// surface; RP-Ehrenfest the mean field with the centroid density (ehrenfest_rpmd.jl:23-43).
template <typename R>
__device__ __forceinline__ double bead_force(const RpArgs& A, const ModelRegs& m, R x, double Ub[4], bool track,
                                             int st, double a, double b, double sx) {
    R v11, v22, v12, d11, d22, d12;
    model1d_eval<R>(m, x, v11, v22, v12, d11, d22, d12);
    if (A.method == 0) return (double)d11;
    if (A.method == 1) {
        // the active surface's gradient (U^T dV U)_ss without the eigenvectors: E = m -+ rad with
        // m = (V11 + V22) / 2, h = (V11 - V22) / 2, rad = sqrt(h^2 + V12^2), so
        // dE/dx = dm -+ (h dh + V12 dV12) / rad, in the bead precision
        const R h = (R)0.5 * (v11 - v22), dh = (R)0.5 * (d11 - d22), dm = (R)0.5 * (d11 + d22);
        const R rad = sqrt(h * h + v12 * v12);
        const R g = (rad > (R)0) ? (h * dh + v12 * d12) / rad : (R)0;
        return (double)((st == 0) ? dm - g : dm + g);
    }
    double E0, E1, G00, G11, G01;
    eigen2((double)v11, (double)v12, (double)v22, E0, E1, Ub, track);
    rotate_derivative(Ub, (double)d11, (double)d22, (double)d12, G00, G11, G01);
    if (A.method == 1) return (st == 0) ? G00 : G11;
    return G00 * a + G11 * b + 2.0 * G01 * sx;
}

template <int P, typename R>
__device__ __forceinline__ void rp_body(const RpArgs* __restrict__ Ap) {
    constexpr int NB = RpShape<P, R>::block, NR = RpShape<P, R>::nr;
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
    const bool elec = A.method != 0;
    // electronics lane: ring polymer e = threadIdx.x of the workgroup
    const int e = threadIdx.x;
    const bool eln = elec && e < NR;
    const int64_t erp = (int64_t)blockIdx.x * NR + e;
    const bool evalid = eln && erp < n;
    const int64_t eid = evalid ? erp : 0;
    double tr = 1.0;
    Sig2 s = {1.0, 0.0, 0.0};
    int est = 0, nh = 0;
    double Uc[4] = {1.0, 0.0, 0.0, 1.0}, Ub[4] = {1.0, 0.0, 0.0, 1.0};
    double dEc = 0.0, wc = 0.0, tc = 0.0, t = A.t0;
    // active state / density of the step start as the bead lanes see them
    int st = 0;
    double a = 1.0, b = 0.0, sx = 0.0;
    if (elec) {
        if (l == 0) x_rc[j] = q / (R)sqP;
        __syncthreads();
        if (eln) {
            double ea = 1.0, eb = 0.0, ex = 0.0, ey = 0.0;
            if (evalid) {
                const RpArgs& B = *rp_opaque(Ap);
                if (B.sig0) {
                    ea = B.sig0[eid]; eb = B.sig0[n + eid]; ex = B.sig0[2 * n + eid]; ey = B.sig0[3 * n + eid];
                } else {
                    ea = (B.initial_state == 0) ? 1.0 : 0.0; eb = 1.0 - ea;
                }
                est = B.state0 ? B.state0[eid] : -1;
                R v11, v22, v12, d11, d22, d12;
                model1d_eval<R>(m, x_rc[e], v11, v22, v12, d11, d22, d12);
                double E0c, E1c, G00c, G11c, G01c;
                eigen2((double)v11, (double)v12, (double)v22, E0c, E1c, Uc, false);
                rotate_derivative(Uc, (double)d11, (double)d22, (double)d12, G00c, G11c, G01c);
                if (A.method == 1 && est < 0) {  // fssh.jl:54
                    const double u = philox_uniform((uint64_t)eid, 0u, STREAM_STATE, A.key) * (ea + eb);
                    est = (u < ea) ? 0 : 1;
                }
            }
            tr = ea + eb;
            s.q = ea - eb; s.x = ex; s.y = ey;
            x_st[e] = est; x_a[e] = ea; x_b[e] = eb; x_sx[e] = ex; x_dp[e] = 0.0;
        }
        __syncthreads();
        st = x_st[j]; a = x_a[j]; b = x_b[j]; sx = x_sx[j];
    }
    // initial accelerations in modes
    R acc = (R)(-bead_force<R>(A, m, rb, Ub, false, st, a, b, sx) / A.mass);
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
        double pa0, pa1;
        if (B.method == 1) { pa0 = (est == 0) ? 1.0 : 0.0; pa1 = 1.0 - pa0; }
        else { pa0 = 0.5 * (tr + s.q); pa1 = 0.5 * (tr - s.q); }
        B.save_pop_ad[((int64_t)slot * 2) * n + eid] = pa0;
        B.save_pop_ad[((int64_t)slot * 2 + 1) * n + eid] = pa1;
        B.save_pop_di[((int64_t)slot * 2) * n + eid] = Uc[0] * Uc[0] * pa0 + Uc[1] * Uc[1] * pa1 + 2.0 * Uc[0] * Uc[1] * s.x;
        B.save_pop_di[((int64_t)slot * 2 + 1) * n + eid] = Uc[2] * Uc[2] * pa0 + Uc[3] * Uc[3] * pa1 + 2.0 * Uc[2] * Uc[3] * s.x;
        B.save_state[(int64_t)slot * n + eid] = est;
    };
    save_centroid(0);
    save_elec(0);
    const R hdt = (R)(0.5 * A.dt);
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
        acc = (R)(-bead_force<R>(A, m, rb, Ub, true, st, a, b, sx) / A.mass);
        an = hartley<P, R>(acc, H, lds, gbase);
        p = p + hdt * an;                                        // step_B!
        const double tn = t + A.dt;
        if (elec) {
            if (l == 0) { x_rc[j] = q / (R)sqP; x_vc[j] = (double)p / sqP; }
            __syncthreads();
            if (eln) {
                double dp = 0.0;
                if (evalid) {
                    const R rc = x_rc[e];
                    const double vc = x_vc[e];
                    R v11, v22, v12, d11, d22, d12;
                    model1d_eval<R>(m, rc, v11, v22, v12, d11, d22, d12);
                    double E0c, E1c, G00c, G11c, G01c;
                    eigen2((double)v11, (double)v12, (double)v22, E0c, E1c, Uc, true);
                    rotate_derivative(Uc, (double)d11, (double)d22, (double)d12, G00c, G11c, G01c);
                    const double gap = E1c - E0c;
                    const double d01 = G01c / gap;
                    const double w = d01 * vc;
                    const double dEn = E0c - E1c;
                    const double inv = 1.0 / (tn - tc);
                    Interp2 Pi;
                    Pi.dE0 = dEc; Pi.dE1 = (dEn - dEc) * inv; Pi.w0 = wc; Pi.w1 = (w - wc) * inv;
                    propagate2(Pi, s, t, tc, A.dt, A.nsub);
                    dEc = dEn; wc = w; tc = tn;
                    // the save precedes the hop callback (a save time holds the state before the step's
                    // callbacks; the hop's velocity change reaches the beads at the next step)
                    if (A.save_every > 0 && ((step + 1) % A.save_every) == 0) save_elec((step + 1) / A.save_every);
                    if (A.method == 1) {
                        const double sss = 0.5 * (tr + ((est == 0) ? s.q : -s.q));
                        const double ratio = s.x / sss;
                        const double dsm = (est == 0) ? d01 : -d01;
                        double pr = 2.0 * vc * ratio * dsm * A.dt;
                        if (pr > 0.0) {
                            pr = fmin(pr, 1.0);
                            const double rnd = philox_uniform((uint64_t)eid, (uint32_t)step, STREAM_HOP, A.key);
                            if (pr > rnd) {
                                const double d = -dsm;  // coupling[new, old]
                                const double dE = (est == 0) ? gap : -gap;
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
                                if (ok) { est = 1 - est; ++nh; }
                            }
                        }
                    }
                }
                x_st[e] = est;
                x_a[e] = 0.5 * (tr + s.q); x_b[e] = 0.5 * (tr - s.q); x_sx[e] = s.x;
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
        if (!elec && l == 0) {  // RPMD: no electronic state
            B.sig[id] = 1.0; B.sig[n + id] = 0.0; B.sig[2 * n + id] = 0.0; B.sig[3 * n + id] = 0.0;
            B.state[id] = 0; B.steps[id] = (int32_t)B.nsteps; B.nhops[id] = 0;
        }
    }
    if (evalid) {
        B.sig[eid] = 0.5 * (tr + s.q);
        B.sig[n + eid] = 0.5 * (tr - s.q);
        B.sig[2 * n + eid] = s.x;
        B.sig[3 * n + eid] = s.y;
        B.state[eid] = est;
        B.steps[eid] = (int32_t)B.nsteps;
        B.nhops[eid] = nh;
    }
}

#define NQCD_RP_KERNELS(P)                                                                                  \
    __global__ __launch_bounds__((RpShape<P, float>::block)) void rp_kernel_f32_##P(const RpArgs* __restrict__ Ap) { \
        rp_body<P, float>(Ap);                                                                              \
    }                                                                                                       \
    __global__ __launch_bounds__((RpShape<P, double>::block)) void rp_kernel_f64_##P(const RpArgs* __restrict__ Ap) { \
        rp_body<P, double>(Ap);                                                                             \
    }
NQCD_RP_KERNELS(1)
NQCD_RP_KERNELS(2)
NQCD_RP_KERNELS(4)
NQCD_RP_KERNELS(8)
NQCD_RP_KERNELS(16)
NQCD_RP_KERNELS(32)
NQCD_RP_KERNELS(64)
#undef NQCD_RP_KERNELS

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

// Workgroup size and ring polymers per workgroup of the (P, precision) kernel.
static void rp_shape(int P, int precision, int& block, int& nr) {
#define NQCD_CASE(PP)                                                                                         \
    case PP:                                                                                                  \
        block = precision == 0 ? (RpShape<PP, double>::block) : (RpShape<PP, float>::block);                  \
        nr = precision == 0 ? (RpShape<PP, double>::nr) : (RpShape<PP, float>::nr);                           \
        return;
    switch (P) {
        NQCD_CASE(1) NQCD_CASE(2) NQCD_CASE(4) NQCD_CASE(8) NQCD_CASE(16) NQCD_CASE(32) NQCD_CASE(64)
    default: block = 0; nr = 0;
    }
#undef NQCD_CASE
}

static hipError_t rp_launch(int P, int precision, int grid, hipStream_t s, const RpArgs* Adev) {
#define NQCD_CASE(PP)                                                                                         \
    case PP:                                                                                                  \
        if (precision == 0)                                                                                   \
            hipLaunchKernelGGL(rp_kernel_f64_##PP, dim3(grid), dim3((RpShape<PP, double>::block)), 0, s, Adev); \
        else hipLaunchKernelGGL(rp_kernel_f32_##PP, dim3(grid), dim3((RpShape<PP, float>::block)), 0, s, Adev); \
        break;
    switch (P) {
        NQCD_CASE(1) NQCD_CASE(2) NQCD_CASE(4) NQCD_CASE(8) NQCD_CASE(16) NQCD_CASE(32) NQCD_CASE(64)
    default: return hipErrorInvalidValue;
    }
#undef NQCD_CASE
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
    rp_shape(P, h->prob.precision, block, nr);
    if (nr < 1) return fail(h, NQCD_ERR_UNSUPPORTED, "nbeads");
    const int grid = (int)((n + nr - 1) / nr);
    HIPCHK(h, hipEventRecord(h->ev0, h->stream));
    HIPCHK(h, rp_launch(P, h->prob.precision, grid, h->stream, h->args_dev.as<const RpArgs>()));
    HIPCHK(h, hipEventRecord(h->ev1, h->stream));
    h->stats.block = block;
    h->stats.grid = grid;
    h->stats.launches = 1;
    return NQCD_OK;
}

}  // namespace nqcd
