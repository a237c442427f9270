// nqcd_device.h -- device building blocks shared by the trajectory kernels.
//
// Everything here runs per lane (one trajectory per lane) in registers:
// counter-based Philox draws, the closed-form model Hamiltonians, the 2x2
// symmetric eigen-decomposition with eigenvector tracking, the nonadiabatic
// coupling, the Tsit5 density-matrix propagation and the fewest-switches hop.
#pragma once
#include <type_traits>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nqcd {

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon et al., SC'11).  Counter (traj_lo, traj_hi, step, stream),
// key (key_lo, key_hi).  Replaces the global rand() of fssh.jl:112.
// ---------------------------------------------------------------------------
enum { STREAM_INIT = 0, STREAM_HOP = 1, STREAM_STATE = 2, STREAM_OCCUPATION = 3 };

// Round-key bump on the scalar unit.  The key is wavefront-uniform; computing the
// schedule in place (instead of letting it be hoisted into ten SGPR pairs that are
// later spilled to VGPR lanes) keeps the round-key work on the co-issued SALU.
__device__ __forceinline__ uint32_t key_bump0(uint32_t k) {
    uint32_t r;
    k = __builtin_amdgcn_readfirstlane(k);
    asm volatile("s_add_u32 %0, %1, 0x9e3779b9" : "=s"(r) : "s"(k) : "scc");
    return r;
}
__device__ __forceinline__ uint32_t key_bump1(uint32_t k) {
    uint32_t r;
    k = __builtin_amdgcn_readfirstlane(k);
    asm volatile("s_add_u32 %0, %1, 0xbb67ae85" : "=s"(r) : "s"(k) : "scc");
    return r;
}

// k0, k1 must be wavefront-uniform (the run's Philox key).
__device__ __forceinline__ void philox4x32_10(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3,
                                              uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        if (round > 0) { k0 = key_bump0(k0); k1 = key_bump1(k1); }
        const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
        const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
}

__device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
    const uint64_t x = ((((uint64_t)a) << 32) | (uint64_t)b) >> 11;
    return (double)x * 0x1.0p-53;
}

__device__ __forceinline__ double philox_uniform(uint64_t traj, uint32_t step, uint32_t stream, uint64_t key) {
    uint32_t c0 = (uint32_t)traj, c1 = (uint32_t)(traj >> 32), c2 = step, c3 = stream;
    philox4x32_10(c0, c1, c2, c3, (uint32_t)key, (uint32_t)(key >> 32));
    return u53(c0, c1);
}

__device__ __forceinline__ void philox_normal2(uint64_t traj, uint32_t step, uint32_t stream, uint64_t key,
                                               double& z0, double& z1) {
    uint32_t c0 = (uint32_t)traj, c1 = (uint32_t)(traj >> 32), c2 = step, c3 = stream;
    philox4x32_10(c0, c1, c2, c3, (uint32_t)key, (uint32_t)(key >> 32));
    const double u1 = u53(c0, c1), u2 = u53(c2, c3);
    const double rad = sqrt(-2.0 * log(1.0 - u1));
    double s, c;
    sincos(6.283185307179586 * u2, &s, &c);
    z0 = rad * c;
    z1 = rad * s;
}

// A wavefront-uniform value (a launch argument read through a pointer) moved to scalar registers,
// so a loop that uses it needs no memory access for it.
__device__ __forceinline__ int32_t uniform_sgpr(int32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint64_t uniform_sgpr(uint64_t x) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ int64_t uniform_sgpr(int64_t x) { return (int64_t)uniform_sgpr((uint64_t)x); }
__device__ __forceinline__ double uniform_sgpr(double x) {
    return __longlong_as_double((long long)uniform_sgpr((uint64_t)__double_as_longlong(x)));
}

// ---------------------------------------------------------------------------
// Transcendentals by precision: FP64 path uses ocml's correctly-rounded-ish
// double routines; the FP32 path uses the hardware exp.
// ---------------------------------------------------------------------------
// TerminatingCallback predicates (DynamicsUtils/callbacks.jl:29) of nqcd_term_kind: 1 |x| > value,
// 2 x > value, 3 x < value, 0 none.
__device__ __forceinline__ bool term_hit(int kind, double x, double value) {
    return (kind == 1) ? fabs(x) > value : (kind == 2) ? x > value : (kind == 3) ? x < value : false;
}

// Simulation cell of the cell callbacks (DynamicsUtils/callbacks.jl:4-22), NQCBase's PeriodicCell: the
// lattice vectors are the columns of a (column-major, leading dimension 3), inv = a^-1, per flags the
// periodic dimensions; nd coordinates per atom.
struct CellDesc {
    int32_t mode;    // nqcd_cell_mode: 0 none, 1 CellBoundaryCallback (wrap), 2 TerminateCellCallback
    int32_t nd;
    int32_t per[3];
    int32_t pad;
    double a[9];
    double inv[9];
};

// Julia's mod(f, 1) for a Float64 (rem, then + 1 where the remainder is negative; +0 for a zero remainder)
__host__ __device__ __forceinline__ double jl_mod1(double f) {
    const double r = fmod(f, 1.0);
    if (r == 0.0) return 0.0;
    return (r > 0.0) ? r : r + 1.0;
}

// check_atoms_in_cell negated: some atom has a fractional coordinate f = inv x outside [0, 1].  x holds NA
// atoms of ND coordinates, the coordinate fastest ((ndofs, natoms) column-major); indices are compile-time
// constants after unrolling, so register arrays stay in registers.
template <int ND, int NA, typename R>
__device__ __forceinline__ bool cell_outside(const CellDesc& c, const R* x) {
    bool out = false;
#pragma unroll
    for (int at = 0; at < NA; ++at)
#pragma unroll
        for (int j = 0; j < ND; ++j) {
            double f = 0.0;
#pragma unroll
            for (int k = 0; k < ND; ++k) f += c.inv[j + 3 * k] * (double)x[at * ND + k];
            out = out || (f > 1.0) || (f < 0.0);
        }
    return out;
}

// apply_cell_boundaries!: every atom to f = inv x, f_j -> mod(f_j, 1) on the periodic dimensions, x = a f.
template <int ND, int NA, typename R>
__device__ __forceinline__ void cell_wrap(const CellDesc& c, R* x) {
#pragma unroll
    for (int at = 0; at < NA; ++at) {
        double f[ND];
#pragma unroll
        for (int j = 0; j < ND; ++j) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < ND; ++k) s += c.inv[j + 3 * k] * (double)x[at * ND + k];
            f[j] = c.per[j] ? jl_mod1(s) : s;
        }
#pragma unroll
        for (int j = 0; j < ND; ++j) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < ND; ++k) s += c.a[j + 3 * k] * f[k];
            x[at * ND + j] = (R)s;
        }
    }
}

// The cell callback after a step (after the hop callback): wrap when an atom left the cell
// (CellBoundaryCallback); returns true when the trajectory ends here (TerminateCellCallback).
template <int ND, int NA, typename R>
__device__ __forceinline__ bool cell_callback(const CellDesc& c, R* x) {
    if (c.mode == 0) return false;
    if (!cell_outside<ND, NA, R>(c, x)) return false;
    if (c.mode == 2) return true;
    cell_wrap<ND, NA, R>(c, x);
    return false;
}

__device__ __forceinline__ double fexp(double x) { return exp(x); }
__device__ __forceinline__ float fexp(float x) { return __expf(x); }
__device__ __forceinline__ double fabs_(double x) { return fabs(x); }
__device__ __forceinline__ float fabs_(float x) { return fabsf(x); }

// ---------------------------------------------------------------------------
// One-DOF two-state models (NQCModels Tully 1/2/3, double well + Gaussian
// coupling).  Outputs V = [[v11, v12],[v12, v22]] and dV/dx.
// ---------------------------------------------------------------------------
enum { MODEL_TULLY1 = 0, MODEL_TULLY2 = 1, MODEL_TULLY3 = 2, MODEL_SPIN_BOSON = 3, MODEL_DW_GAUSSIAN = 4,
       MODEL_HARMONIC = 5 };

struct ModelSlot {
    int32_t kind;
    int32_t pad;
    double p[8];
};

// A model slot held in registers (one per lane): kind and the first five parameters.
struct ModelRegs {
    int32_t kind;
    double p[5];
};

__device__ __forceinline__ ModelRegs model_regs(const ModelSlot* slots, int mid) {
    ModelRegs m;
    m.kind = slots[mid].kind;
#pragma unroll
    for (int i = 0; i < 5; ++i) m.p[i] = slots[mid].p[i];
    return m;
}

__device__ __forceinline__ ModelRegs model_regs(const ModelSlot& s) {
    ModelRegs m;
    m.kind = s.kind;
#pragma unroll
    for (int i = 0; i < 5; ++i) m.p[i] = s.p[i];
    return m;
}

// One model kind, selected at compile time (the per-kind bodies of model1d_eval, for callers that
// evaluate many points under one dispatch).
template <int K, typename R, class M>
__device__ __forceinline__ void model1d_kind(const M& m, R x, R& v11, R& v22, R& v12,
                                             R& d11, R& d22, R& d12) {
    if constexpr (K == MODEL_TULLY1) {
        const R A = (R)m.p[0], B = (R)m.p[1], C = (R)m.p[2], D = (R)m.p[3];
        const R e = fexp(-B * fabs_(x));
        const R sgn = (x > (R)0) ? (R)1 : ((x < (R)0) ? (R)-1 : (R)0);
        v11 = A * ((R)1 - e) * sgn;
        v22 = -v11;
        d11 = A * B * e;
        d22 = -d11;
        v12 = C * fexp(-D * x * x);
        d12 = (R)-2 * D * x * v12;
    } else if constexpr (K == MODEL_TULLY2) {
        const R A = (R)m.p[0], B = (R)m.p[1], C = (R)m.p[2], D = (R)m.p[3], E0 = (R)m.p[4];
        const R g = A * fexp(-B * x * x);
        v11 = (R)0; d11 = (R)0;
        v22 = -g + E0;
        d22 = (R)2 * B * x * g;
        v12 = C * fexp(-D * x * x);
        d12 = (R)-2 * D * x * v12;
    } else if constexpr (K == MODEL_TULLY3) {
        const R A = (R)m.p[0], B = (R)m.p[1], C = (R)m.p[2];
        v11 = A; v22 = -A; d11 = (R)0; d22 = (R)0;
        if (x < (R)0) { const R e = fexp(C * x); v12 = B * e; d12 = B * C * e; }
        else { const R e = fexp(-C * x); v12 = B * ((R)2 - e); d12 = B * C * e; }
    } else if constexpr (K == MODEL_DW_GAUSSIAN) {
        const R a = (R)m.p[0], b = (R)m.p[1], off = (R)m.p[2], C = (R)m.p[3], D = (R)m.p[4];
        const R x2 = x * x;
        v11 = a * x2 * x2 - b * x2;
        d11 = (R)4 * a * x2 * x - (R)2 * b * x;
        v22 = v11 + off; d22 = d11;
        v12 = C * fexp(-D * x2);
        d12 = (R)-2 * D * x * v12;
    } else if constexpr (K == MODEL_HARMONIC) {  // single state: V = m w^2 (x - x0)^2 / 2
        const R mm = (R)m.p[0], w = (R)m.p[1], x0 = (R)m.p[2];
        const R dx = x - x0;
        v11 = (R)0.5 * mm * w * w * dx * dx; d11 = mm * w * w * dx;
        v22 = v11; d22 = d11; v12 = (R)0; d12 = (R)0;
    } else {
        v11 = v22 = v12 = d11 = d22 = d12 = (R)0;
    }
}

// Calls f(std::integral_constant<int, K>{}) for the model kind K of m (one dispatch for a whole
// block of evaluations); unknown kinds take K = -1, whose model1d_kind is zero.
template <class M, class F>
__device__ __forceinline__ void model1d_dispatch(const M& m, F&& f) {
    switch (m.kind) {
    case MODEL_TULLY1: f(std::integral_constant<int, MODEL_TULLY1>{}); break;
    case MODEL_TULLY2: f(std::integral_constant<int, MODEL_TULLY2>{}); break;
    case MODEL_TULLY3: f(std::integral_constant<int, MODEL_TULLY3>{}); break;
    case MODEL_DW_GAUSSIAN: f(std::integral_constant<int, MODEL_DW_GAUSSIAN>{}); break;
    case MODEL_HARMONIC: f(std::integral_constant<int, MODEL_HARMONIC>{}); break;
    default: f(std::integral_constant<int, -1>{}); break;
    }
}

// m: a ModelRegs, or a ModelLane (below) whose parameters are already in the working precision
template <typename R, class M>
__device__ __forceinline__ void model1d_eval(const M& m, R x, R& v11, R& v22, R& v12,
                                             R& d11, R& d22, R& d12) {
    model1d_dispatch(m, [&](auto k) { model1d_kind<decltype(k)::value, R>(m, x, v11, v22, v12, d11, d22, d12); });
}

// A model's kind and parameters in the working precision (converted once, where a kernel keeps them
// per lane).
template <typename R>
struct ModelLane {
    int32_t kind;
    R p[5];
};

// Tully's models 1-3 without a branch on the kind: the two exponentials they use are shared,
// e1 = exp(-B|x|) (model 1), exp(-B x^2) (model 2), exp(-C|x|) (model 3, which is exp(C x) for x < 0 and
// exp(-C x) above), e2 = exp(-D x^2) (models 1 and 2), and each output is selected per lane.  Every value
// is formed with the same operations in the same order as model1d_kind's, so the results are bitwise
// those of the per-kind bodies; a wavefront whose lanes mix the three models then evaluates two
// exponentials instead of the five of its three branches.  Other kinds go through model1d_eval.
template <typename R>
__device__ __forceinline__ void model1d_eval_lane(const ModelLane<R>& m, R x, R& v11, R& v22, R& v12, R& d11, R& d22,
                                                  R& d12) {
    if (m.kind > MODEL_TULLY3) {
        model1d_eval<R>(m, x, v11, v22, v12, d11, d22, d12);
        return;
    }
    const bool t1 = m.kind == MODEL_TULLY1, t2 = m.kind == MODEL_TULLY2;
    const R A = m.p[0], B = m.p[1], C = m.p[2], D = m.p[3], E0 = m.p[4];
    // model 3 keeps its coupling amplitude in p[1] and its rate in p[2]
    const R k1 = t1 || t2 ? B : C;
    const R a1 = t2 ? -k1 * x * x : -k1 * fabs_(x);
    const R e1 = fexp(a1);
    const R e2 = fexp(-D * x * x);
    const R sgn = (x > (R)0) ? (R)1 : ((x < (R)0) ? (R)-1 : (R)0);
    const R g = A * e1;
    const R v11_1 = A * ((R)1 - e1) * sgn, d11_1 = A * B * e1;
    const R v12_12 = C * e2;
    v11 = t1 ? v11_1 : (t2 ? (R)0 : A);
    d11 = t1 ? d11_1 : (R)0;
    v22 = t1 ? -v11_1 : (t2 ? -g + E0 : -A);
    d22 = t1 ? -d11_1 : (t2 ? (R)2 * B * x * g : (R)0);
    v12 = (t1 || t2) ? v12_12 : (x < (R)0 ? B * e1 : B * ((R)2 - e1));
    d12 = (t1 || t2) ? (R)-2 * D * x * v12_12 : B * C * e1;
}

template <typename R>
__device__ __forceinline__ ModelLane<R> model_lane(const ModelRegs& m) {
    ModelLane<R> L;
    L.kind = m.kind;
#pragma unroll
    for (int i = 0; i < 5; ++i) L.p[i] = (R)m.p[i];
    return L;
}

// ---------------------------------------------------------------------------
// 2x2 real symmetric eigen-decomposition, eigenvalues ascending.  U is stored
// row-major (U[0]=U00, U[1]=U01, U[2]=U10, U[3]=U11); column j is the eigenvector
// of E_j.  A fresh decomposition makes the largest-magnitude component of each
// column positive; with a previous U, a column whose overlap with the previous
// one is negative is flipped (eigenvector tracking).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void eigen2(double a, double c, double d, double& E0, double& E1, double U[4],
                                       bool track) {
    const double mean = 0.5 * (a + d), h = 0.5 * (a - d);
    const double rad = sqrt(h * h + c * c);
    E0 = mean - rad;
    E1 = mean + rad;
    double p, q;  // eigenvector of E1
    if (h >= 0.0) { p = h + rad; q = c; } else { p = c; q = rad - h; }
    const double nrm2 = p * p + q * q;
    if (nrm2 > 0.0) { const double inv = 1.0 / sqrt(nrm2); p *= inv; q *= inv; } else { p = 1.0; q = 0.0; }
    // column 1 = (p, q), column 0 = (-q, p)
    double u00 = -q, u10 = p, u01 = p, u11 = q;
    if (track) {
        if (u00 * U[0] + u10 * U[2] < 0.0) { u00 = -u00; u10 = -u10; }
        if (u01 * U[1] + u11 * U[3] < 0.0) { u01 = -u01; u11 = -u11; }
    } else {
        const double m0 = (fabs(u10) > fabs(u00)) ? u10 : u00;
        if (m0 < 0.0) { u00 = -u00; u10 = -u10; }
        const double m1 = (fabs(u11) > fabs(u01)) ? u11 : u01;
        if (m1 < 0.0) { u01 = -u01; u11 = -u11; }
    }
    U[0] = u00; U[1] = u01; U[2] = u10; U[3] = u11;
}

// G = U^T dV U for dV = [[d11, d12],[d12, d22]]; returns G00, G11, G01.
__device__ __forceinline__ void rotate_derivative(const double U[4], double d11, double d22, double d12,
                                                  double& G00, double& G11, double& G01) {
    const double u00 = U[0], u01 = U[1], u10 = U[2], u11 = U[3];
    G00 = u00 * u00 * d11 + 2.0 * u00 * u10 * d12 + u10 * u10 * d22;
    G11 = u01 * u01 * d11 + 2.0 * u01 * u11 * d12 + u11 * u11 * d22;
    G01 = u00 * u01 * d11 + (u00 * u11 + u10 * u01) * d12 + u10 * u11 * d22;
}

// ---------------------------------------------------------------------------
// Tsit5 (Tsitouras 2011) coefficients.
// ---------------------------------------------------------------------------
namespace ts5 {
constexpr double c1 = 0.161, c2 = 0.327, c3 = 0.9, c4 = 0.9800255409045097;
constexpr double a21 = 0.161;
constexpr double a31 = -0.008480655492356989, a32 = 0.335480655492357;
constexpr double a41 = 2.897153057105493, a42 = -6.359448489975075, a43 = 4.3622954328695815;
constexpr double a51 = 5.325864828439257, a52 = -11.748883564062828, a53 = 7.4955393428898365,
                 a54 = -0.09249506636175525;
constexpr double a61 = 5.86145544294642, a62 = -12.92096931784711, a63 = 8.159367898576159,
                 a64 = -0.071584973281401, a65 = -0.028269050394068383;
constexpr double a71 = 0.09646076681806523, a72 = 0.01, a73 = 0.4798896504144996, a74 = 1.379008574103742,
                 a75 = -3.290069515436081, a76 = 2.324710524099774;
}  // namespace ts5

// Linear interpolation of the electronic parameters between the two ends of the
// nuclear step (DoubleBuffer, electronic_dynamics.jl:15-79): dE = E0 - E1 and
// w = (v.d)_{01}, each value(tau) = v0 + v1 * (tau - tc).
struct Interp2 {
    double dE0, dE1, w0, w1;
};

// Hermitian 2x2 density sigma = [[a, x+iy],[x-iy, b]] carried as q = a - b, x, y with the
// trace a + b held fixed.  Under V = [[E0, -iw],[iw, E1]], d sigma/dt = -i[V, sigma] reads
//   dq/dt = -4 w x,   dx/dt = dE y + w q,   dy/dt = -dE x,   dE = E0 - E1.
struct Sig2 {
    double q, x, y;
};

// One Tsit5 step of the density carried as (p, x, y) with p = q / 2, so that
//   dp/dt = -G x,   dx/dt = dE y + G p,   dy/dt = -dE x,   G = 2 w
// (one product per right-hand side fewer than the q form; every quantity is the q form's scaled
// by a power of two, so the rounding is the same).  The stage derivatives are carried
// pre-multiplied by the step h (K = h f): at stage time tau0 + c h the coefficients are
// h dE = he0 + he1 c and h G = hg0 + hg1 c.
__device__ __forceinline__ void tsit5_step2(double he0, double he1, double hg0, double hg1, Sig2& u) {
    using namespace ts5;
#define NQCD_K(c, Q, X, Y, KQ, KX, KY)             \
    {                                              \
        const double e = fma(he1, (c), he0);       \
        const double g = fma(hg1, (c), hg0);       \
        KQ = -(g * (X));                           \
        KX = fma(e, (Y), g * (Q));                 \
        KY = -(e * (X));                           \
    }
    double q1, x1, y1, q2, x2, y2, q3, x3, y3, q4, x4, y4, q5, x5, y5, q6, x6, y6;
    q1 = -(hg0 * u.x); x1 = fma(he0, u.y, hg0 * u.q); y1 = -(he0 * u.x);
    double tq = fma(a21, q1, u.q), tx = fma(a21, x1, u.x), ty = fma(a21, y1, u.y);
    NQCD_K(c1, tq, tx, ty, q2, x2, y2);
    tq = fma(a32, q2, fma(a31, q1, u.q)); tx = fma(a32, x2, fma(a31, x1, u.x)); ty = fma(a32, y2, fma(a31, y1, u.y));
    NQCD_K(c2, tq, tx, ty, q3, x3, y3);
    tq = fma(a43, q3, fma(a42, q2, fma(a41, q1, u.q)));
    tx = fma(a43, x3, fma(a42, x2, fma(a41, x1, u.x)));
    ty = fma(a43, y3, fma(a42, y2, fma(a41, y1, u.y)));
    NQCD_K(c3, tq, tx, ty, q4, x4, y4);
    tq = fma(a54, q4, fma(a53, q3, fma(a52, q2, fma(a51, q1, u.q))));
    tx = fma(a54, x4, fma(a53, x3, fma(a52, x2, fma(a51, x1, u.x))));
    ty = fma(a54, y4, fma(a53, y3, fma(a52, y2, fma(a51, y1, u.y))));
    NQCD_K(c4, tq, tx, ty, q5, x5, y5);
    tq = fma(a65, q5, fma(a64, q4, fma(a63, q3, fma(a62, q2, fma(a61, q1, u.q)))));
    tx = fma(a65, x5, fma(a64, x4, fma(a63, x3, fma(a62, x2, fma(a61, x1, u.x)))));
    ty = fma(a65, y5, fma(a64, y4, fma(a63, y3, fma(a62, y2, fma(a61, y1, u.y)))));
    NQCD_K(1.0, tq, tx, ty, q6, x6, y6);
    u.q = fma(a76, q6, fma(a75, q5, fma(a74, q4, fma(a73, q3, fma(a72, q2, fma(a71, q1, u.q))))));
    u.x = fma(a76, x6, fma(a75, x5, fma(a74, x4, fma(a73, x3, fma(a72, x2, fma(a71, x1, u.x))))));
    u.y = fma(a76, y6, fma(a75, y5, fma(a74, y4, fma(a73, y3, fma(a72, y2, fma(a71, y1, u.y))))));
#undef NQCD_K
}

// step!(electronic_integrator, dt, true) from time t: nsub sub-steps of dt/nsub,
// each clipped so the last lands on t + dt (bab_electronics.jl:40-43,87-90).  Sub-steps before the
// last are never clipped (t + dt - tau >= 2 h there), so the coefficients come from products with
// h formed once; a clipped last sub-step (hs < h, by rounding) scales them by r = hs / h.
__device__ __forceinline__ void propagate2(const Interp2& P, Sig2& u, double t, double tc, double dt, int nsub) {
    const double h = dt / (double)nsub;
    const double hdE0 = h * P.dE0, hdE1 = h * P.dE1, hG0 = h * (2.0 * P.w0), hG1 = h * (2.0 * P.w1);
    const double he1 = h * hdE1, hg1 = h * hG1;
    Sig2 v = {0.5 * u.q, u.x, u.y};
    double tau = t;
    for (int s = 0; s < nsub - 1; ++s) {
        const double s0 = tau - tc;
        tsit5_step2(fma(hdE1, s0, hdE0), he1, fma(hG1, s0, hG0), hg1, v);
        tau += h;
    }
    const double hs = fmin(h, (t + dt) - tau), s0 = tau - tc;
    const double r = (hs == h) ? 1.0 : hs / h;
    tsit5_step2(r * fma(hdE1, s0, hdE0), (r * r) * he1, r * fma(hG1, s0, hG0), (r * r) * hg1, v);
    u.q = 2.0 * v.q;
    u.x = v.x;
    u.y = v.y;
}

// ---------------------------------------------------------------------------
// Dense complex 2x2 formulation (the reference's literal data flow: A = -iW,
// diag(A) = E, du = -i(A u - u A) on all eight real components).  Used by the
// step-per-launch plain kernel.
// ---------------------------------------------------------------------------
struct C2 {
    double re[4], im[4];
};

struct InterpDense {
    double E0c, E1c, E0n, E1n, Wc, Wn, tc, inv;  // W = (v.d)_{01}; W_{10} = -W
};

__device__ __forceinline__ void rhs_dense(const InterpDense& P, double tau, const C2& u, C2& du) {
    const double loc = (tau - P.tc) * P.inv;
    const double E0 = P.E0c + (P.E0n - P.E0c) * loc;
    const double E1 = P.E1c + (P.E1n - P.E1c) * loc;
    const double w = P.Wc + (P.Wn - P.Wc) * loc;
    // A = [[E0, -i w],[ i w, E1]] as complex entries
    double Are[4] = {E0, 0.0, 0.0, E1};
    double Aim[4] = {0.0, -w, w, 0.0};
    double cre[4], cim[4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            double sr = 0.0, si = 0.0;
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                // A[i,k] * u[k,j] - u[i,k] * A[k,j]
                sr += Are[i * 2 + k] * u.re[k * 2 + j] - Aim[i * 2 + k] * u.im[k * 2 + j];
                si += Are[i * 2 + k] * u.im[k * 2 + j] + Aim[i * 2 + k] * u.re[k * 2 + j];
                sr -= u.re[i * 2 + k] * Are[k * 2 + j] - u.im[i * 2 + k] * Aim[k * 2 + j];
                si -= u.re[i * 2 + k] * Aim[k * 2 + j] + u.im[i * 2 + k] * Are[k * 2 + j];
            }
            cre[i * 2 + j] = sr;
            cim[i * 2 + j] = si;
        }
#pragma unroll
    for (int J = 0; J < 4; ++J) { du.re[J] = cim[J]; du.im[J] = -cre[J]; }  // * (-i)
}

__device__ __forceinline__ void tsit5_step_dense(const InterpDense& P, C2& u, double tau, double h) {
    using namespace ts5;
    C2 k1, k2, k3, k4, k5, k6, t;
    rhs_dense(P, tau, u, k1);
#pragma unroll
    for (int J = 0; J < 4; ++J) { t.re[J] = u.re[J] + h * (a21 * k1.re[J]); t.im[J] = u.im[J] + h * (a21 * k1.im[J]); }
    rhs_dense(P, tau + c1 * h, t, k2);
#pragma unroll
    for (int J = 0; J < 4; ++J) {
        t.re[J] = u.re[J] + h * (a31 * k1.re[J] + a32 * k2.re[J]);
        t.im[J] = u.im[J] + h * (a31 * k1.im[J] + a32 * k2.im[J]);
    }
    rhs_dense(P, tau + c2 * h, t, k3);
#pragma unroll
    for (int J = 0; J < 4; ++J) {
        t.re[J] = u.re[J] + h * (a41 * k1.re[J] + a42 * k2.re[J] + a43 * k3.re[J]);
        t.im[J] = u.im[J] + h * (a41 * k1.im[J] + a42 * k2.im[J] + a43 * k3.im[J]);
    }
    rhs_dense(P, tau + c3 * h, t, k4);
#pragma unroll
    for (int J = 0; J < 4; ++J) {
        t.re[J] = u.re[J] + h * (a51 * k1.re[J] + a52 * k2.re[J] + a53 * k3.re[J] + a54 * k4.re[J]);
        t.im[J] = u.im[J] + h * (a51 * k1.im[J] + a52 * k2.im[J] + a53 * k3.im[J] + a54 * k4.im[J]);
    }
    rhs_dense(P, tau + c4 * h, t, k5);
#pragma unroll
    for (int J = 0; J < 4; ++J) {
        t.re[J] = u.re[J] + h * (a61 * k1.re[J] + a62 * k2.re[J] + a63 * k3.re[J] + a64 * k4.re[J] + a65 * k5.re[J]);
        t.im[J] = u.im[J] + h * (a61 * k1.im[J] + a62 * k2.im[J] + a63 * k3.im[J] + a64 * k4.im[J] + a65 * k5.im[J]);
    }
    rhs_dense(P, tau + h, t, k6);
#pragma unroll
    for (int J = 0; J < 4; ++J) {
        u.re[J] = u.re[J] + h * (a71 * k1.re[J] + a72 * k2.re[J] + a73 * k3.re[J] + a74 * k4.re[J] +
                                 a75 * k5.re[J] + a76 * k6.re[J]);
        u.im[J] = u.im[J] + h * (a71 * k1.im[J] + a72 * k2.im[J] + a73 * k3.im[J] + a74 * k4.im[J] +
                                 a75 * k5.im[J] + a76 * k6.im[J]);
    }
}

__device__ __forceinline__ void propagate_dense(const InterpDense& P, C2& u, double t, double dt, int nsub) {
    const double h = dt / (double)nsub;
    const double tend = t + dt;
    double tau = t;
    for (int s = 0; s < nsub; ++s) {
        const double hs = fmin(h, tend - tau);
        tsit5_step_dense(P, u, tau, hs);
        tau = (s == nsub - 1) ? tend : tau + hs;
    }
}

// ---------------------------------------------------------------------------
// Velocity rescaling along d (surface_hopping.jl:64-99), D flattened DOFs.
// Returns true when the hop is accepted.
// ---------------------------------------------------------------------------
template <int D, typename R>
__device__ __forceinline__ bool rescale_velocity(int rescaling, R v[D], const double inv_mass[D],
                                                 const double dcoup[D], double dEnergy) {
    if (rescaling == 2) return true;  // :off
    double a = 0.0, b = 0.0;
#pragma unroll
    for (int I = 0; I < D; ++I) { a += dcoup[I] * dcoup[I] * inv_mass[I]; b += dcoup[I] * (double)v[I]; }
    a *= 0.5;
    const double disc = b * b - 4.0 * a * dEnergy;
    if (disc < 0.0) {
        if (rescaling == 1) {  // frustrated_hop_invert_velocity! (:155-164)
            double nrm = 0.0;
#pragma unroll
            for (int I = 0; I < D; ++I) nrm += dcoup[I] * dcoup[I];
            nrm = sqrt(nrm);
            double g = 0.0;
#pragma unroll
            for (int I = 0; I < D; ++I) g += (double)v[I] * (dcoup[I] / nrm);
#pragma unroll
            for (int I = 0; I < D; ++I) v[I] = (R)((double)v[I] - 2.0 * g * (dcoup[I] / nrm));
        }
        return false;
    }
    const double root = sqrt(disc);
    const double gamma = (b < 0.0) ? (b + root) / (2.0 * a) : (b - root) / (2.0 * a);
#pragma unroll
    for (int I = 0; I < D; ++I) v[I] = (R)((double)v[I] - gamma * dcoup[I] * inv_mass[I]);
    return true;
}

// Wave-level helpers (wave64).
__device__ __forceinline__ int lane_id() { return __lane_id(); }

}  // namespace nqcd
