// nqcd_oracle.cpp -- FP64 CPU restatement of NQCDynamics.jl ensemble trajectory
// integration (FSSH and Ehrenfest with BABwithTsit5, hopping, rescaling, outputs).
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP library
// and the timed CPU baseline.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load it.  Nothing in the product links it.
//
// Every function cites the reference file:line it restates (paths relative to the
// NQCDynamics.jl v1.0.5 source tree).  Where behaviour lives in third-party
// Julia packages absent from the reference tree (NQCModels, NQCCalculators,
// OrdinaryDiffEq's Tsit5, RingPolymerArrays, NQCDistributions) the published
// algorithm is restated and the choice is named in the function comment.
//
// Numerics: IEEE FP64, no FMA contraction (build with -ffp-contract=off).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

using cd = std::complex<double>;

#define OR_API extern "C" __attribute__((visibility("default")))

// ---------------------------------------------------------------------------
// Philox4x32-10 counter-based RNG (Salmon, Moraes, Dror, Shaw, SC'11).
// Replaces Julia's global rand() at fssh.jl:112 and iesh.jl:393 so that the CPU
// and GPU draw identical hop uniforms.  Counter = (traj_lo, traj_hi, step, stream),
// key = (key_lo, key_hi).
// ---------------------------------------------------------------------------
static inline void philox10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        if (round > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline double u53(uint32_t a, uint32_t b) {
    const uint64_t x = (((uint64_t)a << 32) | (uint64_t)b) >> 11;
    return (double)x * 0x1.0p-53;
}

static inline void philox_draw(uint64_t traj, uint32_t step, uint32_t stream, uint64_t key, uint32_t o[4]) {
    const uint32_t ctr[4] = {(uint32_t)traj, (uint32_t)(traj >> 32), step, stream};
    const uint32_t k[2] = {(uint32_t)key, (uint32_t)(key >> 32)};
    philox10(ctr, k, o);
}

// One uniform in [0,1) per (traj, step, stream): the hop draw of fssh.jl:112.
static inline double uniform01(uint64_t traj, uint32_t step, uint32_t stream, uint64_t key) {
    uint32_t o[4];
    philox_draw(traj, step, stream, key, o);
    return u53(o[0], o[1]);
}

// Two standard normals per counter (Box-Muller on the two 53-bit uniforms).
static inline void normal2(uint64_t traj, uint32_t step, uint32_t stream, uint64_t key, double* z0, double* z1) {
    uint32_t o[4];
    philox_draw(traj, step, stream, key, o);
    const double u1 = u53(o[0], o[1]);
    const double u2 = u53(o[2], o[3]);
    const double rad = std::sqrt(-2.0 * std::log(1.0 - u1));
    const double ang = 6.283185307179586 * u2;
    *z0 = rad * std::cos(ang);
    *z1 = rad * std::sin(ang);
}

enum { STREAM_INIT = 0, STREAM_HOP = 1, STREAM_STATE = 2, STREAM_OCCUPATION = 3 };

OR_API void oracle_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) { philox10(ctr, key, out); }
OR_API double oracle_uniform(uint64_t traj, uint32_t step, uint32_t stream, uint64_t key) {
    return uniform01(traj, step, stream, key);
}
OR_API void oracle_normal2(uint64_t traj, uint32_t step, uint32_t stream, uint64_t key, double* z) {
    normal2(traj, step, stream, key, z, z + 1);
}

// ---------------------------------------------------------------------------
// Models (NQCModels is not in the reference tree; closed forms from the model
// definitions: Tully, J. Chem. Phys. 93, 1061 (1990) for TULLY1/2/3;
// docs/src/NQCModels/systembathmodels.md:21-27 for the spin-boson matrix).
// V is the state-dependent (electronic) part, V0 the state-independent part.
// ---------------------------------------------------------------------------
enum {
    MODEL_TULLY1 = 0,
    MODEL_TULLY2 = 1,
    MODEL_TULLY3 = 2,
    MODEL_SPIN_BOSON = 3,
    MODEL_DW_GAUSSIAN = 4,
    MODEL_HARMONIC = 5,
    MODEL_DIATOM_SURFACE = 7,
    MODEL_MORSE_N = 9,
    MODEL_ANDERSON_HOLSTEIN_1 = 10,
};

struct OModelDesc {
    int32_t kind;
    int32_t nstates;
    int32_t ndof;        // flattened ndofs*natoms
    int32_t nbath;       // spin-boson bath modes (ndof == nbath)
    double p[16];
    const double* omega; // nbath
    const double* cbath; // nbath
    const double* ext;   // MODEL_MORSE_N: [D, beta, r0, c] per state, then [A, a, r] per pair i < j (row-major)
};

static void model_eval(const OModelDesc& m, const double* r, double* V, double* dV, double* V0, double* dV0) {
    const int n = m.nstates, D = m.ndof;
    std::fill(V, V + n * n, 0.0);
    std::fill(dV, dV + (size_t)D * n * n, 0.0);
    *V0 = 0.0;
    std::fill(dV0, dV0 + D, 0.0);
    switch (m.kind) {
    case MODEL_TULLY1: {  // V11 = A(1-exp(-B|x|))sign(x), V22 = -V11, V12 = C exp(-D x^2)
        const double A = m.p[0], B = m.p[1], C = m.p[2], Dd = m.p[3];
        const double x = r[0];
        const double e = std::exp(-B * std::fabs(x));
        const double sgn = (x > 0) ? 1.0 : ((x < 0) ? -1.0 : 0.0);
        const double v11 = A * (1.0 - e) * sgn;
        const double dv11 = A * B * e;
        const double v12 = C * std::exp(-Dd * x * x);
        const double dv12 = -2.0 * Dd * x * v12;
        V[0] = v11; V[3] = -v11; V[1] = V[2] = v12;
        dV[0] = dv11; dV[3] = -dv11; dV[1] = dV[2] = dv12;
        break;
    }
    case MODEL_TULLY2: {  // V11 = 0, V22 = -A exp(-B x^2) + E0, V12 = C exp(-D x^2)
        const double A = m.p[0], B = m.p[1], C = m.p[2], Dd = m.p[3], E0 = m.p[4];
        const double x = r[0];
        const double g = A * std::exp(-B * x * x);
        const double v12 = C * std::exp(-Dd * x * x);
        V[0] = 0.0; V[3] = -g + E0; V[1] = V[2] = v12;
        dV[0] = 0.0; dV[3] = 2.0 * B * x * g; dV[1] = dV[2] = -2.0 * Dd * x * v12;
        break;
    }
    case MODEL_TULLY3: {  // V11 = A, V22 = -A, V12 = B e^{Cx} (x<0), B(2 - e^{-Cx}) (x>=0)
        const double A = m.p[0], B = m.p[1], C = m.p[2];
        const double x = r[0];
        double v12, dv12;
        if (x < 0) { const double e = std::exp(C * x); v12 = B * e; dv12 = B * C * e; }
        else { const double e = std::exp(-C * x); v12 = B * (2.0 - e); dv12 = B * C * e; }
        V[0] = A; V[3] = -A; V[1] = V[2] = v12;
        dV[1] = dV[2] = dv12;
        break;
    }
    case MODEL_SPIN_BOSON: {  // [[eps + c.x, Delta],[Delta, -eps - c.x]] + sum w^2 x^2 / 2
        const double eps = m.p[0], Delta = m.p[1];
        double cx = 0.0, v0 = 0.0;
        for (int j = 0; j < m.nbath; ++j) {
            cx += m.cbath[j] * r[j];
            v0 += 0.5 * m.omega[j] * m.omega[j] * r[j] * r[j];
            dV[(size_t)j * 4 + 0] = m.cbath[j];
            dV[(size_t)j * 4 + 3] = -m.cbath[j];
            dV0[j] = m.omega[j] * m.omega[j] * r[j];
        }
        V[0] = eps + cx; V[3] = -eps - cx; V[1] = V[2] = Delta;
        *V0 = v0;
        break;
    }
    case MODEL_DW_GAUSSIAN: {  // V11 = a x^4 - b x^2, V22 = V11 + off, V12 = C exp(-D x^2)
        const double a = m.p[0], b = m.p[1], off = m.p[2], C = m.p[3], Dd = m.p[4];
        const double x = r[0];
        const double w = a * x * x * x * x - b * x * x;
        const double dw = 4.0 * a * x * x * x - 2.0 * b * x;
        const double v12 = C * std::exp(-Dd * x * x);
        if (n == 1) { V[0] = w; dV[0] = dw; break; }
        V[0] = w; V[3] = w + off; V[1] = V[2] = v12;
        dV[0] = dw; dV[3] = dw; dV[1] = dV[2] = -2.0 * Dd * x * v12;
        break;
    }
    case MODEL_HARMONIC: {  // single state: V = m w^2 (x - x0)^2 / 2
        const double mass = m.p[0], w = m.p[1], x0 = m.p[2];
        for (int j = 0; j < D; ++j) {
            V[0] += 0.5 * mass * w * w * (r[j] - x0) * (r[j] - x0);
            dV[j] = mass * w * w * (r[j] - x0);
        }
        break;
    }
    case MODEL_DIATOM_SURFACE: {
        // Diatomic molecule above a surface, coordinates (x1, y1, z1, x2, y2, z2); bond length
        // b = |R1 - R2|, height Z = mu z1 + (1 - mu) z2.  Two Morse diabats and a Gaussian coupling:
        //   V11 = Dn (1 - e^{-an (b - bn)})^2 + Aw e^{-bw Z}
        //   V22 = Di (1 - e^{-ai (b - bi)})^2 + Ds (1 - e^{-as (Z - Zs)})^2 + Delta
        //   V12 = C e^{-g (Z - Zc)^2}
        const double* q = m.p;
        const double Dn = q[0], an = q[1], bn = q[2], Aw = q[3], bw = q[4], Di = q[5], ai = q[6], bi = q[7];
        const double Ds = q[8], as = q[9], Zs = q[10], Delta = q[11], C = q[12], g = q[13], Zc = q[14], mu = q[15];
        const double dx = r[0] - r[3], dy = r[1] - r[4], dz = r[2] - r[5];
        const double b = std::sqrt(dx * dx + dy * dy + dz * dz);
        const double Z = mu * r[2] + (1.0 - mu) * r[5];
        auto morse = [](double D, double a, double q0, double x, double& f, double& df) {
            const double e = std::exp(-a * (x - q0));
            f = D * (1.0 - e) * (1.0 - e);
            df = 2.0 * D * a * e * (1.0 - e);
        };
        double mn, dmn, mi, dmi, ms, dms;
        morse(Dn, an, bn, b, mn, dmn);
        morse(Di, ai, bi, b, mi, dmi);
        morse(Ds, as, Zs, Z, ms, dms);
        const double wall = Aw * std::exp(-bw * Z);
        const double v12 = C * std::exp(-g * (Z - Zc) * (Z - Zc));
        V[0] = mn + wall; V[3] = mi + ms + Delta; V[1] = V[2] = v12;
        const double db11 = dmn, dZ11 = -bw * wall, db22 = dmi, dZ22 = dms, dZ12 = -2.0 * g * (Z - Zc) * v12;
        const double ub[6] = {dx / b, dy / b, dz / b, -dx / b, -dy / b, -dz / b};
        const double uZ[6] = {0.0, 0.0, mu, 0.0, 0.0, 1.0 - mu};
        for (int j = 0; j < 6; ++j) {
            dV[(size_t)j * 4 + 0] = db11 * ub[j] + dZ11 * uZ[j];
            dV[(size_t)j * 4 + 3] = db22 * ub[j] + dZ22 * uZ[j];
            dV[(size_t)j * 4 + 1] = dV[(size_t)j * 4 + 2] = dZ12 * uZ[j];
        }
        break;
    }
    case MODEL_MORSE_N: {
        // n-state Morse model of Coker and Xiao (J. Chem. Phys. 102, 496 (1995); NQCModels'
        // ThreeStateMorse for n = 3), one coordinate:
        //   V_ii = D_i (1 - e^{-beta_i (x - r_i)})^2 + c_i,   V_ij = A_ij e^{-a_ij (x - r_ij)^2}
        const double x = r[0];
        const double* q = m.ext;
        for (int i = 0; i < n; ++i) {
            const double Dm = q[4 * i], be = q[4 * i + 1], r0 = q[4 * i + 2], c = q[4 * i + 3];
            const double e = std::exp(-be * (x - r0));
            V[i * n + i] = Dm * (1.0 - e) * (1.0 - e) + c;
            dV[i * n + i] = 2.0 * Dm * be * e * (1.0 - e);
        }
        int pidx = 0;
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j, ++pidx) {
                const double* c = q + 4 * n + 3 * pidx;
                const double v = c[0] * std::exp(-c[1] * (x - c[2]) * (x - c[2]));
                V[i * n + j] = V[j * n + i] = v;
                dV[i * n + j] = dV[j * n + i] = -2.0 * c[1] * (x - c[2]) * v;
            }
        break;
    }
    case MODEL_ANDERSON_HOLSTEIN_1: {
        // One-mode Anderson-Holstein impurity (the CME / BCME model of test/Dynamics/cme.jl:27-55, Dou,
        // Nitzan, Subotnik, J. Chem. Phys. 142, 084110 (2015)): V11 = k x^2/2, V22 = V11 + Ed + c x,
        // V12 = const.  params: k, Ed, c, V12.
        const double k = m.p[0], Ed = m.p[1], c = m.p[2], v12 = m.p[3];
        const double x = r[0];
        V[0] = 0.5 * k * x * x; V[3] = V[0] + Ed + c * x; V[1] = V[2] = v12;
        dV[0] = k * x; dV[3] = k * x + c; dV[1] = dV[2] = 0.0;
        break;
    }
    default: break;
    }
}

OR_API void oracle_model_eval(const OModelDesc* m, const double* r, double* V, double* dV, double* V0, double* dV0) {
    model_eval(*m, r, V, dV, V0, dV0);
}

// ---------------------------------------------------------------------------
// Real symmetric eigen-decomposition (stands in for NQCCalculators' eigen; the
// reference calls LAPACK through it).  n == 2: closed form via the rotation
// angle; n > 2: cyclic Jacobi.  Eigenvalues ascending.  Sign convention of a
// fresh decomposition: the largest-magnitude component of each column is
// positive (ties -> first).  Subsequent decompositions flip columns whose
// overlap with the previous eigenvector is negative (eigenvector tracking).
// ---------------------------------------------------------------------------
static void canonical_sign(int n, double* U) {
    for (int j = 0; j < n; ++j) {
        int imax = 0; double amax = -1.0;
        for (int i = 0; i < n; ++i) { const double a = std::fabs(U[i * n + j]); if (a > amax) { amax = a; imax = i; } }
        if (U[imax * n + j] < 0) for (int i = 0; i < n; ++i) U[i * n + j] = -U[i * n + j];
    }
}

static void track_sign(int n, double* U, const double* Uprev) {
    for (int j = 0; j < n; ++j) {
        double s = 0.0;
        for (int i = 0; i < n; ++i) s += U[i * n + j] * Uprev[i * n + j];
        if (s < 0) for (int i = 0; i < n; ++i) U[i * n + j] = -U[i * n + j];
    }
}

static void eigen2(const double* H, double* E, double* U) {
    const double a = H[0], b = H[1], d = H[3];
    const double mean = 0.5 * (a + d), half = 0.5 * (a - d);
    const double rad = std::hypot(half, b);
    const double th = 0.5 * std::atan2(b, half);
    const double c = std::cos(th), s = std::sin(th);
    E[0] = mean - rad; E[1] = mean + rad;
    U[0] = -s; U[2] = c;   // column 0: eigenvector of E0
    U[1] = c;  U[3] = s;   // column 1: eigenvector of E1
}

static void eigen_jacobi(int n, const double* H, double* E, double* U) {
    std::vector<double> A(H, H + n * n);
    for (int i = 0; i < n * n; ++i) U[i] = 0.0;
    for (int i = 0; i < n; ++i) U[i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) { const double a2 = A[i * n + j] * A[i * n + j]; tot += a2; if (i != j) off += a2; }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int p = 0; p < n - 1; ++p) for (int q = p + 1; q < n; ++q) {
            const double apq = A[p * n + q];
            if (apq == 0.0) continue;
            const double theta = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
            const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
            const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
            for (int k = 0; k < n; ++k) {  // A <- A J (columns p, q)
                const double akp = A[k * n + p], akq = A[k * n + q];
                A[k * n + p] = c * akp - s * akq; A[k * n + q] = s * akp + c * akq;
            }
            for (int k = 0; k < n; ++k) {  // A <- J^T A (rows p, q)
                const double apk = A[p * n + k], aqk = A[q * n + k];
                A[p * n + k] = c * apk - s * aqk; A[q * n + k] = s * apk + c * aqk;
            }
            for (int k = 0; k < n; ++k) {
                const double ukp = U[k * n + p], ukq = U[k * n + q];
                U[k * n + p] = c * ukp - s * ukq; U[k * n + q] = s * ukp + c * ukq;
            }
        }
    }
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return A[x * n + x] < A[y * n + y]; });
    std::vector<double> Us(U, U + n * n);
    for (int j = 0; j < n; ++j) { E[j] = A[idx[j] * n + idx[j]]; for (int i = 0; i < n; ++i) U[i * n + j] = Us[i * n + idx[j]]; }
}

static void eigen_sym(int n, const double* H, double* E, double* U, const double* Uprev) {
    if (n == 1) { E[0] = H[0]; U[0] = 1.0; return; }
    if (n == 2) eigen2(H, E, U); else eigen_jacobi(n, H, E, U);
    if (Uprev) track_sign(n, U, Uprev); else canonical_sign(n, U);
}

OR_API void oracle_eigen_sym(int32_t n, const double* H, double* E, double* U, const double* Uprev) {
    eigen_sym(n, H, E, U, Uprev);
}

// Adiabatic derivative G_I = U^T dV_I U and nonadiabatic coupling
// d_I[i][j] = G_I[i][j] / (E_j - E_i) (i != j), 0 on the diagonal.
// (NQCCalculators get_adiabatic_derivative / get_nonadiabatic_coupling, called
// at bab_electronics.jl:83-84 and fssh.jl:68.)
static void adiabatic_quantities(int n, int D, const double* E, const double* U, const double* dV, double* G, double* d) {
    std::vector<double> tmp(n * n);
    for (int I = 0; I < D; ++I) {
        const double* A = dV + (size_t)I * n * n;
        double* g = G + (size_t)I * n * n;
        double* dd = d + (size_t)I * n * n;
        for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) {  // tmp = A U
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += A[i * n + k] * U[k * n + j];
            tmp[i * n + j] = s;
        }
        for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) {  // g = U^T tmp
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += U[k * n + i] * tmp[k * n + j];
            g[i * n + j] = s;
        }
        for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j)
            dd[i * n + j] = (i == j) ? 0.0 : g[i * n + j] / (E[j] - E[i]);
    }
}

OR_API void oracle_adiabatic_quantities(int32_t n, int32_t D, const double* E, const double* U, const double* dV, double* G, double* d) {
    adiabatic_quantities(n, D, E, U, dV, G, d);
}

// ---------------------------------------------------------------------------
// Density-matrix electronic propagation: electronic_dynamics.jl:15-79,104-140,
// density_matrix_dynamics.jl:29-33, integrated by Tsit5 (OrdinaryDiffEq 6.103,
// coefficients from Tsitouras, Comput. Math. Appl. 62, 770 (2011)) with 5 fixed
// sub-steps of dt/5 (bab_electronics.jl:40-43), the last one clipped to t+dt.
// ---------------------------------------------------------------------------
namespace tsit5 {
static const double c1 = 0.161, c2 = 0.327, c3 = 0.9, c4 = 0.9800255409045097;
static const double a21 = 0.161;
static const double a31 = -0.008480655492356989, a32 = 0.335480655492357;
static const double a41 = 2.897153057105493, a42 = -6.359448489975075, a43 = 4.3622954328695815;
static const double a51 = 5.325864828439257, a52 = -11.748883564062828, a53 = 7.4955393428898365, a54 = -0.09249506636175525;
static const double a61 = 5.86145544294642, a62 = -12.92096931784711, a63 = 8.159367898576159, a64 = -0.071584973281401, a65 = -0.028269050394068383;
static const double a71 = 0.09646076681806523, a72 = 0.01, a73 = 0.4798896504144996, a74 = 1.379008574103742, a75 = -3.290069515436081, a76 = 2.324710524099774;
}

// DoubleBuffer of electronic_dynamics.jl:15-53: "cur" holds (E, W, t) at the
// start of the nuclear step, "next" at its end; W = sum_I d_I v_I.
struct EBuf {
    std::vector<double> E, W;
    double t = 0.0;
};

struct EParams {
    int n;
    EBuf cur, next;
};

// update_parameters! (electronic_dynamics.jl:38-53): swap, then fill next.
static void update_parameters(EParams& P, const double* E, const double* d, const double* v, int D, double t) {
    std::swap(P.cur, P.next);
    const int n = P.n;
    for (int i = 0; i < n; ++i) P.next.E[i] = E[i];
    std::fill(P.next.W.begin(), P.next.W.end(), 0.0);
    for (int I = 0; I < D; ++I)
        for (int J = 0; J < n * n; ++J) P.next.W[J] += d[(size_t)I * n * n + J] * v[I];
    P.next.t = t;
}

// f! of electronic_dynamics.jl:106-116: A = -i W(tau); diag(A) = E(tau); du = -i(A u - u A).
static void density_rhs(const EParams& P, const cd* u, double tau, cd* du, cd* A, cd* tmp) {
    const int n = P.n;
    double loc = (tau - P.cur.t) / (P.next.t - P.cur.t);
    if (std::isnan(loc)) loc = 0.0;
    for (int J = 0; J < n * n; ++J) {
        const double w = P.cur.W[J] + (P.next.W[J] - P.cur.W[J]) * loc;
        A[J] = cd(0.0, -w);  // (w + 0i) * (-1im)
    }
    for (int i = 0; i < n; ++i) A[i * n + i] = cd(P.cur.E[i] + (P.next.E[i] - P.cur.E[i]) * loc, 0.0);
    for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) {  // tmp = u A
        cd s = 0.0;
        for (int k = 0; k < n; ++k) s += u[i * n + k] * A[k * n + j];
        tmp[i * n + j] = s;
    }
    for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) {  // du = A u - tmp, then * -i
        cd s = 0.0;
        for (int k = 0; k < n; ++k) s += A[i * n + k] * u[k * n + j];
        const cd c = s - tmp[i * n + j];
        du[i * n + j] = cd(c.imag(), -c.real());
    }
}

static void tsit5_step(const EParams& P, cd* u, double tau, double h, std::vector<cd>& ws) {
    using namespace tsit5;
    const int nn = P.n * P.n;
    cd* k1 = ws.data(); cd* k2 = k1 + nn; cd* k3 = k2 + nn; cd* k4 = k3 + nn; cd* k5 = k4 + nn; cd* k6 = k5 + nn;
    cd* tmp = k6 + nn; cd* A = tmp + nn; cd* t2 = A + nn;
    density_rhs(P, u, tau, k1, A, t2);
    for (int J = 0; J < nn; ++J) tmp[J] = u[J] + h * (a21 * k1[J]);
    density_rhs(P, tmp, tau + c1 * h, k2, A, t2);
    for (int J = 0; J < nn; ++J) tmp[J] = u[J] + h * (a31 * k1[J] + a32 * k2[J]);
    density_rhs(P, tmp, tau + c2 * h, k3, A, t2);
    for (int J = 0; J < nn; ++J) tmp[J] = u[J] + h * (a41 * k1[J] + a42 * k2[J] + a43 * k3[J]);
    density_rhs(P, tmp, tau + c3 * h, k4, A, t2);
    for (int J = 0; J < nn; ++J) tmp[J] = u[J] + h * (a51 * k1[J] + a52 * k2[J] + a53 * k3[J] + a54 * k4[J]);
    density_rhs(P, tmp, tau + c4 * h, k5, A, t2);
    for (int J = 0; J < nn; ++J) tmp[J] = u[J] + h * (a61 * k1[J] + a62 * k2[J] + a63 * k3[J] + a64 * k4[J] + a65 * k5[J]);
    density_rhs(P, tmp, tau + h, k6, A, t2);
    for (int J = 0; J < nn; ++J)
        u[J] = u[J] + h * (a71 * k1[J] + a72 * k2[J] + a73 * k3[J] + a74 * k4[J] + a75 * k5[J] + a76 * k6[J]);
}

// step!(electronic_integrator, dt, true) after set_ut!(.., sigma, t) (bab_electronics.jl:87-90).
static void propagate_density(const EParams& P, cd* u, double t, double dt, int nsub, std::vector<cd>& ws) {
    const double h = dt / nsub;
    const double tend = t + dt;
    double tau = t;
    for (int s = 0; s < nsub; ++s) {
        double hs = std::min(h, tend - tau);
        tsit5_step(P, u, tau, hs, ws);
        tau = (s == nsub - 1) ? tend : tau + hs;
    }
}

OR_API void oracle_propagate_density(int32_t n, double* sig_re, double* sig_im,
                                     const double* E0, const double* W0, double t0,
                                     const double* E1, const double* W1, double t1,
                                     double t, double dt, int32_t nsub) {
    EParams P; P.n = n;
    P.cur.E.assign(E0, E0 + n); P.cur.W.assign(W0, W0 + n * n); P.cur.t = t0;
    P.next.E.assign(E1, E1 + n); P.next.W.assign(W1, W1 + n * n); P.next.t = t1;
    std::vector<cd> u(n * n), ws(9 * n * n);
    for (int J = 0; J < n * n; ++J) u[J] = cd(sig_re[J], sig_im[J]);
    propagate_density(P, u.data(), t, dt, nsub, ws);
    for (int J = 0; J < n * n; ++J) { sig_re[J] = u[J].real(); sig_im[J] = u[J].imag(); }
}

// ---------------------------------------------------------------------------
// Surface hopping: fssh.jl:96-121, surface_hopping.jl:64-168.
// ---------------------------------------------------------------------------
// fewest_switches_probability! (fssh.jl:96-108): clamp then cumsum.
static void fssh_probability(int n, int D, const double* v, const cd* sig, int s, const double* d, double dt, double* prob) {
    for (int m = 0; m < n; ++m) prob[m] = 0.0;
    for (int m = 0; m < n; ++m) {
        if (m == s) continue;
        const double ratio = (sig[m * n + s] / sig[s * n + s]).real();
        for (int I = 0; I < D; ++I) prob[m] += 2.0 * v[I] * ratio * d[(size_t)I * n * n + s * n + m] * dt;
    }
    for (int m = 0; m < n; ++m) prob[m] = std::min(std::max(prob[m], 0.0), 1.0);
    for (int m = 1; m < n; ++m) prob[m] += prob[m - 1];
}

// select_new_state (fssh.jl:110-121): first m != s with cumulative prob > r.
static int fssh_select(int n, const double* prob, int s, double rnd) {
    for (int m = 0; m < n; ++m) {
        if (m == s) continue;
        if (prob[m] > rnd) return m;
    }
    return s;
}

enum { RESCALE_STANDARD = 0, RESCALE_VINVERSION = 1, RESCALE_OFF = 2 };

// rescale_velocity! (surface_hopping.jl:64-99) for a flat DOF vector.
// Returns 1 if the hop is accepted.  d = coupling[new, old] (surface_hopping.jl:106-108).
static int rescale_velocity(int n, int D, int rescaling, double* v, const double* mass, const double* E,
                            const double* dcoup, int new_state, int old_state) {
    if (rescaling == RESCALE_OFF) return 1;
    std::vector<double> d(D);
    for (int I = 0; I < D; ++I) d[I] = dcoup[(size_t)I * n * n + new_state * n + old_state];
    double a = 0.0, b = 0.0;
    for (int I = 0; I < D; ++I) { a += d[I] * d[I] / mass[I]; b += d[I] * v[I]; }
    a /= 2.0;
    const double c = E[new_state] - E[old_state];
    const double disc = b * b - 4.0 * a * c;
    if (disc < 0) {
        if (rescaling == RESCALE_VINVERSION) {  // frustrated_hop_invert_velocity! (:155-164)
            double nrm = 0.0;
            for (int I = 0; I < D; ++I) nrm += d[I] * d[I];
            nrm = std::sqrt(nrm);
            double g = 0.0;
            for (int I = 0; I < D; ++I) g += v[I] * (d[I] / nrm);
            for (int I = 0; I < D; ++I) v[I] -= 2.0 * g * (d[I] / nrm);
        }
        return 0;
    }
    const double root = std::sqrt(disc);
    const double gamma = (b < 0) ? (b + root) / (2.0 * a) : (b - root) / (2.0 * a);
    for (int I = 0; I < D; ++I) v[I] -= gamma * d[I] / mass[I];
    return 1;
}

OR_API void oracle_fssh_probability(int32_t n, int32_t D, const double* v, const double* sig_re, const double* sig_im,
                                    int32_t s, const double* d, double dt, double* prob) {
    std::vector<cd> sig(n * n);
    for (int J = 0; J < n * n; ++J) sig[J] = cd(sig_re[J], sig_im[J]);
    fssh_probability(n, D, v, sig.data(), s, d, dt, prob);
}
OR_API int32_t oracle_fssh_select(int32_t n, const double* prob, int32_t s, double rnd) { return fssh_select(n, prob, s, rnd); }
OR_API int32_t oracle_rescale_velocity(int32_t n, int32_t D, int32_t rescaling, double* v, const double* mass,
                                       const double* E, const double* d, int32_t new_state, int32_t old_state) {
    return rescale_velocity(n, D, rescaling, v, mass, E, d, new_state, old_state);
}

// ---------------------------------------------------------------------------
// Per-trajectory state and the BABwithTsit5 step (bab_electronics.jl:48-91).
// ---------------------------------------------------------------------------
enum { METHOD_FSSH = 0, METHOD_EHRENFEST = 1 };

// ---------------------------------------------------------------------------
// Simulation cell of the cell callbacks (DynamicsUtils/callbacks.jl:4-22).  The cell type and its two
// functions live in NQCBase (Project.toml: NQCBase = "1", not in the reference tree); restated from its
// published definition: PeriodicCell(vectors) keeps the lattice vectors as the columns of `vectors`, its
// inverse and per-dimension periodicity; check_atoms_in_cell is true when every fractional coordinate
// inverse * r of every atom lies in [0, 1]; apply_cell_boundaries! maps each atom to inverse * r, takes
// Julia's mod(., 1) of the periodic components and multiplies back by `vectors`.  outside_cell is the
// negated check; CellBoundaryCallback applies the wrap and TerminateCellCallback terminates when an atom
// is outside after a step.
// ---------------------------------------------------------------------------
struct OCell {
    int32_t mode;      // 0 none, 1 CellBoundaryCallback, 2 TerminateCellCallback
    int32_t nd;        // dimensions per atom
    int32_t per[3];
    int32_t pad;
    double a[9];       // column-major, leading dimension 3
    double inv[9];
};

// Julia's mod(x, 1) for a Float64: r = rem(x, 1); r == 0 -> +0; r + 1 where r < 0
static double jl_mod1(double x) {
    const double r = std::fmod(x, 1.0);
    if (r == 0.0) return 0.0;
    return (r > 0.0) ? r : r + 1.0;
}

static bool ocell_outside(const OCell& c, const double* x, int D) {
    for (int at = 0; at + c.nd <= D; at += c.nd)
        for (int j = 0; j < c.nd; ++j) {
            double f = 0.0;
            for (int k = 0; k < c.nd; ++k) f += c.inv[j + 3 * k] * x[at + k];
            if (f > 1.0 || f < 0.0) return true;
        }
    return false;
}

static void ocell_wrap(const OCell& c, double* x, int D) {
    for (int at = 0; at + c.nd <= D; at += c.nd) {
        double f[3];
        for (int j = 0; j < c.nd; ++j) {
            double s = 0.0;
            for (int k = 0; k < c.nd; ++k) s += c.inv[j + 3 * k] * x[at + k];
            f[j] = c.per[j] ? jl_mod1(s) : s;
        }
        for (int j = 0; j < c.nd; ++j) {
            double s = 0.0;
            for (int k = 0; k < c.nd; ++k) s += c.a[j + 3 * k] * f[k];
            x[at + j] = s;
        }
    }
}

OR_API int32_t oracle_cell_apply(const OCell* c, int32_t D, double* x) {
    const bool out = ocell_outside(*c, x, D);
    if (out && c->mode == 1) ocell_wrap(*c, x, D);
    return out ? 1 : 0;
}

struct ORunDesc {
    int32_t method;
    int32_t nstates;
    int32_t ndof;
    int32_t rescaling;
    double dt;
    double t0;
    int64_t nsteps;       // maximum number of nuclear steps
    double term_x;        // terminate when |r[term_dof]| > term_x (disabled when <= 0)
    int32_t save_every;   // save slot every this many steps (slot 0 = t0); 0 = none
    int32_t nsave;        // number of save slots per trajectory
    uint64_t key;
    int64_t traj_offset;  // global id of trajectory 0 of this call
    int32_t nsub;         // electronic sub-steps per nuclear step (5)
    int32_t term_dof;     // flattened dof index of the termination predicate
    int32_t term_kind;    // 0: |x| > term_x when term_x > 0; 2: x > term_x; 3: x < term_x
    int32_t pad;
    OCell cell;           // cell callback after the hop callback (mode 0: none)
};

// TerminatingCallback(func) (DynamicsUtils/callbacks.jl:29) for the predicates on one coordinate.
static bool term_hit(const ORunDesc& R, double x) {
    if (R.term_kind == 2) return x > R.term_x;
    if (R.term_kind == 3) return x < R.term_x;
    return R.term_x > 0 && std::fabs(x) > R.term_x;
}

struct Cache {
    int n, D;
    std::vector<double> V, dV, dV0, E, U, Uprev, G, d;
    double V0 = 0.0;
    bool have_prev = false;
    void init(int n_, int D_) {
        n = n_; D = D_;
        V.assign(n * n, 0); dV.assign((size_t)D * n * n, 0); dV0.assign(D, 0);
        E.assign(n, 0); U.assign(n * n, 0); Uprev.assign(n * n, 0);
        G.assign((size_t)D * n * n, 0); d.assign((size_t)D * n * n, 0);
        have_prev = false;
    }
    // update_cache!(cache, r): model, eigen with tracking, U^T dV U, NAC.
    void update(const OModelDesc& m, const double* r) {
        model_eval(m, r, V.data(), dV.data(), &V0, dV0.data());
        if (have_prev) Uprev = U;
        eigen_sym(n, V.data(), E.data(), U.data(), have_prev ? Uprev.data() : nullptr);
        have_prev = true;
        adiabatic_quantities(n, D, E.data(), U.data(), dV.data(), G.data(), d.data());
    }
};

// FSSH acceleration (fssh.jl:67-74): -dE_s/dR / m, with the state-independent
// derivative of system-bath models added (QuantumModel convention).
static void accel_fssh(const Cache& c, int s, const double* mass, double* k) {
    const int n = c.n;
    for (int I = 0; I < c.D; ++I) k[I] = -(c.G[(size_t)I * n * n + s * n + s] + c.dV0[I]) / mass[I];
}

// Ehrenfest acceleration (ehrenfest.jl:50-68): -(dV0 + sum_nm G_nm Re sigma_nm) / m.
static void accel_ehrenfest(const Cache& c, const cd* sig, const double* mass, double* k) {
    const int n = c.n;
    for (int I = 0; I < c.D; ++I) {
        double f = -c.dV0[I];
        for (int m = 0; m < n; ++m) for (int nn = 0; nn < n; ++nn)
            f -= c.G[(size_t)I * n * n + nn * n + m] * sig[nn * n + m].real();
        k[I] = f / mass[I];
    }
}

// Diabatic population estimators: FSSH (fssh.jl:132-142) and Ehrenfest (ehrenfest.jl:75-83).
static void diabatic_population(int method, int n, const double* U, const cd* sig, int s, double* out) {
    std::vector<double> S(n * n);
    for (int J = 0; J < n * n; ++J) S[J] = sig[J].real();
    if (method == METHOD_FSSH) {
        for (int i = 0; i < n; ++i) S[i * n + i] = 0.0;
        S[s * n + s] = 1.0;
    }
    for (int i = 0; i < n; ++i) {  // diag(U S U^T)
        double p = 0.0;
        for (int a = 0; a < n; ++a) for (int b = 0; b < n; ++b) p += U[i * n + a] * S[a * n + b] * U[i * n + b];
        out[i] = p;
    }
}

static void adiabatic_population(int method, int n, const cd* sig, int s, double* out) {
    if (method == METHOD_FSSH) { for (int i = 0; i < n; ++i) out[i] = (i == s) ? 1.0 : 0.0; }
    else { for (int i = 0; i < n; ++i) out[i] = sig[i * n + i].real(); }
}

struct TrajOut {
    double* r; double* v; double* sig_re; double* sig_im;
    int32_t* state; int64_t* steps; int32_t* nhops;
    double* pop_ad; double* pop_di; int32_t* save_state;
    double* ekin = nullptr; double* epot = nullptr;  // per save slot, or null
};

static void run_one(const ORunDesc& R, const OModelDesc& m, const double* mass, int64_t i, uint64_t gid,
                    const double* r0, const double* v0, const double* sr0, const double* si0, int32_t st0,
                    const TrajOut& O) {
    const int n = R.nstates, D = R.ndof, nn = n * n;
    std::vector<double> r(r0, r0 + D), v(v0, v0 + D), k(D), vtmp(D);
    std::vector<cd> sig(nn), ws(9 * nn);
    for (int J = 0; J < nn; ++J) sig[J] = cd(sr0[J], si0[J]);
    Cache c; c.init(n, D);
    c.update(m, r.data());  // initialize! (bab_electronics.jl:48-59)
    int s = st0;
    if (R.method == METHOD_FSSH && s < 0) {  // fssh.jl:54 sample(Weights(diag(real.(sigma))))
        double tot = 0.0;
        for (int a = 0; a < n; ++a) tot += sig[a * n + a].real();
        const double u = uniform01(gid, 0, STREAM_STATE, R.key) * tot;
        double cum = 0.0; s = n - 1;
        for (int a = 0; a < n; ++a) { cum += sig[a * n + a].real(); if (u < cum) { s = a; break; } }
    }
    if (R.method == METHOD_FSSH) accel_fssh(c, s, mass, k.data());
    else accel_ehrenfest(c, sig.data(), mass, k.data());
    EParams P; P.n = n;
    P.cur.E.assign(n, 0.0); P.cur.W.assign(nn, 0.0); P.cur.t = 0.0;  // Q1: zero buffers at t = 0
    P.next = P.cur;
    double t = R.t0;
    int32_t nhops = 0;
    auto save = [&](int slot) {
        if (slot < 0 || slot >= R.nsave) return;
        if (O.pop_ad) adiabatic_population(R.method, n, sig.data(), s, O.pop_ad + ((size_t)i * R.nsave + slot) * n);
        if (O.pop_di) diabatic_population(R.method, n, c.U.data(), sig.data(), s, O.pop_di + ((size_t)i * R.nsave + slot) * n);
        if (O.save_state) O.save_state[(size_t)i * R.nsave + slot] = s;
        if (O.ekin) {  // classical_kinetic_energy; classical_potential_energy (fssh.jl:150-154, ehrenfest.jl:85-95)
            double ek = 0.0;
            for (int I = 0; I < D; ++I) ek += 0.5 * mass[I] * v[I] * v[I];
            double ep;
            if (R.method == METHOD_FSSH) ep = c.E[s];
            else {
                ep = c.V0;
                for (int a = 0; a < n; ++a) ep += sig[(size_t)a * n + a].real() * c.E[a];
            }
            O.ekin[(size_t)i * R.nsave + slot] = ek;
            O.epot[(size_t)i * R.nsave + slot] = ep;
        }
    };
    save(0);
    int64_t step = 0;
    std::vector<double> prob(n);
    while (step < R.nsteps) {
        // TerminatingCallback is a DiscreteCallback: its condition is evaluated after each step and
        // not at t0 (DynamicsUtils/callbacks.jl:15-29), so a trajectory that starts past the
        // boundary takes one step
        if (step > 0 && (term_hit(R, r[R.term_dof]) || (R.cell.mode == 2 && ocell_outside(R.cell, r.data(), D)))) break;
        std::vector<cd> sigprev = sig;
        for (int I = 0; I < D; ++I) vtmp[I] = v[I] + (R.dt / 2) * k[I];   // step_B!
        for (int I = 0; I < D; ++I) r[I] = r[I] + R.dt * vtmp[I];         // step_A!
        c.update(m, r.data());
        if (R.method == METHOD_FSSH) accel_fssh(c, s, mass, k.data());
        else accel_ehrenfest(c, sigprev.data(), mass, k.data());
        for (int I = 0; I < D; ++I) v[I] = vtmp[I] + (R.dt / 2) * k[I];
        update_parameters(P, c.E.data(), c.d.data(), v.data(), D, t + R.dt);
        propagate_density(P, sig.data(), t, R.dt, R.nsub, ws);
        t = t + R.dt;
        // A save time's values are those before the step's discrete callbacks: DiffEqBase's
        // apply_discrete_callback! saves the saveat points due before it calls affect!, and
        // handle_callbacks! saves after the callbacks only when none fired, so either way the slot holds
        // the state before the hop (a hop at a save time shows in the next slot).
        if (R.save_every > 0 && (step + 1) % R.save_every == 0) save((int)((step + 1) / R.save_every));
        if (R.method == METHOD_FSSH) {  // HoppingCallback after the step (surface_hopping.jl:2-16)
            fssh_probability(n, D, v.data(), sig.data(), s, c.d.data(), R.dt, prob.data());
            const double rnd = uniform01(gid, (uint32_t)step, STREAM_HOP, R.key);
            const int ns = fssh_select(n, prob.data(), s, rnd);
            if (ns != s) {
                if (rescale_velocity(n, D, R.rescaling, v.data(), mass, c.E.data(), c.d.data(), ns, s)) { s = ns; ++nhops; }
            }
            // Q4: k is not recomputed after an accepted hop.
        }
        // CellBoundaryCallback after the hop callback (the problem's callbacks precede the user's); the
        // BABwithTsit5 cache keeps the step-end acceleration (bab_electronics.jl:61-74)
        if (R.cell.mode == 1 && ocell_outside(R.cell, r.data(), D)) ocell_wrap(R.cell, r.data(), D);
        ++step;
    }
    for (int I = 0; I < D; ++I) { O.r[(size_t)i * D + I] = r[I]; O.v[(size_t)i * D + I] = v[I]; }
    for (int J = 0; J < nn; ++J) { O.sig_re[(size_t)i * nn + J] = sig[J].real(); O.sig_im[(size_t)i * nn + J] = sig[J].imag(); }
    O.state[i] = s; O.steps[i] = step;
    if (O.nhops) O.nhops[i] = nhops;
}

template <class F>
static void parallel_for(int64_t n, int nthreads, F f) {
    if (nthreads <= 1 || n < 2) { for (int64_t i = 0; i < n; ++i) f(i); return; }
    std::atomic<int64_t> next(0);
    std::vector<std::thread> th;
    const int64_t chunk = std::max<int64_t>(1, n / (nthreads * 16));
    for (int t = 0; t < nthreads; ++t) th.emplace_back([&]() {
        for (;;) {
            const int64_t b = next.fetch_add(chunk);
            if (b >= n) break;
            const int64_t e = std::min(n, b + chunk);
            for (int64_t i = b; i < e; ++i) f(i);
        }
    });
    for (auto& x : th) x.join();
}

// Ensemble driver (the EnsembleThreads analogue): each trajectory independent.  The Philox
// counters use the global trajectory id traj_ids[i] (or traj_offset + i when traj_ids is null).
OR_API int32_t oracle_run_ensemble(const ORunDesc* R, const OModelDesc* models, const double* mass, int64_t ntraj,
                                   const int32_t* model_id, const double* r0, const double* v0,
                                   const double* sr0, const double* si0, const int32_t* st0,
                                   double* r, double* v, double* sr, double* si, int32_t* state, int64_t* steps,
                                   int32_t* nhops, double* pop_ad, double* pop_di, int32_t* save_state, int32_t nthreads,
                                   const int64_t* traj_ids) {
    const int D = R->ndof, nn = R->nstates * R->nstates;
    TrajOut O{r, v, sr, si, state, steps, nhops, pop_ad, pop_di, save_state};
    parallel_for(ntraj, nthreads, [&](int64_t i) {
        const int mid = model_id ? model_id[i] : 0;
        const uint64_t gid = traj_ids ? (uint64_t)traj_ids[i] : (uint64_t)(R->traj_offset + i);
        run_one(*R, models[mid], mass, i, gid, r0 + (size_t)i * D, v0 + (size_t)i * D, sr0 + (size_t)i * nn, si0 + (size_t)i * nn,
                st0 ? st0[i] : 0, O);
    });
    return 0;
}

// oracle_run_ensemble with the kinetic and classical potential energy of every save slot.
OR_API int32_t oracle_run_ensemble_energies(const ORunDesc* R, const OModelDesc* models, const double* mass, int64_t ntraj,
                                            const int32_t* model_id, const double* r0, const double* v0,
                                            const double* sr0, const double* si0, const int32_t* st0,
                                            double* r, double* v, double* sr, double* si, int32_t* state, int64_t* steps,
                                            int32_t* nhops, double* pop_ad, double* pop_di, int32_t* save_state,
                                            double* ekin, double* epot, int32_t nthreads) {
    const int D = R->ndof, nn = R->nstates * R->nstates;
    TrajOut O{r, v, sr, si, state, steps, nhops, pop_ad, pop_di, save_state, ekin, epot};
    parallel_for(ntraj, nthreads, [&](int64_t i) {
        const int mid = model_id ? model_id[i] : 0;
        run_one(*R, models[mid], mass, i, (uint64_t)(R->traj_offset + i), r0 + (size_t)i * D, v0 + (size_t)i * D,
                sr0 + (size_t)i * nn, si0 + (size_t)i * nn, st0 ? st0[i] : 0, O);
    });
    return 0;
}

// Initial-condition sampling used by the workloads (stream 0 normals):
// position x ~ N(mu_x, sx) per DOF, velocity v ~ N(mu_v, sv) per DOF.
// Draw j of trajectory g is philox counter (g, j, STREAM_INIT): DOF I uses
// counter j = I and takes z0 for r, z1 for v.
OR_API void oracle_sample_gaussian(int64_t ntraj, int64_t traj_offset, int32_t D, const double* mu_r, const double* sd_r,
                                   const double* mu_v, const double* sd_v, uint64_t key, double* r, double* v,
                                   const int64_t* traj_ids) {
    for (int64_t i = 0; i < ntraj; ++i) {
        const uint64_t g = traj_ids ? (uint64_t)traj_ids[i] : (uint64_t)(traj_offset + i);
        for (int I = 0; I < D; ++I) {
            double z0, z1;
            normal2(g, (uint32_t)I, STREAM_INIT, key, &z0, &z1);
            r[(size_t)i * D + I] = mu_r[(size_t)i * D + I] + sd_r[(size_t)i * D + I] * z0;
            v[(size_t)i * D + I] = mu_v[(size_t)i * D + I] + sd_v[(size_t)i * D + I] * z1;
        }
    }
}

// Adiabatic initial density from a diabatic pure state (density_matrix_dynamics.jl:37-46,64-75):
// sigma_ad = U^T rho_diab U with U from the eigen-decomposition at r0.
OR_API void oracle_diabatic_to_adiabatic(const OModelDesc* m, int64_t ntraj, const double* r0, int32_t diabatic_state,
                                         double* sig_re, double* sig_im) {
    const int n = m->nstates, D = m->ndof, nn = n * n;
    for (int64_t i = 0; i < ntraj; ++i) {
        Cache c; c.init(n, D);
        c.update(*m, r0 + (size_t)i * D);
        for (int a = 0; a < n; ++a) for (int b = 0; b < n; ++b) {
            sig_re[(size_t)i * nn + a * n + b] = c.U[diabatic_state * n + a] * c.U[diabatic_state * n + b];
            sig_im[(size_t)i * nn + a * n + b] = 0.0;
        }
    }
}

// ---------------------------------------------------------------------------
// Ring polymers: ring_polymer.jl:6-107 (frequencies, Cayley propagator), steps.jl:10-39
// (step_C!), bcb.jl:81-116 (BCB, RPMD), bcb_electronics.jl:53-97 (BCBwithTsit5: RPSH and
// RP-Ehrenfest with centroid electronics), rpsh.jl:30-50 (bead-broadcast rescaling).
// One degree of freedom per bead.  RingPolymerArrays (external) supplies the orthogonal
// normal-mode matrix; any orthonormal basis whose column k spans the spring eigenspace of
// w_k = 2 w_n sin(k pi / P) gives the same bead-space B-C-B step, and the real DFT basis
// is used here.
// ---------------------------------------------------------------------------
enum { RP_CLASSICAL = 0, RP_FSSH = 1, RP_EHRENFEST = 2 };

struct ORPDesc {
    int32_t method;     // RP_*
    int32_t nstates;
    int32_t nbeads;
    int32_t rescaling;
    double dt;
    double t0;
    int64_t nsteps;
    int32_t save_every;
    int32_t nsave;
    uint64_t key;
    double temperature; // w_n = nbeads * temperature
    int32_t nsub;
    int32_t pad;
};

// Column k of the real DFT basis (j = bead index).
static void normal_mode_matrix(int P, double* U) {
    for (int j = 0; j < P; ++j)
        for (int k = 0; k < P; ++k) {
            double u;
            if (k == 0) u = 1.0 / std::sqrt((double)P);
            else if (2 * k < P) u = std::sqrt(2.0 / P) * std::cos(2.0 * M_PI * j * k / P);
            else if (2 * k == P) u = ((j % 2) ? -1.0 : 1.0) / std::sqrt((double)P);
            else u = std::sqrt(2.0 / P) * std::sin(2.0 * M_PI * j * k / P);
            U[j * P + k] = u;
        }
}

static double matsubara(int P, double wn, int k) { return 2.0 * wn * std::sin(k * M_PI / P); }

// cay(dt A) = (I - dt A/2)^{-1} (I + dt A/2), A = [0 1; -w^2 0] (ring_polymer.jl:71-82), half = false.
static void cayley(double w, double dt, double C[4]) {
    const double u = dt * dt * w * w / 4.0;
    const double s = 1.0 / (1.0 + u);
    C[0] = (1.0 - u) * s; C[1] = dt * s; C[2] = -dt * w * w * s; C[3] = (1.0 - u) * s;
}

OR_API void oracle_rp_normal_modes(int32_t P, double wn, double dt, double* U, double* w, double* C) {
    normal_mode_matrix(P, U);
    for (int k = 0; k < P; ++k) { w[k] = matsubara(P, wn, k); cayley(w[k], dt, C + 4 * k); }
}

struct RPTrajOut {
    double* r; double* v; double* sig_re; double* sig_im; int32_t* state; int32_t* nhops;
    double* centroid; double* pop_ad; double* pop_di; int32_t* save_state;
};

static void rp_run_one(const ORPDesc& R, const OModelDesc& m, double mass, int64_t i, uint64_t gid,
                       const double* r0, const double* v0, const double* sr0, const double* si0, int32_t st0,
                       const RPTrajOut& O) {
    const int P = R.nbeads, n = R.nstates, nn = n * n;
    const double wn = P * R.temperature;
    std::vector<double> U(P * P), w(P), C(4 * P);
    normal_mode_matrix(P, U.data());
    for (int k = 0; k < P; ++k) { w[k] = matsubara(P, wn, k); cayley(w[k], R.dt, &C[4 * k]); }
    std::vector<double> r(r0, r0 + P), v(v0, v0 + P), a(P), vtmp(P), rn(P), vn(P);
    std::vector<cd> sig(nn), ws(9 * nn);
    for (int J = 0; J < nn; ++J) sig[J] = sr0 ? cd(sr0[J], si0[J]) : cd(0.0, 0.0);
    std::vector<Cache> bead(P);
    for (int b = 0; b < P; ++b) bead[b].init(n, 1);
    Cache cen; cen.init(n, 1);
    int s = st0;
    auto centroid_of = [&](const std::vector<double>& x) { double c = 0.0; for (int b = 0; b < P; ++b) c += x[b]; return c / P; };
    // per-bead acceleration (no springs: they are inside the C step)
    auto accel = [&](const std::vector<cd>& sg) {
        for (int b = 0; b < P; ++b) {
            bead[b].update(m, &r[b]);
            double f;
            if (R.method == RP_CLASSICAL) f = bead[b].dV[0];  // single-state model: dV/dx
            else if (R.method == RP_FSSH) f = bead[b].G[s * n + s] + bead[b].dV0[0];
            else {
                f = bead[b].dV0[0];
                for (int mm = 0; mm < n; ++mm) for (int q = 0; q < n; ++q) f += bead[b].G[q * n + mm] * sg[q * n + mm].real();
            }
            a[b] = -f / mass;
        }
    };
    double rc = centroid_of(r);
    if (R.method != RP_CLASSICAL) cen.update(m, &rc);
    if (R.method == RP_FSSH && s < 0) {
        double tot = 0.0;
        for (int q = 0; q < n; ++q) tot += sig[q * n + q].real();
        const double u = uniform01(gid, 0, STREAM_STATE, R.key) * tot;
        double cum = 0.0; s = n - 1;
        for (int q = 0; q < n; ++q) { cum += sig[q * n + q].real(); if (u < cum) { s = q; break; } }
    }
    accel(sig);
    EParams Pe; Pe.n = n;
    Pe.cur.E.assign(n, 0.0); Pe.cur.W.assign(nn, 0.0); Pe.cur.t = 0.0; Pe.next = Pe.cur;
    double t = R.t0;
    int32_t nh = 0;
    std::vector<double> prob(n);
    auto save = [&](int slot) {
        if (slot < 0 || slot >= R.nsave) return;
        const size_t o = (size_t)i * R.nsave + slot;
        if (O.centroid) O.centroid[o] = centroid_of(r);
        if (R.method == RP_CLASSICAL) return;
        if (O.pop_ad) adiabatic_population(R.method == RP_FSSH ? METHOD_FSSH : METHOD_EHRENFEST, n, sig.data(), s, O.pop_ad + o * n);
        if (O.pop_di) diabatic_population(R.method == RP_FSSH ? METHOD_FSSH : METHOD_EHRENFEST, n, cen.U.data(), sig.data(), s, O.pop_di + o * n);
        if (O.save_state) O.save_state[o] = s;
    };
    save(0);
    for (int64_t step = 0; step < R.nsteps; ++step) {
        std::vector<cd> sigprev = sig;
        for (int b = 0; b < P; ++b) vtmp[b] = v[b] + (R.dt / 2) * a[b];                  // step_B!
        for (int k = 0; k < P; ++k) {                                                    // to normal modes
            double x = 0.0, y = 0.0;
            for (int b = 0; b < P; ++b) { x += U[b * P + k] * r[b]; y += U[b * P + k] * vtmp[b]; }
            rn[k] = x; vn[k] = y;
        }
        for (int k = 0; k < P; ++k) {                                                    // step_C!
            const double x = C[4 * k] * rn[k] + C[4 * k + 1] * vn[k];
            const double y = C[4 * k + 2] * rn[k] + C[4 * k + 3] * vn[k];
            rn[k] = x; vn[k] = y;
        }
        for (int b = 0; b < P; ++b) {                                                    // from normal modes
            double x = 0.0, y = 0.0;
            for (int k = 0; k < P; ++k) { x += U[b * P + k] * rn[k]; y += U[b * P + k] * vn[k]; }
            r[b] = x; vtmp[b] = y;
        }
        accel(sigprev);
        for (int b = 0; b < P; ++b) v[b] = vtmp[b] + (R.dt / 2) * a[b];                  // step_B!
        if (R.method != RP_CLASSICAL) {
            rc = centroid_of(r);
            const double vc = centroid_of(v);
            cen.update(m, &rc);
            update_parameters(Pe, cen.E.data(), cen.d.data(), &vc, 1, t + R.dt);
            propagate_density(Pe, sig.data(), t, R.dt, R.nsub, ws);
        }
        t = t + R.dt;
        // the save precedes the hop callback (the rule of run_one)
        if (R.save_every > 0 && (step + 1) % R.save_every == 0) save((int)((step + 1) / R.save_every));
        if (R.method == RP_FSSH) {
            const double vc = centroid_of(v);
            fssh_probability(n, 1, &vc, sig.data(), s, cen.d.data(), R.dt, prob.data());
            const double rnd = uniform01(gid, (uint32_t)step, STREAM_HOP, R.key);
            const int ns = fssh_select(n, prob.data(), s, rnd);
            if (ns != s) {
                bool ok = true;
                if (R.rescaling != RESCALE_OFF) {
                    const double d = cen.d[ns * n + s];                    // coupling[new, old]
                    const double aq = d * d / mass / 2.0;
                    const double bq = d * vc;
                    const double cq = cen.E[ns] - cen.E[s];
                    const double disc = bq * bq - 4.0 * aq * cq;
                    if (disc < 0) {
                        ok = false;
                        if (R.rescaling == RESCALE_VINVERSION) {          // rpsh.jl:39-50
                            const double dn = d / std::fabs(d);
                            const double g = vc * dn;
                            for (int b = 0; b < P; ++b) v[b] -= 2.0 * g * dn;
                        }
                    } else {
                        const double root = std::sqrt(disc);
                        const double gamma = (bq < 0) ? (bq + root) / (2.0 * aq) : (bq - root) / (2.0 * aq);
                        for (int b = 0; b < P; ++b) v[b] -= gamma * d / mass;  // rpsh.jl:30-37
                    }
                }
                if (ok) { s = ns; ++nh; }
            }
        }
    }
    for (int b = 0; b < P; ++b) { O.r[(size_t)i * P + b] = r[b]; O.v[(size_t)i * P + b] = v[b]; }
    if (R.method != RP_CLASSICAL) {
        for (int J = 0; J < nn; ++J) { O.sig_re[(size_t)i * nn + J] = sig[J].real(); O.sig_im[(size_t)i * nn + J] = sig[J].imag(); }
        O.state[i] = s;
        O.nhops[i] = nh;
    }
}

OR_API int32_t oracle_rp_run_ensemble(const ORPDesc* R, const OModelDesc* m, double mass, int64_t ntraj,
                                      const double* r0, const double* v0, const double* sr0, const double* si0,
                                      const int32_t* st0, double* r, double* v, double* sr, double* si, int32_t* state,
                                      int32_t* nhops, double* centroid, double* pop_ad, double* pop_di,
                                      int32_t* save_state, int32_t nthreads) {
    const int P = R->nbeads, nn = R->nstates * R->nstates;
    RPTrajOut O{r, v, sr, si, state, nhops, centroid, pop_ad, pop_di, save_state};
    parallel_for(ntraj, nthreads, [&](int64_t i) {
        rp_run_one(*R, *m, mass, i, (uint64_t)i, r0 + (size_t)i * P, v0 + (size_t)i * P,
                   sr0 ? sr0 + (size_t)i * nn : nullptr, si0 ? si0 + (size_t)i * nn : nullptr, st0 ? st0[i] : 0, O);
    });
    return 0;
}

// ---------------------------------------------------------------------------
// Path-integral Monte Carlo for one-dof ring polymers (MetropolisHastings.jl:58-82,131-148,170-258):
// one Metropolis-Hastings chain per trajectory, for one quantum atom with no fixed atoms, so every
// step is (run_main_loop!, :170-195):
//   centroid move: the centroid normal mode shifted by (u - 1/2) Delta / sqrt(ndofs) (:208-229), i.e.
//     every bead by that over sqrt(P), then assess_proposal!;
//   pass_length = P / segment_length times: segment_length draws of a mode k in 2:P (1-based), each
//     resampled from Normal(0, sigma_k) (sample_mode!, :240-248), then assess_proposal!;
//   assess_proposal! once more on the last proposal (:194).
// assess_proposal! accepts when min(1, exp(-(E' - E)/w_n)) > u (:260-285) with E the classical
// potential summed over beads (DynamicsUtils.jl:143-147: element [1] of the model's potential, the
// first diabat of a two-state model), w_n = P T.  sigma_k = sqrt(w_n/m)/(m w_k^2/2) as the reference
// writes it (normal_mode_springs = w_k^2/2, ring_polymer.jl:59), or with width = 1 the free ring
// polymer's sqrt(w_n/m)/w_k.  The normal modes are the Hartley basis of the device kernels (column k
// spans the w_k eigenspace, as RingPolymerArrays' modes do; the chain's stationary law only sees the
// eigenspaces).  Julia's rand calls become Philox stream 4: counter (traj, step, 4 + 16 sub) with
// sub 0 the centroid move (u53(o0,o1) the shift, u53(o2,o3) the acceptance), sub 1 + j (L + 1) + s
// segment draw s of mode pass j (mode 1 + (o0 (P-1)) >> 32, z = sqrt(-2 log(1 - u53(o1,o2)))
// cos(2 pi o3 2^-32)), sub 1 + j (L + 1) + L its acceptance u53(o0,o1), sub 1 + J (L + 1) the last
// re-assessment.  Velocities: v_b ~ Normal(0, sqrt(w_n/m)) from counter (traj, bead, stream 0), z1.
// ---------------------------------------------------------------------------
enum { STREAM_PIMC = 4 };

struct OPimcDesc {
    int32_t nbeads;
    int32_t segment_length;   // max(1, round(segment * nbeads)), <= nbeads
    int64_t steps;            // n_atoms * passes
    double mass;
    double temperature;       // w_n = nbeads * temperature
    double delta;             // centroid move width Delta
    int32_t width;            // 0: the reference's sigma_k, 1: the free ring polymer's
    int32_t reassess;         // 1: the reference's extra assess_proposal! of the last proposal (:194)
    uint64_t key;
};

static void hartley_matrix(int P, double* H) {
    const double s = 1.0 / std::sqrt((double)P);
    for (int j = 0; j < P; ++j)
        for (int k = 0; k < P; ++k) {
            const double a = 2.0 * M_PI * (double)((j * k) % P) / (double)P;
            H[j * P + k] = (std::cos(a) + std::sin(a)) * s;
        }
}

static inline void pimc_draw(uint64_t traj, int64_t step, uint32_t sub, uint64_t key, uint32_t o[4]) {
    philox_draw(traj, (uint32_t)step, STREAM_PIMC + 16u * sub, key, o);
}

static double pimc_energy(const OModelDesc& m, const double* x, int P) {
    double E = 0.0;
    std::vector<double> V(m.nstates * m.nstates), dV(m.nstates * m.nstates * m.ndof), dV0(m.ndof);
    double V0;
    for (int b = 0; b < P; ++b) {
        model_eval(m, x + b, V.data(), dV.data(), &V0, dV0.data());
        E += V[0] + V0;
    }
    return E;
}

static void pimc_one(const OPimcDesc& D, const OModelDesc& m, uint64_t traj, const double* H, double* x, double* v,
                     int64_t* accepted) {
    const int P = D.nbeads, L = D.segment_length, J = P / L;
    const double wn = P * D.temperature, m_ = D.mass;
    std::vector<double> xp(P), y(P), sig(P, 0.0);
    for (int k = 1; k < P; ++k) {
        const double wk = matsubara(P, wn, k);
        sig[k] = D.width == 0 ? std::sqrt(wn / m_) / (0.5 * wk * wk * m_) : std::sqrt(wn / m_) / wk;
    }
    double Ei = pimc_energy(m, x, P), Ep = Ei;
    int64_t acc = 0;
    auto assess = [&](double u) {
        Ep = pimc_energy(m, xp.data(), P);
        if (std::min(1.0, std::exp(-(Ep - Ei) / wn)) > u) {
            Ei = Ep;
            std::copy(xp.begin(), xp.end(), x);
            ++acc;
        }
    };
    uint32_t o[4];
    for (int64_t it = 0; it < D.steps; ++it) {
        pimc_draw(traj, it, 0, D.key, o);
        const double shift = (u53(o[0], o[1]) - 0.5) * D.delta / std::sqrt(1.0);
        for (int b = 0; b < P; ++b) xp[b] = x[b] + shift * H[b * P + 0];
        assess(u53(o[2], o[3]));
        if (P > 1) {
            for (int j = 0; j < J; ++j) {
                for (int k = 0; k < P; ++k) {  // to normal modes
                    double a = 0.0;
                    for (int b = 0; b < P; ++b) a += H[k * P + b] * x[b];
                    y[k] = a;
                }
                for (int s = 0; s < L; ++s) {
                    pimc_draw(traj, it, 1u + (uint32_t)(j * (L + 1) + s), D.key, o);
                    const int k = 1 + (int)(((uint64_t)o[0] * (uint64_t)(P - 1)) >> 32);
                    const double z = std::sqrt(-2.0 * std::log(1.0 - u53(o[1], o[2]))) *
                                     std::cos(6.283185307179586 * ((double)o[3] * 0x1.0p-32));
                    y[k] = sig[k] * z;
                }
                for (int b = 0; b < P; ++b) {  // back to beads
                    double a = 0.0;
                    for (int k = 0; k < P; ++k) a += H[b * P + k] * y[k];
                    xp[b] = a;
                }
                pimc_draw(traj, it, 1u + (uint32_t)(j * (L + 1) + L), D.key, o);
                assess(u53(o[0], o[1]));
            }
        }
        if (D.reassess) {
            pimc_draw(traj, it, 1u + (uint32_t)(J * (L + 1)), D.key, o);
            assess(u53(o[0], o[1]));
        }
    }
    for (int b = 0; b < P; ++b) {
        double z0, z1;
        normal2(traj, (uint32_t)b, STREAM_INIT, D.key, &z0, &z1);
        v[b] = std::sqrt(wn / m_) * z1;
    }
    *accepted = acc;
}

// x (ntraj, P): starting beads in, final beads out; v (ntraj, P) out; accepted (ntraj) out.
OR_API int32_t oracle_pimc(const OPimcDesc* D, const OModelDesc* m, int64_t ntraj, double* x, double* v,
                           int64_t* accepted, int32_t nthreads) {
    const int P = D->nbeads;
    if (P < 1 || D->segment_length < 1 || D->segment_length > P) return -1;
    std::vector<double> H((size_t)P * P);
    hartley_matrix(P, H.data());
    parallel_for(ntraj, nthreads, [&](int64_t i) {
        pimc_one(*D, *m, (uint64_t)i, H.data(), x + (size_t)i * P, v + (size_t)i * P, accepted + i);
    });
    return 0;
}

// ---------------------------------------------------------------------------
// Classical master equation surface hopping (cme.jl:26-182): CME (Dou, Nitzan, Subotnik 2015) and
// BCME (broadened, Dou and Subotnik 2016) on a two-state diabatic model of one coordinate.
//   motion!: dr = v, dv = a(r, state) (cme.jl:7-18), integrated by Tsit5 (select_algorithm,
//     IntegrationAlgorithms.jl:68).  The reference's solve adapts the step to abstol / reltol; here
//     the step is fixed at dt (OrdinaryDiffEq's Tsit5 tableau, no controller).
//   CME force: -dV_ss/dx / m (cme.jl:96-105); BCME adds f dh - dh n(h) - (dGamma/Gamma) n2(h)
//     (cme.jl:137-160) with n, n2 the Lorentzian-broadened Fermi integrals over the bandwidth
//     (cme.jl:167-182; QuadGK's adaptive Gauss-Kronrod, restated here with a 7/15-point pair and a
//     tighter tolerance).
//   HoppingCallback after each step (surface_hopping.jl:2-16): Gamma = 2 pi V12^2, f = fermi(V22 - V11,
//     mu, beta), P = Gamma f dt from state 1 and Gamma (1 - f) dt from state 2 (cme.jl:26-41, dt the
//     proposed step, here dt); one uniform, hop when u < P (cme.jl:43-53); no rescaling (cme.jl:55-57).
// ---------------------------------------------------------------------------
struct OCmeDesc {
    int32_t method;        // 0 CME, 1 BCME
    int32_t save_every;
    int32_t nsave;
    int32_t pad;
    double dt, t0, mass, kT, mu, bandwidth;
    int64_t nsteps;
    uint64_t key;
    OCell cell;
};

static inline double fermi_fn(double e, double mu, double beta) { return 1.0 / (1.0 + std::exp(beta * (e - mu))); }

// Adaptive Gauss-Kronrod (7/15) on [a, b] to |err| <= max(atol, rtol |I|), bisecting the worst panel.
template <class F>
static double quad_gk(F f, double a, double b, double rtol) {
    static const double xk[8] = {0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                                 0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                                 0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                                 0.207784955007898467600689403773245, 0.0};
    static const double wk[8] = {0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                                 0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                                 0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                                 0.204432940075298892414161999234649, 0.209482141084727828012999174891714};
    static const double wg[4] = {0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                                 0.381830050505118944950369775488975, 0.417959183673469387755102040816327};
    struct Seg { double a, b, I, E; };
    auto eval = [&](double lo, double hi) {
        const double c = 0.5 * (lo + hi), h = 0.5 * (hi - lo);
        double fk = 0.0, fg = 0.0;
        const double f0 = f(c);
        fk = wk[7] * f0; fg = wg[3] * f0;
        for (int j = 0; j < 7; ++j) {
            const double y = f(c - h * xk[j]) + f(c + h * xk[j]);
            fk += wk[j] * y;
            if (j % 2 == 1) fg += wg[j / 2] * y;
        }
        return Seg{lo, hi, fk * h, std::fabs((fk - fg) * h)};
    };
    std::vector<Seg> segs{eval(a, b)};
    for (int it = 0; it < 4000; ++it) {
        double I = 0.0, E = 0.0;
        size_t worst = 0;
        for (size_t k = 0; k < segs.size(); ++k) {
            I += segs[k].I; E += segs[k].E;
            if (segs[k].E > segs[worst].E) worst = k;
        }
        if (E <= rtol * std::fabs(I) || E < 1e-300) return I;
        const Seg w = segs[worst];
        const double mid = 0.5 * (w.a + w.b);
        segs[worst] = eval(w.a, mid);
        segs.push_back(eval(mid, w.b));
    }
    double I = 0.0;
    for (const Seg& g : segs) I += g.I;
    return I;
}

// evaluate_broadening / evaluate_force_broadening (cme.jl:167-182)
static double cme_broadening(double h, double mu, double beta, double Gam, double W, int force) {
    auto A = [&](double e) { return 1.0 / M_PI * (Gam / 2.0) / ((e - h) * (e - h) + (Gam / 2.0) * (Gam / 2.0)); };
    auto k = [&](double e) { return (force ? (e - h) : 1.0) * A(e) * fermi_fn(e, mu, beta); };
    // the Lorentzian peak and the Fermi edge are panel boundaries, so the adaptive rule sees both
    std::vector<double> pts{mu - W / 2, mu + W / 2};
    for (double p : {h, mu}) if (p > mu - W / 2 && p < mu + W / 2) pts.push_back(p);
    std::sort(pts.begin(), pts.end());
    double I = 0.0;
    for (size_t j = 0; j + 1 < pts.size(); ++j) I += quad_gk(k, pts[j], pts[j + 1], 1e-13);
    return I;
}

static void cme_accel(const OCmeDesc& D, const OModelDesc& m, double x, int s, double* a, double* Vout) {
    double V[4], dV[4], V0, dV0;
    model_eval(m, &x, V, dV, &V0, &dV0);
    double f = -dV[s * 3];
    if (D.method == 1) {
        const double Gam = 2.0 * M_PI * V[1] * V[1];
        const double beta = 1.0 / D.kT;
        const double h = V[3] - V[0];
        const double n = cme_broadening(h, D.mu, beta, Gam, D.bandwidth, 0);
        const double n2 = cme_broadening(h, D.mu, beta, Gam, D.bandwidth, 1);
        const double fe = fermi_fn(h, D.mu, beta);
        const double dh = dV[3] - dV[0];
        const double dGam = 4.0 * M_PI * V[1] * dV[1];
        f += fe * dh - dh * n - dGam / Gam * n2;
    }
    *a = f / D.mass;
    if (Vout) { Vout[0] = V[0]; Vout[1] = V[1]; Vout[2] = V[3]; }
}

// One trajectory; x, v in/out; saves the position, kinetic energy, potential V_ss and state per slot.
static void cme_one(const OCmeDesc& D, const OModelDesc& m, uint64_t gid, double& x, double& v, int& s, int64_t& steps,
                    int32_t& nh, double* sx, double* sek, double* sep, int32_t* sst) {
    using namespace tsit5;
    auto save = [&](int slot) {
        if (slot >= D.nsave) return;
        double a, Vv[3];
        cme_accel(D, m, x, s, &a, Vv);
        sx[slot] = x; sek[slot] = 0.5 * D.mass * v * v; sep[slot] = s == 0 ? Vv[0] : Vv[2]; sst[slot] = s;
    };
    save(0);
    nh = 0;
    for (int64_t step = 0; step < D.nsteps; ++step) {
        const double h = D.dt;
        double k1x, k1v, k2x, k2v, k3x, k3v, k4x, k4v, k5x, k5v, k6x, k6v, a;
        k1x = v; cme_accel(D, m, x, s, &a, nullptr); k1v = a;
        double tx = x + h * (a21 * k1x), tv = v + h * (a21 * k1v);
        k2x = tv; cme_accel(D, m, tx, s, &a, nullptr); k2v = a;
        tx = x + h * (a31 * k1x + a32 * k2x); tv = v + h * (a31 * k1v + a32 * k2v);
        k3x = tv; cme_accel(D, m, tx, s, &a, nullptr); k3v = a;
        tx = x + h * (a41 * k1x + a42 * k2x + a43 * k3x); tv = v + h * (a41 * k1v + a42 * k2v + a43 * k3v);
        k4x = tv; cme_accel(D, m, tx, s, &a, nullptr); k4v = a;
        tx = x + h * (a51 * k1x + a52 * k2x + a53 * k3x + a54 * k4x);
        tv = v + h * (a51 * k1v + a52 * k2v + a53 * k3v + a54 * k4v);
        k5x = tv; cme_accel(D, m, tx, s, &a, nullptr); k5v = a;
        tx = x + h * (a61 * k1x + a62 * k2x + a63 * k3x + a64 * k4x + a65 * k5x);
        tv = v + h * (a61 * k1v + a62 * k2v + a63 * k3v + a64 * k4v + a65 * k5v);
        k6x = tv; cme_accel(D, m, tx, s, &a, nullptr); k6v = a;
        x = x + h * (a71 * k1x + a72 * k2x + a73 * k3x + a74 * k4x + a75 * k5x + a76 * k6x);
        v = v + h * (a71 * k1v + a72 * k2v + a73 * k3v + a74 * k4v + a75 * k5v + a76 * k6v);
        steps = step + 1;
        // the save precedes the hop callback (the rule of run_one)
        if (D.save_every > 0 && steps % D.save_every == 0) save((int)(steps / D.save_every));
        // HoppingCallback (surface_hopping.jl:2-16) with the CME probability (cme.jl:26-53)
        double Vv[3], aa;
        cme_accel(D, m, x, s, &aa, Vv);
        const double Gam = 2.0 * M_PI * Vv[1] * Vv[1];
        const double fe = fermi_fn(Vv[2] - Vv[0], D.mu, 1.0 / D.kT);
        const double P = (s == 0) ? Gam * fe * D.dt : Gam * (1.0 - fe) * D.dt;
        const double u = uniform01(gid, (uint32_t)step, STREAM_HOP, D.key);
        if (u < P) { s = 1 - s; ++nh; }
        if (D.cell.mode != 0 && ocell_outside(D.cell, &x, 1)) {   // cell callback after the hop callback
            if (D.cell.mode == 2) break;
            ocell_wrap(D.cell, &x, 1);
        }
    }
}

// ---------------------------------------------------------------------------
// Classical dynamics, Simulation{Classical} (classical.jl:40-86) on a single-state model of one
// coordinate: OrdinaryDiffEq's VelocityVerlet (classical.jl:86) on the DynamicalODEProblem(acceleration!,
// velocity!, v, r) (classical.jl:80-84).  VelocityVerlet's perform_step! (OrdinaryDiffEq, a dependency
// absent from the reference tree) evaluates a = f1(v, x) at the step start, then x' = x + dt v +
// dt^2 (a / 2), a' = f1(v, x'), v' = v + dt (a / 2 + a' / 2); the start value equals the previous end
// value unless a callback moved x.  Saves precede the callbacks; the cell callback and the termination
// predicate follow the step.
// ---------------------------------------------------------------------------
struct OClsDesc {
    int32_t save_every, nsave, term_kind, pad;
    double term_x, dt, mass;
    int64_t nsteps;
    OCell cell;
};

OR_API int32_t oracle_classical_run(const OClsDesc* D, const OModelDesc* m, int64_t ntraj, double* x, double* v,
                                    int64_t* steps, double* sx, double* sv, double* sek, double* sep, int32_t nthreads) {
    const int ns = D->nsave;
    parallel_for(ntraj, nthreads, [&](int64_t i) {
        double xi = x[i], vi = v[i];
        auto accel = [&](double r, double& V) {
            double Vm[1], dV[1], V0, dV0[1];
            model_eval(*m, &r, Vm, dV, &V0, dV0);
            V = Vm[0] + V0;
            return -(dV[0] + dV0[0]) / D->mass;
        };
        double V;
        double a = accel(xi, V);
        auto save = [&](int slot) {
            if (slot >= ns) return;
            sx[(size_t)i * ns + slot] = xi; sv[(size_t)i * ns + slot] = vi;
            sek[(size_t)i * ns + slot] = 0.5 * D->mass * vi * vi; sep[(size_t)i * ns + slot] = V;
        };
        save(0);
        int64_t step = 0;
        while (step < D->nsteps) {
            const double xn = xi + D->dt * vi + (D->dt * D->dt) * (0.5 * a);
            const double an = accel(xn, V);
            vi = vi + D->dt * (0.5 * a + 0.5 * an);
            xi = xn;
            a = an;
            if (D->save_every > 0 && (step + 1) % D->save_every == 0) save((int)((step + 1) / D->save_every));
            bool end = false;
            if (D->cell.mode != 0 && ocell_outside(D->cell, &xi, 1)) {
                if (D->cell.mode == 2) end = true;
                else { ocell_wrap(D->cell, &xi, 1); a = accel(xi, V); }
            }
            ++step;
            if (end) break;
            if (D->term_kind == 1 && std::fabs(xi) > D->term_x) break;
            if (D->term_kind == 2 && xi > D->term_x) break;
            if (D->term_kind == 3 && xi < D->term_x) break;
        }
        x[i] = xi; v[i] = vi; steps[i] = step;
    });
    return 0;
}

OR_API int32_t oracle_cme_run(const OCmeDesc* D, const OModelDesc* m, int64_t ntraj, double* x, double* v,
                              int32_t* state, int32_t* nhops, double* sx, double* sek, double* sep, int32_t* sst,
                              int32_t nthreads) {
    parallel_for(ntraj, nthreads, [&](int64_t i) {
        int s = state[i];
        int64_t steps = 0;
        cme_one(*D, *m, (uint64_t)i, x[i], v[i], s, steps, nhops[i], sx + (size_t)i * D->nsave, sek + (size_t)i * D->nsave,
                sep + (size_t)i * D->nsave, sst + (size_t)i * D->nsave);
        state[i] = s;
    });
    return 0;
}

OR_API double oracle_cme_broadening(double h, double mu, double beta, double Gam, double W, int32_t force) {
    return cme_broadening(h, mu, beta, Gam, W, force);
}

// ---------------------------------------------------------------------------
// Spin mapping (SpinMappingW, spin_mapping.jl:1-131) integrated by MInt (mint.jl:42-181), one
// coordinate.  With gamma = 2 (sqrt(n + 1) - 1) / n (spin_mapping.jl:4-9) the Hamiltonian is
//   H = m v^2 / 2 + V0 + sum_i V_ii (X_i^2 + P_i^2 - gamma) / 2 + sum_{i<j} V_ij (X_i X_j + P_i P_j)
// (classical_potential_energy, spin_mapping.jl:102-118), X the mapping positions, P the momenta.
// MInt step (mint.jl:42-86): r' = r + dt/2 v; at r' the eigenpairs (w, Z) of V; X, P <- C X - D P,
// C P + D X with C = Z diag(cos(w dt)) Z^T, D = Z diag(sin(-w dt)) Z^T (mint.jl:88-124); then with
// W = Z^T dV Z: Gamma_ii = W_ii dt, Gamma_ji = sin((w_i - w_j) dt) W_ji / (w_i - w_j), Xi_ji =
// (1 - cos((w_i - w_j) dt)) W_ji / (w_i - w_j) (j != i), E = Z Gamma Z^T, F = Z Xi Z^T
// (mint.jl:132-162), force = (X^T E X + P^T E P) / 2 - X^T F P (mint.jl:164-181),
// v <- v - force / m - dV0 / m dt + tr(dV) gamma / m dt / 2, and r'' = r' + dt/2 v.  set_xi! writes
// only the off-diagonal entries of the cache matrix that transform_matrix! then overwrites with
// Z Xi Z^T, so each step's Xi carries the previous step's transformed diagonal (zero at a
// trajectory's first step); restated as the reference computes it.
//
// eCMM (method 1, cmm.jl:1-120) has H = m v^2 / 2 + (q^T V q + p^T V p - 2 gamma tr V) / 2
// (classical_potential_energy, cmm.jl:84-94) with its own gamma (keyword, default 0) and is integrated
// by the same MInt step, whose force term adds tr(dV) gamma / m dt / 2 for either method (mint.jl:85):
// for eCMM that is half of what its Hamiltonian's -gamma tr V gives, which matters only where
// tr dV != 0; restated as the reference computes it.  Populations: spin mapping (X^2 + P^2 - gamma)/2;
// eCMM the inverse mapping kernel (1 + F) / (1 + F gamma)^2 (X^2 + P^2)/2 - (1 - gamma) / (1 + F gamma)
// (diabatic_population, cmm.jl:96-112).
// ---------------------------------------------------------------------------
struct OMapDesc {
    int32_t nstates, save_every, nsave, method;   // method: 0 SpinMappingW, 1 eCMM
    double dt, mass;
    int64_t nsteps;
    double gamma;                                 // eCMM's gamma (unused for spin mapping)
};

static double mapping_gamma(int n) { return 2.0 / n * (std::sqrt((double)n + 1.0) - 1.0); }

// gamma of the method: spin mapping's from n, eCMM's as given
static double map_gamma(const OMapDesc& D) { return D.method == 1 ? D.gamma : mapping_gamma(D.nstates); }
// the constant each diagonal V_ii carries in H: V_ii (X^2 + P^2 - c) / 2
static double map_diag_const(const OMapDesc& D) { return D.method == 1 ? 2.0 * D.gamma : mapping_gamma(D.nstates); }

static double map_population(const OMapDesc& D, double x, double p) {
    if (D.method == 1) {   // inverse_mapping_kernel, cmm.jl:99-104
        const int F = D.nstates;
        const double g = D.gamma;
        return (1.0 + F) / ((1.0 + F * g) * (1.0 + F * g)) * (x * x + p * p) / 2.0 - (1.0 - g) / (1.0 + F * g);
    }
    return (x * x + p * p - mapping_gamma(D.nstates)) / 2.0;
}

// H at (r, v, X, P): the classical Hamiltonian of the mapping variables.
static double mapping_energy(const OMapDesc& D, const OModelDesc& m, double r, double v, const double* X, const double* P) {
    const int n = D.nstates;
    std::vector<double> V(n * n), dV(n * n);
    double V0, dV0;
    model_eval(m, &r, V.data(), dV.data(), &V0, &dV0);
    const double g = map_diag_const(D);
    double e = 0.5 * D.mass * v * v + V0;
    for (int i = 0; i < n; ++i) {
        e += V[i * n + i] * (X[i] * X[i] + P[i] * P[i] - g) / 2.0;
        for (int j = i + 1; j < n; ++j) e += V[j * n + i] * (X[i] * X[j] + P[i] * P[j]);
    }
    return e;
}

// motion! (spin_mapping.jl:56-100): dr = v, dv = -(dV0 + sum_i dV_ii (X_i^2 + P_i^2 - gamma)/2
// + sum_{i<j} dV_ji (X_i X_j + P_i P_j)) / m, dX = V P, dP = -V X.  u = [r, v, X(n), P(n)].
static void mapping_motion(const OMapDesc& D, const OModelDesc& m, const double* u, double* du) {
    const int n = D.nstates;
    const double r = u[0], v = u[1];
    const double* X = u + 2;
    const double* P = u + 2 + n;
    std::vector<double> V(n * n), dV(n * n);
    double V0, dV0;
    model_eval(m, &r, V.data(), dV.data(), &V0, &dV0);
    const double g = map_diag_const(D);   // eCMM: -(q^T D q + p^T D p - 2 gamma tr D) / 2, cmm.jl:55-70
    double f = -dV0;
    for (int i = 0; i < n; ++i) {
        f -= dV[i * n + i] * (X[i] * X[i] + P[i] * P[i] - g) / 2.0;
        for (int j = i + 1; j < n; ++j) f -= dV[j * n + i] * (X[i] * X[j] + P[i] * P[j]);
    }
    du[0] = v;
    du[1] = f / D.mass;
    for (int i = 0; i < n; ++i) {
        double sx = 0.0, sp = 0.0;
        for (int j = 0; j < n; ++j) { sx += V[i * n + j] * P[j]; sp += V[i * n + j] * X[j]; }
        du[2 + i] = sx;
        du[2 + n + i] = -sp;
    }
}

// M <- Z M Z^T (transform_matrix!, mint.jl:126-130)
static void mapping_transform(int n, const double* Z, double* M) {
    std::vector<double> T(n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += M[i * n + k] * Z[j * n + k];
            T[i * n + j] = s;
        }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += Z[i * n + k] * T[k * n + j];
            M[i * n + j] = s;
        }
}

// One MInt step; Xi holds the cache matrix Xi between steps (its diagonal carries over).
static void mint_step(const OMapDesc& D, const OModelDesc& m, double& r, double& v, double* X, double* P, double* Xi) {
    const int n = D.nstates;
    const double dt = D.dt, g = map_gamma(D);
    const double rt = r + dt / 2.0 * v;
    std::vector<double> V(n * n), dV(n * n), w(n), Z(n * n), C(n * n, 0.0), Dm(n * n, 0.0), tmp(n), tmp2(n);
    double V0, dV0;
    model_eval(m, &rt, V.data(), dV.data(), &V0, &dV0);
    eigen_sym(n, V.data(), w.data(), Z.data(), nullptr);
    for (int i = 0; i < n; ++i) { C[i * n + i] = std::cos(w[i] * dt); Dm[i * n + i] = std::sin(-w[i] * dt); }
    mapping_transform(n, Z.data(), C.data());
    mapping_transform(n, Z.data(), Dm.data());
    for (int i = 0; i < n; ++i) {   // tmp = C X - D P, tmp2 = C P + D X
        double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
        for (int j = 0; j < n; ++j) { a += C[i * n + j] * X[j]; b += Dm[i * n + j] * P[j]; c += C[i * n + j] * P[j]; d += Dm[i * n + j] * X[j]; }
        tmp[i] = a - b;
        tmp2[i] = c + d;
    }
    for (int i = 0; i < n; ++i) { X[i] = tmp[i]; P[i] = tmp2[i]; }
    // W = Z^T dV Z
    std::vector<double> W(n * n), G(n * n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int a = 0; a < n; ++a)
                for (int b = 0; b < n; ++b) s += Z[a * n + i] * dV[a * n + b] * Z[b * n + j];
            W[i * n + j] = s;
        }
    for (int i = 0; i < n; ++i) {   // set_gamma! / set_xi! (column i, row j)
        G[i * n + i] = W[i * n + i] * dt;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            const double dl = w[i] - w[j];
            G[j * n + i] = std::sin(dl * dt) * W[j * n + i] / dl;
            Xi[j * n + i] = (1.0 - std::cos(dl * dt)) * W[j * n + i] / dl;
        }
    }
    mapping_transform(n, Z.data(), G.data());
    mapping_transform(n, Z.data(), Xi);
    double f = 0.0;   // get_mapping_nuclear_force
    for (int i = 0; i < n; ++i) {
        double eq = 0.0, ep = 0.0, fp = 0.0;
        for (int j = 0; j < n; ++j) { eq += G[i * n + j] * X[j]; ep += G[i * n + j] * P[j]; fp += Xi[i * n + j] * P[j]; }
        f += X[i] * eq + P[i] * ep - 2.0 * X[i] * fp;
    }
    f /= 2.0;
    double trdV = 0.0;
    for (int i = 0; i < n; ++i) trdV += dV[i * n + i];
    v -= f / D.mass;
    v -= dV0 / D.mass * dt;
    v += trdV * g / D.mass * dt / 2.0;
    r = rt + dt / 2.0 * v;
}

// Initial mapping variables of a pure diabatic state (spin_mapping.jl:18-37): theta_i uniform in
// [0, 2 pi) from the Philox block (trajectory, i, stream 4); radius sqrt(2 + gamma) for the occupied
// state, sqrt(gamma) for the others; the mapping positions take the sines and the momenta the cosines
// (DynamicsVariables sets qmap = P and pmap = X of its cos / sin pair).
OR_API void oracle_mapping_sample(int32_t n, int64_t ntraj, const int32_t* state, uint64_t key, const int64_t* traj_ids,
                                  double* X, double* P) {
    const double g = mapping_gamma(n);
    for (int64_t t = 0; t < ntraj; ++t) {
        const uint64_t id = traj_ids ? (uint64_t)traj_ids[t] : (uint64_t)t;
        for (int i = 0; i < n; ++i) {
            const double th = 2.0 * M_PI * uniform01(id, (uint32_t)i, 4u, key);
            const double rad = std::sqrt(i == state[t] ? 2.0 + g : g);
            X[t * n + i] = rad * std::sin(th);
            P[t * n + i] = rad * std::cos(th);
        }
    }
}

// eCMM's initial mapping variables (cmm.jl:26-43): a point uniform on the sphere of radius sqrt(2 + 2 F gamma)
// in 2F dimensions (2F normals normalised: the pair (z_2k, z_2k+1) from the Philox block (trajectory, k,
// stream 4)); the first F coordinates are the positions q, the last F the momenta p.  (The reference
// splits at div(F, 2) + 1, which is F only for F = 2; for other F its q and p lengths differ from F.)
OR_API void oracle_ecmm_sample(int32_t n, int64_t ntraj, double gamma, uint64_t key, const int64_t* traj_ids, double* X,
                               double* P) {
    const double R = std::sqrt(2.0 + 2.0 * n * gamma);
    std::vector<double> z(2 * n);
    for (int64_t t = 0; t < ntraj; ++t) {
        const uint64_t id = traj_ids ? (uint64_t)traj_ids[t] : (uint64_t)t;
        double s = 0.0;
        for (int k = 0; k < n; ++k) {
            normal2(id, (uint32_t)k, 4u, key, &z[2 * k], &z[2 * k + 1]);
            s += z[2 * k] * z[2 * k] + z[2 * k + 1] * z[2 * k + 1];
        }
        const double sc = R / std::sqrt(s);
        for (int i = 0; i < n; ++i) { X[t * n + i] = z[i] * sc; P[t * n + i] = z[n + i] * sc; }
    }
}

OR_API double oracle_mapping_energy(const OMapDesc* D, const OModelDesc* m, const double* u) {
    return mapping_energy(*D, *m, u[0], u[1], u + 2, u + 2 + D->nstates);
}

OR_API void oracle_mapping_motion(const OMapDesc* D, const OModelDesc* m, const double* u, double* du) {
    mapping_motion(*D, *m, u, du);
}

// Trajectories by MInt: x, v (ntraj), X, P ([traj][n]) in place; per save slot the diabatic
// populations (spin mapping: (X_i^2 + P_i^2 - gamma) / 2, spin_mapping.jl:120-131; eCMM: the inverse
// mapping kernel, cmm.jl:99-112), the kinetic and the potential energy (H - m v^2 / 2).
// sinit (ntraj, n): the initial estimator of the population correlation at t0 (initial_diabatic_population:
// the population itself for spin mapping, Estimators.jl:121; eCMM's mapping kernel (X^2 + P^2)/2 - gamma,
// cmm.jl:106-110).
OR_API int32_t oracle_mapping_run(const OMapDesc* D, const OModelDesc* m, int64_t ntraj, double* x, double* v,
                                  double* X, double* P, double* spop, double* sek, double* sep, double* sinit,
                                  int32_t nthreads) {
    const int n = D->nstates, ns = D->nsave;
    parallel_for(ntraj, nthreads, [&](int64_t t) {
        double* Xt = X + t * n;
        double* Pt = P + t * n;
        std::vector<double> Xi(n * n, 0.0);
        auto save = [&](int slot) {
            if (slot >= ns) return;
            for (int i = 0; i < n; ++i) spop[((size_t)t * ns + slot) * n + i] = map_population(*D, Xt[i], Pt[i]);
            const double ek = 0.5 * D->mass * v[t] * v[t];
            sek[(size_t)t * ns + slot] = ek;
            sep[(size_t)t * ns + slot] = mapping_energy(*D, *m, x[t], v[t], Xt, Pt) - ek;
        };
        for (int i = 0; i < n; ++i)
            sinit[(size_t)t * n + i] = D->method == 1 ? (Xt[i] * Xt[i] + Pt[i] * Pt[i]) / 2.0 - D->gamma
                                                      : map_population(*D, Xt[i], Pt[i]);
        save(0);
        for (int64_t step = 0; step < D->nsteps; ++step) {
            mint_step(*D, *m, x[t], v[t], Xt, Pt, Xi.data());
            if (D->save_every > 0 && (step + 1) % D->save_every == 0) save((int)((step + 1) / D->save_every));
        }
    });
    return 0;
}

// ---------------------------------------------------------------------------
// Ring-polymer mapping dynamics by RingPolymerMInt (ringpolymer_mint.jl:28-131): NRPMD (nrpmd.jl:1-139,
// method 0) and ring-polymer eCMM (rpecmm.jl:1-103, method 1), one coordinate per bead, F states.
// Per bead b the Hamiltonian holds (p_b^T V~ p_b + q_b^T V~ q_b) / 2 + Vbar with Vbar = tr V / F and
// V~ = V - Vbar I (classical_potential_energy, nrpmd.jl:123-139, rpecmm.jl:62-78), plus the kinetic
// energy and the springs m w_n^2 / 2 sum_b (r_b - r_b+1)^2 (ring_polymer.jl:89-107).  A step: the free
// ring polymer's half step in normal modes with sqrt(cay(dt A)) (ring_polymer.jl:71-82, half = true),
// which for mode frequency w is [[1, dt/2], [-w^2 dt/2, 1]] / sqrt(1 + dt^2 w^2 / 4); at the new beads the
// mapping variables rotated by the traceless eigenvalues w_k - Vbar (propagate_mapping_variables!,
// ringpolymer_mint.jl:88-101); the force from Gamma and Xi of W~ = Z^T (dV - Dbar I) Z with Dbar = tr dV / F
// (Xi built fresh with a zero diagonal, get_gamma / get_xi :111-129) and v_b -= force / m + Dbar / m dt
// (:65-80); then the second half step.  The normal modes are the real DFT basis (any orthonormal basis of
// the spring eigenspaces gives the same bead-space step).
// ---------------------------------------------------------------------------
struct ORpMapDesc {
    int32_t nstates, nbeads, method, save_every;   // method: 0 NRPMD, 1 ring-polymer eCMM
    int32_t nsave, pad;
    double dt, mass, temperature, gamma;
    int64_t nsteps;
};

// Potential part (without springs) of bead b: (p^T V~ p + q^T V~ q) / 2 + Vbar.
static double rpmap_bead_potential(int F, const OModelDesc& m, double r, const double* q, const double* p) {
    std::vector<double> V(F * F), dV(F * F);
    double V0, dV0;
    model_eval(m, &r, V.data(), dV.data(), &V0, &dV0);
    double tr = 0.0;
    for (int i = 0; i < F; ++i) tr += V[i * F + i];
    const double Vb = tr / F;
    double e = Vb + V0;
    for (int i = 0; i < F; ++i)
        for (int j = 0; j < F; ++j) {
            const double vt = V[i * F + j] - (i == j ? Vb : 0.0);
            e += 0.5 * vt * (p[i] * p[j] + q[i] * q[j]);
        }
    return e;
}

static double rpmap_spring(const ORpMapDesc& D, const double* r) {
    const int P = D.nbeads;
    const double wn = P * D.temperature;
    double e = 0.0;
    if (P > 1)
        for (int b = 0; b < P; ++b) { const double d = r[b] - r[(b + 1) % P]; e += d * d; }
    return e * D.mass * wn * wn / 2.0;
}

// classical_hamiltonian (DynamicsUtils.jl:74-79) at u = [r(P), v(P), q(P*F), p(P*F)] (bead-major).
static double rpmap_energy(const ORpMapDesc& D, const OModelDesc& m, const double* u) {
    const int P = D.nbeads, F = D.nstates;
    const double* r = u; const double* v = u + P; const double* q = u + 2 * P; const double* p = u + 2 * P + P * F;
    double e = rpmap_spring(D, r);
    for (int b = 0; b < P; ++b) e += 0.5 * D.mass * v[b] * v[b] + rpmap_bead_potential(F, m, r[b], q + b * F, p + b * F);
    return e;
}

// motion! (nrpmd.jl:67-121, rpecmm.jl:36-60): dr = v; dv = -((q^T D~ q + p^T D~ p) / 2 + Dbar) / m plus the
// springs (apply_interbead_coupling!); dq = V~ p, dp = -V~ q per bead.
static void rpmap_motion(const ORpMapDesc& D, const OModelDesc& m, const double* u, double* du) {
    const int P = D.nbeads, F = D.nstates;
    const double wn = P * D.temperature;
    const double* r = u; const double* v = u + P; const double* q = u + 2 * P; const double* p = u + 2 * P + P * F;
    std::vector<double> V(F * F), dV(F * F);
    for (int b = 0; b < P; ++b) {
        double V0, dV0;
        model_eval(m, &r[b], V.data(), dV.data(), &V0, &dV0);
        double trV = 0.0, trD = 0.0;
        for (int i = 0; i < F; ++i) { trV += V[i * F + i]; trD += dV[i * F + i]; }
        const double Vb = trV / F, Db = trD / F;
        const double* qb = q + b * F; const double* pb = p + b * F;
        double f = 0.0;
        for (int i = 0; i < F; ++i)
            for (int j = 0; j < F; ++j) f += (dV[i * F + j] - (i == j ? Db : 0.0)) * (qb[i] * qb[j] + pb[i] * pb[j]);
        double a = -(0.5 * f + Db + dV0) / D.mass;
        if (P > 1) a -= wn * wn * (2.0 * r[b] - r[(b + 1) % P] - r[(b + P - 1) % P]);
        du[b] = v[b];
        du[P + b] = a;
        for (int i = 0; i < F; ++i) {
            double sx = 0.0, sp = 0.0;
            for (int j = 0; j < F; ++j) {
                const double vt = V[i * F + j] - (i == j ? Vb : 0.0);
                sx += vt * pb[j]; sp += vt * qb[j];
            }
            du[2 * P + b * F + i] = sx;
            du[2 * P + P * F + b * F + i] = -sp;
        }
    }
}

static void rpmap_half_cayley(double w, double dt, double Ch[4]) {
    const double s = 1.0 / std::sqrt(1.0 + dt * dt * w * w / 4.0);
    Ch[0] = s; Ch[1] = 0.5 * dt * s; Ch[2] = -0.5 * dt * w * w * s; Ch[3] = s;
}

// One RingPolymerMInt step on one ring polymer (r, v: P beads; q, p: P*F bead-major).
static void rpmap_step(const ORpMapDesc& D, const OModelDesc& m, const std::vector<double>& U,
                       const std::vector<double>& Ch, double* r, double* v, double* q, double* p) {
    const int P = D.nbeads, F = D.nstates;
    const double dt = D.dt;
    std::vector<double> rn(P), vn(P);
    auto half = [&]() {
        for (int k = 0; k < P; ++k) {
            double x = 0.0, y = 0.0;
            for (int b = 0; b < P; ++b) { x += U[b * P + k] * r[b]; y += U[b * P + k] * v[b]; }
            rn[k] = Ch[4 * k] * x + Ch[4 * k + 1] * y;
            vn[k] = Ch[4 * k + 2] * x + Ch[4 * k + 3] * y;
        }
        for (int b = 0; b < P; ++b) {
            double x = 0.0, y = 0.0;
            for (int k = 0; k < P; ++k) { x += U[b * P + k] * rn[k]; y += U[b * P + k] * vn[k]; }
            r[b] = x; v[b] = y;
        }
    };
    half();
    std::vector<double> V(F * F), dV(F * F), w(F), Z(F * F), C(F * F), Dm(F * F), W(F * F), G(F * F), X(F * F), tq(F), tp(F);
    for (int b = 0; b < P; ++b) {
        double V0, dV0;
        model_eval(m, &r[b], V.data(), dV.data(), &V0, &dV0);
        eigen_sym(F, V.data(), w.data(), Z.data(), nullptr);
        double trV = 0.0, trD = 0.0;
        for (int i = 0; i < F; ++i) { trV += V[i * F + i]; trD += dV[i * F + i]; }
        const double Vb = trV / F, Db = trD / F;
        double* qb = q + b * F; double* pb = p + b * F;
        // C = Z cos((w - Vbar) dt) Z^T, D = Z sin(-(w - Vbar) dt) Z^T (get_C_propagator / get_D_propagator)
        std::fill(C.begin(), C.end(), 0.0); std::fill(Dm.begin(), Dm.end(), 0.0);
        for (int i = 0; i < F; ++i) { C[i * F + i] = std::cos((w[i] - Vb) * dt); Dm[i * F + i] = std::sin(-(w[i] - Vb) * dt); }
        mapping_transform(F, Z.data(), C.data());
        mapping_transform(F, Z.data(), Dm.data());
        for (int i = 0; i < F; ++i) {
            double a = 0.0, bb = 0.0, c = 0.0, d = 0.0;
            for (int j = 0; j < F; ++j) { a += C[i * F + j] * qb[j]; bb += Dm[i * F + j] * pb[j]; c += C[i * F + j] * pb[j]; d += Dm[i * F + j] * qb[j]; }
            tq[i] = a - bb; tp[i] = c + d;
        }
        for (int i = 0; i < F; ++i) { qb[i] = tq[i]; pb[i] = tp[i]; }
        // traceless adiabatic derivative W~ = Z^T (dV - Dbar I) Z; Gamma, Xi (element [j, i] from the pair (i, j))
        for (int i = 0; i < F; ++i)
            for (int j = 0; j < F; ++j) {
                double sacc = 0.0;
                for (int a = 0; a < F; ++a)
                    for (int c = 0; c < F; ++c) sacc += Z[a * F + i] * (dV[a * F + c] - (a == c ? Db : 0.0)) * Z[c * F + j];
                W[i * F + j] = sacc;
            }
        for (int i = 0; i < F; ++i)
            for (int j = 0; j < F; ++j) {
                const double li = w[i] - Vb, lj = w[j] - Vb;
                if (i == j) { G[j * F + i] = W[i * F + j] * dt; X[j * F + i] = 0.0; }
                else {
                    G[j * F + i] = std::sin((li - lj) * dt) * W[i * F + j] / (li - lj);
                    X[j * F + i] = (1.0 - std::cos((li - lj) * dt)) * W[i * F + j] / (li - lj);
                }
            }
        mapping_transform(F, Z.data(), G.data());
        mapping_transform(F, Z.data(), X.data());
        double f = 0.0;
        for (int i = 0; i < F; ++i) {
            double eq = 0.0, ep = 0.0, fp = 0.0;
            for (int j = 0; j < F; ++j) { eq += G[i * F + j] * qb[j]; ep += G[i * F + j] * pb[j]; fp += X[i * F + j] * pb[j]; }
            f += 0.5 * (qb[i] * eq + pb[i] * ep) - qb[i] * fp;
        }
        v[b] -= f / D.mass;
        v[b] -= Db / D.mass * dt;
    }
    half();
}

// Initial mapping variables of ring polymers.  NRPMD (nrpmd.jl:49-65): theta uniform per (state, bead) from
// the Philox block (trajectory, b F + i, stream 4); q = R cos theta, p = R sin theta with R = sqrt(2 + 2 gamma)
// on the occupied state and sqrt(2 gamma) elsewhere.  Ring-polymer eCMM (rpecmm.jl:11-25): a point uniform on
// the sphere of radius sqrt(2 + 2 F gamma N) in 2 F N dimensions (normals z_2k, z_2k+1 from the block
// (trajectory, k, stream 4), normalised); q the first F N coordinates and p the rest, bead-major.
OR_API void oracle_rpmap_sample(const ORpMapDesc* D, int64_t ntraj, const int32_t* state, uint64_t key,
                                const int64_t* traj_ids, double* q, double* p) {
    const int P = D->nbeads, F = D->nstates, FN = F * P;
    std::vector<double> z(2 * FN);
    for (int64_t t = 0; t < ntraj; ++t) {
        const uint64_t id = traj_ids ? (uint64_t)traj_ids[t] : (uint64_t)t;
        if (D->method == 0) {
            for (int b = 0; b < P; ++b)
                for (int i = 0; i < F; ++i) {
                    const double th = 2.0 * M_PI * uniform01(id, (uint32_t)(b * F + i), 4u, key);
                    const double R = std::sqrt(i == state[t] ? 2.0 + 2.0 * D->gamma : 2.0 * D->gamma);
                    q[(t * P + b) * F + i] = R * std::cos(th);
                    p[(t * P + b) * F + i] = R * std::sin(th);
                }
        } else {
            double s = 0.0;
            for (int k = 0; k < FN; ++k) {
                normal2(id, (uint32_t)k, 4u, key, &z[2 * k], &z[2 * k + 1]);
                s += z[2 * k] * z[2 * k] + z[2 * k + 1] * z[2 * k + 1];
            }
            const double sc = std::sqrt(2.0 + 2.0 * F * D->gamma * P) / std::sqrt(s);
            for (int j = 0; j < FN; ++j) { q[t * FN + j] = z[j] * sc; p[t * FN + j] = z[FN + j] * sc; }
        }
    }
}

OR_API double oracle_rpmap_energy(const ORpMapDesc* D, const OModelDesc* m, const double* u) { return rpmap_energy(*D, *m, u); }
OR_API void oracle_rpmap_motion(const ORpMapDesc* D, const OModelDesc* m, const double* u, double* du) { rpmap_motion(*D, *m, u, du); }

// Ring polymers by RingPolymerMInt: r, v (ntraj, P), q, p (ntraj, P, F) in place; per save the diabatic
// populations (NRPMD: the bead mean of (q^2 + p^2)/2 - gamma, nrpmd.jl:110-121; eCMM: the bead sum of the
// inverse mapping kernel with N beads, rpecmm.jl:80-94), the kinetic energy, the potential energy with the
// springs, and the centroid.
// sinit (ntraj, F): the initial estimator at t0 (NRPMD: its population, Estimators.jl:121; eCMM: the bead sum of
// the mapping kernel (q^2 + p^2)/2 - gamma, rpecmm.jl:95-103).
OR_API int32_t oracle_rpmap_run(const ORpMapDesc* D, const OModelDesc* m, int64_t ntraj, double* r, double* v, double* q,
                                double* p, double* spop, double* sek, double* sep, double* scen, double* sinit,
                                int32_t nthreads) {
    const int P = D->nbeads, F = D->nstates, ns = D->nsave;
    const double wn = P * D->temperature;
    std::vector<double> U(P * P), Ch(4 * P);
    normal_mode_matrix(P, U.data());
    for (int k = 0; k < P; ++k) rpmap_half_cayley(matsubara(P, wn, k), D->dt, &Ch[4 * k]);
    const double g = D->gamma, NF = (double)P * F;
    parallel_for(ntraj, nthreads, [&](int64_t t) {
        double* rt = r + t * P; double* vt = v + t * P;
        double* qt = q + t * P * F; double* pt = p + t * P * F;
        auto save = [&](int slot) {
            if (slot >= ns) return;
            double ek = 0.0, ep = rpmap_spring(*D, rt), cen = 0.0;
            for (int b = 0; b < P; ++b) {
                ek += 0.5 * D->mass * vt[b] * vt[b];
                ep += rpmap_bead_potential(F, *m, rt[b], qt + b * F, pt + b * F);
                cen += rt[b];
            }
            for (int i = 0; i < F; ++i) {
                double pop = 0.0;
                for (int b = 0; b < P; ++b) {
                    const double x = qt[b * F + i], y = pt[b * F + i];
                    if (D->method == 0) pop += (x * x + y * y) / 2.0 - g;
                    else pop += (1.0 + NF) / ((1.0 + NF * g) * (1.0 + NF * g)) * (x * x + y * y) / 2.0 - (1.0 - g) / (1.0 + NF * g);
                }
                spop[((size_t)t * ns + slot) * F + i] = D->method == 0 ? pop / P : pop;
            }
            sek[(size_t)t * ns + slot] = ek;
            sep[(size_t)t * ns + slot] = ep;
            scen[(size_t)t * ns + slot] = cen / P;
        };
        for (int i = 0; i < F; ++i) {
            double k = 0.0;
            for (int b = 0; b < P; ++b) k += (qt[b * F + i] * qt[b * F + i] + pt[b * F + i] * pt[b * F + i]) / 2.0 - g;
            sinit[(size_t)t * F + i] = D->method == 0 ? k / P : k;
        }
        save(0);
        for (int64_t step = 0; step < D->nsteps; ++step) {
            rpmap_step(*D, *m, U, Ch, rt, vt, qt, pt);
            if (D->save_every > 0 && (step + 1) % D->save_every == 0) save((int)((step + 1) / D->save_every));
        }
    });
    return 0;
}

// ---------------------------------------------------------------------------
// Independent-electron surface hopping (IESH) and EhrenfestNA on a Newns-Anderson
// (impurity + discretised bath) model.
//   Model (NQCModels AndersonHolstein with a MiaoSubotnik impurity and a TrapezoidalRule
//   bath; docs/src/NQCModels/systembathmodels.md "Newns-Anderson model"):
//     U0(x) = m w^2 x^2 / 2 (state independent), h(x) = m w^2 (x - g)^2 / 2 + dG - U0(x),
//     V = [[h, V_1 .. V_M], [V_k, eps_k delta_kl]], eps_k = a + (k-1)(b-a)/(M-1),
//     V_k = sqrt(Gamma / 2 pi) sqrt((b-a)/(M-1)).
//   Step: VerletwithElectronics (verlet_with_electronics.jl:42-69): B, A, update_cache!,
//   acceleration from the occupied states (iesh.jl:190-207), B, wavefunction propagation
//   psi <- U exp(-i Lambda dt) U^H psi with H = diag(E) - i v.d at (r', v')
//   (wavefunction_dynamics.jl:15-58).  Hop (iesh.jl:231-335,390-407): one uniform for the
//   estimate and the selection, determinants by LU (FastDeterminant.jl:20-22).
// ---------------------------------------------------------------------------
struct ONADesc {
    double mw2;      // m_model * omega^2
    double g, dG, Gamma, a, b;
    int32_t M;       // bath states (N = M + 1)
    int32_t pad;
};

static void na_matrices(const ONADesc& md, double x, std::vector<double>& V, double& dh, double& U0, double& dU0) {
    const int N = md.M + 1;
    V.assign((size_t)N * N, 0.0);
    const double de = (md.b - md.a) / (md.M - 1);
    const double Vk = std::sqrt(md.Gamma / (2.0 * M_PI)) * std::sqrt(de);
    U0 = 0.5 * md.mw2 * x * x;
    dU0 = md.mw2 * x;
    const double U1 = 0.5 * md.mw2 * (x - md.g) * (x - md.g) + md.dG;
    V[0] = U1 - U0;
    dh = md.mw2 * (x - md.g) - md.mw2 * x;
    for (int k = 1; k < N; ++k) {
        V[(size_t)k * N + k] = md.a + (k - 1) * de;
        V[k] = V[(size_t)k * N] = Vk;
    }
}

// Complex Hermitian cyclic Jacobi (stands in for LAPACK zheevr); eigenvalues ascending,
// U columns orthonormal eigenvectors.
static void eigen_hermitian(int n, std::vector<cd> A, std::vector<double>& E, std::vector<cd>& U) {
    U.assign((size_t)n * n, cd(0.0, 0.0));
    for (int i = 0; i < n; ++i) U[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) {
            const double a2 = std::norm(A[(size_t)i * n + j]);
            tot += a2; if (i != j) off += a2;
        }
        if (off <= 1e-32 * tot || off == 0.0) break;
        for (int p = 0; p < n - 1; ++p) for (int q = p + 1; q < n; ++q) {
            const cd apq = A[(size_t)p * n + q];
            const double mag = std::abs(apq);
            if (mag == 0.0) continue;
            const cd ph = apq / mag;  // a_pq = |a_pq| e^{i phi}
            const double app = A[(size_t)p * n + p].real(), aqq = A[(size_t)q * n + q].real();
            const double theta = (aqq - app) / (2.0 * mag);
            const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
            const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
            // J = [[c, s ph], [-s conj(ph), c]] acting on columns p, q: A <- J^H A J
            for (int k = 0; k < n; ++k) {
                const cd akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                A[(size_t)k * n + p] = c * akp - s * std::conj(ph) * akq;
                A[(size_t)k * n + q] = s * ph * akp + c * akq;
            }
            for (int k = 0; k < n; ++k) {
                const cd apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                A[(size_t)p * n + k] = c * apk - s * ph * aqk;
                A[(size_t)q * n + k] = s * std::conj(ph) * apk + c * aqk;
            }
            for (int k = 0; k < n; ++k) {
                const cd ukp = U[(size_t)k * n + p], ukq = U[(size_t)k * n + q];
                U[(size_t)k * n + p] = c * ukp - s * std::conj(ph) * ukq;
                U[(size_t)k * n + q] = s * ph * ukp + c * ukq;
            }
        }
    }
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return A[(size_t)x * n + x].real() < A[(size_t)y * n + y].real(); });
    std::vector<cd> Us(U);
    E.assign(n, 0.0);
    for (int j = 0; j < n; ++j) {
        E[j] = A[(size_t)idx[j] * n + idx[j]].real();
        for (int i = 0; i < n; ++i) U[(size_t)i * n + j] = Us[(size_t)i * n + idx[j]];
    }
}

OR_API void oracle_eigen_hermitian(int32_t n, const double* Hre, const double* Him, double* E, double* Ure, double* Uim) {
    std::vector<cd> A((size_t)n * n);
    for (int i = 0; i < n * n; ++i) A[i] = cd(Hre[i], Him[i]);
    std::vector<double> e; std::vector<cd> U;
    eigen_hermitian(n, A, e, U);
    for (int i = 0; i < n; ++i) E[i] = e[i];
    for (int i = 0; i < n * n; ++i) { Ure[i] = U[i].real(); Uim[i] = U[i].imag(); }
}

// LU determinant with partial pivoting (FastDeterminant.det!, FastDeterminant.jl:20-22).
static cd det_lu(int n, std::vector<cd> A) {
    cd det = 1.0;
    for (int k = 0; k < n; ++k) {
        int piv = k; double best = std::abs(A[(size_t)k * n + k]);
        for (int i = k + 1; i < n; ++i) { const double a = std::abs(A[(size_t)i * n + k]); if (a > best) { best = a; piv = i; } }
        if (best == 0.0) return 0.0;
        if (piv != k) { for (int j = 0; j < n; ++j) std::swap(A[(size_t)k * n + j], A[(size_t)piv * n + j]); det = -det; }
        const cd d = A[(size_t)k * n + k];
        det *= d;
        for (int i = k + 1; i < n; ++i) {
            const cd f = A[(size_t)i * n + k] / d;
            for (int j = k + 1; j < n; ++j) A[(size_t)i * n + j] -= f * A[(size_t)k * n + j];
        }
    }
    return det;
}

OR_API void oracle_det(int32_t n, const double* re, const double* im, double* out) {
    std::vector<cd> A((size_t)n * n);
    for (int i = 0; i < n * n; ++i) A[i] = cd(re[i], im[i]);
    const cd d = det_lu(n, A);
    out[0] = d.real(); out[1] = d.imag();
}

struct OIEDesc {
    int32_t method;      // 0 = IESH, 1 = EhrenfestNA
    int32_t nelec;
    int32_t rescaling;
    int32_t estimate;    // estimate_probability (iesh.jl:251-254)
    double dt, t0, mass;
    int64_t nsteps;
    uint64_t key;
    int32_t edc;         // DecoherenceCorrectionEDC applied after each step (iesh.jl:424-441)
    int32_t pad;
    double edc_C;
    int32_t save_every, nsave;
};

// IESH hopping probabilities prob[m * Ne + n] for electron n -> unoccupied state m
// (iesh.jl:231-283): S_ji = psi[occ_j, i], A_kk = |det S|^2, v_dot_d[m, n] = -v d[m, occ_n],
// P = 2 dt Re(det S conj(det S_{n<-m})) v_dot_d / A_kk, each clamped to [0, 1].  Returns false when
// the early-exit estimate (iesh.jl:241-244) is below the uniform, leaving prob all zero.
static bool ie_probabilities(int N, int Ne, const std::vector<cd>& psi, const std::vector<int>& occ,
                             const std::vector<double>& d, double v, double dt, bool estimate, double rnd,
                             std::vector<double>& prob, bool clamp = true) {
    prob.assign((size_t)N * Ne, 0.0);
    std::vector<cd> S((size_t)Ne * Ne);
    for (int j = 0; j < Ne; ++j) for (int e = 0; e < Ne; ++e) S[(size_t)j * Ne + e] = psi[(size_t)occ[j] * Ne + e];
    const cd det_cur = det_lu(Ne, S);
    const double Akk = std::norm(det_cur);
    const double pref = 2.0 * dt / Akk;
    std::vector<char> isocc(N, 0);
    for (int j = 0; j < Ne; ++j) isocc[occ[j]] = 1;
    double sumvd = 0.0;
    for (int n = 0; n < Ne; ++n)
        for (int m = 0; m < N; ++m)
            if (!isocc[m]) sumvd += std::fabs(-v * d[(size_t)m * N + occ[n]]);
    if (estimate) {
        const double est = pref * sumvd * (std::fabs(det_cur.real()) + std::fabs(det_cur.imag()));
        if (est < rnd) return false;
    }
    for (int n = 0; n < Ne; ++n)
        for (int m = 0; m < N; ++m) {
            if (isocc[m]) continue;
            std::vector<cd> Sn(S);
            for (int e = 0; e < Ne; ++e) Sn[(size_t)n * Ne + e] = psi[(size_t)m * Ne + e];
            const cd Akj = det_cur * std::conj(det_lu(Ne, Sn));
            double p = pref * Akj.real() * (-v * d[(size_t)m * N + occ[n]]);
            if (clamp) p = std::min(std::max(p, 0.0), 1.0);
            prob[(size_t)m * Ne + n] = p;
        }
    return true;
}

OR_API void oracle_iesh_probabilities(int32_t N, int32_t Ne, const double* psi_re, const double* psi_im,
                                      const int32_t* occ, const double* d, double v, double dt, int32_t clamp,
                                      double* prob) {
    std::vector<cd> psi((size_t)N * Ne);
    for (int k = 0; k < N * Ne; ++k) psi[k] = cd(psi_re[k], psi_im[k]);
    std::vector<int> o(occ, occ + Ne);
    std::vector<double> dd(d, d + (size_t)N * N), p;
    ie_probabilities(N, Ne, psi, o, dd, v, dt, false, 0.0, p, clamp != 0);
    std::copy(p.begin(), p.end(), prob);
}

static void ie_run_one(const OIEDesc& R, const ONADesc& md, int64_t i, uint64_t gid, double r0, double v0,
                       const int32_t* occ0, const double* psi_re0, const double* psi_im0,
                       double* r_out, double* v_out, double* psi_re, double* psi_im, int32_t* occ_out,
                       int32_t* nhops_out, double* pop_di) {
    const int N = md.M + 1, Ne = R.nelec;
    std::vector<double> V, E(N), U((size_t)N * N), Up((size_t)N * N), G((size_t)N * N), d((size_t)N * N);
    double dh, U0, dU0;
    std::vector<cd> psi((size_t)N * Ne);
    for (int k = 0; k < N * Ne; ++k) psi[k] = cd(psi_re0[(size_t)i * N * Ne + k], psi_im0[(size_t)i * N * Ne + k]);
    std::vector<int> occ(occ0 + (size_t)i * Ne, occ0 + (size_t)i * Ne + Ne);
    bool have = false;
    double x = r0, v = v0, kacc;
    auto update = [&]() {
        na_matrices(md, x, V, dh, U0, dU0);
        if (have) Up = U;
        eigen_sym(N, V.data(), E.data(), U.data(), have ? Up.data() : nullptr);
        if (!have) {
            // Initial gauge: impurity component positive.  The impurity component 1/sqrt(1 + sum V^2/(E-eps_k)^2)
            // never vanishes, so this gauge is continuous in x and the overlap tracking below keeps it.
            // (The reference takes whatever signs zheevr/dsyevr return first, then tracks them.)
            for (int c = 0; c < N; ++c)
                if (U[c] < 0) for (int a = 0; a < N; ++a) U[(size_t)a * N + c] = -U[(size_t)a * N + c];
        }
        have = true;
        std::vector<double> dV((size_t)N * N, 0.0);
        dV[0] = dh;
        adiabatic_quantities(N, 1, E.data(), U.data(), dV.data(), G.data(), d.data());
    };
    auto accel = [&]() {
        double f = dU0;
        if (R.method == 0) { for (int e = 0; e < Ne; ++e) f += G[(size_t)occ[e] * N + occ[e]]; }
        else {  // EhrenfestNA (ehrenfest_na.jl:72-86): sum over electrons of psi^H G psi
            for (int e = 0; e < Ne; ++e)
                for (int a = 0; a < N; ++a) for (int b = 0; b < N; ++b)
                    f += (std::conj(psi[(size_t)a * Ne + e]) * G[(size_t)a * N + b] * psi[(size_t)b * Ne + e]).real();
        }
        return -f / R.mass;
    };
    auto save_pop = [&](int slot) {
        if (!pop_di || slot >= R.nsave) return;
        double* out = pop_di + ((size_t)i * R.nsave + slot) * N;
        for (int a = 0; a < N; ++a) {  // diabatic state a; U[a * N + b] = component a of eigenvector b
            double p = 0.0;
            if (R.method == 1) {  // EhrenfestNA mean field: sum_e |(U psi)_{a e}|^2
                for (int e = 0; e < Ne; ++e) {
                    cd sum = 0.0;
                    for (int b = 0; b < N; ++b) sum += U[(size_t)a * N + b] * psi[(size_t)b * Ne + e];
                    p += std::norm(sum);
                }
            } else {  // iesh_diabatic_population (iesh.jl:351-369): diag(U rho_e U^T), rho_e = offdiag(c c^T) + e_s e_s^T
                for (int e = 0; e < Ne; ++e) {
                    double s1 = 0.0, s2 = 0.0;
                    for (int b = 0; b < N; ++b) {
                        const double c = psi[(size_t)b * Ne + e].real(), u = U[(size_t)a * N + b];
                        s1 += u * c;
                        s2 += u * u * c * c;
                    }
                    const double us = U[(size_t)a * N + occ[e]];
                    p += s1 * s1 - s2 + us * us;
                }
            }
            out[a] = p;
        }
    };
    update();
    kacc = accel();
    save_pop(0);
    int nh = 0;
    double t = R.t0;
    std::vector<cd> H((size_t)N * N), P((size_t)N * N), tmp((size_t)N * Ne);
    std::vector<double> Eh; std::vector<cd> Uh;
    std::vector<cd> S((size_t)Ne * Ne);
    for (int64_t step = 0; step < R.nsteps; ++step) {
        const double vt = v + (R.dt / 2) * kacc;
        x = x + R.dt * vt;
        update();
        if (R.method == 1) {
            // EhrenfestNA uses psi of the step start (verlet_with_electronics.jl:76-79)
        }
        kacc = accel();
        v = vt + (R.dt / 2) * kacc;
        // H = diag(E) - i d v; P = U e^{-i Lambda dt} U^H; psi <- P psi
        for (int a = 0; a < N; ++a) for (int b = 0; b < N; ++b)
            H[(size_t)a * N + b] = cd((a == b) ? E[a] : 0.0, 0.0) - cd(0.0, 1.0) * (d[(size_t)a * N + b] * v);
        eigen_hermitian(N, H, Eh, Uh);
        for (int a = 0; a < N; ++a) for (int b = 0; b < N; ++b) {
            cd s = 0.0;
            for (int k = 0; k < N; ++k) s += Uh[(size_t)a * N + k] * std::exp(cd(0.0, -Eh[k] * R.dt)) * std::conj(Uh[(size_t)b * N + k]);
            P[(size_t)a * N + b] = s;
        }
        for (int a = 0; a < N; ++a) for (int e = 0; e < Ne; ++e) {
            cd s = 0.0;
            for (int b = 0; b < N; ++b) s += P[(size_t)a * N + b] * psi[(size_t)b * Ne + e];
            tmp[(size_t)a * Ne + e] = s;
        }
        psi = tmp;
        t += R.dt;
        // the save precedes the hop and decoherence callbacks (the rule of run_one)
        if (R.save_every > 0 && (step + 1) % R.save_every == 0) save_pop((int)((step + 1) / R.save_every));
        if (R.method == 0) {
            const double rnd = uniform01(gid, (uint32_t)step, STREAM_HOP, R.key);
            // hopping probabilities (iesh.jl:231-283), selection (iesh.jl:318-335)
            std::vector<double> prob;
            const bool evaluate = ie_probabilities(N, Ne, psi, occ, d, v, R.dt, R.estimate != 0, rnd, prob);
            if (evaluate) {
                std::vector<char> isocc(N, 0);
                for (int j = 0; j < Ne; ++j) isocc[occ[j]] = 1;
                double cum = 0.0;
                int hn = -1, hm = -1;
                for (int n = 0; n < Ne && hn < 0; ++n)
                    for (int m = 0; m < N; ++m) {
                        if (isocc[m]) continue;
                        cum += prob[(size_t)m * Ne + n];
                        if (rnd < cum) { hn = n; hm = m; break; }
                    }
                if (hn >= 0) {
                    const int old = occ[hn];
                    bool ok = true;
                    if (R.rescaling != RESCALE_OFF) {
                        const double dd = d[(size_t)hm * N + old];  // coupling[new, old]
                        const double aq = dd * dd / R.mass / 2.0, bq = dd * v, cq = E[hm] - E[old];
                        const double disc = bq * bq - 4.0 * aq * cq;
                        if (disc < 0) {
                            ok = false;
                            if (R.rescaling == RESCALE_VINVERSION) v -= 2.0 * (v * (dd / std::fabs(dd))) * (dd / std::fabs(dd));
                        } else {
                            const double root = std::sqrt(disc);
                            const double gam = (bq < 0) ? (bq + root) / (2.0 * aq) : (bq - root) / (2.0 * aq);
                            v -= gam * dd / R.mass;
                        }
                    }
                    if (ok) { occ[hn] = hm; ++nh; }
                }
            }
            if (R.edc) {  // Granucci-Persico EDC on each electron (decoherence_corrections.jl:21-39)
                const double Ekin = 0.5 * R.mass * v * v;
                for (int e = 0; e < Ne; ++e) {
                    const int s = occ[e];
                    double nrm = 0.0;
                    for (int a = 0; a < N; ++a) {
                        if (a == s) continue;
                        const double tau = (1.0 + R.edc_C / Ekin) / std::fabs(E[a] - E[s]);
                        psi[(size_t)a * Ne + e] *= std::exp(-R.dt / tau);
                        nrm += std::norm(psi[(size_t)a * Ne + e]);
                    }
                    const cd cm = psi[(size_t)s * Ne + e];
                    psi[(size_t)s * Ne + e] = cm * std::sqrt((1.0 - nrm) / std::norm(cm));
                }
            }
        }
    }
    r_out[i] = x; v_out[i] = v;
    for (int k = 0; k < N * Ne; ++k) { psi_re[(size_t)i * N * Ne + k] = psi[k].real(); psi_im[(size_t)i * N * Ne + k] = psi[k].imag(); }
    for (int e = 0; e < Ne; ++e) occ_out[(size_t)i * Ne + e] = occ[e];
    nhops_out[i] = nh;
}

// Ring-polymer IESH / EhrenfestNA with BCBWavefunction (bcb_wavefunction.jl:37-69), literally:
// initialize! takes the bead accelerations at r (rpiesh.jl:20-38 with the occupations,
// rpehrenfest_na.jl:14-40 with psi); perform_step!: vtmp = v + dt/2 k; r, vtmp to normal modes,
// step_C! (Cayley, half = false), back; k = accelerations at the new beads (EhrenfestNA with psi of
// the step start); v = vtmp + dt/2 k; then psi <- exp(-i dt H) psi with H = diag(E) - i d v at the
// centroid of the step-start positions and velocities (get_quantum_propagator(sim, vprev, rprev),
// wavefunction_dynamics.jl:27-58 with the centroid quantities of SurfaceHoppingMethods.jl:85-103).
// The IESH hop (iesh.jl:390-407) uses the centroid eigenpairs, coupling and velocity at the step end;
// an accepted hop rescales every bead by gamma d / m and a frustrated one with vinversion reverses the
// centroid velocity along d (rpsh.jl:24-44).  Eigenvectors in the impurity-positive gauge at every
// bead and at the centroid.  r0, v0, r_out, v_out: (ntraj, P); pop_di from the centroid eigenvectors.
static void ie_rp_run_one(const OIEDesc& R, const ONADesc& md, int P, double temperature, int64_t i, uint64_t gid,
                          const double* r0, const double* v0, const int32_t* occ0, const double* psi_re0,
                          const double* psi_im0, double* r_out, double* v_out, double* psi_re, double* psi_im,
                          int32_t* occ_out, int32_t* nhops_out, double* pop_di) {
    const int N = md.M + 1, Ne = R.nelec;
    std::vector<double> Tm((size_t)P * P), wk(P), Cy((size_t)4 * P);
    normal_mode_matrix(P, Tm.data());
    for (int k = 0; k < P; ++k) { wk[k] = matsubara(P, P * temperature, k); cayley(wk[k], R.dt, Cy.data() + 4 * k); }
    std::vector<double> r(r0 + (size_t)i * P, r0 + (size_t)i * P + P), v(v0 + (size_t)i * P, v0 + (size_t)i * P + P);
    std::vector<cd> psi((size_t)N * Ne);
    for (int k = 0; k < N * Ne; ++k) psi[k] = cd(psi_re0[(size_t)i * N * Ne + k], psi_im0[(size_t)i * N * Ne + k]);
    std::vector<int> occ(occ0 + (size_t)i * Ne, occ0 + (size_t)i * Ne + Ne);
    // eigenpairs, U^T dV U and the couplings at position x, impurity components positive
    auto eig = [&](double x, std::vector<double>& E, std::vector<double>& U, std::vector<double>& G, std::vector<double>& d,
                   double& dU0) {
        std::vector<double> V; double dh, U0;
        na_matrices(md, x, V, dh, U0, dU0);
        E.assign(N, 0.0); U.assign((size_t)N * N, 0.0); G.assign((size_t)N * N, 0.0); d.assign((size_t)N * N, 0.0);
        eigen_sym(N, V.data(), E.data(), U.data(), nullptr);
        for (int c = 0; c < N; ++c)
            if (U[c] < 0) for (int a = 0; a < N; ++a) U[(size_t)a * N + c] = -U[(size_t)a * N + c];
        std::vector<double> dV((size_t)N * N, 0.0);
        dV[0] = dh;
        adiabatic_quantities(N, 1, E.data(), U.data(), dV.data(), G.data(), d.data());
    };
    auto mean = [&](const std::vector<double>& a) { double s = 0.0; for (int b = 0; b < P; ++b) s += a[b]; return s / P; };
    std::vector<double> Eb, Ub, Gb, db, k(P);
    auto accels = [&](const std::vector<double>& rr, const std::vector<cd>& ps) {
        for (int b = 0; b < P; ++b) {
            double dU0;
            eig(rr[b], Eb, Ub, Gb, db, dU0);
            double f = dU0;
            if (R.method == 0) { for (int e = 0; e < Ne; ++e) f += Gb[(size_t)occ[e] * N + occ[e]]; }
            else {
                for (int e = 0; e < Ne; ++e)
                    for (int a = 0; a < N; ++a) for (int c = 0; c < N; ++c)
                        f += (std::conj(ps[(size_t)a * Ne + e]) * Gb[(size_t)a * N + c] * ps[(size_t)c * Ne + e]).real();
            }
            k[b] = -f / R.mass;
        }
    };
    std::vector<double> E, U, G, d;
    double dU0c;
    eig(mean(r), E, U, G, d, dU0c);
    auto save_pop = [&](int slot) {
        if (!pop_di || slot >= R.nsave) return;
        double* out = pop_di + ((size_t)i * R.nsave + slot) * N;
        for (int a = 0; a < N; ++a) {
            double pp = 0.0;
            if (R.method == 1) {
                for (int e = 0; e < Ne; ++e) {
                    cd sum = 0.0;
                    for (int b = 0; b < N; ++b) sum += U[(size_t)a * N + b] * psi[(size_t)b * Ne + e];
                    pp += std::norm(sum);
                }
            } else {
                for (int e = 0; e < Ne; ++e) {
                    double s1 = 0.0, s2 = 0.0;
                    for (int b = 0; b < N; ++b) {
                        const double c = psi[(size_t)b * Ne + e].real(), u = U[(size_t)a * N + b];
                        s1 += u * c;
                        s2 += u * u * c * c;
                    }
                    const double us = U[(size_t)a * N + occ[e]];
                    pp += s1 * s1 - s2 + us * us;
                }
            }
            out[a] = pp;
        }
    };
    accels(r, psi);
    save_pop(0);
    int nh = 0;
    std::vector<cd> H((size_t)N * N), Pm((size_t)N * N), tmp((size_t)N * Ne);
    std::vector<double> Eh, q(P), pmode(P), vt(P);
    std::vector<cd> Uh;
    for (int64_t step = 0; step < R.nsteps; ++step) {
        const std::vector<double> E0 = E, d0 = d;  // centroid quantities of the step start
        const double vc0 = mean(v);
        const std::vector<cd> psi0 = psi;
        for (int b = 0; b < P; ++b) vt[b] = v[b] + (R.dt / 2) * k[b];                       // step_B!
        for (int m = 0; m < P; ++m) {                                                         // to normal modes
            double a1 = 0.0, a2 = 0.0;
            for (int b = 0; b < P; ++b) { a1 += Tm[(size_t)b * P + m] * r[b]; a2 += Tm[(size_t)b * P + m] * vt[b]; }
            q[m] = a1; pmode[m] = a2;
        }
        for (int m = 0; m < P; ++m) {                                                         // step_C!
            const double* c = Cy.data() + 4 * m;
            const double qn = c[0] * q[m] + c[1] * pmode[m], pn = c[2] * q[m] + c[3] * pmode[m];
            q[m] = qn; pmode[m] = pn;
        }
        for (int b = 0; b < P; ++b) {                                                         // back
            double a1 = 0.0, a2 = 0.0;
            for (int m = 0; m < P; ++m) { a1 += Tm[(size_t)b * P + m] * q[m]; a2 += Tm[(size_t)b * P + m] * pmode[m]; }
            r[b] = a1; vt[b] = a2;
        }
        accels(r, psi0);
        for (int b = 0; b < P; ++b) v[b] = vt[b] + (R.dt / 2) * k[b];                        // step_B!
        // psi <- exp(-i dt H(rprev, vprev)) psi
        for (int a = 0; a < N; ++a) for (int b = 0; b < N; ++b)
            H[(size_t)a * N + b] = cd((a == b) ? E0[a] : 0.0, 0.0) - cd(0.0, 1.0) * (d0[(size_t)a * N + b] * vc0);
        eigen_hermitian(N, H, Eh, Uh);
        for (int a = 0; a < N; ++a) for (int b = 0; b < N; ++b) {
            cd sm = 0.0;
            for (int kk = 0; kk < N; ++kk) sm += Uh[(size_t)a * N + kk] * std::exp(cd(0.0, -Eh[kk] * R.dt)) * std::conj(Uh[(size_t)b * N + kk]);
            Pm[(size_t)a * N + b] = sm;
        }
        for (int a = 0; a < N; ++a) for (int e = 0; e < Ne; ++e) {
            cd sm = 0.0;
            for (int b = 0; b < N; ++b) sm += Pm[(size_t)a * N + b] * psi0[(size_t)b * Ne + e];
            tmp[(size_t)a * Ne + e] = sm;
        }
        psi = tmp;
        eig(mean(r), E, U, G, d, dU0c);                                                       // centroid at the step end
        // the save precedes the hop callback (the rule of run_one)
        if (R.save_every > 0 && (step + 1) % R.save_every == 0) save_pop((int)((step + 1) / R.save_every));
        if (R.method == 0) {
            const double rnd = uniform01(gid, (uint32_t)step, STREAM_HOP, R.key);
            const double vc = mean(v);
            std::vector<double> prob;
            const bool evaluate = ie_probabilities(N, Ne, psi, occ, d, vc, R.dt, R.estimate != 0, rnd, prob);
            if (evaluate) {
                std::vector<char> isocc(N, 0);
                for (int j = 0; j < Ne; ++j) isocc[occ[j]] = 1;
                double cum = 0.0;
                int hn = -1, hm = -1;
                for (int n = 0; n < Ne && hn < 0; ++n)
                    for (int m = 0; m < N; ++m) {
                        if (isocc[m]) continue;
                        cum += prob[(size_t)m * Ne + n];
                        if (rnd < cum) { hn = n; hm = m; break; }
                    }
                if (hn >= 0) {
                    const int old = occ[hn];
                    bool ok = true;
                    if (R.rescaling != RESCALE_OFF) {
                        const double dd = d[(size_t)hm * N + old];
                        const double aq = dd * dd / R.mass / 2.0, bq = dd * vc, cq = E[hm] - E[old];
                        const double disc = bq * bq - 4.0 * aq * cq;
                        if (disc < 0) {
                            ok = false;
                            if (R.rescaling == RESCALE_VINVERSION) {
                                const double dn = dd / std::fabs(dd), g2 = vc * dn;
                                for (int b = 0; b < P; ++b) v[b] -= 2.0 * g2 * dn;
                            }
                        } else {
                            const double root = std::sqrt(disc);
                            const double gam = (bq < 0) ? (bq + root) / (2.0 * aq) : (bq - root) / (2.0 * aq);
                            for (int b = 0; b < P; ++b) v[b] -= gam * dd / R.mass;
                        }
                    }
                    if (ok) { occ[hn] = hm; ++nh; }
                }
            }
        }
    }
    for (int b = 0; b < P; ++b) { r_out[(size_t)i * P + b] = r[b]; v_out[(size_t)i * P + b] = v[b]; }
    for (int kk = 0; kk < N * Ne; ++kk) { psi_re[(size_t)i * N * Ne + kk] = psi[kk].real(); psi_im[(size_t)i * N * Ne + kk] = psi[kk].imag(); }
    for (int e = 0; e < Ne; ++e) occ_out[(size_t)i * Ne + e] = occ[e];
    nhops_out[i] = nh;
}

OR_API int32_t oracle_rp_iesh_run_ensemble(const OIEDesc* R, const ONADesc* md, int32_t P, double temperature,
                                           int64_t ntraj, const double* r0, const double* v0, const int32_t* occ0,
                                           const double* psi_re0, const double* psi_im0, double* r, double* v,
                                           double* psi_re, double* psi_im, int32_t* occ, int32_t* nhops, double* pop_di,
                                           int32_t nthreads) {
    parallel_for(ntraj, nthreads, [&](int64_t i) {
        ie_rp_run_one(*R, *md, P, temperature, i, (uint64_t)i, r0, v0, occ0, psi_re0, psi_im0, r, v, psi_re, psi_im,
                      occ, nhops, pop_di);
    });
    return 0;
}

OR_API int32_t oracle_iesh_run_ensemble(const OIEDesc* R, const ONADesc* md, int64_t ntraj, const double* r0,
                                        const double* v0, const int32_t* occ0, const double* psi_re0,
                                        const double* psi_im0, double* r, double* v, double* psi_re, double* psi_im,
                                        int32_t* occ, int32_t* nhops, double* pop_di, int32_t nthreads) {
    parallel_for(ntraj, nthreads, [&](int64_t i) {
        ie_run_one(*R, *md, i, (uint64_t)i, r0[i], v0[i], occ0, psi_re0, psi_im0, r, v, psi_re, psi_im, occ, nhops, pop_di);
    });
    return 0;
}

// sample_fermi_dirac_distribution with the Boltzmann factor (DynamicsUtils.jl:191-205), literally:
// state = the first n_e entries of `avail`; n_avail * n_e times pick a random slot of `state`
// (rand(eachindex(state))), a random element of setdiff(avail, state) in the order of `avail`, and
// replace the slot's state when exp(-beta (E_j - E_i)) > rand(); finally sort.  Julia's rand calls
// are replaced by Philox counter (traj, move, STREAM_OCCUPATION) = (o0, o1, o2, o3): slot
// (o0 n_e) >> 32, element (o1 n_free) >> 32, uniform u53(o2, o3).  With beta = inf the factor is
// IEEE's: 0 above, inf below, NaN (never accepted) for equal energies.  Returns 0, or -1 when
// fewer than n_e + 1 states are available (setdiff would be empty, as rand() of it throws).
OR_API int32_t oracle_fermi_dirac_occupations(const double* E, int32_t Ne, const int32_t* avail, int32_t navail,
                                              double beta, uint64_t traj, uint64_t key, int32_t* occ) {
    if (navail <= Ne) return -1;
    std::vector<int32_t> state(avail, avail + Ne), rest;
    const int64_t nmoves = (int64_t)navail * Ne;
    for (int64_t it = 0; it < nmoves; ++it) {
        uint32_t o[4];
        philox_draw(traj, (uint32_t)it, STREAM_OCCUPATION, key, o);
        const int ci = (int)(((uint64_t)o[0] * (uint64_t)Ne) >> 32);
        rest.clear();
        for (int32_t k = 0; k < navail; ++k)
            if (std::find(state.begin(), state.end(), avail[k]) == state.end()) rest.push_back(avail[k]);
        const int kk = (int)(((uint64_t)o[1] * (uint64_t)rest.size()) >> 32);
        const int i = state[ci], j = rest[kk];
        const double prob = std::exp(-beta * (E[j] - E[i]));
        if (prob > u53(o[2], o[3])) state[ci] = j;
    }
    std::sort(state.begin(), state.end());
    std::copy(state.begin(), state.end(), occ);
    return 0;
}

OR_API void oracle_na_matrix(const ONADesc* md, double x, double* Vout) {
    std::vector<double> V; double dh, U0, dU0;
    na_matrices(*md, x, V, dh, U0, dU0);
    std::copy(V.begin(), V.end(), Vout);
}

// DynamicsVariables(sim, v, r, FermiDiracState{Diabatic}) for IESH (iesh.jl:138-183), literally:
// diabatic_state = the chain above over the diagonal of the model matrix (Ediab); then for each
// electron i, adiabatic_density = U^T diagm(e_d) U (U[a * N + b] = component a of eigenvector b),
// adiabatic_population = its diagonal, s = sample(Weights(adiabatic_population)) until s is not
// yet in adiabatic_state, psi[:, i] = adiabatic_density[:, 1] ./ sqrt(adiabatic_population[1]);
// finally sort!(adiabatic_state) (psi's columns keep their order).  StatsBase's weighted sample
// draws t = rand() * sum(w) and returns the first index whose running sum reaches t (the last at
// most); rand() is Philox counter (traj, 2^31 + k, STREAM_OCCUPATION) for the k-th draw of the
// trajectory, u53(o0, o1).  psi is N x Ne row-major (real).  Returns 0, -1 as the chain, or -2
// after 2^20 draws for one electron.
OR_API int32_t oracle_fermi_dirac_diabatic(const double* Ediab, const double* U, int32_t N, int32_t Ne,
                                           const int32_t* avail, int32_t navail, double beta, uint64_t traj,
                                           uint64_t key, int32_t* occ, double* psi) {
    std::vector<int32_t> dstate(Ne);
    if (oracle_fermi_dirac_occupations(Ediab, Ne, avail, navail, beta, traj, key, dstate.data()) != 0) return -1;
    std::vector<int32_t> ad(Ne, -1);
    std::vector<double> dens((size_t)N * N), pop(N);
    uint32_t draw = 0;
    for (int i = 0; i < Ne; ++i) {
        const int d = dstate[i];
        for (int b = 0; b < N; ++b)
            for (int c = 0; c < N; ++c) {
                double acc = 0.0;
                for (int a = 0; a < N; ++a) acc += U[(size_t)a * N + b] * (a == d ? 1.0 : 0.0) * U[(size_t)a * N + c];
                dens[(size_t)b * N + c] = acc;
            }
        double sw = 0.0;
        for (int b = 0; b < N; ++b) { pop[b] = dens[(size_t)b * N + b]; sw += pop[b]; }
        int pick = -1;
        for (int tries = 0; tries < (1 << 20) && pick < 0; ++tries) {
            uint32_t o[4];
            philox_draw(traj, 0x80000000u + draw++, STREAM_OCCUPATION, key, o);
            const double t = u53(o[0], o[1]) * sw;
            int k = 0;
            double cw = pop[0];
            while (cw < t && k < N - 1) { ++k; cw += pop[k]; }
            if (std::find(ad.begin(), ad.end(), k) == ad.end()) pick = k;
        }
        if (pick < 0) return -2;
        ad[i] = pick;
        for (int b = 0; b < N; ++b) psi[(size_t)b * Ne + i] = dens[(size_t)b * N + 0] / std::sqrt(pop[0]);
    }
    std::sort(ad.begin(), ad.end());
    std::copy(ad.begin(), ad.end(), occ);
    return 0;
}

OR_API void oracle_na_eigen(const ONADesc* md, double x, double* E, double* U, double* dh) {
    std::vector<double> V; double U0, dU0;
    na_matrices(*md, x, V, *dh, U0, dU0);
    const int N = md->M + 1;
    eigen_sym(N, V.data(), E, U, nullptr);
}
