// iesh.hip -- independent-electron surface hopping (IESH) and EhrenfestNA on the Newns-Anderson
// impurity model, one trajectory per workgroup.
//
// Reference path: VerletwithElectronics (src/DynamicsMethods/IntegrationAlgorithms/
// verlet_with_electronics.jl:42-69), the wavefunction propagator (src/DynamicsUtils/
// wavefunction_dynamics.jl:15-58), the IESH force and hop (src/DynamicsMethods/SurfaceHoppingMethods/
// iesh.jl:190-207,231-335,390-441) and the EhrenfestNA force (src/DynamicsMethods/EhrenfestMethods/
// ehrenfest_na.jl:72-86).
//
// Model: N = M + 1 states, impurity h(x) = U1 - U0 coupled with a constant V to M bath levels
// eps_k = a + k (b - a)/(M - 1), V = sqrt(Gamma (b - a)/(M - 1) / 2 pi).  The Hamiltonian is an
// arrowhead matrix, so per step:
//   * eigen: root i of the secular function f(l) = l - h - V^2 sum_k 1/(l - eps_k) lies in
//     (eps_{i-1}, eps_i); rational-model iteration from the previous step's root, one thread per
//     root (3 to 5 evaluations of f per root and step).
//     The impurity component of eigenvector i is u_i = f'(E_i)^(-1/2) > 0 (a continuous gauge).
//   * dV/dx only touches the impurity entry, so U^T dV U = dh u u^T and the nonadiabatic coupling
//     is d_ij = dh u_i u_j / (E_j - E_i).
//   * propagator: psi <- exp(-i dt (diag(E) - i v d)) psi, applied as a Taylor series whose length
//     follows from the bound ||dt H|| <= dt max|E| + ||dt v d||_F (remainder below 1e-17); each term is one real GEMM
//     (-dt v dh W) [Re psi | Im psi] on FP64 MFMA 16x16x4 plus the diagonal part.  The reference
//     diagonalises H with zheevr and forms U exp(-i Lambda dt) U^H; both give exp(-i H dt) to
//     rounding.
//   * IESH hop: S = psi[occ, :], det S by LU (partial pivoting) on one wavefront, the reference's
//     early-exit estimate, and on the (rare) steps that pass it the replaced-row determinants by the
//     determinant lemma det(S with row n <- psi_m) = det(S) (psi_m S^-1)_n, so
//     P(n -> m) = 2 dt Re((psi_m S^-1)_n) v_dot_d[m, n]; cumulative selection electrons outer,
//     unoccupied states inner (iesh.jl:318-335).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "ensemble.h"
#include "handle.h"
#include "iesh.h"
#include "nqcd_device.h"

namespace nqcd {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kIeMaxNe = 64;
constexpr int kIeMaxStates = 128;
constexpr size_t kIeLdsBytes = 160 * 1024;  // LDS of one gfx950 compute unit

struct IeshArgs {
    int64_t ntraj;
    int64_t nsteps;
    int32_t N, M, Ne, cw;         // states, bath levels, electrons, psi columns in use (16 * column tiles)
    int32_t sb_len, pad0;         // doubles of the S / probability workspace in LDS
    int32_t method;               // 0 = IESH, 1 = EhrenfestNA
    int32_t rescaling, estimate, edc;
    int32_t save_every, nsave;
    int32_t P, pad1;              // ring-polymer beads (1: no ring polymer)
    double dt, mass, edc_C;
    double wn;                    // ring-polymer frequency P kT
    uint64_t key;
    double mw2, g, dG, V, a, de;  // model
    const double* r0; const double* v0; const int32_t* occ0; const double* psi0;  // psi0 may be null
    double* r; double* v; double* psi; int32_t* occ; int32_t* nhops; int32_t* steps;
    double* save_di;              // (nsave, N, ntraj)
    double* save_ad;              // (nsave, N, ntraj)
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Sum over the workgroup, same value returned to every thread; fixed reduction order.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < NT; ++j) s += red[j];
    return s;
}

template <int NT>
__device__ __forceinline__ double block_max(double v, double* red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = red[0];
#pragma unroll
    for (int j = 1; j < NT; ++j) s = fmax(s, red[j]);
    return s;
}

struct Scalars { double h, dh, dU0; };

__device__ __forceinline__ Scalars na_scalars(const IeshArgs& A, double x) {
    Scalars s;
    const double U0 = 0.5 * A.mw2 * x * x;
    const double U1 = 0.5 * A.mw2 * (x - A.g) * (x - A.g) + A.dG;
    s.h = U1 - U0;
    s.dh = A.mw2 * (x - A.g) - A.mw2 * x;
    s.dU0 = A.mw2 * x;
    return s;
}

// 1/x from the hardware estimate and two Newton-Raphson corrections (within an ulp).
__device__ __forceinline__ double rcp_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// Root i of the secular function f(l) = l - h - V^2 sum_k 1/(l - eps_k), which increases from
// -inf to +inf on its interval (lo, hi) = (eps_{i-1}, eps_i) (the outer roots have one finite end).
// Each iteration evaluates f, f' once and steps to the root of a two-pole rational model that
// matches f and f' at the current point ("middle way", R.-C. Li, LAPACK Working Note 89):
//   F(y) = c + bl / (y - eps_{i-1}) + bh / (y - eps_i),
// where the poles below and above the interval carry their own parts of f' and the slope of the
// linear term is shared between them; the model's root is a quadratic's, and a step that leaves
// the bracket (kept from the signs of f) bisects instead.  Stops when |f| is at its rounding level,
// when the step is a few ulps or when the bracket is narrower than tol.  Returns E_i and
// f'(E_i) = 1 + V^2 sum_k 1/(E_i - eps_k)^2.
// G lanes (1, 2 or 4 consecutive lanes, sub = lane mod G) share a root: lane sub sums the poles
// k = sub, sub + G, ..., and the partial sums are combined by DPP exchanges whose results have
// the same bits in every lane of the group, so the group iterates in lockstep.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
    const long long bits = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(bits & 0xffffffffll), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(bits >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

template <int G>
__device__ __forceinline__ double group_sum(double v) {
    if (G >= 2) v += dpp_f64<0xB1>(v);  // quad_perm [1,0,3,2]
    if (G >= 4) v += dpp_f64<0x4E>(v);  // quad_perm [2,3,0,1]
    return v;
}

template <int G = 1>
__device__ double secular_root(const double* eps, int M, int i, double V2, double h, double lo, double hi,
                               double guess, double tol, double& fp, int sub = 0) {
    const bool has_lo = i > 0, has_hi = i < M;
    const double dl = has_lo ? eps[i - 1] : lo, dh = has_hi ? eps[i] : hi;
    const double wl = has_lo ? (has_hi ? 0.5 : 1.0) : 0.0;  // share of the linear term's slope at the lower pole
    double lam = (guess > lo && guess < hi) ? guess : 0.5 * (lo + hi);
    fp = 1.0;
    for (int it = 0; it < 100; ++it) {
        // q_k > 0 below the interval, < 0 above it
        double sl = 0.0, sl2 = 0.0, sh = 0.0, sh2 = 0.0;
        for (int k = sub; k < M; k += G) {
            const double q = rcp_nr(lam - eps[k]);
            const double q2 = q * q;
            const bool below = k < i;
            sl += below ? q : 0.0;
            sh += below ? 0.0 : q;
            sl2 += below ? q2 : 0.0;
            sh2 += below ? 0.0 : q2;
        }
        if (G > 1) {
            sl = group_sum<G>(sl);
            sh = group_sum<G>(sh);
            sl2 = group_sum<G>(sl2);
            sh2 = group_sum<G>(sh2);
        }
        const double f = lam - h - V2 * (sl + sh);
        fp = 1.0 + V2 * (sl2 + sh2);
        if (fabs(f) <= 8.0 * __DBL_EPSILON__ * (fabs(lam) + fabs(h) + V2 * (sl - sh))) break;
        if (f > 0.0) hi = lam;
        else lo = lam;
        const double bl = has_lo ? -(V2 * sl2 + wl) * (lam - dl) * (lam - dl) : 0.0;
        const double bh = has_hi ? -(V2 * sh2 + (1.0 - wl)) * (lam - dh) * (lam - dh) : 0.0;
        const double c = f - (has_lo ? bl / (lam - dl) : 0.0) - (has_hi ? bh / (lam - dh) : 0.0);
        double nl;
        if (has_lo && has_hi) {
            // c t (t - D) + bl (t - D) + bh t = 0 with t = y - dl, D = dh - dl; root in (0, D)
            const double D = dh - dl;
            const double a1 = bl + bh - c * D, a0 = -bl * D;
            const double sq = sqrt(fmax(a1 * a1 - 4.0 * c * a0, 0.0));
            const double den = a1 > 0.0 ? -a1 - sq : -a1 + sq;
            nl = dl + den / (2.0 * c);
            if (!(nl >= lo && nl <= hi)) nl = dl + 2.0 * a0 / den;
        } else if (has_hi) {
            nl = dh - bh / c;
        } else {
            nl = dl - bl / c;
        }
        if (fabs(nl - lam) <= 4.0 * __DBL_EPSILON__ * fabs(lam) || (hi - lo) <= tol) break;
        if (!(nl > lo && nl < hi)) nl = 0.5 * (lo + hi);
        lam = nl;
    }
    return lam;
}

// Orders this wavefront's LDS accesses: its lanes exchange data through LDS without a workgroup
// barrier (a wavefront's LDS instructions execute in program order).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// One step of a wavefront arg-max over (mag, idx) with the DPP lane pattern CTRL (larger mag
// wins, the lower idx on ties).
// Eigenvalue ir of the Newns-Anderson matrix at impurity energy h: the secular root on its
// bracket (the outer brackets reach past the band and h by the coupling norm zn = sqrt(M) V);
// warm starts from `guess`, a cold start from below the bracket.  Also returns f'(E_ir).
template <int G = 1>
__device__ __forceinline__ double na_root(const double* eps, int M, double a, double V2, double zn, double h, int ir,
                                          bool warm, double guess, double& fp, int sub = 0) {
    const double lo = (ir == 0) ? fmin(h, eps[0]) - zn - 1e-3 * (zn + 1e-300) : eps[ir - 1];
    const double hi = (ir == M) ? fmax(h, eps[M - 1]) + zn + 1e-3 * (zn + 1e-300) : eps[ir];
    const double scale = fmax(fmax(fabs(a), fabs(eps[M - 1])), fabs(h)) + zn;
    return secular_root<G>(eps, M, ir, V2, h, lo, hi, warm ? guess : lo - 1.0, 4.0e-16 * scale, fp, sub);
}

template <int CTRL>
__device__ __forceinline__ void argmax_dpp(double& mag, int& idx) {
    const long long bits = __double_as_longlong(mag);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(bits & 0xffffffffll), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(bits >> 32), CTRL, 0xF, 0xF, false);
    const int oi = __builtin_amdgcn_update_dpp(0, idx, CTRL, 0xF, 0xF, false);
    const double om = __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
    if (om > mag || (om == mag && oi < idx)) { mag = om; idx = oi; }
}

__device__ __forceinline__ double readlane_f64(double x, int lane) {
    const long long bits = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)(bits & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Wavefront-uniform arg-max of (mag, idx), the lowest idx among equal maxima: butterflies inside
// each row of 16 lanes (quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror, row_mirror), then the
// four row results compared in lane order.
__device__ __forceinline__ void wave_argmax(double& mag, int& idx) {
    argmax_dpp<0xB1>(mag, idx);
    argmax_dpp<0x4E>(mag, idx);
    argmax_dpp<0x141>(mag, idx);
    argmax_dpp<0x140>(mag, idx);
    double bm = readlane_f64(mag, 0);
    int bi = __builtin_amdgcn_readlane(idx, 0);
#pragma unroll
    for (int row = 1; row < 4; ++row) {
        const double m = readlane_f64(mag, 16 * row);
        const int i = __builtin_amdgcn_readlane(idx, 16 * row);
        if (m > bm || (m == bm && i < bi)) { bm = m; bi = i; }
    }
    mag = bm;
    idx = bi;
}

// det of the Ne x Ne complex matrix held as real / imaginary planes with row stride LD, by LU with
// partial pivoting (largest |.| in the column, lowest row on ties) on one wavefront: lane r holds
// row r, rows are eliminated in place and never moved, and the sign of the pivot permutation
// comes from counting, at each step, the remaining rows above the pivot row.  Each lane keeps the
// current column of its row in registers; the rest of the row is updated through LDS in chunks of
// eight columns whose loads are all issued before its stores.  The matrix is overwritten.
// Returns false for a singular matrix (a zero pivot column).  Called by all 64 lanes.
__device__ bool wave_det(double* Sre, double* Sim, int LD, int Ne, double& dre, double& dim) {
    constexpr int kChunk = 8;
    const int r = __lane_id();
    bool used = r >= Ne;
    double cr = used ? 0.0 : Sre[r * LD], ci = used ? 0.0 : Sim[r * LD];  // S[r][p]
    double ar = 1.0, ai = 0.0;
    int inv = 0;
    for (int p = 0; p < Ne; ++p) {
        double mag = used ? -1.0 : cr * cr + ci * ci;
        int idx = r;
        wave_argmax(mag, idx);
        if (!(mag > 0.0)) { dre = 0.0; dim = 0.0; return false; }
        const int q = idx;
        inv += __popcll(__ballot(!used && r < q));
        const double pr = readlane_f64(cr, q), pi = readlane_f64(ci, q);
        {
            const double nr = ar * pr - ai * pi, ni = ar * pi + ai * pr;
            ar = nr;
            ai = ni;
        }
        if (r == q) used = true;
        wave_lds_sync();  // the pivot row's stores of the previous step are visible
        if (!used) {
            const double ip = 1.0 / mag;
            const double fr = (cr * pr + ci * pi) * ip, fi = (ci * pr - cr * pi) * ip;  // S[r][p] / pivot
            for (int c0 = p + 1; c0 < Ne; c0 += kChunk) {
                double yr[kChunk], yi[kChunk], xr[kChunk], xi[kChunk];
#pragma unroll
                for (int k = 0; k < kChunk; ++k) {
                    const int c = c0 + k < Ne ? c0 + k : Ne - 1;
                    yr[k] = Sre[q * LD + c];
                    yi[k] = Sim[q * LD + c];
                    xr[k] = Sre[r * LD + c];
                    xi[k] = Sim[r * LD + c];
                }
#pragma unroll
                for (int k = 0; k < kChunk; ++k) {
                    xr[k] -= fr * yr[k] - fi * yi[k];
                    xi[k] -= fr * yi[k] + fi * yr[k];
                }
                if (c0 == p + 1) { cr = xr[0]; ci = xi[0]; }
#pragma unroll
                for (int k = 0; k < kChunk; ++k) {
                    if (c0 + k < Ne) {
                        Sre[r * LD + c0 + k] = xr[k];
                        Sim[r * LD + c0 + k] = xi[k];
                    }
                }
            }
        }
        wave_lds_sync();
    }
    dre = (inv & 1) ? -ar : ar;
    dim = (inv & 1) ? -ai : ai;
    return true;
}

// RPB: ring polymer (BCBWavefunction), a separate instantiation so the one-bead kernels carry none of it
template <int NT, int CTM, bool RPB>
__global__ __launch_bounds__(64 * NT) __attribute__((amdgpu_waves_per_eu(2))) void iesh_kernel(const IeshArgs A) {
    constexpr int NP = 16 * NT;  // padded state count (rows of psi / W tiles)
    constexpr int B = 64 * NT;
    constexpr int KF = 4 * NT;   // A-fragments per lane (K = NP in steps of 4)
    extern __shared__ double smem[];
    constexpr int CW = 16 * CTM; // row stride of psi in LDS (compile time, so LDS offsets are immediates)
    const int N = A.N, M = A.M, Ne = A.Ne, CT = A.cw / 16;
    double* Tb = smem;                       // NP x CW: psi (Re, Im interleaved per electron) / Taylor term
    double* Sb = Tb + NP * CW;               // Ne x Ne complex S, inverted in place
    double* E = Sb + A.sb_len;               // NP adiabatic energies
    double* U = E + NP;                      // NP impurity components of the eigenvectors
    double* eps = U + NP;                    // NP bath levels
    double* red = eps + NP;                  // NT reduction slots
    double* pcol = red + 8;                  // max(NP, 2 kIeMaxNe) per-state / per-electron temporaries
    double* part = pcol + (NP > 2 * kIeMaxNe ? NP : 2 * kIeMaxNe);  // 8 kIeMaxNe partial sums / probabilities
    double* scal = part + 8 * kIeMaxNe;      // 8 broadcast scalars
    int* occ = reinterpret_cast<int*>(scal + 8);  // kIeMaxNe occupied states
    int* isocc = occ + kIeMaxNe;                  // NP electron index occupying the state, -1 if empty
    int* piv = isocc + NP;                        // kIeMaxNe pivot rows
    int* unocc = piv + kIeMaxNe;                  // NP unoccupied states, ascending
    // ring polymers (P > 1): bead positions, velocities, accelerations, normal-mode pair, and the
    // eigenpairs of one bead (these arrays are allocated only when P > 1)
    const int P = RPB ? A.P : 1;
    constexpr bool rp = RPB;
    double* brl = reinterpret_cast<double*>(unocc + NP);
    double* bvl = brl + 64;
    double* bal = bvl + 64;
    double* bq = bal + 64;
    double* bp = bq + 64;
    double* Eb = bp + 64;
    double* Ub = Eb + NP;

    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    const double dt = A.dt;
    const double V2 = A.V * A.V;
    for (int k = tid; k < NP; k += B) eps[k] = (k < M) ? A.a + k * A.de : 0.0;
    __syncthreads();
    const double zn = sqrt((double)M) * A.V;

    for (int64_t traj = blockIdx.x; traj < A.ntraj; traj += gridDim.x) {
        __syncthreads();
        // ---- load the trajectory --------------------------------------------------------------
        for (int idx = tid; idx < NP * CW; idx += B) {
            const int row = idx / CW, col = idx - row * CW;
            double val = 0.0;
            if (A.psi0 && row < N && col < 2 * Ne) val = A.psi0[(((size_t)traj * N + row) * Ne + (col >> 1)) * 2 + (col & 1)];
            Tb[idx] = val;
        }
        for (int s = tid; s < NP; s += B) isocc[s] = -1;
        if (tid < Ne) occ[tid] = A.occ0[(size_t)traj * Ne + tid];
        __syncthreads();
        if (tid < Ne) {
            isocc[occ[tid]] = tid;
            if (!A.psi0) Tb[occ[tid] * CW + 2 * tid] = 1.0;  // psi[state_e, e] = 1 (iesh.jl:174-180)
        }
        // The determinant bound of the hop estimate (below) needs orthonormal electron orbitals:
        // true for psi[state_e, e] = 1, checked for a given psi; propagation keeps it, EDC does not.
        bool det_bound = A.method == 0 && A.estimate && !A.edc;
        if (det_bound && A.psi0) {
            __syncthreads();
            double dev = 0.0;
            for (int k = tid; k < Ne * Ne; k += B) {
                const int e1 = k / Ne, e2 = k - e1 * Ne;
                double gr = (e1 == e2) ? -1.0 : 0.0, gi = 0.0;
                for (int a = 0; a < N; ++a) {
                    const double xr = Tb[a * CW + 2 * e1], xi = Tb[a * CW + 2 * e1 + 1];
                    const double yr = Tb[a * CW + 2 * e2], yi = Tb[a * CW + 2 * e2 + 1];
                    gr += xr * yr + xi * yi;
                    gi += xr * yi - xi * yr;
                }
                dev = fmax(dev, fmax(fabs(gr), fabs(gi)));
            }
            det_bound = block_max<NT>(dev, red) <= 1e-10;
            __syncthreads();
        }
        auto bead_mean = [&](const double* a) {
            double m = 0.0;
            for (int b = 0; b < P; ++b) m += a[b];
            return m / P;
        };
        double x, v;
        if (rp) {
            if (tid < P) {
                brl[tid] = A.r0[(size_t)traj * P + tid];
                bvl[tid] = A.v0[(size_t)traj * P + tid];
            }
            __syncthreads();
            x = bead_mean(brl);
            v = bead_mean(bvl);
        } else {
            x = A.r0[traj];
            v = A.v0[traj];
        }
        double Eprev = -1.0;
        bool warm = false;
        int nh = 0;

        // eigen at x (all threads); leaves E, U in LDS and returns the model scalars
        auto eigen = [&](double xx) -> Scalars {
            const Scalars s = na_scalars(A, xx);
            // four lanes per root, roots dealt round-robin over the wavefronts (16 NT >= N slots)
            const int ir = (l >> 2) * NT + w, sub = l & 3;
            if (ir < N) {
                double fp;
                const double lam = na_root<4>(eps, M, A.a, V2, zn, s.h, ir, warm, Eprev, fp, sub);
                if (sub == 0) {
                    E[ir] = lam;
                    U[ir] = 1.0 / sqrt(fp);
                }
                Eprev = lam;
            } else if (ir < NP && sub == 0) {
                E[ir] = 0.0;
                U[ir] = 0.0;
            }
            warm = true;
            __syncthreads();
            return s;
        };
        // eigenpairs at a bead position into Eb, Ub (cold start: no roots are kept per bead)
        auto bead_eigen = [&](double xx) -> Scalars {
            const Scalars s = na_scalars(A, xx);
            const int ir = (l >> 2) * NT + w, sub = l & 3;
            if (ir < N) {
                double fp;
                const double lam = na_root<4>(eps, M, A.a, V2, zn, s.h, ir, false, 0.0, fp, sub);
                if (sub == 0) {
                    Eb[ir] = lam;
                    Ub[ir] = 1.0 / sqrt(fp);
                }
            }
            __syncthreads();
            return s;
        };
        // acceleration at the eigenpairs whose impurity components are Uv, with the wavefunction in Tb
        // (iesh.jl:190-207, ehrenfest_na.jl:72-86; per bead rpiesh.jl:20-38, rpehrenfest_na.jl:14-40)
        auto accel = [&](const Scalars& s, const double* Uv) -> double {
            if (A.method == 0) {
                if (tid == 0) {
                    double f = s.dU0;
                    for (int e = 0; e < Ne; ++e) { const double u = Uv[occ[e]]; f += u * s.dh * u; }
                    scal[0] = f;
                }
            } else {
                if (tid < Ne) {
                    double re = 0.0, im = 0.0;
                    for (int a = 0; a < N; ++a) {
                        re += Uv[a] * Tb[a * CW + 2 * tid];
                        im += Uv[a] * Tb[a * CW + 2 * tid + 1];
                    }
                    pcol[tid] = s.dh * (re * re + im * im);
                }
                __syncthreads();
                if (tid == 0) {
                    double f = s.dU0;
                    for (int e = 0; e < Ne; ++e) f += pcol[e];
                    scal[0] = f;
                }
            }
            __syncthreads();
            const double f = scal[0];
            __syncthreads();
            return -f / A.mass;
        };
        // Diabatic populations: component a of eigenvector b is U_ab = u_b (a = 0, impurity) or
        // V u_b / (E_b - eps_{a-1}) (bath level a - 1).  IESH (iesh.jl:351-369): per electron
        // diag(U rho U^T), rho = offdiag(c c^T) + e_s e_s^T with c = Re psi_e, i.e.
        // (sum_b U_ab c_b)^2 - sum_b U_ab^2 c_b^2 + U_as^2; EhrenfestNA: sum_e |(U psi)_{a e}|^2.
        auto save = [&](int slot) {
            if (slot >= A.nsave) return;
            if (A.method == 0) {
                for (int b = tid; b < N; b += B) {
                    double wb = 0.0;
                    for (int e = 0; e < Ne; ++e) { const double c = Tb[b * CW + 2 * e]; wb += c * c; }
                    pcol[b] = wb;
                }
                __syncthreads();
            }
            // four lanes per diabatic state a (4 N <= 64 NT lanes): lane sub takes the 4-electron
            // chunks sub, sub + 4, ... and every fourth term of the IESH corrections; the quad's
            // partial sums are combined by DPP (same bits in all four lanes)
            {
                const int a = tid >> 2, sub = tid & 3;
                double pa = 0.0;
                if (a < N) {
                    const double ea = (a == 0) ? 0.0 : eps[a - 1];
                    for (int e0 = 4 * sub; e0 < Ne; e0 += 16) {
                        double sr[4], si[4];
#pragma unroll
                        for (int q = 0; q < 4; ++q) { sr[q] = 0.0; si[q] = 0.0; }
                        for (int b = 0; b < N; ++b) {
                            const double u = (a == 0) ? U[b] : A.V * U[b] / (E[b] - ea);
                            const double* row = Tb + b * CW + 2 * e0;
#pragma unroll
                            for (int q = 0; q < 4; ++q) {
                                if (e0 + q < Ne) {
                                    sr[q] += u * row[2 * q];
                                    if (A.method != 0) si[q] += u * row[2 * q + 1];
                                }
                            }
                        }
#pragma unroll
                        for (int q = 0; q < 4; ++q) pa += sr[q] * sr[q] + si[q] * si[q];
                    }
                    if (A.method == 0) {
                        for (int b = sub; b < N; b += 4) {
                            const double u = (a == 0) ? U[b] : A.V * U[b] / (E[b] - ea);
                            pa -= u * u * pcol[b];
                        }
                        for (int e = sub; e < Ne; e += 4) {
                            const int sb = occ[e];
                            const double u = (a == 0) ? U[sb] : A.V * U[sb] / (E[sb] - ea);
                            pa += u * u;
                        }
                    }
                }
                pa = group_sum<4>(pa);
                if (a < N && sub == 0) {
                    A.save_di[((size_t)slot * N + a) * A.ntraj + traj] = pa;
                    double pad;
                    if (A.method == 0) pad = isocc[a] >= 0 ? 1.0 : 0.0;  // one-hot (iesh.jl:371-375)
                    else {                                                // ehrenfest_na.jl:119-126
                        pad = 0.0;
                        for (int e = 0; e < Ne; ++e) {
                            const double re = Tb[a * CW + 2 * e], im = Tb[a * CW + 2 * e + 1];
                            pad += re * re + im * im;
                        }
                    }
                    A.save_ad[((size_t)slot * N + a) * A.ntraj + traj] = pad;
                }
            }
            __syncthreads();
        };

        // bead accelerations into bal (one bead at a time over the whole workgroup)
        auto bead_accels = [&]() {
            for (int b = 0; b < P; ++b) {
                const Scalars s = bead_eigen(brl[b]);
                const double a = accel(s, Ub);
                if (tid == 0) bal[b] = a;
            }
            __syncthreads();
        };
        // y = H x for the two bead vectors (H: the discrete Hartley transform, orthogonal and its own
        // inverse; modes l and P - l share the frequency, so the Cayley step per Hartley mode is the
        // free ring polymer's), one bead per lane of wavefront 0
        auto hartley2 = [&](const double* x1, const double* x2, double* y1, double* y2) {
            if (tid < P) {
                double a1 = 0.0, a2 = 0.0;
                for (int k = 0; k < P; ++k) {
                    double sn, cs;
                    sincospi(2.0 * (double)((tid * k) % P) / (double)P, &sn, &cs);
                    a1 += (cs + sn) * x1[k];
                    a2 += (cs + sn) * x2[k];
                }
                const double sP = 1.0 / sqrt((double)P);
                y1[tid] = a1 * sP;
                y2[tid] = a2 * sP;
            }
            __syncthreads();
        };

        Scalars sc = eigen(x);
        double kacc = 0.0;
        if (rp) bead_accels();
        else kacc = accel(sc, U);
        save(0);

        for (int64_t step = 0; step < A.nsteps; ++step) {
            if (!rp) {
                // ---- nuclei: B A (update) B (verlet_with_electronics.jl:58-66) ----------------------
                const double vt = v + (dt / 2) * kacc;
                x = x + dt * vt;
                sc = eigen(x);
                kacc = accel(sc, U);  // EhrenfestNA uses psi of the step start
                v = vt + (dt / 2) * kacc;
            } else {
                // ---- ring polymer: B, C in normal modes, B (bcb_wavefunction.jl:46-66); the bead
                // forces see psi of the step start, and the wavefunction step below the centroid
                // eigenpairs and velocity of the step start (E, U, sc, v are not updated here) ---
                if (tid < P) bvl[tid] += (dt / 2) * bal[tid];
                __syncthreads();
                hartley2(brl, bvl, bq, bp);
                if (tid < P) {  // Cayley map of mode tid, half = false (ring_polymer.jl:71-82)
                    const double wk = 2.0 * A.wn * sinpi((double)tid / (double)P);
                    const double uu = dt * dt * wk * wk / 4.0, sc1 = 1.0 / (1.0 + uu);
                    const double c11 = (1.0 - uu) * sc1, c12 = dt * sc1, c21 = -dt * wk * wk * sc1;
                    const double qn = c11 * bq[tid] + c12 * bp[tid], pn = c21 * bq[tid] + c11 * bp[tid];
                    bq[tid] = qn;
                    bp[tid] = pn;
                }
                __syncthreads();
                hartley2(bq, bp, brl, bvl);
                bead_accels();
                if (tid < P) bvl[tid] += (dt / 2) * bal[tid];
                __syncthreads();
            }

            // ---- wavefunction: psi <- exp(-i dt H) psi, H = diag(E) - i v d ---------------------------
            {
                const double sW = dt * v * sc.dh;
                const int ia = 16 * w + (l & 15);
                double Af[KF];
                double loc = 0.0;
#pragma unroll
                for (int kk = 0; kk < KF; ++kk) {
                    const int k = 4 * kk + (l >> 4);
                    double val = 0.0;
                    if (ia < N && k < N && ia != k) val = -sW * (U[ia] * U[k] / (E[k] - E[ia]));
                    Af[kk] = val;
                    loc += val * val;
                }
                // ||dt H||_2 <= dt max|E_i| + ||dt v d||_F
                const double fro = sqrt(block_sum<NT>(loc, red));
                const double theta = dt * block_max<NT>(tid < N ? fabs(E[tid]) : 0.0, red) + fro;
                // Sub-steps of norm <= 0.5, at most 1024 of them; a non-finite bound leaves psi as is.
                const bool finite = theta <= 512.0;
                const int ns = (finite && theta > 0.5) ? (int)ceil(theta / 0.5) : 1;
                const double ths = finite ? theta / ns : 0.0, eth = exp(ths);
                int K = 0;
                double term = 1.0;
                while (finite && K < 40 && term * ths / (K + 1) * eth > 1e-17) { ++K; term *= ths / K; }
                const double inv_ns = 1.0 / ns;
                if (ns > 1) {
#pragma unroll
                    for (int kk = 0; kk < KF; ++kk) Af[kk] *= inv_ns;
                }
                double dg[4];
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int ir = 16 * w + 4 * r + (l >> 4);
                    dg[r] = (ir < N) ? dt * E[ir] * inv_ns : 0.0;
                }
                d4 P[CTM];
#pragma unroll
                for (int ct = 0; ct < CTM; ++ct) {
                    if (ct < CT) {
#pragma unroll
                        for (int r = 0; r < 4; ++r) P[ct][r] = Tb[(16 * w + 4 * r + (l >> 4)) * CW + 16 * ct + (l & 15)];
                    }
                }
                // Column tiles are independent in the GEMM: each tile of the next term is computed,
                // then written back over the same tile after every wave has read it.
                const double* bcol = Tb + (l >> 4) * CW + (l & 15);
                for (int sub = 0; sub < ns; ++sub) {
                    for (int k = 1; k <= K; ++k) {
                        const double ik = 1.0 / k;
                        const bool last = (k == K);
#pragma unroll
                        for (int ct = 0; ct < CTM; ++ct) {
                            if (ct < CT) {
                                d4 Y = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
                                for (int kk = 0; kk < KF; ++kk)
                                    Y = __builtin_amdgcn_mfma_f64_16x16x4f64(Af[kk], bcol[4 * kk * CW + 16 * ct], Y, 0, 0, 0);
                                const int j = 16 * ct + (l & 15);
#pragma unroll
                                for (int r = 0; r < 4; ++r) {
                                    const double part = Tb[(16 * w + 4 * r + (l >> 4)) * CW + (j ^ 1)];
                                    Y[r] = (Y[r] + ((j & 1) ? -dg[r] : dg[r]) * part) * ik;
                                    P[ct][r] += Y[r];
                                }
                                __syncthreads();
#pragma unroll
                                for (int r = 0; r < 4; ++r) Tb[(16 * w + 4 * r + (l >> 4)) * CW + j] = last ? P[ct][r] : Y[r];
                            }
                        }
                        __syncthreads();
                    }
                    if (K == 0) break;
                }
            }

            if (rp) {  // centroid at the step end: eigenpairs and velocity for the hop and the next step
                x = bead_mean(brl);
                v = bead_mean(bvl);
                sc = eigen(x);
            }

            // the save precedes the hop and decoherence callbacks (a save time holds the state before the
            // step's callbacks: DiffEqBase saves the due saveat points before affect!)
            if (A.save_every > 0 && (step + 1) % A.save_every == 0) save((int)((step + 1) / A.save_every));

            // ---- IESH hop (iesh.jl:390-407) -------------------------------------------------------------
            if (A.method == 0) {
                const double rnd = philox_uniform((uint64_t)traj, (uint32_t)step, STREAM_HOP, A.key);
                const int LD = Ne | 1, Nu = N - Ne;
                double* Sre = Sb;
                double* Sim = Sb + Ne * LD;
                // sum |v_dot_d[m, n]| over unoccupied m and electrons n (iesh.jl:285-298), and
                // t2 = ||psi[unoccupied, :]||_F^2, the weight of the orbitals outside the occupied rows
                double lvd = 0.0, lt2 = 0.0;
                for (int k = tid; k < N * Ne; k += B) {
                    const int m = k / Ne, n = k - m * Ne;
                    if (isocc[m] < 0) {
                        const int o = occ[n];
                        lvd += fabs(v * (sc.dh * (U[m] * U[o] / (E[o] - E[m]))));
                        const double pr = Tb[m * CW + 2 * n], pi = Tb[m * CW + 2 * n + 1];
                        lt2 += pr * pr + pi * pi;
                    }
                }
                const double sumvd = block_sum<NT>(lvd, red);
                const double t2 = block_sum<NT>(lt2, red);
                // The estimate 2 dt sumvd (|Re det S| + |Im det S|) / |det S|^2 (iesh.jl:250-253) is at
                // most 2 dt sumvd sqrt(2) / |det S|, and for orthonormal orbitals (S^H S + T^H T = I,
                // T = psi[unoccupied, :]) |det S|^2 = prod(1 - s_i^2(T)) >= 1 - t2.  Where that bound
                // is already below the uniform, the reference's estimate is too: no hop, and the LU is
                // skipped (1e-6 absorbs the orbitals' rounding-level loss of orthonormality; 1e-8 and
                // the 1e-4 floor, which bounds cond(S) by 100, absorb the rounding of the estimate).
                bool skip = false;
                if (det_bound) {
                    const double L2 = 1.0 - t2 - 1e-6;
                    skip = L2 > 1e-4 && 2.0 * dt * sumvd * 1.4142135623730951 * (1.0 + 1e-8) < rnd * sqrt(L2);
                }
                bool evaluate = false;
                if (!skip) {
                    // S_je = psi[occ_j, e] (iesh.jl:309-316) as real / imaginary planes, row stride LD (odd)
                    for (int idx = tid; idx < Ne * Ne; idx += B) {
                        const int j = idx / Ne, e = idx - j * Ne;
                        Sre[j * LD + e] = Tb[occ[j] * CW + 2 * e];
                        Sim[j * LD + e] = Tb[occ[j] * CW + 2 * e + 1];
                    }
                    if (tid == 0) {  // unoccupied states in ascending order (DynamicsUtils.jl:162-171)
                        int c = 0;
                        for (int m = 0; m < N; ++m)
                            if (isocc[m] < 0) unocc[c++] = m;
                    }
                    __syncthreads();
                    // det S (FastDeterminant.jl:20-22) on wavefront 0
                    if (w == 0) {
                        double dr, di;
                        const bool ok = wave_det(Sre, Sim, LD, Ne, dr, di);
                        if (l == 0) {
                            scal[1] = dr;
                            scal[2] = di;
                            scal[3] = ok ? 0.0 : 1.0;
                        }
                    }
                    __syncthreads();
                    const double dre = scal[1], dim = scal[2];
                    const bool singular = scal[3] != 0.0;
                    const double Akk = dre * dre + dim * dim;
                    const double pref = 2.0 * dt / Akk;
                    const double est = pref * sumvd * (fabs(dre) + fabs(dim));
                    evaluate = !singular && Akk > 0.0 && (!A.estimate || !(est < rnd));
                }
                if (evaluate) {
                    __syncthreads();  // wavefront 0's determinant work on Sb is complete
                    for (int idx = tid; idx < Ne * Ne; idx += B) {
                        const int j = idx / Ne, e = idx - j * Ne;
                        Sb[2 * idx] = Tb[occ[j] * CW + 2 * e];
                        Sb[2 * idx + 1] = Tb[occ[j] * CW + 2 * e + 1];
                    }
                    // S^-1: Gauss-Jordan inversion of S (interleaved complex, row-major) in place with partial pivoting
                    bool gj_singular = false;
                    for (int p = 0; p < Ne; ++p) {
                        __syncthreads();
                        if (w == 0) {
                            double val = -1.0;
                            int idx = l;
                            if (l >= p && l < Ne) val = hypot(Sb[2 * (l * Ne + p)], Sb[2 * (l * Ne + p) + 1]);
#pragma unroll
                            for (int o = 32; o > 0; o >>= 1) {
                                const double ov = __shfl_xor(val, o);
                                const int oi = __shfl_xor(idx, o);
                                if (ov > val || (ov == val && oi < idx)) { val = ov; idx = oi; }
                            }
                            if (l == 0) {
                                piv[p] = (val > 0.0) ? idx : -1;
                                scal[1] = Sb[2 * (idx * Ne + p)];
                                scal[2] = Sb[2 * (idx * Ne + p) + 1];
                                scal[3] = Sb[2 * (p * Ne + p)];  // moves to row idx with the interchange
                                scal[4] = Sb[2 * (p * Ne + p) + 1];
                            }
                        }
                        __syncthreads();
                        const int q = piv[p];
                        if (q < 0) { gj_singular = true; break; }
                        const double pr = scal[1], pi = scal[2];
                        const double den = pr * pr + pi * pi;
                        const double ir = pr / den, ii = -pi / den;  // 1 / pivot
                        // column p after the row swap, for the elimination
                        for (int r = tid; r < Ne; r += B) {
                            if (r == p) continue;
                            pcol[2 * r] = (r == q) ? scal[3] : Sb[2 * (r * Ne + p)];
                            pcol[2 * r + 1] = (r == q) ? scal[4] : Sb[2 * (r * Ne + p) + 1];
                        }
                        // swap rows p and q, then scale the pivot row (its pivot entry becomes 1 / pivot)
                        for (int c = tid; c < Ne; c += B) {
                            const double ar = Sb[2 * (p * Ne + c)], ai = Sb[2 * (p * Ne + c) + 1];
                            double br = Sb[2 * (q * Ne + c)], bi = Sb[2 * (q * Ne + c) + 1];
                            if (q != p) { Sb[2 * (q * Ne + c)] = ar; Sb[2 * (q * Ne + c) + 1] = ai; }
                            if (c == p) { br = 1.0; bi = 0.0; }
                            Sb[2 * (p * Ne + c)] = br * ir - bi * ii;
                            Sb[2 * (p * Ne + c) + 1] = br * ii + bi * ir;
                        }
                        __syncthreads();
                        for (int idx = tid; idx < Ne * Ne; idx += B) {
                            const int r = idx / Ne, c = idx - r * Ne;
                            if (r == p) continue;
                            const double fr = pcol[2 * r], fi = pcol[2 * r + 1];
                            const double xr = (c == p) ? 0.0 : Sb[2 * idx], xi = (c == p) ? 0.0 : Sb[2 * idx + 1];
                            const double yr = Sb[2 * (p * Ne + c)], yi = Sb[2 * (p * Ne + c) + 1];
                            Sb[2 * idx] = xr - (fr * yr - fi * yi);
                            Sb[2 * idx + 1] = xi - (fr * yi + fi * yr);
                        }
                    }
                    __syncthreads();
                    if (!gj_singular) {  // undo the row interchanges as column interchanges: Sb = S^-1
                        for (int p = Ne - 1; p >= 0; --p) {
                            const int q = piv[p];
                            if (q != p) {
                                for (int r = tid; r < Ne; r += B) {
                                    const double ar = Sb[2 * (r * Ne + p)], ai = Sb[2 * (r * Ne + p) + 1];
                                    Sb[2 * (r * Ne + p)] = Sb[2 * (r * Ne + q)];
                                    Sb[2 * (r * Ne + p) + 1] = Sb[2 * (r * Ne + q) + 1];
                                    Sb[2 * (r * Ne + q)] = ar;
                                    Sb[2 * (r * Ne + q) + 1] = ai;
                                }
                            }
                            __syncthreads();
                        }
                    }
                    // P(n -> m) = 2 dt Re((psi_m S^-1)_n) v_dot_d[m, n]: det(S with row n replaced by
                    // psi_m) = det(S) (psi_m S^-1)_n, so Re(det S conj(det S_{n<-m})) / |det S|^2 is that real part.
                    // pass 1: per-electron totals, partial sums over interleaved chunks of unoccupied states
                    const int nch = max(1, min(8, B / Ne));
                    if (tid < Ne * nch) {
                        const int n = tid % Ne, ch = tid / Ne;
                        const int o = occ[n];
                        double part_sum = 0.0;
                        for (int mi = gj_singular ? Nu : ch; mi < Nu; mi += nch) {  // no hop from a singular S
                            const int m = unocc[mi];
                            double zr = 0.0;
                            for (int e = 0; e < Ne; ++e)
                                zr += Tb[m * CW + 2 * e] * Sb[2 * (e * Ne + n)] - Tb[m * CW + 2 * e + 1] * Sb[2 * (e * Ne + n) + 1];
                            const double vd = -v * (sc.dh * (U[m] * U[o] / (E[o] - E[m])));
                            part_sum += fmin(fmax(2.0 * dt * zr * vd, 0.0), 1.0);
                        }
                        part[n * 8 + ch] = part_sum;
                    }
                    __syncthreads();
                    for (int n = tid; n < Ne; n += B) {
                        double sn = 0.0;
                        for (int ch = 0; ch < nch; ++ch) sn += part[n * 8 + ch];
                        pcol[n] = sn;
                    }
                    __syncthreads();
                    if (tid == 0) {  // the electron whose cumulative total passes the uniform, if any
                        double cum = 0.0;
                        int hn = -1;
                        for (int n = 0; n < Ne; ++n) {
                            if (rnd < cum + pcol[n]) { hn = n; break; }
                            cum += pcol[n];
                        }
                        scal[3] = (double)hn;
                        scal[4] = cum;
                    }
                    __syncthreads();
                    const int hsel = (int)scal[3];
                    if (hsel >= 0) {  // pass 2: that electron's probabilities, in unoccupied order
                        const int o = occ[hsel];
                        for (int mi = tid; mi < Nu; mi += B) {
                            const int m = unocc[mi];
                            double zr = 0.0;
                            for (int e = 0; e < Ne; ++e)
                                zr += Tb[m * CW + 2 * e] * Sb[2 * (e * Ne + hsel)] - Tb[m * CW + 2 * e + 1] * Sb[2 * (e * Ne + hsel) + 1];
                            const double vd = -v * (sc.dh * (U[m] * U[o] / (E[o] - E[m])));
                            part[mi] = fmin(fmax(2.0 * dt * zr * vd, 0.0), 1.0);
                        }
                    }
                    __syncthreads();
                    if (tid == 0) {  // cumulative selection, electrons outer, unoccupied inner (iesh.jl:318-335)
                        int hn = -1, hm = -1;
                        if (hsel >= 0) {
                            double cum = scal[4];
                            for (int mi = 0; mi < Nu; ++mi) {
                                cum += part[mi];
                                if (rnd < cum) { hn = hsel; hm = unocc[mi]; break; }
                            }
                        }
                        scal[5] = v;
                        scal[6] = 0.0;
                        if (hn >= 0) {  // rescale along d[new, old] (surface_hopping.jl:65-92)
                            const int old = occ[hn];
                            bool ok = true;
                            double vn = v;
                            if (A.rescaling != NQCD_RESCALE_OFF) {
                                const double dd = sc.dh * (U[hm] * U[old] / (E[old] - E[hm]));
                                const double aq = dd * dd / A.mass / 2.0, bq = dd * v, cq = E[hm] - E[old];
                                const double disc = bq * bq - 4.0 * aq * cq;
                                if (disc < 0.0) {
                                    ok = false;
                                    if (A.rescaling == NQCD_RESCALE_VINVERSION) {
                                        const double inc = 2.0 * (v * (dd / fabs(dd))) * (dd / fabs(dd));
                                        vn = v - inc;
                                        scal[6] = -inc;
                                    }
                                } else {
                                    const double root = sqrt(disc);
                                    const double gam = (bq < 0.0) ? (bq + root) / (2.0 * aq) : (bq - root) / (2.0 * aq);
                                    vn = v - gam * dd / A.mass;
                                    scal[6] = -(gam * dd / A.mass);
                                }
                            }
                            if (ok) {
                                isocc[old] = -1;
                                isocc[hm] = hn;
                                occ[hn] = hm;
                                ++nh;
                            }
                            scal[5] = vn;
                        }
                    }
                    __syncthreads();
                    v = scal[5];
                    if (rp) {  // every bead takes the centroid's change (rpsh.jl:24-44)
                        const double inc = scal[6];
                        __syncthreads();
                        if (tid < P) bvl[tid] = bvl[tid] - (-inc);
                        __syncthreads();
                        v = bead_mean(bvl);
                    }
                    __syncthreads();
                }
                // ---- Granucci-Persico EDC per electron (decoherence_corrections.jl:21-39, iesh.jl:424-433)
                if (A.edc) {
                    const double Ekin = 0.5 * A.mass * v * v;
                    if (tid < Ne) {
                        const int s = occ[tid];
                        double nrm = 0.0;
                        for (int a = 0; a < N; ++a) {
                            if (a == s) continue;
                            const double tau = (1.0 + A.edc_C / Ekin) / fabs(E[a] - E[s]);
                            const double f = exp(-dt / tau);
                            const double re = Tb[a * CW + 2 * tid] * f, im = Tb[a * CW + 2 * tid + 1] * f;
                            Tb[a * CW + 2 * tid] = re;
                            Tb[a * CW + 2 * tid + 1] = im;
                            nrm += re * re + im * im;
                        }
                        const double cr = Tb[s * CW + 2 * tid], ci = Tb[s * CW + 2 * tid + 1];
                        const double fac = sqrt((1.0 - nrm) / (cr * cr + ci * ci));
                        Tb[s * CW + 2 * tid] = cr * fac;
                        Tb[s * CW + 2 * tid + 1] = ci * fac;
                    }
                    __syncthreads();
                }
            }
        }

        // ---- store --------------------------------------------------------------------------------
        if (rp) {
            if (tid < P) {
                A.r[(size_t)traj * P + tid] = brl[tid];
                A.v[(size_t)traj * P + tid] = bvl[tid];
            }
        } else if (tid == 0) {
            A.r[traj] = x;
            A.v[traj] = v;
        }
        if (tid == 0) {
            A.nhops[traj] = nh;
            A.steps[traj] = (int32_t)A.nsteps;
        }
        if (tid < Ne) A.occ[(size_t)traj * Ne + tid] = occ[tid];
        for (int idx = tid; idx < N * 2 * Ne; idx += B) {
            const int row = idx / (2 * Ne), col = idx - row * 2 * Ne;
            A.psi[(size_t)traj * N * 2 * Ne + idx] = Tb[row * CW + col];
        }
    }
}

// Determinants of nmat complex Ne x Ne matrices (re, im planes, row-major) by wave_det, one
// wavefront per matrix: the test hook of the IESH estimate's determinant.
__global__ __launch_bounds__(64) void wave_det_kernel(const double* __restrict__ re, const double* __restrict__ im,
                                                      int Ne, double* __restrict__ out) {
    extern __shared__ double S[];  // 2 Ne LD
    const int LD = Ne | 1;
    const size_t base = (size_t)blockIdx.x * Ne * Ne;
    for (int idx = threadIdx.x; idx < Ne * Ne; idx += 64) {
        const int j = idx / Ne, e = idx - j * Ne;
        S[j * LD + e] = re[base + idx];
        S[Ne * LD + j * LD + e] = im[base + idx];
    }
    __syncthreads();
    double dr, di;
    const bool ok = wave_det(S, S + Ne * LD, LD, Ne, dr, di);
    if (threadIdx.x == 0) {
        out[3 * blockIdx.x] = dr;
        out[3 * blockIdx.x + 1] = di;
        out[3 * blockIdx.x + 2] = ok ? 1.0 : 0.0;
    }
}

// The energies the Fermi-Dirac occupation chain weighs, at each trajectory's initial position: the
// adiabatic energies (iesh.jl:113-114), or for the diabatic basis the diagonal of the model matrix,
// [h(x0), eps_0, ..., eps_{M-1}] (iesh.jl:150-153), followed then by the adiabatic energies E_b and
// the impurity components u_b of the eigenvectors (row stride 3N; N for the adiabatic basis).  One
// wavefront per trajectory, root ir on lane ir % 64.
__global__ __launch_bounds__(256) void ie_initial_energies(const IeshArgs A, int diabatic, double* __restrict__ W) {
    __shared__ double eps[kIeMaxStates];
    for (int k = threadIdx.x; k < A.M; k += blockDim.x) eps[k] = A.a + k * A.de;
    __syncthreads();
    const int64_t traj = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (traj >= A.ntraj) return;
    double x0 = A.r0[traj];
    if (A.P > 1) {  // the centroid (the hopping eigenvalues of a ring polymer, SurfaceHoppingMethods.jl:85-87)
        x0 = 0.0;
        for (int b = 0; b < A.P; ++b) x0 += A.r0[(size_t)traj * A.P + b];
        x0 /= A.P;
    }
    const Scalars s = na_scalars(A, x0);
    const double zn = sqrt((double)A.M) * A.V;
    double* row = W + traj * (diabatic ? 3 * A.N : A.N);
    for (int ir = threadIdx.x & 63; ir < A.N; ir += 64) {
        double fp;
        const double lam = na_root(eps, A.M, A.a, A.V * A.V, zn, s.h, ir, false, 0.0, fp);
        if (diabatic) {
            row[ir] = (ir == 0) ? s.h : eps[ir - 1];
            row[A.N + ir] = lam;
            row[2 * A.N + ir] = 1.0 / sqrt(fp);
        } else {
            row[ir] = lam;
        }
    }
}

// Position of the k-th set bit (k < popcount) of a 64-bit word.
__device__ __forceinline__ int select_bit64(uint64_t w, int k) {
    int pos = 0;
#pragma unroll
    for (int sh = 32; sh > 0; sh >>= 1) {
        const int c = __popcll(w & ((1ull << sh) - 1));
        if (k >= c) { k -= c; w >>= sh; pos += sh; }
    }
    return pos;
}

// Fermi-Dirac occupations by the Boltzmann-factor Metropolis chain of
// sample_fermi_dirac_distribution (DynamicsUtils.jl:191-205), one lane per trajectory: start from
// the n_e lowest available states, then n_avail * n_e moves; move `it` draws Philox counter (traj,
// it, STREAM_OCCUPATION) = (o0, o1, o2, o3), takes the occupied slot (o0 n_e) >> 32 of the state
// array and the unoccupied available state number (o1 n_free) >> 32 in ascending order, and
// accepts when exp(-beta (E_j - E_i)) > u53(o2, o3) (beta = inf: when E_j < E_i).  The slot array
// lives in LDS ([slot][lane]); the occupied set is a 128-bit mask, so the sorted result is its set
// bits in order.
__global__ __launch_bounds__(256) void ie_fermi_dirac_chain(int64_t ntraj, int N, int Ne, uint64_t avail_lo,
                                                            uint64_t avail_hi, double beta, uint64_t key,
                                                            const double* __restrict__ E0, int stride,
                                                            int32_t* __restrict__ occ0) {
    extern __shared__ int slots[];  // Ne x blockDim.x
    const int tid = threadIdx.x, B = blockDim.x;
    const int64_t traj = (int64_t)blockIdx.x * B + tid;
    if (traj >= ntraj) return;
    const int navail = __popcll(avail_lo) + __popcll(avail_hi), nfree = navail - Ne;
    const double* E = E0 + traj * stride;
    uint64_t occ_lo = 0, occ_hi = 0;
    for (int e = 0; e < Ne; ++e) {  // the first Ne available states
        const uint64_t flo = avail_lo & ~occ_lo, fhi = avail_hi & ~occ_hi;
        const int st = flo ? __ffsll((unsigned long long)flo) - 1 : 64 + __ffsll((unsigned long long)fhi) - 1;
        slots[e * B + tid] = st;
        if (st < 64) occ_lo |= 1ull << st; else occ_hi |= 1ull << (st - 64);
    }
    const bool cold = isinf(beta);
    const int64_t nmoves = (int64_t)navail * Ne;
    for (int64_t it = 0; it < nmoves; ++it) {
        uint32_t c0 = (uint32_t)traj, c1 = (uint32_t)(traj >> 32), c2 = (uint32_t)it, c3 = STREAM_OCCUPATION;
        philox4x32_10(c0, c1, c2, c3, (uint32_t)key, (uint32_t)(key >> 32));
        const int ci = (int)(((uint64_t)c0 * (uint64_t)Ne) >> 32);
        const int kk = (int)(((uint64_t)c1 * (uint64_t)nfree) >> 32);
        const uint64_t flo = avail_lo & ~occ_lo, fhi = avail_hi & ~occ_hi;
        const int nlo = __popcll(flo);
        const int j = kk < nlo ? select_bit64(flo, kk) : 64 + select_bit64(fhi, kk - nlo);
        const int i = slots[ci * B + tid];
        const double dE = E[j] - E[i];
        const bool accept = cold ? (dE < 0.0) : (exp(-beta * dE) > u53(c2, c3));
        if (accept) {
            slots[ci * B + tid] = j;
            if (i < 64) occ_lo &= ~(1ull << i); else occ_hi &= ~(1ull << (i - 64));
            if (j < 64) occ_lo |= 1ull << j; else occ_hi |= 1ull << (j - 64);
        }
    }
    int e = 0;
    for (uint64_t w = occ_lo; w; w &= w - 1) occ0[traj * Ne + e++] = __ffsll((unsigned long long)w) - 1;
    for (uint64_t w = occ_hi; w; w &= w - 1) occ0[traj * Ne + e++] = 64 + __ffsll((unsigned long long)w) - 1;
}

// Diabatic Fermi-Dirac occupations to IESH orbitals (iesh.jl:159-182), one lane per trajectory:
// for each sampled diabatic state d (ascending), an adiabatic state is drawn with weights U_db^2
// (U_db = component d of eigenvector b: u_b for the impurity, V u_b / (E_b - eps_{d-1}) for a bath
// level) until one not yet taken comes up -- StatsBase's sample(Weights(w)): t = rand() * sum(w), the
// first b whose running sum reaches t -- and the electron's orbital is row d of U times the sign of
// U_d0, psi_b = U_db U_d0 / sqrt(U_d0^2) (real).  Draw k of a trajectory is Philox counter
// (traj, 2^31 + k, STREAM_OCCUPATION), u = u53(o0, o1).  The orbitals stay in sampling order and
// the occupations are sorted, as the reference leaves them.  A trajectory that needs more than
// 4096 draws for one electron (free states of vanishing weight) gets occupation -1.
__global__ __launch_bounds__(256) void ie_diabatic_orbitals(int64_t ntraj, int N, int Ne, double V, double a, double de,
                                                            uint64_t key, const double* __restrict__ W,
                                                            const int32_t* __restrict__ dstate,
                                                            int32_t* __restrict__ occ0, double* __restrict__ psi0) {
    const int64_t traj = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (traj >= ntraj) return;
    const double* E = W + traj * 3 * N + N;
    const double* u = E + N;
    uint64_t tk_lo = 0, tk_hi = 0;
    uint32_t draw = 0;
    bool failed = false;
    for (int i = 0; i < Ne && !failed; ++i) {
        const int d = dstate[traj * Ne + i];
        const double ed = (d == 0) ? 0.0 : a + (d - 1) * de;
        auto Ud = [&](int b) { return (d == 0) ? u[b] : V * u[b] / (E[b] - ed); };
        double sw = 0.0;
        for (int b = 0; b < N; ++b) { const double x = Ud(b); sw += x * x; }
        int pick = -1;
        for (int tries = 0; tries < 4096 && pick < 0; ++tries) {
            uint32_t c0 = (uint32_t)traj, c1 = (uint32_t)(traj >> 32), c2 = 0x80000000u + draw++, c3 = STREAM_OCCUPATION;
            philox4x32_10(c0, c1, c2, c3, (uint32_t)key, (uint32_t)(key >> 32));
            const double t = u53(c0, c1) * sw;
            int k = 0;
            double x = Ud(0), cw = x * x;
            while (cw < t && k < N - 1) {
                ++k;
                x = Ud(k);
                cw += x * x;
            }
            const bool taken = (k < 64) ? (tk_lo >> k) & 1 : (tk_hi >> (k - 64)) & 1;
            if (!taken) pick = k;
        }
        if (pick < 0) { failed = true; break; }
        if (pick < 64) tk_lo |= 1ull << pick; else tk_hi |= 1ull << (pick - 64);
        const double u0 = Ud(0), nrm = sqrt(u0 * u0);
        for (int b = 0; b < N; ++b) {
            double* p = psi0 + (((size_t)traj * N + b) * Ne + i) * 2;
            p[0] = (Ud(b) * u0) / nrm;
            p[1] = 0.0;
        }
    }
    if (failed) {
        for (int e = 0; e < Ne; ++e) occ0[traj * Ne + e] = -1;
        return;
    }
    int e = 0;
    for (uint64_t w = tk_lo; w; w &= w - 1) occ0[traj * Ne + e++] = __ffsll((unsigned long long)w) - 1;
    for (uint64_t w = tk_hi; w; w &= w - 1) occ0[traj * Ne + e++] = 64 + __ffsll((unsigned long long)w) - 1;
}

// S as two planes with an odd row stride (determinant), or interleaved Ne x Ne (inverse)
int ie_sb_len(int N, int Ne) { (void)N; return 2 * Ne * (Ne | 1); }

size_t ie_lds_bytes(int NT, int cw, int N, int Ne, int P) {
    const int NP = 16 * NT;
    return sizeof(double) * ((size_t)NP * cw + ie_sb_len(N, Ne) + 3 * NP + 8 + std::max(NP, 2 * kIeMaxNe) + 8 * kIeMaxNe + 8) +
           sizeof(int) * (2 * kIeMaxNe + 2 * NP) + (P > 1 ? sizeof(double) * (5 * 64 + 2 * (size_t)NP) : 0);
}

typedef void (*IeKernel)(const IeshArgs);

IeKernel ie_pick(int NT, bool wide, bool rp) {
#define NQCD_IE(nt)                                                                                    \
    case nt:                                                                                           \
        if (rp) return wide ? (IeKernel)iesh_kernel<nt, 8, true> : (IeKernel)iesh_kernel<nt, nt, true>; \
        return wide ? (IeKernel)iesh_kernel<nt, 8, false> : (IeKernel)iesh_kernel<nt, nt, false>;
    switch (NT) {
        NQCD_IE(1) NQCD_IE(2) NQCD_IE(3) NQCD_IE(4) NQCD_IE(5) NQCD_IE(6) NQCD_IE(7) NQCD_IE(8)
    default: return nullptr;
    }
#undef NQCD_IE
}

// Model parameters in the handle's slot 0: m, omega, g, dG, Gamma, a, b, M.
struct IeModel { double mw2, g, dG, V, a, de; int M; };

IeModel ie_model(const nqcd_handle_t* h) {
    const double* p = h->models[0].p;
    IeModel m;
    m.M = (int)p[7];
    m.mw2 = p[0] * p[1] * p[1];
    m.g = p[2];
    m.dG = p[3];
    m.a = p[5];
    m.de = (p[6] - p[5]) / (m.M - 1);
    m.V = std::sqrt(p[4] / (2.0 * M_PI)) * std::sqrt(m.de);
    return m;
}

}  // namespace

int32_t ie_check_problem(nqcd_handle_t* h) {
    const nqcd_problem_desc& d = h->prob;
    if (d.nbeads <= 1 && d.algorithm != NQCD_ALG_VERLET_ELECTRONICS)
        return fail(h, NQCD_ERR_UNSUPPORTED, "IESH / EhrenfestNA integrate with VerletwithElectronics");
    if (d.nbeads > 1 && d.algorithm != NQCD_ALG_BCB_WAVEFUNCTION)
        return fail(h, NQCD_ERR_UNSUPPORTED, "ring-polymer IESH / EhrenfestNA integrate with BCBWavefunction");
    if (d.ndofs * d.natoms != 1)
        return fail(h, NQCD_ERR_UNSUPPORTED, "IESH / EhrenfestNA: one degree of freedom");
    if (d.nbeads > 64) return fail(h, NQCD_ERR_UNSUPPORTED, "ring-polymer IESH / EhrenfestNA: at most 64 beads");
    if (d.nbeads > 1 && !(d.temperature > 0.0))
        return fail(h, NQCD_ERR_ARG, "ring polymers need a positive temperature");
    if (d.nbeads > 1 && d.decoherence != NQCD_DECOHERENCE_NONE)
        return fail(h, NQCD_ERR_UNSUPPORTED, "the EDC decoherence correction is not defined for ring polymers");
    if (d.n_states < 3 || d.n_states > kIeMaxStates)
        return fail(h, NQCD_ERR_UNSUPPORTED, "IESH / EhrenfestNA: n_states must be in 3..128");
    if (d.n_electrons < 1 || d.n_electrons >= d.n_states || d.n_electrons > kIeMaxNe)
        return fail(h, NQCD_ERR_ARG, "n_electrons must be in 1..min(64, n_states - 1)");
    if (d.precision != NQCD_PREC_FP64) return fail(h, NQCD_ERR_UNSUPPORTED, "IESH / EhrenfestNA run in FP64");
    if (d.decoherence < 0 || d.decoherence > 1) return fail(h, NQCD_ERR_ARG, "unknown decoherence correction");
    if (d.decoherence == NQCD_DECOHERENCE_EDC && d.method != NQCD_METHOD_IESH)
        return fail(h, NQCD_ERR_UNSUPPORTED, "the EDC decoherence correction applies to IESH");
    return NQCD_OK;
}

int32_t ie_set_model(nqcd_handle_t* h, int32_t slot, const double* p, int64_t n) {
    if (slot != 0) return fail(h, NQCD_ERR_UNSUPPORTED, "the Newns-Anderson model must use slot 0");
    if (n != 8) return fail(h, NQCD_ERR_ARG, "Newns-Anderson params: m, omega, g, dG, Gamma, a, b, M");
    const double M = p[7];
    if (!(M >= 2 && M == std::floor(M) && M + 1 <= kIeMaxStates))
        return fail(h, NQCD_ERR_ARG, "Newns-Anderson: M must be an integer in 2..127");
    if (!(p[4] > 0.0)) return fail(h, NQCD_ERR_ARG, "Newns-Anderson: Gamma must be positive");
    if (!(p[6] > p[5])) return fail(h, NQCD_ERR_ARG, "Newns-Anderson: the band needs b > a");
    if (!(p[0] > 0.0)) return fail(h, NQCD_ERR_ARG, "Newns-Anderson: m must be positive");
    ModelSlot m{};
    m.kind = NQCD_MODEL_NEWNS_ANDERSON;
    for (int i = 0; i < 8; ++i) m.p[i] = p[i];
    h->models[0] = m;
    h->model_set[0] = true;
    return NQCD_OK;
}

static int32_t ie_check_occ(nqcd_handle_t* h, int64_t ntraj, const int32_t* occ) {
    const int N = h->prob.n_states, Ne = h->prob.n_electrons;
    std::vector<char> seen(N);
    for (int64_t i = 0; i < ntraj; ++i) {
        std::fill(seen.begin(), seen.end(), 0);
        for (int e = 0; e < Ne; ++e) {
            const int32_t s = occ[(size_t)i * Ne + e];
            if (s < 0 || s >= N) return fail(h, NQCD_ERR_ARG, "occupied state out of range");
            if (seen[s]) return fail(h, NQCD_ERR_ARG, "an adiabatic state is occupied twice");
            seen[s] = 1;
        }
    }
    return NQCD_OK;
}

int32_t ie_load_initial(nqcd_handle_t* h, int64_t ntraj, const double* r, const double* v, const double* psi_re,
                        const double* psi_im, const int32_t* occ) {
    const int N = h->prob.n_states, Ne = h->prob.n_electrons;
    if (!occ) return fail(h, NQCD_ERR_ARG, "IESH / EhrenfestNA need the occupied states (n_electrons per trajectory)");
    if ((psi_re == nullptr) != (psi_im == nullptr)) return fail(h, NQCD_ERR_ARG, "give both psi parts or neither");
    if (int32_t rc = ie_check_occ(h, ntraj, occ)) return rc;
    h->ntraj = ntraj;
    const int P = h->prob.nbeads > 1 ? h->prob.nbeads : 1;  // (nbeads, ntraj) column-major for ring polymers
    HIPCHK(h, h->r0.ensure(sizeof(double) * P * ntraj));
    HIPCHK(h, h->v0.ensure(sizeof(double) * P * ntraj));
    HIPCHK(h, h->state0.ensure(sizeof(int32_t) * Ne * ntraj));
    HIPCHK(h, hipMemcpy(h->r0.p, r, sizeof(double) * P * ntraj, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->v0.p, v, sizeof(double) * P * ntraj, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->state0.p, occ, sizeof(int32_t) * Ne * ntraj, hipMemcpyHostToDevice));
    h->ie_psi_given = psi_re != nullptr;
    if (psi_re) {  // (N, Ne, ntraj) column-major -> per trajectory rows of (Re, Im) pairs
        std::vector<double> t((size_t)2 * N * Ne * ntraj);
        for (int64_t i = 0; i < ntraj; ++i)
            for (int a = 0; a < N; ++a)
                for (int e = 0; e < Ne; ++e) {
                    const size_t src = ((size_t)i * Ne + e) * N + a;
                    const size_t dst = (((size_t)i * N + a) * Ne + e) * 2;
                    t[dst] = psi_re[src];
                    t[dst + 1] = psi_im[src];
                }
        HIPCHK(h, h->sig0.ensure(sizeof(double) * t.size()));
        HIPCHK(h, hipMemcpy(h->sig0.p, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice));
    }
    return NQCD_OK;
}

int32_t ie_sample_initial(nqcd_handle_t* h, const nqcd_sampler_desc* S, uint64_t seed) {
    if (h->prob.nbeads > 1)
        return fail(h, NQCD_ERR_UNSUPPORTED, "ring-polymer IESH / EhrenfestNA take their beads from nqcd_mi355x_load_initial");
    if (S->kind != NQCD_SAMPLER_GAUSSIAN || S->nparams != 4 || !S->params)
        return fail(h, NQCD_ERR_ARG, "IESH / EhrenfestNA sampler: GAUSSIAN with mu_r, sd_r, mu_v, sd_v");
    const int Ne = h->prob.n_electrons;
    const int64_t n = S->ntraj;
    h->ntraj = n;
    HIPCHK(h, h->r0.ensure(sizeof(double) * n));
    HIPCHK(h, h->v0.ensure(sizeof(double) * n));
    HIPCHK(h, h->small.ensure(sizeof(double) * 4));
    HIPCHK(h, hipMemcpy(h->small.p, S->params, sizeof(double) * 4, hipMemcpyHostToDevice));
    HIPCHK(h, launch_sample_gaussian(n, 1, h->small.as<double>(), seed, h->r0.as<double>(), h->v0.as<double>(), h->stream));
    // lowest n_electrons adiabatic states occupied: the zero-temperature limit of the
    // Fermi-Dirac occupation sampling (DynamicsUtils.jl:176-208)
    std::vector<int32_t> occ((size_t)Ne * n);
    for (int64_t i = 0; i < n; ++i)
        for (int e = 0; e < Ne; ++e) occ[(size_t)i * Ne + e] = e;
    HIPCHK(h, h->state0.ensure(sizeof(int32_t) * occ.size()));
    HIPCHK(h, hipMemcpyAsync(h->state0.p, occ.data(), sizeof(int32_t) * occ.size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->ie_psi_given = false;
    return NQCD_OK;
}

int32_t ie_sample_fermi_dirac(nqcd_handle_t* h, double beta, const int32_t* avail, int32_t navail, int32_t basis,
                              uint64_t seed) {
    if (!h->have_initial || h->ntraj < 1)
        return fail(h, NQCD_ERR_STATE, "load or sample the nuclear initial conditions first");
    if (!h->model_set[0] || h->models[0].kind != NQCD_MODEL_NEWNS_ANDERSON)
        return fail(h, NQCD_ERR_STATE, "Fermi-Dirac occupations need the Newns-Anderson model in slot 0");
    if (!(beta >= 0.0)) return fail(h, NQCD_ERR_ARG, "beta must be >= 0 (inf for zero temperature)");
    if (basis != 0 && basis != 1) return fail(h, NQCD_ERR_ARG, "basis must be 0 (adiabatic) or 1 (diabatic)");
    if (basis == 1 && h->prob.method != NQCD_METHOD_IESH)
        return fail(h, NQCD_ERR_UNSUPPORTED, "diabatic Fermi-Dirac occupations are defined for IESH (iesh.jl:138)");
    const IeModel md = ie_model(h);
    const int N = h->prob.n_states, Ne = h->prob.n_electrons;
    if (md.M + 1 != N) return fail(h, NQCD_ERR_ARG, "n_states must equal M + 1 for the Newns-Anderson model");
    if (N > kIeMaxStates) return fail(h, NQCD_ERR_UNSUPPORTED, "Fermi-Dirac occupations support up to 128 states");
    uint64_t am[2] = {0, 0};
    if (!avail) {
        for (int s = 0; s < N; ++s) am[s >> 6] |= 1ull << (s & 63);
    } else {
        for (int32_t k = 0; k < navail; ++k) {
            const int s = avail[k];
            if (s < 0 || s >= N) return fail(h, NQCD_ERR_ARG, "available state out of range");
            if (am[s >> 6] >> (s & 63) & 1) return fail(h, NQCD_ERR_ARG, "available states repeat");
            am[s >> 6] |= 1ull << (s & 63);
        }
    }
    const int na = __builtin_popcountll(am[0]) + __builtin_popcountll(am[1]);
    if (na <= Ne) return fail(h, NQCD_ERR_ARG, "need more available states than electrons");
    const int64_t n = h->ntraj;
    IeshArgs A{};
    A.ntraj = n;
    A.N = N; A.M = md.M; A.Ne = Ne;
    A.mw2 = md.mw2; A.g = md.g; A.dG = md.dG; A.V = md.V; A.a = md.a; A.de = md.de;
    A.r0 = h->r0.as<double>();
    A.P = h->prob.nbeads > 1 ? h->prob.nbeads : 1;
    const int stride = basis ? 3 * N : N;
    HIPCHK(h, h->tmp.ensure(sizeof(double) * (size_t)stride * n));
    hipLaunchKernelGGL(ie_initial_energies, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, h->stream, A, basis,
                       h->tmp.as<double>());
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, h->state0.ensure(sizeof(int32_t) * (size_t)Ne * n));
    HIPCHK(h, h->state.ensure(sizeof(int32_t) * (size_t)Ne * n));
    int32_t* chain_out = basis ? h->state.as<int32_t>() : h->state0.as<int32_t>();  // diabatic states: scratch
    const int block = 256;
    hipLaunchKernelGGL(ie_fermi_dirac_chain, dim3((unsigned)((n + block - 1) / block)), dim3(block),
                       sizeof(int) * (size_t)Ne * block, h->stream, n, N, Ne, am[0], am[1], beta, seed,
                       h->tmp.as<double>(), stride, chain_out);
    HIPCHK(h, hipGetLastError());
    if (basis) {
        HIPCHK(h, h->sig0.ensure(sizeof(double) * 2 * (size_t)N * Ne * n));
        hipLaunchKernelGGL(ie_diabatic_orbitals, dim3((unsigned)((n + block - 1) / block)), dim3(block), 0, h->stream,
                           n, N, Ne, md.V, md.a, md.de, seed, h->tmp.as<double>(), h->state.as<int32_t>(),
                           h->state0.as<int32_t>(), h->sig0.as<double>());
        HIPCHK(h, hipGetLastError());
        HIPCHK(h, hipStreamSynchronize(h->stream));
        std::vector<int32_t> occ((size_t)Ne * n);
        HIPCHK(h, hipMemcpy(occ.data(), h->state0.p, sizeof(int32_t) * occ.size(), hipMemcpyDeviceToHost));
        for (int32_t o : occ)
            if (o < 0) return fail(h, NQCD_ERR_ARG, "diabatic Fermi-Dirac occupations: no free adiabatic state of nonzero weight");
        h->ie_psi_given = true;
        return NQCD_OK;
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->ie_psi_given = false;
    return NQCD_OK;
}

int32_t ie_run(nqcd_handle_t* h, double t0, int64_t nsteps, const nqcd_term_desc* term, uint64_t key) {
    (void)t0;
    if (term && term->kind != NQCD_TERM_NONE)
        return fail(h, NQCD_ERR_UNSUPPORTED, "termination predicates are not supported for IESH / EhrenfestNA");
    if (!h->model_set[0] || h->models[0].kind != NQCD_MODEL_NEWNS_ANDERSON)
        return fail(h, NQCD_ERR_STATE, "IESH / EhrenfestNA need the Newns-Anderson model in slot 0");
    const IeModel md = ie_model(h);
    const int N = h->prob.n_states, Ne = h->prob.n_electrons;
    if (md.M + 1 != N) return fail(h, NQCD_ERR_ARG, "n_states must equal M + 1 for the Newns-Anderson model");
    const int64_t n = h->ntraj;
    const int NT = (N + 15) / 16;
    const int CT = (2 * Ne + 15) / 16;
    const bool wide = CT > NT;
    IeKernel kern = ie_pick(NT, wide, h->prob.nbeads > 1);
    if (!kern) return fail(h, NQCD_ERR_UNSUPPORTED, "n_states");
    const int P = h->prob.nbeads > 1 ? h->prob.nbeads : 1;
    const size_t lds = ie_lds_bytes(NT, 16 * (wide ? 8 : NT), N, Ne, P);
    if (lds > kIeLdsBytes)
        return fail(h, NQCD_ERR_UNSUPPORTED,
                    "IESH / EhrenfestNA with " + std::to_string(N) + " states and " + std::to_string(Ne) +
                        " electrons needs " + std::to_string(lds) + " bytes of LDS per workgroup (" +
                        std::to_string(kIeLdsBytes) + " available)");
    HIPCHK(h, hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const int block = 64 * NT;
    int bpc = 0;
    HIPCHK(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, (const void*)kern, block, lds));
    if (bpc < 1) return fail(h, NQCD_ERR_UNSUPPORTED, "IESH kernel does not fit on a compute unit");
    int64_t grid = (int64_t)bpc * h->cu_count;
    if (grid > n) grid = n;

    HIPCHK(h, h->r.ensure(sizeof(double) * P * n));
    HIPCHK(h, h->v.ensure(sizeof(double) * P * n));
    HIPCHK(h, h->sig.ensure(sizeof(double) * 2 * N * Ne * n));
    HIPCHK(h, h->state.ensure(sizeof(int32_t) * Ne * n));
    HIPCHK(h, h->steps.ensure(sizeof(int32_t) * n));
    HIPCHK(h, h->nhops.ensure(sizeof(int32_t) * n));
    HIPCHK(h, h->save_ad.ensure(sizeof(double) * N * h->nsave * n));
    HIPCHK(h, h->save_di.ensure(sizeof(double) * N * h->nsave * n));
    HIPCHK(h, launch_zero(h->save_ad.p, sizeof(double) * N * h->nsave * n, h->stream));
    HIPCHK(h, launch_zero(h->save_di.p, sizeof(double) * N * h->nsave * n, h->stream));

    IeshArgs A{};
    A.ntraj = n;
    A.nsteps = nsteps;
    A.N = N; A.M = md.M; A.Ne = Ne; A.cw = 16 * CT;
    A.sb_len = ie_sb_len(N, Ne);
    A.method = (h->prob.method == NQCD_METHOD_IESH) ? 0 : 1;
    A.rescaling = h->prob.rescaling;
    A.estimate = h->prob.disable_estimate ? 0 : 1;
    A.edc = h->prob.decoherence == NQCD_DECOHERENCE_EDC;
    A.save_every = h->save_every;
    A.nsave = h->nsave;
    A.dt = h->prob.dt;
    A.P = P;
    A.wn = P * h->prob.temperature;  // omega_n = P kT (ring_polymer.jl:6-39)
    A.mass = h->masses[0];
    A.edc_C = h->prob.edc_C > 0.0 ? h->prob.edc_C : 0.1;
    A.key = key;
    A.mw2 = md.mw2; A.g = md.g; A.dG = md.dG; A.V = md.V; A.a = md.a; A.de = md.de;
    A.r0 = h->r0.as<double>(); A.v0 = h->v0.as<double>(); A.occ0 = h->state0.as<int32_t>();
    A.psi0 = h->ie_psi_given ? h->sig0.as<double>() : nullptr;
    A.r = h->r.as<double>(); A.v = h->v.as<double>(); A.psi = h->sig.as<double>(); A.occ = h->state.as<int32_t>();
    A.nhops = h->nhops.as<int32_t>(); A.steps = h->steps.as<int32_t>();
    A.save_di = h->save_di.as<double>(); A.save_ad = h->save_ad.as<double>();
    HIPCHK(h, hipEventRecord(h->ev0, h->stream));
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(block), lds, h->stream, A);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipEventRecord(h->ev1, h->stream));
    h->stats.block = block;
    h->stats.grid = (int)grid;
    h->stats.launches = 1;
    return NQCD_OK;
}

int32_t ie_download(nqcd_handle_t* h, int32_t field, void* out, int64_t bytes) {
    const int64_t nt = h->ntraj;
    const int N = h->prob.n_states, Ne = h->prob.n_electrons, ns = h->nsave;
    switch (field) {
    case NQCD_FIELD_R:
    case NQCD_FIELD_V:  // (nbeads, ntraj) for ring polymers
        if (bytes != (int64_t)sizeof(double) * nt * (h->prob.nbeads > 1 ? h->prob.nbeads : 1))
            return fail(h, NQCD_ERR_ARG, "bytes mismatch");
        HIPCHK(h, hipMemcpy(out, field == NQCD_FIELD_R ? h->r.p : h->v.p, bytes, hipMemcpyDeviceToHost));
        return NQCD_OK;
    case NQCD_FIELD_SIGMA_RE:
    case NQCD_FIELD_SIGMA_IM: {  // (N, Ne, ntraj) column-major
        if (bytes != (int64_t)sizeof(double) * N * Ne * nt) return fail(h, NQCD_ERR_ARG, "bytes mismatch");
        std::vector<double> t((size_t)2 * N * Ne * nt);
        HIPCHK(h, hipMemcpy(t.data(), h->sig.p, sizeof(double) * t.size(), hipMemcpyDeviceToHost));
        double* o = static_cast<double*>(out);
        const int c = field == NQCD_FIELD_SIGMA_RE ? 0 : 1;
        for (int64_t i = 0; i < nt; ++i)
            for (int a = 0; a < N; ++a)
                for (int e = 0; e < Ne; ++e) o[((size_t)i * Ne + e) * N + a] = t[(((size_t)i * N + a) * Ne + e) * 2 + c];
        return NQCD_OK;
    }
    case NQCD_FIELD_STATE:
        if (bytes != (int64_t)sizeof(int32_t) * Ne * nt) return fail(h, NQCD_ERR_ARG, "bytes mismatch");
        HIPCHK(h, hipMemcpy(out, h->state.p, bytes, hipMemcpyDeviceToHost));
        return NQCD_OK;
    case NQCD_FIELD_STEPS:
    case NQCD_FIELD_NHOPS:
        if (bytes != (int64_t)sizeof(int32_t) * nt) return fail(h, NQCD_ERR_ARG, "bytes mismatch");
        HIPCHK(h, hipMemcpy(out, field == NQCD_FIELD_STEPS ? h->steps.p : h->nhops.p, bytes, hipMemcpyDeviceToHost));
        return NQCD_OK;
    case NQCD_FIELD_SAVE_POP_AD:
    case NQCD_FIELD_SAVE_POP_DI: {  // (N, nsave, ntraj)
        if (bytes != (int64_t)sizeof(double) * N * ns * nt) return fail(h, NQCD_ERR_ARG, "bytes mismatch");
        std::vector<double> t((size_t)N * ns * nt);
        HIPCHK(h, hipMemcpy(t.data(), field == NQCD_FIELD_SAVE_POP_AD ? h->save_ad.p : h->save_di.p,
                            sizeof(double) * t.size(), hipMemcpyDeviceToHost));
        double* o = static_cast<double*>(out);
        for (int64_t i = 0; i < nt; ++i)
            for (int sl = 0; sl < ns; ++sl)
                for (int a = 0; a < N; ++a) o[((size_t)i * ns + sl) * N + a] = t[((size_t)sl * N + a) * nt + i];
        return NQCD_OK;
    }
    default: return fail(h, NQCD_ERR_UNSUPPORTED, "field not available for IESH / EhrenfestNA");
    }
}

int32_t ie_test_det(nqcd_handle_t* h, const double* re, const double* im, int32_t Ne, int32_t nmat, double* out) {
    if (Ne < 1 || Ne > kIeMaxNe) return fail(h, NQCD_ERR_ARG, "1 <= n <= 64 required");
    if (nmat < 1 || nmat > 65536) return fail(h, NQCD_ERR_ARG, "1 <= nmat <= 65536 required");
    HIPCHK(h, hipSetDevice(h->device));
    const size_t n = (size_t)nmat * Ne * Ne;
    DevBuf a, b, o;
    HIPCHK(h, a.ensure(sizeof(double) * n));
    HIPCHK(h, b.ensure(sizeof(double) * n));
    HIPCHK(h, o.ensure(sizeof(double) * 3 * nmat));
    hipError_t e = hipMemcpy(a.p, re, sizeof(double) * n, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b.p, im, sizeof(double) * n, hipMemcpyHostToDevice);
    const size_t lds = sizeof(double) * 2 * Ne * (Ne | 1);
    if (e == hipSuccess) e = hipFuncSetAttribute((const void*)wave_det_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(wave_det_kernel, dim3(nmat), dim3(64), lds, h->stream, a.as<double>(), b.as<double>(), Ne, o.as<double>());
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(out, o.p, sizeof(double) * 3 * nmat, hipMemcpyDeviceToHost);
    a.release(); b.release(); o.release();
    if (e != hipSuccess) return fail(h, NQCD_ERR_DEVICE, std::string("determinant test: ") + hipGetErrorString(e));
    return NQCD_OK;
}

}  // namespace nqcd
