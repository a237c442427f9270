// fssh_n_kern.h -- device templates of the n-state FSSH / Ehrenfest kernel (fssh_n.hip describes the
// method and the layout); included by the two translation units that instantiate the state counts.
#pragma once
#include "fssh_n.h"

namespace nqcd {

template <class T>
__device__ __forceinline__ const T* ns_opaque(const T* p) {
    asm volatile("" : "+s"(p));
    return p;
}

// index of the upper pair (i, j), i < j, of an NM x NM matrix
template <int NM>
__host__ __device__ constexpr int pidx(int i, int j) {
    return i * NM - i * (i + 1) / 2 + (j - i - 1);
}

template <int NM, typename R>
struct LaneN {
    static constexpr int NP = NM * (NM - 1) / 2;
    R r, v, k;
    double a[NM], x[NP > 0 ? NP : 1], y[NP > 0 ? NP : 1];  // sigma: diagonal, upper real / imaginary parts
    double U[NM * NM];                                     // eigenvectors, row-major, column j = state j
    double Ec[NM], Wc[NP > 0 ? NP : 1];                    // electronic buffer at the step start
    double tc, t;
    const double* tab;                                     // MORSE_N table of the trajectory's model, or null
    ModelRegs m;
    int64_t tid;
    int32_t n, st, nh, step;
};

// Model matrix V (diagonal vd, upper vo) and derivative (dd, dof) of the trajectory's model at x.
template <int NM, typename R>
__device__ __forceinline__ void ns_model(const LaneN<NM, R>& L, R x, double vd[NM], double vo[], double dd[NM],
                                         double dof[]) {
    constexpr int NP = LaneN<NM, R>::NP;
#pragma unroll
    for (int i = 0; i < NM; ++i) { vd[i] = 0.0; dd[i] = 0.0; }
#pragma unroll
    for (int p = 0; p < NP; ++p) { vo[p] = 0.0; dof[p] = 0.0; }
    if (!L.tab) {  // two-state closed-form models (Tully 1/2/3, double well)
        R v11, v22, v12, d11, d22, d12;
        model1d_eval<R>(L.m, x, v11, v22, v12, d11, d22, d12);
        vd[0] = (double)v11; dd[0] = (double)d11;
        if (NM > 1) { vd[1] = (double)v22; dd[1] = (double)d22; vo[0] = (double)v12; dof[0] = (double)d12; }
        return;
    }
    // n-state Morse (Coker and Xiao, J. Chem. Phys. 102, 496 (1995)):
    //   V_ii = D_i (1 - e^{-beta_i (x - r_i)})^2 + c_i,   V_ij = A_ij e^{-a_ij (x - r_ij)^2}
    const double xd = (double)x;
    const double* q = L.tab;
    const int n = L.n;
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        if (i < n) {
            const double Dm = q[4 * i], be = q[4 * i + 1], r0 = q[4 * i + 2], c = q[4 * i + 3];
            const double e = exp(-be * (xd - r0));
            vd[i] = Dm * (1.0 - e) * (1.0 - e) + c;
            dd[i] = 2.0 * Dm * be * e * (1.0 - e);
        }
    }
    // pairs of the n-state table are numbered for n, those of the registers for NM
    int tp = 0;
#pragma unroll
    for (int i = 0; i < NM; ++i)
#pragma unroll
        for (int j = i + 1; j < NM; ++j) {
            if (j < n) {
                const double* c = q + 4 * n + 3 * tp;
                const double u = xd - c[2];
                const double w = c[0] * exp(-c[1] * u * u);
                vo[pidx<NM>(i, j)] = w;
                dof[pidx<NM>(i, j)] = -2.0 * c[1] * u * w;
                ++tp;
            }
        }
}

// Cyclic Jacobi on the leading n x n block of the symmetric matrix (vd, vo): eigenvalues ascending in
// E, eigenvectors in U (columns), signs by the fresh rule or tracked against Uprev (= U on entry).
// warm: the sweeps start from the basis U holds on entry (an orthogonal matrix that is the identity outside
// the leading block, such as the previous step's eigenvectors): A = U^T V U and Q = U, so a basis close to
// the new one needs one or two sweeps to meet the convergence test instead of four or five.  From the
// identity the warm start is the cold one bit for bit (products with exact ones and zeros).
template <int NM>
__device__ __forceinline__ void ns_eigen(int n, const double vd[NM], const double vo[], double E[NM], double U[NM * NM],
                                         bool track, bool warm = false) {
    constexpr int NP = NM * (NM - 1) / 2;
    double A[NM * NM], Q[NM * NM];
#pragma unroll
    for (int i = 0; i < NM; ++i)
#pragma unroll
        for (int j = 0; j < NM; ++j) {
            A[i * NM + j] = (i == j) ? vd[i] : (i < j ? vo[pidx<NM>(i, j)] : vo[pidx<NM>(j, i)]);
            Q[i * NM + j] = warm ? U[i * NM + j] : ((i == j) ? 1.0 : 0.0);
        }
    if (warm) {
        double T[NM * NM];
#pragma unroll
        for (int i = 0; i < NM; ++i)  // T = V U
#pragma unroll
            for (int j = 0; j < NM; ++j) {
                double t = 0.0;
#pragma unroll
                for (int k = 0; k < NM; ++k) t += A[i * NM + k] * U[k * NM + j];
                T[i * NM + j] = t;
            }
#pragma unroll
        for (int i = 0; i < NM; ++i)  // A = U^T T, the upper triangle mirrored
#pragma unroll
            for (int j = i; j < NM; ++j) {
                double t = 0.0;
#pragma unroll
                for (int k = 0; k < NM; ++k) t += U[k * NM + i] * T[k * NM + j];
                A[i * NM + j] = t;
                A[j * NM + i] = t;
            }
    }
    (void)NP;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
#pragma unroll
        for (int i = 0; i < NM; ++i)
#pragma unroll
            for (int j = 0; j < NM; ++j) {
                const double a2 = (i < n && j < n) ? A[i * NM + j] * A[i * NM + j] : 0.0;
                tot += a2;
                if (i != j) off += a2;
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
#pragma unroll
        for (int p = 0; p < NM - 1; ++p)
#pragma unroll
            for (int q = p + 1; q < NM; ++q) {
                const double apq = A[p * NM + q];
                if (q < n && apq != 0.0) {
                    const double theta = (A[q * NM + q] - A[p * NM + p]) / (2.0 * apq);
                    const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                    const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
                    for (int k = 0; k < NM; ++k) {  // A <- A J (columns p, q)
                        const double akp = A[k * NM + p], akq = A[k * NM + q];
                        A[k * NM + p] = c * akp - s * akq;
                        A[k * NM + q] = s * akp + c * akq;
                    }
#pragma unroll
                    for (int k = 0; k < NM; ++k) {  // A <- J^T A (rows p, q)
                        const double apk = A[p * NM + k], aqk = A[q * NM + k];
                        A[p * NM + k] = c * apk - s * aqk;
                        A[q * NM + k] = s * apk + c * aqk;
                    }
#pragma unroll
                    for (int k = 0; k < NM; ++k) {
                        const double ukp = Q[k * NM + p], ukq = Q[k * NM + q];
                        Q[k * NM + p] = c * ukp - s * ukq;
                        Q[k * NM + q] = s * ukp + c * ukq;
                    }
                }
            }
    }
    if (warm) {
        // one Newton-Schulz step Q <- Q (3 I - Q^T Q) / 2: a basis carried from step to step would otherwise
        // lose orthogonality by rounding, step after step
        double G[NM * NM], Qn[NM * NM];
#pragma unroll
        for (int i = 0; i < NM; ++i)
#pragma unroll
            for (int j = i; j < NM; ++j) {
                double g = 0.0;
#pragma unroll
                for (int k = 0; k < NM; ++k) g += Q[k * NM + i] * Q[k * NM + j];
                G[i * NM + j] = ((i == j) ? 1.5 : 0.0) - 0.5 * g;
                G[j * NM + i] = G[i * NM + j];
            }
#pragma unroll
        for (int i = 0; i < NM; ++i)
#pragma unroll
            for (int j = 0; j < NM; ++j) {
                double t = 0.0;
#pragma unroll
                for (int k = 0; k < NM; ++k) t += Q[i * NM + k] * G[k * NM + j];
                Qn[i * NM + j] = t;
            }
#pragma unroll
        for (int i = 0; i < NM * NM; ++i) Q[i] = Qn[i];
    }
#pragma unroll
    for (int i = 0; i < NM; ++i) E[i] = A[i * NM + i];
    // ascending order over the first n (a stable exchange sort on fixed indices)
#pragma unroll
    for (int pass = 0; pass < NM - 1; ++pass)
#pragma unroll
        for (int j = 0; j < NM - 1 - pass; ++j) {
            if (j + 1 < n && E[j + 1] < E[j]) {
                const double te = E[j]; E[j] = E[j + 1]; E[j + 1] = te;
#pragma unroll
                for (int k = 0; k < NM; ++k) {
                    const double tq = Q[k * NM + j]; Q[k * NM + j] = Q[k * NM + j + 1]; Q[k * NM + j + 1] = tq;
                }
            }
        }
#pragma unroll
    for (int j = 0; j < NM; ++j) {
        double flip;
        if (track) {
            double s = 0.0;
#pragma unroll
            for (int i = 0; i < NM; ++i) s += Q[i * NM + j] * U[i * NM + j];
            flip = (s < 0.0) ? -1.0 : 1.0;
        } else {
            double amax = -1.0, vmax = 0.0;
#pragma unroll
            for (int i = 0; i < NM; ++i) {
                const double av = fabs(Q[i * NM + j]);
                if (i < n && av > amax) { amax = av; vmax = Q[i * NM + j]; }
            }
            flip = (vmax < 0.0) ? -1.0 : 1.0;
        }
#pragma unroll
        for (int i = 0; i < NM; ++i) U[i * NM + j] = flip * Q[i * NM + j];
    }
#pragma unroll
    for (int i = 0; i < NM; ++i) if (i >= n) E[i] = 0.0;
}

// G = U^T dV U (upper triangle with diagonal): Gd[i] = G_ii, Go[p(i,j)] = G_ij.
template <int NM>
__device__ __forceinline__ void ns_rotate(const double U[NM * NM], const double dd[NM], const double dof[], double Gd[NM],
                                          double Go[]) {
    double T[NM * NM];  // T = dV U
#pragma unroll
    for (int a = 0; a < NM; ++a)
#pragma unroll
        for (int j = 0; j < NM; ++j) {
            double s = 0.0;
#pragma unroll
            for (int b = 0; b < NM; ++b) {
                const double dab = (a == b) ? dd[a] : (a < b ? dof[pidx<NM>(a, b)] : dof[pidx<NM>(b, a)]);
                s += dab * U[b * NM + j];
            }
            T[a * NM + j] = s;
        }
#pragma unroll
    for (int i = 0; i < NM; ++i)
#pragma unroll
        for (int j = i; j < NM; ++j) {
            double s = 0.0;
#pragma unroll
            for (int a = 0; a < NM; ++a) s += U[a * NM + i] * T[a * NM + j];
            if (i == j) Gd[i] = s; else Go[pidx<NM>(i, j)] = s;
        }
}

// Right-hand side of the density equation, K = f(a, x, y) at energies e and couplings w (upper).
template <int NM>
__device__ __forceinline__ void ns_rhs(const double e[NM], const double w[], const double a[NM], const double x[],
                                       const double y[], double ka[NM], double kx[], double ky[]) {
    auto W = [&](int i, int k) -> double { return i == k ? 0.0 : (i < k ? w[pidx<NM>(i, k)] : -w[pidx<NM>(k, i)]); };
    auto X = [&](int k, int j) -> double { return k == j ? a[k] : (k < j ? x[pidx<NM>(k, j)] : x[pidx<NM>(j, k)]); };
    auto Y = [&](int k, int j) -> double { return k == j ? 0.0 : (k < j ? y[pidx<NM>(k, j)] : -y[pidx<NM>(j, k)]); };
    double M[NM * NM], N[NM * NM];
#pragma unroll
    for (int i = 0; i < NM; ++i)
#pragma unroll
        for (int j = 0; j < NM; ++j) {
            double sm = 0.0, sn = 0.0;
#pragma unroll
            for (int k = 0; k < NM; ++k) {
                if (k != i) { sm += W(i, k) * X(k, j); sn += W(i, k) * Y(k, j); }
            }
            M[i * NM + j] = sm;
            N[i * NM + j] = sn;
        }
#pragma unroll
    for (int i = 0; i < NM; ++i) ka[i] = -2.0 * M[i * NM + i];
#pragma unroll
    for (int i = 0; i < NM; ++i)
#pragma unroll
        for (int j = i + 1; j < NM; ++j) {
            const int p = pidx<NM>(i, j);
            const double de = e[i] - e[j];
            kx[p] = de * y[p] - (M[i * NM + j] + M[j * NM + i]);
            ky[p] = -de * x[p] - (N[i * NM + j] - N[j * NM + i]);
        }
}

// One Tsit5 step of size h from tau (electronic_dynamics.jl:104-130 integrated by Tsit5, the stage
// arithmetic of the oracle: u + h (sum a_ij k_j)); the parameters at stage time tau + c h are the
// linear interpolation e = Ec + (En - Ec) loc, loc = (tau - tc) / (tn - tc).
template <int NM>
struct NsInterp {
    static constexpr int NP = NM * (NM - 1) / 2 > 0 ? NM * (NM - 1) / 2 : 1;
    double Ec[NM], En[NM], Wc[NP], Wn[NP];
    double tc, tn;
};

template <int NM>
__device__ __forceinline__ void ns_params(const NsInterp<NM>& I, double tau, double e[NM], double w[]) {
    constexpr int NP = NM * (NM - 1) / 2;
    double loc = (tau - I.tc) / (I.tn - I.tc);
    if (loc != loc) loc = 0.0;
#pragma unroll
    for (int i = 0; i < NM; ++i) e[i] = I.Ec[i] + (I.En[i] - I.Ec[i]) * loc;
#pragma unroll
    for (int p = 0; p < NP; ++p) w[p] = I.Wc[p] + (I.Wn[p] - I.Wc[p]) * loc;
}

template <int NM>
__device__ __forceinline__ void ns_tsit5(const NsInterp<NM>& I, double tau, double h, double a[NM], double x[], double y[]) {
    using namespace ts5;
    constexpr int NP = NM * (NM - 1) / 2;
    constexpr int NQ = NP > 0 ? NP : 1;
    double e[NM], w[NQ];
    double ka[6][NM], kx[6][NQ], ky[6][NQ];
    double ta[NM], tx[NQ], ty[NQ];
    const double cs[6] = {0.0, c1, c2, c3, c4, 1.0};
    const double A[6][5] = {{0, 0, 0, 0, 0},
                            {a21, 0, 0, 0, 0},
                            {a31, a32, 0, 0, 0},
                            {a41, a42, a43, 0, 0},
                            {a51, a52, a53, a54, 0},
                            {a61, a62, a63, a64, a65}};
#pragma unroll
    for (int st = 0; st < 6; ++st) {
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < st; ++j) s = (j == 0) ? A[st][0] * ka[0][i] : s + A[st][j] * ka[j][i];
            ta[i] = (st == 0) ? a[i] : a[i] + h * s;
        }
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            double sx = 0.0, sy = 0.0;
#pragma unroll
            for (int j = 0; j < st; ++j) {
                sx = (j == 0) ? A[st][0] * kx[0][p] : sx + A[st][j] * kx[j][p];
                sy = (j == 0) ? A[st][0] * ky[0][p] : sy + A[st][j] * ky[j][p];
            }
            tx[p] = (st == 0) ? x[p] : x[p] + h * sx;
            ty[p] = (st == 0) ? y[p] : y[p] + h * sy;
        }
        ns_params<NM>(I, tau + cs[st] * h, e, w);
        ns_rhs<NM>(e, w, ta, tx, ty, ka[st], kx[st], ky[st]);
    }
    const double b[6] = {a71, a72, a73, a74, a75, a76};
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        double s = b[0] * ka[0][i];
#pragma unroll
        for (int j = 1; j < 6; ++j) s += b[j] * ka[j][i];
        a[i] = a[i] + h * s;
    }
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        double sx = b[0] * kx[0][p], sy = b[0] * ky[0][p];
#pragma unroll
        for (int j = 1; j < 6; ++j) { sx += b[j] * kx[j][p]; sy += b[j] * ky[j][p]; }
        x[p] = x[p] + h * sx;
        y[p] = y[p] + h * sy;
    }
}

template <int NM, typename R>
__device__ __forceinline__ void ns_accel(const FsshNArgs& A, LaneN<NM, R>& L, const double Gd[NM], const double Go[],
                                         const double* aa, const double* xx) {
    constexpr int NP = LaneN<NM, R>::NP;
    double f;
    if (A.method == 0) {
        f = 0.0;
#pragma unroll
        for (int i = 0; i < NM; ++i) f = (i == L.st) ? Gd[i] : f;
    } else {  // sum_ij Re sigma_ij G_ij
        f = 0.0;
#pragma unroll
        for (int i = 0; i < NM; ++i) f += Gd[i] * aa[i];
#pragma unroll
        for (int p = 0; p < NP; ++p) f += 2.0 * Go[p] * xx[p];
    }
    L.k = (R)(-f * A.inv_mass);
}

template <int NM, typename R>
__device__ __forceinline__ void ns_save(const FsshNArgs* Ap, const LaneN<NM, R>& L, int slot) {
    const FsshNArgs& A = *ns_opaque(Ap);
    if (slot >= A.nsave) return;
    const int64_t nt = A.ntraj;
    double pa[NM];
#pragma unroll
    for (int i = 0; i < NM; ++i) pa[i] = (A.method == 0) ? ((i == L.st) ? 1.0 : 0.0) : L.a[i];
    if (A.save_pop_ad)
#pragma unroll
        for (int i = 0; i < NM; ++i) A.save_pop_ad[((int64_t)slot * NM + i) * nt + L.tid] = pa[i];
    if (A.save_pop_di) {
        // diag(U S U^T), S = Re sigma, with the diagonal replaced by the active state for FSSH
        // (fssh.jl:132-142, ehrenfest.jl:75-83)
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            double p = 0.0;
#pragma unroll
            for (int q = 0; q < NM; ++q) p += L.U[i * NM + q] * L.U[i * NM + q] * pa[q];
#pragma unroll
            for (int q = 0; q < NM; ++q)
#pragma unroll
                for (int r = q + 1; r < NM; ++r) p += 2.0 * L.U[i * NM + q] * L.U[i * NM + r] * L.x[pidx<NM>(q, r)];
            A.save_pop_di[((int64_t)slot * NM + i) * nt + L.tid] = p;
        }
    }
    if (A.save_state) A.save_state[(int64_t)slot * nt + L.tid] = L.st;
}

template <int NM, typename R>
__device__ __forceinline__ void ns_load(const FsshNArgs* Ap, LaneN<NM, R>& L, int64_t tid) {
    const FsshNArgs& A = *ns_opaque(Ap);
    constexpr int NP = LaneN<NM, R>::NP;
    const int64_t nt = A.ntraj;
    L.tid = tid;
    const int mid = A.model_id ? A.model_id[tid] : 0;
    L.n = A.slot_n[mid];
    L.tab = A.slot_tab[mid] >= 0 ? A.mtab + A.slot_tab[mid] : nullptr;
    L.m = model_regs(A.models, mid);
    L.r = (R)A.r0[tid];
    L.v = (R)A.v0[tid];
    double vd[NM], vo[NP > 0 ? NP : 1], dd[NM], dof[NP > 0 ? NP : 1], E[NM], Gd[NM], Go[NP > 0 ? NP : 1];
    ns_model<NM, R>(L, L.r, vd, vo, dd, dof);
    ns_eigen<NM>(L.n, vd, vo, E, L.U, false);                     // initialize! (bab_electronics.jl:48-59)
    ns_rotate<NM>(L.U, dd, dof, Gd, Go);
    if (A.init_diabatic >= 0) {  // sigma = U^T e_d e_d^T U (density_matrix_dynamics.jl:37-46,64-75)
        const int d = A.init_diabatic;
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            double ud = 0.0;
#pragma unroll
            for (int q = 0; q < NM; ++q) ud = (q == d) ? L.U[q * NM + i] : ud;
            L.a[i] = ud * ud;
        }
#pragma unroll
        for (int i = 0; i < NM; ++i)
#pragma unroll
            for (int j = i + 1; j < NM; ++j) {
                double ui = 0.0, uj = 0.0;
#pragma unroll
                for (int q = 0; q < NM; ++q) { ui = (q == d) ? L.U[q * NM + i] : ui; uj = (q == d) ? L.U[q * NM + j] : uj; }
                L.x[pidx<NM>(i, j)] = ui * uj;
                L.y[pidx<NM>(i, j)] = 0.0;
            }
    } else {
#pragma unroll
        for (int i = 0; i < NM; ++i) L.a[i] = A.sig0_re[(int64_t)(i + NM * i) * nt + tid];
#pragma unroll
        for (int i = 0; i < NM; ++i)
#pragma unroll
            for (int j = i + 1; j < NM; ++j) {  // sigma_ij = element (i, j) of the column-major (NM, NM) matrix
                L.x[pidx<NM>(i, j)] = A.sig0_re[(int64_t)(i + NM * j) * nt + tid];
                L.y[pidx<NM>(i, j)] = A.sig0_im[(int64_t)(i + NM * j) * nt + tid];
            }
    }
    int st = A.state0 ? A.state0[tid] : -1;
    if (A.method == 0 && st < 0) {  // fssh.jl:54: sample(Weights(diag(real.(sigma))))
        double tot = 0.0;
#pragma unroll
        for (int i = 0; i < NM; ++i) tot += L.a[i];
        const double u = philox_uniform((uint64_t)tid, 0u, STREAM_STATE, A.key) * tot;
        double cum = 0.0;
        st = L.n - 1;
        bool found = false;
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            cum += L.a[i];
            if (!found && i < L.n && u < cum) { st = i; found = true; }
        }
    }
    L.st = st;
    ns_accel<NM, R>(A, L, Gd, Go, L.a, L.x);
#pragma unroll
    for (int i = 0; i < NM; ++i) L.Ec[i] = 0.0;
#pragma unroll
    for (int p = 0; p < NP; ++p) L.Wc[p] = 0.0;
    L.tc = 0.0;  // zero buffers stamped t = 0 (electronic_dynamics.jl:118-127)
    L.t = A.t0;
    L.nh = 0;
    L.step = 0;
    ns_save<NM, R>(Ap, L, 0);
}

template <int NM, typename R>
__device__ __forceinline__ void ns_finish(const FsshNArgs* Ap, const LaneN<NM, R>& L) {
    const FsshNArgs& A = *ns_opaque(Ap);
    const int64_t nt = A.ntraj;
    A.r[L.tid] = (double)L.r;
    A.v[L.tid] = (double)L.v;
#pragma unroll
    for (int i = 0; i < NM; ++i)
#pragma unroll
        for (int j = 0; j < NM; ++j) {
            double re, im;
            if (i == j) { re = L.a[i]; im = 0.0; }
            else if (i < j) { re = L.x[pidx<NM>(i, j)]; im = L.y[pidx<NM>(i, j)]; }
            else { re = L.x[pidx<NM>(j, i)]; im = -L.y[pidx<NM>(j, i)]; }
            A.sig_re[(int64_t)(i + NM * j) * nt + L.tid] = re;
            A.sig_im[(int64_t)(i + NM * j) * nt + L.tid] = im;
        }
    A.state[L.tid] = L.st;
    A.steps[L.tid] = L.step;
    A.nhops[L.tid] = L.nh;
}

template <int NM, typename R>
__device__ __forceinline__ void ns_step(const FsshNArgs* Ap, const FsshNArgs& A, LaneN<NM, R>& L) {
    constexpr int NP = LaneN<NM, R>::NP;
    constexpr int NQ = NP > 0 ? NP : 1;
    const R hdt = (R)(0.5 * A.dt), dtR = (R)A.dt;
    const R vt = L.v + hdt * L.k;                                     // step_B!
    L.r = L.r + dtR * vt;                                             // step_A!
    double vd[NM], vo[NQ], dd[NM], dof[NQ], E[NM], Gd[NM], Go[NQ];
    ns_model<NM, R>(L, L.r, vd, vo, dd, dof);                         // update_cache!
    ns_eigen<NM>(L.n, vd, vo, E, L.U, true);
    ns_rotate<NM>(L.U, dd, dof, Gd, Go);
    ns_accel<NM, R>(A, L, Gd, Go, L.a, L.x);                          // sigma of the step start
    L.v = vt + hdt * L.k;                                             // step_B!
    // nonadiabatic couplings d_ij = G_ij / (E_j - E_i) and W = v d on the trajectory's n states
    double d[NQ];
    NsInterp<NM> I;
#pragma unroll
    for (int i = 0; i < NM; ++i)
#pragma unroll
        for (int j = i + 1; j < NM; ++j) {
            const int p = pidx<NM>(i, j);
            d[p] = (j < L.n) ? Go[p] / (E[j] - E[i]) : 0.0;
            I.Wn[p] = d[p] * (double)L.v;
            I.Wc[p] = L.Wc[p];
        }
#pragma unroll
    for (int i = 0; i < NM; ++i) { I.Ec[i] = L.Ec[i]; I.En[i] = E[i]; }
    const double tn = L.t + A.dt;
    I.tc = L.tc; I.tn = tn;
    {   // step!(electronic_integrator, dt, true): nsub sub-steps clipped to t + dt (bab_electronics.jl:87-90)
        const double h = A.dt / (double)A.nsub;
        double tau = L.t;
        for (int s = 0; s < A.nsub; ++s) {
            const double hs = fmin(h, tn - tau);
            ns_tsit5<NM>(I, tau, hs, L.a, L.x, L.y);
            tau = (s == A.nsub - 1) ? tn : tau + hs;
        }
    }
#pragma unroll
    for (int i = 0; i < NM; ++i) L.Ec[i] = E[i];
#pragma unroll
    for (int p = 0; p < NP; ++p) L.Wc[p] = I.Wn[p];
    L.tc = tn;
    L.t = tn;
    // the save precedes the hop callback (a save time holds the state before the step's callbacks)
    if (A.save_every > 0 && ((L.step + 1) % A.save_every) == 0) ns_save<NM, R>(Ap, L, (L.step + 1) / A.save_every);
    if (A.method == 0) {  // HoppingCallback (surface_hopping.jl:2-16), fssh.jl:96-121
        const int s = L.st;
        double ass = 0.0;
#pragma unroll
        for (int i = 0; i < NM; ++i) ass = (i == s) ? L.a[i] : ass;
        double cum[NM];
        double run = 0.0;
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            double pm = 0.0;
            if (m != s && m < L.n) {
                double xms = 0.0, dsm = 0.0;  // Re sigma_ms and d_sm
#pragma unroll
                for (int q = 0; q < NM; ++q) {
                    if (q == s) {
                        if (m < q) { xms = L.x[pidx<NM>(m, q)]; dsm = -d[pidx<NM>(m, q)]; }
                        else if (m > q) { xms = L.x[pidx<NM>(q, m)]; dsm = d[pidx<NM>(q, m)]; }
                    }
                }
                pm = 2.0 * (double)L.v * (xms / ass) * dsm * A.dt;
                pm = fmin(fmax(pm, 0.0), 1.0);
            }
            run += pm;
            cum[m] = run;
        }
        if (run > 0.0) {
            const double rnd = philox_uniform((uint64_t)L.tid, (uint32_t)L.step, STREAM_HOP, A.key);
            int ns = s;
            bool found = false;
#pragma unroll
            for (int m = 0; m < NM; ++m)
                if (!found && m != s && m < L.n && cum[m] > rnd) { ns = m; found = true; }
            if (found) {
                double dno = 0.0, dE = 0.0;  // coupling[new, old], E_new - E_old
#pragma unroll
                for (int i = 0; i < NM; ++i)
#pragma unroll
                    for (int j = i + 1; j < NM; ++j) {
                        if (i == ns && j == s) dno = d[pidx<NM>(i, j)];
                        if (i == s && j == ns) dno = -d[pidx<NM>(i, j)];
                    }
#pragma unroll
                for (int i = 0; i < NM; ++i) { if (i == ns) dE += E[i]; if (i == s) dE -= E[i]; }
                R vv[1] = {L.v};
                const double dc[1] = {dno};
                const double im[1] = {A.inv_mass};
                if (rescale_velocity<1, R>(A.rescaling, vv, im, dc, dE)) { L.st = ns; ++L.nh; }
                L.v = vv[0];
            }
        }
    }
    if (A.cell.mode == 1) {  // CellBoundaryCallback after the hop callback; the step-end acceleration is kept
        R rr[1] = {L.r};
        if (cell_outside<1, 1, R>(A.cell, rr)) { cell_wrap<1, 1, R>(A.cell, rr); L.r = rr[0]; }
    }
    ++L.step;
}

template <int NM, typename R>
__device__ __forceinline__ void ns_body(const FsshNArgs* __restrict__ Ap) {
    const FsshNArgs& A = *Ap;
    const int lane = __lane_id();
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    LaneN<NM, R> L;
    bool active = false, exhausted = false;
    for (;;) {
        const uint64_t want = __ballot(!active && !exhausted);
        if (want) {
            const int cnt = __popcll(want);
            const int leader = __ffsll((unsigned long long)want) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(ns_opaque(Ap)->queue, (unsigned long long)cnt);
            base = __shfl(base, leader);
            if (!active && !exhausted) {
                const unsigned long long q = base + (unsigned long long)__popcll(want & below);
                if (q < (unsigned long long)A.ntraj) {
                    const int32_t* order = ns_opaque(Ap)->order;
                    ns_load<NM, R>(Ap, L, order ? (int64_t)order[q] : (int64_t)q);
                    if (A.nsteps <= 0) ns_finish<NM, R>(Ap, L); else active = true;
                } else {
                    exhausted = true;
                }
            }
        }
        if (__ballot(active) == 0) break;
        if (active) {
            ns_step<NM, R>(Ap, A, L);
            // TerminatingCallback after each step, not at t0 (callbacks.jl:15-29)
            const R rr[1] = {L.r};
            if (L.step >= A.nsteps || (A.cell.mode == 2 && cell_outside<1, 1, R>(A.cell, rr)) ||
                term_hit(A.term_kind, (double)L.r, A.term_value)) {
                ns_finish<NM, R>(Ap, L);
                active = false;
            }
        }
    }
}

#define NQCD_NS_KERNELS(NM)                                                                              \
    __global__ __launch_bounds__(256) void fssh_n_f64_##NM(const FsshNArgs* __restrict__ Ap) {           \
        ns_body<NM, double>(Ap);                                                                         \
    }                                                                                                    \
    __global__ __launch_bounds__(256) void fssh_n_f32_##NM(const FsshNArgs* __restrict__ Ap) {           \
        ns_body<NM, float>(Ap);                                                                          \
    }

}  // namespace nqcd
