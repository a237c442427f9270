// mapping.hip -- spin-mapping dynamics (Simulation{SpinMappingW}, spin_mapping.jl:1-131) integrated by
// MInt (mint.jl:42-181), one coordinate, two to four electronic states, FP64, one trajectory per lane.
//
// With gamma = 2 (sqrt(n + 1) - 1) / n the mapping Hamiltonian of positions X and momenta P is
//   H = m v^2 / 2 + sum_i V_ii (X_i^2 + P_i^2 - gamma) / 2 + sum_{i<j} V_ij (X_i X_j + P_i P_j)
// (classical_potential_energy, spin_mapping.jl:102-118; the models here have no state-independent
// part).  An MInt step:
//   r' = r + dt/2 v;  (w, Z) the eigenpairs of V(r');  a = Z^T X, b = Z^T P;
//   a_k <- cos(w_k dt) a_k + sin(w_k dt) b_k,  b_k <- cos(w_k dt) b_k - sin(w_k dt) a_k
//     (X <- C X - D P, P <- C P + D X with C = Z cos(w dt) Z^T, D = Z sin(-w dt) Z^T, mint.jl:88-124,
//     applied in the eigenbasis);
//   W = Z^T dV Z;  Gamma_ii = W_ii dt,  Gamma_ij = sin(D_ij dt) W_ij / D_ij,
//   Xi_ij = (1 - cos(D_ij dt)) W_ij / D_ij with D_ij = w_j - w_i (i != j)  (mint.jl:132-162);
//   force = (X^T E X + P^T E P) / 2 - X^T F P with E = Z Gamma Z^T, F = Z Xi Z^T  (mint.jl:164-181),
//     evaluated in the eigenbasis as (a^T Gamma a + b^T Gamma b) / 2 - a^T Xi b;
//   v <- v - force / m + tr(dV) gamma dt / (2 m);  r <- r' + dt/2 v.
// Xi's off-diagonal part is antisymmetric (D_ji = -D_ij, W symmetric).  The reference's set_xi! leaves
// the diagonal of its cache matrix at the previous step's diagonal of F; F's diagonal is the quadratic
// form of Xi in the rows of Z, to which the antisymmetric part contributes nothing, so starting from a
// zero matrix that diagonal stays zero and the carried term is dropped here.
//
// Initial mapping variables (spin_mapping.jl:15-33): theta_i = 2 pi u with u from Philox stream 4 at
// counter (trajectory, i), radius sqrt(2 + gamma) on the occupied diabatic state and sqrt(gamma) on
// the others, X_i = radius sin(theta_i), P_i = radius cos(theta_i) (the reference stores the sines as
// the mapping positions); or uploaded by nqcd_mi355x_load_mapping.
//
// eCMM (Simulation{eCMM}, cmm.jl:1-120) runs the same kernel: its Hamiltonian
// (q^T V q + p^T V p - 2 gamma tr V) / 2 is the form above with 2 gamma in place of spin mapping's gamma on
// the diagonal; MInt's trace term is tr(dV) gamma dt / 2m for either method (mint.jl:85), which for eCMM
// is half of its Hamiltonian's and is kept as the reference computes it; its initial variables are a point
// uniform on the 2n-sphere of radius sqrt(2 + 2 n gamma) (2n Philox normals, pair k from stream 4 at
// counter (trajectory, k), normalised; the first n coordinates are the positions); its diabatic
// population is the inverse mapping kernel (1 + n) / (1 + n gamma)^2 (X^2 + P^2) / 2 - (1 - gamma) /
// (1 + n gamma) (cmm.jl:96-112).
//
// Device layout: X, P [state][traj]; diabatic populations (spin mapping (X_i^2 + P_i^2 - gamma) / 2,
// spin_mapping.jl:120-131) [slot][state][traj]; kinetic and potential energy [slot][traj].
#include "mapping.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "ensemble.h"
#include "fssh_n_kern.h"

namespace nqcd {

constexpr uint32_t kMapStream = 4;

struct MapArgs {
    int64_t ntraj, nsteps;
    int32_t save_every, nsave;
    int32_t slot_tab[kMaxModels];   // offset of a MORSE_N slot's table in mtab, -1 for closed-form models
    int32_t ecmm, pad;              // 1: eCMM's initial variables (sphere) instead of spin mapping's
    double dt, inv_mass, mass;
    double gamma;                   // constant of the diagonal terms V_ii (X_i^2 + P_i^2 - gamma) / 2 of H
    double gamma_force;             // gamma of MInt's trace term tr(dV) gamma dt / 2m
    double pop_a, pop_b;            // population estimator pop_a (X_i^2 + P_i^2) / 2 - pop_b
    double radius;                  // eCMM: radius of the initial sphere
    double occ_radius, free_radius; // spin mapping: initial radii of the occupied and the other states
    uint64_t key;
    const ModelSlot* models;
    const double* mtab;
    const double* r0;
    const double* v0;
    const double* X0;        // [state][traj] uploaded mapping variables, or null (drawn)
    const double* P0;
    const int32_t* state0;   // occupied diabatic state per trajectory
    const int32_t* model_id; // or null (slot 0)
    double* r;
    double* v;
    double* X;
    double* P;
    int32_t* steps;          // nuclear steps taken (every trajectory runs to t1)
    double* save_pop;        // [slot][state][traj]
    double* save_ekin;       // [slot][traj]
    double* save_epot;
    double* init_pop;        // [state][traj] initial estimator (X^2 + P^2) / 2 - k0_b of the population correlation
    double k0_b;
};

// Potential part of H at coordinate x for mapping variables X, P.
template <int NM>
__device__ __forceinline__ double map_potential(const LaneN<NM, double>& L, double x, const double X[NM],
                                                const double P[NM], double gamma) {
    constexpr int NP = NM * (NM - 1) / 2;
    double vd[NM], vo[NP], dd[NM], dof[NP];
    ns_model<NM, double>(L, x, vd, vo, dd, dof);
    double e = 0.0;
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        e += vd[i] * (X[i] * X[i] + P[i] * P[i] - gamma) * 0.5;
#pragma unroll
        for (int j = i + 1; j < NM; ++j) e += vo[pidx<NM>(i, j)] * (X[i] * X[j] + P[i] * P[j]);
    }
    return e;
}

// One step of the mapping variables at coordinate x and the nuclear force over the step (MInt, mint.jl:42-90):
// the eigenpairs (w, Z) of V(x); X, P rotated in the eigenbasis by the angles (w_k - shift) dt; the force
// (a^T Gamma a + b^T Gamma b) / 2 - a^T Xi b of the rotated variables with Gamma, Xi from
// W = Z^T (dV - dshift I) Z.  TRACELESS (RingPolymerMInt, ringpolymer_mint.jl:65-101): shift = tr V / NM and
// dshift = tr dV / NM, the traceless eigenvalues and derivative; otherwise both zero.  Returns f and
// tr = tr dV.  Z holds the eigenvectors on return; above two states the Jacobi sweeps start from the Z it holds
// on entry (the previous step's, or the identity), which the step's results do not depend on beyond rounding.
template <int NM, bool TRACELESS>
__device__ __forceinline__ void map_electronic_step(const LaneN<NM, double>& L, double x, double X[NM], double P[NM],
                                                    double dt, double Z[NM * NM], double& f, double& tr) {
    constexpr int NP = NM * (NM - 1) / 2;
    double vd[NM], vo[NP > 0 ? NP : 1], dd[NM], dof[NP > 0 ? NP : 1];
    ns_model<NM, double>(L, x, vd, vo, dd, dof);
    double w[NM];
    if constexpr (NM == 2) eigen2(vd[0], vo[0], vd[1], w[0], w[1], Z, false);
    else ns_eigen<NM>(NM, vd, vo, w, Z, false, true);
    double vbar = 0.0, dbar = 0.0;
    tr = 0.0;
#pragma unroll
    for (int i = 0; i < NM; ++i) { tr += dd[i]; vbar += vd[i]; }
    if constexpr (TRACELESS) { vbar *= 1.0 / NM; dbar = tr * (1.0 / NM); }
    else vbar = 0.0;
    // mapping variables in the eigenbasis, rotated by exp(-i (w - vbar) dt)
    double a[NM], b[NM], sk[NM], ck[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k) {
        double sa = 0.0, sb = 0.0;
#pragma unroll
        for (int i = 0; i < NM; ++i) { sa += Z[i * NM + k] * X[i]; sb += Z[i * NM + k] * P[i]; }
        sincos((w[k] - vbar) * dt, &sk[k], &ck[k]);
        a[k] = ck[k] * sa + sk[k] * sb;
        b[k] = ck[k] * sb - sk[k] * sa;
    }
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        double sx = 0.0, sp = 0.0;
#pragma unroll
        for (int k = 0; k < NM; ++k) { sx += Z[i * NM + k] * a[k]; sp += Z[i * NM + k] * b[k]; }
        X[i] = sx;
        P[i] = sp;
    }
    // nuclear force from the rotated variables
    double Wd[NM], Wo[NP > 0 ? NP : 1];
    ns_rotate<NM>(Z, dd, dof, Wd, Wo);
    f = 0.0;
#pragma unroll
    for (int i = 0; i < NM; ++i) {
        f += 0.5 * (Wd[i] - dbar) * dt * (a[i] * a[i] + b[i] * b[i]);
#pragma unroll
        for (int j = i + 1; j < NM; ++j) {
            // sin and cos of (w_j - w_i) dt from those of the rotation angles
            const double dl = w[j] - w[i];
            const double s = sk[j] * ck[i] - ck[j] * sk[i], c = ck[j] * ck[i] + sk[j] * sk[i];
            const double wij = Wo[pidx<NM>(i, j)] / dl;
            f += s * wij * (a[i] * a[j] + b[i] * b[j]) - (1.0 - c) * wij * (a[i] * b[j] - a[j] * b[i]);
        }
    }
}

template <int NM>
__global__ __launch_bounds__(256) void mapping_kernel(const MapArgs A) {
    const int64_t nt = A.ntraj;
    const double dt = A.dt, hdt = 0.5 * A.dt, g = A.gamma, gf = A.gamma_force, im = A.inv_mass;
    for (int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; tid < nt; tid += (int64_t)gridDim.x * blockDim.x) {
        const int mid = A.model_id ? A.model_id[tid] : 0;
        LaneN<NM, double> L;
        L.n = NM;
        L.tab = A.slot_tab[mid] >= 0 ? A.mtab + A.slot_tab[mid] : nullptr;
        L.m = model_regs(A.models, mid);
        double x = A.r0[tid], v = A.v0[tid];
        double X[NM], P[NM];
        if (A.X0) {
#pragma unroll
            for (int i = 0; i < NM; ++i) { X[i] = A.X0[(int64_t)i * nt + tid]; P[i] = A.P0[(int64_t)i * nt + tid]; }
        } else if (A.ecmm) {
            double z[2 * NM], nrm = 0.0;
#pragma unroll
            for (int k = 0; k < NM; ++k) {
                philox_normal2((uint64_t)tid, (uint32_t)k, kMapStream, A.key, z[2 * k], z[2 * k + 1]);
                nrm += z[2 * k] * z[2 * k] + z[2 * k + 1] * z[2 * k + 1];
            }
            const double sc = A.radius / sqrt(nrm);
#pragma unroll
            for (int i = 0; i < NM; ++i) { X[i] = z[i] * sc; P[i] = z[NM + i] * sc; }
        } else {
            const int st = A.state0[tid];
#pragma unroll
            for (int i = 0; i < NM; ++i) {
                const double th = 6.283185307179586 * philox_uniform((uint64_t)tid, (uint32_t)i, kMapStream, A.key);
                const double rad = i == st ? A.occ_radius : A.free_radius;
                double s, c;
                sincos(th, &s, &c);
                X[i] = rad * s;
                P[i] = rad * c;
            }
        }
#pragma unroll
        for (int i = 0; i < NM; ++i) A.init_pop[(int64_t)i * nt + tid] = (X[i] * X[i] + P[i] * P[i]) * 0.5 - A.k0_b;
        auto save = [&](int slot) {
#pragma unroll
            for (int i = 0; i < NM; ++i)
                A.save_pop[((int64_t)slot * NM + i) * nt + tid] = A.pop_a * ((X[i] * X[i] + P[i] * P[i]) * 0.5) - A.pop_b;
            A.save_ekin[(int64_t)slot * nt + tid] = 0.5 * A.mass * v * v;
            A.save_epot[(int64_t)slot * nt + tid] = map_potential<NM>(L, x, X, P, g);
        };
        save(0);
        int slot = 1, until = A.save_every;
        double Z[NM * NM];
#pragma unroll
        for (int i = 0; i < NM * NM; ++i) Z[i] = (i % (NM + 1) == 0) ? 1.0 : 0.0;
        for (int64_t step = 0; step < A.nsteps; ++step) {
            const double rt = x + hdt * v;
            double f, tr;
            map_electronic_step<NM, false>(L, rt, X, P, dt, Z, f, tr);
            v = v - f * im + tr * gf * dt * 0.5 * im;
            x = rt + hdt * v;
            if (A.save_every > 0 && step + 1 == until) {
                if (slot < A.nsave) save(slot);
                ++slot;
                until += A.save_every;
            }
        }
        A.r[tid] = x;
        A.v[tid] = v;
        A.steps[tid] = (int32_t)A.nsteps;
#pragma unroll
        for (int i = 0; i < NM; ++i) { A.X[(int64_t)i * nt + tid] = X[i]; A.P[(int64_t)i * nt + tid] = P[i]; }
    }
}

template __global__ void mapping_kernel<2>(const MapArgs);
template __global__ void mapping_kernel<3>(const MapArgs);
template __global__ void mapping_kernel<4>(const MapArgs);

// ---------------------------------------------------------------------------
// Ring polymers by RingPolymerMInt (ringpolymer_mint.jl:28-131): NRPMD (nrpmd.jl) and ring-polymer eCMM
// (rpecmm.jl).  One bead per lane: a 64-lane workgroup holds 64 / P ring polymers of P beads; lane b of a ring
// polymer holds normal mode b (position and velocity) and bead b's mapping variables.  A step: every lane
// applies the free ring polymer's half step sqrt(cay(dt A)) = [[1, dt/2], [-w_b^2 dt/2, 1]] / sqrt(1 +
// dt^2 w_b^2 / 4) to its mode (ring_polymer.jl:71-82); the bead positions r_b = sum_k U[b][k] q_k come
// through LDS; each lane runs map_electronic_step (traceless) at its bead and forms its bead's velocity
// change -(force + Dbar dt) / m; the changes go back to the modes through LDS (U is orthogonal); the second
// half step.  U is the real DFT basis (column k spans the spring eigenspace of w_k = 2 w_n sin(k pi / P)),
// kept in LDS.  Device layout: r, v [bead][ring polymer]; X, P [bead * n + state][ring polymer]; saves
// [slot][state][ring polymer] and [slot][ring polymer].
// ---------------------------------------------------------------------------
struct RpMapArgs {
    int64_t nrp, nsteps;
    int32_t nbeads, save_every, nsave, ecmm;
    int32_t slot_tab[kMaxModels];
    double dt, mass, inv_mass, wn;
    double pop_a, pop_b, pop_scale;   // population: pop_scale sum_b [pop_a (X^2 + P^2) / 2 - pop_b]
    double occ_radius, free_radius;   // NRPMD initial radii
    double radius;                    // ring-polymer eCMM: radius of the initial sphere
    uint64_t key;
    const ModelSlot* models;
    const double* mtab;
    const double* r0;
    const double* v0;
    const double* X0;                 // or null (drawn)
    const double* P0;
    const int32_t* state0;
    const int32_t* model_id;
    double* r;
    double* v;
    double* X;
    double* Pmap;
    int32_t* steps;
    double* save_pop;
    double* save_ekin;                // kinetic energy
    double* save_epot;                // potential energy with the springs
    double* save_centroid;
    double* init_pop;                 // [state][ring polymer] k0_scale sum_b [(X^2 + P^2) / 2 - k0_b] at t0
    double k0_b, k0_scale;
};

// sum over the lanes g P .. g P + P - 1 of val, returned to every lane of the group (LDS buf of 64 doubles)
__device__ __forceinline__ double rpmap_group_sum(double* buf, int l, int g, int P, double val) {
    buf[l] = val;
    __syncthreads();
    double s = 0.0;
    for (int j = 0; j < P; ++j) s += buf[g * P + j];
    __syncthreads();
    return s;
}

// y_b = sum_k U[b][k] x_k (to beads), or with `to_modes` y_k = sum_b U[b][k] x_b, for the lane's index
__device__ __forceinline__ double rpmap_transform(double* buf, const double* U, int l, int g, int P, int idx, double x,
                                                  bool to_modes) {
    buf[l] = x;
    __syncthreads();
    double s = 0.0;
    for (int j = 0; j < P; ++j) s += (to_modes ? U[j * P + idx] : U[idx * P + j]) * buf[g * P + j];
    __syncthreads();
    return s;
}

template <int NM>
__global__ __launch_bounds__(64) void rpmap_kernel(const RpMapArgs A) {
    extern __shared__ double lds[];
    const int P = A.nbeads;
    double* U = lds;                 // P * P
    double* buf = lds + P * P;       // 64
    const int l = threadIdx.x;
    const int G = 64 / P;            // ring polymers per workgroup
    const int g0 = l / P, b = l % P;
    // lanes beyond the last whole ring polymer (64 not a multiple of P) shadow group 0: they read its
    // exchange slots and store nothing
    const int g = g0 < G ? g0 : 0;
    // the real DFT basis U[bead][mode]
    for (int e = l; e < P * P; e += 64) {
        const int j = e / P, k = e % P;
        double u;
        if (k == 0) u = 1.0 / sqrt((double)P);
        else if (2 * k < P) u = sqrt(2.0 / P) * cospi(2.0 * (double)((j * k) % P) / P);
        else if (2 * k == P) u = ((j % 2) ? -1.0 : 1.0) / sqrt((double)P);
        else u = sqrt(2.0 / P) * sinpi(2.0 * (double)((j * k) % P) / P);
        U[e] = u;
    }
    __syncthreads();
    const double dt = A.dt, im = A.inv_mass, m = A.mass;
    // half step of mode b
    const double wk = 2.0 * A.wn * sinpi((double)b / P);
    const double hs = 1.0 / sqrt(1.0 + dt * dt * wk * wk / 4.0);
    const double c12 = 0.5 * dt * hs, c21 = -0.5 * dt * wk * wk * hs;
    const int64_t nrp = A.nrp;
    for (int64_t base = (int64_t)blockIdx.x * G; base < nrp; base += (int64_t)gridDim.x * G) {
        const int64_t rp = base + g;
        const bool on = g0 < G && rp < nrp;
        const int64_t rq = on ? rp : 0;   // loads of idle lanes read ring polymer 0 and are not stored
        const int mid = A.model_id ? A.model_id[rq] : 0;
        LaneN<NM, double> L;
        L.n = NM;
        L.tab = A.slot_tab[mid] >= 0 ? A.mtab + A.slot_tab[mid] : nullptr;
        L.m = model_regs(A.models, mid);
        const int bb = on ? b : 0;
        double q = rpmap_transform(buf, U, l, g, P, b, A.r0[(int64_t)bb * nrp + rq], true);
        double p = rpmap_transform(buf, U, l, g, P, b, A.v0[(int64_t)bb * nrp + rq], true);
        double X[NM], Pm[NM];
        if (A.X0) {
#pragma unroll
            for (int i = 0; i < NM; ++i) {
                X[i] = A.X0[(int64_t)(bb * NM + i) * nrp + rq];
                Pm[i] = A.P0[(int64_t)(bb * NM + i) * nrp + rq];
            }
        } else if (A.ecmm) {
            // coordinates b NM + i (positions) and P NM + b NM + i (momenta) of the 2 NM P normals
            const int FN = NM * P;
            double ss = 0.0;
#pragma unroll
            for (int i = 0; i < NM; ++i) {
                double z0, z1;
                const int jq = bb * NM + i, jp = FN + bb * NM + i;
                philox_normal2((uint64_t)rq, (uint32_t)(jq / 2), kMapStream, A.key, z0, z1);
                X[i] = (jq & 1) ? z1 : z0;
                philox_normal2((uint64_t)rq, (uint32_t)(jp / 2), kMapStream, A.key, z0, z1);
                Pm[i] = (jp & 1) ? z1 : z0;
                ss += X[i] * X[i] + Pm[i] * Pm[i];
            }
            const double tot = rpmap_group_sum(buf, l, g, P, ss);
            const double sc = A.radius / sqrt(tot);
#pragma unroll
            for (int i = 0; i < NM; ++i) { X[i] *= sc; Pm[i] *= sc; }
        } else {
            const int st = A.state0[rq];
#pragma unroll
            for (int i = 0; i < NM; ++i) {
                const double th = 6.283185307179586 * philox_uniform((uint64_t)rq, (uint32_t)(bb * NM + i), kMapStream, A.key);
                const double rad = i == st ? A.occ_radius : A.free_radius;
                double sn, cs;
                sincos(th, &sn, &cs);
                X[i] = rad * cs;
                Pm[i] = rad * sn;
            }
        }
#pragma unroll
        for (int i = 0; i < NM; ++i) {
            const double k0 = A.k0_scale * rpmap_group_sum(buf, l, g, P, (X[i] * X[i] + Pm[i] * Pm[i]) * 0.5 - A.k0_b);
            if (on && b == 0) A.init_pop[(int64_t)i * nrp + rp] = k0;
        }
        auto save = [&](int slot) {
            const double rb = rpmap_transform(buf, U, l, g, P, b, q, false);
            double pot = map_potential<NM>(L, rb, X, Pm, 0.0);
            // Vbar + (p^T V~ p + q^T V~ q) / 2 = map_potential with gamma 0 less Vbar (sum X^2 + P^2 - 2) / 2
            {
                double vd[NM], vo[NM * (NM - 1) / 2 > 0 ? NM * (NM - 1) / 2 : 1], dd[NM], dof[NM * (NM - 1) / 2 > 0 ? NM * (NM - 1) / 2 : 1];
                ns_model<NM, double>(L, rb, vd, vo, dd, dof);
                double vb = 0.0, n2 = 0.0;
#pragma unroll
                for (int i = 0; i < NM; ++i) { vb += vd[i]; n2 += X[i] * X[i] + Pm[i] * Pm[i]; }
                vb *= 1.0 / NM;
                pot -= vb * (0.5 * n2 - 1.0);
            }
            const double ek = 0.5 * m * p * p, es = 0.5 * m * wk * wk * q * q;
            const double ekin = rpmap_group_sum(buf, l, g, P, ek);
            const double epot = rpmap_group_sum(buf, l, g, P, pot + es);
            const double cen = rpmap_group_sum(buf, l, g, P, rb);
            double pops[NM];
#pragma unroll
            for (int i = 0; i < NM; ++i)
                pops[i] = A.pop_scale * rpmap_group_sum(buf, l, g, P, A.pop_a * ((X[i] * X[i] + Pm[i] * Pm[i]) * 0.5) - A.pop_b);
            if (on && b == 0) {
#pragma unroll
                for (int i = 0; i < NM; ++i) A.save_pop[((int64_t)slot * NM + i) * nrp + rp] = pops[i];
                A.save_ekin[(int64_t)slot * nrp + rp] = ekin;
                A.save_epot[(int64_t)slot * nrp + rp] = epot;
                A.save_centroid[(int64_t)slot * nrp + rp] = cen / P;
            }
        };
        save(0);
        int slot = 1, until = A.save_every;
        double Z[NM * NM];
#pragma unroll
        for (int i = 0; i < NM * NM; ++i) Z[i] = (i % (NM + 1) == 0) ? 1.0 : 0.0;
        for (int64_t step = 0; step < A.nsteps; ++step) {
            double q1 = hs * q + c12 * p, p1 = c21 * q + hs * p;
            const double rb = rpmap_transform(buf, U, l, g, P, b, q1, false);
            double f, tr;
            map_electronic_step<NM, true>(L, rb, X, Pm, dt, Z, f, tr);
            const double dvb = -(f + tr * (1.0 / NM) * dt) * im;
            p1 += rpmap_transform(buf, U, l, g, P, b, dvb, true);
            q = hs * q1 + c12 * p1;
            p = c21 * q1 + hs * p1;
            if (A.save_every > 0 && step + 1 == until) {
                if (slot < A.nsave) save(slot);
                ++slot;
                until += A.save_every;
            }
        }
        const double rb = rpmap_transform(buf, U, l, g, P, b, q, false);
        const double vb = rpmap_transform(buf, U, l, g, P, b, p, false);
        if (on) {
            A.r[(int64_t)b * nrp + rp] = rb;
            A.v[(int64_t)b * nrp + rp] = vb;
#pragma unroll
            for (int i = 0; i < NM; ++i) {
                A.X[(int64_t)(b * NM + i) * nrp + rp] = X[i];
                A.Pmap[(int64_t)(b * NM + i) * nrp + rp] = Pm[i];
            }
            if (b == 0) A.steps[rp] = (int32_t)A.nsteps;
        }
    }
}

template __global__ void rpmap_kernel<2>(const RpMapArgs);
template __global__ void rpmap_kernel<3>(const RpMapArgs);
template __global__ void rpmap_kernel<4>(const RpMapArgs);

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
bool map_is_problem(const nqcd_handle_t* h) {
    return h->prob.method == NQCD_METHOD_SPIN_MAPPING || h->prob.method == NQCD_METHOD_ECMM ||
           h->prob.method == NQCD_METHOD_NRPMD;
}

// ring polymers by RingPolymerMInt (NRPMD, ring-polymer eCMM)
static bool map_rp(const nqcd_handle_t* h) { return h->prob.algorithm == NQCD_ALG_RP_MINT; }
// beads per trajectory (1 for MInt)
static int map_beads(const nqcd_handle_t* h) { return map_rp(h) ? h->prob.nbeads : 1; }

static double map_gamma(int n) { return 2.0 / n * (std::sqrt((double)n + 1.0) - 1.0); }

static int map_slot_states(const nqcd_handle_t* h, int s) {
    return h->models[s].kind == NQCD_MODEL_MORSE_N ? (int)h->models[s].p[0] : 2;
}

int32_t map_check_problem(nqcd_handle_t* h) {
    const nqcd_problem_desc& d = h->prob;
    if (d.algorithm == NQCD_ALG_RP_MINT) {
        if (d.method != NQCD_METHOD_NRPMD && d.method != NQCD_METHOD_ECMM)
            return fail(h, NQCD_ERR_UNSUPPORTED, "RingPolymerMInt integrates RingPolymerSimulation{NRPMD | eCMM}");
        if (d.nbeads < 1 || d.nbeads > 64) return fail(h, NQCD_ERR_UNSUPPORTED, "NRPMD / ring-polymer eCMM: 1 to 64 beads");
        if (!(d.temperature > 0.0)) return fail(h, NQCD_ERR_ARG, "ring polymers need temperature > 0");
        if (d.method == NQCD_METHOD_NRPMD && !(d.mapping_gamma >= 0.0 && std::isfinite(d.mapping_gamma)))
            return fail(h, NQCD_ERR_ARG, "NRPMD: mapping_gamma must be finite and >= 0");
        if (d.method == NQCD_METHOD_ECMM &&
            (!std::isfinite(d.mapping_gamma) || 2.0 + 2.0 * d.n_states * d.mapping_gamma * d.nbeads <= 0.0))
            return fail(h, NQCD_ERR_ARG, "ring-polymer eCMM: mapping_gamma must be finite with 2 + 2 n gamma nbeads > 0");
    } else if (d.algorithm != NQCD_ALG_MINT || d.method == NQCD_METHOD_NRPMD || d.nbeads != 1) {
        return fail(h, NQCD_ERR_UNSUPPORTED, "Simulation{SpinMappingW | eCMM} integrates with MInt, "
                                             "RingPolymerSimulation{NRPMD | eCMM} with RingPolymerMInt");
    }
    if (d.n_states < 2 || d.n_states > kMapMaxStates || d.ndofs * d.natoms != 1)
        return fail(h, NQCD_ERR_UNSUPPORTED, "spin mapping / eCMM / NRPMD: 2 to 4 states, one nuclear coordinate");
    if (d.precision != NQCD_PREC_FP64) return fail(h, NQCD_ERR_UNSUPPORTED, "spin mapping / eCMM run in FP64");
    if (d.algorithm == NQCD_ALG_MINT && d.method == NQCD_METHOD_ECMM &&
        (!std::isfinite(d.mapping_gamma) || 2.0 + 2.0 * d.n_states * d.mapping_gamma <= 0.0))
        return fail(h, NQCD_ERR_ARG, "eCMM: mapping_gamma must be finite with 2 + 2 n gamma > 0");
    return NQCD_OK;
}

static int32_t map_check_models(nqcd_handle_t* h) {
    for (int s = 0; s < kMaxModels; ++s) {
        if (!h->model_set[s]) continue;
        const int k = h->models[s].kind;
        if (k != NQCD_MODEL_MORSE_N && k != NQCD_MODEL_TULLY1 && k != NQCD_MODEL_TULLY2 && k != NQCD_MODEL_TULLY3 &&
            k != NQCD_MODEL_DW_GAUSSIAN)
            return fail(h, NQCD_ERR_UNSUPPORTED, "spin mapping models: TULLY1/2/3, DW_GAUSSIAN, MORSE_N");
        if (map_slot_states(h, s) != h->prob.n_states)
            return fail(h, NQCD_ERR_ARG, "spin mapping: every model slot must have n_states states");
    }
    return NQCD_OK;
}

int32_t map_load_initial(nqcd_handle_t* h, int64_t ntraj, const double* r, const double* v, const int32_t* state,
                         const int32_t* model_id) {
    if (int32_t rc = map_check_models(h)) return rc;
    std::vector<int32_t> zeros;
    if (!state && h->prob.method == NQCD_METHOD_ECMM) {   // eCMM's draw does not depend on a state (cmm.jl:26)
        zeros.assign(ntraj, 0);
        state = zeros.data();
    }
    if (!state) return fail(h, NQCD_ERR_ARG, "spin mapping / NRPMD need the occupied diabatic state of each trajectory");
    h->ntraj = ntraj;
    const int B = map_beads(h);
    // positions / velocities (B, ntraj) column-major -> [bead][traj]
    std::vector<double> tr((size_t)B * ntraj), tv((size_t)B * ntraj);
    for (int64_t t = 0; t < ntraj; ++t)
        for (int b = 0; b < B; ++b) {
            tr[(size_t)b * ntraj + t] = r[(size_t)t * B + b];
            tv[(size_t)b * ntraj + t] = v[(size_t)t * B + b];
        }
    HIPCHK(h, h->r0.ensure(sizeof(double) * B * ntraj));
    HIPCHK(h, h->v0.ensure(sizeof(double) * B * ntraj));
    HIPCHK(h, h->state0.ensure(sizeof(int32_t) * ntraj));
    HIPCHK(h, hipMemcpyAsync(h->r0.p, tr.data(), sizeof(double) * B * ntraj, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->v0.p, tv.data(), sizeof(double) * B * ntraj, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipMemcpyAsync(h->state0.p, state, sizeof(int32_t) * ntraj, hipMemcpyHostToDevice, h->stream));
    h->model_id_given = model_id != nullptr;
    if (model_id) {
        HIPCHK(h, h->model_id.ensure(sizeof(int32_t) * ntraj));
        HIPCHK(h, hipMemcpyAsync(h->model_id.p, model_id, sizeof(int32_t) * ntraj, hipMemcpyHostToDevice, h->stream));
    }
    h->map_given = false;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return NQCD_OK;
}

int32_t map_sample_initial(nqcd_handle_t* h, const nqcd_sampler_desc* S, uint64_t seed) {
    if (int32_t rc = map_check_models(h)) return rc;
    if (S->kind != NQCD_SAMPLER_GAUSSIAN || S->nparams != 4 || !S->params)
        return fail(h, NQCD_ERR_UNSUPPORTED, "spin mapping samples with GAUSSIAN (params mu_r, sd_r, mu_v, sd_v)");
    if (S->state_basis != 1 && h->prob.method != NQCD_METHOD_ECMM)
        return fail(h, NQCD_ERR_UNSUPPORTED, "spin mapping / NRPMD start from a diabatic pure state (spin_mapping.jl:15, "
                                             "nrpmd.jl:49)");
    if (!h->model_set[0]) return fail(h, NQCD_ERR_STATE, "model slot 0 is not set");
    if (S->initial_state < 0 || S->initial_state >= h->prob.n_states) return fail(h, NQCD_ERR_ARG, "initial_state out of range");
    const int64_t ntraj = S->ntraj;
    h->ntraj = ntraj;
    const int B = map_beads(h);
    // every bead drawn independently from the same Gaussian (params repeated per bead)
    std::vector<double> prm((size_t)4 * B);
    for (int b = 0; b < B; ++b)
        for (int k = 0; k < 4; ++k) prm[(size_t)4 * b + k] = S->params[k];
    HIPCHK(h, h->r0.ensure(sizeof(double) * B * ntraj));
    HIPCHK(h, h->v0.ensure(sizeof(double) * B * ntraj));
    HIPCHK(h, h->small.ensure(sizeof(double) * prm.size()));
    HIPCHK(h, hipMemcpyAsync(h->small.p, prm.data(), sizeof(double) * prm.size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, launch_sample_gaussian(ntraj, B, h->small.as<double>(), seed, h->r0.as<double>(), h->v0.as<double>(),
                                     h->stream));
    std::vector<int32_t> st(ntraj, S->initial_state);
    HIPCHK(h, h->state0.ensure(sizeof(int32_t) * ntraj));
    HIPCHK(h, hipMemcpyAsync(h->state0.p, st.data(), sizeof(int32_t) * ntraj, hipMemcpyHostToDevice, h->stream));
    h->model_id_given = false;
    h->map_given = false;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return NQCD_OK;
}

int32_t map_load_variables(nqcd_handle_t* h, int64_t ntraj, const double* X, const double* P) {
    if (!X && !P) { h->map_given = false; return NQCD_OK; }
    if (!X || !P) return fail(h, NQCD_ERR_ARG, "mapping positions and momenta are given together");
    if (ntraj != h->ntraj) return fail(h, NQCD_ERR_ARG, "ntraj differs from the loaded initial conditions");
    const int n = h->prob.n_states, B = map_beads(h), rows = n * B;
    // (n, B, ntraj) column-major -> [bead * n + state][traj]
    std::vector<double> t((size_t)2 * rows * ntraj);
    for (int64_t k = 0; k < ntraj; ++k)
        for (int i = 0; i < rows; ++i) {
            t[(size_t)i * ntraj + k] = X[(size_t)k * rows + i];
            t[(size_t)(rows + i) * ntraj + k] = P[(size_t)k * rows + i];
        }
    HIPCHK(h, h->sig0.ensure(sizeof(double) * t.size()));
    HIPCHK(h, hipMemcpy(h->sig0.p, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice));
    h->map_given = true;
    return NQCD_OK;
}

// RingPolymerMInt launch: 64-lane workgroups of 64 / P ring polymers, U and the exchange slots in LDS.
static int32_t map_run_rp(nqcd_handle_t* h, int64_t nsteps, uint64_t key, const int32_t* slot_tab) {
    const int n = h->prob.n_states, P = h->prob.nbeads;
    const int64_t nt = h->ntraj;
    const bool ecmm = h->prob.method == NQCD_METHOD_ECMM;
    const double g = h->prob.mapping_gamma, NF = (double)n * P;
    RpMapArgs A{};
    A.nrp = nt;
    A.nsteps = nsteps;
    A.nbeads = P;
    A.save_every = h->save_every;
    A.nsave = h->nsave;
    A.ecmm = ecmm ? 1 : 0;
    for (int s = 0; s < kMaxModels; ++s) A.slot_tab[s] = slot_tab[s];
    A.dt = h->prob.dt;
    A.mass = h->masses[0];
    A.inv_mass = h->inv_mass[0];
    A.wn = P * h->prob.temperature;
    // NRPMD: bead mean of (X^2 + P^2)/2 - gamma (nrpmd.jl:110-121); eCMM: bead sum of the inverse mapping
    // kernel with N = P (rpecmm.jl:80-94)
    A.pop_a = ecmm ? (1.0 + NF) / ((1.0 + NF * g) * (1.0 + NF * g)) : 1.0;
    A.pop_b = ecmm ? (1.0 - g) / (1.0 + NF * g) : g;
    A.pop_scale = ecmm ? 1.0 : 1.0 / P;
    A.occ_radius = std::sqrt(2.0 + 2.0 * g);
    A.free_radius = std::sqrt(2.0 * g);
    A.radius = std::sqrt(2.0 + 2.0 * NF * g);
    A.key = key;
    A.models = h->models_dev.as<ModelSlot>();
    A.mtab = h->ns_tab_dev.as<double>();
    A.r0 = h->r0.as<double>();
    A.v0 = h->v0.as<double>();
    A.X0 = h->map_given ? h->sig0.as<double>() : nullptr;
    A.P0 = h->map_given ? h->sig0.as<double>() + (size_t)n * P * nt : nullptr;
    A.state0 = h->state0.as<int32_t>();
    A.model_id = h->model_id_given ? h->model_id.as<int32_t>() : nullptr;
    A.r = h->r.as<double>();
    A.v = h->v.as<double>();
    A.X = h->sig.as<double>();
    A.Pmap = h->sig.as<double>() + (size_t)n * P * nt;
    A.steps = h->steps.as<int32_t>();
    A.save_pop = h->save_di.as<double>();
    A.save_ekin = h->save_ekin.as<double>();
    A.save_epot = h->save_epot.as<double>();
    A.save_centroid = h->save_centroid.as<double>();
    // initial estimator of the population correlation: NRPMD's population itself (Estimators.jl:121), eCMM's
    // bead-summed mapping kernel (rpecmm.jl:95-103)
    A.init_pop = h->final_di.as<double>();
    A.k0_b = g;
    A.k0_scale = ecmm ? 1.0 : 1.0 / P;
    const int block = 64, G = 64 / P;
    const size_t lds = sizeof(double) * ((size_t)P * P + 64);
    int bpc = 0;
    switch (n) {
    case 2: HIPCHK(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, rpmap_kernel<2>, block, lds)); break;
    case 3: HIPCHK(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, rpmap_kernel<3>, block, lds)); break;
    default: HIPCHK(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, rpmap_kernel<4>, block, lds)); break;
    }
    if (bpc < 1) bpc = 1;
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((nt + G - 1) / G, (int64_t)bpc * h->cu_count));
    HIPCHK(h, hipEventRecord(h->ev0, h->stream));
    switch (n) {
    case 2: hipLaunchKernelGGL(rpmap_kernel<2>, dim3((unsigned)grid), dim3(block), lds, h->stream, A); break;
    case 3: hipLaunchKernelGGL(rpmap_kernel<3>, dim3((unsigned)grid), dim3(block), lds, h->stream, A); break;
    default: hipLaunchKernelGGL(rpmap_kernel<4>, dim3((unsigned)grid), dim3(block), lds, h->stream, A); break;
    }
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipEventRecord(h->ev1, h->stream));
    h->stats.block = block;
    h->stats.grid = (int)grid;
    h->stats.launches = 1;
    return NQCD_OK;
}

int32_t map_run(nqcd_handle_t* h, double t0, int64_t nsteps, uint64_t key) {
    (void)t0;
    if (int32_t rc = map_check_models(h)) return rc;
    if (!h->model_set[0] && !h->model_id_given) return fail(h, NQCD_ERR_STATE, "model slot 0 is not set");
    const int n = h->prob.n_states, B = map_beads(h);
    const int64_t nt = h->ntraj;
    const int ns = h->nsave;
    HIPCHK(h, h->r.ensure(sizeof(double) * B * nt));
    HIPCHK(h, h->v.ensure(sizeof(double) * B * nt));
    HIPCHK(h, h->sig.ensure(sizeof(double) * 2 * n * B * nt));
    HIPCHK(h, h->steps.ensure(sizeof(int32_t) * nt));
    HIPCHK(h, h->final_di.ensure(sizeof(double) * n * nt));
    if (map_rp(h)) {
        HIPCHK(h, h->save_centroid.ensure(sizeof(double) * ns * nt));
        HIPCHK(h, launch_zero(h->save_centroid.p, sizeof(double) * ns * nt, h->stream));
    }
    HIPCHK(h, h->save_di.ensure(sizeof(double) * n * ns * nt));
    HIPCHK(h, h->save_ekin.ensure(sizeof(double) * ns * nt));
    HIPCHK(h, h->save_epot.ensure(sizeof(double) * ns * nt));
    HIPCHK(h, launch_zero(h->save_di.p, sizeof(double) * n * ns * nt, h->stream));
    HIPCHK(h, launch_zero(h->save_ekin.p, sizeof(double) * ns * nt, h->stream));
    HIPCHK(h, launch_zero(h->save_epot.p, sizeof(double) * ns * nt, h->stream));
    MapArgs A{};
    std::vector<double> tab;
    for (int s = 0; s < kMaxModels; ++s) {
        A.slot_tab[s] = -1;
        if (h->model_set[s] && h->models[s].kind == NQCD_MODEL_MORSE_N) {
            A.slot_tab[s] = (int32_t)tab.size();
            tab.insert(tab.end(), h->ns_tab[s].begin(), h->ns_tab[s].end());
        }
    }
    if (tab.empty()) tab.push_back(0.0);
    HIPCHK(h, h->ns_tab_dev.ensure(sizeof(double) * tab.size()));
    HIPCHK(h, hipMemcpyAsync(h->ns_tab_dev.p, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, h->models_dev.ensure(sizeof(ModelSlot) * kMaxModels));
    HIPCHK(h, hipMemcpyAsync(h->models_dev.p, h->models, sizeof(ModelSlot) * kMaxModels, hipMemcpyHostToDevice, h->stream));
    if (map_rp(h)) return map_run_rp(h, nsteps, key, A.slot_tab);
    A.ntraj = nt;
    A.nsteps = nsteps;
    A.save_every = h->save_every;
    A.nsave = ns;
    A.dt = h->prob.dt;
    A.mass = h->masses[0];
    A.inv_mass = h->inv_mass[0];
    const bool ecmm = h->prob.method == NQCD_METHOD_ECMM;
    const double ge = h->prob.mapping_gamma, gs = map_gamma(n);
    A.ecmm = ecmm ? 1 : 0;
    A.gamma = ecmm ? 2.0 * ge : gs;
    A.gamma_force = ecmm ? ge : gs;
    A.pop_a = ecmm ? (1.0 + n) / ((1.0 + n * ge) * (1.0 + n * ge)) : 1.0;
    A.pop_b = ecmm ? (1.0 - ge) / (1.0 + n * ge) : 0.5 * gs;
    A.radius = std::sqrt(2.0 + 2.0 * n * ge);
    A.occ_radius = std::sqrt(2.0 + gs);
    A.free_radius = std::sqrt(gs);
    A.key = key;
    A.models = h->models_dev.as<ModelSlot>();
    A.mtab = h->ns_tab_dev.as<double>();
    A.r0 = h->r0.as<double>();
    A.v0 = h->v0.as<double>();
    A.X0 = h->map_given ? h->sig0.as<double>() : nullptr;
    A.P0 = h->map_given ? h->sig0.as<double>() + (size_t)n * nt : nullptr;
    A.state0 = h->state0.as<int32_t>();
    A.model_id = h->model_id_given ? h->model_id.as<int32_t>() : nullptr;
    A.r = h->r.as<double>();
    A.v = h->v.as<double>();
    A.X = h->sig.as<double>();
    A.P = h->sig.as<double>() + (size_t)n * nt;
    A.steps = h->steps.as<int32_t>();
    A.save_pop = h->save_di.as<double>();
    A.save_ekin = h->save_ekin.as<double>();
    A.save_epot = h->save_epot.as<double>();
    // initial estimator of the population correlation: spin mapping's population (Estimators.jl:121), eCMM's
    // mapping kernel (X^2 + P^2) / 2 - gamma (cmm.jl:106-110)
    A.init_pop = h->final_di.as<double>();
    A.k0_b = ecmm ? ge : 0.5 * gs;
    const int block = 256;
    int bpc = 0;
    switch (n) {
    case 2: HIPCHK(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, mapping_kernel<2>, block, 0)); break;
    case 3: HIPCHK(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, mapping_kernel<3>, block, 0)); break;
    default: HIPCHK(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, mapping_kernel<4>, block, 0)); break;
    }
    if (bpc < 1) bpc = 1;
    const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((nt + block - 1) / block, (int64_t)bpc * h->cu_count));
    HIPCHK(h, hipEventRecord(h->ev0, h->stream));
    switch (n) {
    case 2: hipLaunchKernelGGL(mapping_kernel<2>, dim3((unsigned)grid), dim3(block), 0, h->stream, A); break;
    case 3: hipLaunchKernelGGL(mapping_kernel<3>, dim3((unsigned)grid), dim3(block), 0, h->stream, A); break;
    default: hipLaunchKernelGGL(mapping_kernel<4>, dim3((unsigned)grid), dim3(block), 0, h->stream, A); break;
    }
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipEventRecord(h->ev1, h->stream));
    h->stats.block = block;
    h->stats.grid = (int)grid;
    h->stats.launches = 1;
    return NQCD_OK;
}

int32_t map_download(nqcd_handle_t* h, int32_t field, void* out, int64_t bytes) {
    const int n = h->prob.n_states, B = map_beads(h);
    const int64_t nt = h->ntraj;
    const int ns = h->nsave;
    auto need = [&](int64_t b) -> int32_t {
        return b == bytes ? NQCD_OK : fail(h, NQCD_ERR_ARG, "bytes does not match the field size (" + std::to_string(b) + ")");
    };
    std::vector<double> tmp;
    auto fetch = [&](const void* src, size_t count) -> hipError_t {
        tmp.resize(count);
        return hipMemcpy(tmp.data(), src, sizeof(double) * count, hipMemcpyDeviceToHost);
    };
    double* o = static_cast<double*>(out);
    switch (field) {
    case NQCD_FIELD_R:
    case NQCD_FIELD_V: {   // (nbeads, ntraj)
        if (int32_t rc = need(8 * (int64_t)B * nt)) return rc;
        HIPCHK(h, fetch(field == NQCD_FIELD_R ? h->r.p : h->v.p, (size_t)B * nt));
        for (int64_t t = 0; t < nt; ++t)
            for (int b = 0; b < B; ++b) o[(size_t)t * B + b] = tmp[(size_t)b * nt + t];
        return NQCD_OK;
    }
    case NQCD_FIELD_SAVE_CENTROID: {
        if (!map_rp(h)) return fail(h, NQCD_ERR_UNSUPPORTED, "centroids exist for ring polymers");
        if (int32_t rc = need(8 * (int64_t)ns * nt)) return rc;
        HIPCHK(h, fetch(h->save_centroid.p, (size_t)ns * nt));
        for (int64_t t = 0; t < nt; ++t)
            for (int s = 0; s < ns; ++s) o[(size_t)t * ns + s] = tmp[(size_t)s * nt + t];
        return NQCD_OK;
    }
    case NQCD_FIELD_STEPS: {
        if (int32_t rc = need(4 * nt)) return rc;
        HIPCHK(h, hipMemcpy(out, h->steps.p, 4 * nt, hipMemcpyDeviceToHost));
        return NQCD_OK;
    }
    case NQCD_FIELD_MAP_X:
    case NQCD_FIELD_MAP_P: {   // (n, nbeads, ntraj)
        const int rows = n * B;
        if (int32_t rc = need(8 * (int64_t)rows * nt)) return rc;
        HIPCHK(h, fetch(h->sig.as<double>() + (field == NQCD_FIELD_MAP_P ? (size_t)rows * nt : 0), (size_t)rows * nt));
        for (int64_t t = 0; t < nt; ++t)
            for (int i = 0; i < rows; ++i) o[(size_t)t * rows + i] = tmp[(size_t)i * nt + t];
        return NQCD_OK;
    }
    case NQCD_FIELD_INIT_POP_DI: {
        if (int32_t rc = need(8 * (int64_t)n * nt)) return rc;
        HIPCHK(h, fetch(h->final_di.p, (size_t)n * nt));
        for (int64_t t = 0; t < nt; ++t)
            for (int i = 0; i < n; ++i) o[(size_t)t * n + i] = tmp[(size_t)i * nt + t];
        return NQCD_OK;
    }
    case NQCD_FIELD_SAVE_POP_DI: {
        if (int32_t rc = need(8 * (int64_t)n * ns * nt)) return rc;
        HIPCHK(h, fetch(h->save_di.p, (size_t)n * ns * nt));
        for (int64_t t = 0; t < nt; ++t)
            for (int s = 0; s < ns; ++s)
                for (int i = 0; i < n; ++i) o[((size_t)t * ns + s) * n + i] = tmp[((size_t)s * n + i) * nt + t];
        return NQCD_OK;
    }
    case NQCD_FIELD_SAVE_EKIN:
    case NQCD_FIELD_SAVE_EPOT: {
        if (int32_t rc = need(8 * (int64_t)ns * nt)) return rc;
        HIPCHK(h, fetch(field == NQCD_FIELD_SAVE_EKIN ? h->save_ekin.p : h->save_epot.p, (size_t)ns * nt));
        for (int64_t t = 0; t < nt; ++t)
            for (int s = 0; s < ns; ++s) o[(size_t)t * ns + s] = tmp[(size_t)s * nt + t];
        return NQCD_OK;
    }
    default:
        return fail(h, NQCD_ERR_UNSUPPORTED, "mapping fields: R, V, STEPS, MAP_X, MAP_P, INIT_POP_DI, SAVE_POP_DI, SAVE_EKIN, SAVE_EPOT (and SAVE_CENTROID for "
                                             "ring polymers)");
    }
}

}  // namespace nqcd
