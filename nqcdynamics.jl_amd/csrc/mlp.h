// mlp.h -- the MLP surrogate of a two-state diabatic Hamiltonian: 6 -> 128 -> 128 -> 3, tanh,
// BF16 MFMA (16x16x32, FP32 accumulation) for the 128-wide products, FP32 MFMA (16x16x4) for the
// first layer.
//
// Inputs: the six-number descriptor of a diatom configuration x = (x1, y1, z1, x2, y2, z2),
//   q = (b, z1, z2, cos(theta), X, Y),  b = |r1 - r2|,  cos(theta) = (z1 - z2) / b,
// (X, Y) the bond midpoint parallel to the surface: smooth wherever b > 0, and invariant under
// rotations of the bond about the surface normal (a flat, uncorrugated surface).  The network sees
// q normalised by the training data's mean and standard deviation; dH/dx is the chain rule through q.
//
// Register layout of a hidden layer inside a wavefront: the 64 lanes evaluate 16 samples at a
// time (a "sample tile"); lane l (group g = l / 16, column j = l % 16) holds units
// u = 16 t + 4 g + r (t = 0..7, r = 0..3) of sample j.  That is the accumulator layout of
// v_mfma_f32_16x16x32_bf16 with units along M and samples along N, and it is also, unchanged, the
// B operand of the next product once the K axis of each 32-unit step q is ordered as
//   k = 8 g + e  <->  unit 32 q + 4 g + e (e < 4),  32 q + 16 + 4 g + (e - 4) (e >= 4).
// The packed weights store W2 and W3 with their columns in that order, so hidden activations never
// leave registers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nqcd {

constexpr int kMlpH = 128;
constexpr int kMlpIn = 6;
constexpr int kMlpOut = 3;
constexpr int kMlpParams = kMlpH * kMlpIn + kMlpH + kMlpH * kMlpH + kMlpH + kMlpOut * kMlpH + kMlpOut;  // 17795
// FP32 master parameter vector: W1 (row-major 128 x 6), b1, W2 (128 x 128), b2, W3 (3 x 128), b3.
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kMlpH * kMlpIn, kOffW2 = kOffB1 + kMlpH;
constexpr int kOffB2 = kOffW2 + kMlpH * kMlpH, kOffW3 = kOffB2 + kMlpH, kOffB3 = kOffW3 + kMlpOut * kMlpH;

// Packed member (bytes), the layout staged into LDS by the kernels:
//   [0, 4096)      float w1b[128][8]: W1[u][0..5], b1[u], 0
//   [4096, 4608)   float b2[128]
//   [4608, 4672)   float b3[16] (3 used)
//   [4672, 37440)  bf16 W2p[128][128], columns in the permuted K order
//   [37440, 41536) bf16 W3p[16][128], rows 3..15 zero, columns permuted
//   [41536, 74304) bf16 W2Tp[128][128] = W2^T, columns in the permuted K order (input gradients)
//   [74304, 78400) bf16 W1Tp[16][128] = W1^T, rows 6..15 zero, columns permuted (input gradients)
//   [78400, 79936) float W3f[3][128], columns permuted (delta2 of the input gradient)
// In the four BF16 matrices the 16-byte chunks of each row are XOR-swizzled by the row index
// (mlp_swz): a fragment read has the 16 lanes of a lane group on rows 16 t + j, one chunk each, and
// unswizzled rows 256 bytes apart would put all 16 in the same LDS banks.
constexpr int kPkW1B = 0, kPkB2 = 4096, kPkB3 = 4608, kPkW2 = 4672, kPkW3 = 37440, kPkW2T = 41536, kPkW1T = 74304,
              kPkW3F = 78400, kPkBytes = 79936;

struct MlpNorm {
    float in_shift[kMlpIn], in_scale[kMlpIn];
    float out_shift[kMlpOut], out_scale[kMlpOut];
};

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <class T>
__host__ __device__ inline void mlp_descriptor(const T x[kMlpIn], T q[kMlpIn]) {
    const T dx = x[0] - x[3], dy = x[1] - x[4], dz = x[2] - x[5];
    const T b = sqrt(dx * dx + dy * dy + dz * dz);
    q[0] = b;
    q[1] = x[2];
    q[2] = x[5];
    q[3] = dz / b;
    q[4] = T(0.5) * (x[0] + x[3]);
    q[5] = T(0.5) * (x[1] + x[4]);
}

// dH/dx from g = dH/dq at x (the transpose of dq/dx applied to g).
template <class T>
__host__ __device__ inline void mlp_descriptor_pullback(const T x[kMlpIn], const T g[kMlpIn], T out[kMlpIn]) {
    const T d[3] = {x[0] - x[3], x[1] - x[4], x[2] - x[5]};
    const T ib = T(1) / sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const T c = d[2] * ib;
    T w[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) w[a] = (g[0] - g[3] * c * ib) * d[a] * ib;  // db/dr1 = d / b, dc/dr1 = (e_z - c d / b) / b
    w[2] += g[3] * ib;
    out[0] = w[0] + T(0.5) * g[4];
    out[1] = w[1] + T(0.5) * g[5];
    out[2] = w[2] + g[1];
    out[3] = -w[0] + T(0.5) * g[4];
    out[4] = -w[1] + T(0.5) * g[5];
    out[5] = -w[2] + g[2];
}

// Element (row, permuted column c) of a packed 128-column BF16 matrix, and the bf16x8 fragment
// holding chunk ch (8 columns) of a row.
__host__ __device__ constexpr int mlp_swz(int row, int c) { return row * 128 + ((((c >> 3) ^ (row & 15))) << 3) + (c & 7); }
__host__ __device__ constexpr int mlp_frag(int row, int ch) { return row * 16 + (ch ^ (row & 15)); }

// unit index of position k = 8 g + e inside the 32-unit K step q
__host__ __device__ constexpr int mlp_perm_unit(int q, int k) {
    return 32 * q + ((k & 7) < 4 ? 4 * (k >> 3) + (k & 7) : 16 + 4 * (k >> 3) + (k & 7) - 4);
}

// A zero the optimiser cannot see through: offsetting a pointer by it at a point makes the loads
// through that pointer issue there rather than be hoisted out of the surrounding loop (which would
// keep every weight fragment live), while the pointer keeps its LDS address space.
__device__ __forceinline__ int mlp_opaque_zero() {
    int z = 0;
    asm volatile("" : "+s"(z));
    return z;
}

// tanh(a) = 1 - 2 / (e^{2a} + 1) with the hardware exponential and reciprocal (about 1 ulp each,
// far below the BF16 rounding the activations get next); saturates to +-1 for large |a|.
__device__ __forceinline__ float mlp_tanh(float a) { return 1.0f - 2.0f * __builtin_amdgcn_rcpf(__expf(2.0f * a) + 1.0f); }

__device__ __forceinline__ f32x4 mfma_bf16(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ bf16x8 pack8(const float* lo, const float* hi) {
    bf16x8 v;
#pragma unroll
    for (int e = 0; e < 4; ++e) { v[e] = (__bf16)lo[e]; v[4 + e] = (__bf16)hi[e]; }
    return v;
}

// One member on one sample tile.  xs: this lane's sample (column j) normalised inputs.
// Returns, in lanes 0..15 (sample j), o[0..2]; with kGrad also G[o][i] = d o / d xs[i] in every
// lane of sample j (lanes j, j + 16, j + 32, j + 48).  The gradient is reverse mode: per output o,
// delta2 = W3[o] * s2, delta1 = (W2^T delta2) * s1 (s = 1 - h^2), G[o] = W1^T delta1, both products on
// BF16 MFMA with delta2, delta1 and W1 rounded to BF16; the three outputs share every W2^T
// fragment read.  Optional outputs for training: h1, s1 = 1 - h1^2,
// h2, s2 of this lane's units.
template <bool kGrad>
__device__ __forceinline__ void mlp_tile(const unsigned char* __restrict__ W, const float xs[kMlpIn], float o[kMlpOut],
                                         float G[kMlpOut][kMlpIn], float* h1_out = nullptr, float* s1_out = nullptr,
                                         float* h2_out = nullptr, float* s2_out = nullptr) {
    const int l = __lane_id(), g = l >> 4, j = l & 15;
    const float* w1b = reinterpret_cast<const float*>(W + kPkW1B);
    const float* b2 = reinterpret_cast<const float*>(W + kPkB2);
    const float* b3 = reinterpret_cast<const float*>(W + kPkB3);
    const bf16x8* W2p = reinterpret_cast<const bf16x8*>(W + kPkW2);
    const bf16x8* W3p = reinterpret_cast<const bf16x8*>(W + kPkW3);
    // layer 1 on FP32 MFMA (16x16x4): units along M, the augmented input (x0..x5, 1, 0) along K in
    // two steps of 4, samples along N; lane (g, j) supplies W1aug[16 t + j][4 ks + g] and input
    // 4 ks + g of sample j, and receives units 16 t + 4 g + r -- the layout of h1 below
    const float xb0 = (g == 0) ? xs[0] : (g == 1) ? xs[1] : (g == 2) ? xs[2] : xs[3];
    const float xb1 = (g == 0) ? xs[4] : (g == 1) ? xs[5] : (g == 2) ? 1.0f : 0.0f;
    float h1[32];
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        f32x4 a = f32x4{0.f, 0.f, 0.f, 0.f};
        a = __builtin_amdgcn_mfma_f32_16x16x4f32(w1b[8 * (16 * t + j) + g], xb0, a, 0, 0, 0);
        a = __builtin_amdgcn_mfma_f32_16x16x4f32(w1b[8 * (16 * t + j) + 4 + g], xb1, a, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) h1[4 * t + r] = mlp_tanh(a[r]);
    }
    // layer 2 and the output layer; with kGrad, delta2 of each output in the B layout, built as
    // each K step's units come out: lane (g, j), K step q, e < 4: unit 32 q + 4 g + e, e >= 4:
    // unit 32 q + 16 + 4 g + e - 4, the W3p row's permuted columns 32 q + 8 g + e
    bf16x8 D[kMlpOut][4];
    {
        bf16x8 B[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) B[q] = pack8(&h1[4 * (2 * q)], &h1[4 * (2 * q + 1)]);
        f32x4 acc3 = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int q3 = 0; q3 < 4; ++q3) {
            float val[2][4], sv[2][4];
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                const int ot = 2 * q3 + half;
                f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kq = 0; kq < 4; ++kq) acc = mfma_bf16(W2p[mlp_frag(16 * ot + j, 4 * kq + g)], B[kq], acc);
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int idx = 4 * ot + r;
                    const float h = mlp_tanh(acc[r] + b2[16 * ot + 4 * g + r]);
                    sv[half][r] = 1.0f - h * h;
                    val[half][r] = h;
                    if (h2_out) { h2_out[idx] = h; s2_out[idx] = sv[half][r]; }
                }
            }
            const bf16x8 w3o = W3p[mlp_frag(j, 4 * q3 + g)];
            acc3 = mfma_bf16(w3o, pack8(val[0], val[1]), acc3);
            if constexpr (kGrad) {
#pragma unroll
                for (int oo = 0; oo < kMlpOut; ++oo) {
                    const float* w3f = reinterpret_cast<const float*>(W + kPkW3F) + oo * kMlpH + 32 * q3 + 8 * g;
                    const float4 wa = *reinterpret_cast<const float4*>(w3f);
                    const float4 wb = *reinterpret_cast<const float4*>(w3f + 4);
                    const float lo[4] = {wa.x * sv[0][0], wa.y * sv[0][1], wa.z * sv[0][2], wa.w * sv[0][3]};
                    const float hi[4] = {wb.x * sv[1][0], wb.y * sv[1][1], wb.z * sv[1][2], wb.w * sv[1][3]};
                    D[oo][q3] = pack8(lo, hi);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < kMlpOut; ++r) o[r] = acc3[r] + b3[r];
    }
    if (h1_out) {
#pragma unroll
        for (int k = 0; k < 32; ++k) { h1_out[k] = h1[k]; s1_out[k] = 1.0f - h1[k] * h1[k]; }
    }
    if constexpr (kGrad) {
        const bf16x8* W2Tp = reinterpret_cast<const bf16x8*>(W + kPkW2T);
        const bf16x8* W1Tp = reinterpret_cast<const bf16x8*>(W + kPkW1T);
        // K step q covers units 32 q .. 32 q + 31 = the two 16-unit tiles 2 q, 2 q + 1: W2^T delta2 on
        // those tiles (the three outputs share each W2^T fragment), times s1, packed to BF16 is the
        // B operand of G += W1^T delta1 for that K step.  The q loop is unrolled with the weight
        // pointers laundered per step, so the loads stay inside their step.
        f32x4 gacc[kMlpOut];
#pragma unroll
        for (int oo = 0; oo < kMlpOut; ++oo) gacc[oo] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const bf16x8* W2Tq = W2Tp + mlp_opaque_zero();
            float v[kMlpOut][2][4];
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                const int ut = 2 * q + half;
                f32x4 acc[kMlpOut];
#pragma unroll
                for (int oo = 0; oo < kMlpOut; ++oo) acc[oo] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int kq = 0; kq < 4; ++kq) {
                    const bf16x8 A = W2Tq[mlp_frag(16 * ut + j, 4 * kq + g)];
#pragma unroll
                    for (int oo = 0; oo < kMlpOut; ++oo) acc[oo] = mfma_bf16(A, D[oo][kq], acc[oo]);
                }
                // acc[oo][r] = (W2^T delta2)[u] for u = 16 ut + 4 g + r, sample j
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float s1 = 1.0f - h1[4 * ut + r] * h1[4 * ut + r];
#pragma unroll
                    for (int oo = 0; oo < kMlpOut; ++oo) v[oo][half][r] = acc[oo][r] * s1;
                }
            }
            const bf16x8 A1 = (W1Tp + mlp_opaque_zero())[mlp_frag(j, 4 * q + g)];
#pragma unroll
            for (int oo = 0; oo < kMlpOut; ++oo) gacc[oo] = mfma_bf16(A1, pack8(v[oo][0], v[oo][1]), gacc[oo]);
        }
        // gacc[oo][r] in lane (g, j) is d o / d xs[4 g + r] of sample j: gather inputs 0..5 into
        // every lane of sample j
#pragma unroll
        for (int oo = 0; oo < kMlpOut; ++oo) {
#pragma unroll
            for (int i = 0; i < kMlpIn; ++i) G[oo][i] = __shfl(gacc[oo][i & 3], j + 16 * (i >> 2));
        }
    }
}

}  // namespace nqcd
