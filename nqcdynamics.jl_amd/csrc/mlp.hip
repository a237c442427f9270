// mlp.hip -- training and batched evaluation of the MLP surrogate (mlp.h) on the GPU.
//
// Training: Adam on the mean weighted squared error of the normalised (H11, H22, H12) targets.
// Each step: mlp_train_grad (one wavefront per 32 samples: forward through mlp_tile, backward
// with BF16 MFMA for delta1 = W2^T delta2 and dW2 = delta2 h1^T, FP32 elsewhere; per-group partial
// gradients), mlp_reduce_adam (gradient = fixed-order sum of the partials, Adam on the FP32
// master parameters), mlp_pack (BF16 packed copy for the next forward).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "handle.h"
#include "mlp.h"
#include "mlp_host.h"

namespace nqcd {
namespace {

constexpr int kTrainSamples = 32;  // samples per training workgroup (two sample tiles)
constexpr int kTrainThreads = 128; // one wavefront per sample tile
static_assert(kTrainThreads / 64 == kTrainSamples / 16, "one wavefront per 16-sample tile");
constexpr int kPackBlocks = 64;

// Packs member blockIdx.y: P and out advance by one member per blockIdx.y.
__global__ __launch_bounds__(256) void mlp_pack(const float* __restrict__ P, unsigned char* __restrict__ out) {
    P += (size_t)blockIdx.y * kMlpParams;
    out += (size_t)blockIdx.y * kPkBytes;
    float* w1b = reinterpret_cast<float*>(out + kPkW1B);
    float* b2 = reinterpret_cast<float*>(out + kPkB2);
    float* b3 = reinterpret_cast<float*>(out + kPkB3);
    __bf16* W2p = reinterpret_cast<__bf16*>(out + kPkW2);
    __bf16* W3p = reinterpret_cast<__bf16*>(out + kPkW3);
    const int tid = blockIdx.x * blockDim.x + threadIdx.x, B = gridDim.x * blockDim.x;
    for (int i = tid; i < kMlpH * 8; i += B) {
        const int u = i >> 3, c = i & 7;
        w1b[i] = (c < kMlpIn) ? P[kOffW1 + u * kMlpIn + c] : (c == kMlpIn ? P[kOffB1 + u] : 0.0f);
    }
    for (int u = tid; u < kMlpH; u += B) b2[u] = P[kOffB2 + u];
    for (int o = tid; o < 16; o += B) b3[o] = (o < kMlpOut) ? P[kOffB3 + o] : 0.0f;
    for (int i = tid; i < kMlpH * kMlpH; i += B) {
        const int row = i >> 7, c = i & 127;
        W2p[mlp_swz(row, c)] = (__bf16)P[kOffW2 + row * kMlpH + mlp_perm_unit(c >> 5, c & 31)];
    }
    for (int i = tid; i < 16 * kMlpH; i += B) {
        const int row = i >> 7, c = i & 127;
        W3p[mlp_swz(row, c)] = (row < kMlpOut) ? (__bf16)P[kOffW3 + row * kMlpH + mlp_perm_unit(c >> 5, c & 31)] : (__bf16)0.0f;
    }
    __bf16* W2Tp = reinterpret_cast<__bf16*>(out + kPkW2T);
    for (int i = tid; i < kMlpH * kMlpH; i += B) {
        const int u = i >> 7, c = i & 127;
        W2Tp[mlp_swz(u, c)] = (__bf16)P[kOffW2 + mlp_perm_unit(c >> 5, c & 31) * kMlpH + u];
    }
    float* W3f = reinterpret_cast<float*>(out + kPkW3F);
    for (int i = tid; i < kMlpOut * kMlpH; i += B) {
        const int o = i >> 7, c = i & 127;
        W3f[i] = (float)(__bf16)P[kOffW3 + o * kMlpH + mlp_perm_unit(c >> 5, c & 31)];  // the BF16 value, held as FP32
    }
    __bf16* W1Tp = reinterpret_cast<__bf16*>(out + kPkW1T);
    for (int i = tid; i < 16 * kMlpH; i += B) {
        const int in = i >> 7, c = i & 127;
        W1Tp[mlp_swz(in, c)] = (in < kMlpIn) ? (__bf16)P[kOffW1 + mlp_perm_unit(c >> 5, c & 31) * kMlpIn + in] : (__bf16)0.0f;
    }
}

__device__ __forceinline__ void stage_member(const unsigned char* __restrict__ src, unsigned char* dst) {
    const uint4* s = reinterpret_cast<const uint4*>(src);
    uint4* d = reinterpret_cast<uint4*>(dst);
    for (int i = threadIdx.x; i < kPkBytes / 16; i += blockDim.x) d[i] = s[i];
}

// Batched evaluation: H (n, 3) and dH/dx (n, 3, 6) of one member at raw coordinates X (n, 6).
__global__ __launch_bounds__(256) void mlp_eval_kernel(const unsigned char* __restrict__ packed, MlpNorm nrm, int64_t n,
                                                       const float* __restrict__ X, float* __restrict__ Hout,
                                                       float* __restrict__ dH) {
    extern __shared__ __align__(16) unsigned char lds[];
    stage_member(packed, lds);
    __syncthreads();
    const int l = __lane_id(), j = l & 15, wave = threadIdx.x >> 6;
    const int64_t base = ((int64_t)blockIdx.x * (blockDim.x >> 6) + wave) * 64;
    for (int T = 0; T < 4; ++T) {
        const int64_t s = base + 16 * T + j;
        float x[kMlpIn], xs[kMlpIn];
#pragma unroll
        for (int i = 0; i < kMlpIn; ++i) x[i] = (s < n) ? X[s * kMlpIn + i] : (i == 0 ? 1.0f : 0.0f);
        mlp_descriptor(x, xs);
#pragma unroll
        for (int i = 0; i < kMlpIn; ++i) xs[i] = (xs[i] - nrm.in_shift[i]) * nrm.in_scale[i];
        float o[kMlpOut], J[kMlpOut][kMlpIn];
        mlp_tile<true>(lds, xs, o, J);
        if (l < 16 && s < n) {
#pragma unroll
            for (int r = 0; r < kMlpOut; ++r) {
                Hout[s * kMlpOut + r] = o[r] * nrm.out_scale[r] + nrm.out_shift[r];
                float gq[kMlpIn], gx[kMlpIn];
#pragma unroll
                for (int i = 0; i < kMlpIn; ++i) gq[i] = J[r][i] * nrm.out_scale[r] * nrm.in_scale[i];
                mlp_descriptor_pullback(x, gq, gx);
#pragma unroll
                for (int i = 0; i < kMlpIn; ++i) dH[(s * kMlpOut + r) * kMlpIn + i] = gx[i];
            }
        }
    }
}

// One workgroup of two wavefronts per 32 samples of the minibatch: partial gradient (kMlpParams
// floats, master order) and partial loss of those samples.  Each wavefront takes one 16-sample
// tile through the network, then the two split the parameter sums over the 32 samples.
struct TrainLds {
    unsigned char member[kPkBytes];        // the packed member, W2^T included
    float W3[kMlpOut * kMlpH];
    __bf16 h1q[kMlpH * kTrainSamples];     // [unit][sample]
    float d2[kMlpH * kTrainSamples];       // [unit][sample]
    float h2q[kMlpH * kTrainSamples];
    float d1[kMlpH * kTrainSamples];
    float x[kTrainSamples * kMlpIn];
    float d3[kTrainSamples * kMlpOut];
    float loss[kTrainSamples];
};

// Member blockIdx.y (the members train side by side in one launch): its packed copy, master
// parameters, minibatch indices (idx + member * idx_stride), partial gradients and losses.
__global__ __launch_bounds__(128) void mlp_train_grad(const unsigned char* __restrict__ packed, const float* __restrict__ P,
                                                     const float* __restrict__ Xn, const float* __restrict__ Tn,
                                                     const int32_t* __restrict__ idx, int64_t idx_stride, int32_t nb,
                                                     float w0, float w1, float w2, float* __restrict__ partial,
                                                     float* __restrict__ loss_part) {
    const int mem = blockIdx.y;
    packed += (size_t)mem * kPkBytes;
    P += (size_t)mem * kMlpParams;
    idx += mem * idx_stride;
    partial += (size_t)mem * gridDim.x * kMlpParams;
    loss_part += (size_t)mem * gridDim.x;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    TrainLds& S = *reinterpret_cast<TrainLds*>(lds_raw);
    const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63, g = l >> 4, j = l & 15;
    stage_member(packed, S.member);
    for (int i = tid; i < kMlpOut * kMlpH; i += kTrainThreads) S.W3[i] = P[kOffW3 + i];
    __syncthreads();
    const float wl[3] = {w0, w1, w2};
    const int g0 = blockIdx.x * kTrainSamples;
    {
        const int T = wv;  // this wavefront's sample tile
        const int ls = 16 * T + j;                  // local sample of this lane's column
        const bool valid = g0 + ls < nb;
        const int s = valid ? idx[g0 + ls] : 0;
        float xs[kMlpIn];
#pragma unroll
        for (int i = 0; i < kMlpIn; ++i) xs[i] = valid ? Xn[(int64_t)s * kMlpIn + i] : 0.0f;
        float o[kMlpOut], J[kMlpOut][kMlpIn], h1[32], s1[32], h2[32], s2[32];
        mlp_tile<false>(S.member, xs, o, J, h1, s1, h2, s2);
        // output residuals (sample j lives in lane j of group 0)
        float d3[kMlpOut];
        float lo = 0.0f;
#pragma unroll
        for (int r = 0; r < kMlpOut; ++r) {
            const float t = valid ? Tn[(int64_t)s * kMlpOut + r] : 0.0f;
            const float res = valid ? o[r] - t : 0.0f;
            lo += wl[r] * res * res;
            d3[r] = 2.0f * wl[r] * res / (float)nb;
        }
#pragma unroll
        for (int r = 0; r < kMlpOut; ++r) d3[r] = __shfl(d3[r], j);  // sample j's residual, from lane j
        if (g == 0) {
            S.loss[ls] = lo / (float)nb;
#pragma unroll
            for (int r = 0; r < kMlpOut; ++r) S.d3[ls * kMlpOut + r] = d3[r];
#pragma unroll
            for (int i = 0; i < kMlpIn; ++i) S.x[ls * kMlpIn + i] = xs[i];
        }
        // delta2 = s2 * (W3^T delta3); stage h1q, h2q, delta2 as [unit][sample]
        float d2[32];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int u = 16 * t + 4 * g + r, k = 4 * t + r;
                float a = 0.0f;
#pragma unroll
                for (int o3 = 0; o3 < kMlpOut; ++o3) a += S.W3[o3 * kMlpH + u] * d3[o3];
                d2[k] = s2[k] * a;
                S.d2[u * kTrainSamples + ls] = d2[k];
                S.h1q[u * kTrainSamples + ls] = (__bf16)h1[k];
                S.h2q[u * kTrainSamples + ls] = (float)(__bf16)h2[k];
            }
        }
        // delta1 = s1 * (W2^T delta2) on MFMA (units of W2's input along M, samples along N)
        bf16x8 Bd[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) Bd[q] = pack8(&d2[4 * (2 * q)], &d2[4 * (2 * q + 1)]);
        const bf16x8* W2Tp = reinterpret_cast<const bf16x8*>(S.member + kPkW2T);
#pragma unroll
        for (int ut = 0; ut < 8; ++ut) {
            f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int q = 0; q < 4; ++q) acc = mfma_bf16(W2Tp[mlp_frag(16 * ut + j, 4 * q + g)], Bd[q], acc);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int u = 16 * ut + 4 * g + r;
                S.d1[u * kTrainSamples + ls] = s1[4 * ut + r] * acc[r];
            }
        }
    }
    __syncthreads();
    float* out = partial + (int64_t)blockIdx.x * kMlpParams;
    // dW2 = delta2 h1q^T on MFMA, K = the 32 samples: A = delta2 [v][s], B = h1q [u][s]
#pragma unroll 1
    for (int vt = wv; vt < 8; vt += kTrainThreads / 64) {
        float av[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) av[e] = S.d2[(16 * vt + j) * kTrainSamples + 8 * g + e];
        const bf16x8 A = pack8(av, av + 4);
#pragma unroll 1
        for (int ut = 0; ut < 8; ++ut) {
            const bf16x8 Bh = *reinterpret_cast<const bf16x8*>(&S.h1q[(16 * ut + j) * kTrainSamples + 8 * g]);
            const f32x4 acc = mfma_bf16(A, Bh, f32x4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
            for (int r = 0; r < 4; ++r) out[kOffW2 + (16 * vt + 4 * g + r) * kMlpH + 16 * ut + j] = acc[r];
        }
    }
    // FP32 sums over the samples for the small parameter blocks
    for (int p = tid; p < kMlpOut * kMlpH; p += kTrainThreads) {
        const int o3 = p / kMlpH, u = p % kMlpH;
        float a = 0.0f;
        for (int s = 0; s < kTrainSamples; ++s) a += S.d3[s * kMlpOut + o3] * S.h2q[u * kTrainSamples + s];
        out[kOffW3 + p] = a;
    }
    for (int p = tid; p < kMlpH * kMlpIn; p += kTrainThreads) {
        const int u = p / kMlpIn, i = p % kMlpIn;
        float a = 0.0f;
        for (int s = 0; s < kTrainSamples; ++s) a += S.d1[u * kTrainSamples + s] * S.x[s * kMlpIn + i];
        out[kOffW1 + p] = a;
    }
    for (int u = tid; u < kMlpH; u += kTrainThreads) {
        float a1 = 0.0f, a2 = 0.0f;
        for (int s = 0; s < kTrainSamples; ++s) { a1 += S.d1[u * kTrainSamples + s]; a2 += S.d2[u * kTrainSamples + s]; }
        out[kOffB1 + u] = a1;
        out[kOffB2 + u] = a2;
    }
    if (tid < kMlpOut) {
        float a = 0.0f;
        for (int s = 0; s < kTrainSamples; ++s) a += S.d3[s * kMlpOut + tid];
        out[kOffB3 + tid] = a;
    }
    if (tid == 0) {
        float a = 0.0f;
        for (int s = 0; s < kTrainSamples; ++s) a += S.loss[s];
        loss_part[blockIdx.x] = a;
    }
}

// x[0] + x[stride] + ... + x[(n - 1) stride], added in that order; the loads of 16 terms are
// issued together, so a sum costs n / 16 memory latencies instead of n.
__device__ __forceinline__ float ordered_sum(const float* __restrict__ x, int64_t stride, int n) {
    float a = 0.0f;
    int k = 0;
    for (; k + 16 <= n; k += 16) {
        float t[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) t[q] = x[(k + q) * stride];
#pragma unroll
        for (int q = 0; q < 16; ++q) a += t[q];
    }
    for (; k < n; ++k) a += x[k * stride];
    return a;
}

// Member blockIdx.y: fixed-order sum of its ngroups partial gradients, then Adam.
__global__ __launch_bounds__(256) void mlp_reduce_adam(const float* __restrict__ partial, int32_t ngroups, float* P,
                                                       float* m, float* v, float lr, float b1, float b2, float eps,
                                                       float b1t, float b2t, int32_t apply, float* grad_out,
                                                       const float* __restrict__ loss_part, float* loss_out,
                                                       int64_t loss_stride) {
    const int mem = blockIdx.y;
    partial += (size_t)mem * ngroups * kMlpParams;
    P += (size_t)mem * kMlpParams;
    m += (size_t)mem * kMlpParams;
    v += (size_t)mem * kMlpParams;
    if (grad_out) grad_out += (size_t)mem * kMlpParams;
    loss_part += (size_t)mem * ngroups;
    loss_out += mem * loss_stride;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p == 0) *loss_out = ordered_sum(loss_part, 1, ngroups);
    if (p >= kMlpParams) return;
    const float gsum = ordered_sum(partial + p, kMlpParams, ngroups);
    if (grad_out) grad_out[p] = gsum;
    if (!apply) return;
    const float mm = b1 * m[p] + (1.0f - b1) * gsum;
    const float vv = b2 * v[p] + (1.0f - b2) * gsum * gsum;
    m[p] = mm;
    v[p] = vv;
    P[p] -= lr * (mm / (1.0f - b1t)) / (sqrtf(vv / (1.0f - b2t)) + eps);
}

size_t train_lds_bytes() { return sizeof(TrainLds); }

}  // namespace

// ---------------------------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------------------------
int32_t mlp_ensure(nqcd_handle_t* h) {
    HIPCHK(h, h->mlp_master.ensure(sizeof(float) * kMlpParams * kMlpMembers));
    HIPCHK(h, h->mlp_packed.ensure((size_t)kPkBytes * kMlpMembers));
    return NQCD_OK;
}

int32_t mlp_pack_all(nqcd_handle_t* h) {
    hipLaunchKernelGGL(mlp_pack, dim3(kPackBlocks, kMlpMembers), dim3(256), 0, h->stream, h->mlp_master.as<float>(),
                       h->mlp_packed.as<unsigned char>());
    HIPCHK(h, hipGetLastError());
    return NQCD_OK;
}

static void init_member(std::vector<float>& P, uint64_t seed) {
    std::mt19937_64 rng(seed);
    std::normal_distribution<float> nd(0.0f, 1.0f);
    P.assign(kMlpParams, 0.0f);
    for (int i = 0; i < kMlpH * kMlpIn; ++i) P[kOffW1 + i] = nd(rng) / std::sqrt((float)kMlpIn);
    for (int i = 0; i < kMlpH * kMlpH; ++i) P[kOffW2 + i] = nd(rng) / std::sqrt((float)kMlpH);
    for (int i = 0; i < kMlpOut * kMlpH; ++i) P[kOffW3 + i] = nd(rng) / std::sqrt((float)kMlpH);
}

// Descriptors of the configurations X (n, 6), in double precision.
static std::vector<double> descriptors(const double* X, int64_t n) {
    std::vector<double> q((size_t)n * kMlpIn);
    for (int64_t k = 0; k < n; ++k) mlp_descriptor(X + k * kMlpIn, q.data() + k * kMlpIn);
    return q;
}

static bool bonds_positive(const double* X, int64_t n) {
    for (int64_t k = 0; k < n; ++k) {
        const double* x = X + k * kMlpIn;
        const double d2 = (x[0] - x[3]) * (x[0] - x[3]) + (x[1] - x[4]) * (x[1] - x[4]) + (x[2] - x[5]) * (x[2] - x[5]);
        if (!(d2 > 0.0)) return false;
    }
    return true;
}

static void set_norm_from_data(nqcd_handle_t* h, const std::vector<double>& Q, const double* Y, int64_t n) {
    MlpNorm& N = h->mlp_norm;
    for (int i = 0; i < kMlpIn; ++i) {
        double s = 0.0, s2 = 0.0;
        for (int64_t k = 0; k < n; ++k) { const double x = Q[k * kMlpIn + i]; s += x; s2 += x * x; }
        const double mu = s / n, var = std::max(s2 / n - mu * mu, 1e-12);
        N.in_shift[i] = (float)mu;
        N.in_scale[i] = (float)(1.0 / std::sqrt(var));
    }
    for (int r = 0; r < kMlpOut; ++r) {
        double s = 0.0, s2 = 0.0;
        for (int64_t k = 0; k < n; ++k) { const double y = Y[k * kMlpOut + r]; s += y; s2 += y * y; }
        const double mu = s / n, var = std::max(s2 / n - mu * mu, 1e-16);
        N.out_shift[r] = (float)mu;
        N.out_scale[r] = (float)std::sqrt(var);
    }
}

// Normalised descriptors and targets on the device.
static int32_t upload_data(nqcd_handle_t* h, const std::vector<double>& Q, const double* Y, int64_t n) {
    const MlpNorm& N = h->mlp_norm;
    std::vector<float> xn((size_t)n * kMlpIn), tn((size_t)n * kMlpOut);
    for (int64_t k = 0; k < n; ++k) {
        for (int i = 0; i < kMlpIn; ++i) xn[k * kMlpIn + i] = (float)((Q[k * kMlpIn + i] - N.in_shift[i]) * N.in_scale[i]);
        for (int r = 0; r < kMlpOut; ++r) tn[k * kMlpOut + r] = (float)((Y[k * kMlpOut + r] - N.out_shift[r]) / N.out_scale[r]);
    }
    HIPCHK(h, h->mlp_x.ensure(sizeof(float) * xn.size()));
    HIPCHK(h, h->mlp_t.ensure(sizeof(float) * tn.size()));
    HIPCHK(h, hipMemcpy(h->mlp_x.p, xn.data(), sizeof(float) * xn.size(), hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(h->mlp_t.p, tn.data(), sizeof(float) * tn.size(), hipMemcpyHostToDevice));
    return NQCD_OK;
}

// One gradient evaluation (and, when apply, one Adam step) of members k0 .. k0 + nmem - 1 side by
// side: member k0 + i takes its samples at idx_dev + i * idx_stride and writes its loss to
// loss_dev + i * loss_stride.
static int32_t train_step(nqcd_handle_t* h, int k0, int nmem, const int32_t* idx_dev, int64_t idx_stride, int32_t nb,
                          const float w[3], bool apply, int64_t t, float lr, float* grad_dev, float* loss_dev,
                          int64_t loss_stride) {
    const int ngroups = (nb + kTrainSamples - 1) / kTrainSamples;
    HIPCHK(h, h->mlp_partial.ensure(sizeof(float) * (size_t)nmem * ngroups * kMlpParams));
    HIPCHK(h, h->mlp_losspart.ensure(sizeof(float) * (size_t)nmem * ngroups));
    float* P = h->mlp_master.as<float>() + (size_t)k0 * kMlpParams;
    unsigned char* packed = h->mlp_packed.as<unsigned char>() + (size_t)k0 * kPkBytes;
    const size_t lds = train_lds_bytes();
    HIPCHK(h, hipFuncSetAttribute((const void*)mlp_train_grad, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(mlp_train_grad, dim3(ngroups, nmem), dim3(kTrainThreads), lds, h->stream, packed, P, h->mlp_x.as<float>(),
                       h->mlp_t.as<float>(), idx_dev, idx_stride, nb, w[0], w[1], w[2], h->mlp_partial.as<float>(),
                       h->mlp_losspart.as<float>());
    HIPCHK(h, hipGetLastError());
    const float b1 = 0.9f, b2 = 0.999f;
    hipLaunchKernelGGL(mlp_reduce_adam, dim3((kMlpParams + 255) / 256, nmem), dim3(256), 0, h->stream,
                       h->mlp_partial.as<float>(), ngroups, P, h->mlp_adam_m.as<float>() + (size_t)k0 * kMlpParams,
                       h->mlp_adam_v.as<float>() + (size_t)k0 * kMlpParams, lr, b1, b2, 1e-8f,
                       (float)std::pow(b1, (double)t), (float)std::pow(b2, (double)t), apply ? 1 : 0, grad_dev,
                       h->mlp_losspart.as<float>(), loss_dev, loss_stride);
    HIPCHK(h, hipGetLastError());
    if (apply) {
        hipLaunchKernelGGL(mlp_pack, dim3(kPackBlocks, nmem), dim3(256), 0, h->stream, P, packed);
        HIPCHK(h, hipGetLastError());
    }
    return NQCD_OK;
}

int32_t mlp_train(nqcd_handle_t* h, const double* X, const double* Y, int64_t n, const nqcd_train_desc* d,
                  double* loss_history) {
    if (n < kTrainSamples) return fail(h, NQCD_ERR_ARG, "training needs at least 32 samples");
    if (n > (int64_t)1 << 30) return fail(h, NQCD_ERR_ARG, "too many training samples");
    if (d->steps < 1 || d->batch < kTrainSamples || d->batch > n) return fail(h, NQCD_ERR_ARG, "steps >= 1, 32 <= batch <= n");
    if (!(d->lr > 0.0) || !(d->lr_final > 0.0)) return fail(h, NQCD_ERR_ARG, "learning rates must be positive");
    if (!bonds_positive(X, n)) return fail(h, NQCD_ERR_ARG, "a configuration has coincident atoms");
    HIPCHK(h, hipSetDevice(h->device));
    if (int32_t rc = mlp_ensure(h)) return rc;
    const std::vector<double> Q = descriptors(X, n);
    set_norm_from_data(h, Q, Y, n);
    if (int32_t rc = upload_data(h, Q, Y, n)) return rc;
    HIPCHK(h, h->mlp_adam_m.ensure(sizeof(float) * kMlpParams * kMlpMembers));
    HIPCHK(h, h->mlp_adam_v.ensure(sizeof(float) * kMlpParams * kMlpMembers));
    HIPCHK(h, hipMemsetAsync(h->mlp_adam_m.p, 0, sizeof(float) * kMlpParams * kMlpMembers, h->stream));
    HIPCHK(h, hipMemsetAsync(h->mlp_adam_v.p, 0, sizeof(float) * kMlpParams * kMlpMembers, h->stream));
    for (int k = 0; k < kMlpMembers; ++k) {
        std::vector<float> P;
        init_member(P, d->seed + 0x9E3779B97F4A7C15ull * (uint64_t)(k + 1));
        HIPCHK(h, hipMemcpy(h->mlp_master.as<float>() + (size_t)k * kMlpParams, P.data(), sizeof(float) * kMlpParams,
                            hipMemcpyHostToDevice));
    }
    if (int32_t rc = mlp_pack_all(h)) return rc;
    // one shuffled pass over the data per epoch, per member (its own order); the members step side by side
    HIPCHK(h, h->mlp_idx.ensure(sizeof(int32_t) * n * kMlpMembers));
    HIPCHK(h, h->mlp_lossbuf.ensure(sizeof(float) * d->steps * kMlpMembers));
    const float w[3] = {(float)d->loss_weights[0], (float)d->loss_weights[1], (float)d->loss_weights[2]};
    std::vector<std::mt19937_64> rngs;
    for (int k = 0; k < kMlpMembers; ++k) rngs.emplace_back(d->seed ^ (0xA5A5A5A5ull + (uint64_t)k));
    std::vector<int32_t> perm((size_t)n * kMlpMembers);
    int64_t pos = n;
    for (int64_t t = 1; t <= d->steps; ++t) {
        if (pos + d->batch > n) {
            for (int k = 0; k < kMlpMembers; ++k) {
                int32_t* pk = perm.data() + (size_t)k * n;
                for (int64_t i = 0; i < n; ++i) pk[i] = (int32_t)i;
                std::shuffle(pk, pk + n, rngs[k]);
            }
            HIPCHK(h, hipMemcpyAsync(h->mlp_idx.p, perm.data(), sizeof(int32_t) * perm.size(), hipMemcpyHostToDevice, h->stream));
            HIPCHK(h, hipStreamSynchronize(h->stream));
            pos = 0;
        }
        const double frac = (double)(t - 1) / std::max<int64_t>(1, d->steps - 1);
        const float lr = (float)(d->lr * std::pow(d->lr_final / d->lr, frac));
        if (int32_t rc = train_step(h, 0, kMlpMembers, h->mlp_idx.as<int32_t>() + pos, n, d->batch, w, true, t, lr, nullptr,
                                    h->mlp_lossbuf.as<float>() + (t - 1), d->steps))
            return rc;
        pos += d->batch;
    }
    std::vector<float> lh((size_t)d->steps * kMlpMembers);
    HIPCHK(h, hipMemcpy(lh.data(), h->mlp_lossbuf.p, sizeof(float) * lh.size(), hipMemcpyDeviceToHost));
    if (loss_history)
        for (size_t i = 0; i < lh.size(); ++i) loss_history[i] = lh[i];
    h->mlp_ready = true;
    return NQCD_OK;
}

int32_t mlp_gradient(nqcd_handle_t* h, int32_t member, const double* X, const double* Y, int64_t n, const double w[3],
                     double* grad, double* loss) {
    if (!h->mlp_ready) return fail(h, NQCD_ERR_STATE, "no surrogate: train or load it first");
    if (member < 0 || member >= kMlpMembers) return fail(h, NQCD_ERR_ARG, "member out of range");
    if (n < 1 || n > (1 << 24)) return fail(h, NQCD_ERR_ARG, "n out of range");
    if (!bonds_positive(X, n)) return fail(h, NQCD_ERR_ARG, "a configuration has coincident atoms");
    HIPCHK(h, hipSetDevice(h->device));
    if (int32_t rc = upload_data(h, descriptors(X, n), Y, n)) return rc;
    std::vector<int32_t> idx(n);
    for (int64_t i = 0; i < n; ++i) idx[i] = (int32_t)i;
    HIPCHK(h, h->mlp_idx.ensure(sizeof(int32_t) * n));
    HIPCHK(h, hipMemcpy(h->mlp_idx.p, idx.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    HIPCHK(h, h->mlp_grad.ensure(sizeof(float) * (kMlpParams + 1)));
    const float wf[3] = {(float)w[0], (float)w[1], (float)w[2]};
    if (int32_t rc = train_step(h, member, 1, h->mlp_idx.as<int32_t>(), 0, (int32_t)n, wf, false, 1, 0.0f,
                                h->mlp_grad.as<float>(), h->mlp_grad.as<float>() + kMlpParams, 0))
        return rc;
    std::vector<float> gv(kMlpParams + 1);
    HIPCHK(h, hipMemcpy(gv.data(), h->mlp_grad.p, sizeof(float) * gv.size(), hipMemcpyDeviceToHost));
    for (int p = 0; p < kMlpParams; ++p) grad[p] = gv[p];
    *loss = gv[kMlpParams];
    return NQCD_OK;
}

int32_t mlp_eval(nqcd_handle_t* h, int32_t member, const double* X, int64_t n, double* Hout, double* dH) {
    if (!h->mlp_ready) return fail(h, NQCD_ERR_STATE, "no surrogate: train or load it first");
    if (member < 0 || member >= kMlpMembers) return fail(h, NQCD_ERR_ARG, "member out of range");
    if (n < 1 || n > (1 << 26)) return fail(h, NQCD_ERR_ARG, "n out of range");
    if (!bonds_positive(X, n)) return fail(h, NQCD_ERR_ARG, "a configuration has coincident atoms");
    HIPCHK(h, hipSetDevice(h->device));
    std::vector<float> xf((size_t)n * kMlpIn);
    for (int64_t i = 0; i < n * kMlpIn; ++i) xf[i] = (float)X[i];
    DevBuf xb, hb, db;
    HIPCHK(h, xb.ensure(sizeof(float) * xf.size()));
    HIPCHK(h, hb.ensure(sizeof(float) * n * kMlpOut));
    HIPCHK(h, db.ensure(sizeof(float) * n * kMlpOut * kMlpIn));
    hipError_t e = hipMemcpy(xb.p, xf.data(), sizeof(float) * xf.size(), hipMemcpyHostToDevice);
    const int waves = 4;
    const int64_t grid = (n + 64 * waves - 1) / (64 * waves);
    if (e == hipSuccess) e = hipFuncSetAttribute((const void*)mlp_eval_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, kPkBytes);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(mlp_eval_kernel, dim3((unsigned)grid), dim3(64 * waves), kPkBytes, h->stream,
                           h->mlp_packed.as<unsigned char>() + (size_t)member * kPkBytes, h->mlp_norm, n, xb.as<float>(),
                           hb.as<float>(), db.as<float>());
        e = hipGetLastError();
    }
    std::vector<float> hf((size_t)n * kMlpOut), df((size_t)n * kMlpOut * kMlpIn);
    if (e == hipSuccess) e = hipMemcpy(hf.data(), hb.p, sizeof(float) * hf.size(), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(df.data(), db.p, sizeof(float) * df.size(), hipMemcpyDeviceToHost);
    xb.release(); hb.release(); db.release();
    if (e != hipSuccess) return fail(h, NQCD_ERR_DEVICE, std::string("surrogate evaluation: ") + hipGetErrorString(e));
    for (size_t i = 0; i < hf.size(); ++i) Hout[i] = hf[i];
    if (dH)
        for (size_t i = 0; i < df.size(); ++i) dH[i] = df[i];
    return NQCD_OK;
}

int32_t mlp_set_params(nqcd_handle_t* h, int32_t member, const float* P, const float* norm) {
    if (member < 0 || member >= kMlpMembers) return fail(h, NQCD_ERR_ARG, "member out of range");
    HIPCHK(h, hipSetDevice(h->device));
    if (int32_t rc = mlp_ensure(h)) return rc;
    HIPCHK(h, hipMemcpy(h->mlp_master.as<float>() + (size_t)member * kMlpParams, P, sizeof(float) * kMlpParams,
                        hipMemcpyHostToDevice));
    if (norm) std::memcpy(&h->mlp_norm, norm, sizeof(MlpNorm));
    hipLaunchKernelGGL(mlp_pack, dim3(kPackBlocks, 1), dim3(256), 0, h->stream, h->mlp_master.as<float>() + (size_t)member * kMlpParams,
                       h->mlp_packed.as<unsigned char>() + (size_t)member * kPkBytes);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->mlp_ready = true;
    return NQCD_OK;
}

int32_t mlp_get_params(nqcd_handle_t* h, int32_t member, float* P, float* norm) {
    if (!h->mlp_ready) return fail(h, NQCD_ERR_STATE, "no surrogate: train or load it first");
    if (member < 0 || member >= kMlpMembers) return fail(h, NQCD_ERR_ARG, "member out of range");
    HIPCHK(h, hipSetDevice(h->device));
    HIPCHK(h, hipMemcpy(P, h->mlp_master.as<float>() + (size_t)member * kMlpParams, sizeof(float) * kMlpParams,
                        hipMemcpyDeviceToHost));
    if (norm) std::memcpy(norm, &h->mlp_norm, sizeof(MlpNorm));
    return NQCD_OK;
}

}  // namespace nqcd
