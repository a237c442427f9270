// ensemble.hip -- device-side initial-condition sampling and deterministic
// ensemble reductions.
//
// Sampling stands in for RandomSelection/sample_distribution
// (src/Ensembles/selections.jl:70-99) for the recognised distributions; the
// reductions stand in for MeanReduction/SumReduction over the outputs of
// DynamicsOutputs.jl / TimeCorrelationFunctions.jl (src/Ensembles/reductions.jl:8-52).
// Reductions run in a fixed order (chunked tree sums, then an ordered sum of
// chunk partials) so results do not depend on the device or launch shape.
#include <algorithm>
#include <mutex>

#include "ensemble.h"

namespace nqcd {

// x ~ N(x0, p0/(20 sqrt 2)), v = p0/m; model_id = traj % nm; p0 = p_lo + (traj/nm) % n_p.
__global__ __launch_bounds__(256) void sample_tully_grid(int64_t ntraj, int32_t nm, double x0, double p_lo, int32_t n_p,
                                                         double mass, uint64_t key, double* r0, double* v0,
                                                         int32_t* model_id) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntraj) return;
    const int32_t m = (int32_t)(i % nm);
    const double p0 = p_lo + (double)((i / nm) % n_p);
    double z0, z1;
    philox_normal2((uint64_t)i, 0u, STREAM_INIT, key, z0, z1);
    r0[i] = x0 + (p0 / (20.0 * 1.4142135623730951)) * z0;
    v0[i] = p0 / mass;
    model_id[i] = m;
}

// Per-dof Gaussian: r = mu_r + sd_r z0, v = mu_v + sd_v z1 with (z0, z1) from counter (traj, dof, 0).
__global__ __launch_bounds__(256) void sample_gaussian(int64_t ntraj, int32_t D, const double* prm, uint64_t key,
                                                       double* r0, double* v0) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntraj) return;
    for (int I = 0; I < D; ++I) {
        double z0, z1;
        philox_normal2((uint64_t)i, (uint32_t)I, STREAM_INIT, key, z0, z1);
        r0[(int64_t)I * ntraj + i] = prm[4 * I + 0] + prm[4 * I + 1] * z0;
        v0[(int64_t)I * ntraj + i] = prm[4 * I + 2] + prm[4 * I + 3] * z1;
    }
}

// Pure electronic state in the adiabatic or diabatic basis for two-state one-dof
// models (density_matrix_dynamics.jl:37-75): sigma_ad = U^T rho U.
__global__ __launch_bounds__(256) void init_density2(int64_t ntraj, int32_t state, int32_t diabatic, const ModelSlot* models,
                                                     const int32_t* model_id, const double* r0, double* sig0) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntraj) return;
    double a, b, x;
    if (!diabatic) {
        a = (state == 0) ? 1.0 : 0.0; b = 1.0 - a; x = 0.0;
    } else {
        const ModelRegs m = model_regs(models, model_id ? model_id[i] : 0);
        double v11, v22, v12, d11, d22, d12, E0, E1, U[4];
        model1d_eval<double>(m, r0[i], v11, v22, v12, d11, d22, d12);
        eigen2(v11, v12, v22, E0, E1, U, false);
        const double u0 = U[state * 2 + 0], u1 = U[state * 2 + 1];  // row `state` of U
        a = u0 * u0; b = u1 * u1; x = u0 * u1;
    }
    sig0[i] = a; sig0[ntraj + i] = b; sig0[2 * ntraj + i] = x; sig0[3 * ntraj + i] = 0.0;
}

// ---------------------------------------------------------------------------
// Per-trajectory output rows for the two-state kernels.
// ---------------------------------------------------------------------------
// Number of rows of an output computed from the final state or the save slots.
__host__ __device__ inline int32_t output2_rows(int32_t kind, int32_t nsave, int32_t nst = 2) {
    return kind == OUTK_FINAL_ADIABATIC ? 2 : (kind == OUTK_SCATTERING_1D || kind == OUTK_SCATTERING_1D_DIABATIC) ? 4
                                                                                                     : nst * nst * nsave;
}

// Row `row` of trajectory i (one expression shared by the materialising and the fused reductions,
// so both sum the same values in the same order).
__device__ __forceinline__ double output2_value(const OutputSrc& S, int64_t ntraj, int32_t row, int64_t i) {
    if (S.kind == OUTK_FINAL_ADIABATIC || S.kind == OUTK_SCATTERING_1D || S.kind == OUTK_SCATTERING_1D_DIABATIC) {
        double p0, p1;
        if (S.kind == OUTK_SCATTERING_1D_DIABATIC) { p0 = S.final_pop_di[i]; p1 = S.final_pop_di[ntraj + i]; }
        else if (S.method == 0) { p0 = (S.state[i] == 0) ? 1.0 : 0.0; p1 = 1.0 - p0; }
        else { p0 = S.sig[i]; p1 = S.sig[ntraj + i]; }
        if (S.kind == OUTK_FINAL_ADIABATIC) return row == 0 ? p0 : p1;
        // OutputStateResolvedScattering1D (adiabatic or diabatic populations of the final state): rows
        // [reflection(2), transmission(2)] (DynamicsOutputs.jl:313-338)
        const bool trans = S.r[(int64_t)S.xdof * ntraj + i] > 0.0;
        const double p = (row & 1) ? p1 : p0;
        return (row < 2) ? (trans ? 0.0 : p) : (trans ? p : 0.0);
    }
    // OUTK_POP_CORRELATION: row = n^2 slot + (a + n b) holds norm K_a(0) P_b(t_slot), column-major n x n per slot
    // (TimeCorrelationFunctions.jl:26-40,86-88), K the initial estimator (the slot-0 population unless given)
    const int32_t n = S.nst, nn = n * n;
    const int32_t slot = row / nn, ab = row - slot * nn, a = ab % n, b = ab / n;
    const double* K0 = S.init_pop ? S.init_pop : S.save_pop_di;
    return S.norm * K0[(int64_t)a * ntraj + i] * S.save_pop_di[((int64_t)slot * n + b) * ntraj + i];
}

__global__ __launch_bounds__(256) void fill_output2(OutputSrc S, int64_t ntraj, double* out) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ntraj) return;
    const int32_t rows = output2_rows(S.kind, S.nsave, S.nst);
    for (int32_t row = 0; row < rows; ++row) out[(int64_t)row * ntraj + i] = output2_value(S, ntraj, row, i);
}

// Deterministic two-pass row reduction: rows x ntraj -> rows.  Pass 1 sums chunk `blockIdx.x` of row
// row0 + blockIdx.y with a tree in LDS; pass 2 sums a row's chunk partials in order.  The source is a
// materialised array (data, row-major [row][traj]) or, with data null, the output rows of S computed
// on the fly.
__global__ __launch_bounds__(256) void reduce_rows_pass1(const double* data, OutputSrc S, int64_t ntraj, int32_t nchunk,
                                                         int32_t row0, double* partial) {
    __shared__ double sm[256];
    const int row = row0 + (int)blockIdx.y, chunk = blockIdx.x;
    const int64_t len = (ntraj + nchunk - 1) / nchunk;
    const int64_t b = (int64_t)chunk * len;
    const int64_t e = (b + len < ntraj) ? b + len : ntraj;
    double s = 0.0;
    if (data) {
        const double* d = data + (int64_t)row * ntraj;
        for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) s += d[i];
    } else {
        for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) s += output2_value(S, ntraj, row, i);
    }
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(int64_t)row * nchunk + chunk] = sm[0];
}

// out[row] = scale * (sum of row's partials) [+ scale * (sum of row's partials in partial2)]
__global__ void reduce_rows_pass2(const double* partial, const double* partial2, int32_t rows, int32_t nchunk, double scale,
                                  double* out) {
    const int row = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (row >= rows) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += partial[(int64_t)row * nchunk + c];
    double v = s * scale;
    if (partial2) {
#pragma clang fp contract(off)
        // a rounded product, then the sum, as the materialising path adds two rounded rows on the host
        double s2 = 0.0;
        for (int c = 0; c < nchunk; ++c) s2 += partial2[(int64_t)row * nchunk + c];
        const double w = s2 * scale;
        v = v + w;
    }
    out[row] = v;
}

static inline int grid_for(int64_t n, int block) { return (int)((n + block - 1) / block); }

hipError_t launch_sample_tully_grid(int64_t ntraj, int32_t nm, double x0, double p_lo, int32_t n_p, double mass,
                                    uint64_t key, double* r0, double* v0, int32_t* model_id, hipStream_t s) {
    hipLaunchKernelGGL(sample_tully_grid, dim3(grid_for(ntraj, 256)), dim3(256), 0, s, ntraj, nm, x0, p_lo, n_p, mass,
                       key, r0, v0, model_id);
    return hipGetLastError();
}

hipError_t launch_sample_gaussian(int64_t ntraj, int32_t D, const double* prm, uint64_t key, double* r0, double* v0,
                                  hipStream_t s) {
    hipLaunchKernelGGL(sample_gaussian, dim3(grid_for(ntraj, 256)), dim3(256), 0, s, ntraj, D, prm, key, r0, v0);
    return hipGetLastError();
}

hipError_t launch_init_density2(int64_t ntraj, int32_t state, int32_t diabatic, const ModelSlot* models,
                                const int32_t* model_id, const double* r0, double* sig0, hipStream_t s) {
    hipLaunchKernelGGL(init_density2, dim3(grid_for(ntraj, 256)), dim3(256), 0, s, ntraj, state, diabatic, models,
                       model_id, r0, sig0);
    return hipGetLastError();
}

hipError_t launch_fill_output2(const OutputSrc& S, int64_t ntraj, double* out, hipStream_t s) {
    hipLaunchKernelGGL(fill_output2, dim3(grid_for(ntraj, 256)), dim3(256), 0, s, S, ntraj, out);
    return hipGetLastError();
}

int32_t fill_output2_rows(int32_t kind, int32_t nsave) { return output2_rows(kind, nsave); }

int reduce_nchunk(int64_t ntraj) {
    int64_t c = (ntraj + 4095) / 4096;
    if (c < 1) c = 1;
    if (c > 1024) c = 1024;
    return (int)c;
}

// Pass 1 over row blocks of at most 65535 (the grid's y limit), then pass 2.
static hipError_t reduce_rows_launch(const double* data, const OutputSrc& S, int32_t rows, int64_t ntraj, double* partial,
                                     hipStream_t s) {
    const int nchunk = reduce_nchunk(ntraj);
    for (int32_t row0 = 0; row0 < rows; row0 += 65535) {
        const int32_t nr = std::min<int32_t>(65535, rows - row0);
        hipLaunchKernelGGL(reduce_rows_pass1, dim3(nchunk, nr), dim3(256), 0, s, data, S, ntraj, nchunk, row0, partial);
    }
    return hipGetLastError();
}

hipError_t launch_reduce_rows(const double* data, int32_t rows, int64_t ntraj, double* partial, double scale, double* out,
                              hipStream_t s) {
    hipError_t e = reduce_rows_launch(data, OutputSrc{}, rows, ntraj, partial, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(reduce_rows_pass2, dim3(grid_for(rows, 64)), dim3(64), 0, s, partial, (const double*)nullptr, rows,
                       reduce_nchunk(ntraj), scale, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Allocation-free reduction.  Its partial sums and result live in `scratch`, one array per device
// that the library holds until the process ends; host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (the array's name is a host-side stand-in).
// Every call on a device shares that device's array, so calls that overlap on one device race.
// Pass 1 writes every partial it reads and pass 2 writes every result it returns, so no part of the
// array has to start at zero.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[kReduceScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t reduce_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;  // the caller makes `device` current first
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = static_cast<double*>(p);
    return hipSuccess;
}

double* reduce_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

hipError_t launch_reduce_rows_noalloc(const double* data, const double* data2, const OutputSrc& S, int32_t rows,
                                      int64_t ntraj, double scale, double* scratch_ptr, hipStream_t s) {
    const int nchunk = reduce_nchunk(ntraj);
    double* partial = scratch_ptr;
    double* partial2 = data2 ? scratch_ptr + (size_t)rows * nchunk : nullptr;
    double* out = scratch_ptr + (size_t)rows * nchunk * (data2 ? 2 : 1);
    hipError_t e = reduce_rows_launch(data, S, rows, ntraj, partial, s);
    if (e != hipSuccess) return e;
    if (data2) {
        e = reduce_rows_launch(data2, S, rows, ntraj, partial2, s);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(reduce_rows_pass2, dim3(grid_for(rows, 64)), dim3(64), 0, s, (const double*)partial,
                       (const double*)partial2, rows, nchunk, scale, out);
    return hipGetLastError();
}

size_t reduce_noalloc_doubles(int32_t rows, int64_t ntraj, bool two_sources) {
    return (size_t)rows * reduce_nchunk(ntraj) * (two_sources ? 2 : 1) + (size_t)rows;
}

__global__ __launch_bounds__(256) void zero_words(uint32_t* __restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}

// Sum of n int32 values into out[0]: each wave sums a strided share, reduces it with shuffles and
// adds it to out[0] with one 64-bit atomic (integer addition, so the result does not depend on the
// order); the host then copies 8 bytes instead of the whole array.
__global__ void sum_i32_clear(long long* out) { out[0] = 0; }

__global__ __launch_bounds__(256) void sum_i32(const int32_t* __restrict__ x, int64_t n, long long* __restrict__ out) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    long long a = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) a += x[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o);
    if ((threadIdx.x & 63) == 0 && a != 0) atomicAdd(reinterpret_cast<unsigned long long*>(out), (unsigned long long)a);
}

hipError_t launch_sum_i32(const int32_t* x, int64_t n, long long* out, hipStream_t s) {
    hipLaunchKernelGGL(sum_i32_clear, dim3(1), dim3(1), 0, s, out);
    const int64_t blocks = std::min<int64_t>(std::max<int64_t>((n + 255) / 256, 1), 1024);
    hipLaunchKernelGGL(sum_i32, dim3((unsigned)blocks), dim3(256), 0, s, x, n, out);
    return hipGetLastError();
}

hipError_t launch_zero(void* p, size_t bytes, hipStream_t s) {
    const size_t n = bytes / 4;
    if (!p || n == 0) return hipSuccess;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(zero_words, dim3((unsigned)blocks), dim3(256), 0, s, reinterpret_cast<uint32_t*>(p), n);
    return hipGetLastError();
}

}  // namespace nqcd
