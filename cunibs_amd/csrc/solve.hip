// Block PCG for k right-hand sides sharing one operator (solver/amg/solver.cpp:84-198 solves the
// same systems as k lockstep scalar chains; this is a true block method instead).
//
// Precision. The operator, the V-cycle and every (n, K) block vector are FP32. FP64 holds the
// solution X, the k x k Gram matrices and their pivoted Cholesky, the block coefficients, and the
// true residual B - A X that an outer iterative-refinement loop recomputes with the FP64 operator.
// Convergence is judged on that true residual, per column, against ||b|| (solver.cpp:237-246).
//
// Iteration (A-orthonormal search blocks; breakdown-free by dropping dependent directions):
//     S = A W,  G = W^T S,  pivoted Cholesky of G -> T (K x r) with T^T G T = I
//     P = W T,  Q = S T,  alpha = P^T R,  X += P alpha,  R -= Q alpha
//     Z = M R (one V-cycle),  C = Q^T Z,  W = Z - P C
// A column whose recursive residual meets its target freezes (its alpha column is zero).
//
// Every FP64 reduction is a fixed two-stage tree: kGramBlocks contiguous row chunks, each summed
// in a fixed lane/wave order, then the chunk partials summed in chunk order. Nothing depends on
// scheduling, so a solve is bit-reproducible run to run.
//
// Kernels read a device `done` flag and return at once when it is set, so the host can replay
// the captured iteration graph several times between convergence reads.
#include <algorithm>
#include <cstring>

#include "solve.hpp"

namespace cn {

int padded_width(int k) {
    for (int w : kBlockWidths)
        if (w >= k) return w;
    throw InvalidArgument("block width exceeds the largest compiled width (32)");
}

namespace {

template <int K>
constexpr int lanes_per_row() {
    return K >= 8 ? K : 8;
}

// Workgroups go to the 8 XCDs round-robin. Renumber them so that each XCD receives one contiguous
// range of logical blocks, which keeps a row range's column gathers inside one XCD's L2.
__device__ __forceinline__ int logical_block() {
    const int per = static_cast<int>(gridDim.x) >> 3;  // the grid is a multiple of 8
    return static_cast<int>(blockIdx.x & 7u) * per + static_cast<int>(blockIdx.x >> 3);
}

__device__ __forceinline__ bool finished(const int* state) { return state != nullptr && state[0] != 0; }

unsigned xcd_grid(std::int64_t blocks) { return static_cast<unsigned>((blocks + 7) / 8 * 8); }

// Sum of `acc` over the row groups of a wave that share a lane-in-row index (xor over offsets
// L, 2L, ..., 32): afterwards every lane holds its column's wave total.
template <int L, typename T>
__device__ __forceinline__ T wave_rows_sum(T v) {
#pragma unroll
    for (int off = L; off < kWave; off <<= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Sequential, wave-ordered accumulation of per-wave K-vectors into a block partial. `g` holds
// the wave's totals in lanes [0, K) (column = lane) for rows a = 0..K-1.
template <int K, int WAVES>
__device__ __forceinline__ void block_gram_store(const double (&g)[K], double* __restrict__ out) {
    __shared__ double acc[K * K];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
    for (int e = threadIdx.x; e < K * K; e += blockDim.x) acc[e] = 0.0;
    __syncthreads();
    for (int w = 0; w < WAVES; ++w) {
        if (wave == w && lane < K) {
#pragma unroll
            for (int a = 0; a < K; ++a) acc[a * K + lane] += g[a];
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < K * K; e += blockDim.x) out[e] = acc[e];
}

template <int K, int WAVES>
__device__ __forceinline__ void block_vec_store(double v, double* __restrict__ out) {
    __shared__ double acc[K];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
    if (threadIdx.x < K) acc[threadIdx.x] = 0.0;
    __syncthreads();
    for (int w = 0; w < WAVES; ++w) {
        if (wave == w && lane < K) acc[lane] += v;
        __syncthreads();
    }
    if (threadIdx.x < K) out[threadIdx.x] = acc[threadIdx.x];
}

constexpr int kGramThreads = 1024;
constexpr int kGramWaves = kGramThreads / kWave;

// ---------------------------------------------------------------------------------------------
// SpMM: Y = A X, X and Y (n, K) row-major FP32. L lanes per row: lane column c = l % K; for K < 8
// the L/K sub-lanes stride the row's nonzeros and are combined by a fixed xor tree. With GRAM the
// kernel also accumulates W^T Y in FP64 (W = X here), one K x K partial per logical block.
// ---------------------------------------------------------------------------------------------
template <int K, bool GRAM>
__global__ void __launch_bounds__(kGramThreads) spmm_kernel(
    std::int64_t n, const std::int32_t* __restrict__ rp, const std::int32_t* __restrict__ ci,
    const float* __restrict__ av, const float* __restrict__ X, float* __restrict__ Y,
    double* __restrict__ partials, std::int64_t rows_per_block, const int* __restrict__ state) {
    if (finished(state)) return;
    constexpr int L = lanes_per_row<K>();
    constexpr int SUB = L / K;
    constexpr int RPW = kWave / L;
    const int lb = logical_block();
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
    const int l = lane % L, c = l % K, sub = l / K;
    const int group = lane - l;
    const std::int64_t lo = static_cast<std::int64_t>(lb) * rows_per_block;
    const std::int64_t hi = min(lo + rows_per_block, n);
    double g[K];
    if constexpr (GRAM) {
#pragma unroll
        for (int a = 0; a < K; ++a) g[a] = 0.0;
    }
    for (std::int64_t r0 = lo + static_cast<std::int64_t>(wave) * RPW; r0 < hi; r0 += kGramWaves * RPW) {
        const std::int64_t row = r0 + lane / L;
        const bool valid = row < hi;
        float acc = 0.0f;
        if (valid) {
            const int e = rp[row + 1];
            for (int j = rp[row] + sub; j < e; j += SUB)
                acc = fmaf(av[j], X[static_cast<std::int64_t>(ci[j]) * K + c], acc);
        }
#pragma unroll
        for (int off = K; off < L; off <<= 1) acc += __shfl_xor(acc, off, kWave);
        if (valid && sub == 0) Y[row * K + c] = acc;
        if constexpr (GRAM) {
            const float w_own = valid ? X[row * K + c] : 0.0f;
            const double s = (valid && sub == 0) ? static_cast<double>(acc) : 0.0;
#pragma unroll
            for (int a = 0; a < K; ++a) {
                const float wa = __shfl(w_own, group + a, kWave);
                g[a] = fma(static_cast<double>(wa), s, g[a]);
            }
        }
    }
    if constexpr (GRAM) {
#pragma unroll
        for (int a = 0; a < K; ++a) g[a] = wave_rows_sum<L>(g[a]);
        block_gram_store<K, kGramWaves>(g, partials + static_cast<std::int64_t>(lb) * K * K);
    }
}

// ---------------------------------------------------------------------------------------------
// FP64 true residual: R64 = B - A X64 with the FP64 operator, plus per-column ||R64||^2 partials.
// B is the FP32 right-hand side (exact in FP64).
// ---------------------------------------------------------------------------------------------
template <int K>
__global__ void __launch_bounds__(kGramThreads) residual64_kernel(
    std::int64_t n, const std::int32_t* __restrict__ rp, const std::int32_t* __restrict__ ci,
    const double* __restrict__ av, const float* __restrict__ B, const double* __restrict__ X,
    double* __restrict__ R, double* __restrict__ partials, std::int64_t rows_per_block, bool x_zero) {
    constexpr int L = lanes_per_row<K>();
    constexpr int SUB = L / K;
    constexpr int RPW = kWave / L;
    const int lb = logical_block();
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
    const int l = lane % L, c = l % K, sub = l / K;
    const std::int64_t lo = static_cast<std::int64_t>(lb) * rows_per_block;
    const std::int64_t hi = min(lo + rows_per_block, n);
    double nrm = 0.0;
    for (std::int64_t r0 = lo + static_cast<std::int64_t>(wave) * RPW; r0 < hi; r0 += kGramWaves * RPW) {
        const std::int64_t row = r0 + lane / L;
        const bool valid = row < hi;
        double acc = 0.0;
        if (valid && !x_zero) {
            const int e = rp[row + 1];
            for (int j = rp[row] + sub; j < e; j += SUB)
                acc = fma(av[j], X[static_cast<std::int64_t>(ci[j]) * K + c], acc);
        }
#pragma unroll
        for (int off = K; off < L; off <<= 1) acc += __shfl_xor(acc, off, kWave);
        if (valid && sub == 0) {
            const double r = static_cast<double>(B[row * K + c]) - acc;
            R[row * K + c] = r;
            nrm = fma(r, r, nrm);
        }
    }
    nrm = wave_rows_sum<L>(nrm);
    block_vec_store<K, kGramWaves>(nrm, partials + static_cast<std::int64_t>(lb) * K);
}

// ---------------------------------------------------------------------------------------------
// V-cycle kernels (solver/amg/vcycle.cu:124-311), K-wide, FP32.
// ---------------------------------------------------------------------------------------------
constexpr int kRowThreads = 256;
constexpr int kRowWaves = kRowThreads / kWave;
constexpr int kRowIters = 4;  // row groups each wave walks per block

template <int K>
constexpr std::int64_t rows_per_row_block() {
    return static_cast<std::int64_t>(kRowWaves) * (kWave / lanes_per_row<K>()) * kRowIters;
}

// MODE 0: residual  out = b - A x
// MODE 1: Chebyshev out = c_cur x + c_prev x_prev + alpha dinv (b - A x); out may alias x_prev.
template <int K, int MODE>
__global__ void __launch_bounds__(kRowThreads) smoother_kernel(
    std::int64_t n, const std::int32_t* __restrict__ rp, const std::int32_t* __restrict__ ci,
    const float* __restrict__ av, const float* __restrict__ dinv, const float* __restrict__ b,
    const float* __restrict__ x, const float* x_prev, float c_cur, float c_prev, float alpha, float* out,
    const int* __restrict__ state) {
    if (finished(state)) return;
    constexpr int L = lanes_per_row<K>();
    constexpr int SUB = L / K;
    constexpr int RPW = kWave / L;
    const int lb = logical_block();
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
    const int l = lane % L, c = l % K, sub = l / K;
    const std::int64_t lo = static_cast<std::int64_t>(lb) * rows_per_row_block<K>();
    const std::int64_t hi = min(lo + rows_per_row_block<K>(), n);
    for (std::int64_t r0 = lo + static_cast<std::int64_t>(wave) * RPW; r0 < hi; r0 += kRowWaves * RPW) {
        const std::int64_t row = r0 + lane / L;
        const bool valid = row < hi;
        float acc = 0.0f;
        if (valid) {
            const int e = rp[row + 1];
            for (int j = rp[row] + sub; j < e; j += SUB)
                acc = fmaf(av[j], x[static_cast<std::int64_t>(ci[j]) * K + c], acc);
        }
#pragma unroll
        for (int off = K; off < L; off <<= 1) acc += __shfl_xor(acc, off, kWave);
        if (valid && sub == 0) {
            const std::int64_t at = row * K + c;
            const float rhs = b[at] - acc;
            if constexpr (MODE == 0) {
                out[at] = rhs;
            } else {
                const float d = alpha * dinv[row];
                const float cur = x[at];
                const float prev = x_prev[at];  // read before the store below: out may alias x_prev
                out[at] = fmaf(d, rhs, fmaf(c_prev, prev, c_cur * cur));
            }
        }
    }
}

template <int K>
__global__ void jacobi_zero_kernel(std::int64_t n, const float* __restrict__ dinv, const float* __restrict__ b,
                                   float alpha, float* __restrict__ x, const int* __restrict__ state) {
    if (finished(state)) return;
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n * K) x[i] = alpha * dinv[i / K] * b[i];
}

// b_coarse = R r_fine, one wavefront per coarse row (R rows are the widest in the hierarchy).
template <int K>
__global__ void __launch_bounds__(kRowThreads) restrict_kernel(
    std::int64_t n_coarse, const std::int32_t* __restrict__ rp, const std::int32_t* __restrict__ ci,
    const float* __restrict__ rv, const float* __restrict__ fine, float* __restrict__ out,
    const int* __restrict__ state) {
    if (finished(state)) return;
    constexpr int SUB = kWave / K;
    const std::int64_t row = static_cast<std::int64_t>(blockIdx.x) * kRowWaves + (threadIdx.x >> 6);
    const int lane = threadIdx.x & (kWave - 1);
    const int c = lane % K, sub = lane / K;
    float acc = 0.0f;
    if (row < n_coarse) {
        const int e = rp[row + 1];
        for (int j = rp[row] + sub; j < e; j += SUB)
            acc = fmaf(rv[j], fine[static_cast<std::int64_t>(ci[j]) * K + c], acc);
    }
#pragma unroll
    for (int off = K; off < kWave; off <<= 1) acc += __shfl_xor(acc, off, kWave);
    if (row < n_coarse && sub == 0) out[row * K + c] = acc;
}

// x_fine += P x_coarse.
template <int K>
__global__ void __launch_bounds__(kRowThreads) prolongate_kernel(
    std::int64_t n, const std::int32_t* __restrict__ rp, const std::int32_t* __restrict__ ci,
    const float* __restrict__ pv, const float* __restrict__ coarse, float* __restrict__ fine,
    const int* __restrict__ state) {
    if (finished(state)) return;
    constexpr int L = lanes_per_row<K>();
    constexpr int SUB = L / K;
    constexpr int RPW = kWave / L;
    const int lb = logical_block();
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
    const int l = lane % L, c = l % K, sub = l / K;
    const std::int64_t lo = static_cast<std::int64_t>(lb) * rows_per_row_block<K>();
    const std::int64_t hi = min(lo + rows_per_row_block<K>(), n);
    for (std::int64_t r0 = lo + static_cast<std::int64_t>(wave) * RPW; r0 < hi; r0 += kRowWaves * RPW) {
        const std::int64_t row = r0 + lane / L;
        const bool valid = row < hi;
        float acc = 0.0f;
        if (valid) {
            const int e = rp[row + 1];
            for (int j = rp[row] + sub; j < e; j += SUB)
                acc = fmaf(pv[j], coarse[static_cast<std::int64_t>(ci[j]) * K + c], acc);
        }
#pragma unroll
        for (int off = K; off < L; off <<= 1) acc += __shfl_xor(acc, off, kWave);
        if (valid && sub == 0) fine[row * K + c] += acc;
    }
}

// Coarsest level in FP64: x = Ainv b, one wavefront per row of the dense inverse.
template <int K>
__global__ void __launch_bounds__(kRowThreads) coarse_kernel(int n, const double* __restrict__ inv,
                                                             const float* __restrict__ b, float* __restrict__ x,
                                                             const int* __restrict__ state) {
    if (finished(state)) return;
    const int row = static_cast<int>(blockIdx.x) * kRowWaves + static_cast<int>(threadIdx.x >> 6);
    const int lane = threadIdx.x & (kWave - 1);
    double acc[K];
#pragma unroll
    for (int c = 0; c < K; ++c) acc[c] = 0.0;
    if (row < n) {
        const double* a = inv + static_cast<std::int64_t>(row) * n;
        for (int j = lane; j < n; j += kWave) {
            const double v = a[j];
#pragma unroll
            for (int c = 0; c < K; ++c) acc[c] = fma(v, static_cast<double>(b[static_cast<std::int64_t>(j) * K + c]), acc[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < K; ++c)
        for (int off = kWave / 2; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off, kWave);
    if (row < n && lane < K) {
        double v = 0.0;
#pragma unroll
        for (int c = 0; c < K; ++c)
            if (c == lane) v = acc[c];
        x[static_cast<std::int64_t>(row) * K + lane] = static_cast<float>(v);
    }
}

// ---------------------------------------------------------------------------------------------
// Block-vector kernels of the PCG iteration.
// ---------------------------------------------------------------------------------------------

// P = W T, Q = S T (T: K x K, FP64 on the device, applied in FP32), partials of alpha = P^T R.
template <int K>
__global__ void __launch_bounds__(kGramThreads) form_pq_kernel(
    std::int64_t n, const float* __restrict__ W, const float* __restrict__ S, const double* __restrict__ T,
    const float* __restrict__ R, float* __restrict__ P, float* __restrict__ Q, double* __restrict__ partials,
    std::int64_t rows_per_block, const int* __restrict__ state) {
    if (finished(state)) return;
    constexpr int L = lanes_per_row<K>();
    constexpr int SUB = L / K;
    constexpr int RPW = kWave / L;
    __shared__ float t[K * K];
    for (int e = threadIdx.x; e < K * K; e += blockDim.x) t[e] = static_cast<float>(T[e]);
    __syncthreads();
    const int lb = logical_block();
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
    const int l = lane % L, c = l % K, sub = l / K;
    const int group = lane - l;
    const std::int64_t lo = static_cast<std::int64_t>(lb) * rows_per_block;
    const std::int64_t hi = min(lo + rows_per_block, n);
    double g[K];
#pragma unroll
    for (int a = 0; a < K; ++a) g[a] = 0.0;
    for (std::int64_t r0 = lo + static_cast<std::int64_t>(wave) * RPW; r0 < hi; r0 += kGramWaves * RPW) {
        const std::int64_t row = r0 + lane / L;
        const bool valid = row < hi;
        const std::int64_t at = row * K + c;
        const float w_own = valid ? W[at] : 0.0f;
        const float s_own = valid ? S[at] : 0.0f;
        float p = 0.0f, q = 0.0f;
#pragma unroll
        for (int a = 0; a < K; ++a) {
            const float wa = __shfl(w_own, group + a, kWave);
            const float sa = __shfl(s_own, group + a, kWave);
            p = fmaf(wa, t[a * K + c], p);
            q = fmaf(sa, t[a * K + c], q);
        }
        const float r_own = valid ? R[at] : 0.0f;
        if (valid && sub == 0) {
            P[at] = p;
            Q[at] = q;
        }
        // alpha[b][c] += P[row][b] * R[row][c]
        const double rr = (valid && sub == 0) ? static_cast<double>(r_own) : 0.0;
#pragma unroll
        for (int b = 0; b < K; ++b) {
            const float pb = __shfl(p, group + b, kWave);
            g[b] = fma(static_cast<double>(pb), rr, g[b]);
        }
    }
#pragma unroll
    for (int a = 0; a < K; ++a) g[a] = wave_rows_sum<L>(g[a]);
    block_gram_store<K, kGramWaves>(g, partials + static_cast<std::int64_t>(lb) * K * K);
}

// X64 += P alpha (FP64), R -= Q alpha (FP32), partials of ||R_c||^2. Frozen columns are skipped.
template <int K>
__global__ void __launch_bounds__(kGramThreads) update_xr_kernel(
    std::int64_t n, const float* __restrict__ P, const float* __restrict__ Q, const double* __restrict__ alpha,
    double* __restrict__ X, float* __restrict__ R, double* __restrict__ partials, std::int64_t rows_per_block,
    const int* __restrict__ state) {
    if (finished(state)) return;
    constexpr int L = lanes_per_row<K>();
    constexpr int RPW = kWave / L;
    constexpr int SUB = L / K;
    __shared__ double al[K * K];
    for (int e = threadIdx.x; e < K * K; e += blockDim.x) al[e] = alpha[e];
    __syncthreads();
    const int lb = logical_block();
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
    const int l = lane % L, c = l % K, sub = l / K;
    const int group = lane - l;
    const bool frozen = state[2 + c] != 0;
    const std::int64_t lo = static_cast<std::int64_t>(lb) * rows_per_block;
    const std::int64_t hi = min(lo + rows_per_block, n);
    double nrm = 0.0;
    for (std::int64_t r0 = lo + static_cast<std::int64_t>(wave) * RPW; r0 < hi; r0 += kGramWaves * RPW) {
        const std::int64_t row = r0 + lane / L;
        const bool valid = row < hi;
        const std::int64_t at = row * K + c;
        const float p_own = valid ? P[at] : 0.0f;
        const float q_own = valid ? Q[at] : 0.0f;
        double dx = 0.0;
        float dr = 0.0f;
#pragma unroll
        for (int b = 0; b < K; ++b) {
            const float pb = __shfl(p_own, group + b, kWave);
            const float qb = __shfl(q_own, group + b, kWave);
            const double a = al[b * K + c];
            dx = fma(static_cast<double>(pb), a, dx);
            dr = fmaf(qb, static_cast<float>(a), dr);
        }
        if (valid && sub == 0) {
            float r = R[at];
            if (!frozen) {
                X[at] += dx;
                r -= dr;
                R[at] = r;
            }
            nrm = fma(static_cast<double>(r), static_cast<double>(r), nrm);
        }
    }
    nrm = wave_rows_sum<L>(nrm);
    block_vec_store<K, kGramWaves>(nrm, partials + static_cast<std::int64_t>(lb) * K);
}

// partials of C = Q^T Z.
template <int K>
__global__ void __launch_bounds__(kGramThreads) gram_kernel(std::int64_t n, const float* __restrict__ A,
                                                            const float* __restrict__ B,
                                                            double* __restrict__ partials,
                                                            std::int64_t rows_per_block,
                                                            const int* __restrict__ state) {
    if (finished(state)) return;
    constexpr int L = lanes_per_row<K>();
    constexpr int RPW = kWave / L;
    constexpr int SUB = L / K;
    const int lb = logical_block();
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
    const int l = lane % L, c = l % K, sub = l / K;
    const int group = lane - l;
    const std::int64_t lo = static_cast<std::int64_t>(lb) * rows_per_block;
    const std::int64_t hi = min(lo + rows_per_block, n);
    double g[K];
#pragma unroll
    for (int a = 0; a < K; ++a) g[a] = 0.0;
    for (std::int64_t r0 = lo + static_cast<std::int64_t>(wave) * RPW; r0 < hi; r0 += kGramWaves * RPW) {
        const std::int64_t row = r0 + lane / L;
        const bool valid = row < hi;
        const std::int64_t at = row * K + c;
        const float a_own = valid ? A[at] : 0.0f;
        const double bb = (valid && sub == 0) ? static_cast<double>(B[at]) : 0.0;
#pragma unroll
        for (int a = 0; a < K; ++a) {
            const float aa = __shfl(a_own, group + a, kWave);
            g[a] = fma(static_cast<double>(aa), bb, g[a]);
        }
    }
#pragma unroll
    for (int a = 0; a < K; ++a) g[a] = wave_rows_sum<L>(g[a]);
    block_gram_store<K, kGramWaves>(g, partials + static_cast<std::int64_t>(lb) * K * K);
}

// W = Z - P C.
template <int K>
__global__ void __launch_bounds__(kRowThreads) update_w_kernel(std::int64_t n, const float* __restrict__ Z,
                                                               const float* __restrict__ P,
                                                               const double* __restrict__ C,
                                                               float* __restrict__ W,
                                                               const int* __restrict__ state) {
    if (finished(state)) return;
    constexpr int L = lanes_per_row<K>();
    __shared__ float cm[K * K];
    for (int e = threadIdx.x; e < K * K; e += blockDim.x) cm[e] = static_cast<float>(C[e]);
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int l = lane % L, c = l % K, sub = l / K;
    const int group = lane - l;
    const std::int64_t row = (static_cast<std::int64_t>(blockIdx.x) * kRowThreads + threadIdx.x) / L;
    const bool valid = row < n;
    const std::int64_t at = row * K + c;
    const float p_own = valid ? P[at] : 0.0f;
    float acc = valid ? Z[at] : 0.0f;
#pragma unroll
    for (int b = 0; b < K; ++b) {
        const float pb = __shfl(p_own, group + b, kWave);
        acc = fmaf(-pb, cm[b * K + c], acc);
    }
    if (valid && sub == 0) W[at] = acc;
}

__global__ void cast_f64_f32_kernel(std::int64_t count, const double* __restrict__ in, float* __restrict__ out) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) out[i] = static_cast<float>(in[i]);
}

__global__ void zero_f64_kernel(std::int64_t count, double* __restrict__ out) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) out[i] = 0.0;
}

// ---------------------------------------------------------------------------------------------
// One-block kernels: fixed-order reduction of the chunk partials, then the k x k work.
// ---------------------------------------------------------------------------------------------
constexpr int kSmallThreads = 1024;

// out[e] = sum_b partials[b * E + e], b ascending, split over chunks of blocks then chunks in order.
template <int E>
__device__ void reduce_partials(const double* __restrict__ partials, int nb, double* out_lds) {
    constexpr int Q = (kSmallThreads / E) > 0 ? (kSmallThreads / E) : 1;
    __shared__ double chunk[Q * E];
    const int t = threadIdx.x;
    if (t < Q * E) {
        const int e = t % E, q = t / E;
        const int b0 = (nb * q) / Q, b1 = (nb * (q + 1)) / Q;
        double s = 0.0;
        for (int b = b0; b < b1; ++b) s += partials[static_cast<std::int64_t>(b) * E + e];
        chunk[q * E + e] = s;
    }
    __syncthreads();
    if (t < E) {
        double s = 0.0;
        for (int q = 0; q < Q; ++q) s += chunk[q * E + t];
        out_lds[t] = s;
    }
    __syncthreads();
}

// G -> T (K x K, columns beyond the rank zero) by pivoted Cholesky of the unit-diagonal scaling.
template <int K>
__global__ void __launch_bounds__(kSmallThreads) chol_kernel(const double* __restrict__ partials, int nb,
                                                             double drop_tol, double* __restrict__ G_out,
                                                             double* __restrict__ T, int* __restrict__ state) {
    if (finished(state)) return;
    __shared__ double A[K * K];
    __shared__ double Lc[K * K];  // Lc[j * K + i]: column j of L at original row i
    __shared__ double sc[K];
    __shared__ double Lp[K * K];
    __shared__ double Li[K * K];
    __shared__ int idx[K];
    __shared__ int rem[K];
    __shared__ int rank_s, stop;
    reduce_partials<K * K>(partials, nb, A);
    const int t = threadIdx.x;
    if (t < K * K) G_out[t] = A[t];
    if (t < K) {
        const double d = A[t * K + t];
        const bool ok = d > 0.0 && isfinite(d);
        sc[t] = ok ? 1.0 / sqrt(d) : 0.0;
        rem[t] = ok ? 1 : 0;
    }
    __syncthreads();
    if (t < K * K) {
        const int i = t / K, j = t % K;
        A[t] = A[t] * sc[i] * sc[j];
        Lc[t] = 0.0;
        T[t] = 0.0;
    }
    if (t == 0) {
        rank_s = 0;
        stop = 0;
    }
    __syncthreads();
    for (int j = 0; j < K; ++j) {
        if (t == 0) {
            int p = -1;
            double best = 0.0;
            for (int i = 0; i < K; ++i)
                if (rem[i] && A[i * K + i] > best) {
                    best = A[i * K + i];
                    p = i;
                }
            if (p < 0 || !(best > drop_tol)) {
                stop = 1;
            } else {
                idx[j] = p;
                rem[p] = 0;
                rank_s = j + 1;
            }
        }
        __syncthreads();
        if (stop) break;
        const int p = idx[j];
        const double ljj = sqrt(A[p * K + p]);
        if (t < K) Lc[j * K + t] = (t == p) ? ljj : (rem[t] ? A[t * K + p] / ljj : 0.0);
        __syncthreads();
        if (t < K * K) {
            const int i = t / K, m = t % K;
            if (rem[i] && rem[m]) A[t] -= Lc[j * K + i] * Lc[j * K + m];
        }
        __syncthreads();
    }
    const int r = rank_s;
    // Lp (r x r, lower): Lp[a][b] = Lc[b][idx[a]]
    if (t < K * K) {
        const int a = t / K, b = t % K;
        Lp[t] = (a < r && b < r && b <= a) ? Lc[b * K + idx[a]] : 0.0;
        Li[t] = 0.0;
    }
    __syncthreads();
    // Li = Lp^-1 column by column (thread b solves Lp y = e_b).
    if (t < r) {
        const int b = t;
        Li[b * K + b] = 1.0 / Lp[b * K + b];
        for (int a = b + 1; a < r; ++a) {
            double s = 0.0;
            for (int q = b; q < a; ++q) s += Lp[a * K + q] * Li[q * K + b];
            Li[a * K + b] = -s / Lp[a * K + a];
        }
    }
    __syncthreads();
    // T[idx[a]][b] = sc[idx[a]] * Li[b][a]
    if (t < K * K) {
        const int a = t / K, b = t % K;
        if (a < r && b < r && b >= a) T[idx[a] * K + b] = sc[idx[a]] * Li[b * K + a];
    }
}

template <int K>
__global__ void __launch_bounds__(kSmallThreads) matrix_kernel(const double* __restrict__ partials, int nb,
                                                               double* __restrict__ out, int* __restrict__ state) {
    if (finished(state)) return;
    __shared__ double A[K * K];
    reduce_partials<K * K>(partials, nb, A);
    if (threadIdx.x < K * K) out[threadIdx.x] = A[threadIdx.x];
}

// ||R_c|| against the column targets; freezes converged columns, counts the iteration.
template <int K>
__global__ void __launch_bounds__(kSmallThreads) converge_kernel(const double* __restrict__ partials, int nb,
                                                                 const double* __restrict__ target,
                                                                 double* __restrict__ norms,
                                                                 int* __restrict__ state) {
    if (finished(state)) return;
    __shared__ double v[K];
    __shared__ int all;
    reduce_partials<K>(partials, nb, v);
    const int t = threadIdx.x;
    if (t == 0) all = 1;
    __syncthreads();
    const int it = state[1] + 1;
    if (t < K) {
        const double rn = sqrt(v[t]);
        norms[t] = rn;
        if (!state[2 + t] && rn <= target[t]) {
            state[2 + t] = 1;
            state[2 + K + t] = it;
        }
    }
    __syncthreads();
    if (t < K && !state[2 + t]) atomicAnd(&all, 0);
    __syncthreads();
    if (t == 0) {
        state[1] = it;
        if (all) state[0] = 1;
    }
}

template <int K>
__global__ void column_norms_kernel(const double* __restrict__ partials, int nb, double* __restrict__ out) {
    __shared__ double v[K];
    reduce_partials<K>(partials, nb, v);
    if (threadIdx.x < K) out[threadIdx.x] = sqrt(v[threadIdx.x]);
}

}  // namespace

// ----------------------------------------------------------------------------------------------
// Host driver
// ----------------------------------------------------------------------------------------------

BlockSolver::BlockSolver(const Operator& op, const Hierarchy& h, hipStream_t s) : op_(op), h_(h), s_(s) {
    check(hipHostMalloc(reinterpret_cast<void**>(&host_state_), sizeof(int) * (2 + 2 * kMaxBlock)),
          "pinned solver state");
}

BlockSolver::~BlockSolver() {
    if (graph_) (void)hipGraphExecDestroy(graph_);
    if (host_state_) (void)hipHostFree(host_state_);
}

void BlockSolver::ensure(int K) {
    if (cap_k_ >= K) return;
    const std::size_t nk = static_cast<std::size_t>(n()) * K;
    R_.reset(nk, "R");
    Z_.reset(nk, "Z");
    W_.reset(nk, "W");
    S_.reset(nk, "S");
    P_.reset(nk, "P");
    Q_.reset(nk, "Q");
    R64_.reset(nk, "R64");
    X64_.reset(nk, "X64");
    lb_.clear();
    lx_.clear();
    lr_.clear();
    for (std::size_t i = 0; i < h_.levels.size(); ++i) {
        const std::size_t ln = static_cast<std::size_t>(h_.levels[i].a.n) * K;
        lb_.emplace_back(i == 0 ? 1 : ln, "level b");
        lx_.emplace_back(ln, "level x");
        lr_.emplace_back(ln, "level r");
    }
    const std::size_t nc = static_cast<std::size_t>(std::max<std::int64_t>(h_.n_coarse(), 1)) * K;
    cb_.reset(nc, "coarse b");
    cx_.reset(nc, "coarse x");
    partials_.reset(static_cast<std::size_t>(kGramBlocks) * K * K, "gram partials");
    G_.reset(K * K, "G");
    T_.reset(K * K, "T");
    C_.reset(K * K, "C");
    alpha_.reset(K * K, "alpha");
    norms_.reset(K, "norms");
    target_.reset(K, "targets");
    state_.reset(2 + 2 * kMaxBlock, "solver state");
    cap_k_ = K;
    if (graph_) {
        (void)hipGraphExecDestroy(graph_);
        graph_ = nullptr;
    }
}

template <int K>
static void launch_smoother(const CsrView& a, const float* dinv, const float* b, const float* x, const float* xp,
                            float cc, float cp, float al, float* out, int mode, const int* state, hipStream_t s) {
    const unsigned g = xcd_grid((a.n + rows_per_row_block<K>() - 1) / rows_per_row_block<K>());
    if (!g) return;
    if (mode == 0)
        smoother_kernel<K, 0><<<g, kRowThreads, 0, s>>>(a.n, a.ptr, a.col, a.val, dinv, b, x, xp, cc, cp, al, out, state);
    else
        smoother_kernel<K, 1><<<g, kRowThreads, 0, s>>>(a.n, a.ptr, a.col, a.val, dinv, b, x, xp, cc, cp, al, out, state);
}

// One zero-guess V-cycle (solver/amg/vcycle.cu:560-648). `state` lets captured replays skip it.
template <int K>
static void run_vcycle(const Hierarchy& h, std::vector<DeviceBuffer<float>>& lb, std::vector<DeviceBuffer<float>>& lx,
                       std::vector<DeviceBuffer<float>>& lr, float* cb, float* cx, const float* B, float* X,
                       const int* state, hipStream_t s) {
    const Cheby& ch = h.cheb;
    const int nl = static_cast<int>(h.levels.size());
    if (nl == 0) {
        coarse_kernel<K><<<grid_for(h.n_coarse(), kRowWaves), kRowThreads, 0, s>>>(
            static_cast<int>(h.n_coarse()), h.coarse_inv.get(), B, X, state);
        return;
    }
    auto smooth = [&](const Level& L, const float* bi, const float* x0, float* first, float* other,
                      float* final_out) -> float* {
        auto target = [&](int st) { return (st + 1 == ch.degree && final_out) ? final_out : ((st % 2) ? other : first); };
        if (x0 == nullptr) {
            const std::int64_t tot = L.a.n * K;
            jacobi_zero_kernel<K><<<grid_for(tot), kBlock, 0, s>>>(L.a.n, L.dinv.get(), bi, ch.alpha0, target(0), state);
        } else {
            launch_smoother<K>(L.a, L.dinv.get(), bi, x0, x0, 1.0f, 0.0f, ch.alpha0, target(0), 1, state, s);
        }
        float* cur = target(0);
        const float* prev = (x0 == nullptr) ? cur : other;
        for (int st = 1; st < ch.degree; ++st) {
            float* out = target(st);
            const float cprev = (st == 1 && x0 == nullptr) ? 0.0f : ch.c_prev[st - 1];
            launch_smoother<K>(L.a, L.dinv.get(), bi, cur, prev, ch.c_cur[st - 1], cprev, ch.alpha[st - 1], out, 1,
                               state, s);
            prev = cur;
            cur = out;
        }
        return cur;
    };
    const bool odd = (ch.degree % 2) != 0;
    for (int i = 0; i < nl; ++i) {
        const Level& L = h.levels[i];
        const float* bi = (i == 0) ? B : lb[i].get();
        smooth(L, bi, nullptr, odd ? lx[i].get() : lr[i].get(), odd ? lr[i].get() : lx[i].get(), nullptr);
        launch_smoother<K>(L.a, L.dinv.get(), bi, lx[i].get(), lx[i].get(), 0.f, 0.f, 0.f, lr[i].get(), 0, state, s);
        float* next_b = (i + 1 < nl) ? lb[i + 1].get() : cb;
        restrict_kernel<K><<<grid_for(L.n_coarse, kRowWaves), kRowThreads, 0, s>>>(
            L.n_coarse, L.r.ptr.get(), L.r.col.get(), L.r.val.get(), lr[i].get(), next_b, state);
    }
    coarse_kernel<K><<<grid_for(h.n_coarse(), kRowWaves), kRowThreads, 0, s>>>(static_cast<int>(h.n_coarse()),
                                                                                h.coarse_inv.get(), cb, cx, state);
    const float* xc = cx;
    for (int i = nl - 1; i >= 0; --i) {
        const Level& L = h.levels[i];
        const float* bi = (i == 0) ? B : lb[i].get();
        const unsigned g = xcd_grid((L.a.n + rows_per_row_block<K>() - 1) / rows_per_row_block<K>());
        prolongate_kernel<K><<<g, kRowThreads, 0, s>>>(L.a.n, L.p.ptr.get(), L.p.col.get(), L.p.val.get(), xc,
                                                       lx[i].get(), state);
        xc = smooth(L, bi, lx[i].get(), lr[i].get(), lx[i].get(), (i == 0) ? X : nullptr);
    }
}

template <int K>
void BlockSolver::vcycle_k(const float* B, float* X) {
    // The coarse buffers are FP64-sized; the V-cycle uses them as FP32 storage.
    run_vcycle<K>(h_, lb_, lx_, lr_, reinterpret_cast<float*>(cb_.get()), reinterpret_cast<float*>(cx_.get()), B, X,
                  nullptr, s_);
}

static std::int64_t gram_rows(std::int64_t n) { return (n + kGramBlocks - 1) / kGramBlocks; }

template <int K>
void BlockSolver::iteration_k(const SolveOptions& opt) {
    const std::int64_t nn = n();
    const std::int64_t rpb = gram_rows(nn);
    const int* st = state_.get();
    int* stw = state_.get();
    // S = A W, G partials
    spmm_kernel<K, true><<<kGramBlocks, kGramThreads, 0, s_>>>(nn, op_.a32.ptr, op_.a32.col, op_.a32.val, W_.get(),
                                                               S_.get(), partials_.get(), rpb, st);
    chol_kernel<K><<<1, kSmallThreads, 0, s_>>>(partials_.get(), kGramBlocks, opt.drop_tol, G_.get(), T_.get(), stw);
    form_pq_kernel<K><<<kGramBlocks, kGramThreads, 0, s_>>>(nn, W_.get(), S_.get(), T_.get(), R_.get(), P_.get(),
                                                            Q_.get(), partials_.get(), rpb, st);
    matrix_kernel<K><<<1, kSmallThreads, 0, s_>>>(partials_.get(), kGramBlocks, alpha_.get(), stw);
    update_xr_kernel<K><<<kGramBlocks, kGramThreads, 0, s_>>>(nn, P_.get(), Q_.get(), alpha_.get(), X64_.get(), R_.get(),
                                                              partials_.get(), rpb, st);
    converge_kernel<K><<<1, kSmallThreads, 0, s_>>>(partials_.get(), kGramBlocks, target_.get(), norms_.get(), stw);
    run_vcycle<K>(h_, lb_, lx_, lr_, reinterpret_cast<float*>(cb_.get()), reinterpret_cast<float*>(cx_.get()),
                  R_.get(), Z_.get(), st, s_);
    gram_kernel<K><<<kGramBlocks, kGramThreads, 0, s_>>>(nn, Q_.get(), Z_.get(), partials_.get(), rpb, st);
    matrix_kernel<K><<<1, kSmallThreads, 0, s_>>>(partials_.get(), kGramBlocks, C_.get(), stw);
    update_w_kernel<K><<<grid_for(nn * lanes_per_row<K>(), kRowThreads), kRowThreads, 0, s_>>>(
        nn, Z_.get(), P_.get(), C_.get(), W_.get(), st);
}

template <int K>
SolveStats BlockSolver::solve_k(const float* B, double* X, const double* X0, int k, const SolveOptions& opt) {
    ensure(K);
    const std::int64_t nn = n();
    const std::int64_t nk = nn * K;
    const std::int64_t rpb = gram_rows(nn);
    SolveStats stats;
    stats.column_iterations.assign(k, 0);
    stats.relative_residual.assign(k, 0.0);
    double* const Xi = X64_.get();  // solver-owned, so a captured graph never names a caller buffer

    // ||b_c||: the FP64 residual kernel with X = 0 gives R64 = B and its column norms.
    std::vector<double> bnorm(K), rnorm(K);
    auto true_residual = [&](bool zero) {
        residual64_kernel<K><<<kGramBlocks, kGramThreads, 0, s_>>>(nn, op_.a32.ptr, op_.a32.col, op_.val64, B, Xi,
                                                                   R64_.get(), partials_.get(), rpb, zero);
        column_norms_kernel<K><<<1, kSmallThreads, 0, s_>>>(partials_.get(), kGramBlocks, norms_.get());
        d2h(rnorm.data(), norms_.get(), K, s_);
        sync(s_, "true residual");
    };
    true_residual(true);  // x_zero: R64 = B, so these are ||b_c||
    bnorm = rnorm;
    if (X0 != nullptr) {
        d2d(Xi, X0, static_cast<std::size_t>(nk), s_);
        true_residual(false);
    } else {
        zero_f64_kernel<<<grid_for(nk), kBlock, 0, s_>>>(nk, Xi);
    }
    for (int c = 0; c < k; ++c) require(bnorm[c] > 0.0, "zero right-hand-side column " + std::to_string(c));

    std::vector<int> host_state(2 + 2 * kMaxBlock, 0);
    int total_iters = 0;
    std::vector<int> col_iters(K, 0);
    for (int round = 0; round < opt.max_refine; ++round) {
        bool all_ok = true;
        for (int c = 0; c < k; ++c) {
            stats.relative_residual[c] = rnorm[c] / bnorm[c];
            if (!(stats.relative_residual[c] <= opt.tolerance)) all_ok = false;
        }
        if (all_ok || total_iters >= opt.max_iters) break;
        ++stats.refinements;
        // Inner solve on the current true residual.
        std::vector<double> tgt(K, 0.0);
        std::fill(host_state.begin(), host_state.end(), 0);
        host_state[1] = total_iters;
        bool any_active = false;
        for (int c = 0; c < K; ++c) {
            const bool active = c < k && !(stats.relative_residual[c] <= opt.tolerance);
            tgt[c] = active ? opt.inner_margin * opt.tolerance * bnorm[c] : 0.0;
            host_state[2 + c] = active ? 0 : 1;
            any_active |= active;
        }
        if (!any_active) break;
        h2d(target_.get(), tgt.data(), K, s_);
        h2d(state_.get(), host_state.data(), host_state.size(), s_);
        cast_f64_f32_kernel<<<grid_for(nk), kBlock, 0, s_>>>(nk, R64_.get(), R_.get());
        run_vcycle<K>(h_, lb_, lx_, lr_, reinterpret_cast<float*>(cb_.get()), reinterpret_cast<float*>(cx_.get()),
                      R_.get(), Z_.get(), nullptr, s_);
        ++stats.vcycles;
        d2d(W_.get(), Z_.get(), static_cast<std::size_t>(nk), s_);
        if (opt.use_graph && (graph_ == nullptr || graph_k_ != K)) {
            if (graph_) {
                (void)hipGraphExecDestroy(graph_);
                graph_ = nullptr;
            }
            hipGraph_t g = nullptr;
            check(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal), "begin capture");
            iteration_k<K>(opt);
            check(hipStreamEndCapture(s_, &g), "end capture");
            check(hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0), "graph instantiate");
            (void)hipGraphDestroy(g);
            graph_k_ = K;
        }
        // Bounded by the launches the host issues, not only by the device counter.
        int launched = total_iters;
        while (launched < opt.max_iters) {
            const int batch = std::min(opt.check_every, opt.max_iters - launched);
            for (int b = 0; b < batch; ++b) {
                if (opt.use_graph)
                    check(hipGraphLaunch(graph_, s_), "graph launch");
                else
                    iteration_k<K>(opt);
            }
            launched += batch;
            d2h(host_state_, state_.get(), 2 + 2 * K, s_);
            sync(s_, "convergence read");
            total_iters = host_state_[1];
            if (host_state_[0]) break;
        }
        total_iters = std::max(total_iters, std::min(launched, opt.max_iters));
        if (!host_state_[0]) total_iters = launched;
        for (int c = 0; c < K; ++c)
            if (host_state[2 + c] == 0 && host_state_[2 + K + c] > 0) col_iters[c] = host_state_[2 + K + c];
        stats.vcycles += total_iters - host_state[1];
        true_residual(false);
    }
    for (int c = 0; c < k; ++c) {
        stats.relative_residual[c] = rnorm[c] / bnorm[c];
        stats.column_iterations[c] = col_iters[c];
    }
    stats.iterations = total_iters;
    d2d(X, Xi, static_cast<std::size_t>(nk), s_);
    sync(s_, "solution copy");
    return stats;
}

SolveStats BlockSolver::solve(const float* B, double* X, const double* X0, int k, const SolveOptions& opt) {
    require(k >= 1 && k <= kMaxBlock, "block width must be in [1, 32]");
    switch (padded_width(k)) {
        case 1: return solve_k<1>(B, X, X0, k, opt);
        case 2: return solve_k<2>(B, X, X0, k, opt);
        case 4: return solve_k<4>(B, X, X0, k, opt);
        case 8: return solve_k<8>(B, X, X0, k, opt);
        case 16: return solve_k<16>(B, X, X0, k, opt);
        default: return solve_k<32>(B, X, X0, k, opt);
    }
}

void BlockSolver::spmm(const float* X, float* Y, int K) {
    const std::int64_t nn = n();
    const std::int64_t rpb = gram_rows(nn);
    switch (K) {
        case 1: spmm_kernel<1, false><<<kGramBlocks, kGramThreads, 0, s_>>>(nn, op_.a32.ptr, op_.a32.col, op_.a32.val, X, Y, nullptr, rpb, nullptr); break;
        case 2: spmm_kernel<2, false><<<kGramBlocks, kGramThreads, 0, s_>>>(nn, op_.a32.ptr, op_.a32.col, op_.a32.val, X, Y, nullptr, rpb, nullptr); break;
        case 4: spmm_kernel<4, false><<<kGramBlocks, kGramThreads, 0, s_>>>(nn, op_.a32.ptr, op_.a32.col, op_.a32.val, X, Y, nullptr, rpb, nullptr); break;
        case 8: spmm_kernel<8, false><<<kGramBlocks, kGramThreads, 0, s_>>>(nn, op_.a32.ptr, op_.a32.col, op_.a32.val, X, Y, nullptr, rpb, nullptr); break;
        case 16: spmm_kernel<16, false><<<kGramBlocks, kGramThreads, 0, s_>>>(nn, op_.a32.ptr, op_.a32.col, op_.a32.val, X, Y, nullptr, rpb, nullptr); break;
        case 32: spmm_kernel<32, false><<<kGramBlocks, kGramThreads, 0, s_>>>(nn, op_.a32.ptr, op_.a32.col, op_.a32.val, X, Y, nullptr, rpb, nullptr); break;
        default: throw InvalidArgument("SpMM width must be one of 1, 2, 4, 8, 16, 32");
    }
}

namespace {
__global__ void fill_pattern_kernel(std::int64_t count, float* __restrict__ out) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) out[i] = 1.0f + 0.5f * sinf(0.001f * static_cast<float>(i % 100003));
}
}  // namespace

double BlockSolver::bench_spmm(int K, int reps, bool fused_gram) {
    require(padded_width(K) == K, "K must be a compiled block width");
    ensure(K);
    const std::int64_t nn = n(), nk = nn * K, rpb = gram_rows(nn);
    fill_pattern_kernel<<<grid_for(nk), kBlock, 0, s_>>>(nk, W_.get());
    auto launch = [&] {
        if (!fused_gram) {
            spmm(W_.get(), S_.get(), K);
            return;
        }
        const auto& a = op_.a32;
        switch (K) {
            case 1: spmm_kernel<1, true><<<kGramBlocks, kGramThreads, 0, s_>>>(nn, a.ptr, a.col, a.val, W_.get(), S_.get(), partials_.get(), rpb, nullptr); break;
            case 2: spmm_kernel<2, true><<<kGramBlocks, kGramThreads, 0, s_>>>(nn, a.ptr, a.col, a.val, W_.get(), S_.get(), partials_.get(), rpb, nullptr); break;
            case 4: spmm_kernel<4, true><<<kGramBlocks, kGramThreads, 0, s_>>>(nn, a.ptr, a.col, a.val, W_.get(), S_.get(), partials_.get(), rpb, nullptr); break;
            case 8: spmm_kernel<8, true><<<kGramBlocks, kGramThreads, 0, s_>>>(nn, a.ptr, a.col, a.val, W_.get(), S_.get(), partials_.get(), rpb, nullptr); break;
            case 16: spmm_kernel<16, true><<<kGramBlocks, kGramThreads, 0, s_>>>(nn, a.ptr, a.col, a.val, W_.get(), S_.get(), partials_.get(), rpb, nullptr); break;
            default: spmm_kernel<32, true><<<kGramBlocks, kGramThreads, 0, s_>>>(nn, a.ptr, a.col, a.val, W_.get(), S_.get(), partials_.get(), rpb, nullptr); break;
        }
    };
    launch();  // warm
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "event");
    check(hipEventCreate(&e1), "event");
    check(hipEventRecord(e0, s_), "event record");
    for (int r = 0; r < reps; ++r) launch();
    check(hipEventRecord(e1, s_), "event record");
    check(hipEventSynchronize(e1), "event sync");
    float ms = 0.f;
    check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return static_cast<double>(ms) / reps;
}

double BlockSolver::bench_vcycle(int K, int reps) {
    require(padded_width(K) == K, "K must be a compiled block width");
    ensure(K);
    const std::int64_t nk = n() * K;
    fill_pattern_kernel<<<grid_for(nk), kBlock, 0, s_>>>(nk, R_.get());
    vcycle(R_.get(), Z_.get(), K);
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "event");
    check(hipEventCreate(&e1), "event");
    check(hipEventRecord(e0, s_), "event record");
    for (int r = 0; r < reps; ++r) vcycle(R_.get(), Z_.get(), K);
    check(hipEventRecord(e1, s_), "event record");
    check(hipEventSynchronize(e1), "event sync");
    float ms = 0.f;
    check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return static_cast<double>(ms) / reps;
}

void BlockSolver::vcycle(const float* B, float* X, int K) {
    ensure(K);
    switch (K) {
        case 1: vcycle_k<1>(B, X); break;
        case 2: vcycle_k<2>(B, X); break;
        case 4: vcycle_k<4>(B, X); break;
        case 8: vcycle_k<8>(B, X); break;
        case 16: vcycle_k<16>(B, X); break;
        case 32: vcycle_k<32>(B, X); break;
        default: throw InvalidArgument("V-cycle width must be one of 1, 2, 4, 8, 16, 32");
    }
}

}  // namespace cn
