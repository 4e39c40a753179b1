// Block PCG for k right-hand sides sharing one operator (solver/amg/solver.cpp:84-198 solves the
// same systems as k lockstep scalar chains; this is a true block method instead).
//
// Precision. The operator, the V-cycle and every (n, K) block vector are FP32. FP64 holds the
// solution X, the K x K Gram matrices and their pivoted Cholesky, the block coefficients, and the
// true residual B - A X that an outer iterative-refinement loop recomputes with the FP64 operator.
// Convergence is judged on that true residual, per column, against ||b|| (solver.cpp:237-246).
//
// Iteration, written on the pair (W, S = A W) so that no FP32 rounding of a derived search block
// enters the recursive residual:
//     S = A W                         (differenced SpMM, fused with G = W^T S and H = W^T R)
//     T: pivoted Cholesky of G, T^T G T = I on the retained directions (P = W T is A-orthonormal)
//     a = T T^T H                     (K x K; zero in the columns that have frozen)
//     X += W a,  R -= S a
//     Z = M R                         (one V-cycle)
//     E = S^T Z,  W <- Z - W (T T^T E)
// This is block CG with A-orthonormal blocks P = W T (X += P (P^T R), W' = Z - P (P^T A Z)),
// breakdown-free because directions whose Schur pivot falls below drop_tol are dropped.
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

// Lanes per row of the fine-level SpMM: K lanes (one per column), at least 8.
template <int K>
constexpr int spmm_lanes() {
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

// Sum over the row groups of a wave that share a lane-in-row index (xor over L, 2L, ..., 32).
template <int L, typename T>
__device__ __forceinline__ T wave_rows_sum(T v) {
#pragma unroll
    for (int off = L; off < kWave; off <<= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Wave-ordered accumulation of per-wave K-vectors into one block partial: `g[a]` of lane c < K
// is entry (a, c).
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
    __syncthreads();
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

// sum_{j in [j, e) step SUB} av[j] * X[ci[j] * K + c], four gathers in flight per step; the fold
// order is the plain loop's.
template <int K, int SUB>
__device__ __forceinline__ float row_gather(const std::int32_t* __restrict__ ci, const float* __restrict__ av,
                                            int j, int e, const float* __restrict__ X, int c) {
    float acc = 0.0f;
    for (; j + 3 * SUB < e; j += 4 * SUB) {
        const std::int64_t c0 = ci[j], c1 = ci[j + SUB], c2 = ci[j + 2 * SUB], c3 = ci[j + 3 * SUB];
        const float a0 = av[j], a1 = av[j + SUB], a2 = av[j + 2 * SUB], a3 = av[j + 3 * SUB];
        const float x0 = X[c0 * K + c], x1 = X[c1 * K + c], x2 = X[c2 * K + c], x3 = X[c3 * K + c];
        acc = fmaf(a0, x0, acc);
        acc = fmaf(a1, x1, acc);
        acc = fmaf(a2, x2, acc);
        acc = fmaf(a3, x3, acc);
    }
    for (; j < e; j += SUB) acc = fmaf(av[j], X[static_cast<std::int64_t>(ci[j]) * K + c], acc);
    return acc;
}

// The same over the differenced off-diagonal operator: sum av[j] * (X[ci[j]] - xi).
template <int K, int SUB>
__device__ __forceinline__ float row_gather_diff(const std::int32_t* __restrict__ ci, const float* __restrict__ av,
                                                 int j, int e, const float* __restrict__ X, int c, float xi) {
    float acc = 0.0f;
    for (; j + 3 * SUB < e; j += 4 * SUB) {
        const std::int64_t c0 = ci[j], c1 = ci[j + SUB], c2 = ci[j + 2 * SUB], c3 = ci[j + 3 * SUB];
        const float a0 = av[j], a1 = av[j + SUB], a2 = av[j + 2 * SUB], a3 = av[j + 3 * SUB];
        const float x0 = X[c0 * K + c], x1 = X[c1 * K + c], x2 = X[c2 * K + c], x3 = X[c3 * K + c];
        acc = fmaf(a0, x0 - xi, acc);
        acc = fmaf(a1, x1 - xi, acc);
        acc = fmaf(a2, x2 - xi, acc);
        acc = fmaf(a3, x3 - xi, acc);
    }
    for (; j < e; j += SUB) acc = fmaf(av[j], X[static_cast<std::int64_t>(ci[j]) * K + c] - xi, acc);
    return acc;
}

constexpr int kGramThreads = 1024;
constexpr int kGramWaves = kGramThreads / kWave;

// ---------------------------------------------------------------------------------------------
// Plain SpMM: Y = A X over the FP32 CSR, X and Y (n, K) row-major.
// ---------------------------------------------------------------------------------------------
template <int K>
__global__ void __launch_bounds__(kGramThreads) spmm_kernel(std::int64_t n, const std::int32_t* __restrict__ rp,
                                                            const std::int32_t* __restrict__ ci,
                                                            const float* __restrict__ av, const float* __restrict__ X,
                                                            float* __restrict__ Y, std::int64_t rows_per_block) {
    constexpr int L = spmm_lanes<K>();
    constexpr int SUB = L / K;
    constexpr int RPW = kWave / L;
    const int lb = logical_block();
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
    const int l = lane % L, c = l % K, sub = l / K;
    const std::int64_t lo = static_cast<std::int64_t>(lb) * rows_per_block;
    const std::int64_t hi = min(lo + rows_per_block, n);
    for (std::int64_t r0 = lo + static_cast<std::int64_t>(wave) * RPW; r0 < hi; r0 += kGramWaves * RPW) {
        const std::int64_t row = r0 + lane / L;
        const bool valid = row < hi;
        float acc = valid ? row_gather<K, SUB>(ci, av, rp[row] + sub, rp[row + 1], X, c) : 0.0f;
#pragma unroll
        for (int off = K; off < L; off <<= 1) acc += __shfl_xor(acc, off, kWave);
        if (valid && sub == 0) Y[row * K + c] = acc;
    }
}

// ---------------------------------------------------------------------------------------------
// Block-PCG SpMM: S = A W over the differenced operator, with the FP64 partials of G = W^T S and
// H = W^T R in one pass. partials[lb] = [G (K x K) | H (K x K)]. With FUSED false only S is formed
// (the differenced SpMM on its own).
// ---------------------------------------------------------------------------------------------
template <int K, bool FUSED>
__global__ void __launch_bounds__(kGramThreads) spmm_block_kernel(
    std::int64_t n, const std::int32_t* __restrict__ rp, const std::int32_t* __restrict__ ci,
    const float* __restrict__ av, const float* __restrict__ rowsum, const float* __restrict__ W,
    const float* __restrict__ R, float* __restrict__ S, double* __restrict__ partials,
    std::int64_t rows_per_block, const int* __restrict__ state) {
    if (finished(state)) return;
    constexpr int L = spmm_lanes<K>();
    constexpr int SUB = L / K;
    constexpr int RPW = kWave / L;
    const int lb = logical_block();
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
    const int l = lane % L, c = l % K, sub = l / K;
    const int group = lane - l;
    const std::int64_t lo = static_cast<std::int64_t>(lb) * rows_per_block;
    const std::int64_t hi = min(lo + rows_per_block, n);
    double g[K], h[K];
    if constexpr (FUSED) {
#pragma unroll
        for (int a = 0; a < K; ++a) g[a] = h[a] = 0.0;
    }
    for (std::int64_t r0 = lo + static_cast<std::int64_t>(wave) * RPW; r0 < hi; r0 += kGramWaves * RPW) {
        const std::int64_t row = r0 + lane / L;
        const bool valid = row < hi;
        const std::int64_t at = row * K + c;
        const float xi = valid ? W[at] : 0.0f;
        float acc = valid ? row_gather_diff<K, SUB>(ci, av, rp[row] + sub, rp[row + 1], W, c, xi) : 0.0f;
#pragma unroll
        for (int off = K; off < L; off <<= 1) acc += __shfl_xor(acc, off, kWave);
        const float s = valid ? fmaf(rowsum[row], xi, acc) : 0.0f;
        if (valid && sub == 0) S[at] = s;
        if constexpr (FUSED) {
            const bool own = valid && sub == 0;
            const double sd = own ? static_cast<double>(s) : 0.0;
            const double rd = own ? static_cast<double>(R[at]) : 0.0;
#pragma unroll
            for (int a = 0; a < K; ++a) {
                const double wa = static_cast<double>(__shfl(xi, group + a, kWave));
                g[a] = fma(wa, sd, g[a]);
                h[a] = fma(wa, rd, h[a]);
            }
        }
    }
    if constexpr (FUSED) {
#pragma unroll
        for (int a = 0; a < K; ++a) {
            g[a] = wave_rows_sum<L>(g[a]);
            h[a] = wave_rows_sum<L>(h[a]);
        }
        double* out = partials + static_cast<std::int64_t>(lb) * 2 * K * K;
        block_gram_store<K, kGramWaves>(g, out);
        block_gram_store<K, kGramWaves>(h, out + K * K);
    }
}

// FP64 true residual R64 = B - A X64 with the FP64 operator, and per-column ||R64||^2 partials.
template <int K>
__global__ void __launch_bounds__(kGramThreads) residual64_kernel(
    std::int64_t n, const std::int32_t* __restrict__ rp, const std::int32_t* __restrict__ ci,
    const double* __restrict__ av, const float* __restrict__ B, const double* __restrict__ X,
    double* __restrict__ R, double* __restrict__ partials, std::int64_t rows_per_block, bool x_zero) {
    constexpr int L = spmm_lanes<K>();
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
// V-cycle kernels (solver/amg/vcycle.cu:124-311), K-wide, FP32. SUB lanes per column split each
// row's nonzeros, chosen per level from its shape (wide coarse rows take more lanes).
// ---------------------------------------------------------------------------------------------
constexpr int kRowThreads = 256;
constexpr int kRowWaves = kRowThreads / kWave;
constexpr int kRowIters = 4;  // row groups each wave walks per block

template <int K, int SUB>
constexpr std::int64_t rows_per_row_block() {
    return static_cast<std::int64_t>(kRowWaves) * (kWave / (K * SUB)) * kRowIters;
}

// MODE 0: residual  out = b - A x
// MODE 1: Chebyshev out = c_cur x + c_prev x_prev + alpha dinv (b - A x); out may alias x_prev.
template <int K, int SUB, int MODE>
__global__ void __launch_bounds__(kRowThreads) smoother_kernel(
    std::int64_t n, const std::int32_t* __restrict__ rp, const std::int32_t* __restrict__ ci,
    const float* __restrict__ av, const float* __restrict__ dinv, const float* __restrict__ b,
    const float* __restrict__ x, const float* x_prev, float c_cur, float c_prev, float alpha, float* out,
    const int* __restrict__ state) {
    if (finished(state)) return;
    constexpr int L = K * SUB;
    constexpr int RPW = kWave / L;
    const int lb = logical_block();
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
    const int l = lane % L, c = l % K, sub = l / K;
    const std::int64_t lo = static_cast<std::int64_t>(lb) * rows_per_row_block<K, SUB>();
    const std::int64_t hi = min(lo + rows_per_row_block<K, SUB>(), n);
    for (std::int64_t r0 = lo + static_cast<std::int64_t>(wave) * RPW; r0 < hi; r0 += kRowWaves * RPW) {
        const std::int64_t row = r0 + lane / L;
        const bool valid = row < hi;
        float acc = valid ? row_gather<K, SUB>(ci, av, rp[row] + sub, rp[row + 1], x, c) : 0.0f;
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
    float acc = (row < n_coarse) ? row_gather<K, SUB>(ci, rv, rp[row] + sub, rp[row + 1], fine, c) : 0.0f;
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
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n * K) return;
    const std::int64_t row = t / K;
    const int c = static_cast<int>(t % K);
    fine[t] += row_gather<K, 1>(ci, pv, rp[row], rp[row + 1], coarse, c);
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

// X64 += W a (FP64), R -= S a (FP32), partials of ||R_c||^2. Frozen columns have a zero column
// in `a`, which leaves X and R exactly unchanged.
template <int K>
__global__ void __launch_bounds__(kGramThreads) update_xr_kernel(
    std::int64_t n, const float* __restrict__ W, const float* __restrict__ S, const double* __restrict__ a,
    double* __restrict__ X, float* __restrict__ R, double* __restrict__ partials, std::int64_t rows_per_block,
    const int* __restrict__ state) {
    if (finished(state)) return;
    constexpr int L = spmm_lanes<K>();
    constexpr int RPW = kWave / L;
    constexpr int SUB = L / K;
    __shared__ double al[K * K];
    __shared__ float alf[K * K];
    for (int e = threadIdx.x; e < K * K; e += blockDim.x) {
        al[e] = a[e];
        alf[e] = static_cast<float>(a[e]);
    }
    __syncthreads();
    const int lb = logical_block();
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
    const int l = lane % L, c = l % K, sub = l / K;
    const int group = lane - l;
    const std::int64_t lo = static_cast<std::int64_t>(lb) * rows_per_block;
    const std::int64_t hi = min(lo + rows_per_block, n);
    double nrm = 0.0;
    for (std::int64_t r0 = lo + static_cast<std::int64_t>(wave) * RPW; r0 < hi; r0 += kGramWaves * RPW) {
        const std::int64_t row = r0 + lane / L;
        const bool valid = row < hi;
        const std::int64_t at = row * K + c;
        const float w_own = valid ? W[at] : 0.0f;
        const float s_own = valid ? S[at] : 0.0f;
        double dx = 0.0;
        float dr = 0.0f;
#pragma unroll
        for (int b = 0; b < K; ++b) {
            const float wb = __shfl(w_own, group + b, kWave);
            const float sb = __shfl(s_own, group + b, kWave);
            dx = fma(static_cast<double>(wb), al[b * K + c], dx);
            dr = fmaf(sb, alf[b * K + c], dr);
        }
        if (valid && sub == 0) {
            X[at] += dx;
            const float r = R[at] - dr;
            R[at] = r;
            nrm = fma(static_cast<double>(r), static_cast<double>(r), nrm);
        }
    }
    nrm = wave_rows_sum<L>(nrm);
    block_vec_store<K, kGramWaves>(nrm, partials + static_cast<std::int64_t>(lb) * K);
}

// partials of E = S^T Z.
template <int K>
__global__ void __launch_bounds__(kGramThreads) gram_kernel(std::int64_t n, const float* __restrict__ A,
                                                            const float* __restrict__ B,
                                                            double* __restrict__ partials,
                                                            std::int64_t rows_per_block,
                                                            const int* __restrict__ state) {
    if (finished(state)) return;
    constexpr int L = spmm_lanes<K>();
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

// W <- Z - W D, in place (each lane reads its row's W before any lane of the row writes it).
template <int K>
__global__ void __launch_bounds__(kRowThreads) update_w_kernel(std::int64_t n, const float* __restrict__ Z,
                                                               const double* __restrict__ D, float* W,
                                                               const int* __restrict__ state) {
    if (finished(state)) return;
    constexpr int L = spmm_lanes<K>();
    __shared__ float dm[K * K];
    for (int e = threadIdx.x; e < K * K; e += blockDim.x) dm[e] = static_cast<float>(D[e]);
    __syncthreads();
    const int lane = threadIdx.x & (kWave - 1);
    const int l = lane % L, c = l % K, sub = l / K;
    const int group = lane - l;
    const std::int64_t row = (static_cast<std::int64_t>(blockIdx.x) * kRowThreads + threadIdx.x) / L;
    const bool valid = row < n;
    const std::int64_t at = row * K + c;
    const float w_own = valid ? W[at] : 0.0f;
    float acc = valid ? Z[at] : 0.0f;
#pragma unroll
    for (int b = 0; b < K; ++b) {
        const float wb = __shfl(w_own, group + b, kWave);
        acc = fmaf(-wb, dm[b * K + c], acc);
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

__global__ void fill_pattern_kernel(std::int64_t count, float* __restrict__ out) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) out[i] = 1.0f + 0.5f * sinf(0.001f * static_cast<float>(i % 100003));
}

// ---------------------------------------------------------------------------------------------
// One-block kernels: fixed-order reduction of the chunk partials, then the K x K work.
// ---------------------------------------------------------------------------------------------
constexpr int kSmallThreads = 1024;

// out[e] = sum_b partials[b * stride + off + e], b ascending: chunks of blocks, then chunks in order.
template <int E>
__device__ void reduce_partials(const double* __restrict__ partials, int nb, int stride, int off, double* out_lds) {
    constexpr int Q = (kSmallThreads / E) > 0 ? (kSmallThreads / E) : 1;
    __shared__ double chunk[Q * E];
    const int t = threadIdx.x;
    if (t < Q * E) {
        const int e = t % E, q = t / E;
        const int b0 = (nb * q) / Q, b1 = (nb * (q + 1)) / Q;
        double s = 0.0;
        for (int b = b0; b < b1; ++b) s += partials[static_cast<std::int64_t>(b) * stride + off + e];
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

// G, H -> T (pivoted Cholesky of the unit-diagonal scaling of G), TT = T T^T, a = TT H with the
// frozen columns zeroed.
template <int K>
__global__ void __launch_bounds__(kSmallThreads) coef_kernel(const double* __restrict__ partials, int nb,
                                                             double drop_tol, double* __restrict__ G_out,
                                                             double* __restrict__ T, double* __restrict__ TT,
                                                             double* __restrict__ AL, int* __restrict__ state) {
    if (finished(state)) return;
    __shared__ double A[K * K];   // scaled G, then its Schur complements; then Lp
    __shared__ double H[K * K];
    __shared__ double Lc[K * K];  // Lc[j * K + i]: column j of L at original row i; then T T^T
    __shared__ double sc[K];
    __shared__ double Li[K * K];
    __shared__ double Tm[K * K];
    double* const Lp = A;   // A is dead once the factorisation stops
    double* const TTm = Lc; // Lc is dead once Lp has been gathered from it
    __shared__ int idx[K];
    __shared__ int rem[K];
    __shared__ int rank_s, stop;
    reduce_partials<K * K>(partials, nb, 2 * K * K, 0, A);
    reduce_partials<K * K>(partials, nb, 2 * K * K, K * K, H);
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
        Tm[t] = 0.0;
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
    if (t < K * K) {
        const int a = t / K, b = t % K;
        Lp[t] = (a < r && b < r && b <= a) ? Lc[b * K + idx[a]] : 0.0;
        Li[t] = 0.0;
    }
    __syncthreads();
    if (t < r) {  // Li = Lp^-1, thread b solves Lp y = e_b
        const int b = t;
        Li[b * K + b] = 1.0 / Lp[b * K + b];
        for (int a = b + 1; a < r; ++a) {
            double s = 0.0;
            for (int q = b; q < a; ++q) s += Lp[a * K + q] * Li[q * K + b];
            Li[a * K + b] = -s / Lp[a * K + a];
        }
    }
    __syncthreads();
    if (t < K * K) {  // T[idx[a]][b] = sc[idx[a]] * Li[b][a]
        const int a = t / K, b = t % K;
        if (a < r && b < r && b >= a) Tm[idx[a] * K + b] = sc[idx[a]] * Li[b * K + a];
    }
    __syncthreads();
    if (t < K * K) {  // TT = T T^T
        const int i = t / K, j = t % K;
        double s = 0.0;
        for (int b = 0; b < K; ++b) s += Tm[i * K + b] * Tm[j * K + b];
        TTm[t] = s;
        T[t] = Tm[t];
        TT[t] = s;
    }
    __syncthreads();
    if (t < K * K) {  // a = TT H, frozen columns zero
        const int i = t / K, c = t % K;
        double s = 0.0;
        for (int b = 0; b < K; ++b) s += TTm[i * K + b] * H[b * K + c];
        AL[t] = state[2 + c] ? 0.0 : s;
    }
}

// D = TT E.
template <int K>
__global__ void __launch_bounds__(kSmallThreads) dmat_kernel(const double* __restrict__ partials, int nb,
                                                             const double* __restrict__ TT, double* __restrict__ D,
                                                             int* __restrict__ state) {
    if (finished(state)) return;
    __shared__ double E[K * K];
    reduce_partials<K * K>(partials, nb, K * K, 0, E);
    const int t = threadIdx.x;
    if (t < K * K) {
        const int i = t / K, c = t % K;
        double s = 0.0;
        for (int b = 0; b < K; ++b) s += TT[i * K + b] * E[b * K + c];
        D[t] = s;
    }
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
    reduce_partials<K>(partials, nb, K, 0, v);
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
    reduce_partials<K>(partials, nb, K, 0, v);
    if (threadIdx.x < K) out[threadIdx.x] = sqrt(v[threadIdx.x]);
}

std::int64_t gram_rows(std::int64_t n) { return (n + kGramBlocks - 1) / kGramBlocks; }

// Lanes per column for a level's row kernels: enough that each lane walks a few nonzeros, more
// when the level is too small to fill the machine (solver/amg/vcycle.cu:104-120 makes the same
// trade for its threads-per-row).
int choose_sub(std::int64_t n_rows, std::int64_t nnz, int K, int n_cu) {
    const std::int64_t per_row = n_rows > 0 ? (nnz + n_rows - 1) / n_rows : 1;
    int sub = 1;
    while (K * sub * 2 <= kWave) {
        const bool pays = per_row >= 4 * (sub * 2);
        const bool starved = n_rows * K * sub < static_cast<std::int64_t>(n_cu) * 8 * kWave;
        if (!pays && !starved) break;
        sub *= 2;
    }
    return sub;
}

template <int K, int MODE>
void launch_smoother_sub(int sub, const CsrView& a, const float* dinv, const float* b, const float* x, const float* xp,
                         float cc, float cp, float al, float* out, const int* state, hipStream_t s) {
    auto go = [&](auto subc) {
        constexpr int SUB = decltype(subc)::value;
        if constexpr (K * SUB <= kWave) {
            const unsigned g = xcd_grid((a.n + rows_per_row_block<K, SUB>() - 1) / rows_per_row_block<K, SUB>());
            if (g) smoother_kernel<K, SUB, MODE><<<g, kRowThreads, 0, s>>>(a.n, a.ptr, a.col, a.val, dinv, b, x, xp, cc, cp, al, out, state);
        }
    };
    switch (sub) {
        case 1: go(std::integral_constant<int, 1>{}); break;
        case 2: go(std::integral_constant<int, 2>{}); break;
        case 4: go(std::integral_constant<int, 4>{}); break;
        case 8: go(std::integral_constant<int, 8>{}); break;
        case 16: go(std::integral_constant<int, 16>{}); break;
        case 32: go(std::integral_constant<int, 32>{}); break;
        default: go(std::integral_constant<int, 64>{}); break;
    }
}

// One zero-guess V-cycle (solver/amg/vcycle.cu:560-648). `state` lets captured replays skip it.
template <int K>
void run_vcycle(const Hierarchy& h, const std::vector<int>& subs, std::vector<DeviceBuffer<float>>& lb,
                std::vector<DeviceBuffer<float>>& lx, std::vector<DeviceBuffer<float>>& lr, float* cb, float* cx,
                const float* B, float* X, const int* state, hipStream_t s) {
    const Cheby& ch = h.cheb;
    const int nl = static_cast<int>(h.levels.size());
    if (nl == 0) {
        coarse_kernel<K><<<grid_for(h.n_coarse(), kRowWaves), kRowThreads, 0, s>>>(
            static_cast<int>(h.n_coarse()), h.coarse_inv.get(), B, X, state);
        return;
    }
    auto smooth = [&](int li, const float* bi, const float* x0, float* first, float* other, float* final_out) -> float* {
        const Level& L = h.levels[li];
        const int sub = subs[li];
        auto target = [&](int st) { return (st + 1 == ch.degree && final_out) ? final_out : ((st % 2) ? other : first); };
        if (x0 == nullptr) {
            jacobi_zero_kernel<K><<<grid_for(L.a.n * K), kBlock, 0, s>>>(L.a.n, L.dinv.get(), bi, ch.alpha0, target(0), state);
        } else {
            launch_smoother_sub<K, 1>(sub, L.a, L.dinv.get(), bi, x0, x0, 1.0f, 0.0f, ch.alpha0, target(0), state, s);
        }
        float* cur = target(0);
        const float* prev = (x0 == nullptr) ? cur : other;
        for (int st = 1; st < ch.degree; ++st) {
            float* out = target(st);
            const float cprev = (st == 1 && x0 == nullptr) ? 0.0f : ch.c_prev[st - 1];
            launch_smoother_sub<K, 1>(sub, L.a, L.dinv.get(), bi, cur, prev, ch.c_cur[st - 1], cprev, ch.alpha[st - 1],
                                      out, state, s);
            prev = cur;
            cur = out;
        }
        return cur;
    };
    const bool odd = (ch.degree % 2) != 0;
    for (int i = 0; i < nl; ++i) {
        const Level& L = h.levels[i];
        const float* bi = (i == 0) ? B : lb[i].get();
        smooth(i, bi, nullptr, odd ? lx[i].get() : lr[i].get(), odd ? lr[i].get() : lx[i].get(), nullptr);
        launch_smoother_sub<K, 0>(subs[i], L.a, L.dinv.get(), bi, lx[i].get(), lx[i].get(), 0.f, 0.f, 0.f, lr[i].get(),
                                  state, s);
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
        prolongate_kernel<K><<<grid_for(L.a.n * K), kBlock, 0, s>>>(L.a.n, L.p.ptr.get(), L.p.col.get(), L.p.val.get(),
                                                                    xc, lx[i].get(), state);
        xc = smooth(i, bi, lx[i].get(), lr[i].get(), lx[i].get(), (i == 0) ? X : nullptr);
    }
}

template <typename F>
void dispatch_k(int K, F&& f) {
    switch (K) {
        case 1: f(std::integral_constant<int, 1>{}); break;
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 4: f(std::integral_constant<int, 4>{}); break;
        case 8: f(std::integral_constant<int, 8>{}); break;
        case 16: f(std::integral_constant<int, 16>{}); break;
        case 32: f(std::integral_constant<int, 32>{}); break;
        default: throw InvalidArgument("block width must be one of 1, 2, 4, 8, 16, 32");
    }
}

}  // namespace

// ----------------------------------------------------------------------------------------------
// Host driver
// ----------------------------------------------------------------------------------------------

BlockSolver::BlockSolver(const Operator& op, const Hierarchy& h, hipStream_t s) : op_(op), h_(h), s_(s) {
    check(hipHostMalloc(reinterpret_cast<void**>(&host_state_), sizeof(int) * (2 + 2 * kMaxBlock)),
          "pinned solver state");
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
        n_cu_ = cus;
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
    partials_.reset(static_cast<std::size_t>(kGramBlocks) * 2 * K * K, "gram partials");
    G_.reset(K * K, "G");
    T_.reset(K * K, "T");
    TT_.reset(K * K, "TT");
    AL_.reset(K * K, "alpha");
    D_.reset(K * K, "D");
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
void BlockSolver::vcycle_k(const float* B, float* X, const int* state) {
    // Lanes per column of each level's row kernels: a property of the level and the width alone.
    std::vector<int> subs(h_.levels.size());
    for (std::size_t i = 0; i < h_.levels.size(); ++i)
        subs[i] = choose_sub(h_.levels[i].a.n, h_.levels[i].a.nnz, K, n_cu_);
    run_vcycle<K>(h_, subs, lb_, lx_, lr_, cb_.get(), cx_.get(), B, X, state, s_);
}

template <int K>
void BlockSolver::spmm_block_k(const int* state) {
    const std::int64_t nn = n();
    spmm_block_kernel<K, true><<<kGramBlocks, kGramThreads, 0, s_>>>(nn, op_.off.ptr, op_.off.col, op_.off.val,
                                                                     op_.rowsum, W_.get(), R_.get(), S_.get(),
                                                                     partials_.get(), gram_rows(nn), state);
}

template <int K>
void BlockSolver::iteration_k(const SolveOptions& opt) {
    const std::int64_t nn = n();
    const std::int64_t rpb = gram_rows(nn);
    const int* st = state_.get();
    int* stw = state_.get();
    spmm_block_k<K>(st);
    coef_kernel<K><<<1, kSmallThreads, 0, s_>>>(partials_.get(), kGramBlocks, opt.drop_tol, G_.get(), T_.get(),
                                                TT_.get(), AL_.get(), stw);
    update_xr_kernel<K><<<kGramBlocks, kGramThreads, 0, s_>>>(nn, W_.get(), S_.get(), AL_.get(), X64_.get(), R_.get(),
                                                              partials_.get(), rpb, st);
    converge_kernel<K><<<1, kSmallThreads, 0, s_>>>(partials_.get(), kGramBlocks, target_.get(), norms_.get(), stw);
    vcycle_k<K>(R_.get(), Z_.get(), st);
    gram_kernel<K><<<kGramBlocks, kGramThreads, 0, s_>>>(nn, S_.get(), Z_.get(), partials_.get(), rpb, st);
    dmat_kernel<K><<<1, kSmallThreads, 0, s_>>>(partials_.get(), kGramBlocks, TT_.get(), D_.get(), stw);
    update_w_kernel<K><<<grid_for(nn * spmm_lanes<K>(), kRowThreads), kRowThreads, 0, s_>>>(nn, Z_.get(), D_.get(),
                                                                                           W_.get(), st);
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
        // Inner solve on the current true residual (a restart: W = Z).
        std::vector<double> tgt(K, 0.0);
        std::fill(host_state.begin(), host_state.end(), 0);
        host_state[1] = total_iters;
        for (int c = 0; c < K; ++c) {
            const bool active = c < k && !(stats.relative_residual[c] <= opt.tolerance);
            tgt[c] = active ? opt.inner_margin * opt.tolerance * bnorm[c] : 0.0;
            host_state[2 + c] = active ? 0 : 1;
        }
        h2d(target_.get(), tgt.data(), K, s_);
        h2d(state_.get(), host_state.data(), host_state.size(), s_);
        cast_f64_f32_kernel<<<grid_for(nk), kBlock, 0, s_>>>(nk, R64_.get(), R_.get());
        vcycle_k<K>(R_.get(), Z_.get(), nullptr);
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
        bool done = false;
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
            if (host_state_[0]) {
                done = true;
                break;
            }
        }
        if (!done) total_iters = launched;
        for (int c = 0; c < K; ++c)
            if (host_state[2 + c] == 0 && host_state_[2 + K + c] > 0) col_iters[c] = host_state_[2 + K + c];
        stats.vcycles += total_iters - host_state[1] + 1;
        true_residual(false);
        double worst = 0.0;
        for (int c = 0; c < k; ++c) worst = std::max(worst, rnorm[c] / bnorm[c]);
        if (stats.refinements < kMaxRounds) {
            stats.round_iterations[stats.refinements] = total_iters;
            stats.round_residual[stats.refinements] = worst;
        }
        ++stats.refinements;
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
    SolveStats out;
    dispatch_k(padded_width(k), [&](auto kc) { out = solve_k<decltype(kc)::value>(B, X, X0, k, opt); });
    return out;
}

void BlockSolver::spmm(const float* X, float* Y, int K) {
    const std::int64_t nn = n();
    dispatch_k(K, [&](auto kc) {
        constexpr int KK = decltype(kc)::value;
        spmm_kernel<KK><<<kGramBlocks, kGramThreads, 0, s_>>>(nn, op_.a32.ptr, op_.a32.col, op_.a32.val, X, Y,
                                                              gram_rows(nn));
    });
}

void BlockSolver::spmm_diff(const float* X, float* Y, int K) {
    const std::int64_t nn = n();
    dispatch_k(K, [&](auto kc) {
        constexpr int KK = decltype(kc)::value;
        spmm_block_kernel<KK, false><<<kGramBlocks, kGramThreads, 0, s_>>>(nn, op_.off.ptr, op_.off.col, op_.off.val,
                                                                           op_.rowsum, X, nullptr, Y, nullptr,
                                                                           gram_rows(nn), nullptr);
    });
}

void BlockSolver::vcycle(const float* B, float* X, int K) {
    ensure(K);
    dispatch_k(K, [&](auto kc) { vcycle_k<decltype(kc)::value>(B, X, nullptr); });
}

double BlockSolver::bench_spmm(int K, int reps, bool fused) {
    require(padded_width(K) == K, "K must be a compiled block width");
    ensure(K);
    const std::int64_t nk = n() * K;
    fill_pattern_kernel<<<grid_for(nk), kBlock, 0, s_>>>(nk, W_.get());
    fill_pattern_kernel<<<grid_for(nk), kBlock, 0, s_>>>(nk, R_.get());
    auto launch = [&] {
        if (fused)
            dispatch_k(K, [&](auto kc) { spmm_block_k<decltype(kc)::value>(nullptr); });
        else
            spmm(W_.get(), S_.get(), K);
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

}  // namespace cn
