// Block PCG for k right-hand sides sharing one operator (solver/amg/solver.cpp:84-198 solves the
// same systems as k lockstep scalar chains; this is a true block method instead).
//
// Precision. The operator, the V-cycle and every (n, K) block vector are FP32. FP64 holds the
// solution X, the K x K Gram matrices and their pivoted inverses, the block coefficients, and the
// true residual B - A X that an outer iterative-refinement loop recomputes with the FP64 operator.
// Convergence is judged on that true residual, per column, against ||b|| (solver.cpp:237-246).
//
// Iteration, written on the pair (W, S = A W) so that no FP32 rounding of a derived search block
// enters the recursive residual:
//     S = A W                         (differenced SpMM, fused with G = W^T S and H = W^T R)
//     TT = G^-1 on the retained directions (diagonal pivoting, pivots below drop_tol dropped)
//     a = TT H                        (K x K; zero in the columns that have frozen)
//     X += W a,  R -= S a
//     Z = M R                         (one V-cycle)
//     E = S^T Z,  W <- Z - W (TT E)
// This is block CG with A-orthonormal blocks P = W T for any T with T T^T = TT
// (X += P (P^T R), W' = Z - P (P^T A Z)),
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
#include <mutex>
#include <type_traits>

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

// One thread per ELL slot: slot j of row r takes the row's j-th CSR entry, or (r, 0) past its end.
__global__ void ell_fill_kernel(std::int64_t n, const std::int32_t* __restrict__ rp, const std::int32_t* __restrict__ ci,
                                const float* __restrict__ av, int w, int2* __restrict__ out) {
    const std::int64_t slots = (n + kEllSlice - 1) / kEllSlice * kEllSlice * w;
    for (std::int64_t t = blockIdx.x * static_cast<std::int64_t>(blockDim.x) + threadIdx.x; t < slots;
         t += static_cast<std::int64_t>(gridDim.x) * blockDim.x) {
        const std::int64_t slice = t / (static_cast<std::int64_t>(kEllSlice) * w);
        const std::int64_t rem = t - slice * kEllSlice * w;
        const int j = static_cast<int>(rem / kEllSlice);
        const std::int64_t r = slice * kEllSlice + rem % kEllSlice;
        int2 pv = make_int2(static_cast<int>(r < n ? r : 0), __float_as_int(0.0f));
        if (r < n) {
            const int p = rp[r] + j;
            if (p < rp[r + 1]) pv = make_int2(ci[p], __float_as_int(av[p]));
        }
        out[t] = pv;
    }
}

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

// ---------------------------------------------------------------------------------------------
// Vectorised row kernels. Each lane carries V consecutive columns of the (n, K) operands (16-byte
// loads for K >= 4), so a row needs G = K / V lane groups; SUB sub-lanes per group split the row's
// nonzeros and are folded by a fixed xor tree. L = G * SUB lanes per row.
// ---------------------------------------------------------------------------------------------
template <int K>
constexpr int vec_width() {
    return K >= 4 ? 4 : K;
}

template <int V>
__device__ __forceinline__ void vload(const float* __restrict__ p, float (&v)[V]) {
    if constexpr (V == 4) {
        const float4 t = *reinterpret_cast<const float4*>(p);
        v[0] = t.x;
        v[1] = t.y;
        v[2] = t.z;
        v[3] = t.w;
    } else if constexpr (V == 2) {
        const float2 t = *reinterpret_cast<const float2*>(p);
        v[0] = t.x;
        v[1] = t.y;
    } else {
        v[0] = *p;
    }
}

template <int V>
__device__ __forceinline__ void vstore(float* p, const float (&v)[V]) {
    if constexpr (V == 4) {
        *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
    } else if constexpr (V == 2) {
        *reinterpret_cast<float2*>(p) = make_float2(v[0], v[1]);
    } else {
        *p = v[0];
    }
}

// acc[v] = sum over j in [j, e) step SUB of av[j] * (X[ci[j] * K + c0 + v] (- xi[v] when DIFF)),
// folded in the plain loop's order. Entries go in batches of U: the U index/value loads, then the
// U gathers, are all issued before the first use, so a typical row is two memory round trips.
// Every load is unconditional: an entry past the row's end reads the row's last entry again and
// gets weight 0, so no branch splits the batch (a branch per entry serialises the loads).
// Requires e > j implies e - 1 is a valid entry.
template <int SUB>
constexpr int gather_batch() {
    return SUB == 4 ? 4 : 8;
}

// With GS, the gathered rows are scaled on the fly: X_j -> (gs0 * gscale[j]) X_j (the Jacobi
// iterate from zero, alpha0 dinv_j b_j, without writing it out).
template <int K, int V, int SUB, bool DIFF, int U = gather_batch<SUB>(), bool GS = false>
__device__ __forceinline__ void row_gather_v(const std::int32_t* __restrict__ ci, const float* __restrict__ av, int j,
                                             int e, const float* __restrict__ X, int c0, const float (&xi)[V],
                                             float (&acc)[V], const float* __restrict__ gscale = nullptr,
                                             float gs0 = 0.0f) {
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v] = 0.0f;
    for (; j < e; j += U * SUB) {
        int cc[U];
        float aa[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            // (Non-temporal loads of the operator were measured: they slow every row kernel by
            // 20-75 %, so the operator goes through the caches like everything else.)
            const int jj = min(j + u * SUB, e - 1);
            cc[u] = ci[jj];
            aa[u] = av[jj];
        }
        float xx[U][V];
        float gw[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            vload<V>(X + static_cast<std::int64_t>(cc[u]) * K + c0, xx[u]);
            if constexpr (GS) gw[u] = gscale[cc[u]];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const float a = (j + u * SUB < e) ? aa[u] : 0.0f;
            if constexpr (GS) {
                const float w = gs0 * gw[u];
#pragma unroll
                for (int v = 0; v < V; ++v) acc[v] = fmaf(a, w * xx[u][v], acc[v]);
            } else {
#pragma unroll
                for (int v = 0; v < V; ++v) acc[v] = fmaf(a, DIFF ? xx[u][v] - xi[v] : xx[u][v], acc[v]);
            }
        }
    }
}

// The same over K matrices that share one pattern: column c reads its own value av[j * K + c].
template <int K, int V, int SUB, int U = 4>
__device__ __forceinline__ void row_gather_multi(const std::int32_t* __restrict__ ci, const float* __restrict__ av,
                                                 int j, int e, const float* __restrict__ X, int c0,
                                                 const float (&xi)[V], float (&acc)[V]) {
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v] = 0.0f;
    for (; j < e; j += U * SUB) {
        int cc[U];
        float aa[U][V], xx[U][V];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int jj = min(j + u * SUB, e - 1);
            cc[u] = ci[jj];
            vload<V>(av + static_cast<std::int64_t>(jj) * K + c0, aa[u]);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) vload<V>(X + static_cast<std::int64_t>(cc[u]) * K + c0, xx[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool ok = j + u * SUB < e;
#pragma unroll
            for (int v = 0; v < V; ++v) acc[v] = fmaf(ok ? aa[u][v] : 0.0f, xx[u][v] - xi[v], acc[v]);
        }
    }
}

// row_gather_v over a sliced ELL row: slot j of the row at pr[base + j * kEllSlice], slots
// [j, e) step SUB in batches of U. The slots follow the CSR entries in order and a padding slot
// (the row itself, value 0) adds an exact zero, so the sums are those of the CSR form. No row
// pointers: the slot loads depend on nothing, and a wavefront's rows read one contiguous run per
// slot.
template <int K, int V, int SUB, bool DIFF, int U = gather_batch<SUB>(), bool GS = false>
__device__ __forceinline__ void ell_gather_v(const int2* __restrict__ pr, std::int64_t base, int j, int e,
                                             const float* __restrict__ X, int c0, const float (&xi)[V],
                                             float (&acc)[V], const float* __restrict__ gscale = nullptr,
                                             float gs0 = 0.0f) {
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v] = 0.0f;
    for (; j < e; j += U * SUB) {
        int cc[U];
        float aa[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int jj = min(j + u * SUB, e - 1);
            const int2 pv = pr[base + static_cast<std::int64_t>(jj) * kEllSlice];
            cc[u] = pv.x;
            aa[u] = __int_as_float(pv.y);
        }
        float xx[U][V];
        float gw[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            vload<V>(X + static_cast<std::int64_t>(cc[u]) * K + c0, xx[u]);
            if constexpr (GS) gw[u] = gscale[cc[u]];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const float a = (j + u * SUB < e) ? aa[u] : 0.0f;
            if constexpr (GS) {
                const float w = gs0 * gw[u];
#pragma unroll
                for (int v = 0; v < V; ++v) acc[v] = fmaf(a, w * xx[u][v], acc[v]);
            } else {
#pragma unroll
                for (int v = 0; v < V; ++v) acc[v] = fmaf(a, DIFF ? xx[u][v] - xi[v] : xx[u][v], acc[v]);
            }
        }
    }
}

constexpr int kRowThreads = 256;
constexpr int kRowWaves = kRowThreads / kWave;
// Row groups each wave walks per block. One: a second pass of the loop would have to wait for the
// first pass's stores (out may alias x_prev, so the compiler orders them), a full memory round trip.
constexpr int kRowIters = 1;

template <int K, int SUB>
constexpr std::int64_t rows_per_vblock() {
    return static_cast<std::int64_t>(kRowWaves) * (kWave / ((K / vec_width<K>()) * SUB)) * kRowIters;
}

enum RowMode : int {
    kResidual = 0, kCheby = 1, kSpmm = 2, kSpmmDiff = 3, kSpmmDiffMulti = 4, kChebyZero = 5, kProlong = 6
};

// One kernel for every SpMV-shaped pass:
//   kResidual  out = b - A x
//   kCheby     out = c_cur x + c_prev x_prev + alpha dinv (b - A x); out may alias x_prev
//   kSpmm      out = A x
//   kSpmmDiff  out = sum_{j != i} a_ij (x_j - x_i) + s_i x_i over the differenced operator, with
//              `diag` holding the row sums s_i
//   kSpmmDiffMulti  the same for K matrices sharing the pattern: values (nnz, K) and row sums
//              (n, K), column c of x multiplied by matrix c
//   kProlong   x_prev += P x over the prolongator's rows (x the coarse vector, x_prev the fine one,
//              updated in place by the row's own lanes)
//   kChebyZero the first two smoothing steps from a zero guess in one pass: with
//              x1 = alpha0 dinv b (formed on the fly for the row and its neighbours, never
//              stored), out = c_cur x1 + alpha dinv (b - A x1); `x` is b and c_prev carries alpha0
// ELL: the rows come from the sliced ELL copy (ci holds its pairs, ell_w slots a row; rp and av
// unused), the same sums as the CSR form.
template <int K, int SUB, int MODE, bool ELL = false>
__global__ void __launch_bounds__(kRowThreads) rowk_kernel(
    std::int64_t n, const std::int32_t* __restrict__ rp, const std::int32_t* __restrict__ ci,
    const float* __restrict__ av, const float* __restrict__ diag, const float* __restrict__ b,
    const float* __restrict__ x, const float* x_prev, float c_cur, float c_prev, float alpha, float* out,
    const int* __restrict__ state, unsigned long long* __restrict__ stamps, int ell_w) {
    if (finished(state)) return;
    constexpr int V = vec_width<K>();
    constexpr int G = K / V;
    constexpr int L = G * SUB;
    constexpr int RPW = kWave / L;
    static_assert(kRowIters == 1, "one row group per wavefront");
    const int lb = logical_block();
    // The block's start time stays in a register until the end: a store here would make the first
    // loads wait for it.
    const unsigned long long t_start = stamps != nullptr ? wall_clock64() : 0ull;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
    const int l = lane % L, grp = l % G, sub = l / G;
    const int c0 = grp * V;
    const std::int64_t lo = static_cast<std::int64_t>(lb) * rows_per_vblock<K, SUB>();
    const std::int64_t hi = min(lo + rows_per_vblock<K, SUB>(), n);
    const std::int64_t r0 = lo + static_cast<std::int64_t>(wave) * RPW;
    if (r0 < hi) {
        const std::int64_t row = r0 + lane / L;
        const bool valid = row < hi;
        const bool lead = valid && sub == 0;
        // Lanes past the last row read the last row (no branch around the loads) and gather nothing.
        const std::int64_t rr = valid ? row : hi - 1;
        const std::int64_t at = rr * K + c0;
        // Row pointers first: the gather waits on them, and loads complete in issue order. Both are
        // loaded unconditionally (a load under a condition becomes a branch and a wait).
        int j0, e;
        std::int64_t ebase = 0;
        if constexpr (ELL) {
            j0 = sub;
            e = valid ? ell_w : j0;
            ebase = (rr / kEllSlice) * kEllSlice * static_cast<std::int64_t>(ell_w) + rr % kEllSlice;
        } else {
            const int p0 = rp[rr], p1 = rp[rr + 1];
            j0 = p0 + sub;
            e = valid ? p1 : j0;
        }
        const int2* const pr = reinterpret_cast<const int2*>(ci);
        float xi[V], acc[V];
#pragma unroll
        for (int v = 0; v < V; ++v) xi[v] = 0.0f;
        if constexpr (MODE == kSpmmDiff || MODE == kSpmmDiffMulti) vload<V>(x + at, xi);
        // The row's own operands go out before the gather, so their latency hides behind it
        // (x_prev is read before the store below either way: out may alias it).
        float bb[V], cur[V], prev[V], sv[V];
        float dg = 0.0f;
#pragma unroll
        for (int v = 0; v < V; ++v) bb[v] = cur[v] = prev[v] = sv[v] = 0.0f;
        if constexpr (MODE == kResidual || MODE == kCheby || MODE == kChebyZero) vload<V>(b + at, bb);
        if constexpr (MODE == kCheby) {
            vload<V>(x + at, cur);
            vload<V>(x_prev + at, prev);
        }
        if constexpr (MODE == kProlong) vload<V>(x_prev + at, prev);  // out is x_prev
        if constexpr (MODE == kCheby || MODE == kSpmmDiff || MODE == kChebyZero) dg = diag[rr];
        if constexpr (MODE == kSpmmDiffMulti) vload<V>(diag + at, sv);
        if constexpr (MODE == kSpmmDiffMulti) {
            static_assert(!ELL, "the multi-matrix values follow the CSR entries");
            row_gather_multi<K, V, SUB>(ci, av, j0, e, x, c0, xi, acc);
        } else if constexpr (MODE == kChebyZero) {
            if constexpr (ELL)
                ell_gather_v<K, V, SUB, false, gather_batch<SUB>(), true>(pr, ebase, j0, e, x, c0, xi, acc, diag, c_prev);
            else
                row_gather_v<K, V, SUB, false, gather_batch<SUB>(), true>(ci, av, j0, e, x, c0, xi, acc, diag, c_prev);
        } else {
            if constexpr (ELL)
                ell_gather_v<K, V, SUB, MODE == kSpmmDiff>(pr, ebase, j0, e, x, c0, xi, acc);
            else
                row_gather_v<K, V, SUB, MODE == kSpmmDiff>(ci, av, j0, e, x, c0, xi, acc);
        }
#pragma unroll
        for (int off = G; off < L; off <<= 1) {
#pragma unroll
            for (int v = 0; v < V; ++v) acc[v] += __shfl_xor(acc[v], off, kWave);
        }
        float o[V];
        if constexpr (MODE == kSpmm) {
#pragma unroll
            for (int v = 0; v < V; ++v) o[v] = acc[v];
        } else if constexpr (MODE == kProlong) {
#pragma unroll
            for (int v = 0; v < V; ++v) o[v] = prev[v] + acc[v];
        } else if constexpr (MODE == kSpmmDiff) {
#pragma unroll
            for (int v = 0; v < V; ++v) o[v] = fmaf(dg, xi[v], acc[v]);
        } else if constexpr (MODE == kSpmmDiffMulti) {
#pragma unroll
            for (int v = 0; v < V; ++v) o[v] = fmaf(sv[v], xi[v], acc[v]);
        } else if constexpr (MODE == kResidual) {
#pragma unroll
            for (int v = 0; v < V; ++v) o[v] = bb[v] - acc[v];
        } else if constexpr (MODE == kChebyZero) {
            const float d = alpha * dg, w = c_prev * dg;  // c_prev carries alpha0
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const float x1 = w * bb[v];
                o[v] = fmaf(d, bb[v] - acc[v], fmaf(0.0f, x1, c_cur * x1));
            }
        } else {
            const float d = alpha * dg;
#pragma unroll
            for (int v = 0; v < V; ++v) o[v] = fmaf(d, bb[v] - acc[v], fmaf(c_prev, prev[v], c_cur * cur[v]));
        }
        if (lead) vstore<V>(out + at, o);
    }
    if (stamps != nullptr) {  // the block's end: after its last wavefront's stores
        __syncthreads();
        if (threadIdx.x == 0) {
            stamps[2 * lb] = t_start;
            stamps[2 * lb + 1] = wall_clock64();
        }
    }
}

// In-solve timing on the device wall clock. A stamp is 0 when the solve had already finished, so
// the accumulator counts only work that ran.
__global__ void stamp_kernel(unsigned long long* __restrict__ slot, const int* __restrict__ state) {
    *slot = finished(state) ? 0ull : wall_clock64();
}

// sums[0] += SpMM span (earliest block start to latest block end over `blocks` block stamps),
// sums[1] += V-cycle span, sums[2] += iteration span, sums[3] += 1, in wall-clock ticks; only for
// iterations that ran in full.
__global__ void __launch_bounds__(kBlock) timing_accumulate_kernel(const unsigned long long* __restrict__ spmm,
                                                                   int blocks, const unsigned long long* __restrict__ marks,
                                                                   double* __restrict__ sums) {
    __shared__ unsigned long long lo[kBlock], hi[kBlock];
    unsigned long long a = ~0ull, b = 0ull;
    for (int i = threadIdx.x; i < blocks; i += kBlock) {
        a = min(a, spmm[2 * i]);
        b = max(b, spmm[2 * i + 1]);
    }
    lo[threadIdx.x] = a;
    hi[threadIdx.x] = b;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            lo[threadIdx.x] = min(lo[threadIdx.x], lo[threadIdx.x + w]);
            hi[threadIdx.x] = max(hi[threadIdx.x], hi[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && marks[0] != 0 && marks[1] != 0 && marks[2] != 0 && marks[3] != 0 && hi[0] > lo[0]) {
        sums[0] += static_cast<double>(hi[0] - lo[0]);
        sums[1] += static_cast<double>(marks[2] - marks[1]);
        sums[2] += static_cast<double>(marks[3] - marks[0]);
        sums[3] += 1.0;
    }
}

// G1 = A^T B1 (and G2 = A^T B2), FP64 partials of one logical block: lane a of a row group holds
// column a of A and accumulates row a of G from the B rows, which every lane of the group reads
// whole (a broadcast). No cross-lane traffic until the block reduction.
// Block size of the streaming block-vector kernels: as many waves as the accumulators' registers
// allow, since each wave walks its rows with one load round trip per step.
template <int K, bool TWO>
constexpr int gram_threads() {
    return (TWO ? 2 : 1) * K <= 16 ? 1024 : ((TWO ? 2 : 1) * K <= 32 ? 512 : 256);
}

template <int K, bool TWO>
__global__ void __launch_bounds__((gram_threads<K, TWO>())) gram_rows_kernel(
    std::int64_t n, const float* __restrict__ A, const float* __restrict__ B1, const float* __restrict__ B2,
    double* __restrict__ partials, std::int64_t rows_per_block, const int* __restrict__ state) {
    if (finished(state)) return;
    constexpr int RPW = kWave / K;
    constexpr int WAVES = gram_threads<K, TWO>() / kWave;
    constexpr int V = vec_width<K>();
    const int lb = logical_block();
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
    const int a = lane % K, rg = lane / K;
    const std::int64_t lo = static_cast<std::int64_t>(lb) * rows_per_block;
    const std::int64_t hi = min(lo + rows_per_block, n);
    double g1[K], g2[K];
#pragma unroll
    for (int c = 0; c < K; ++c) g1[c] = g2[c] = 0.0;
    for (std::int64_t r0 = lo + static_cast<std::int64_t>(wave) * RPW; r0 < hi; r0 += WAVES * RPW) {
        const std::int64_t row = r0 + rg;
        if (row < hi) {
            const double av = static_cast<double>(A[row * K + a]);
#pragma unroll
            for (int c = 0; c < K; c += V) {
                float bv[V];
                vload<V>(B1 + row * K + c, bv);
#pragma unroll
                for (int v = 0; v < V; ++v) g1[c + v] = fma(av, static_cast<double>(bv[v]), g1[c + v]);
                if constexpr (TWO) {
                    vload<V>(B2 + row * K + c, bv);
#pragma unroll
                    for (int v = 0; v < V; ++v) g2[c + v] = fma(av, static_cast<double>(bv[v]), g2[c + v]);
                }
            }
        }
    }
#pragma unroll
    for (int c = 0; c < K; ++c) {
        g1[c] = wave_rows_sum<K>(g1[c]);
        if constexpr (TWO) g2[c] = wave_rows_sum<K>(g2[c]);
    }
    // Block partial in wave order: lanes a < K of each wave add row a.
    __shared__ double acc[(TWO ? 2 : 1) * K * K];
    for (int e = threadIdx.x; e < (TWO ? 2 : 1) * K * K; e += blockDim.x) acc[e] = 0.0;
    __syncthreads();
    for (int w = 0; w < WAVES; ++w) {
        if (wave == w && lane < K) {
#pragma unroll
            for (int c = 0; c < K; ++c) {
                acc[a * K + c] += g1[c];
                if constexpr (TWO) acc[K * K + a * K + c] += g2[c];
            }
        }
        __syncthreads();
    }
    double* out = partials + static_cast<std::int64_t>(lb) * (TWO ? 2 : 1) * K * K;
    for (int e = threadIdx.x; e < (TWO ? 2 : 1) * K * K; e += blockDim.x) out[e] = acc[e];
}

typedef double doublex4_t __attribute__((ext_vector_type(4)));

// The same partials for K = 16 and 32 on the FP64 matrix cores: v_mfma_f64_16x16x4f64 takes a
// 16x4 block of A^T and a 4x16 block of B (lane l holds column l % 16 of data row l / 16, one
// coalesced dword load per operand: 4 rows x 64 B) and accumulates the 16x16 tile in FP64, so
// the products of the FP32 inputs are exact and each input is converted once. Accumulator
// register v of lane l holds tile row 4 v + l / 16 of tile column l % 16 (the FP64 layout, unlike
// the FP32 one's rows 4 (l / 16) + v). U quads of rows per step keep 3 U loads in flight.
// Threads per Gram workgroup. The pipelined sweep keeps one persistent N-body workgroup on every
// CU; above 16 columns a 1024-thread Gram workgroup (16 wave slots, 111 VGPRs a lane with two
// products) rarely found a CU with room beside it (1.46 ms in the sweep against 0.1 ms alone), and
// four-wave workgroups fill around it. Up to 16 columns 1024 threads fit and hide more latency
// (43 µs alone against 84 µs in four-wave workgroups, 105 against 97 µs in the sweep).
template <int K>
constexpr int gram_mfma_threads() {
    return K > 16 ? 256 : 1024;
}

template <int K, bool TWO>
__global__ void __launch_bounds__(gram_mfma_threads<K>()) gram_mfma_kernel(
    std::int64_t n, const float* __restrict__ A, const float* __restrict__ B1, const float* __restrict__ B2,
    double* __restrict__ partials, std::int64_t rows_per_block, const int* __restrict__ state) {
    static_assert(K == 16 || K == 32, "matrix-core Gram tiles are 16 wide");
    if (finished(state)) return;
    // The FP32 -> FP64 conversions feeding the matrix cores are vector-ALU work, which in the
    // pipelined sweep competes with the VALU-bound N-body on the same SIMDs; a higher wave priority
    // lets them issue first (+0.7-1.0 % on the 1.5 mm sweep, same-box pairs in either order).
    __builtin_amdgcn_s_setprio(3);
    constexpr int H = K / 16;          // 16-column halves of a row
    constexpr int NB = TWO ? 2 : 1;
    constexpr int U = 8 / H;           // row quads per step
    constexpr int WAVES = gram_mfma_threads<K>() / kWave;
    const int lb = logical_block();
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
    const int col = lane & 15, quad = lane >> 4;
    const std::int64_t lo = static_cast<std::int64_t>(lb) * rows_per_block;
    const std::int64_t hi = min(lo + rows_per_block, n);
    doublex4_t acc[NB][H][H];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int x = 0; x < H; ++x)
#pragma unroll
            for (int y = 0; y < H; ++y) acc[b][x][y] = doublex4_t{0.0, 0.0, 0.0, 0.0};
    for (std::int64_t r0 = lo + static_cast<std::int64_t>(wave) * (4 * U); r0 < hi; r0 += WAVES * 4 * U) {
        float av[U][H], bv[NB][U][H];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const std::int64_t row = r0 + 4 * u + quad;
            const bool ok = row < hi;
            const std::int64_t at = (ok ? row : hi - 1) * K + col;  // clamped: loads stay unconditional
#pragma unroll
            for (int h = 0; h < H; ++h) {
                av[u][h] = ok ? A[at + 16 * h] : 0.0f;
                bv[0][u][h] = B1[at + 16 * h];
                if constexpr (TWO) bv[1][u][h] = B2[at + 16 * h];
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int x = 0; x < H; ++x) {
                const double a = static_cast<double>(av[u][x]);
#pragma unroll
                for (int b = 0; b < NB; ++b)
#pragma unroll
                    for (int y = 0; y < H; ++y)
                        acc[b][x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, static_cast<double>(bv[b][u][y]),
                                                                            acc[b][x][y], 0, 0, 0);
            }
    }
    // Block partial in wave order.
    __shared__ double red[NB * K * K];
    for (int e = threadIdx.x; e < NB * K * K; e += blockDim.x) red[e] = 0.0;
    __syncthreads();
    for (int w = 0; w < WAVES; ++w) {
        if (wave == w) {
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int x = 0; x < H; ++x)
#pragma unroll
                    for (int y = 0; y < H; ++y)
#pragma unroll
                        for (int v = 0; v < 4; ++v)
                            red[(b * K + 16 * x + 4 * v + quad) * K + 16 * y + col] += acc[b][x][y][v];
        }
        __syncthreads();
    }
    double* out = partials + static_cast<std::int64_t>(lb) * NB * K * K;
    for (int e = threadIdx.x; e < NB * K * K; e += blockDim.x) out[e] = red[e];
}

// G1 = A^T B1 (and G2 = A^T B2) partials, kGramBlocks logical blocks of `rows_per_block` rows.
template <int K, bool TWO>
void launch_gram(std::int64_t n, const float* A, const float* B1, const float* B2, double* partials,
                 std::int64_t rows_per_block, const int* state, hipStream_t s) {
    if constexpr (K >= 16)
        gram_mfma_kernel<K, TWO><<<kGramBlocks, gram_mfma_threads<K>(), 0, s>>>(n, A, B1, B2, partials,
                                                                                rows_per_block, state);
    else
        gram_rows_kernel<K, TWO><<<kGramBlocks, gram_threads<K, TWO>(), 0, s>>>(n, A, B1, B2, partials,
                                                                              rows_per_block, state);
}

constexpr int kGramThreads = 1024;
constexpr int kGramWaves = kGramThreads / kWave;

// ---------------------------------------------------------------------------------------------
// V-cycle kernels (solver/amg/vcycle.cu:124-311), K-wide, FP32. The smoother and residual passes
// are rowk_kernel; SUB is chosen per level from its shape (wide coarse rows take more lanes).
// ---------------------------------------------------------------------------------------------
template <int K>
__global__ void jacobi_zero_kernel(std::int64_t n, const float* __restrict__ dinv, const float* __restrict__ b,
                                   float alpha, float* __restrict__ x, const int* __restrict__ state) {
    if (finished(state)) return;
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n * K) x[i] = alpha * dinv[i / K] * b[i];
}

// Coarsest level in FP64: x = Ainv b, one workgroup per row of the dense inverse. Lane l of
// wavefront w takes column l % K of the entries j = Q w + l / K (mod 4 Q, Q = 64 / K) of the row
// in FP64; a column's partial sums fold by a fixed xor tree within each wavefront, then over the
// wavefronts in order.
template <int K>
__global__ void __launch_bounds__(kRowThreads) coarse_kernel(int n, const double* __restrict__ inv,
                                                             const float* __restrict__ b, float* __restrict__ x,
                                                             const int* __restrict__ state) {
    static_assert(K <= kWave, "a row's columns fit a wavefront");
    constexpr int Q = kWave / K;
    if (finished(state)) return;
    __shared__ double part[kRowWaves][K];
    const int row = static_cast<int>(blockIdx.x);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & (kWave - 1);
    const int c = lane % K, q = lane / K;
    const double* a = inv + static_cast<std::int64_t>(row) * n;
    double acc = 0.0;
    for (int j = wave * Q + q; j < n; j += kRowWaves * Q)
        acc = fma(a[j], static_cast<double>(b[static_cast<std::int64_t>(j) * K + c]), acc);
#pragma unroll
    for (int off = K; off < kWave; off <<= 1) acc += __shfl_xor(acc, off, kWave);
    if (q == 0) part[wave][c] = acc;
    __syncthreads();
    if (threadIdx.x < K) {
        double s = part[0][threadIdx.x];
#pragma unroll
        for (int w = 1; w < kRowWaves; ++w) s += part[w][threadIdx.x];
        x[static_cast<std::int64_t>(row) * K + threadIdx.x] = static_cast<float>(s);
    }
}

typedef float floatx4_t __attribute__((ext_vector_type(4)));

// Dense tail: x = M b for the (n, n) FP32 operator M of the V-cycle from a level down, on the FP32
// matrix cores. M is stored in 16 x 4 tiles in the operand order of v_mfma_f32_16x16x4f32 (lane l
// holds row l % 16, column l / 16 of a tile): tile (row tile rt, column quad q) is the 64 floats at
// (rt * nq + q) * 64, rows and columns past n zero. A block computes one 16-row tile: each of its
// wavefronts accumulates a contiguous share of the column quads into 16 x 16 FP32 tiles (one per 16
// columns of b), one coalesced 256-byte load of M and one of b per product, the loads of kTailBatch
// products issued before the first of them; the shares fold in wavefront order, so the sums are
// fixed. Accumulator v of lane l is tile row 4 (l / 16) + v, column l % 16.
constexpr int kTailThreads = 1024;
constexpr int kTailBatch = 16;

template <int K>
__global__ void __launch_bounds__(kTailThreads) dense_tail_kernel(int n, int nq, const float* __restrict__ Mt,
                                                                  const float* __restrict__ b, float* __restrict__ x,
                                                                  const int* __restrict__ state) {
    constexpr int H = K > 16 ? K / 16 : 1;  // 16-column tiles of b
    constexpr int WAVES = kTailThreads / kWave;
    if (finished(state)) return;
    __shared__ float part[WAVES][H][4][kWave];
    const int rt = static_cast<int>(blockIdx.x);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & (kWave - 1);
    const int col = lane & 15, quad = lane >> 4;
    const int per = (nq + WAVES - 1) / WAVES;
    const int q0 = wave * per, q1 = min(q0 + per, nq);
    floatx4_t acc[H];
#pragma unroll
    for (int h = 0; h < H; ++h) acc[h] = floatx4_t{0.f, 0.f, 0.f, 0.f};
    const float* a = Mt + static_cast<std::int64_t>(rt) * nq * kWave + lane;
    for (int q = q0; q < q1; q += kTailBatch) {
        float av[kTailBatch], bv[kTailBatch][H];
#pragma unroll
        for (int u = 0; u < kTailBatch; ++u) {
            // Past the share's end: the last quad again, with M taken as zero (loads stay unconditional).
            const int qq = min(q + u, q1 - 1);
            av[u] = a[static_cast<std::int64_t>(qq) * kWave];
            const int k = min(4 * qq + quad, n - 1);  // rows of b past n meet zero columns of M
#pragma unroll
            for (int h = 0; h < H; ++h) {
                const int j = 16 * h + col;
                bv[u][h] = (j < K) ? b[static_cast<std::int64_t>(k) * K + j] : 0.0f;
            }
        }
#pragma unroll
        for (int u = 0; u < kTailBatch; ++u) {
            const float m = (q + u < q1) ? av[u] : 0.0f;
#pragma unroll
            for (int h = 0; h < H; ++h) acc[h] = __builtin_amdgcn_mfma_f32_16x16x4f32(m, bv[u][h], acc[h], 0, 0, 0);
        }
    }
#pragma unroll
    for (int h = 0; h < H; ++h)
#pragma unroll
        for (int v = 0; v < 4; ++v) part[wave][h][v][lane] = acc[h][v];
    __syncthreads();
    if (wave == 0) {
#pragma unroll
        for (int h = 0; h < H; ++h) {
            const int j = 16 * h + col;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                float s = part[0][h][v][lane];
                for (int w = 1; w < WAVES; ++w) s += part[w][h][v][lane];
                const int row = 16 * rt + 4 * quad + v;
                if (row < n && j < K) x[static_cast<std::int64_t>(row) * K + j] = s;
            }
        }
    }
}

// Column block [c0, c0 + K) of the identity at n rows, (n, K) row-major.
template <int K>
__global__ void identity_cols_kernel(int n, int c0, float* __restrict__ e) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < static_cast<std::int64_t>(n) * K) e[i] = (i / K == c0 + i % K) ? 1.0f : 0.0f;
}

// Columns c0 .. c0 + K - 1 (those below n) of the tiled M from the (n, K) block X.
template <int K>
__global__ void scatter_cols_kernel(int n, int nq, int c0, const float* __restrict__ X, float* __restrict__ Mt) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= static_cast<std::int64_t>(n) * K) return;
    const int r = static_cast<int>(i / K);
    const int c = c0 + static_cast<int>(i % K);
    if (c < n)
        Mt[(static_cast<std::int64_t>(r / 16) * nq + c / 4) * kWave + (c % 4) * 16 + r % 16] = X[i];
}

// ---------------------------------------------------------------------------------------------
// Block-vector kernels of the PCG iteration.
// ---------------------------------------------------------------------------------------------

// Per-column partials of a (K)-vector held as V entries by the lanes of column group `grp`:
// summed over the wave's rows, then added wave by wave into the block's slot.
template <int K, int V, int WAVES>
__device__ __forceinline__ void block_cols_store(double (&v)[V], int grp, double* __restrict__ out) {
    constexpr int G = K / V;
    __shared__ double acc[K];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
#pragma unroll
    for (int u = 0; u < V; ++u) v[u] = wave_rows_sum<G>(v[u]);
    if (threadIdx.x < K) acc[threadIdx.x] = 0.0;
    __syncthreads();
    for (int w = 0; w < WAVES; ++w) {
        if (wave == w && lane < G) {
#pragma unroll
            for (int u = 0; u < V; ++u) acc[grp * V + u] += v[u];
        }
        __syncthreads();
    }
    if (threadIdx.x < K) out[threadIdx.x] = acc[threadIdx.x];
}

// FP64 true residual R64 = B - A X64 with the FP64 operator (MULTI: A_c = sum_t sigma[c, t] A_t
// from the per-tissue FP64 values tv (T, nnz), combined per entry in tissue order), and per-column
// ||R64||^2 partials. A lane carries V = 2 consecutive columns (16-byte loads of X64), G = K / V
// lanes per row; SUB sub-lanes split the row's entries and are folded by a fixed xor tree. Each
// sub-lane loads a batch of kResBatch entries (indices clamped to the row, values zeroed past it)
// before their X64 gathers, so the gathers of a batch are in flight together.
template <int K>
constexpr int res_vec() {
    return K >= 2 ? 2 : 1;
}
template <int K>
constexpr int res_sub() {
    return (K / res_vec<K>()) >= 16 ? 1 : 16 / (K / res_vec<K>());
}
constexpr int kResBatch = 4;

template <int K, bool MULTI>
__global__ void __launch_bounds__(kGramThreads) residual64_kernel(
    std::int64_t n, const std::int32_t* __restrict__ rp, const std::int32_t* __restrict__ ci,
    const double* __restrict__ av, std::int64_t nnz, int T, const double* __restrict__ sigma,
    const float* __restrict__ B, const double* __restrict__ X, double* __restrict__ R, double* __restrict__ partials,
    std::int64_t rows_per_block, bool x_zero) {
    constexpr int V = res_vec<K>();
    constexpr int G = K / V;
    constexpr int SUB = res_sub<K>();
    constexpr int L = G * SUB;
    constexpr int RPW = kWave / L;
    const int lb = logical_block();
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
    const int l = lane % L, grp = l % G, sub = l / G, c0 = grp * V;
    const std::int64_t lo = static_cast<std::int64_t>(lb) * rows_per_block;
    const std::int64_t hi = min(lo + rows_per_block, n);
    double nrm[V];
#pragma unroll
    for (int v = 0; v < V; ++v) nrm[v] = 0.0;
    for (std::int64_t r0 = lo + static_cast<std::int64_t>(wave) * RPW; r0 < hi; r0 += kGramWaves * RPW) {
        const std::int64_t row = r0 + lane / L;
        const bool valid = row < hi;
        double acc[V];
#pragma unroll
        for (int v = 0; v < V; ++v) acc[v] = 0.0;
        if (valid && !x_zero) {
            const int b = rp[row], e = rp[row + 1];
            for (int j0 = b + sub; j0 < e; j0 += SUB * kResBatch) {
                int col[kResBatch];
                double a[kResBatch][V];
#pragma unroll
                for (int u = 0; u < kResBatch; ++u) {
                    const int j = j0 + SUB * u;
                    const bool ok = j < e;
                    const int jj = ok ? j : e - 1;
                    col[u] = ci[jj];
                    if constexpr (MULTI) {
#pragma unroll
                        for (int v = 0; v < V; ++v) a[u][v] = 0.0;
                        for (int t = 0; t < T; ++t) {
                            const double tv = av[static_cast<std::int64_t>(t) * nnz + jj];
#pragma unroll
                            for (int v = 0; v < V; ++v) a[u][v] = fma(sigma[(c0 + v) * T + t], tv, a[u][v]);
                        }
#pragma unroll
                        for (int v = 0; v < V; ++v) a[u][v] = ok ? a[u][v] : 0.0;
                    } else {
                        const double x = ok ? av[jj] : 0.0;
#pragma unroll
                        for (int v = 0; v < V; ++v) a[u][v] = x;
                    }
                }
                double xv[kResBatch][V];
#pragma unroll
                for (int u = 0; u < kResBatch; ++u) {
                    const double* xp = X + static_cast<std::int64_t>(col[u]) * K + c0;
                    if constexpr (V == 2) {
                        const double2 t2 = *reinterpret_cast<const double2*>(xp);
                        xv[u][0] = t2.x;
                        xv[u][1] = t2.y;
                    } else {
                        xv[u][0] = xp[0];
                    }
                }
#pragma unroll
                for (int u = 0; u < kResBatch; ++u)
#pragma unroll
                    for (int v = 0; v < V; ++v) acc[v] = fma(a[u][v], xv[u][v], acc[v]);
            }
        }
#pragma unroll
        for (int v = 0; v < V; ++v)
#pragma unroll
            for (int off = G; off < L; off <<= 1) acc[v] += __shfl_xor(acc[v], off, kWave);
        if (valid && sub == 0) {
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const double r = static_cast<double>(B[row * K + c0 + v]) - acc[v];
                R[row * K + c0 + v] = r;
                nrm[v] = fma(r, r, nrm[v]);
            }
        }
    }
    block_cols_store<K, V, kGramWaves>(nrm, grp, partials + static_cast<std::int64_t>(lb) * K);
}

// R -= S a (FP32), partials of ||R_c||^2; the solution update X += W a is deferred to the end of
// the window (flush_x_kernel). Frozen columns have a zero column in `a`, which leaves R exactly
// unchanged. A lane owns V columns of one row and reads the row's S whole (the other lanes of the
// row read the same cache lines).
template <int K>
constexpr int xr_threads() {
    return K <= 16 ? 1024 : 512;
}

template <int K>
__global__ void __launch_bounds__(xr_threads<K>()) update_r_kernel(
    std::int64_t n, const float* __restrict__ S, const double* __restrict__ a, float* __restrict__ R,
    double* __restrict__ partials, std::int64_t rows_per_block, const int* __restrict__ state) {
    if (finished(state)) return;
    constexpr int V = vec_width<K>();
    constexpr int G = K / V;
    constexpr int RPW = kWave / G;
    // Above 16 columns the step S a is formed in FP64 from the FP64 coefficients, as the deferred X
    // update uses them: a wide block's Gram is worse conditioned, its coefficients larger with more
    // cancellation, and FP32 products left a 32-column block's recursion 1.5x off the true
    // residual on the 1.5 mm head (a stall at 1.19e-6 and a second round; 28 iterations instead of
    // 22). Up to 16 columns the FP32 form meets the true residual within a few per cent.
    using Acc = std::conditional_t<(K > 16), double, float>;
    __shared__ Acc alf[K * K];
    for (int e = threadIdx.x; e < K * K; e += blockDim.x) alf[e] = static_cast<Acc>(a[e]);
    __syncthreads();
    const int lb = logical_block();
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x >> 6;
    const int grp = lane % G, c0 = grp * V;
    const std::int64_t lo = static_cast<std::int64_t>(lb) * rows_per_block;
    const std::int64_t hi = min(lo + rows_per_block, n);
    double nrm[V];
#pragma unroll
    for (int v = 0; v < V; ++v) nrm[v] = 0.0;
    constexpr int WAVES = xr_threads<K>() / kWave;
    for (std::int64_t r0 = lo + static_cast<std::int64_t>(wave) * RPW; r0 < hi; r0 += WAVES * RPW) {
        const std::int64_t row = r0 + lane / G;
        if (row >= hi) continue;
        Acc dr[V];
#pragma unroll
        for (int v = 0; v < V; ++v) dr[v] = Acc(0);
#pragma unroll
        for (int b0 = 0; b0 < K; b0 += V) {
            float sb[V];
            vload<V>(S + row * K + b0, sb);
#pragma unroll
            for (int u = 0; u < V; ++u) {
#pragma unroll
                for (int v = 0; v < V; ++v) dr[v] = fma(static_cast<Acc>(sb[u]), alf[(b0 + u) * K + c0 + v], dr[v]);
            }
        }
        const std::int64_t at = row * K + c0;
        float rr[V];
        vload<V>(R + at, rr);
#pragma unroll
        for (int v = 0; v < V; ++v) {
            rr[v] = static_cast<float>(static_cast<Acc>(rr[v]) - dr[v]);
            nrm[v] = fma(static_cast<double>(rr[v]), static_cast<double>(rr[v]), nrm[v]);
        }
        vstore<V>(R + at, rr);
    }
    block_cols_store<K, V, WAVES>(nrm, grp, partials + static_cast<std::int64_t>(lb) * K);
}

// X64 += sum over the window's slots j < L whose iteration ran of W_j a_j (FP64), in slot order.
// Runs whether or not the solve has finished: the iterations before the finish are still owed.
// Launched in the window's last iteration before its W update.
template <int K>
__global__ void __launch_bounds__(kRowThreads) flush_x_kernel(std::int64_t n, const float* __restrict__ W0,
                                                              const float* __restrict__ W1,
                                                              const float* __restrict__ W2,
                                                              const float* __restrict__ W3,
                                                              const double* __restrict__ AL,
                                                              const int* __restrict__ ran, int L,
                                                              double* __restrict__ X) {
    static_assert(kLag == 4, "four window slots");
    constexpr int V = vec_width<K>();
    constexpr int G = K / V;
    __shared__ double al[kLag * K * K];
    __shared__ int use[kLag];
    for (int e = threadIdx.x; e < kLag * K * K; e += blockDim.x) al[e] = AL[e];
    if (threadIdx.x < kLag) use[threadIdx.x] = (static_cast<int>(threadIdx.x) < L) && ran[threadIdx.x] != 0;
    __syncthreads();
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * kRowThreads + threadIdx.x;
    const std::int64_t row = t / G;
    if (row >= n) return;
    const int c0 = static_cast<int>(t % G) * V;
    const float* const ws[kLag] = {W0, W1, W2, W3};
    double dx[V];
#pragma unroll
    for (int v = 0; v < V; ++v) dx[v] = 0.0;
#pragma unroll
    for (int j = 0; j < kLag; ++j) {
        if (!use[j]) continue;
        const double* a = al + j * K * K;
#pragma unroll
        for (int b0 = 0; b0 < K; b0 += V) {
            float wb[V];
            vload<V>(ws[j] + row * K + b0, wb);
#pragma unroll
            for (int u = 0; u < V; ++u) {
#pragma unroll
                for (int v = 0; v < V; ++v) dx[v] = fma(static_cast<double>(wb[u]), a[(b0 + u) * K + c0 + v], dx[v]);
            }
        }
    }
    const std::int64_t at = row * K + c0;
#pragma unroll
    for (int v = 0; v < V; ++v) X[at + v] += dx[v];
}

__global__ void clear_ran_kernel(int* __restrict__ ran) {
    if (threadIdx.x < kLag) ran[threadIdx.x] = 0;
}

// W_out <- Z - W D (the next slot of the window's ring, or W itself): every lane of a row loads
// the row's W before any of them stores (the stores depend on the loads, and a row's lanes share
// one wavefront), so W_out may be W.
template <int K>
__global__ void __launch_bounds__(kRowThreads) update_w_kernel(std::int64_t n, const float* __restrict__ Z,
                                                               const double* __restrict__ D, const float* W,
                                                               float* W_out, const int* __restrict__ state) {
    if (finished(state)) return;
    constexpr int V = vec_width<K>();
    constexpr int G = K / V;
    __shared__ float dm[K * K];
    for (int e = threadIdx.x; e < K * K; e += blockDim.x) dm[e] = static_cast<float>(D[e]);
    __syncthreads();
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * kRowThreads + threadIdx.x;
    const std::int64_t row = t / G;
    const bool valid = row < n;
    const int c0 = static_cast<int>(t % G) * V;
    float acc[V];
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v] = 0.0f;
    if (valid) vload<V>(Z + row * K + c0, acc);
#pragma unroll
    for (int b0 = 0; b0 < K; b0 += V) {
        float wb[V];
#pragma unroll
        for (int u = 0; u < V; ++u) wb[u] = 0.0f;
        if (valid) vload<V>(W + row * K + b0, wb);
#pragma unroll
        for (int u = 0; u < V; ++u) {
#pragma unroll
            for (int v = 0; v < V; ++v) acc[v] = fmaf(-wb[u], dm[(b0 + u) * K + c0 + v], acc[v]);
        }
    }
    if (valid) vstore<V>(W_out + row * K + c0, acc);
}

__global__ void cast_f64_f32_kernel(std::int64_t count, const double* __restrict__ in, float* __restrict__ out) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) out[i] = static_cast<float>(in[i]);
}

__global__ void zero_f32_kernel(std::int64_t count, float* __restrict__ out) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) out[i] = 0.0f;
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
        int b = b0;
        for (; b + 8 <= b1; b += 8) {  // eight loads in flight, added in ascending order
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = partials[static_cast<std::int64_t>(b + u) * stride + off + e];
#pragma unroll
            for (int u = 0; u < 8; ++u) s += v[u];
        }
        for (; b < b1; ++b) s += partials[static_cast<std::int64_t>(b) * stride + off + e];
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

// The same sum over many blocks: out[e] = sum_b partials[b * stride + e] for e < E, one wavefront
// per output; lane l adds b = l, l + 64, ... in order (all its loads issued together), then a fixed
// xor tree over the lanes. Run-to-run deterministic like the one-block form, and a single memory
// round trip instead of one per block of eight partials.
constexpr int kReduceThreads = 256;

__global__ void __launch_bounds__(kReduceThreads) reduce_cols_kernel(const double* __restrict__ partials, int nb,
                                                                     int stride, int E, double* __restrict__ out,
                                                                     const int* __restrict__ state) {
    __builtin_amdgcn_s_setprio(3);  // a short chain of dependent steps: issue ahead of the sweep's N-body
    if (finished(state)) return;
    const int w = static_cast<int>((static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & (kWave - 1);
    if (w >= E) return;  // whole wavefronts
    constexpr int kPer = kGramBlocks / kWave;
    double v[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
        const int b = lane + u * kWave;
        v[u] = b < nb ? partials[static_cast<std::int64_t>(b) * stride + w] : 0.0;
    }
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < kPer; ++u) acc += v[u];
    for (int b = lane + kPer * kWave; b < nb; b += kWave) acc += partials[static_cast<std::int64_t>(b) * stride + w];
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, kWave);
    if (lane == 0) out[w] = acc;
}

void reduce_cols(const double* partials, int nb, int E, double* out, const int* state, hipStream_t s) {
    const unsigned g = static_cast<unsigned>((static_cast<std::int64_t>(E) * kWave + kReduceThreads - 1) / kReduceThreads);
    reduce_cols_kernel<<<g, kReduceThreads, 0, s>>>(partials, nb, E, E, out, state);
}

template <int E>
__device__ void load_reduced(const double* __restrict__ red, int off, double* out_lds) {
    for (int e = threadIdx.x; e < E; e += blockDim.x) out_lds[e] = red[off + e];
    __syncthreads();
}

// Symmetric sweeps of an M x M matrix held by one wavefront, lane l owning entries e = l + 64 u
// (row e / M, column e % M) in a[u], its diagonal mirrored in diag (LDS, written by the caller
// before a barrier). The largest remaining diagonal among `cand` (lowest index on ties) is swept
// while it exceeds drop_tol: p -> -1/a_pp on the diagonal, a_ip -> a_ip / a_pp in row and column
// p, a_ij -> a_ij - a_ip a_pj / a_pp elsewhere, with the pivot row passed through LDS (prow) and
// every product formed so that a symmetric matrix stays bitwise symmetric. Returns the swept set;
// its block of a then holds minus the inverse of that principal block, and the unswept diagonal
// the Schur complements the next pivot was chosen from (Goodnight, "A tutorial on the SWEEP
// operator", 1979).
template <int M, int U>
__device__ unsigned sweep_pivots(double (&a)[U], double* diag, double* prow, unsigned cand, double drop_tol) {
    static_assert(M <= 32, "the pivot sets are 32-bit masks");
    const int t = threadIdx.x;
    unsigned swept = 0;
    for (int step = 0; step < M; ++step) {
        double v = -1.0;
        int p = M;
        for (int i = 0; i < M; ++i) {
            const double di = diag[i];
            if (((cand >> i) & 1u) && di > v) {
                v = di;
                p = i;
            }
        }
        if (p >= M || !(v > drop_tol)) break;
        cand &= ~(1u << p);
        swept |= 1u << p;
        const double inv = 1.0 / v;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = t + u * kWave;
            if (e < M * M && e / M == p) prow[e % M] = a[u];  // row p = column p before the sweep
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = t + u * kWave, i = e / M, j = e % M;
            if (e >= M * M) continue;
            if (i == p && j == p)
                a[u] = -inv;
            else if (i == p || j == p)
                a[u] = a[u] * inv;
            else
                a[u] = fma(-(prow[i] * prow[j]), inv, a[u]);
            if (i == j) diag[i] = a[u];
        }
        __syncthreads();
    }
    return swept;
}

// G, H -> TT, a = TT H with the frozen columns zeroed. TT is the inverse of G on the retained
// directions and zero elsewhere. G = W^T A W is symmetric, but its FP64 sums of FP32 products are
// not exactly (an entry small against |W| |A W| keeps a large relative rounding error on each
// side), so the kernel works on the symmetric part (G + G^T) / 2, with every product formed in an
// order that keeps the matrix bitwise symmetric through the sweeps: TT then stays symmetric
// positive definite, as the block recurrence needs. With S = diag(G)^-1/2 and A = S G S, the
// directions are chosen one at a time as the largest remaining Schur-complement diagonal of A
// (lowest index on ties) while it exceeds drop_tol, which is the pivot order and drop rule of a
// diagonally pivoted Cholesky of A, and TT = S (A_PP)^-1 S on the chosen set P, A_PP^-1 from the
// symmetric sweeps (sweep_pivots). One wavefront; lane l keeps entries l, l + 64, ... in
// registers, and a step is a pivot scan of the diagonal (every lane reads it from LDS and finds
// the same pivot), one division, the pivot row through LDS, and one fused update per entry.
template <int K>
__global__ void __launch_bounds__(kWave) coef_kernel(const double* __restrict__ partials, double drop_tol,
                                                     double* __restrict__ TT, double* __restrict__ AL,
                                                     int* __restrict__ state, int independent, int* __restrict__ ran) {
    __builtin_amdgcn_s_setprio(3);  // a short chain of dependent steps: issue ahead of the sweep's N-body
    if (finished(state)) return;
    if (threadIdx.x == 0) *ran = 1;  // this window slot's coefficients are to be applied
    constexpr int kE = K * K;
    constexpr int U = (kE + kWave - 1) / kWave;
    __shared__ double H[kE];
    __shared__ double M[kE];  // TT, for the product with H
    __shared__ double prow[K];
    __shared__ double diag[K];
    __shared__ double sc[K];
    const int t = threadIdx.x;
    // Independent columns (K separate CG chains) keep only the diagonals of G and H.
    if (t < K) {
        const double d = partials[t * K + t];
        sc[t] = d > 0.0 && isfinite(d) ? 1.0 / sqrt(d) : 0.0;
    }
    for (int e = t; e < kE; e += kWave) H[e] = independent && (e / K) != (e % K) ? 0.0 : partials[kE + e];
    __syncthreads();
    unsigned cand = 0;  // rows that may still be pivots: a positive, finite diagonal in G
    for (int i = 0; i < K; ++i)
        if (sc[i] > 0.0) cand |= 1u << i;
    double a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int e = t + u * kWave, i = e / K, j = e % K;
        const bool keep = e < kE && ((cand >> i) & (cand >> j) & 1u) && !(independent && i != j);
        a[u] = keep ? (0.5 * (partials[e] + partials[j * K + i])) * (sc[i] * sc[j]) : 0.0;
        if (e < kE && i == j) diag[i] = a[u];
    }
    __syncthreads();
    const unsigned swept = sweep_pivots<K, U>(a, diag, prow, cand, drop_tol);
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int e = t + u * kWave, i = e / K, j = e % K;
        if (e >= kE) continue;
        const double tt = ((swept >> i) & (swept >> j) & 1u) ? -a[u] * (sc[i] * sc[j]) : 0.0;
        M[e] = tt;
        TT[e] = tt;
    }
    __syncthreads();
    for (int e = t; e < kE; e += kWave) {  // a = TT H, frozen columns zero
        const int i = e / K, c = e % K;
        double s = 0.0;
        for (int b = 0; b < K; ++b) s += M[i * K + b] * H[b * K + c];
        AL[e] = state[2 + c] ? 0.0 : s;
    }
}

// Initial guess of a block in the span of NB earlier blocks' solutions X_i (K columns each, M =
// NB K basis vectors, X_0 the older), the Galerkin projection that minimises the A-norm error over
// that span (Fischer, "Projection techniques for iterative solution of Ax = b with successive
// right-hand sides", 1998): x0 = X c with (X^T A X) c = X^T b. The basis blocks' right-hand sides
// B_i = A X_i (to their solve tolerance) give X^T A X without a product by A: D_i = X_i^T B_i on
// the diagonal, Cx = X_0^T B_1 above it and its transpose below (A is symmetric). gb holds
// [b^T X_0, .., b^T X_{NB-1}] (K x K blocks). C (M x K) = G^+ g on the directions the pivoted
// sweeps of the symmetric part keep above drop_tol, so near-dependent basis vectors drop out;
// `pending` receives X_{NB-1}^T b, the cross block of the next basis once b's block joins it.
template <int K, int NB>
__global__ void __launch_bounds__(kWave) galerkin_solve_kernel(const double* __restrict__ D0,
                                                               const double* __restrict__ D1,
                                                               const double* __restrict__ Cx,
                                                               const double* __restrict__ gb, double drop_tol,
                                                               double* __restrict__ C, double* __restrict__ pending) {
    constexpr int M = NB * K;
    constexpr int kE = M * M;
    constexpr int U = (kE + kWave - 1) / kWave;
    constexpr int KK = K * K;
    __shared__ double Gs[kE];
    __shared__ double g[M * K];
    __shared__ double Tm[kE];
    __shared__ double prow[M];
    __shared__ double diag[M];
    __shared__ double sc[M];
    const int t = threadIdx.x;
    for (int e = t; e < kE; e += kWave) {
        const int I = e / M, J = e % M, bi = I / K, bj = J / K, p = I % K, q = J % K;
        double v;
        if (bi == bj)
            v = (bi == 0 ? D0 : D1)[p * K + q];
        else
            v = bi == 0 ? Cx[p * K + q] : Cx[q * K + p];
        Gs[e] = v;
    }
    for (int e = t; e < M * K; e += kWave) {  // g[iK + p][c] = (b^T X_i)[c][p]
        const int I = e / K, c = e % K;
        g[e] = gb[(I / K) * KK + c * K + I % K];
    }
    for (int e = t; e < KK; e += kWave) pending[e] = gb[(NB - 1) * KK + (e % K) * K + e / K];
    __syncthreads();
    if (t < M) {
        const double d = Gs[t * M + t];
        sc[t] = d > 0.0 && isfinite(d) ? 1.0 / sqrt(d) : 0.0;
    }
    __syncthreads();
    unsigned cand = 0;
    for (int i = 0; i < M; ++i)
        if (sc[i] > 0.0) cand |= 1u << i;
    double a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int e = t + u * kWave, i = e / M, j = e % M;
        const bool keep = e < kE && ((cand >> i) & (cand >> j) & 1u);
        a[u] = keep ? (0.5 * (Gs[e] + Gs[j * M + i])) * (sc[i] * sc[j]) : 0.0;
        if (e < kE && i == j) diag[i] = a[u];
    }
    __syncthreads();
    const unsigned swept = sweep_pivots<M, U>(a, diag, prow, cand, drop_tol);
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const int e = t + u * kWave, i = e / M, j = e % M;
        if (e < kE) Tm[e] = ((swept >> i) & (swept >> j) & 1u) ? -a[u] * (sc[i] * sc[j]) : 0.0;
    }
    __syncthreads();
    for (int e = t; e < M * K; e += kWave) {
        const int I = e / K, c = e % K;
        double s = 0.0;
        for (int J = 0; J < M; ++J) s += Tm[I * M + J] * g[J * K + c];
        C[e] = s;
    }
}

// x0[r][c] = sum over i, p of X_i[r][p] C[iK + p][c]: one thread per row and group of V columns,
// the coefficients in LDS (FP32 products: the result is only a starting point for the solve).
template <int K, int NB>
__global__ void __launch_bounds__(kBlock) galerkin_expand_kernel(std::int64_t n, const float* __restrict__ X0,
                                                                 const float* __restrict__ X1,
                                                                 const double* __restrict__ C,
                                                                 double* __restrict__ x) {
    constexpr int V = vec_width<K>(), G = K / V;
    __shared__ float cs[NB * K * K];
    for (int e = threadIdx.x; e < NB * K * K; e += kBlock) cs[e] = static_cast<float>(C[e]);
    __syncthreads();
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    const std::int64_t r = t / G;
    if (r >= n) return;
    const int c0 = static_cast<int>(t % G) * V;
    float acc[V];
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v] = 0.0f;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const float* xr = (i == 0 ? X0 : X1) + r * K;
#pragma unroll
        for (int p0 = 0; p0 < K; p0 += V) {
            float xv[V];
            vload<V>(xr + p0, xv);
#pragma unroll
            for (int q = 0; q < V; ++q)
#pragma unroll
                for (int v = 0; v < V; ++v) acc[v] = fmaf(xv[q], cs[(i * K + p0 + q) * K + c0 + v], acc[v]);
        }
    }
    double* out = x + r * K + c0;
#pragma unroll
    for (int v = 0; v < V; ++v) out[v] = acc[v];
}

// D = TT E.
template <int K>
__global__ void __launch_bounds__(kSmallThreads) dmat_kernel(const double* __restrict__ partials, int nb,
                                                             const double* __restrict__ TT, double* __restrict__ D,
                                                             int* __restrict__ state, int independent) {
    __builtin_amdgcn_s_setprio(3);  // a short chain of dependent steps: issue ahead of the sweep's N-body
    if (finished(state)) return;
    __shared__ double E[K * K];
    load_reduced<K * K>(partials, 0, E);
    if (independent) {
        for (int e = threadIdx.x; e < K * K; e += blockDim.x)
            if ((e / K) != (e % K)) E[e] = 0.0;
        __syncthreads();
    }
    for (int e = threadIdx.x; e < K * K; e += blockDim.x) {
        const int i = e / K, c = e % K;
        double s = 0.0;
        for (int b = 0; b < K; ++b) s += TT[i * K + b] * E[b * K + c];
        D[e] = s;
    }
}

// ||R_c|| against the column targets; freezes converged columns, counts the iteration.
template <int K>
__global__ void __launch_bounds__(kSmallThreads) converge_kernel(const double* __restrict__ partials, int nb,
                                                                 const double* __restrict__ target,
                                                                 double* __restrict__ norms,
                                                                 int* __restrict__ state) {
    __builtin_amdgcn_s_setprio(3);  // a short chain of dependent steps: issue ahead of the sweep's N-body
    static_assert(K <= kWave, "one lane per column");
    if (finished(state)) return;
    __shared__ double v[K];
    __shared__ int all;
    load_reduced<K>(partials, 0, v);
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

template <int E>
__global__ void __launch_bounds__(kSmallThreads) sum_partials_kernel(const double* __restrict__ partials, int nb,
                                                                     double* __restrict__ out) {
    __shared__ double v[E];
    reduce_partials<E>(partials, nb, E, 0, v);
    for (int e = threadIdx.x; e < E; e += blockDim.x) out[e] = v[e];
}

template <int K>
__global__ void column_norms_kernel(const double* __restrict__ partials, int nb, double* __restrict__ out) {
    __shared__ double v[K];
    reduce_partials<K>(partials, nb, K, 0, v);
    if (threadIdx.x < K) out[threadIdx.x] = sqrt(v[threadIdx.x]);
}

// The first round of a solve, set on the device so that no host read precedes it: a column is
// active when it is a real (unpadded) column with b_c != 0 whose starting residual (rnorm, or b
// itself from zero when rnorm is null) is above the tolerance; an active column's target is margin
// x tolerance x ||b_c||, an inactive one is frozen from the start; the converged-at counters start
// at zero and the done flag is set when no column is active.
template <int K>
__global__ void __launch_bounds__(kWave) round_start_kernel(int k, double tolerance, double margin,
                                                            const double* __restrict__ bnorm,
                                                            const double* __restrict__ rnorm,
                                                            double* __restrict__ target, int* __restrict__ state) {
    __shared__ int active_cols;
    const int t = threadIdx.x;
    for (int i = t; i < 2 + 2 * kMaxBlock; i += kWave) state[i] = 0;
    bool active = false;
    if (t < K) {
        active = t < k && bnorm[t] > 0.0 && (rnorm == nullptr || !(rnorm[t] / bnorm[t] <= tolerance));
        target[t] = active ? margin * tolerance * bnorm[t] : 0.0;
    }
    if (t == 0) active_cols = 0;
    __syncthreads();
    if (t < K) {
        state[2 + t] = active ? 0 : 1;
        if (active) atomicAdd(&active_cols, 1);
    }
    __syncthreads();
    if (t == 0) state[0] = active_cols == 0;
}

std::int64_t gram_rows(std::int64_t n) { return (n + kGramBlocks - 1) / kGramBlocks; }

// Sub-lanes per column group (G = K / V groups per row) for a level's row kernels: enough that
// each lane walks a few nonzeros, more when the level is too small to fill the machine
// (solver/amg/vcycle.cu:104-120 makes the same trade for its threads-per-row).
int choose_sub(std::int64_t n_rows, std::int64_t nnz, int G, int n_cu) {
    const std::int64_t per_row = n_rows > 0 ? (nnz + n_rows - 1) / n_rows : 1;
    int sub = 1;
    while (G * sub * 2 <= kWave) {
        const bool pays = per_row >= 4 * (sub * 2);
        const bool starved = n_rows * G * sub < static_cast<std::int64_t>(n_cu) * 8 * kWave;
        if (!pays && !starved) break;
        sub *= 2;
    }
    return sub;
}

template <int K>
constexpr int col_groups() {
    return K / vec_width<K>();
}

// rowk_kernel<K, sub, MODE> over the rows of `a`; `diag` is dinv (kCheby) or the row sums
// (kSpmmDiff).
template <int K, int MODE>
unsigned launch_rows(int sub, const CsrView& a, const float* diag, const float* b, const float* x, const float* xp,
                     float cc, float cp, float al, float* out, const int* state, hipStream_t s,
                     unsigned long long* stamps = nullptr) {
    unsigned grid = 0;
    auto go = [&](auto subc) {
        constexpr int SUB = decltype(subc)::value;
        if constexpr (col_groups<K>() * SUB <= kWave) {
            const unsigned g = xcd_grid((a.n + rows_per_vblock<K, SUB>() - 1) / rows_per_vblock<K, SUB>());
            grid = g;
            if (g == 0) return;
            if constexpr (MODE != kSpmmDiffMulti) {
                if (a.ell != nullptr && K <= a.ell_max_k) {
                    rowk_kernel<K, SUB, MODE, true><<<g, kRowThreads, 0, s>>>(
                        a.n, nullptr, reinterpret_cast<const std::int32_t*>(a.ell), nullptr, diag, b, x, xp, cc, cp, al,
                        out, state, stamps, a.ell_w);
                    return;
                }
            }
            rowk_kernel<K, SUB, MODE><<<g, kRowThreads, 0, s>>>(a.n, a.ptr, a.col, a.val, diag, b, x, xp, cc, cp, al,
                                                                out, state, stamps, 0);
        } else {
            throw InvalidArgument("sub-lanes per column group exceed a wavefront");
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
    return grid;
}

// One zero-guess V-cycle (solver/amg/vcycle.cu:560-648) from level `first` (B and X have that
// level's rows). With a dense tail at level t >= first, the cycle below level t - 1 is one product
// with the tail's operator: the restriction into level t, x_t = M b_t, the prolongation from it.
// `state` lets captured replays skip it.
template <int K>
void run_vcycle(const Hierarchy& h, const std::vector<int>& subs, std::vector<DeviceBuffer<float>>& lb,
                std::vector<DeviceBuffer<float>>& lx, std::vector<DeviceBuffer<float>>& lr, float* cb, float* cx,
                const float* B, float* X, const int* state, int n_cu, hipStream_t s, int first = 0,
                const DenseTail* tail = nullptr) {
    const Cheby& ch = h.cheb;
    const int nl = static_cast<int>(h.levels.size());
    const bool dense = tail != nullptr && tail->level >= first && tail->level < nl;
    const int stop = dense ? tail->level : nl;  // levels first .. stop - 1 smooth sparsely
    auto apply_tail = [&](const float* bt, float* xt) {
        const unsigned g = static_cast<unsigned>((tail->n + 15) / 16);
        dense_tail_kernel<K><<<g, kTailThreads, 0, s>>>(tail->n, tail->nq, tail->m.get(), bt, xt, state);
    };
    if (dense && stop == first) {
        apply_tail(B, X);
        return;
    }
    if (nl == first) {
        coarse_kernel<K><<<static_cast<unsigned>(h.n_coarse()), kRowThreads, 0, s>>>(
            static_cast<int>(h.n_coarse()), h.coarse_inv.get(), B, X, state);
        return;
    }
    auto smooth = [&](int li, const float* bi, const float* x0, float* first_buf, float* other,
                      float* final_out) -> float* {
        const Level& L = h.levels[li];
        const int sub = subs[li];
        auto target = [&](int st) {
            return (st + 1 == ch.degree && final_out) ? final_out : ((st % 2) ? other : first_buf);
        };
        if (x0 == nullptr && ch.degree == 2) {
            // Jacobi from zero and the Chebyshev step after it, in one pass (the Jacobi iterate is
            // only ever read by that step).
            float* out = target(1);
            launch_rows<K, kChebyZero>(sub, L.a, L.dinv.get(), bi, bi, nullptr, ch.c_cur[0], ch.alpha0, ch.alpha[0],
                                       out, state, s);
            return out;
        }
        if (x0 == nullptr) {
            jacobi_zero_kernel<K><<<grid_for(L.a.n * K), kBlock, 0, s>>>(L.a.n, L.dinv.get(), bi, ch.alpha0, target(0), state);
        } else {
            launch_rows<K, kCheby>(sub, L.a, L.dinv.get(), bi, x0, x0, 1.0f, 0.0f, ch.alpha0, target(0), state, s);
        }
        float* cur = target(0);
        const float* prev = (x0 == nullptr) ? cur : other;
        for (int st = 1; st < ch.degree; ++st) {
            float* out = target(st);
            const float cprev = (st == 1 && x0 == nullptr) ? 0.0f : ch.c_prev[st - 1];
            launch_rows<K, kCheby>(sub, L.a, L.dinv.get(), bi, cur, prev, ch.c_cur[st - 1], cprev, ch.alpha[st - 1],
                                      out, state, s);
            prev = cur;
            cur = out;
        }
        return cur;
    };
    const bool odd = (ch.degree % 2) != 0;
    for (int i = first; i < stop; ++i) {
        const Level& L = h.levels[i];
        const float* bi = (i == first) ? B : lb[i].get();
        smooth(i, bi, nullptr, odd ? lx[i].get() : lr[i].get(), odd ? lr[i].get() : lx[i].get(), nullptr);
        launch_rows<K, kResidual>(subs[i], L.a, L.dinv.get(), bi, lx[i].get(), lx[i].get(), 0.f, 0.f, 0.f, lr[i].get(),
                                  state, s);
        float* next_b = (i + 1 < nl) ? lb[i + 1].get() : cb;
        const CsrView R = L.r.view();
        launch_rows<K, kSpmm>(choose_sub(R.n, R.nnz, col_groups<K>(), n_cu), R, nullptr, nullptr, lr[i].get(),
                              nullptr, 0.f, 0.f, 0.f, next_b, state, s);
    }
    const float* xc;
    if (dense) {
        apply_tail(lb[stop].get(), lx[stop].get());
        xc = lx[stop].get();
    } else {
        coarse_kernel<K><<<static_cast<unsigned>(h.n_coarse()), kRowThreads, 0, s>>>(
            static_cast<int>(h.n_coarse()), h.coarse_inv.get(), cb, cx, state);
        xc = cx;
    }
    for (int i = stop - 1; i >= first; --i) {
        const Level& L = h.levels[i];
        const float* bi = (i == first) ? B : lb[i].get();
        const CsrView P = L.p.view();
        launch_rows<K, kProlong>(choose_sub(P.n, P.nnz, col_groups<K>(), n_cu), P, nullptr, nullptr, xc, lx[i].get(),
                                 0.f, 0.f, 0.f, lx[i].get(), state, s);
        xc = smooth(i, bi, lx[i].get(), lr[i].get(), lx[i].get(), (i == first) ? X : nullptr);
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
    drop_graphs();
    for (auto g : retired_) (void)hipGraphExecDestroy(g);
    if (host_state_) (void)hipHostFree(host_state_);
}

void BlockSolver::set_multi(const MultiOperator* m) {
    if (m) require(m->n_tissue >= 1 && m->n_tissue <= kMaxTissues, "tissue count must be in [1, 16]");
    multi_ = m;  // the solve picks (or captures) the graph set of this mode
}

static bool same_layout(const MultiOperator& a, const MultiOperator& b) {
    return a.width == b.width && a.off_vals == b.off_vals && a.rowsum == b.rowsum &&
           a.tissue_vals == b.tissue_vals && a.n_tissue == b.n_tissue && a.sigma == b.sigma;
}

BlockSolver::GraphSet& BlockSolver::graph_set() {
    ++use_clock_;
    GraphSet* pick = nullptr;
    for (auto& g : sets_)
        if (g.in_use && g.multi == (multi_ != nullptr) && (!multi_ || same_layout(g.key, *multi_))) pick = &g;
    if (pick == nullptr) {
        pick = &sets_[0];
        for (auto& g : sets_) {
            if (!g.in_use) {
                pick = &g;
                break;
            }
            if (g.last_use < pick->last_use) pick = &g;
        }
        for (int i = 0; i < kNumBlockWidths; ++i) {
            retire(pick->win[i]);
            retire(pick->restart[i]);
            for (auto& x : pick->one[i]) retire(x);
        }
        *pick = GraphSet{};
        pick->in_use = true;
        pick->multi = multi_ != nullptr;
        if (multi_) pick->key = *multi_;
    }
    pick->last_use = use_clock_;
    return *pick;
}

void BlockSolver::retire(hipGraphExec_t& g) {
    if (g == nullptr) return;
    retired_.push_back(g);
    g = nullptr;
    if (retired_.size() > kMaxRetired) {
        (void)hipGraphExecDestroy(retired_.front());
        retired_.erase(retired_.begin());
    }
}

int BlockSolver::graph_execs() const {
    int n = static_cast<int>(retired_.size());
    for (const auto& g : sets_)
        for (int i = 0; i < kNumBlockWidths; ++i) {
            n += (g.win[i] != nullptr) + (g.restart[i] != nullptr);
            for (auto x : g.one[i]) n += x != nullptr;
        }
    return n;
}

template <typename F>
bool BlockSolver::capture(hipGraphExec_t& gx, F&& body) {
    gx = nullptr;
    if (fail_captures_ > 0) {
        --fail_captures_;
        return false;
    }
    if (hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    body();
    hipGraph_t g = nullptr;
    const hipError_t ended = hipStreamEndCapture(s_, &g);
    if (ended != hipSuccess || g == nullptr) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return false;
    }
    const hipError_t made = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (made != hipSuccess) {
        gx = nullptr;
        (void)hipGetLastError();
        return false;
    }
    return true;
}

// Replaced executables are retired rather than destroyed at once (a short queue, the oldest
// destroyed beyond its bound).
void BlockSolver::drop_graphs() {
    tail_ready_ = false;  // whatever dropped the graphs may have changed the V-cycle
    for (auto& g : sets_) {
        for (int i = 0; i < kNumBlockWidths; ++i) {
            retire(g.win[i]);
            retire(g.restart[i]);
            for (auto& x : g.one[i]) retire(x);
        }
        g = GraphSet{};
    }
}

void BlockSolver::ensure(int K) {
    if (cap_k_ >= K) return;
    const std::size_t nk = static_cast<std::size_t>(n()) * K;
    R_.reset(nk, "R");
    Z_.reset(nk, "Z");
    W_.reset(nk, "W");
    for (auto& w : Wx_) {
        w.reset(nk, "W ring");
        // finite before any window reads it
        zero_f32_kernel<<<grid_for(static_cast<std::int64_t>(nk)), kBlock, 0, s_>>>(static_cast<std::int64_t>(nk),
                                                                                   w.get());
    }
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
    red_.reset(static_cast<std::size_t>(2) * K * K, "reduced partials");
    G_.reset(K * K, "G");
    TT_.reset(K * K, "TT");
    AL_.reset(static_cast<std::size_t>(kLag) * K * K, "alpha");
    ran_.reset(kLag, "window slots run");
    D_.reset(K * K, "D");
    norms_.reset(K, "norms");
    bnorm_.reset(K, "right-hand-side norms");
    target_.reset(K, "targets");
    state_.reset(2 + 2 * kMaxBlock, "solver state");
    cap_k_ = K;
    drop_graphs();  // the captured iterations named the buffers just replaced
}

std::size_t BlockSolver::stamp_marks() const {
    // room for two stamps per block of the finest row kernel at any tile (4 rows per block at least)
    return 2 * static_cast<std::size_t>(n() / 4 + 64);
}

unsigned long long* BlockSolver::stamps() { return timing_ ? stamps_.get() : nullptr; }

void BlockSolver::mark(int i, const int* state) {
    if (timing_) stamp_kernel<<<1, 1, 0, s_>>>(stamps_.get() + stamp_marks() + i, state);
}

void BlockSolver::set_timing(bool on) {
    if (on && stamps_.get() == nullptr) {
        stamps_.reset(stamp_marks() + 4, "timing stamps");
        t_sums_.reset(4, "timing sums");
        int dev = 0, khz = 0;
        check(hipGetDevice(&dev), "device");
        check(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev), "wall clock rate");
        require(khz > 0, "the device reports no wall-clock rate");
        clock_khz_ = khz;
    }
    if (t_sums_.get() != nullptr) {
        zero_f64_kernel<<<1, kBlock, 0, s_>>>(4, t_sums_.get());
        sync(s_, "timing reset");
    }
    if (on != timing_) drop_graphs();  // the captured iterations are recorded with or without the stamps
    timing_ = on;
}

void BlockSolver::timing(double* out) const {
    double sums[4] = {};
    if (t_sums_.get() != nullptr) {
        d2h(sums, t_sums_.get(), 4, s_);
        sync(s_, "timing read");
    }
    out[0] = sums[3];
    for (int i = 0; i < 3; ++i) out[1 + i] = sums[3] > 0 ? sums[i] / sums[3] / clock_khz_ : 0.0;
}

void BlockSolver::set_fine_sub(int sub) {
    require(sub >= 0 && sub <= kWave && (sub & (sub - 1)) == 0, "sub must be 0 or a power of two <= 64");
    fine_sub_set_ = sub;
    fine_sub_override_ = sub;
    drop_graphs();
}

void BlockSolver::set_coarse_sub(int sub) {
    require(sub >= 0 && sub <= kWave && (sub & (sub - 1)) == 0, "sub must be 0 or a power of two <= 64");
    coarse_sub_ = sub;
    drop_graphs();
}

template <int K>
void BlockSolver::build_tail_k(int t) {
    const int nt = static_cast<int>(h_.levels[t].a.n);
    const int nq = (nt + 3) / 4;
    const std::size_t size = static_cast<std::size_t>((nt + 15) / 16) * nq * kWave;
    std::vector<int> subs(h_.levels.size());
    for (std::size_t i = 0; i < h_.levels.size(); ++i)
        subs[i] = choose_sub(h_.levels[i].a.n, h_.levels[i].a.nnz, col_groups<K>(), n_cu_);
    if (coarse_sub_ > 0)
        for (std::size_t i = 1; i < subs.size(); ++i) subs[i] = std::min(coarse_sub_, kWave / col_groups<K>());
    tail_.m.reset(size, "dense tail");
    check(hipMemsetAsync(tail_.m.get(), 0, sizeof(float) * size, s_), "dense tail clear");
    DeviceBuffer<float> e(static_cast<std::size_t>(nt) * K, "identity block"), xo(static_cast<std::size_t>(nt) * K,
                                                                                   "tail columns");
    const unsigned g = grid_for(static_cast<std::int64_t>(nt) * K);
    for (int c0 = 0; c0 < nt; c0 += K) {
        identity_cols_kernel<K><<<g, kBlock, 0, s_>>>(nt, c0, e.get());
        run_vcycle<K>(h_, subs, lb_, lx_, lr_, cb_.get(), cx_.get(), e.get(), xo.get(), nullptr, n_cu_, s_, t,
                      nullptr);
        scatter_cols_kernel<K><<<g, kBlock, 0, s_>>>(nt, nq, c0, xo.get(), tail_.m.get());
    }
    sync(s_, "dense tail");  // e and xo are freed on return
    tail_.level = t;
    tail_.n = nt;
    tail_.nq = nq;
}

void BlockSolver::ensure_tail() {
    if (tail_ready_) return;
    tail_.level = -1;
    if (tail_on_) {
        int t = -1;
        for (std::size_t i = 0; i < h_.levels.size(); ++i) {
            if (h_.levels[i].a.n <= kDenseTailMaxRows) {
                t = static_cast<int>(i);
                break;
            }
        }
        if (t >= 0) {
            if (cap_k_ < 1) ensure(1);
            // Formed at the widest allocated width up to 16: the columns of M do not depend on it.
            dispatch_k(std::min(cap_k_, 16), [&](auto kc) { build_tail_k<decltype(kc)::value>(t); });
        }
    }
    tail_ready_ = true;
}

template <int K>
void BlockSolver::vcycle_k(const float* B, float* X, const int* state) {
    // Lanes per column of each level's row kernels: a property of the level and the width alone,
    // unless a fine-level count is in force (clamped to what one wavefront holds at this width).
    std::vector<int> subs(h_.levels.size());
    for (std::size_t i = 0; i < h_.levels.size(); ++i)
        subs[i] = choose_sub(h_.levels[i].a.n, h_.levels[i].a.nnz, col_groups<K>(), n_cu_);
    if (!subs.empty() && fine_sub_override_ > 0) subs[0] = std::min(fine_sub_override_, kWave / col_groups<K>());
    if (coarse_sub_ > 0)
        for (std::size_t i = 1; i < subs.size(); ++i) subs[i] = std::min(coarse_sub_, kWave / col_groups<K>());
    run_vcycle<K>(h_, subs, lb_, lx_, lr_, cb_.get(), cx_.get(), B, X, state, n_cu_, s_, 0,
                  (tail_ready_ && tail_.level >= 0) ? &tail_ : nullptr);
}

template <int K>
int BlockSolver::fine_sub() const {
    if (fine_sub_override_ > 0) return std::min(fine_sub_override_, kWave / col_groups<K>());
    return choose_sub(op_.off.n, op_.off.nnz, col_groups<K>(), n_cu_);
}

// S = A W (differenced), then the partials of G = W^T S and H = W^T R.
template <int K>
void BlockSolver::spmm_block_k(const int* state, const float* W) {
    const std::int64_t nn = n();
    if (multi_) {
        require(multi_->width == K, "multi-matrix values were prepared for another block width");
        CsrView off = op_.off;
        off.val = multi_->off_vals;
        off.ell = nullptr;  // per-column values follow the CSR entries
        spmm_grid_ = launch_rows<K, kSpmmDiffMulti>(fine_sub<K>(), off, multi_->rowsum, nullptr, W, nullptr, 0.f,
                                                    0.f, 0.f, S_.get(), state, s_, stamps());
    } else {
        spmm_grid_ = launch_rows<K, kSpmmDiff>(fine_sub<K>(), op_.off, op_.rowsum, nullptr, W, nullptr, 0.f, 0.f,
                                               0.f, S_.get(), state, s_, stamps());
    }
    launch_gram<K, true>(nn, W, S_.get(), R_.get(), partials_.get(), gram_rows(nn), state, s_);
}

template <int K>
void BlockSolver::iteration_k(const SolveOptions& opt, int j) {
    const std::int64_t nn = n();
    const std::int64_t rpb = gram_rows(nn);
    const int* st = state_.get();
    int* stw = state_.get();
    const bool last = j == kLag - 1;
    const float* w_in = wslot(j);
    float* w_out = last ? wslot(0) : wslot(j + 1);
    double* al = AL_.get() + static_cast<std::int64_t>(j) * K * K;
    mark(0, st);
    spmm_block_k<K>(st, w_in);
    reduce_cols(partials_.get(), kGramBlocks, 2 * K * K, red_.get(), st, s_);
    coef_kernel<K><<<1, kWave, 0, s_>>>(red_.get(), opt.drop_tol, TT_.get(), al, stw,
                                        (multi_ || opt.independent) ? 1 : 0, ran_.get() + j);
    update_r_kernel<K><<<kGramBlocks, xr_threads<K>(), 0, s_>>>(nn, S_.get(), al, R_.get(), partials_.get(), rpb, st);
    reduce_cols(partials_.get(), kGramBlocks, K, red_.get(), st, s_);
    // one wavefront: the kernel reads K norms (small workgroups find room beside the sweep's N-body)
    converge_kernel<K><<<1, kWave, 0, s_>>>(red_.get(), kGramBlocks, target_.get(), norms_.get(), stw);
    mark(1, st);
    vcycle_k<K>(R_.get(), Z_.get(), st);
    mark(2, st);
    launch_gram<K, false>(nn, S_.get(), Z_.get(), nullptr, partials_.get(), rpb, st, s_);
    reduce_cols(partials_.get(), kGramBlocks, K * K, red_.get(), st, s_);
    dmat_kernel<K><<<1, kBlock, 0, s_>>>(red_.get(), kGramBlocks, TT_.get(), D_.get(), stw,
                                                (multi_ || opt.independent) ? 1 : 0);
    if (last)  // before this iteration's W update, which overwrites slot 0
        flush_x_kernel<K><<<grid_for(nn * col_groups<K>(), kRowThreads), kRowThreads, 0, s_>>>(
            nn, wslot(0), wslot(1), wslot(2), wslot(3), AL_.get(), ran_.get(), kLag, X64_.get());
    update_w_kernel<K><<<grid_for(nn * col_groups<K>(), kRowThreads), kRowThreads, 0, s_>>>(nn, Z_.get(), D_.get(),
                                                                                           w_in, w_out, st);
    mark(3, st);
    if (timing_)
        timing_accumulate_kernel<<<1, kBlock, 0, s_>>>(stamps_.get(), static_cast<int>(spmm_grid_),
                                                       stamps_.get() + stamp_marks(), t_sums_.get());
}

template <int K>
void BlockSolver::window_k(const SolveOptions& opt, int j0, int L) {
    if (j0 == 0) clear_ran_kernel<<<1, kWave, 0, s_>>>(ran_.get());
    for (int j = j0; j < j0 + L; ++j) iteration_k<K>(opt, j);
}

template <int K>
void BlockSolver::flush_pending_k(int L) {
    const std::int64_t nn = n();
    flush_x_kernel<K><<<grid_for(nn * col_groups<K>(), kRowThreads), kRowThreads, 0, s_>>>(
        nn, wslot(0), wslot(1), wslot(2), wslot(3), AL_.get(), ran_.get(), L, X64_.get());
}

template <int K>
SolveStats BlockSolver::solve_k(const float* B, double* X, const double* X0, int k, const SolveOptions& opt) {
    ensure(K);
    ensure_tail();
    const std::int64_t nn = n();
    const std::int64_t nk = nn * K;
    const std::int64_t rpb = gram_rows(nn);
    SolveStats stats;
    stats.column_iterations.assign(k, 0);
    stats.relative_residual.assign(k, 0.0);
    double* const Xi = X64_.get();  // solver-owned, so a captured graph never names a caller buffer
    const double margin = opt.inner_margin > 0.0 ? opt.inner_margin : margin_for_width(K);

    std::vector<double> bnorm(K), rnorm(K);
    // R64 = B - A X in FP64 and its column norms into `norms`, without a wait.
    auto residual_norms = [&](bool zero, double* norms) {
        if (multi_)
            residual64_kernel<K, true><<<kGramBlocks, kGramThreads, 0, s_>>>(
                nn, op_.a32.ptr, op_.a32.col, multi_->tissue_vals, op_.a32.nnz, multi_->n_tissue, multi_->sigma, B, Xi,
                R64_.get(), partials_.get(), rpb, zero);
        else
            residual64_kernel<K, false><<<kGramBlocks, kGramThreads, 0, s_>>>(
                nn, op_.a32.ptr, op_.a32.col, op_.val64, op_.a32.nnz, 0, nullptr, B, Xi, R64_.get(), partials_.get(), rpb,
                zero);
        column_norms_kernel<K><<<1, kSmallThreads, 0, s_>>>(partials_.get(), kGramBlocks, norms);
    };
    auto true_residual = [&](bool zero) {
        residual_norms(zero, norms_.get());
        d2h(rnorm.data(), norms_.get(), K, s_);
        sync(s_, "true residual");
    };
    auto check_bnorm = [&] {
        for (int c = 0; c < k; ++c) require(bnorm[c] > 0.0, "zero right-hand-side column " + std::to_string(c));
    };
    residual_norms(true, bnorm_.get());  // x_zero: R64 = B, so these are ||b_c||
    if (X0 != nullptr) {
        d2d(Xi, X0, static_cast<std::size_t>(nk), s_);
        residual_norms(false, norms_.get());
    } else {
        zero_f64_kernel<<<grid_for(nk), kBlock, 0, s_>>>(nk, Xi);
    }
    // The first round's targets and flags follow from the norms alone, so they are set on the
    // device and ||b|| comes back with the first convergence read (a column whose guess already
    // meets the tolerance starts frozen; if all do, the round runs no iteration).
    bool bnorm_pending = opt.tolerance < 1.0;
    if (!bnorm_pending) {
        d2h(bnorm.data(), bnorm_.get(), K, s_);
        if (X0 != nullptr) d2h(rnorm.data(), norms_.get(), K, s_);
        sync(s_, "initial norms");
        check_bnorm();
        if (X0 == nullptr) rnorm = bnorm;
    }

    std::vector<int> host_state(2 + 2 * kMaxBlock, 0);
    bool use_graph = opt.use_graph;  // cleared for the rest of this solve if a capture fails
    int total_iters = 0;
    std::vector<int> col_iters(K, 0);
    int slot = 0;
    while (kBlockWidths[slot] != K) ++slot;
    for (int round = 0; round < opt.max_refine; ++round) {
        const bool device_start = round == 0 && bnorm_pending;
        if (!device_start) {
            bool all_ok = true;
            for (int c = 0; c < k; ++c) {
                stats.relative_residual[c] = rnorm[c] / bnorm[c];
                if (!(stats.relative_residual[c] <= opt.tolerance)) all_ok = false;
            }
            if (all_ok || total_iters >= opt.max_iters) break;
        }
        // Inner solve on the current true residual (a restart: W = M R).
        std::fill(host_state.begin(), host_state.end(), 0);
        host_state[1] = total_iters;
        if (device_start) {
            round_start_kernel<K><<<1, kWave, 0, s_>>>(k, opt.tolerance, margin, bnorm_.get(),
                                                      X0 != nullptr ? norms_.get() : nullptr, target_.get(),
                                                      state_.get());
        } else {
            std::vector<double> tgt(K, 0.0);
            for (int c = 0; c < K; ++c) {
                const bool active = c < k && !(stats.relative_residual[c] <= opt.tolerance);
                tgt[c] = active ? margin * opt.tolerance * bnorm[c] : 0.0;
                host_state[2 + c] = active ? 0 : 1;
            }
            const std::size_t tb = sizeof(double) * K, sb = sizeof(int) * host_state.size();
            char* h = static_cast<char*>(stage_.fill(tb + sb));
            std::memcpy(h, tgt.data(), tb);
            std::memcpy(h + tb, host_state.data(), sb);
            stage_.send(target_.get(), 0, tb, s_);
            stage_.send(state_.get(), tb, sb, s_);
            stage_.done(s_);
        }
        // The restart writes M R straight into the first W slot (the iteration's V-cycle writes Z
        // afresh before anything reads it).
        auto restart = [&] {
            cast_f64_f32_kernel<<<grid_for(nk), kBlock, 0, s_>>>(nk, R64_.get(), R_.get());
            vcycle_k<K>(R_.get(), W_.get(), nullptr);
        };
        const int diag = (multi_ || opt.independent) ? 1 : 0;
        hipGraphExec_t win = nullptr, one[kLag] = {};
        if (use_graph) {
            GraphSet& gs = graph_set();
            if (gs.restart[slot] == nullptr && !capture(gs.restart[slot], restart)) use_graph = false;
        }
        if (use_graph)
            check(hipGraphLaunch(graph_set().restart[slot], s_), "graph launch");
        else
            restart();
        if (use_graph) {
            GraphSet& gs = graph_set();
            if (gs.win[slot] == nullptr || gs.drop[slot] != opt.drop_tol || gs.diag[slot] != diag) {
                // a whole window (j0 = 0, L = kLag), then one iteration at each ring position
                bool ok = true;
                for (int v = -1; v < kLag && ok; ++v) {
                    hipGraphExec_t& gx = v < 0 ? gs.win[slot] : gs.one[slot][v];
                    retire(gx);
                    ok = capture(gx, [&] { v < 0 ? window_k<K>(opt, 0, kLag) : window_k<K>(opt, v, 1); });
                }
                if (ok) {
                    gs.drop[slot] = opt.drop_tol;
                    gs.diag[slot] = diag;
                } else {
                    // A failed capture records and runs nothing: this solve launches its iterations
                    // directly; the next solve tries to capture again.
                    retire(gs.restart[slot]);
                    retire(gs.win[slot]);
                    for (auto& x : gs.one[slot]) retire(x);
                    use_graph = false;
                }
            }
            if (use_graph) {
                win = gs.win[slot];
                for (int v = 0; v < kLag; ++v) one[v] = gs.one[slot][v];
            }
        }
        // Bounded by the launches the host issues, not only by the device counter.
        int launched = total_iters;
        bool done = false;
        const int hint = round == 0 ? last_first_round_[slot] : 0;
        int reads = 0;
        int pos = 0;  // ring position of the next iteration: the round's iteration count mod kLag
        while (launched < opt.max_iters) {
            int want = opt.check_every;
            // the first read after whole windows short of the hint, then after every iteration
            if (hint > 0) want = reads == 0 ? std::max(1, (hint - 1) / kLag * kLag) : (reads <= 4 ? 1 : opt.check_every);
            const int batch = std::min(want, opt.max_iters - launched);
            ++reads;
            // whole windows where the ring position is 0 and four or more remain, else one
            // iteration at the current position
            for (int b = 0; b < batch;) {
                const int L = (pos == 0 && batch - b >= kLag) ? kLag : 1;
                if (use_graph)
                    check(hipGraphLaunch(L == kLag ? win : one[pos], s_), "graph launch");
                else
                    window_k<K>(opt, pos, L);
                b += L;
                pos = (pos + L) % kLag;
            }
            launched += batch;
            if (bnorm_pending) d2h(bnorm.data(), bnorm_.get(), K, s_);
            d2h(host_state_, state_.get(), 2 + 2 * K, s_);
            sync(s_, "convergence read");
            if (bnorm_pending) {
                bnorm_pending = false;
                check_bnorm();
            }
            total_iters = host_state_[1];
            if (host_state_[0]) {
                done = true;
                break;
            }
        }
        if (pos != 0) flush_pending_k<K>(pos);  // the round ended inside a window
        if (!done) total_iters = launched;
        if (round == 0) last_first_round_[slot] = total_iters - host_state[1];
        for (int c = 0; c < K; ++c)  // set only for columns that converged in this round
            if (host_state_[2 + K + c] > 0) col_iters[c] = host_state_[2 + K + c];
        stats.vcycles += total_iters - host_state[1] + 1;
        if (bnorm_pending) {  // the round made no convergence read
            bnorm_pending = false;
            d2h(bnorm.data(), bnorm_.get(), K, s_);
            sync(s_, "right-hand-side norms");
            check_bnorm();
        }
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
    dispatch_k(K, [&](auto kc) {
        constexpr int KK = decltype(kc)::value;
        const int sub = fine_sub_override_ > 0 ? std::min(fine_sub_override_, kWave / col_groups<KK>())
                                               : choose_sub(op_.a32.n, op_.a32.nnz, col_groups<KK>(), n_cu_);
        launch_rows<KK, kSpmm>(sub, op_.a32, nullptr, nullptr, X, nullptr, 0.f, 0.f, 0.f, Y, nullptr, s_);
    });
}

void BlockSolver::spmm_diff(const float* X, float* Y, int K) {
    dispatch_k(K, [&](auto kc) {
        constexpr int KK = decltype(kc)::value;
        launch_rows<KK, kSpmmDiff>(fine_sub<KK>(), op_.off, op_.rowsum, nullptr, X, nullptr, 0.f, 0.f, 0.f, Y,
                                   nullptr, s_);
    });
}

void BlockSolver::cross_gram32(const float* A, const float* B, double* out_host) {
    ensure(32);
    const std::int64_t nn = n();
    launch_gram<32, false>(nn, A, B, nullptr, partials_.get(), gram_rows(nn), nullptr, s_);
    sum_partials_kernel<32 * 32><<<1, kSmallThreads, 0, s_>>>(partials_.get(), kGramBlocks, G_.get());
    d2h(out_host, G_.get(), 32 * 32, s_);
    sync(s_, "cross gram");
}

void BlockSolver::gram(const float* A, const float* B1, const float* B2, double* out_host, int K) {
    ensure(K);
    const std::int64_t nn = n();
    const int E = (B2 != nullptr ? 2 : 1) * K * K;
    dispatch_k(K, [&](auto kc) {
        constexpr int KK = decltype(kc)::value;
        if (B2 != nullptr)
            launch_gram<KK, true>(nn, A, B1, B2, partials_.get(), gram_rows(nn), nullptr, s_);
        else
            launch_gram<KK, false>(nn, A, B1, nullptr, partials_.get(), gram_rows(nn), nullptr, s_);
    });
    check(hipGetLastError(), "gram launch");
    reduce_cols(partials_.get(), kGramBlocks, E, red_.get(), nullptr, s_);
    d2h(out_host, red_.get(), E, s_);
    sync(s_, "gram");
}

void BlockSolver::guess_reset(int capacity, int K) {
    require(capacity >= 1 && capacity <= 2 && capacity * K <= 32,
            "the guess basis holds one or two blocks, 32 vectors at most");
    ensure(K);
    const std::size_t nk = static_cast<std::size_t>(n()) * K;
    for (int i = 0; i < capacity; ++i) {
        g_x_[i].ensure(nk, "guess basis");
        g_b_[i].ensure(nk, "guess basis right-hand sides");
    }
    g_mat_.ensure(static_cast<std::size_t>(8) * K * K, "guess Grams");
    g_cap_ = capacity;
    g_k_ = K;
    g_count_ = 0;
    g_newest_ = 0;
    g_pending_ = false;
}

// g_mat_ layout (K x K blocks): D of slot 0 and 1, the cross block, the pending cross block, the
// guess's Grams (two blocks) and its coefficients (two blocks).
void BlockSolver::guess_push(const double* X, const float* B, int K) {
    require(g_cap_ > 0 && K == g_k_, "guess_push: the basis was reset for another width");
    const std::int64_t nn = n();
    const std::int64_t nk = nn * K;
    const std::int64_t E = static_cast<std::int64_t>(K) * K;
    const int slot = g_cap_ == 1 ? 0 : 1 - g_newest_;
    const int older = g_newest_;
    cast_f64_f32_kernel<<<grid_for(nk), kBlock, 0, s_>>>(nk, X, g_x_[slot].get());
    d2d(g_b_[slot].get(), B, static_cast<std::size_t>(nk), s_);
    double* const m = g_mat_.get();
    dispatch_k(K, [&](auto kc) {
        constexpr int KK = decltype(kc)::value;
        const std::int64_t rpb = gram_rows(nn);
        launch_gram<KK, false>(nn, g_x_[slot].get(), g_b_[slot].get(), nullptr, partials_.get(), rpb, nullptr, s_);
        reduce_cols(partials_.get(), kGramBlocks, static_cast<int>(E), m + slot * E, nullptr, s_);
        if (g_cap_ == 2 && g_count_ >= 1) {  // the cross block X_older^T B_new
            if (g_pending_)  // the last guess was for this block: its projection onto the older block
                d2d(m + 2 * E, m + 3 * E, static_cast<std::size_t>(E), s_);
            else {
                launch_gram<KK, false>(nn, g_x_[older].get(), g_b_[slot].get(), nullptr, partials_.get(), rpb,
                                       nullptr, s_);
                reduce_cols(partials_.get(), kGramBlocks, static_cast<int>(E), m + 2 * E, nullptr, s_);
            }
        }
    });
    g_newest_ = slot;
    g_count_ = std::min(g_count_ + 1, g_cap_);
    g_pending_ = false;
}

void BlockSolver::guess(const float* b, int K, double drop_tol, double* x0) {
    require(g_count_ > 0 && K == g_k_, "guess: no basis block of this width");
    const std::int64_t nn = n();
    const std::int64_t E = static_cast<std::int64_t>(K) * K;
    double* const m = g_mat_.get();
    double* const gb = m + 4 * E;
    double* const coef = m + 6 * E;
    const int newest = g_newest_, older = 1 - g_newest_;
    dispatch_k(K, [&](auto kc) {
        constexpr int KK = decltype(kc)::value;
        const std::int64_t rpb = gram_rows(nn);
        const unsigned grid = grid_for(nn * (KK / vec_width<KK>()));
        if (g_count_ == 2) {
            if constexpr (2 * KK <= 32) {
                launch_gram<KK, true>(nn, b, g_x_[older].get(), g_x_[newest].get(), partials_.get(), rpb, nullptr,
                                      s_);
                reduce_cols(partials_.get(), kGramBlocks, static_cast<int>(2 * E), gb, nullptr, s_);
                galerkin_solve_kernel<KK, 2><<<1, kWave, 0, s_>>>(m + older * E, m + newest * E, m + 2 * E, gb,
                                                                  drop_tol, coef, m + 3 * E);
                galerkin_expand_kernel<KK, 2><<<grid, kBlock, 0, s_>>>(nn, g_x_[older].get(), g_x_[newest].get(),
                                                                       coef, x0);
            }
        } else {
            launch_gram<KK, false>(nn, b, g_x_[newest].get(), nullptr, partials_.get(), rpb, nullptr, s_);
            reduce_cols(partials_.get(), kGramBlocks, static_cast<int>(E), gb, nullptr, s_);
            galerkin_solve_kernel<KK, 1><<<1, kWave, 0, s_>>>(m + newest * E, nullptr, nullptr, gb, drop_tol, coef,
                                                              m + 3 * E);
            galerkin_expand_kernel<KK, 1><<<grid, kBlock, 0, s_>>>(nn, g_x_[newest].get(), nullptr, coef, x0);
        }
    });
    check(hipGetLastError(), "Galerkin guess launch");
    g_pending_ = true;
}

void BlockSolver::vcycle(const float* B, float* X, int K) {
    ensure(K);
    ensure_tail();
    dispatch_k(K, [&](auto kc) { vcycle_k<decltype(kc)::value>(B, X, nullptr); });
}

double BlockSolver::bench_spmm(int K, int reps, int mode, int sub) {
    require(padded_width(K) == K, "K must be a compiled block width");
    require(mode >= 0 && mode <= 2, "mode must be 0, 1 or 2");
    require(sub >= 0 && sub <= kWave && (sub & (sub - 1)) == 0, "sub must be 0 or a power of two <= 64");
    require(sub == 0 || (K / (K >= 4 ? 4 : K)) * sub <= kWave, "sub-lanes exceed a wavefront for this K");
    ensure(K);
    const std::int64_t nk = n() * K;
    fill_pattern_kernel<<<grid_for(nk), kBlock, 0, s_>>>(nk, W_.get());
    fill_pattern_kernel<<<grid_for(nk), kBlock, 0, s_>>>(nk, R_.get());
    struct Restore {
        int& v;
        int keep;
        ~Restore() { v = keep; }
    } restore{fine_sub_override_, fine_sub_set_};
    if (sub > 0) fine_sub_override_ = sub;
    auto launch = [&] {
        if (mode == 2)
            dispatch_k(K, [&](auto kc) { spmm_block_k<decltype(kc)::value>(nullptr, W_.get()); });
        else if (mode == 1)
            spmm_diff(W_.get(), S_.get(), K);
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

// Streaming copy out[i] = in[i] over float4s: the calibration kernel of the HBM traffic counters
// (FETCH_SIZE / WRITE_SIZE of a launch that moves a known number of bytes) and the practical
// bandwidth ceiling beside the published peak.
// Each thread moves U float4s, all U loads issued before the first store; the grid covers the
// array once (no loop), so every wave has U loads in flight.
template <int U>
__global__ void __launch_bounds__(kBlock) stream_copy_kernel(std::int64_t n4, const float4* __restrict__ in,
                                                             float4* __restrict__ out) {
    const std::int64_t base = (static_cast<std::int64_t>(blockIdx.x) * U) * blockDim.x + threadIdx.x;
    float4 v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const std::int64_t i = base + static_cast<std::int64_t>(u) * blockDim.x;
        if (i < n4) v[u] = in[i];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
        const std::int64_t i = base + static_cast<std::int64_t>(u) * blockDim.x;
        if (i < n4) out[i] = v[u];
    }
}

double BlockSolver::bench_copy(std::int64_t bytes, int reps) {
    require(bytes >= 16 && bytes % 16 == 0, "bytes must be a positive multiple of 16");
    const std::int64_t n4 = bytes / 16;
    DeviceBuffer<float> a(bytes / 4), b(bytes / 4);
    fill_pattern_kernel<<<grid_for(bytes / 4), kBlock, 0, s_>>>(bytes / 4, a.get());
    constexpr int kCopyU = 4;
    const std::int64_t grid = (n4 + static_cast<std::int64_t>(kBlock) * kCopyU - 1) / (static_cast<std::int64_t>(kBlock) * kCopyU);
    require(grid <= 0x7fffffff, "copy too large");
    auto launch = [&] {
        stream_copy_kernel<kCopyU><<<static_cast<unsigned>(grid), kBlock, 0, s_>>>(
            n4, reinterpret_cast<const float4*>(a.get()), reinterpret_cast<float4*>(b.get()));
    };
    launch();
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

int build_ell(const CsrView& a, DeviceBuffer<int2>& pairs, hipStream_t s) {
    pairs.release();
    if (a.n <= 0) return 0;
    std::vector<std::int32_t> rp(static_cast<std::size_t>(a.n + 1));
    d2h(rp.data(), a.ptr, rp.size(), s);
    sync(s, "row pointers");
    int w = 0;
    for (std::int64_t r = 0; r < a.n; ++r) w = std::max(w, rp[r + 1] - rp[r]);
    if (w == 0 || w > kEllMaxWidth) return 0;
    const std::int64_t slots = (a.n + kEllSlice - 1) / kEllSlice * kEllSlice * w;
    pairs.reset(static_cast<std::size_t>(slots), "ell pairs");
    ell_fill_kernel<<<static_cast<unsigned>(std::min<std::int64_t>((slots + kBlock - 1) / kBlock, 65536)), kBlock, 0, s>>>(
        a.n, a.ptr, a.col, a.val, w, pairs.get());
    check(hipGetLastError(), "ell fill");
    return w;
}

// ----------------------------------------------------------------------------------------------
// Gram entry point that allocates nothing
// ----------------------------------------------------------------------------------------------

// Partials of gram_noalloc: kGramBlocks chunks x 2 K^2 entries at the largest width (8 MiB). Each
// device holds its own copy for the life of the process; every call on a device shares it and
// overwrites all the partials it reads (nothing has to start at zero), so calls that overlap on
// one device race.
constexpr std::size_t kNoAllocScratch = static_cast<std::size_t>(kGramBlocks) * 2 * kMaxBlock * kMaxBlock;
static __device__ __align__(256) double scratch[kNoAllocScratch];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_lock;
double* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (the symbol's host-side name is not a
// device address), looked up once per device.
double* scratch_on_current_device() {
    int dev = 0;
    check(hipGetDevice(&dev), "current device");
    require(dev >= 0 && dev < kMaxDevices, "device index beyond the scratch table");
    std::lock_guard<std::mutex> hold(g_scratch_lock);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "scratch symbol address");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    return g_scratch_ptr[dev];
}
}  // namespace

void gram_noalloc(std::int64_t n, int K, const float* A, const float* B1, const float* B2, double* out,
                  hipStream_t s) {
    require(n >= 1, "n must be at least 1");
    require(K >= 1 && K <= kMaxBlock && padded_width(K) == K, "K must be one of 1, 2, 4, 8, 16, 32");
    require(A != nullptr && B1 != nullptr && out != nullptr, "null operand");
    const int E = (B2 != nullptr ? 2 : 1) * K * K;
    require(static_cast<std::size_t>(kGramBlocks) * E <= kNoAllocScratch, "Gram partials exceed the scratch");
    double* const part = scratch_on_current_device();
    const std::int64_t rpb = gram_rows(n);
    dispatch_k(K, [&](auto kc) {
        constexpr int KK = decltype(kc)::value;
        if (B2 != nullptr)
            launch_gram<KK, true>(n, A, B1, B2, part, rpb, nullptr, s);
        else
            launch_gram<KK, false>(n, A, B1, nullptr, part, rpb, nullptr, s);
    });
    check(hipGetLastError(), "gram launch");
    reduce_cols(part, kGramBlocks, E, out, nullptr, s);
    check(hipGetLastError(), "gram reduction launch");
}

}  // namespace cn
