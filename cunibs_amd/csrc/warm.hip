// Reduced-basis warm start for the block PCG.
//
// The basis W (n_red x S) is a set of solutions the solver itself produced, stored in BF16 row-major
// (row r holds the S basis values of unknown r). The basis is defined as that BF16 array, so every
// product below sees the same W exactly:
//   Galerkin matrix   G  = W^T (A W)                       (FP64, built once; A applied in FP32)
//   projection        c  = P (W^T B),  P = pseudo-inverse of G (supplied by the caller)
//   expansion         x0 = W c                              (BF16 MFMA, FP32 accumulation)
// x0 is the A-norm best approximation of A^-1 B in span(W), up to the rounding of P and of c,
// which is carried as a BF16 high part plus a BF16 low part. PCG then iterates from x0 to the
// same tolerance against ||b||, so the warm start changes only the iteration count.
//
// No reference counterpart: the reference's closest analogue is the Galerkin-projected initial
// guess of its conductivity-UQ loop (uq/conductivity/run.py:154-200).
#include <hip/hip_bf16.h>

#include "warm.hpp"

namespace cn {

namespace {

__device__ __forceinline__ unsigned short bf16_bits(float x) {
    // Round to nearest even (finite inputs).
    const unsigned u = __float_as_uint(x);
    const unsigned r = u + 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<unsigned short>(r >> 16);
}

__device__ __forceinline__ float bf16_value(unsigned short b) { return __uint_as_float(static_cast<unsigned>(b) << 16); }

// basis[r * S + (s0 + c)] = bf16(X[r * K + c]) for c < k: append a block's solutions as columns.
__global__ void append_kernel(const double* __restrict__ X, std::int64_t n, int k, int K, int S, int s0,
                              unsigned short* __restrict__ basis) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n * k) return;
    const std::int64_t r = t / k;
    const int c = static_cast<int>(t % k);
    basis[r * S + s0 + c] = bf16_bits(static_cast<float>(X[r * K + c]));
}

// out (n, 32) FP32 = columns [s0, s0 + 32) of the basis (zero beyond S).
__global__ void gather_chunk_kernel(const unsigned short* __restrict__ basis, std::int64_t n, int S, int s0,
                                    float* __restrict__ out) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n * 32) return;
    const std::int64_t r = t >> 5;
    const int c = static_cast<int>(t & 31);
    out[t] = (s0 + c < S) ? bf16_value(basis[r * S + s0 + c]) : 0.0f;
}

// Partials of W^T B: one thread per basis column s (blockDim = S rounded up to a wave multiple),
// the block walks its row range; B rows are read whole by every thread (a broadcast).
template <int K>
__global__ void __launch_bounds__(512) project_kernel(const unsigned short* __restrict__ basis, std::int64_t n, int S,
                                                      const float* __restrict__ B, std::int64_t rows_per_block,
                                                      double* __restrict__ partials) {
    const int s = threadIdx.x;
    const std::int64_t lo = static_cast<std::int64_t>(blockIdx.x) * rows_per_block;
    const std::int64_t hi = min(lo + rows_per_block, n);
    double acc[K];
#pragma unroll
    for (int c = 0; c < K; ++c) acc[c] = 0.0;
    if (s < S) {
        for (std::int64_t r = lo; r < hi; ++r) {
            const double w = static_cast<double>(bf16_value(basis[r * S + s]));
#pragma unroll
            for (int c = 0; c < K; ++c) acc[c] = fma(w, static_cast<double>(B[r * K + c]), acc[c]);
        }
        double* out = partials + (static_cast<std::int64_t>(blockIdx.x) * S + s) * K;
#pragma unroll
        for (int c = 0; c < K; ++c) out[c] = acc[c];
    }
}

// WtB[s, c] = sum over blocks (ascending) of the partials; then c = P WtB (FP64) split into BF16
// hi/lo halves laid out for the expansion: coef[(half * S16 + s) * 16 + c], zero-padded to
// S16 = ceil(S / 16) * 16 rows and 16 columns.
template <int K>
__global__ void coefficients_kernel(const double* __restrict__ partials, int nb, int S, const double* __restrict__ P,
                                    int k, unsigned short* __restrict__ coef, int S16) {
    extern __shared__ double wtb[];  // S * K
    for (int e = threadIdx.x; e < S * K; e += blockDim.x) {
        double acc = 0.0;
        for (int b = 0; b < nb; ++b) acc += partials[static_cast<std::int64_t>(b) * S * K + e];
        wtb[e] = acc;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < S16 * 16; e += blockDim.x) {
        const int s = e / 16, c = e % 16;
        double v = 0.0;
        if (s < S && c < k) {
            for (int j = 0; j < S; ++j) v = fma(P[static_cast<std::int64_t>(s) * S + j], wtb[j * K + c], v);
        }
        const float vf = static_cast<float>(v);
        const unsigned short hi = bf16_bits(vf);
        const unsigned short lo = bf16_bits(vf - bf16_value(hi));
        coef[e] = hi;
        coef[S16 * 16 + e] = lo;
    }
}

typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4_t;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4_t;

// x0 (n, K) FP64 = W c: one wavefront per 16-row tile, v_mfma_f32_16x16x16bf16_1k over the basis
// in 16-column steps, the high and low coefficient halves accumulated into the same FP32 tile.
// Operand layout of the 16x16x16 BF16 MFMA (CDNA3/4 ISA): lane l supplies A[l % 16][4 (l / 16) + j]
// and B[4 (l / 16) + j][l % 16], j = 0..3, and receives D[4 (l / 16) + j][l % 16].
__global__ void __launch_bounds__(256) expand_kernel(const unsigned short* __restrict__ basis, std::int64_t n, int S,
                                                     int S16, const unsigned short* __restrict__ coef, int k, int K,
                                                     double* __restrict__ x0) {
    const int lane = threadIdx.x & 63;
    const std::int64_t tile = static_cast<std::int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const std::int64_t r0 = tile * 16;
    if (r0 >= n) return;
    const int i = lane & 15, q = lane >> 4;
    const std::int64_t row = r0 + i;
    floatx4_t acc = {0.f, 0.f, 0.f, 0.f};
    for (int s0 = 0; s0 < S16; s0 += 16) {
        bf16x4_t a;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int s = s0 + 4 * q + j;
            a[j] = (row < n && s < S) ? static_cast<short>(basis[row * S + s]) : static_cast<short>(0);
        }
        bf16x4_t bh, bl;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int s = s0 + 4 * q + j;
            bh[j] = static_cast<short>(coef[s * 16 + i]);
            bl[j] = static_cast<short>(coef[S16 * 16 + s * 16 + i]);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bh, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bl, acc, 0, 0, 0);
    }
    // acc[j] = x0[r0 + 4 q + j][i]
    if (i < K) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const std::int64_t r = r0 + 4 * q + j;
            if (r < n) x0[r * K + i] = (i < k) ? static_cast<double>(acc[j]) : 0.0;
        }
    }
}

}  // namespace

void WarmBasis::reset(std::int64_t n_rows, int capacity) {
    require(capacity >= 1 && capacity <= kMaxWarmBasis, "warm-start capacity must be in [1, 512]");
    n = n_rows;
    cap = capacity;
    count = 0;
    ready = false;
    basis.reset(static_cast<std::size_t>(n) * cap, "warm basis");
    check(hipMemset(basis.get(), 0, sizeof(unsigned short) * n * cap), "warm basis clear");
}

int WarmBasis::append(const double* X, int k, int K, hipStream_t s) {
    require(cap > 0, "warm-start basis not reset");
    const int take = std::min(k, cap - count);
    if (take <= 0) return 0;
    if (const unsigned g = grid_for(n * take))
        append_kernel<<<g, kBlock, 0, s>>>(X, n, take, K, cap, count, basis.get());
    count += take;
    ready = false;
    return take;
}

void WarmBasis::chunk(int s0, float* out, hipStream_t s) const {
    if (const unsigned g = grid_for(n * 32)) gather_chunk_kernel<<<g, kBlock, 0, s>>>(basis.get(), n, cap, s0, out);
}

void WarmBasis::set_projector(const double* P_host, int S, hipStream_t s) {
    require(S == count && S >= 1, "projector size must equal the number of basis columns");
    P.reset(static_cast<std::size_t>(S) * S, "warm projector");
    h2d(P.get(), P_host, static_cast<std::size_t>(S) * S, s);
    sync(s, "warm projector");
    ready = true;
}

void WarmBasis::initial_guess(const float* B, int k, int K, double* x0, hipStream_t s) {
    require(ready, "warm start requested but no projector is set");
    const int S = cap;  // row stride; columns >= count are zero in the basis and in P's image
    const int S_used = count;
    const int S16 = (S + 15) / 16 * 16;
    const int nb = kProjectBlocks;
    const std::int64_t rpb = (n + nb - 1) / nb;
    partials.ensure(static_cast<std::size_t>(nb) * S * K, "warm partials");
    coef.ensure(static_cast<std::size_t>(2) * S16 * 16, "warm coefficients");
    Pfull.ensure(static_cast<std::size_t>(S) * S, "warm projector (padded)");
    // P padded to the basis stride: rows/columns >= count are zero.
    if (padded_for != S_used) {
        std::vector<double> host(static_cast<std::size_t>(S) * S, 0.0);
        std::vector<double> small(static_cast<std::size_t>(S_used) * S_used);
        d2h(small.data(), P.get(), small.size(), s);
        sync(s, "projector");
        for (int a = 0; a < S_used; ++a)
            for (int b = 0; b < S_used; ++b) host[static_cast<std::size_t>(a) * S + b] = small[a * S_used + b];
        h2d(Pfull.get(), host.data(), host.size(), s);
        padded_for = S_used;
    }
    const int threads = (S + 63) / 64 * 64;
    auto run = [&](auto kc) {
        constexpr int KK = decltype(kc)::value;
        project_kernel<KK><<<nb, threads, 0, s>>>(basis.get(), n, S, B, rpb, partials.get());
        coefficients_kernel<KK><<<1, 512, sizeof(double) * S * KK, s>>>(partials.get(), nb, S, Pfull.get(), k,
                                                                        coef.get(), S16);
    };
    switch (K) {
        case 1: run(std::integral_constant<int, 1>{}); break;
        case 2: run(std::integral_constant<int, 2>{}); break;
        case 4: run(std::integral_constant<int, 4>{}); break;
        case 8: run(std::integral_constant<int, 8>{}); break;
        case 16: run(std::integral_constant<int, 16>{}); break;
        default: throw InvalidArgument("warm start supports block widths up to 16");
    }
    const std::int64_t tiles = (n + 15) / 16;
    expand_kernel<<<grid_for(tiles, 4), 256, 0, s>>>(basis.get(), n, S, S16, coef.get(), k, K, x0);
}

}  // namespace cn
