// Reduced-basis warm start for the block PCG.
//
// Snapshots W (n_red x S, S <= 512): solutions the solver itself produced, FP32, row-major (row r
// holds the S values of unknown r). From the Galerkin matrix G = W^T A W (FP64) the caller takes
// the leading eigenpairs and hands back M = V_R Lambda_R^(-1/2); the modes
//   Phi = W M          (n_red x R, R <= 256; an MFMA GEMM with W and M as BF16 hi + lo halves)
// are stored as BF16 hi + lo pairs (16 significant bits): a basis rounded to BF16 alone carries
// high-frequency noise whose residual A e, amplified by the operator's spread, is as large as b
// itself, so the pair is what makes the guess worth starting from. The modes are A-orthonormal up
// to their rounding, which the small projector P = (Phi^T A Phi)^-1 (FP64, close to the identity)
// removes. A block then starts from
//   x0 = Phi c,   c = P Phi^T B   (Galerkin projection: the A-norm best point of span(Phi))
//   or c supplied by the caller  (e.g. predicted by the warm-start network from the pose),
// the expansion being a BF16 MFMA GEMM with FP32 accumulation (hi.hi + hi.lo + lo.hi). PCG
// then iterates from x0 to the same tolerance against ||b||, so only the iteration count changes.
//
// No reference counterpart: the reference's closest analogue is the Galerkin-projected initial
// guess of its conductivity-UQ loop (uq/conductivity/run.py:154-200).
#include <cstring>
#include <vector>

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

// basis[r * S + (s0 + c)] = X[r * K + c] (FP32) for c < k: append a block's solutions as columns.
__global__ void append_kernel(const double* __restrict__ X, std::int64_t n, int k, int K, int S, int s0,
                              float* __restrict__ basis) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n * k) return;
    const std::int64_t r = t / k;
    const int c = static_cast<int>(t % k);
    basis[r * S + s0 + c] = static_cast<float>(X[r * K + c]);
}

// A BF16 hi + lo pair: value = hi + lo.
__device__ __forceinline__ float pair_value(const unsigned short* __restrict__ hi, std::int64_t at, std::int64_t lo_off) {
    return bf16_value(hi[at]) + bf16_value(hi[lo_off + at]);
}

// out (n, 32) FP32 = columns [s0, s0 + 32) of the basis (zero beyond S).
__global__ void gather_chunk_kernel(const float* __restrict__ basis, std::int64_t n, int S, int s0,
                                    float* __restrict__ out) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n * 32) return;
    const std::int64_t r = t >> 5;
    const int c = static_cast<int>(t & 31);
    out[t] = (s0 + c < S) ? basis[r * S + s0 + c] : 0.0f;
}

// The same over the modes' hi + lo pairs (row stride R, lo half at offset n * R).
__global__ void gather_modes_kernel(const unsigned short* __restrict__ modes, std::int64_t n, int R, int j0,
                                    float* __restrict__ out) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n * 32) return;
    const std::int64_t r = t >> 5;
    const int c = static_cast<int>(t & 31);
    out[t] = (j0 + c < R) ? pair_value(modes, r * R + j0 + c, n * R) : 0.0f;
}

typedef double doublex4_t __attribute__((ext_vector_type(4)));

// n ushorts from p (n a compile-time count, p aligned to 2 n bytes up to 16) into v.
template <int N>
__device__ __forceinline__ void load_u16(const unsigned short* __restrict__ p, unsigned short (&v)[N]) {
    if constexpr (N % 8 == 0) {
#pragma unroll
        for (int i = 0; i < N / 8; ++i) {
            const uint4 w = reinterpret_cast<const uint4*>(p)[i];
            __builtin_memcpy(&v[8 * i], &w, 16);
        }
    } else if constexpr (N % 4 == 0) {
#pragma unroll
        for (int i = 0; i < N / 4; ++i) {
            const uint2 w = reinterpret_cast<const uint2*>(p)[i];
            __builtin_memcpy(&v[4 * i], &w, 8);
        }
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) v[i] = p[i];
    }
}

// Partials of Phi^T B on the FP64 matrix cores, T = R16 / 16 tiles of 16 modes: v_mfma_f64_16x16x4f64
// with A = 16 modes x 4 rows of Phi^T and B = the 4 rows of the block (lane l: column l % 16 of
// row r0 + l / 16), accumulated in FP64 per tile. Tile t, tile row i is mode i T + t, so lane l
// reads the T consecutive modes i T .. i T + T - 1 (i = l % 16) of its row in 16-byte loads, and
// each value enters as its exact FP32 hi + lo sum. Accumulator register v of lane l holds tile row
// 4 v + l / 16, column l % 16. The block's waves fold into its partial in wave order:
// partials[(block * R16 + mode) * K + c].
constexpr int kProjectThreads = 512;

// Columns c0 .. c0 + Kh - 1 (Kh <= 16) of a block whose rows are ldb values apart.
template <int T>
__global__ void __launch_bounds__(kProjectThreads) project_kernel(const unsigned short* __restrict__ modes,
                                                                  std::int64_t n, const float* __restrict__ B, int ldb,
                                                                  int c0, int K, std::int64_t rows_per_block,
                                                                  double* __restrict__ partials) {
    constexpr int R16 = 16 * T;
    constexpr int kWaves = kProjectThreads / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, quad = lane >> 4;
    const std::int64_t lo = static_cast<std::int64_t>(blockIdx.x) * rows_per_block;
    const std::int64_t hi = min(lo + rows_per_block, n);
    const std::int64_t half = n * R16;
    doublex4_t acc[T];
#pragma unroll
    for (int t = 0; t < T; ++t) acc[t] = doublex4_t{0.0, 0.0, 0.0, 0.0};
    for (std::int64_t r0 = lo + 4 * wave; r0 < hi; r0 += 4 * kWaves) {
        const std::int64_t row = r0 + quad;
        const bool ok = row < hi;
        const std::int64_t rr = ok ? row : hi - 1;  // clamped: the loads stay in bounds
        const double b = (ok && col < K) ? static_cast<double>(B[rr * ldb + c0 + col]) : 0.0;
        unsigned short mh[T], ml[T];
        load_u16<T>(modes + rr * R16 + col * T, mh);
        load_u16<T>(modes + half + rr * R16 + col * T, ml);
#pragma unroll
        for (int t = 0; t < T; ++t) {
            const double a = static_cast<double>(bf16_value(mh[t]) + bf16_value(ml[t]));
            acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
        }
    }
    __shared__ double red[kMaxWarmModes * 16];
    for (int e = threadIdx.x; e < R16 * 16; e += blockDim.x) red[e] = 0.0;
    __syncthreads();
    for (int w = 0; w < kWaves; ++w) {
        if (wave == w) {
#pragma unroll
            for (int t = 0; t < T; ++t)
#pragma unroll
                for (int v = 0; v < 4; ++v) red[((4 * v + quad) * T + t) * 16 + col] += acc[t][v];
        }
        __syncthreads();
    }
    double* out = partials + static_cast<std::int64_t>(blockIdx.x) * R16 * K;
    for (int e = threadIdx.x; e < R16 * K; e += blockDim.x) out[e] = red[(e / K) * 16 + e % K];
}

// out[e] = sum over the nb block partials of entry e (E entries per block): one wavefront per
// entry, lane j adding blocks j, j + 64, ... in order, then a fixed xor tree.
__global__ void sum_blocks_kernel(const double* __restrict__ partials, int nb, int E, double* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int e = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (e >= E) return;
    double acc = 0.0;
    for (int b = lane; b < nb; b += 64) acc += partials[static_cast<std::int64_t>(b) * E + e];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) out[e] = acc;
}

// c = P (Phi^T B) from the reduced R x K product, then split into BF16 hi/lo halves laid out for
// the expansion: coef[(half * R + s) * 16 + col] (R a multiple of 16, columns >= k zero).
__global__ void coefficients_kernel(const double* __restrict__ ptb, int R, int K, const double* __restrict__ P, int k,
                                    unsigned short* __restrict__ coef) {
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < R * 16; e += gridDim.x * blockDim.x) {
        const int sidx = e / 16, c = e % 16;
        double v = 0.0;
        if (c < k) {
            for (int j = 0; j < R; ++j) v = fma(P[static_cast<std::int64_t>(sidx) * R + j], ptb[j * K + c], v);
        }
        const float vf = static_cast<float>(v);
        const unsigned short hi = bf16_bits(vf);
        const unsigned short lo = bf16_bits(vf - bf16_value(hi));
        coef[e] = hi;
        coef[R * 16 + e] = lo;
    }
}

typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4_t;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4_t;

// x0 (n, K) FP64 = Phi c: one wavefront per 16-row tile, v_mfma_f32_16x16x16bf16_1k over the modes
// in T steps of 16; hi.hi, hi.lo and lo.hi products accumulate into the same FP32 tile.
// Operand layout of the 16x16x16 BF16 MFMA (CDNA3/4 ISA): lane l supplies A[l % 16][4 (l / 16) + j]
// and B[4 (l / 16) + j][l % 16], j = 0..3, and receives D[4 (l / 16) + j][l % 16]. At step t the
// product index 4 q + j (q = l / 16) stands for mode 4 T q + 4 t + j, so lane l reads the 4 T
// consecutive modes 4 T q .. of its row (16-byte loads) before the first product; the
// coefficients sit in LDS transposed, so a lane's B fragment is one 8-byte read.
constexpr int kExpandThreads = 256;

// Writes columns c0 .. c0 + K - 1 (K <= 16) of an x0 whose rows are ldx values apart.
template <int T>
__global__ void __launch_bounds__(kExpandThreads) expand_kernel(const unsigned short* __restrict__ modes,
                                                                std::int64_t n, const unsigned short* __restrict__ coef,
                                                                int k, int K, int ldx, int c0, double* __restrict__ x0) {
    constexpr int R16 = 16 * T;
    __shared__ uint2 ct[2][16][kMaxWarmModes / 4];  // [half][column][mode / 4]: 4 modes' BF16 bits
    {
        unsigned short* flat = reinterpret_cast<unsigned short*>(ct);
        for (int e = threadIdx.x; e < 2 * R16 * 16; e += blockDim.x) {
            const int h = e / (R16 * 16), r = e % (R16 * 16), s = r / 16, c = r % 16;
            flat[(h * 16 + c) * kMaxWarmModes + s] = coef[e];
        }
    }
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const std::int64_t tile = static_cast<std::int64_t>(blockIdx.x) * (kExpandThreads / 64) + (threadIdx.x >> 6);
    const std::int64_t r0 = tile * 16;
    if (r0 >= n) return;
    const int i = lane & 15, q = lane >> 4;
    const std::int64_t row = r0 + i;
    const std::int64_t rr = row < n ? row : n - 1;  // clamped: the loads stay in bounds
    unsigned short ah[4 * T], al[4 * T];
    load_u16<4 * T>(modes + rr * R16 + 4 * T * q, ah);
    load_u16<4 * T>(modes + n * R16 + rr * R16 + 4 * T * q, al);
    floatx4_t acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const uint2 bh2 = ct[0][i][T * q + t], bl2 = ct[1][i][T * q + t];
        bf16x4_t a, alo, bh, bl;
        __builtin_memcpy(&a, &ah[4 * t], 8);
        __builtin_memcpy(&alo, &al[4 * t], 8);
        __builtin_memcpy(&bh, &bh2, 8);
        __builtin_memcpy(&bl, &bl2, 8);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bh, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bl, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(alo, bh, acc, 0, 0, 0);
    }
    // acc[j] = x0[r0 + 4 q + j][i]
    if (i < K) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const std::int64_t r = r0 + 4 * q + j;
            if (r < n) x0[r * ldx + c0 + i] = (i < k) ? static_cast<double>(acc[j]) : 0.0;
        }
    }
}

// Calls f with the tile count T = R16 / 16 as a compile-time constant.
template <typename F>
void dispatch_tiles(int R16, F&& f) {
    switch (R16 / 16) {
        case 1: f(std::integral_constant<int, 1>{}); break;
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 3: f(std::integral_constant<int, 3>{}); break;
        case 4: f(std::integral_constant<int, 4>{}); break;
        case 5: f(std::integral_constant<int, 5>{}); break;
        case 6: f(std::integral_constant<int, 6>{}); break;
        case 7: f(std::integral_constant<int, 7>{}); break;
        case 8: f(std::integral_constant<int, 8>{}); break;
        case 9: f(std::integral_constant<int, 9>{}); break;
        case 10: f(std::integral_constant<int, 10>{}); break;
        case 11: f(std::integral_constant<int, 11>{}); break;
        case 12: f(std::integral_constant<int, 12>{}); break;
        case 13: f(std::integral_constant<int, 13>{}); break;
        case 14: f(std::integral_constant<int, 14>{}); break;
        case 15: f(std::integral_constant<int, 15>{}); break;
        case 16: f(std::integral_constant<int, 16>{}); break;
        default: throw InvalidArgument("mode width must be a multiple of 16 in [16, 256]");
    }
}

// Modes (n, R) = W (n, S) M (S, R): one wavefront per 16 x 16 output tile, S stepped by 16, W split
// into BF16 hi/lo on load, M carried as BF16 hi/lo halves mhl[(half * S16 + s) * R + j] (zero rows
// past S); the FP32 result is stored as a BF16 hi + lo pair (lo half at offset n * R).
__global__ void __launch_bounds__(256) modes_kernel(const float* __restrict__ basis, std::int64_t n, int S,
                                                   int S16, const unsigned short* __restrict__ mhl, int R,
                                                   unsigned short* __restrict__ modes) {
    const int lane = threadIdx.x & 63;
    const std::int64_t tile = static_cast<std::int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
    const int tiles_r = R / 16;
    const std::int64_t r0 = (tile / tiles_r) * 16;
    const int j0 = static_cast<int>(tile % tiles_r) * 16;
    if (r0 >= n) return;
    const int i = lane & 15, q = lane >> 4;
    const std::int64_t row = r0 + i;
    floatx4_t acc = {0.f, 0.f, 0.f, 0.f};
    for (int s0 = 0; s0 < S16; s0 += 16) {
        bf16x4_t a, alo, bh, bl;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int sidx = s0 + 4 * q + j;
            const float w = (row < n && sidx < S) ? basis[row * S + sidx] : 0.0f;
            const unsigned short wh = bf16_bits(w);
            a[j] = static_cast<short>(wh);
            alo[j] = static_cast<short>(bf16_bits(w - bf16_value(wh)));
            bh[j] = static_cast<short>(mhl[static_cast<std::int64_t>(sidx) * R + j0 + i]);
            bl[j] = static_cast<short>(mhl[static_cast<std::int64_t>(S16 + sidx) * R + j0 + i]);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bh, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bl, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(alo, bh, acc, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const std::int64_t r = r0 + 4 * q + j;
        if (r < n) {
            const unsigned short hi = bf16_bits(acc[j]);
            modes[r * R + j0 + i] = hi;
            modes[n * R + r * R + j0 + i] = bf16_bits(acc[j] - bf16_value(hi));
        }
    }
}

}  // namespace

void WarmBasis::reset(std::int64_t n_rows, int capacity) {
    require(capacity >= 1 && capacity <= kMaxWarmBasis, "warm-start capacity must be in [1, 2048]");
    n = n_rows;
    cap = capacity;
    count = 0;
    ready = false;
    n_modes = used_modes = 0;
    modes.reset(0);
    basis.reset(static_cast<std::size_t>(n) * cap, "warm basis");
    check(hipMemset(basis.get(), 0, sizeof(float) * n * cap), "warm basis clear");
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

// Host BF16 rounding, the same rule as the device's (round to nearest even).
static unsigned short host_bf16(float x) {
    unsigned u;
    std::memcpy(&u, &x, sizeof u);
    return static_cast<unsigned short>((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

static float host_bf16_value(unsigned short b) {
    const unsigned u = static_cast<unsigned>(b) << 16;
    float x;
    std::memcpy(&x, &u, sizeof x);
    return x;
}

void WarmBasis::set_modes(const double* M_host, int S, int R, hipStream_t s) {
    require(S == count && S >= 1, "the mode matrix must have one row per basis column");
    require(R >= 1 && R <= kMaxWarmModes, "mode count must be in [1, 256]");
    const int S16 = (S + 15) / 16 * 16;
    const int R16 = (R + 15) / 16 * 16;
    std::vector<unsigned short> mhl(static_cast<std::size_t>(2) * S16 * R16, 0);
    for (int a = 0; a < S; ++a)
        for (int j = 0; j < R; ++j) {
            const float v = static_cast<float>(M_host[static_cast<std::size_t>(a) * R + j]);
            const unsigned short hi = host_bf16(v);
            mhl[static_cast<std::size_t>(a) * R16 + j] = hi;
            mhl[static_cast<std::size_t>(S16 + a) * R16 + j] = host_bf16(v - host_bf16_value(hi));
        }
    DeviceBuffer<unsigned short> dm(mhl.size(), "mode matrix");
    h2d(dm.get(), mhl.data(), mhl.size(), s);
    n_modes = R16;
    modes.reset(static_cast<std::size_t>(2) * n * R16, "warm modes");
    const std::int64_t tiles = (n + 15) / 16 * (R16 / 16);
    modes_kernel<<<static_cast<unsigned>((tiles + 3) / 4), 256, 0, s>>>(basis.get(), n, cap, S16, dm.get(), R16,
                                                                       modes.get());
    used_modes = R;
    // P = I until the caller supplies the BF16-rounding correction.
    std::vector<double> eye(static_cast<std::size_t>(R16) * R16, 0.0);
    for (int j = 0; j < R; ++j) eye[static_cast<std::size_t>(j) * R16 + j] = 1.0;
    P.reset(eye.size(), "warm projector");
    h2d(P.get(), eye.data(), eye.size(), s);
    sync(s, "warm modes");
    ready = true;
}

void WarmBasis::set_projector(const double* P_host, int R, hipStream_t s) {
    require(ready && R == used_modes, "the projector must be R x R for the current modes");
    std::vector<double> pad(static_cast<std::size_t>(n_modes) * n_modes, 0.0);
    for (int a = 0; a < R; ++a)
        for (int b = 0; b < R; ++b) pad[static_cast<std::size_t>(a) * n_modes + b] = P_host[static_cast<std::size_t>(a) * R + b];
    h2d(P.get(), pad.data(), pad.size(), s);
    sync(s, "warm projector");
}

void WarmBasis::mode_chunk(int j0, float* out, hipStream_t s) const {
    if (const unsigned g = grid_for(n * 32)) gather_modes_kernel<<<g, kBlock, 0, s>>>(modes.get(), n, n_modes, j0, out);
}

void WarmBasis::expand(int k, int K, double* x0, hipStream_t s, int ldx, int c0) {
    require(n_modes % 16 == 0 && n_modes <= kMaxWarmModes, "mode width must be a multiple of 16, at most 256");
    require(K >= 1 && K <= 16, "the expansion writes at most 16 columns at a time");
    const std::int64_t tiles = (n + 15) / 16;
    dispatch_tiles(n_modes, [&](auto tc) {
        expand_kernel<decltype(tc)::value><<<grid_for(tiles, kExpandThreads / 64), kExpandThreads, 0, s>>>(
            modes.get(), n, coef.get(), k, K, ldx > 0 ? ldx : K, c0, x0);
    });
}

void WarmBasis::project(const float* B, int K, hipStream_t s, int ldb, int c0) {
    require(ready, "warm start requested but no modes are set");
    require(K >= 1 && K <= 16, "the projection reads at most 16 columns at a time");
    require(n_modes % 16 == 0 && n_modes <= kMaxWarmModes, "mode width must be a multiple of 16, at most 256");
    const int R = n_modes;
    const int nb = kProjectBlocks;
    const std::int64_t rpb = (n + nb - 1) / nb;
    partials.ensure(static_cast<std::size_t>(nb) * R * K + static_cast<std::size_t>(R) * K, "warm partials");
    dispatch_tiles(R, [&](auto tc) {
        project_kernel<decltype(tc)::value><<<nb, kProjectThreads, 0, s>>>(modes.get(), n, B, ldb > 0 ? ldb : K, c0,
                                                                           K, rpb, partials.get());
    });
}

void WarmBasis::coefficients(int k, int K, hipStream_t s) {
    require(k >= 0 && k <= K, "k must be in [0, K]");
    const int R = n_modes;
    const int nb = kProjectBlocks;
    double* const ptb = partials.get() + static_cast<std::size_t>(nb) * R * K;
    coef.ensure(static_cast<std::size_t>(2) * R * 16, "warm coefficients");
    sum_blocks_kernel<<<(R * K + 3) / 4, 256, 0, s>>>(partials.get(), nb, R * K, ptb);
    coefficients_kernel<<<(R * 16 + 255) / 256, 256, 0, s>>>(ptb, R, K, P.get(), k, coef.get());
}

void WarmBasis::project_host(const float* B, int k, int K, double* out, hipStream_t s) {
    require(K >= 1 && K <= 32 && k >= 1 && k <= K, "k must be in [1, K], K at most 32");
    const int R = n_modes;
    const int nb = kProjectBlocks;
    std::vector<double> h;
    for (int c0 = 0; c0 < k; c0 += 16) {
        const int Kh = std::min(16, K - c0);
        project(B, Kh, s, K, c0);
        double* const ptb = partials.get() + static_cast<std::size_t>(nb) * R * Kh;
        sum_blocks_kernel<<<(R * Kh + 3) / 4, 256, 0, s>>>(partials.get(), nb, R * Kh, ptb);
        h.resize(static_cast<std::size_t>(R) * Kh);
        d2h(h.data(), ptb, h.size(), s);
        sync(s, "projection readback");
        for (int c = 0; c < Kh && c0 + c < k; ++c)
            for (int j = 0; j < used_modes; ++j)
                out[static_cast<std::size_t>(c0 + c) * used_modes + j] = h[static_cast<std::size_t>(j) * Kh + c];
    }
}

void WarmBasis::initial_guess(const float* B, int k, int K, double* x0, hipStream_t s) {
    require(K >= 1 && K <= 32, "warm start supports block widths up to 32");
    // Blocks wider than 16 go through in halves of 16 columns (the MFMA tiles are 16 wide).
    for (int c0 = 0; c0 < K; c0 += 16) {
        const int Kh = std::min(16, K - c0), kh = std::max(0, std::min(k - c0, Kh));
        project(B, Kh, s, K, c0);
        coefficients(kh, Kh, s);
        expand(kh, Kh, x0, s, K, c0);
    }
}

void WarmBasis::bench(const float* B, int k, int K, double* x0, int reps, double* ms3, hipStream_t s) {
    require(reps >= 1, "reps must be positive");
    require(K >= 1 && K <= 16, "the stage timing covers blocks of up to 16 columns");
    initial_guess(B, k, K, x0, s);  // untimed: buffers and caches
    hipEvent_t ev[4];
    for (auto& e : ev) check(hipEventCreate(&e), "event");
    double acc[3] = {0.0, 0.0, 0.0};
    for (int r = 0; r < reps; ++r) {
        check(hipEventRecord(ev[0], s), "event record");
        project(B, K, s);
        check(hipEventRecord(ev[1], s), "event record");
        coefficients(k, K, s);
        check(hipEventRecord(ev[2], s), "event record");
        expand(k, K, x0, s);
        check(hipEventRecord(ev[3], s), "event record");
        check(hipEventSynchronize(ev[3]), "event sync");
        for (int i = 0; i < 3; ++i) {
            float ms = 0.f;
            check(hipEventElapsedTime(&ms, ev[i], ev[i + 1]), "elapsed");
            acc[i] += ms;
        }
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    for (int i = 0; i < 3; ++i) ms3[i] = acc[i] / reps;
}

void WarmBasis::guess_from_coefficients(const double* c_host, int k, int K, double* x0, hipStream_t s) {
    require(ready, "warm start requested but no modes are set");
    require(k <= K && K <= 32, "supplied-coefficient warm start supports block widths up to 32");
    const int R = n_modes;
    for (int c0 = 0; c0 < K; c0 += 16) {
        const int Kh = std::min(16, K - c0), kh = std::max(0, std::min(k - c0, Kh));
        std::vector<unsigned short> hl(static_cast<std::size_t>(2) * R * 16, 0);
        for (int c = 0; c < kh; ++c)
            for (int j = 0; j < used_modes; ++j) {
                const float v = static_cast<float>(c_host[static_cast<std::size_t>(c0 + c) * used_modes + j]);
                const unsigned short hi = host_bf16(v);
                hl[static_cast<std::size_t>(j) * 16 + c] = hi;
                hl[static_cast<std::size_t>(R + j) * 16 + c] = host_bf16(v - host_bf16_value(hi));
            }
        coef.ensure(hl.size(), "warm coefficients");
        stage.put(coef.get(), hl.data(), hl.size(), s);  // waits for the previous half's copy
        expand(kh, Kh, x0, s, K, c0);
    }
}

}  // namespace cn
