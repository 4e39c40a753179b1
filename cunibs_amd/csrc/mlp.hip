// The learned part of the warm start: a small MLP (inputs -> hidden -> hidden -> outputs, GELU)
// trained on the GPU with the library's own kernels, mapping a placement's pose (and tissue
// conductivities) to coefficients in the reduced basis of solver-generated solutions.
//
// Training runs in FP32: tiled GEMMs on the vector ALUs for the forward pass, the backward pass and
// the weight gradients, an exact GELU (erf form) and Adam on FP32 master weights. Inference uses
// BF16 copies of the weights: each layer is a v_mfma_f32_16x16x16bf16_1k GEMM with FP32
// accumulation, the activations rounded to BF16 between layers.
//
// No reference counterpart (cuNIBS has no learned component).
#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

#include "mlp.hpp"

namespace cn {

namespace {

constexpr int kTile = 64;   // output tile of the FP32 GEMM (64 x 64 per 256-thread block)
constexpr int kTileK = 16;  // K chunk staged in LDS

// C[i, j] = (accumulate ? C[i, j] : 0) + sum_k A(i, k) B(k, j) (+ bias[j]), FP32. A(i, k) is
// A[i * sa_i + k * sa_k] and B(k, j) is B[k * sb_k + j * sb_j], so transposed operands are strides.
// Each thread owns a 4 x 4 block of the tile; k runs in ascending order for every output.
__global__ void __launch_bounds__(256) gemm_f32_kernel(int M, int N, int K, const float* __restrict__ A,
                                                       long sa_i, long sa_k, const float* __restrict__ B, long sb_k,
                                                       long sb_j, float* __restrict__ C, int ldc,
                                                       const float* __restrict__ bias, int accumulate) {
    __shared__ float As[kTileK][kTile + 1];
    __shared__ float Bs[kTileK][kTile + 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int i0 = blockIdx.y * kTile, j0 = blockIdx.x * kTile;
    float acc[4][4] = {};
    for (int k0 = 0; k0 < K; k0 += kTileK) {
        for (int e = threadIdx.x; e < kTileK * kTile; e += 256) {
            const int kk = e / kTile, r = e % kTile;
            const int i = i0 + r, j = j0 + r, k = k0 + kk;
            As[kk][r] = (i < M && k < K) ? A[i * sa_i + k * sa_k] : 0.0f;
            Bs[kk][r] = (j < N && k < K) ? B[k * sb_k + j * sb_j] : 0.0f;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < kTileK; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                a[u] = As[kk][ty * 4 + u];
                b[u] = Bs[kk][tx * 4 + u];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[u][v] = fmaf(a[u], b[v], acc[u][v]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int i = i0 + ty * 4 + u;
        if (i >= M) continue;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int j = j0 + tx * 4 + v;
            if (j >= N) continue;
            float r = acc[u][v];
            if (bias) r += bias[j];
            if (accumulate) r += C[static_cast<long>(i) * ldc + j];
            C[static_cast<long>(i) * ldc + j] = r;
        }
    }
}

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_grad(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

__global__ void gelu_kernel(long count, const float* __restrict__ z, float* __restrict__ h) {
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) h[i] = gelu(z[i]);
}

// dz = dh * gelu'(z)
__global__ void gelu_back_kernel(long count, const float* __restrict__ z, const float* __restrict__ dh,
                                 float* __restrict__ dz) {
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) dz[i] = dh[i] * gelu_grad(z[i]);
}

// Mean squared error over the batch: dy = 2 (y - t) / (B * N); per-row partial losses in FP64.
__global__ void mse_kernel(int B, int N, const float* __restrict__ y, const float* __restrict__ t,
                           float* __restrict__ dy, double* __restrict__ row_loss) {
    const int r = blockIdx.x;
    double acc = 0.0;
    for (int j = threadIdx.x; j < N; j += blockDim.x) {
        const float d = y[static_cast<long>(r) * N + j] - t[static_cast<long>(r) * N + j];
        dy[static_cast<long>(r) * N + j] = 2.0f * d / static_cast<float>(B * N);
        acc += static_cast<double>(d) * d;
    }
    __shared__ double s[256];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) row_loss[r] = s[0];
}

// Bias gradient: db[j] = sum over rows of dz[r, j], rows in order.
__global__ void bias_grad_kernel(int B, int N, const float* __restrict__ dz, float* __restrict__ db) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    float acc = 0.0f;
    for (int r = 0; r < B; ++r) acc += dz[static_cast<long>(r) * N + j];
    db[j] = acc;
}

__device__ __forceinline__ unsigned short to_bf16(float x) {
    const unsigned u = __float_as_uint(x);
    return static_cast<unsigned short>((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// Adam on the FP32 master weights, then the BF16 copy used by inference.
__global__ void adam_kernel(long count, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2,
                            unsigned short* __restrict__ w16) {
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    w16[i] = to_bf16(wi);
}

__global__ void to_bf16_kernel(long count, const float* __restrict__ x, unsigned short* __restrict__ y) {
    const long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) y[i] = to_bf16(x[i]);
}

typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4_t;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4_t;

// Y (M, N) FP32 = act(X (M, K) BF16 . W (K, N) BF16 + bias), one wavefront per 16 x 16 tile,
// K stepped by 16 (M, N, K multiples of 16). Lane l supplies A[l % 16][4 (l / 16) + j],
// B[4 (l / 16) + j][l % 16] and receives D[4 (l / 16) + j][l % 16]. With gelu set, the output is
// GELU-activated and also written as BF16 for the next layer.
__global__ void __launch_bounds__(256) mfma_layer_kernel(int M, int N, int K, const unsigned short* __restrict__ X,
                                                         const unsigned short* __restrict__ W,
                                                         const float* __restrict__ bias, int gelu_act,
                                                         float* __restrict__ Y, unsigned short* __restrict__ Y16) {
    const int lane = threadIdx.x & 63;
    const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int tiles_n = N / 16;
    if (tile >= (M / 16) * tiles_n) return;
    const int i0 = (tile / tiles_n) * 16, j0 = (tile % tiles_n) * 16;
    const int i = lane & 15, q = lane >> 4;
    floatx4_t acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4_t a, b;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            a[j] = static_cast<short>(X[static_cast<long>(i0 + i) * K + k0 + 4 * q + j]);
            b[j] = static_cast<short>(W[static_cast<long>(k0 + 4 * q + j) * N + j0 + i]);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const long r = i0 + 4 * q + j;
        const int c = j0 + i;
        float v = acc[j] + bias[c];
        if (gelu_act) v = gelu(v);
        Y[r * N + c] = v;
        if (Y16) Y16[r * N + c] = to_bf16(v);
    }
}

void gemm(int M, int N, int K, const float* A, long sa_i, long sa_k, const float* B, long sb_k, long sb_j, float* C,
          int ldc, const float* bias, bool accumulate, hipStream_t s) {
    const dim3 grid((N + kTile - 1) / kTile, (M + kTile - 1) / kTile);
    gemm_f32_kernel<<<grid, 256, 0, s>>>(M, N, K, A, sa_i, sa_k, B, sb_k, sb_j, C, ldc, bias, accumulate ? 1 : 0);
}

}  // namespace

Mlp::Mlp(int n_in, int n_hidden, int n_out, unsigned seed, hipStream_t s)
    : n_in_(n_in), n_hidden_(n_hidden), n_out_(n_out), s_(s) {
    require(n_in >= 1 && n_in <= 16, "the MLP takes 1 to 16 inputs");
    require(n_hidden >= 16 && n_hidden % 16 == 0 && n_hidden <= 4096, "hidden width must be a multiple of 16");
    require(n_out >= 16 && n_out % 16 == 0 && n_out <= 4096, "output width must be a multiple of 16");
    const int shapes[3][2] = {{kInPad, n_hidden}, {n_hidden, n_hidden}, {n_hidden, n_out}};
    std::mt19937 rng(seed);
    std::vector<float> host;
    for (int l = 0; l < 3; ++l) {
        const int fan_in = l == 0 ? n_in : shapes[l][0];
        offset_[l] = static_cast<long>(host.size());
        std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / static_cast<float>(fan_in)));
        for (int r = 0; r < shapes[l][0]; ++r)
            for (int c = 0; c < shapes[l][1]; ++c) host.push_back((l == 0 && r >= n_in) ? 0.0f : nd(rng));
        boffset_[l] = static_cast<long>(host.size());
        for (int c = 0; c < shapes[l][1]; ++c) host.push_back(0.0f);
    }
    n_params_ = static_cast<long>(host.size());
    w_.reset(n_params_, "mlp weights");
    g_.reset(n_params_, "mlp grads");
    m_.reset(n_params_, "adam m");
    v_.reset(n_params_, "adam v");
    w16_.reset(n_params_, "mlp bf16 weights");
    set_params(host.data());
    std::vector<float> zeros(n_params_, 0.0f);
    h2d(m_.get(), zeros.data(), n_params_, s_);
    h2d(v_.get(), zeros.data(), n_params_, s_);
    sync(s_, "mlp init");
}

void Mlp::set_params(const float* host) {
    h2d(w_.get(), host, n_params_, s_);
    to_bf16_kernel<<<grid_for(n_params_), kBlock, 0, s_>>>(n_params_, w_.get(), w16_.get());
    sync(s_, "mlp params");
}

void Mlp::get_params(float* host) const {
    d2h(host, w_.get(), n_params_, s_);
    sync(s_, "mlp params");
}

void Mlp::ensure_batch(int B) {
    if (B <= cap_batch_) return;
    const long hb = static_cast<long>(B) * n_hidden_;
    x_.reset(static_cast<long>(B) * kInPad, "mlp x");
    z1_.reset(hb, "z1");
    h1_.reset(hb, "h1");
    z2_.reset(hb, "z2");
    h2_.reset(hb, "h2");
    y_.reset(static_cast<long>(B) * n_out_, "y");
    t_.reset(static_cast<long>(B) * n_out_, "t");
    dy_.reset(static_cast<long>(B) * n_out_, "dy");
    dh_.reset(hb, "dh");
    dz_.reset(hb, "dz");
    loss_.reset(B, "row loss");
    x16_.reset(static_cast<long>(B) * kInPad, "x16");
    h16a_.reset(hb, "h16a");
    h16b_.reset(hb, "h16b");
    cap_batch_ = B;
}

// X (B, n_in) host-padded to (B, kInPad) on upload.
void Mlp::upload_inputs(const float* X, int B) {
    const std::size_t count = static_cast<std::size_t>(B) * kInPad;
    float* pad = static_cast<float*>(stage_.fill(count * sizeof(float)));
    std::fill(pad, pad + count, 0.0f);
    for (int r = 0; r < B; ++r)
        for (int c = 0; c < n_in_; ++c) pad[static_cast<std::size_t>(r) * kInPad + c] = X[static_cast<std::size_t>(r) * n_in_ + c];
    stage_.send(x_.get(), 0, count * sizeof(float), s_);
    stage_.done(s_);
}

double Mlp::forward_backward(int B) {
    const int H = n_hidden_, O = n_out_;
    const float* W1 = w_.get() + offset_[0];
    const float* b1 = w_.get() + boffset_[0];
    const float* W2 = w_.get() + offset_[1];
    const float* b2 = w_.get() + boffset_[1];
    const float* W3 = w_.get() + offset_[2];
    const float* b3 = w_.get() + boffset_[2];
    const long hb = static_cast<long>(B) * H;
    // forward
    gemm(B, H, kInPad, x_.get(), kInPad, 1, W1, H, 1, z1_.get(), H, b1, false, s_);
    gelu_kernel<<<grid_for(hb), kBlock, 0, s_>>>(hb, z1_.get(), h1_.get());
    gemm(B, H, H, h1_.get(), H, 1, W2, H, 1, z2_.get(), H, b2, false, s_);
    gelu_kernel<<<grid_for(hb), kBlock, 0, s_>>>(hb, z2_.get(), h2_.get());
    gemm(B, O, H, h2_.get(), H, 1, W3, O, 1, y_.get(), O, b3, false, s_);
    mse_kernel<<<B, 256, 0, s_>>>(B, O, y_.get(), t_.get(), dy_.get(), loss_.get());
    // backward: dW3 = h2^T dy, db3, dh2 = dy W3^T, dz2, dW2 = h1^T dz2, ...
    float* G = g_.get();
    gemm(H, O, B, h2_.get(), 1, H, dy_.get(), O, 1, G + offset_[2], O, nullptr, false, s_);
    bias_grad_kernel<<<grid_for(O), kBlock, 0, s_>>>(B, O, dy_.get(), G + boffset_[2]);
    gemm(B, H, O, dy_.get(), O, 1, W3, 1, O, dh_.get(), H, nullptr, false, s_);
    gelu_back_kernel<<<grid_for(hb), kBlock, 0, s_>>>(hb, z2_.get(), dh_.get(), dz_.get());
    gemm(H, H, B, h1_.get(), 1, H, dz_.get(), H, 1, G + offset_[1], H, nullptr, false, s_);
    bias_grad_kernel<<<grid_for(H), kBlock, 0, s_>>>(B, H, dz_.get(), G + boffset_[1]);
    gemm(B, H, H, dz_.get(), H, 1, W2, 1, H, dh_.get(), H, nullptr, false, s_);
    gelu_back_kernel<<<grid_for(hb), kBlock, 0, s_>>>(hb, z1_.get(), dh_.get(), dz_.get());
    gemm(kInPad, H, B, x_.get(), 1, kInPad, dz_.get(), H, 1, G + offset_[0], H, nullptr, false, s_);
    bias_grad_kernel<<<grid_for(H), kBlock, 0, s_>>>(B, H, dz_.get(), G + boffset_[0]);
    std::vector<double> rl(B);
    d2h(rl.data(), loss_.get(), B, s_);
    sync(s_, "mlp step");
    double loss = 0.0;
    for (double v : rl) loss += v;
    return loss / (static_cast<double>(B) * O);
}

double Mlp::gradients(const float* X, const float* T, int B, float* grads_host) {
    ensure_batch(B);
    upload_inputs(X, B);
    h2d(t_.get(), T, static_cast<std::size_t>(B) * n_out_, s_);
    const double loss = forward_backward(B);
    d2h(grads_host, g_.get(), n_params_, s_);
    sync(s_, "mlp grads");
    return loss;
}

double Mlp::train(const float* X, const float* T, long N, int epochs, int batch, float lr, unsigned seed) {
    require(N >= 1 && batch >= 1 && epochs >= 1, "bad training arguments");
    const int B = static_cast<int>(std::min<long>(batch, N));
    ensure_batch(B);
    std::mt19937 rng(seed);
    std::vector<long> order(N);
    for (long i = 0; i < N; ++i) order[i] = i;
    std::vector<float> xb(static_cast<std::size_t>(B) * n_in_), tb(static_cast<std::size_t>(B) * n_out_);
    double last = 0.0;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int ep = 0; ep < epochs; ++ep) {
        std::shuffle(order.begin(), order.end(), rng);
        double sum = 0.0;
        int steps = 0;
        for (long at = 0; at + B <= N; at += B) {
            for (int r = 0; r < B; ++r) {
                const long src = order[at + r];
                std::copy(X + src * n_in_, X + (src + 1) * n_in_, xb.begin() + static_cast<long>(r) * n_in_);
                std::copy(T + src * n_out_, T + (src + 1) * n_out_, tb.begin() + static_cast<long>(r) * n_out_);
            }
            upload_inputs(xb.data(), B);
            h2d(t_.get(), tb.data(), tb.size(), s_);
            sum += forward_backward(B);
            ++steps_;
            ++steps;
            const float c1 = 1.0f - std::pow(b1, static_cast<float>(steps_));
            const float c2 = 1.0f - std::pow(b2, static_cast<float>(steps_));
            adam_kernel<<<grid_for(n_params_), kBlock, 0, s_>>>(n_params_, w_.get(), g_.get(), m_.get(), v_.get(), lr,
                                                                 b1, b2, eps, c1, c2, w16_.get());
        }
        last = steps ? sum / steps : 0.0;
    }
    sync(s_, "mlp train");
    return last;
}

void Mlp::predict(const float* X, int B, float* Y_host) {
    const int Bp = (B + 15) / 16 * 16;
    ensure_batch(Bp);
    std::vector<float> pad(static_cast<std::size_t>(Bp) * n_in_, 0.0f);
    std::copy(X, X + static_cast<std::size_t>(B) * n_in_, pad.begin());
    upload_inputs(pad.data(), Bp);
    const long xin = static_cast<long>(Bp) * kInPad;
    to_bf16_kernel<<<grid_for(xin), kBlock, 0, s_>>>(xin, x_.get(), x16_.get());
    const int H = n_hidden_, O = n_out_;
    auto layer = [&](const unsigned short* in, int K, int N, int l, bool act, float* out, unsigned short* out16) {
        const int tiles = (Bp / 16) * (N / 16);
        mfma_layer_kernel<<<(tiles + 3) / 4, 256, 0, s_>>>(Bp, N, K, in, w16_.get() + offset_[l],
                                                           w_.get() + boffset_[l], act ? 1 : 0, out, out16);
    };
    layer(x16_.get(), kInPad, H, 0, true, h1_.get(), h16a_.get());
    layer(h16a_.get(), H, H, 1, true, h2_.get(), h16b_.get());
    layer(h16b_.get(), H, O, 2, false, y_.get(), nullptr);
    d2h(Y_host, y_.get(), static_cast<std::size_t>(B) * O, s_);
    sync(s_, "mlp predict");
}

}  // namespace cn
