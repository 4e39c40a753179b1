// Conductivity uncertainty (the reference's uq/conductivity: per-tissue decomposition
// K(sigma) = sum_t sigma_t K_t, b(sigma) = sum_t sigma_t b_t, assembly.py:37-146; moments of |E|,
// solver/fem/moments.cu:5-15). K conductivity samples share the reduced pattern and go through the
// block solver as K independent CG chains under the nominal preconditioner.
#include "uq.hpp"

namespace cn {

namespace {

__global__ void tissue_scale_kernel(const double* __restrict__ vols, const std::int32_t* __restrict__ tags,
                                    std::int64_t m, int tag, double* __restrict__ scale) {
    const std::int64_t e = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < m) scale[e] = tags[e] == tag ? vols[e] : 0.0;
}

__global__ void tissue_neg_vc_kernel(const float* __restrict__ vols, const std::int32_t* __restrict__ tags,
                                     std::int64_t m, int tag, float* __restrict__ out) {
    const std::int64_t e = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < m) out[e] = tags[e] == tag ? -vols[e] : 0.0f;
}

__global__ void split_offdiag_kernel(const std::int32_t* __restrict__ rp, const std::int32_t* __restrict__ ci,
                                     const double* __restrict__ v, std::int64_t n, double* __restrict__ off,
                                     double* __restrict__ rowsum) {
    const std::int64_t r = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= n) return;
    std::int64_t at = rp[r] - r;  // every reduced row holds its diagonal
    double s = 0.0;
    for (int p = rp[r]; p < rp[r + 1]; ++p) {
        s += v[p];
        if (ci[p] != r) off[at++] = v[p];
    }
    rowsum[r] = s;
}

template <typename T_IN>
__global__ void combine_kernel(const T_IN* __restrict__ in, std::int64_t count, int T, const double* __restrict__ sigma,
                               int k, int K, float* __restrict__ out) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= count * K) return;
    const std::int64_t i = t / K;
    const int c = static_cast<int>(t % K);
    double acc = 0.0;
    if (c < k) {
        for (int q = 0; q < T; ++q) acc = fma(sigma[c * T + q], static_cast<double>(in[q * count + i]), acc);
    }
    out[t] = static_cast<float>(acc);
}

__global__ void moments_kernel(const float* __restrict__ x, std::int64_t m, int k, int K, double* __restrict__ sum,
                               double* __restrict__ sumsq) {
    const std::int64_t e = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= m) return;
    double s = sum[e], q = sumsq[e];
    for (int c = 0; c < k; ++c) {
        const double v = static_cast<double>(x[e * K + c]);
        s += v;
        q = fma(v, v, q);
    }
    sum[e] = s;
    sumsq[e] = q;
}

__global__ void spmm64_kernel(const std::int32_t* __restrict__ rp, const std::int32_t* __restrict__ ci,
                              const double* __restrict__ v, std::int64_t n, const double* __restrict__ X, int k, int K,
                              double* __restrict__ Y) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n * K) return;
    const std::int64_t i = t / K;
    const int c = static_cast<int>(t % K);
    double acc = 0.0;
    if (c < k)
        for (int j = rp[i]; j < rp[i + 1]; ++j) acc = fma(v[j], X[static_cast<std::int64_t>(ci[j]) * K + c], acc);
    Y[t] = acc;
}

constexpr int kGram64Blocks = 1024;

// partials[b][a * k + c] = sum over block b's rows of A[i, a] B[i, c], rows in order.
__global__ void gram64_kernel(const double* __restrict__ A, const double* __restrict__ B, std::int64_t n, int k, int K,
                              double* __restrict__ partials) {
    const std::int64_t rows = (n + gridDim.x - 1) / gridDim.x;
    const std::int64_t lo = static_cast<std::int64_t>(blockIdx.x) * rows;
    const std::int64_t hi = min(lo + rows, n);
    for (int p = threadIdx.x; p < k * k; p += blockDim.x) {
        const int a = p / k, c = p % k;
        double acc = 0.0;
        for (std::int64_t i = lo; i < hi; ++i) acc = fma(A[i * K + a], B[i * K + c], acc);
        partials[static_cast<std::int64_t>(blockIdx.x) * k * k + p] = acc;
    }
}

__global__ void gram64_sum_kernel(const double* __restrict__ partials, int blocks, int kk, double* __restrict__ out) {
    for (int p = threadIdx.x; p < kk; p += blockDim.x) {
        double acc = 0.0;
        for (int b = 0; b < blocks; ++b) acc += partials[static_cast<std::int64_t>(b) * kk + p];
        out[p] = acc;
    }
}

__global__ void sensitivity_rhs_kernel(const std::int32_t* __restrict__ rp, const std::int32_t* __restrict__ ci,
                                       const double* __restrict__ vals, std::int64_t nnz,
                                       const float* __restrict__ b_tissue, const double* __restrict__ x_nom,
                                       std::int64_t n, int T, int K, double* __restrict__ R, float* __restrict__ Bf) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n * K) return;
    const std::int64_t i = t / K;
    const int q = static_cast<int>(t % K);
    double r = 0.0;
    if (q < T) {
        const double* v = vals + static_cast<std::int64_t>(q) * nnz;
        double ax = 0.0;
        for (int j = rp[i]; j < rp[i + 1]; ++j) ax = fma(v[j], x_nom[ci[j]], ax);
        r = static_cast<double>(b_tissue[static_cast<std::int64_t>(q) * n + i]) - ax;
    }
    R[t] = r;
    Bf[t] = static_cast<float>(r);
}

__global__ void expand_guess_kernel(const double* __restrict__ x_nom, const double* __restrict__ S, std::int64_t n,
                                    int T, int KS, const double* __restrict__ coef, int k, int K,
                                    double* __restrict__ X0) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n * K) return;
    const std::int64_t i = t / K;
    const int c = static_cast<int>(t % K);
    double x = 0.0;
    if (c < k) {
        x = x_nom[i];
        for (int q = 0; q < T; ++q) x = fma(S[i * KS + q], coef[q * K + c], x);
    }
    X0[t] = x;
}

}  // namespace

void spmm64(const std::int32_t* rp, const std::int32_t* ci, const double* v, std::int64_t n, const double* X, int k,
            int K, double* Y, hipStream_t s) {
    if (const unsigned g = grid_for(n * K)) spmm64_kernel<<<g, kBlock, 0, s>>>(rp, ci, v, n, X, k, K, Y);
}

void gram64(const double* A, const double* B, std::int64_t n, int k, int K, double* out_host, hipStream_t s) {
    require(k >= 1 && k <= K && K <= 32, "gram width must be in [1, K], K <= 32");
    DeviceBuffer<double> partials(static_cast<std::size_t>(kGram64Blocks) * k * k, "gram64 partials");
    DeviceBuffer<double> out(static_cast<std::size_t>(k) * k, "gram64");
    gram64_kernel<<<kGram64Blocks, kBlock, 0, s>>>(A, B, n, k, K, partials.get());
    gram64_sum_kernel<<<1, kBlock, 0, s>>>(partials.get(), kGram64Blocks, k * k, out.get());
    d2h(out_host, out.get(), static_cast<std::size_t>(k) * k, s);
    sync(s, "gram64");
}

void sensitivity_rhs(const std::int32_t* rp, const std::int32_t* ci, const double* vals, std::int64_t nnz,
                     const float* b_tissue, const double* x_nom, std::int64_t n, int T, int K, double* R, float* Bf,
                     hipStream_t s) {
    if (const unsigned g = grid_for(n * K))
        sensitivity_rhs_kernel<<<g, kBlock, 0, s>>>(rp, ci, vals, nnz, b_tissue, x_nom, n, T, K, R, Bf);
}

void expand_guess(const double* x_nom, const double* S, std::int64_t n, int T, int KS, const double* coef, int k, int K,
                  double* X0, hipStream_t s) {
    if (const unsigned g = grid_for(n * K)) expand_guess_kernel<<<g, kBlock, 0, s>>>(x_nom, S, n, T, KS, coef, k, K, X0);
}

void tissue_scale(const double* vols, const std::int32_t* tags, std::int64_t m, int tag, double* scale, hipStream_t s) {
    if (const unsigned g = grid_for(m)) tissue_scale_kernel<<<g, kBlock, 0, s>>>(vols, tags, m, tag, scale);
}

void tissue_neg_vc(const float* vols, const std::int32_t* tags, std::int64_t m, int tag, float* neg_vc, hipStream_t s) {
    if (const unsigned g = grid_for(m)) tissue_neg_vc_kernel<<<g, kBlock, 0, s>>>(vols, tags, m, tag, neg_vc);
}

void split_offdiag(const std::int32_t* rp, const std::int32_t* ci, const double* v, std::int64_t n, double* off,
                   double* rowsum, hipStream_t s) {
    if (const unsigned g = grid_for(n)) split_offdiag_kernel<<<g, kBlock, 0, s>>>(rp, ci, v, n, off, rowsum);
}

void combine_tissues(const double* in, std::int64_t count, int T, const double* sigma, int k, int K, float* out,
                     hipStream_t s) {
    if (const unsigned g = grid_for(count * K)) combine_kernel<double><<<g, kBlock, 0, s>>>(in, count, T, sigma, k, K, out);
}

void combine_tissues(const float* in, std::int64_t count, int T, const double* sigma, int k, int K, float* out,
                     hipStream_t s) {
    if (const unsigned g = grid_for(count * K)) combine_kernel<float><<<g, kBlock, 0, s>>>(in, count, T, sigma, k, K, out);
}

void accumulate_moments(const float* x, std::int64_t m, int k, int K, double* sum, double* sumsq, hipStream_t s) {
    if (const unsigned g = grid_for(m)) moments_kernel<<<g, kBlock, 0, s>>>(x, m, k, K, sum, sumsq);
}

}  // namespace cn
