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

__global__ void broadcast_kernel(const float* __restrict__ in, std::int64_t count, int K, float* __restrict__ out) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < count * K) out[t] = in[t / K];
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

}  // namespace

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

void broadcast_columns(const float* in, std::int64_t rows_inner, int K, float* out, hipStream_t s) {
    if (const unsigned g = grid_for(rows_inner * K)) broadcast_kernel<<<g, kBlock, 0, s>>>(in, rows_inner, K, out);
}

void accumulate_moments(const float* x, std::int64_t m, int k, int K, double* sum, double* sumsq, hipStream_t s) {
    if (const unsigned g = grid_for(m)) moments_kernel<<<g, kBlock, 0, s>>>(x, m, k, K, sum, sumsq);
}

}  // namespace cn
