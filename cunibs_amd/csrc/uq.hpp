// Conductivity-uncertainty kernels (uq.hip).
#pragma once
#include "kernels.hpp"

namespace cn {

// scale[e] = vols[e] if tags[e] == tag else 0 (FP64): the unit-conductivity stiffness of one tissue.
void tissue_scale(const double* vols, const std::int32_t* tags, std::int64_t m, int tag, double* scale, hipStream_t s);
// neg_vc[e] = -vols[e] if tags[e] == tag else 0 (FP32): the unit-conductivity RHS weights of one tissue.
void tissue_neg_vc(const float* vols, const std::int32_t* tags, std::int64_t m, int tag, float* neg_vc, hipStream_t s);
// Off-diagonal values (row order, FP64) and FP64 row sums of a reduced CSR.
void split_offdiag(const std::int32_t* rp, const std::int32_t* ci, const double* v, std::int64_t n, double* off,
                   double* rowsum, hipStream_t s);
// out[i * K + c] = (float) sum_t sigma[c * T + t] in[t * count + i] for c < k, 0 for k <= c < K.
void combine_tissues(const double* in, std::int64_t count, int T, const double* sigma, int k, int K, float* out,
                     hipStream_t s);
void combine_tissues(const float* in, std::int64_t count, int T, const double* sigma, int k, int K, float* out,
                     hipStream_t s);
// sum[e] += x[e, c], sumsq[e] += x[e, c]^2 over c < k, FP64, columns in order.
void accumulate_moments(const float* x, std::int64_t m, int k, int K, double* sum, double* sumsq, hipStream_t s);

// Y (n, K) FP64 = A X for a reduced CSR with FP64 values and X (n, K) FP64 (columns < k; the
// rest of Y zero).
void spmm64(const std::int32_t* rp, const std::int32_t* ci, const double* v, std::int64_t n, const double* X, int k,
            int K, double* Y, hipStream_t s);
// out (k x k, FP64, host) = A^T B for (n, K) FP64 blocks, columns < k; fixed-order two-stage sum.
void gram64(const double* A, const double* B, std::int64_t n, int k, int K, double* out_host, hipStream_t s);
// Sensitivity right-hand sides of the conductivity draws around x_nom:
// R[i, t] = b_t[i] - (K_t x_nom)[i] (FP64, (n, K)) and Bf = (float) R, for tissues t < T.
void sensitivity_rhs(const std::int32_t* rp, const std::int32_t* ci, const double* vals, std::int64_t nnz,
                     const float* b_tissue, const double* x_nom, std::int64_t n, int T, int K, double* R, float* Bf,
                     hipStream_t s);
// X0[i, c] = x_nom[i] + sum_t S[i, t] coef[t * K + c] for c < k, 0 for k <= c < K (FP64).
void expand_guess(const double* x_nom, const double* S, std::int64_t n, int T, int KS, const double* coef, int k, int K,
                  double* X0, hipStream_t s);

}  // namespace cn
