// Reduced-basis warm start (warm.hip).
#pragma once
#include <vector>

#include "kernels.hpp"

namespace cn {

constexpr int kMaxWarmBasis = 512;
constexpr int kProjectBlocks = 1024;

struct WarmBasis {
    std::int64_t n = 0;  // rows (unknowns of the reduced system)
    int cap = 0;         // basis capacity (row stride of `basis`)
    int count = 0;       // columns filled
    bool ready = false;  // a projector matching `count` has been set
    int padded_for = -1;
    DeviceBuffer<unsigned short> basis;  // (n, cap) BF16 bits, row-major
    DeviceBuffer<double> P, Pfull, partials;
    DeviceBuffer<unsigned short> coef;

    void reset(std::int64_t n_rows, int capacity);
    // Append columns 0..k-1 of the (n, K) row-major FP64 block X; returns how many fitted.
    int append(const double* X, int k, int K, hipStream_t s);
    // Columns [s0, s0 + 32) as an (n, 32) FP32 block (zero past the capacity).
    void chunk(int s0, float* out, hipStream_t s) const;
    // P: (S, S) FP64 host array, S = count: the pseudo-inverse of the Galerkin matrix W^T A W.
    void set_projector(const double* P_host, int S, hipStream_t s);
    // x0 (n, K) FP64 = W P W^T B for the (n, K) FP32 block B (columns >= k of x0 are zero).
    void initial_guess(const float* B, int k, int K, double* x0, hipStream_t s);
};

}  // namespace cn
