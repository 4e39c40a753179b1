// Reduced-basis warm start (warm.hip).
#pragma once
#include <vector>

#include "kernels.hpp"

namespace cn {

constexpr int kMaxWarmBasis = 2048;
constexpr int kMaxWarmModes = 256;
constexpr int kProjectBlocks = 512;

struct WarmBasis {
    std::int64_t n = 0;  // rows (unknowns of the reduced system)
    int cap = 0;         // basis capacity (row stride of `basis`)
    int count = 0;       // columns filled
    bool ready = false;  // a projector matching `count` has been set
    int n_modes = 0;     // mode buffer width (a multiple of 16)
    int used_modes = 0;  // modes set by the caller
    DeviceBuffer<float> basis;           // (n, cap) FP32, row-major
    DeviceBuffer<unsigned short> modes;  // (2, n, n_modes) BF16 hi then lo halves, row-major
    DeviceBuffer<double> P, partials;    // P: (n_modes, n_modes), identity-padded
    DeviceBuffer<unsigned short> coef;
    PinnedStage stage;  // supplied coefficients, copied without a wait

    void reset(std::int64_t n_rows, int capacity);
    // Append columns 0..k-1 of the (n, K) row-major FP64 block X; returns how many fitted.
    int append(const double* X, int k, int K, hipStream_t s);
    // Columns [s0, s0 + 32) as an (n, 32) FP32 block (zero past the capacity).
    void chunk(int s0, float* out, hipStream_t s) const;
    // Modes Phi = W M for M (count x R) FP64 host, row-major; resets P to the identity.
    void set_modes(const double* M_host, int S, int R, hipStream_t s);
    // Columns [j0, j0 + 32) of the modes as an (n, 32) FP32 block.
    void mode_chunk(int j0, float* out, hipStream_t s) const;
    // P (R x R) FP64 host: the inverse of Phi^T A Phi.
    void set_projector(const double* P_host, int R, hipStream_t s);
    // x0 (n, K) FP64 = Phi P Phi^T B for the (n, K) FP32 block B (columns >= k of x0 are zero):
    // project (the block partials of Phi^T B), coefficients (c = P Phi^T B), expand (x0 = Phi c).
    void initial_guess(const float* B, int k, int K, double* x0, hipStream_t s);
    // Columns c0 .. c0 + K - 1 (K <= 16) of B, rows ldb apart (0: K).
    void project(const float* B, int K, hipStream_t s, int ldb = 0, int c0 = 0);
    void coefficients(int k, int K, hipStream_t s);
    // Phi^T B for the first k columns of the (n, K) FP32 block B (K <= 32), into host out (k x R,
    // row-major), before the projector P.
    void project_host(const float* B, int k, int K, double* out, hipStream_t s);
    // Average ms of the three stages of initial_guess over `reps` timed runs (HIP events).
    void bench(const float* B, int k, int K, double* x0, int reps, double* ms3, hipStream_t s);
    // x0 = Phi c for caller-supplied coefficients c (k x R) FP64 host.
    void guess_from_coefficients(const double* c_host, int k, int K, double* x0, hipStream_t s);
    // x0 = Phi c from the coefficients already in `coef`: columns c0 .. c0 + K - 1 (K <= 16) of an
    // x0 whose rows are ldx apart (0: K).
    void expand(int k, int K, double* x0, hipStream_t s, int ldx = 0, int c0 = 0);
};

}  // namespace cn
