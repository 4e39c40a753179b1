// Block-PCG solve on the device: FP32 SpMM and V-cycle, FP64 Gram matrices, pivoted-Cholesky
// A-orthonormalisation of each search block, FP64 iterative refinement of the true residual.
#pragma once

#include <vector>

#include "kernels.hpp"

namespace cn {

constexpr int kBlockWidths[] = {1, 2, 4, 8, 16, 32};
constexpr int kNumBlockWidths = 6;
constexpr int kMaxBlock = 32;

// Partial-sum blocks of every Gram-producing kernel. Fixed, so the association of each FP64
// reduction depends on the row count alone.
constexpr int kGramBlocks = 512;

int padded_width(int k);  // smallest compiled width >= k

struct SolveOptions {
    double tolerance = 1e-6;  // ||b - A x|| / ||b|| per column (FP64 true residual)
    int max_iters = 2000;     // inner iterations over all refinement rounds
    int max_refine = 8;       // outer FP64 refinement rounds
    double inner_margin = 0.5;   // inner target = margin * tol * ||b||
    double drop_tol = 1e-6;      // pivoted Cholesky: relative Schur pivot below which a direction drops
    int check_every = 4;         // graph replays between host convergence reads
    bool use_graph = true;
};

struct SolveStats {
    std::vector<int> column_iterations;    // inner iterations until each column froze
    std::vector<double> relative_residual; // final FP64 true residual per column
    int iterations = 0;                    // inner iterations run (all rounds)
    int refinements = 0;                   // outer rounds run
    int vcycles = 0;
};

// The reduced operator the solver iterates on: FP32 values for SpMM, FP64 for the true residual.
struct Operator {
    CsrView a32;               // FP32 view (ptr, col, val32)
    const double* val64 = nullptr;
};

class BlockSolver {
  public:
    BlockSolver(const Operator& op, const Hierarchy& h, hipStream_t s);
    ~BlockSolver();
    BlockSolver(const BlockSolver&) = delete;
    BlockSolver& operator=(const BlockSolver&) = delete;

    // B: (n, K) row-major FP32 on the device (columns >= k are ignored), X: (n, K) FP64 out,
    // X0: optional (n, K) FP64 initial guess. K = padded_width(k).
    SolveStats solve(const float* B, double* X, const double* X0, int k, const SolveOptions& opt);

    // Kernel-level pieces, exposed for parity tests and benchmarks.
    void spmm(const float* X, float* Y, int K);          // Y = A X (FP32)
    void vcycle(const float* B, float* X, int K);       // X = M B, one zero-guess V-cycle
    // Average device time (ms) over `reps` launches, HIP events on the solver's stream.
    double bench_spmm(int K, int reps, bool fused_gram);
    double bench_vcycle(int K, int reps);
    std::int64_t n() const { return op_.a32.n; }
    hipStream_t stream() const { return s_; }

  private:
    template <int K>
    SolveStats solve_k(const float* B, double* X, const double* X0, int k, const SolveOptions& opt);
    template <int K>
    void vcycle_k(const float* B, float* X);
    template <int K>
    void iteration_k(const SolveOptions& opt);
    void ensure(int K);

    Operator op_;
    const Hierarchy& h_;
    hipStream_t s_;
    int cap_k_ = 0;
    // block vectors (n, K)
    DeviceBuffer<float> R_, Z_, W_, S_, P_, Q_;
    DeviceBuffer<double> R64_, X64_;
    // per-level V-cycle vectors
    std::vector<DeviceBuffer<float>> lb_, lx_, lr_;
    DeviceBuffer<double> cb_, cx_;  // coarsest level (FP64)
    // reductions and small state
    DeviceBuffer<double> partials_;  // kGramBlocks * K * K
    DeviceBuffer<double> G_, T_, C_, alpha_, norms_, target_;
    DeviceBuffer<int> state_;  // [0] done, [1] iteration, [2..2+K) frozen, [2+K..2+2K) col_iters
    int* host_state_ = nullptr;  // pinned
    hipGraphExec_t graph_ = nullptr;
    int graph_k_ = 0;
};

}  // namespace cn
