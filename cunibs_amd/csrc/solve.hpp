// Block-PCG solve on the device: FP32 SpMM and V-cycle, FP64 Gram matrices, pivoted-Cholesky
// A-orthonormalisation of each search block, FP64 iterative refinement of the true residual.
#pragma once

#include <vector>

#include "kernels.hpp"

namespace cn {

constexpr int kBlockWidths[] = {1, 2, 4, 8, 16, 32};
constexpr int kNumBlockWidths = 6;
constexpr int kMaxBlock = 32;
constexpr int kMaxRounds = 8;

// Partial-sum blocks of every Gram-producing kernel. Fixed, so the association of each FP64
// reduction depends on the row count alone.
constexpr int kGramBlocks = 512;

// Iterations whose solution updates X += W a are deferred and applied together: X (FP64) is read
// and written once per window instead of once per iteration.
constexpr int kLag = 4;

int padded_width(int k);  // smallest compiled width >= k

// The sliced ELL copy of `a` (see CsrView::ell) when its longest row has at most kEllMaxWidth
// entries: fills `pairs` and returns the width; else 0 and `pairs` stays empty.
int build_ell(const CsrView& a, DeviceBuffer<int2>& pairs, hipStream_t s);

// G1 = A^T B1 and, with B2, G2 = A^T B2 for (n, K) FP32 device blocks, into out (device, K*K then
// K*K more with B2), the iteration's Gram kernels and fixed-order reduction, launched on s without
// a wait. Allocates nothing: the partials live in a file-scope device array (one per device).
void gram_noalloc(std::int64_t n, int K, const float* A, const float* B1, const float* B2, double* out,
                  hipStream_t s);

// The zero-guess V-cycle from one level down (that level's smoothing, every coarser level and the
// coarse solve) as one dense FP32 operator M (n x n), formed by running that part of the cycle on
// the identity's columns: the solve applies it as one matrix-core product. M is held in the
// product's 16 x 4 operand tiles (nq = ceil(n / 4) column quads per 16-row tile, zero past n).
// Levels of up to kDenseTailMaxRows rows qualify (M at most 64 MB); the first one does.
struct DenseTail {
    int level = -1;  // -1: none
    int n = 0, nq = 0;
    DeviceBuffer<float> m;
};
constexpr int kDenseTailMaxRows = 4096;

// The inner margin by block width. A column above the tolerance after a round starts another
// round, so the margin trades iterations against extra rounds. Up to 16 columns the FP64 true
// residual lands within a few per cent of the recursive one (the differenced SpMM keeps the FP32
// recursion accurate: 8.2e-7 at most over a 768-placement sweep at 0.8); a 32-column block's
// recursion, even with its FP64 residual step, ends nearer the tolerance (9.9e-7 after a round at
// 0.8 on the 1.5 mm head), so it keeps 0.5.
inline double margin_for_width(int K) { return K <= 16 ? 0.8 : 0.5; }

struct SolveOptions {
    double tolerance = 1e-6;    // ||b - A x|| / ||b|| per column (FP64 true residual)
    int max_iters = 2000;       // inner iterations over all refinement rounds
    int max_refine = kMaxRounds;  // outer FP64 refinement rounds
    // Inner target = margin * tol * ||b|| on the recursive residual (margin_for_width when <= 0).
    double inner_margin = 0.0;
    double drop_tol = 1e-6;     // relative Schur pivot of the block Gram below which a direction drops
    int check_every = 4;        // graph replays between host convergence reads
    bool use_graph = true;
    bool independent = false;   // k independent CG chains (diagonal Grams) instead of block CG
};

struct SolveStats {
    std::vector<int> column_iterations;     // inner iterations until each column froze
    std::vector<double> relative_residual;  // final FP64 true residual per column
    int iterations = 0;                     // inner iterations run (all rounds)
    int refinements = 0;                    // outer rounds run
    int vcycles = 0;
    int round_iterations[kMaxRounds] = {};  // cumulative iterations at the end of each round
    double round_residual[kMaxRounds] = {}; // worst true relative residual after each round
};

// The reduced operator the solver iterates on.
//   a32:  FP32 CSR (the V-cycle's finest level and the plain SpMM)
//   off/rowsum: the same operator with the diagonal folded into row sums, applied as
//         y_i = sum_{j != i} a_ij (x_j - x_i) + s_i x_i, which keeps the FP32 product free of the
//         cancellation between a stiffness row's entries (they sum to ~0)
//   val64: FP64 values for the true residual
struct Operator {
    CsrView a32;
    CsrView off;
    const float* rowsum = nullptr;
    const double* val64 = nullptr;
};

constexpr int kMaxTissues = 16;

// K matrices sharing the reduced operator's pattern (conductivity samples): column c of every
// block is solved with A_c = sum_t sigma[c, t] A_t, as K independent CG chains (the block Grams
// restricted to their diagonals) under the shared preconditioner.
struct MultiOperator {
    int width = 0;                     // K the values below were laid out for
    const float* off_vals = nullptr;   // (nnz_off, K) differenced off-diagonal values
    const float* rowsum = nullptr;     // (n, K) row sums
    const double* tissue_vals = nullptr;  // (T, nnz) FP64 per-tissue values on the reduced pattern
    int n_tissue = 0;
    const double* sigma = nullptr;     // (K, T) device
};

class BlockSolver {
  public:
    BlockSolver(const Operator& op, const Hierarchy& h, hipStream_t s);
    ~BlockSolver();
    BlockSolver(const BlockSolver&) = delete;
    BlockSolver& operator=(const BlockSolver&) = delete;

    // B: (n, K) row-major FP32 on the device (columns >= k are zero), X: (n, K) FP64 out,
    // X0: optional (n, K) FP64 initial guess. K = padded_width(k).
    SolveStats solve(const float* B, double* X, const double* X0, int k, const SolveOptions& opt);

    // Kernel-level pieces, exposed for parity tests and benchmarks.
    void spmm(const float* X, float* Y, int K);     // Y = A X (FP32, plain CSR)
    void spmm_diff(const float* X, float* Y, int K);  // Y = A X (FP32, differenced form)
    void vcycle(const float* B, float* X, int K);   // X = M B, one zero-guess V-cycle
    // Multi-matrix mode (nullptr: the single operator). The pointed-to values must outlive the solves.
    void set_multi(const MultiOperator* m);
    // out (32 x 32, FP64, host) = A^T B for (n, 32) FP32 blocks, fixed-order reduction.
    void cross_gram32(const float* A, const float* B, double* out_host);
    // G1 = A^T B1 and, with B2, G2 = A^T B2 (FP64, the iteration's Gram kernels and reduction)
    // into out_host (K*K, then K*K more with B2); A, B1, B2 device (n, K).
    void gram(const float* A, const float* B1, const float* B2, double* out_host, int K);
    // Galerkin initial guesses from earlier blocks' solutions (a sweep's projection basis), all
    // queued on the solver's stream without a host wait. guess_reset empties a basis of up to
    // `capacity` (1 or 2) blocks of width K (capacity K <= 32); guess_push adds a solved block (its
    // FP64 solution and FP32 right-hand side, n x K; the oldest leaves a full basis); guess writes
    // x0 (n x K, FP64), the projection of the solution of A x = b onto the basis. Each push costs
    // one Gram of its block; the cross block of a two-block basis is the projection the last guess
    // already formed when that guess was for the block pushed after it.
    void guess_reset(int capacity, int K);
    void guess_push(const double* X, const float* B, int K);
    void guess(const float* b, int K, double drop_tol, double* x0);
    int guess_blocks() const { return g_count_; }
    // Average device time (ms) over `reps` launches, HIP events on the solver's stream.
    // mode 0: plain SpMM; 1: differenced SpMM (the iteration's S = A W); 2: 1 followed by the Gram
    // pass G = W^T S, H = W^T R. sub > 0 forces the sub-lanes per column group of the fine level.
    double bench_spmm(int K, int reps, int mode, int sub);
    double bench_vcycle(int K, int reps);
    // Average ms of a streaming copy of `bytes` bytes (read `bytes`, write `bytes`).
    double bench_copy(std::int64_t bytes, int reps);
    std::int64_t n() const { return op_.a32.n; }
    hipStream_t stream() const { return s_; }
    // Sub-lanes per column group of the fine level's row kernels (the SpMM tile); 0 = chosen from
    // the level's shape. Drops the captured iterations.
    void set_fine_sub(int sub);
    int fine_sub_setting() const { return fine_sub_set_; }
    void set_coarse_sub(int sub);
    // Apply the V-cycle's bottom levels as one dense product (see DenseTail); on by default.
    // Drops the captured iterations.
    void set_dense_tail(bool on) {
        tail_on_ = on;
        drop_graphs();
    }
    bool dense_tail_setting() const { return tail_on_; }
    // The level the dense tail starts at (-1: none), forming it first if it is due.
    int dense_tail_level() {
        ensure_tail();
        return tail_.level;
    }
    // The differenced operator's sliced ELL copy for the block iteration's SpMM (nullptr: CSR).
    // Drops the captured iterations.
    void set_off_ell(const int2* ell, int w, int max_k) {
        op_.off.ell = ell;
        op_.off.ell_w = w;
        op_.off.ell_max_k = max_k;
        drop_graphs();
    }
    int coarse_sub_setting() const { return coarse_sub_; }
    // Forget every captured iteration (after a change to the hierarchy's smoother).
    void invalidate() { drop_graphs(); }
    // In-solve timing on the device wall clock: with it on, the differenced SpMM's blocks stamp
    // their start and end, stamp kernels mark the V-cycle and the iteration, and an accumulator
    // kernel at the end of each iteration (captured with it) adds the spans of every iteration
    // that ran in full. Switching drops the captured iterations; every call resets the sums.
    void set_timing(bool on);
    // Failure handling test hook: the next n graph captures report failure, so those solves run
    // their iterations as direct launches.
    void fail_next_captures(int n) { fail_captures_ = n; }
    // Captured executables alive (current and retired), for the tests of the graph cache.
    int graph_execs() const;
    bool timing_on() const { return timing_; }
    // out[0] iterations sampled, out[1..3] mean ms of the SpMM, the V-cycle and the iteration.
    void timing(double* out) const;

  private:
    template <int K>
    SolveStats solve_k(const float* B, double* X, const double* X0, int k, const SolveOptions& opt);
    template <int K>
    void vcycle_k(const float* B, float* X, const int* state);
    template <int K>
    // Iteration at ring position j (the round's iteration count mod kLag): reads W slot j and
    // writes slot j + 1, except at the last position, which first applies the window's deferred
    // updates X += W_j a_j and then writes slot 0. Windows are thus aligned to the round's own
    // iteration count, whatever batches the host launches them in.
    void iteration_k(const SolveOptions& opt, int j);
    // the iterations at positions j0 .. j0 + L - 1 (clearing the window's slots when j0 = 0)
    template <int K>
    void window_k(const SolveOptions& opt, int j0, int L);
    // X += the deferred updates of a window cut short after L iterations (the round ended)
    template <int K>
    void flush_pending_k(int L);
    template <int K>
    void spmm_block_k(const int* state, const float* W);
    template <int K>
    int fine_sub() const;
    void ensure(int K);
    // Form the dense tail if it is on and not current (never during a capture: solve entry points
    // call it before any launch).
    void ensure_tail();
    template <int K>
    void build_tail_k(int t);

    Operator op_;
    const Hierarchy& h_;
    hipStream_t s_;
    int cap_k_ = 0;
    int n_cu_ = 256;
    int fine_sub_set_ = 0;       // set_fine_sub's value (0: automatic)
    int fine_sub_override_ = 0;  // the fine level's sub-lane count in force (bench_spmm may force one)
    int coarse_sub_ = 0;         // sub-lanes of the coarser levels' row kernels (0: automatic)
    const MultiOperator* multi_ = nullptr;
    DenseTail tail_;
    bool tail_on_ = true;
    bool tail_ready_ = false;  // tail_ matches the hierarchy (cleared with the captured iterations)
    // block vectors (n, K)
    // W_ is slot 0 of the search-block ring (kLag slots), so the deferred solution update still
    // has every W of its window.
    DeviceBuffer<float> R_, Z_, W_, S_;
    DeviceBuffer<float> Wx_[kLag - 1];
    float* wslot(int j) const { return j == 0 ? W_.get() : Wx_[j - 1].get(); }
    DeviceBuffer<double> R64_, X64_;
    // per-level V-cycle vectors
    std::vector<DeviceBuffer<float>> lb_, lx_, lr_;
    DeviceBuffer<float> cb_, cx_;  // coarsest level
    // reductions and small state
    DeviceBuffer<double> partials_;  // kGramBlocks * 2 * K * K
    DeviceBuffer<double> red_;       // the partials summed over the blocks (2 K^2)
    DeviceBuffer<double> G_, TT_, AL_, D_, norms_, target_;  // AL_: kLag K x K coefficient slots
    DeviceBuffer<double> bnorm_;  // ||b_c||, kept on the device for the first round's targets
    DeviceBuffer<float> g_x_[2], g_b_[2];  // the guess basis: FP32 solutions and their right-hand sides
    DeviceBuffer<double> g_mat_;           // its Grams, cross blocks and the last guess's coefficients
    int g_cap_ = 0, g_k_ = 0, g_count_ = 0, g_newest_ = 0;
    bool g_pending_ = false;
    DeviceBuffer<int> ran_;  // per window slot: the iteration ran (its coefficients are to be applied)
    DeviceBuffer<int> state_;  // [0] done, [1] iteration, [2..2+K) frozen, [2+K..2+2K) col_iters
    int* host_state_ = nullptr;  // pinned
    PinnedStage stage_;          // the round's targets and initial state, copied without a wait
    // One captured iteration per block width, kept for the solver's lifetime: every buffer the
    // iteration names is solver-owned and reallocated only when the capacity grows (which drops
    // them all).
    bool timing_ = false;
    DeviceBuffer<unsigned long long> stamps_;  // SpMM block stamps, then 4 marks
    DeviceBuffer<double> t_sums_;
    double clock_khz_ = 1.0;
    unsigned spmm_grid_ = 0;
    std::size_t stamp_marks() const;
    unsigned long long* stamps();
    void mark(int i, const int* state);
    // Captured iterations, one set per operator mode: the single operator, or one multi-matrix
    // layout (its value buffers and width), so that alternating between them captures nothing
    // again. Per width: a whole window of kLag iterations, one iteration at each ring position, and
    // the restart of a round (the true residual cast to FP32, then W = M R).
    struct GraphSet {
        bool in_use = false;
        bool multi = false;
        MultiOperator key{};
        unsigned long long last_use = 0;
        hipGraphExec_t win[kNumBlockWidths] = {};
        hipGraphExec_t one[kNumBlockWidths][kLag] = {};
        hipGraphExec_t restart[kNumBlockWidths] = {};
        double drop[kNumBlockWidths] = {};  // the drop tolerance each width was captured with
        int diag[kNumBlockWidths] = {};     // and whether its Grams were diagonal
    };
    static constexpr int kGraphSets = 4;
    GraphSet sets_[kGraphSets];
    unsigned long long use_clock_ = 0;
    GraphSet& graph_set();  // the current mode's set; a new mode takes the least recently used one
    void retire(hipGraphExec_t& g);
    // Capture `body` into gx; false (nothing recorded, nothing run) when the capture or the
    // instantiation fails, or when the failure hook asks for it.
    template <typename F>
    bool capture(hipGraphExec_t& gx, F&& body);
    int fail_captures_ = 0;
    // Inner iterations the last solve at each width took in its first round: the next solve runs
    // the whole windows short of that count before its first convergence read, then reads after
    // every iteration for a few, so a sweep of similar blocks neither waits on early reads nor
    // replays finished iterations (each a graph of early-exiting launches).
    int last_first_round_[kNumBlockWidths] = {};
    // Replaced executables wait here and the oldest are destroyed beyond kMaxRetired.
    static constexpr std::size_t kMaxRetired = 64;
    std::vector<hipGraphExec_t> retired_;
    void drop_graphs();
};

}  // namespace cn
