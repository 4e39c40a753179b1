// Smoothed-aggregation AMG setup on the device: l1 diagonal, SIZE_4 pairwise aggregation, the
// tentative and smoothed prolongators, the Galerkin product and the coarsest-level FP64 inverse.
//
// Every sparse product C = A B here folds, for each output entry, the products a_ij * b_jk with
// fmaf in the order j appears in row i of A, starting from zero. That order is a property of the
// inputs alone, so setup is bit-reproducible and matches the CPU oracle exactly.
#include <algorithm>
#include <vector>

#include "kernels.hpp"
#include "prims.hpp"

namespace cn {

namespace {

// ---------------------------------------------------------------------------------------------
// l1 diagonal, solver/amg/l1.cu:15-35.
// ---------------------------------------------------------------------------------------------
__global__ void l1_dinv_kernel(const std::int32_t* __restrict__ rp, const std::int32_t* __restrict__ ci,
                               const float* __restrict__ v, std::int64_t n, float* __restrict__ dinv) {
    const std::int64_t row = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (row >= n) return;
    float acc = 0.f, diag = 0.f;
    for (int p = rp[row]; p < rp[row + 1]; ++p) {
        acc += fabsf(v[p]);
        if (ci[p] == row) diag = v[p];
    }
    if (diag < 0.f) acc = -acc;
    if (acc == 0.f) acc = 1.f;
    dinv[row] = 1.f / acc;
}

// ---------------------------------------------------------------------------------------------
// SIZE_4 aggregation, solver/amg/aggregate.cu:20-278. One thread per row in every pass; each pass
// reads only state no other row of the same pass writes, so the result is schedule-independent.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ unsigned agg_hash(unsigned a, unsigned seed) {
    a ^= seed;
    a = (a + 0x7ed55d16u) + (a << 12);
    a = (a ^ 0xc761c23cu) + (a >> 19);
    a = (a + 0x165667b1u) + (a << 5);
    a = (a ^ 0xd3a2646cu) + (a << 9);
    a = (a + 0xfd7046c5u) + (a << 3);
    a = (a ^ 0xb55a4f09u) + (a >> 16);
    return a;
}

__device__ __forceinline__ bool is_candidate(int col, std::int64_t row, std::int64_t n) {
    return col != row && static_cast<unsigned>(col) < static_cast<unsigned>(n);
}

struct Best {
    float w;
    int col;
};

__device__ __forceinline__ void best_take(Best& acc, float w, int col) {
    if (w > acc.w || (w == acc.w && col > acc.col)) {
        acc.w = w;
        acc.col = col;
    }
}

// Adds the block's count of set predicates to *counter (integer, order-free): each wavefront's
// count by ballot, the wavefronts' counts summed in LDS, one atomic per block (a pass over a
// million rows otherwise queues thousands of atomics on one address). Every thread of the block
// must call it.
__device__ __forceinline__ void count_block(int pred, int* counter) {
    __shared__ int wave_counts[kBlock / kWave];
    const unsigned long long m = __ballot(pred);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & (kWave - 1)) == 0) wave_counts[wave] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int c = 0;
        for (int w = 0; w < static_cast<int>(blockDim.x) / kWave; ++w) c += wave_counts[w];
        if (c) atomicAdd(counter, c);
    }
}

// Matching passes run under device-side control: every pass's kernels return at once when `stop`
// is set, and a one-thread decision after each pass sets it by the rules a host loop would apply
// to the pass's count of unmatched rows, so the passes are queued back to back with no readback.
struct AggLoop {
    int remaining;  // unmatched rows counted by the current pass
    int previous;   // the count after the previous pass (n before the first)
    int stop;
};

__global__ void agg_loop_init_kernel(AggLoop* __restrict__ c, int n) {
    c->remaining = 0;
    c->previous = n;
    c->stop = 0;
}

// Stop when no row is left unmatched or a pass matched none; in the second round also when fewer
// than 5 % of the rows are left (aggregate.cu's loop exits).
__global__ void agg_loop_decide_kernel(AggLoop* __restrict__ c, int n, int second_round) {
    if (c->stop) return;
    const int u = c->remaining;
    bool stop = u == 0 || u == c->previous;
    if (second_round) stop = stop || static_cast<float>(u) / static_cast<float>(n) < 0.05f;
    c->stop = stop ? 1 : 0;
    c->previous = u;
    c->remaining = 0;
}

__global__ void agg_diag_kernel(const std::int32_t* __restrict__ rp, const std::int32_t* __restrict__ ci,
                                const float* __restrict__ v, std::int64_t n, float* __restrict__ diag) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float d = 0.f;
    for (int j = rp[i]; j < rp[i + 1]; ++j)
        if (ci[j] == i) {
            d = v[j];
            break;
        }
    diag[i] = d;
}

__global__ void agg_edge_weight_kernel(const std::int32_t* __restrict__ rp, const std::int32_t* __restrict__ ci,
                                       const float* __restrict__ v, const float* __restrict__ diag,
                                       std::int64_t n, float* __restrict__ w) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float di = fabsf(diag[i]);
    for (int j = rp[i]; j < rp[i + 1]; ++j) {
        const int jc = ci[j];
        if (!is_candidate(jc, i, n)) {
            w[j] = -1.0f;
            continue;
        }
        const float den = fmaxf(di, fabsf(diag[jc]));
        const float ed = fabsf(v[j]) / den;
        const unsigned lo = static_cast<unsigned>(min(static_cast<int>(i), jc));
        const unsigned hi = static_cast<unsigned>(max(static_cast<int>(i), jc));
        const float frac = 1e-5f * static_cast<float>(agg_hash(lo, hi)) / 4294967295.0f;
        w[j] = fmaf(frac, ed, ed);
    }
}

__global__ void agg_init_kernel(std::int64_t n, std::int32_t* __restrict__ agg, std::int32_t* __restrict__ partner,
                                std::int32_t* __restrict__ strongest) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    agg[i] = static_cast<std::int32_t>(i);
    partner[i] = -1;
    strongest[i] = -1;
}

__global__ void agg_find_strongest_nomerge_kernel(const std::int32_t* __restrict__ rp,
                                                  const std::int32_t* __restrict__ ci,
                                                  const float* __restrict__ w, std::int64_t n,
                                                  const std::int32_t* __restrict__ partner,
                                                  std::int32_t* __restrict__ strongest,
                                                  const AggLoop* __restrict__ ctl) {
    if (ctl->stop) return;
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n || partner[t] != -1) return;
    Best b{0.0f, -1};
    for (int j = rp[t]; j < rp[t + 1]; ++j) {
        const int jc = ci[j];
        if (!is_candidate(jc, t, n) || partner[jc] != -1) continue;
        best_take(b, w[j], jc);
    }
    if (b.col != -1) strongest[t] = b.col;
}

__global__ void agg_match_edges_kernel(std::int64_t n, std::int32_t* __restrict__ partner,
                                       std::int32_t* __restrict__ agg, const std::int32_t* __restrict__ strongest,
                                       AggLoop* __restrict__ ctl) {
    if (ctl->stop) return;
    int* const remaining = &ctl->remaining;
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    int unassigned = 0;
    if (t < n && partner[t] == -1) {
        const int pm = strongest[t];
        if (pm != -1 && strongest[pm] == t) {
            partner[t] = pm;
            agg[t] = min(pm, static_cast<int>(t));
        } else {
            unassigned = 1;
        }
    }
    count_block(unassigned, remaining);
}

__global__ void agg_round2_init_kernel(std::int64_t n, std::int32_t* __restrict__ partner,
                                       float* __restrict__ wsn, std::int32_t* __restrict__ aggregated) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    if (partner[t] == -1) partner[t] = static_cast<std::int32_t>(t);
    wsn[t] = -1.0f;
    aggregated[t] = -1;
}

__global__ void agg_find_strongest_store_kernel(const std::int32_t* __restrict__ rp,
                                                const std::int32_t* __restrict__ ci,
                                                const float* __restrict__ w, std::int64_t n,
                                                const std::int32_t* __restrict__ aggregated,
                                                const std::int32_t* __restrict__ agg,
                                                const std::int32_t* __restrict__ partner,
                                                std::int32_t* __restrict__ strongest, float* __restrict__ wsn,
                                                const AggLoop* __restrict__ ctl) {
    if (ctl->stop) return;
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n || aggregated[t] != -1) return;
    const int p = partner[t];
    Best b{0.0f, -1};
    for (int j = rp[t]; j < rp[t + 1]; ++j) {
        const int jc = ci[j];
        if (!is_candidate(jc, t, n)) continue;
        if (aggregated[jc] != -1 || jc == p) continue;
        best_take(b, w[j], jc);
    }
    if (b.col != -1) {
        wsn[t] = b.w;
        strongest[t] = agg[b.col];
    }
}

__global__ void agg_agree_on_proposal_kernel(std::int64_t n, std::int32_t* __restrict__ aggregated,
                                             std::int32_t* strongest, const float* __restrict__ wsn,
                                             const std::int32_t* __restrict__ partner,
                                             const AggLoop* __restrict__ ctl) {
    if (ctl->stop) return;
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n || aggregated[t] != -1) return;
    const int p = partner[t];
    const float mine = wsn[t];
    const float theirs = (p != -1) ? wsn[p] : -1.0f;
    if (mine < 0.0f && theirs < 0.0f) {
        aggregated[t] = 1;
        strongest[t] = -1;
    } else if (mine < theirs) {
        strongest[t] = strongest[p];
    }
}

__global__ void agg_match_aggregates_kernel(std::int64_t n, std::int32_t* __restrict__ agg,
                                            std::int32_t* __restrict__ aggregated,
                                            const std::int32_t* __restrict__ strongest,
                                            AggLoop* __restrict__ ctl) {
    if (ctl->stop) return;
    int* const remaining = &ctl->remaining;
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    int unassigned = 0;
    if (t < n && aggregated[t] == -1) {
        const int pm = strongest[t];
        const int mine = agg[t];
        if (pm != -1 && strongest[pm] == mine) {
            aggregated[t] = 1;
            agg[t] = min(pm, mine);
        } else {
            unassigned = 1;
        }
    }
    count_block(unassigned, remaining);
}

__global__ void agg_merge_existing_kernel(const std::int32_t* __restrict__ rp, const std::int32_t* __restrict__ ci,
                                          const float* __restrict__ w, std::int64_t n,
                                          const std::int32_t* __restrict__ agg,
                                          const std::int32_t* __restrict__ aggregated,
                                          std::int32_t* __restrict__ cand) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n || aggregated[t] != -1) return;
    Best b{0.0f, -1};
    for (int j = rp[t]; j < rp[t + 1]; ++j) {
        const int jc = ci[j];
        if (!is_candidate(jc, t, n) || aggregated[jc] == -1) continue;
        best_take(b, w[j], jc);
    }
    cand[t] = (b.col != -1) ? agg[b.col] : static_cast<std::int32_t>(t);
}

__global__ void agg_join_existing_kernel(std::int64_t n, std::int32_t* __restrict__ agg,
                                         std::int32_t* __restrict__ aggregated,
                                         const std::int32_t* __restrict__ cand) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n && aggregated[t] == -1) {
        agg[t] = cand[t];
        aggregated[t] = 1;
    }
}

// Rows of one aggregate all store 1 here; any of the identical stores may land.
__global__ void agg_mark_used_kernel(std::int64_t n, const std::int32_t* __restrict__ agg,
                                     std::int32_t* __restrict__ used) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) used[agg[t]] = 1;
}

__global__ void agg_relabel_kernel(std::int64_t n, std::int32_t* __restrict__ agg,
                                   const std::int32_t* __restrict__ excl) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) agg[t] = excl[agg[t]];
}

// agg = step[agg]: the composition of two pairwise passes.
__global__ void compose_kernel(std::int32_t* __restrict__ agg, std::int64_t n,
                               const std::int32_t* __restrict__ step) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) agg[t] = step[agg[t]];
}

// ---------------------------------------------------------------------------------------------
// Sparse products. One wavefront per output row. Columns are collected in an LDS hash set; the
// fill pass folds each product into its column's LDS accumulator while walking row i of A in
// order. Within one row of B the columns are distinct, so the 64 lanes never touch the same slot
// in a step, and the steps themselves are ordered by the position in A's row: each output entry
// is folded in exactly the oracle's order. The row is then written out sorted by column.
// ---------------------------------------------------------------------------------------------
__global__ void candidates_kernel(const std::int32_t* __restrict__ ap, const std::int32_t* __restrict__ ac,
                                  const std::int32_t* __restrict__ bp, std::int64_t n_rows,
                                  std::int32_t* __restrict__ cand) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n_rows) return;
    int c = 0;
    for (int p = ap[i]; p < ap[i + 1]; ++p) c += bp[ac[p] + 1] - bp[ac[p]];
    cand[i] = c;
}

__device__ __forceinline__ unsigned hash_slot(int key, unsigned mask) {
    return (static_cast<unsigned>(key) * 2654435761u) & mask;
}

// Inserts key; returns its slot, or -1 when `cap` probes found neither the key nor a free slot
// (the table is full). The probe count is bounded, so a full table cannot spin.
__device__ __forceinline__ int hash_insert(int* keys, int cap, int key, int* inserted) {
    const unsigned mask = static_cast<unsigned>(cap - 1);
    unsigned s = hash_slot(key, mask);
    *inserted = 0;
    for (int probe = 0; probe < cap; ++probe) {
        const int prev = atomicCAS(&keys[s], -1, key);
        if (prev == -1) {
            *inserted = 1;
            return static_cast<int>(s);
        }
        if (prev == key) return static_cast<int>(s);
        s = (s + 1) & mask;
    }
    return -1;
}

__global__ void __launch_bounds__(kWave) spgemm_count_kernel(
    const std::int32_t* __restrict__ ap, const std::int32_t* __restrict__ ac,
    const std::int32_t* __restrict__ bp, const std::int32_t* __restrict__ bc, std::int64_t n_rows,
    int cap, std::int32_t* __restrict__ counts, int* __restrict__ overflow) {
    extern __shared__ int lds[];
    int* keys = lds;
    const std::int64_t row = blockIdx.x;
    if (row >= n_rows) return;
    const int lane = threadIdx.x;
    for (int s = lane; s < cap; s += kWave) keys[s] = -1;
    __syncthreads();
    int mine = 0, full = 0;
    for (int p = ap[row]; p < ap[row + 1]; ++p) {
        const int j = ac[p];
        for (int q = bp[j] + lane; q < bp[j + 1]; q += kWave) {
            int ins = 0;
            if (hash_insert(keys, cap, bc[q], &ins) < 0) full = 1;
            mine += ins;
        }
    }
    for (int off = kWave / 2; off > 0; off >>= 1) mine += __shfl_down(mine, off, kWave);
    if (lane == 0) counts[row] = mine;
    if (full) atomicOr(overflow, 1);
}

__global__ void __launch_bounds__(kWave) spgemm_fill_kernel(
    const std::int32_t* __restrict__ ap, const std::int32_t* __restrict__ ac, const float* __restrict__ av,
    const std::int32_t* __restrict__ bp, const std::int32_t* __restrict__ bc, const float* __restrict__ bv,
    std::int64_t n_rows, int cap, const std::int32_t* __restrict__ cp, std::int32_t* __restrict__ cc,
    float* __restrict__ cv) {
    extern __shared__ int lds[];
    int* keys = lds;
    float* acc = reinterpret_cast<float*>(lds + cap);
    __shared__ int n_used;
    const std::int64_t row = blockIdx.x;
    if (row >= n_rows) return;
    const int lane = threadIdx.x;
    for (int s = lane; s < cap; s += kWave) {
        keys[s] = -1;
        acc[s] = 0.0f;
    }
    __syncthreads();
    for (int p = ap[row]; p < ap[row + 1]; ++p) {
        const int j = ac[p];
        const float a = av[p];
        for (int q = bp[j] + lane; q < bp[j + 1]; q += kWave) {
            int ins = 0;
            // cap >= 2 x this row's distinct count (from the count pass), so the slot exists.
            const int s = hash_insert(keys, cap, bc[q], &ins);
            if (s >= 0) acc[s] = fmaf(a, bv[q], acc[s]);
        }
        __syncthreads();  // step p's folds land before step p+1 reads them
    }
    // Compact the occupied slots to the front of a second region, then place each entry by rank.
    int* ckeys = reinterpret_cast<int*>(acc + cap);
    float* cvals = reinterpret_cast<float*>(ckeys + cap);
    if (lane == 0) n_used = 0;
    __syncthreads();
    for (int s = lane; s < cap; s += kWave) {
        if (keys[s] != -1) {
            const int at = atomicAdd(&n_used, 1);
            ckeys[at] = keys[s];
            cvals[at] = acc[s];
        }
    }
    __syncthreads();
    const int u = n_used;
    const int base = cp[row];
    for (int t = lane; t < u; t += kWave) {
        const int k = ckeys[t];
        int rank = 0;
        for (int o = 0; o < u; ++o) rank += ckeys[o] < k;
        cc[base + rank] = k;
        cv[base + rank] = cvals[t];
    }
}

// ---------------------------------------------------------------------------------------------
// Prolongators and transpose.
// ---------------------------------------------------------------------------------------------
__global__ void agg_counts_kernel(const std::int32_t* __restrict__ agg, std::int64_t n,
                                  std::int32_t* __restrict__ counts) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(&counts[agg[i]], 1);
}

// fem/solve.py:94-107: P_tent[i, agg_i] = 1/sqrt(|agg_i|).
__global__ void tentative_kernel(const std::int32_t* __restrict__ agg, const std::int32_t* __restrict__ counts,
                                 std::int64_t n, std::int32_t* __restrict__ pp, std::int32_t* __restrict__ pc,
                                 float* __restrict__ pv) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i > n) return;
    pp[i] = static_cast<std::int32_t>(i);
    if (i == n) return;
    const int a = agg[i];
    pc[i] = a;
    pv[i] = 1.0f / sqrtf(static_cast<float>(counts[a]));
}

// fem/solve.py:110-126: P = P_tent - (4/3) D^-1 (A P_tent), applied in place over A P_tent.
__global__ void smooth_prolongator_kernel(const std::int32_t* __restrict__ app, const std::int32_t* __restrict__ apc,
                                          float* __restrict__ apv, const float* __restrict__ dinv,
                                          const std::int32_t* __restrict__ agg, const float* __restrict__ tent,
                                          std::int64_t n) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float d = dinv[i];
    const int own = agg[i];
    const float base = tent[i];
    const float damp = 4.0f / 3.0f;
    for (int p = app[i]; p < app[i + 1]; ++p) {
        const float t = damp * (apv[p] * d);
        apv[p] = (apc[p] == own ? base : 0.0f) - t;
    }
}

__global__ void expand_rows_kernel(const std::int32_t* __restrict__ rp, std::int64_t n, std::int32_t* __restrict__ rows) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int p = rp[i]; p < rp[i + 1]; ++p) rows[p] = static_cast<std::int32_t>(i);
}

__global__ void gather_transpose_kernel(const std::int32_t* __restrict__ order, const std::int32_t* __restrict__ rows,
                                        const float* __restrict__ val, std::int64_t nnz,
                                        std::int32_t* __restrict__ tc, float* __restrict__ tv) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nnz) return;
    const int src = order[t];
    tc[t] = rows[src];
    tv[t] = val[src];
}

// ---------------------------------------------------------------------------------------------
// Coarsest level: dense FP64 copy, then an in-place Gauss-Jordan inverse (SPD: no pivoting).
// ---------------------------------------------------------------------------------------------
__global__ void csr_to_dense_kernel(const std::int32_t* __restrict__ rp, const std::int32_t* __restrict__ ci,
                                    const float* __restrict__ v, int n, double* __restrict__ dense) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int j = 0; j < n; ++j) dense[static_cast<std::int64_t>(i) * n + j] = 0.0;
    for (int p = rp[i]; p < rp[i + 1]; ++p) dense[static_cast<std::int64_t>(i) * n + ci[p]] = v[p];
}

// One Gauss-Jordan step on pivot k, in two launches so no thread reads what another writes:
// save the pivot row and column, then update every entry from the saved copies.
__global__ void gj_save_kernel(const double* __restrict__ a, int n, int k, double* __restrict__ row,
                               double* __restrict__ col) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    row[i] = a[static_cast<std::int64_t>(k) * n + i];
    col[i] = a[static_cast<std::int64_t>(i) * n + k];
}

__global__ void gj_update_kernel(double* __restrict__ a, int n, int k, const double* __restrict__ row,
                                 const double* __restrict__ col) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<std::int64_t>(n) * n) return;
    const int i = static_cast<int>(t / n), j = static_cast<int>(t % n);
    const double piv = row[k];
    double out;
    if (i == k && j == k) {
        out = 1.0 / piv;
    } else if (i == k) {
        out = row[j] / piv;
    } else if (j == k) {
        out = -col[i] / piv;
    } else {
        out = a[t] - col[i] * row[j] / piv;
    }
    a[t] = out;
}

}  // namespace

// ----------------------------------------------------------------------------------------------
// Host side
// ----------------------------------------------------------------------------------------------

void l1_dinv(const CsrView& a, float* dinv, hipStream_t s) {
    if (const unsigned g = grid_for(a.n)) l1_dinv_kernel<<<g, kBlock, 0, s>>>(a.ptr, a.col, a.val, a.n, dinv);
}

static int read_int(const int* d, hipStream_t s) {
    int h = 0;
    d2h(&h, d, 1, s);
    sync(s, "counter readback");
    return h;
}

std::int64_t select_size4(const CsrView& a, std::int32_t* agg, hipStream_t s) {
    const std::int64_t n = a.n;
    if (n <= 0) return 0;
    const std::int64_t nnz = a.nnz;
    const unsigned g = grid_for(n);
    DeviceBuffer<float> diag(n, "agg diag"), w(std::max<std::int64_t>(nnz, 1), "agg w"), wsn(n, "agg wsn");
    DeviceBuffer<std::int32_t> partner(n, "partner"), strongest(n, "strongest"), aggregated(n, "aggregated"),
        cand(n, "cand"), used(n + 1, "used"), excl(n + 1, "excl");
    DeviceBuffer<AggLoop> ctl(1, "matching control");
    const int ni = static_cast<int>(n);
    agg_diag_kernel<<<g, kBlock, 0, s>>>(a.ptr, a.col, a.val, n, diag.get());
    if (nnz > 0) agg_edge_weight_kernel<<<g, kBlock, 0, s>>>(a.ptr, a.col, a.val, diag.get(), n, w.get());
    agg_init_kernel<<<g, kBlock, 0, s>>>(n, agg, partner.get(), strongest.get());
    // Round 1: match rows in pairs along mutual strongest edges, at most 16 passes.
    agg_loop_init_kernel<<<1, 1, 0, s>>>(ctl.get(), ni);
    for (int iter = 0; iter <= 15; ++iter) {
        agg_find_strongest_nomerge_kernel<<<g, kBlock, 0, s>>>(a.ptr, a.col, w.get(), n, partner.get(),
                                                               strongest.get(), ctl.get());
        agg_match_edges_kernel<<<g, kBlock, 0, s>>>(n, partner.get(), agg, strongest.get(), ctl.get());
        agg_loop_decide_kernel<<<1, 1, 0, s>>>(ctl.get(), ni, 0);
    }
    agg_round2_init_kernel<<<g, kBlock, 0, s>>>(n, partner.get(), wsn.get(), aggregated.get());
    // Round 2: match the pairs (and singletons) into aggregates of up to four, at most 16 passes.
    agg_loop_init_kernel<<<1, 1, 0, s>>>(ctl.get(), ni);
    for (int iter = 0; iter <= 15; ++iter) {
        agg_find_strongest_store_kernel<<<g, kBlock, 0, s>>>(a.ptr, a.col, w.get(), n, aggregated.get(), agg,
                                                             partner.get(), strongest.get(), wsn.get(), ctl.get());
        agg_agree_on_proposal_kernel<<<g, kBlock, 0, s>>>(n, aggregated.get(), strongest.get(), wsn.get(),
                                                          partner.get(), ctl.get());
        agg_match_aggregates_kernel<<<g, kBlock, 0, s>>>(n, agg, aggregated.get(), strongest.get(), ctl.get());
        agg_loop_decide_kernel<<<1, 1, 0, s>>>(ctl.get(), ni, 1);
    }
    // Rows still unmatched join their strongest aggregated neighbour's aggregate (both kernels
    // touch only unmatched rows, so with none left they change nothing).
    agg_merge_existing_kernel<<<g, kBlock, 0, s>>>(a.ptr, a.col, w.get(), n, agg, aggregated.get(), cand.get());
    agg_join_existing_kernel<<<g, kBlock, 0, s>>>(n, agg, aggregated.get(), cand.get());
    check(hipMemsetAsync(used.get(), 0, sizeof(std::int32_t) * (n + 1), s), "memset used");
    agg_mark_used_kernel<<<g, kBlock, 0, s>>>(n, agg, used.get());
    exclusive_scan_i32(used.get(), excl.get(), n + 1, s);
    agg_relabel_kernel<<<g, kBlock, 0, s>>>(n, agg, excl.get());
    return read_int(excl.get() + n, s);
}

void compose_aggregates(std::int32_t* agg, std::int64_t n, const std::int32_t* step, hipStream_t s) {
    if (const unsigned g = grid_for(n)) compose_kernel<<<g, kBlock, 0, s>>>(agg, n, step);
}

void agg_edge_weights(const CsrView& a, float* w, hipStream_t s) {
    DeviceBuffer<float> diag(std::max<std::int64_t>(a.n, 1), "agg diag");
    if (const unsigned g = grid_for(a.n)) {
        agg_diag_kernel<<<g, kBlock, 0, s>>>(a.ptr, a.col, a.val, a.n, diag.get());
        agg_edge_weight_kernel<<<g, kBlock, 0, s>>>(a.ptr, a.col, a.val, diag.get(), a.n, w);
    }
}

CsrF spgemm(const CsrView& a, const CsrView& b, hipStream_t s) {
    CsrF c;
    c.n = a.n;
    c.n_cols = b.n_cols;
    c.ptr.reset(static_cast<std::size_t>(a.n + 1), "spgemm ptr");
    if (a.n == 0) {
        check(hipMemsetAsync(c.ptr.get(), 0, sizeof(std::int32_t), s), "memset");
        return c;
    }
    DeviceBuffer<std::int32_t> cand(static_cast<std::size_t>(a.n), "spgemm candidates");
    candidates_kernel<<<grid_for(a.n), kBlock, 0, s>>>(a.ptr, a.col, b.ptr, a.n, cand.get());
    // Counting table: keys only, sized from the candidate count (duplicates included) but capped
    // at 64 KB; a row whose distinct columns would not fit raises the overflow flag.
    constexpr int kMaxCountCap = 16384;
    constexpr int kMaxFillCap = 4096;  // 16 B per slot: 64 KB
    const std::int64_t widest =
        std::max<std::int64_t>(1, std::min<std::int64_t>(max_i32(cand.get(), a.n, s), b.n_cols));
    int cap = 64;
    while (cap < 2 * widest && cap < kMaxCountCap) cap <<= 1;
    DeviceBuffer<std::int32_t> counts(static_cast<std::size_t>(a.n + 1), "spgemm counts");
    DeviceBuffer<int> overflow(1, "spgemm overflow");
    check(hipMemsetAsync(counts.get() + a.n, 0, sizeof(std::int32_t), s), "memset");
    check(hipMemsetAsync(overflow.get(), 0, sizeof(int), s), "memset");
    spgemm_count_kernel<<<static_cast<unsigned>(a.n), kWave, static_cast<std::size_t>(cap) * 4, s>>>(
        a.ptr, a.col, b.ptr, b.col, a.n, cap, counts.get(), overflow.get());
    check(hipGetLastError(), "spgemm count launch");
    int hov = 0;
    d2h(&hov, overflow.get(), 1, s);
    sync(s, "spgemm count");
    require(hov == 0, "sparse product: a row has more distinct columns than the LDS table holds");
    const std::int64_t widest_u = std::max<std::int64_t>(1, max_i32(counts.get(), a.n, s));
    cap = 64;
    while (cap < 2 * widest_u) cap <<= 1;
    require(cap <= kMaxFillCap, "sparse product: a row has more distinct columns than the LDS table holds");
    const std::size_t lds = static_cast<std::size_t>(cap) * 16;  // keys, accumulators, compacted pairs
    exclusive_scan_i32(counts.get(), c.ptr.get(), a.n + 1, s);
    std::int32_t nnz = 0;
    d2h(&nnz, c.ptr.get() + a.n, 1, s);
    sync(s, "spgemm nnz");
    c.nnz = nnz;
    c.col.reset(static_cast<std::size_t>(std::max(nnz, 1)), "spgemm col");
    c.val.reset(static_cast<std::size_t>(std::max(nnz, 1)), "spgemm val");
    spgemm_fill_kernel<<<static_cast<unsigned>(a.n), kWave, lds, s>>>(a.ptr, a.col, a.val, b.ptr, b.col, b.val,
                                                                      a.n, cap, c.ptr.get(), c.col.get(),
                                                                      c.val.get());
    check(hipGetLastError(), "spgemm fill launch");
    return c;
}

CsrF transpose(const CsrView& a, hipStream_t s) {
    CsrF t;
    t.n = a.n_cols;
    t.n_cols = a.n;
    t.nnz = a.nnz;
    t.ptr.reset(static_cast<std::size_t>(t.n + 1), "transpose ptr");
    DeviceBuffer<std::int32_t> counts(static_cast<std::size_t>(t.n + 1), "transpose counts");
    check(hipMemsetAsync(counts.get(), 0, sizeof(std::int32_t) * (t.n + 1), s), "memset");
    t.col.reset(static_cast<std::size_t>(std::max<std::int64_t>(a.nnz, 1)), "transpose col");
    t.val.reset(static_cast<std::size_t>(std::max<std::int64_t>(a.nnz, 1)), "transpose val");
    if (a.nnz > 0) {
        agg_counts_kernel<<<grid_for(a.nnz), kBlock, 0, s>>>(a.col, a.nnz, counts.get());
        DeviceBuffer<std::int32_t> rows(static_cast<std::size_t>(a.nnz), "row ids");
        DeviceBuffer<std::int32_t> iota(static_cast<std::size_t>(a.nnz), "iota");
        DeviceBuffer<std::int32_t> order(static_cast<std::size_t>(a.nnz), "order");
        DeviceBuffer<std::uint32_t> ksorted(static_cast<std::size_t>(a.nnz), "keys");
        expand_rows_kernel<<<grid_for(a.n), kBlock, 0, s>>>(a.ptr, a.n, rows.get());
        iota_i32(iota.get(), a.nnz, s);
        int bits = 1;
        while (bits < 32 && (1ll << bits) < a.n_cols) ++bits;
        sort_pairs_u32(reinterpret_cast<const std::uint32_t*>(a.col), ksorted.get(), iota.get(), order.get(),
                       a.nnz, bits, s);
        gather_transpose_kernel<<<grid_for(a.nnz), kBlock, 0, s>>>(order.get(), rows.get(), a.val, a.nnz,
                                                                   t.col.get(), t.val.get());
    }
    exclusive_scan_i32(counts.get(), t.ptr.get(), t.n + 1, s);
    return t;
}

CsrF tentative_prolongator(const std::int32_t* agg, std::int64_t n, std::int64_t n_coarse, hipStream_t s) {
    CsrF p;
    p.n = n;
    p.n_cols = n_coarse;
    p.nnz = n;
    p.ptr.reset(static_cast<std::size_t>(n + 1), "tent ptr");
    p.col.reset(static_cast<std::size_t>(std::max<std::int64_t>(n, 1)), "tent col");
    p.val.reset(static_cast<std::size_t>(std::max<std::int64_t>(n, 1)), "tent val");
    DeviceBuffer<std::int32_t> counts(static_cast<std::size_t>(n_coarse), "aggregate sizes");
    check(hipMemsetAsync(counts.get(), 0, sizeof(std::int32_t) * n_coarse, s), "memset");
    if (n > 0) agg_counts_kernel<<<grid_for(n), kBlock, 0, s>>>(agg, n, counts.get());
    tentative_kernel<<<grid_for(n + 1), kBlock, 0, s>>>(agg, counts.get(), n, p.ptr.get(), p.col.get(), p.val.get());
    sync(s, "tentative prolongator");
    return p;
}

CsrF smoothed_prolongator(const CsrView& a, const std::int32_t* agg, std::int64_t n_coarse, const float* dinv,
                          hipStream_t s) {
    CsrF tent = tentative_prolongator(agg, a.n, n_coarse, s);
    CsrF ap = spgemm(a, tent.view(), s);
    if (const unsigned g = grid_for(a.n))
        smooth_prolongator_kernel<<<g, kBlock, 0, s>>>(ap.ptr.get(), ap.col.get(), ap.val.get(), dinv, agg,
                                                       tent.val.get(), a.n);
    sync(s, "smoothed prolongator");
    return ap;
}

CsrF galerkin(const CsrView& a, const CsrView& p, hipStream_t s) {
    // P^T (A P): A P rows stay narrow, and P^T (A P) rows are at most as wide as the coarse
    // operator's, so every product fits the LDS hash tables.
    CsrF r = transpose(p, s);
    CsrF ap = spgemm(a, p, s);
    CsrF c = spgemm(r.view(), ap.view(), s);
    sync(s, "galerkin");
    return c;
}

void coarse_inverse(const CsrView& a, DeviceBuffer<double>& inv, hipStream_t s) {
    const int n = static_cast<int>(a.n);
    inv.reset(static_cast<std::size_t>(n) * n, "coarse inverse");
    csr_to_dense_kernel<<<grid_for(n), kBlock, 0, s>>>(a.ptr, a.col, a.val, n, inv.get());
    DeviceBuffer<double> row(n, "gj row"), col(n, "gj col");
    const std::int64_t nn = static_cast<std::int64_t>(n) * n;
    for (int k = 0; k < n; ++k) {
        gj_save_kernel<<<grid_for(n), kBlock, 0, s>>>(inv.get(), n, k, row.get(), col.get());
        gj_update_kernel<<<grid_for(nn), kBlock, 0, s>>>(inv.get(), n, k, row.get(), col.get());
    }
    sync(s, "coarse inverse");
}

std::pair<DeviceBuffer<std::int32_t>, std::int64_t> aggregate_level(const CsrView& a, const AmgParams& prm,
                                                                     hipStream_t s) {
    // fem/solve.py:178-209: compose `rounds` pairwise passes into one level's map.
    DeviceBuffer<std::int32_t> agg;
    std::int64_t n_coarse = a.n;
    CsrF cur_owned;
    CsrView cur = a;
    for (int round = 0; round < prm.rounds; ++round) {
        if (cur.n <= prm.min_coarse_rows) break;
        DeviceBuffer<std::int32_t> step(static_cast<std::size_t>(cur.n), "aggregates");
        const std::int64_t nc = select_size4(cur, step.get(), s);
        if (nc == cur.n || nc < prm.min_coarse_rows) break;
        if (round + 1 < prm.rounds) {
            // The intermediate operator is built from this pass's own map (fem/solve.py:208).
            CsrF tent = tentative_prolongator(step.get(), cur.n, nc, s);
            CsrF next = galerkin(cur, tent.view(), s);
            cur_owned = std::move(next);
        }
        if (!agg) {
            agg = std::move(step);
        } else {
            compose_aggregates(agg.get(), a.n, step.get(), s);  // agg = step[agg]
        }
        n_coarse = nc;
        if (round + 1 < prm.rounds) cur = cur_owned.view();
    }
    sync(s, "aggregate level");
    return {std::move(agg), n_coarse};
}

}  // namespace cn

namespace cn {

Cheby chebyshev(int degree, float lower_ratio) {
    require(degree >= 1 && degree <= kMaxDegree, "smoother degree must be in [1, 6]");
    require(lower_ratio > 1.0f, "smoother lower_ratio must exceed 1");
    Cheby c;
    c.degree = degree;
    const float hi = 1.0f;
    const float lo = hi / lower_ratio;
    const float theta = 0.5f * (hi + lo);
    const float delta = 0.5f * (hi - lo);
    const float sigma = theta / delta;
    c.alpha0 = 1.0f / theta;
    float rho = 1.0f / sigma;
    for (int j = 0; j + 1 < degree; ++j) {
        const float rho_next = 1.0f / (2.0f * sigma - rho);
        const float beta = rho_next * rho;
        c.c_cur[j] = 1.0f + beta;
        c.c_prev[j] = -beta;
        c.alpha[j] = 2.0f * rho_next / delta;
        rho = rho_next;
    }
    return c;
}

// fem/solve.py:224-256 (aggregation_levels) and :259-284 (build_native_vcycle).
Hierarchy build_hierarchy_like(const CsrView& a0, const Hierarchy& like, const AmgParams& prm, hipStream_t s) {
    if (like.levels.empty()) {  // the operator itself is the coarsest level
        AmgParams direct = prm;
        direct.max_levels = 0;
        Hierarchy h = build_hierarchy(a0, direct, s);
        h.cheb = like.cheb;
        return h;
    }
    Hierarchy h;
    h.cheb = like.cheb;
    CsrView a = a0;
    CsrF a_next;
    for (std::size_t lv = 0; lv < like.levels.size(); ++lv) {
        const Level& ref = like.levels[lv];
        require(a.n == ref.a.n && a.nnz == ref.a.nnz, "the operator's pattern differs from the hierarchy's");
        Level L;
        if (lv == 0) {
            L.a = a;
        } else {
            L.a_store = std::move(a_next);
            L.a = L.a_store.view();
        }
        L.dinv.reset(static_cast<std::size_t>(L.a.n), "dinv");
        l1_dinv(L.a, L.dinv.get(), s);
        L.p = smoothed_prolongator(L.a, ref.agg.get(), ref.n_coarse, L.dinv.get(), s);
        L.r = transpose(L.p.view(), s);
        L.n_coarse = ref.n_coarse;
        CsrF c = spgemm(L.r.view(), spgemm(L.a, L.p.view(), s).view(), s);
        sync(s, "galerkin");
        h.levels.push_back(std::move(L));
        a_next = std::move(c);
        a = a_next.view();
    }
    h.coarse = std::move(a_next);
    coarse_inverse(h.coarse.view(), h.coarse_inv, s);
    return h;
}

Hierarchy build_hierarchy(const CsrView& a0, const AmgParams& prm, hipStream_t s) {
    Hierarchy h;
    h.cheb = chebyshev(prm.degree, prm.lower_ratio);
    CsrView a = a0;
    CsrF a_next;
    for (int lv = 0; lv < prm.max_levels; ++lv) {
        if (a.n <= prm.min_coarse_rows) break;
        auto [agg, nc] = aggregate_level(a, prm, s);
        if (!agg || nc == a.n || nc < prm.min_coarse_rows) break;
        Level L;
        if (lv == 0) {
            L.a = a;
        } else {
            L.a_store = std::move(a_next);
            L.a = L.a_store.view();
        }
        L.dinv.reset(static_cast<std::size_t>(L.a.n), "dinv");
        l1_dinv(L.a, L.dinv.get(), s);
        L.p = smoothed_prolongator(L.a, agg.get(), nc, L.dinv.get(), s);
        L.r = transpose(L.p.view(), s);
        L.agg = std::move(agg);
        L.n_coarse = nc;
        CsrF c = spgemm(L.r.view(), spgemm(L.a, L.p.view(), s).view(), s);
        sync(s, "galerkin");
        h.levels.push_back(std::move(L));
        a_next = std::move(c);
        a = a_next.view();
    }
    if (h.levels.empty()) {
        // The fine operator itself is the coarsest: copy it so the hierarchy owns what it solves.
        CsrF c;
        c.n = a.n;
        c.n_cols = a.n_cols;
        c.nnz = a.nnz;
        c.ptr.reset(static_cast<std::size_t>(a.n + 1), "coarse ptr");
        c.col.reset(static_cast<std::size_t>(std::max<std::int64_t>(a.nnz, 1)), "coarse col");
        c.val.reset(static_cast<std::size_t>(std::max<std::int64_t>(a.nnz, 1)), "coarse val");
        d2d(c.ptr.get(), a.ptr, static_cast<std::size_t>(a.n + 1), s);
        d2d(c.col.get(), a.col, static_cast<std::size_t>(a.nnz), s);
        d2d(c.val.get(), a.val, static_cast<std::size_t>(a.nnz), s);
        h.coarse = std::move(c);
    } else {
        h.coarse = std::move(a_next);
    }
    require(h.coarse.n <= 8192, "coarsest level too large for the dense solve");
    coarse_inverse(h.coarse.view(), h.coarse_inv, s);
    return h;
}

}  // namespace cn
