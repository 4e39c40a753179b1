/*
 * CPU oracle for the TMS E-field FEM path: a plain-C restatement of the reference algorithm.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load this library, and only as the checker (or as the timed CPU baseline); the product path
 * never calls it.
 *
 * Precision follows the reference where the reference fixes it and is FP64 elsewhere:
 *   - P1 gradients, stiffness, placement frames, skin normals: FP64 (as the reference).
 *   - Aggregation, l1 diagonal, prolongators and Galerkin products: FP32 values (as the reference
 *     builds its hierarchy from the FP32 copy of the reduced stiffness).
 *   - dA/dt, RHS, reconstruction, V-cycle and PCG: FP64, so a solve here to 1e-12 is the FP64
 *     oracle every GPU result is measured against.
 *
 * Arithmetic that the HIP kernels must reproduce bit for bit (gradients, stiffness, l1 diagonal,
 * edge weights, sparse products) spells its fused multiply-adds out with fma()/fmaf(); the file
 * is compiled with -ffp-contract=off so the compiler adds none of its own.
 *
 * Row loops that are independent per row carry OpenMP pragmas so the same code serves as the
 * multi-core CPU baseline; every reduction is either per row (fixed order) or a fixed-partition
 * sum, so results do not depend on the thread count.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_API __attribute__((visibility("default")))

/* ---------------------------------------------------------------------------------------------
 * Morton key. fem/solve.py:299-323 (morton_order, _spread3): quantise to 21 bits per axis over the
 * bounding box, interleave x, y, z into a 63-bit key.
 * ------------------------------------------------------------------------------------------- */
static uint64_t spread3(uint64_t v) {
    v &= 0x1FFFFFull;
    v = (v | (v << 32)) & 0x1F00000000FFFFull;
    v = (v | (v << 16)) & 0x1F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    return (v | (v << 2)) & 0x1249249249249249ull;
}

OR_API void or_morton_keys(const double* pts, int64_t n, uint64_t* keys) {
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            const double v = pts[3 * i + a];
            if (v < lo[a]) lo[a] = v;
            if (v > hi[a]) hi[a] = v;
        }
    double span[3];
    for (int a = 0; a < 3; ++a) {
        span[a] = hi[a] - lo[a];
        if (!(span[a] > 1e-12)) span[a] = 1e-12;
    }
    const double scale = (double)((1u << 21) - 1u);
    for (int64_t i = 0; i < n; ++i) {
        uint64_t key = 0;
        for (int a = 0; a < 3; ++a) {
            const double q = (pts[3 * i + a] - lo[a]) / span[a] * scale;
            key |= spread3((uint64_t)q) << a;
        }
        keys[i] = key;
    }
}

/* ---------------------------------------------------------------------------------------------
 * P1 gradients and volumes. solver/fem/gradient.cu:29-66: G = (T^-1 A)^T through edge cross
 * products over the determinant, x1e3 (mm -> 1/m), vol = |det| * 1e-9 / 6.
 * ------------------------------------------------------------------------------------------- */
static void cross3(const double* a, const double* b, double* o) {
    o[0] = fma(a[1], b[2], -(a[2] * b[1]));
    o[1] = fma(a[2], b[0], -(a[0] * b[2]));
    o[2] = fma(a[0], b[1], -(a[1] * b[0]));
}

OR_API void or_p1_gradients(const double* nodes, const int32_t* tets, int64_t m, double* g,
                            double* vols) {
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m; ++e) {
        const int32_t* t4 = tets + 4 * e;
        double t[3][3];
        for (int i = 0; i < 3; ++i)
            for (int c = 0; c < 3; ++c) t[i][c] = nodes[3 * (int64_t)t4[i + 1] + c] - nodes[3 * (int64_t)t4[0] + c];
        double adj[3][3];
        cross3(t[1], t[2], adj[0]);
        cross3(t[2], t[0], adj[1]);
        cross3(t[0], t[1], adj[2]);
        const double det = fma(t[0][2], adj[0][2], fma(t[0][1], adj[0][1], t[0][0] * adj[0][0]));
        const double inv = 1e3 / det;
        for (int c = 0; c < 3; ++c) {
            const double g1 = adj[0][c] * inv, g2 = adj[1][c] * inv, g3 = adj[2][c] * inv;
            g[12 * e + 0 + c] = -((g1 + g2) + g3);
            g[12 * e + 3 + c] = g1;
            g[12 * e + 6 + c] = g2;
            g[12 * e + 9 + c] = g3;
        }
        vols[e] = fabs(det) * (1e-9 / 6.0);
    }
}

/* ---------------------------------------------------------------------------------------------
 * Stiffness CSR pattern. solver/fem/pattern.cu:65-87,144-214: row i = sorted distinct nodes of
 * the tetrahedra incident to node i (node2corner segment), duplicates dropped.
 * ------------------------------------------------------------------------------------------- */
static int cmp_i32(const void* a, const void* b) {
    const int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

static int row_unique(const int32_t* tets, const int32_t* ptr, const int32_t* idx, int64_t row,
                      int32_t* buf) {
    const int32_t b = ptr[row], e = ptr[row + 1];
    int n = 0;
    for (int32_t p = b; p < e; ++p) {
        const int64_t t = idx[p] >> 2;
        for (int j = 0; j < 4; ++j) buf[n++] = tets[4 * t + j];
    }
    qsort(buf, (size_t)n, sizeof(int32_t), cmp_i32);
    int k = 0;
    for (int i = 0; i < n; ++i)
        if (i == 0 || buf[i] != buf[i - 1]) buf[k++] = buf[i];
    return k;
}

/* indptr (n+1) out; returns nnz. */
OR_API int64_t or_pattern_count(const int32_t* tets, const int32_t* ptr, const int32_t* idx,
                                int64_t n, int32_t* indptr) {
    int32_t widest = 0;
    for (int64_t i = 0; i < n; ++i)
        if (ptr[i + 1] - ptr[i] > widest) widest = ptr[i + 1] - ptr[i];
    int32_t* buf = (int32_t*)malloc(sizeof(int32_t) * (size_t)(4 * widest + 4));
    int64_t total = 0;
    indptr[0] = 0;
    for (int64_t i = 0; i < n; ++i) {
        total += row_unique(tets, ptr, idx, i, buf);
        indptr[i + 1] = (int32_t)total;
    }
    free(buf);
    return total;
}

OR_API void or_pattern_fill(const int32_t* tets, const int32_t* ptr, const int32_t* idx, int64_t n,
                            const int32_t* indptr, int32_t* indices) {
    int32_t widest = 0;
    for (int64_t i = 0; i < n; ++i)
        if (ptr[i + 1] - ptr[i] > widest) widest = ptr[i + 1] - ptr[i];
    int32_t* buf = (int32_t*)malloc(sizeof(int32_t) * (size_t)(4 * widest + 4));
    for (int64_t i = 0; i < n; ++i) {
        const int k = row_unique(tets, ptr, idx, i, buf);
        memcpy(indices + indptr[i], buf, sizeof(int32_t) * (size_t)k);
    }
    free(buf);
}

/* ---------------------------------------------------------------------------------------------
 * Stiffness values. solver/fem/stiffness.cu:32-66: one row at a time in node2corner order, j_loc
 * ascending within an element; A[i, col] += s_e * (g_i . g_j), the column found by bisection.
 * ------------------------------------------------------------------------------------------- */
static int32_t find_col(const int32_t* indices, int32_t lo, int32_t hi, int32_t col) {
    while (lo < hi) {
        const int32_t mid = lo + ((hi - lo) >> 1);
        if (indices[mid] < col)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

OR_API void or_stiffness_values(const double* g, const double* scale, const int32_t* tets,
                                const int32_t* ptr, const int32_t* idx, int64_t n,
                                const int32_t* indptr, const int32_t* indices, double* data) {
#pragma omp parallel for schedule(static)
    for (int64_t row = 0; row < n; ++row) {
        const int32_t rb = indptr[row], re = indptr[row + 1];
        for (int32_t p = rb; p < re; ++p) data[p] = 0.0;
        for (int32_t p = ptr[row]; p < ptr[row + 1]; ++p) {
            const int64_t corner = idx[p];
            const int64_t e = corner >> 2;
            const double s = scale[e];
            const double* gi = g + 3 * corner;
            for (int j = 0; j < 4; ++j) {
                const double* gj = g + 3 * (4 * e + j);
                const double dot = fma(gi[2], gj[2], fma(gi[1], gj[1], gi[0] * gj[0]));
                const int32_t at = find_col(indices, rb, re, tets[4 * e + j]);
                data[at] = fma(s, dot, data[at]);
            }
        }
    }
}

/* ---------------------------------------------------------------------------------------------
 * l1-Jacobi diagonal. solver/amg/l1.cu:15-35: d_i = sign(a_ii) * sum_j |a_ij| (row order),
 * dinv = 1/d, 1 when the sum is zero. FP32 as the reference.
 * ------------------------------------------------------------------------------------------- */
OR_API void or_l1_dinv(int64_t n, const int32_t* indptr, const int32_t* indices, const float* data,
                       float* dinv) {
    for (int64_t row = 0; row < n; ++row) {
        float acc = 0.f, diag = 0.f;
        for (int32_t p = indptr[row]; p < indptr[row + 1]; ++p) {
            acc += fabsf(data[p]);
            if (indices[p] == row) diag = data[p];
        }
        if (diag < 0.f) acc = -acc;
        if (acc == 0.f) acc = 1.f;
        dinv[row] = 1.f / acc;
    }
}

/* ---------------------------------------------------------------------------------------------
 * SIZE_4 pairwise aggregation, one pass. solver/amg/aggregate.cu:20-278 (kernels) and :368-472
 * (driver). Each reference kernel is one loop over the rows here. The kernels read only state
 * that the same kernel does not write at another row (or, for agree_on_proposal, never a slot its
 * owner writes in the same kernel), so a sequential sweep computes what the parallel launch does.
 * ------------------------------------------------------------------------------------------- */
static uint32_t agg_hash(uint32_t a, uint32_t seed) {
    a ^= seed;
    a = (a + 0x7ed55d16u) + (a << 12);
    a = (a ^ 0xc761c23cu) + (a >> 19);
    a = (a + 0x165667b1u) + (a << 5);
    a = (a ^ 0xd3a2646cu) + (a << 9);
    a = (a + 0xfd7046c5u) + (a << 3);
    a = (a ^ 0xb55a4f09u) + (a >> 16);
    return a;
}

static int is_candidate(int32_t col, int64_t row, int64_t n) {
    return col != row && (uint32_t)col < (uint32_t)n;
}

typedef struct {
    float w;
    int32_t col;
} best_t;

static void best_take(best_t* acc, float w, int32_t col) {
    if (w > acc->w || (w == acc->w && col > acc->col)) {
        acc->w = w;
        acc->col = col;
    }
}

/* Edge weights, aggregate.cu:114-134 (with :91-106 for the diagonal). */
OR_API void or_agg_edge_weights(int64_t n, const int32_t* rp, const int32_t* ci, const float* v,
                                float* w) {
    float* diag = (float*)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    for (int64_t i = 0; i < n; ++i) {
        float d = 0.f;
        for (int32_t j = rp[i]; j < rp[i + 1]; ++j)
            if (ci[j] == i) {
                d = v[j];
                break;
            }
        diag[i] = d;
    }
    for (int64_t i = 0; i < n; ++i) {
        const float di = fabsf(diag[i]);
        for (int32_t j = rp[i]; j < rp[i + 1]; ++j) {
            const int32_t jc = ci[j];
            if (!is_candidate(jc, i, n)) {
                w[j] = -1.0f;
                continue;
            }
            const float den = fmaxf(di, fabsf(diag[jc]));
            const float ed = fabsf(v[j]) / den;
            const uint32_t lo = (uint32_t)(i < jc ? i : jc), hi = (uint32_t)(i < jc ? jc : i);
            const float frac = 1e-5f * (float)agg_hash(lo, hi) / 4294967295.0f;
            w[j] = fmaf(frac, ed, ed);
        }
    }
    free(diag);
}

OR_API int64_t or_select_size4(int64_t n, const int32_t* rp, const int32_t* ci, const float* v,
                               int32_t* agg) {
    if (n <= 0) return 0;
    const int64_t nnz = rp[n];
    float* w = (float*)malloc(sizeof(float) * (size_t)(nnz > 0 ? nnz : 1));
    float* wsn = (float*)malloc(sizeof(float) * (size_t)n);
    int32_t* partner = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* strongest = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* aggregated = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* cand = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int32_t* used = (int32_t*)calloc((size_t)n + 1, sizeof(int32_t));
    or_agg_edge_weights(n, rp, ci, v, w);
    for (int64_t i = 0; i < n; ++i) {
        agg[i] = (int32_t)i;
        partner[i] = -1;
        strongest[i] = -1;
    }
    /* Round 1: handshake pairs. */
    int64_t unassigned = n;
    for (int iter = 0; iter <= 15; ++iter) {
        for (int64_t t = 0; t < n; ++t) { /* agg_find_strongest_nomerge_kernel */
            if (partner[t] != -1) continue;
            best_t b = {0.0f, -1};
            for (int32_t j = rp[t]; j < rp[t + 1]; ++j) {
                const int32_t jc = ci[j];
                if (!is_candidate(jc, t, n) || partner[jc] != -1) continue;
                best_take(&b, w[j], jc);
            }
            if (b.col != -1) strongest[t] = b.col;
        }
        int64_t count = 0;
        for (int64_t t = 0; t < n; ++t) { /* agg_match_edges_kernel */
            if (partner[t] != -1) continue;
            const int32_t pm = strongest[t];
            if (pm != -1 && strongest[pm] == t) {
                partner[t] = pm;
                agg[t] = pm < t ? pm : (int32_t)t;
            } else {
                ++count;
            }
        }
        const int64_t previous = unassigned;
        unassigned = count;
        if (unassigned == 0 || unassigned == previous) break;
    }
    for (int64_t t = 0; t < n; ++t)
        if (partner[t] == -1) partner[t] = (int32_t)t;
    for (int64_t t = 0; t < n; ++t) {
        wsn[t] = -1.0f;
        aggregated[t] = -1;
    }
    /* Round 2: pairs into quads. */
    unassigned = n;
    for (int iter = 0; iter <= 15; ++iter) {
        for (int64_t t = 0; t < n; ++t) { /* agg_find_strongest_store_kernel */
            if (aggregated[t] != -1) continue;
            const int32_t p = partner[t];
            best_t b = {0.0f, -1};
            for (int32_t j = rp[t]; j < rp[t + 1]; ++j) {
                const int32_t jc = ci[j];
                if (!is_candidate(jc, t, n)) continue;
                if (aggregated[jc] != -1 || jc == p) continue;
                best_take(&b, w[j], jc);
            }
            if (b.col != -1) {
                wsn[t] = b.w;
                strongest[t] = agg[b.col];
            }
        }
        for (int64_t t = 0; t < n; ++t) { /* agg_agree_on_proposal_kernel */
            if (aggregated[t] != -1) continue;
            const int32_t p = partner[t];
            const float mine = wsn[t];
            const float theirs = (p != -1) ? wsn[p] : -1.0f;
            if (mine < 0.0f && theirs < 0.0f) {
                aggregated[t] = 1;
                strongest[t] = -1;
            } else if (mine < theirs) {
                strongest[t] = strongest[p];
            }
        }
        int64_t count = 0;
        for (int64_t t = 0; t < n; ++t) { /* agg_match_aggregates_kernel */
            if (aggregated[t] != -1) continue;
            const int32_t pm = strongest[t];
            const int32_t mine = agg[t];
            if (pm != -1 && strongest[pm] == mine) {
                aggregated[t] = 1;
                agg[t] = pm < mine ? pm : mine;
            } else {
                ++count;
            }
        }
        const int64_t previous = unassigned;
        unassigned = count;
        if (unassigned == 0 || unassigned == previous ||
            (float)unassigned / (float)n < 0.05f)
            break;
    }
    if (unassigned != 0) {
        for (int64_t t = 0; t < n; ++t) { /* agg_merge_existing_kernel */
            if (aggregated[t] != -1) continue;
            best_t b = {0.0f, -1};
            for (int32_t j = rp[t]; j < rp[t + 1]; ++j) {
                const int32_t jc = ci[j];
                if (!is_candidate(jc, t, n) || aggregated[jc] == -1) continue;
                best_take(&b, w[j], jc);
            }
            cand[t] = (b.col != -1) ? agg[b.col] : (int32_t)t;
        }
        for (int64_t t = 0; t < n; ++t) /* agg_join_existing_kernel */
            if (aggregated[t] == -1) {
                agg[t] = cand[t];
                aggregated[t] = 1;
            }
    }
    /* Order-preserving relabel: mark, exclusive scan, gather. */
    for (int64_t t = 0; t < n; ++t) used[agg[t]] = 1;
    int32_t* excl = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n + 1));
    int32_t run = 0;
    for (int64_t t = 0; t <= n; ++t) {
        excl[t] = run;
        run += used[t];
    }
    for (int64_t t = 0; t < n; ++t) agg[t] = excl[agg[t]];
    const int64_t n_coarse = excl[n];
    free(w);
    free(wsn);
    free(partner);
    free(strongest);
    free(aggregated);
    free(cand);
    free(used);
    free(excl);
    return n_coarse;
}

/* ---------------------------------------------------------------------------------------------
 * Sparse product C = A B in FP32 with a fixed association: for each output entry, the products
 * a_ij * b_jk are folded with fmaf in the order j appears in row i of A, starting from zero. The
 * output rows are sorted by column. The reference forms these products with cupyx SpGEMM
 * (fem/solve.py:119-133) whose association is unspecified; this order is the one the HIP SpGEMM
 * uses, so the two agree bit for bit.
 * ------------------------------------------------------------------------------------------- */
OR_API int64_t or_spgemm_count(int64_t n_rows, int64_t n_cols_b, const int32_t* ap,
                               const int32_t* ac, const int32_t* bp, const int32_t* bc,
                               int32_t* cp) {
    int32_t* mark = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n_cols_b > 0 ? n_cols_b : 1));
    for (int64_t k = 0; k < n_cols_b; ++k) mark[k] = -1;
    int64_t total = 0;
    cp[0] = 0;
    for (int64_t i = 0; i < n_rows; ++i) {
        for (int32_t p = ap[i]; p < ap[i + 1]; ++p) {
            const int32_t j = ac[p];
            for (int32_t q = bp[j]; q < bp[j + 1]; ++q) {
                const int32_t k = bc[q];
                if (mark[k] != (int32_t)i) {
                    mark[k] = (int32_t)i;
                    ++total;
                }
            }
        }
        cp[i + 1] = (int32_t)total;
    }
    free(mark);
    return total;
}

OR_API void or_spgemm_fill(int64_t n_rows, int64_t n_cols_b, const int32_t* ap, const int32_t* ac,
                           const float* av, const int32_t* bp, const int32_t* bc, const float* bv,
                           const int32_t* cp, int32_t* cc, float* cv) {
    int32_t* slot = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n_cols_b > 0 ? n_cols_b : 1));
    for (int64_t k = 0; k < n_cols_b; ++k) slot[k] = -1;
    for (int64_t i = 0; i < n_rows; ++i) {
        const int32_t base = cp[i];
        int32_t len = 0;
        /* Columns first, so the row can be sorted before any value lands. */
        for (int32_t p = ap[i]; p < ap[i + 1]; ++p) {
            const int32_t j = ac[p];
            for (int32_t q = bp[j]; q < bp[j + 1]; ++q) {
                const int32_t k = bc[q];
                if (slot[k] == -1) {
                    slot[k] = 0;
                    cc[base + len++] = k;
                }
            }
        }
        qsort(cc + base, (size_t)len, sizeof(int32_t), cmp_i32);
        for (int32_t t = 0; t < len; ++t) {
            slot[cc[base + t]] = base + t;
            cv[base + t] = 0.0f;
        }
        for (int32_t p = ap[i]; p < ap[i + 1]; ++p) {
            const int32_t j = ac[p];
            const float a = av[p];
            for (int32_t q = bp[j]; q < bp[j + 1]; ++q) {
                const int32_t at = slot[bc[q]];
                cv[at] = fmaf(a, bv[q], cv[at]);
            }
        }
        for (int32_t t = 0; t < len; ++t) slot[cc[base + t]] = -1;
    }
    free(slot);
}

/* ---------------------------------------------------------------------------------------------
 * V-cycle and PCG in FP64 over an FP32 hierarchy. solver/amg/vcycle.cu:560-648 (run_cycle),
 * :241-275 (Chebyshev step), :124-130 (zero-guess Jacobi), :132-149 (residual), :166-219
 * (restrict / prolongate); solver/amg/solver.cpp:200-318 (PCG). The coarsest level is solved with
 * an FP64 Cholesky factor instead of the reference's FP32 dense inverse (fem/solve.py:281).
 * ------------------------------------------------------------------------------------------- */
typedef struct {
    int32_t n_levels; /* smoothed levels; the coarsest operator is separate */
    const int32_t* n; /* rows per smoothed level */
    const int32_t* const* a_ptr;
    const int32_t* const* a_col;
    const float* const* a_val;
    const float* const* dinv;
    const int32_t* const* p_ptr;
    const int32_t* const* p_col;
    const float* const* p_val;
    const int32_t* const* r_ptr;
    const int32_t* const* r_col;
    const float* const* r_val;
    int32_t n_coarse;
    const double* coarse_l; /* row-major lower Cholesky factor of the coarsest operator */
    int32_t degree;
    double alpha0;
    const double* c_cur;
    const double* c_prev;
    const double* alpha;
} or_hier;

static void spmv_f(int64_t n, const int32_t* rp, const int32_t* ci, const float* v, const double* x,
                   double* y) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int32_t j = rp[i]; j < rp[i + 1]; ++j) s += (double)v[j] * x[ci[j]];
        y[i] = s;
    }
}

/* x_out = c_cur x_cur + c_prev x_prev + alpha dinv (b - A x_cur); x_out may alias x_prev. */
static void cheby_step(int64_t n, const int32_t* rp, const int32_t* ci, const float* v,
                       const float* dinv, const double* b, const double* xc, const double* xp,
                       double c_cur, double c_prev, double alpha, double* xo, double* tmp) {
    spmv_f(n, rp, ci, v, xc, tmp);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i)
        xo[i] = c_cur * xc[i] + c_prev * xp[i] + alpha * (double)dinv[i] * (b[i] - tmp[i]);
}

static void coarse_solve(int32_t n, const double* l, const double* b, double* x) {
    for (int32_t i = 0; i < n; ++i) {
        double s = b[i];
        for (int32_t j = 0; j < i; ++j) s -= l[(int64_t)i * n + j] * x[j];
        x[i] = s / l[(int64_t)i * n + i];
    }
    for (int32_t i = n - 1; i >= 0; --i) {
        double s = x[i];
        for (int32_t j = i + 1; j < n; ++j) s -= l[(int64_t)j * n + i] * x[j];
        x[i] = s / l[(int64_t)i * n + i];
    }
}

/* One Chebyshev sweep from x0 (NULL = zero guess) into `out`. */
static void smooth(const or_hier* h, int lv, const double* b, const double* x0, double* out,
                   double* w1, double* w2, double* tmp) {
    const int64_t n = h->n[lv];
    const int32_t *rp = h->a_ptr[lv], *ci = h->a_col[lv];
    const float *v = h->a_val[lv], *dinv = h->dinv[lv];
    double* cur = w1;
    double* prev = w2;
    if (x0 == NULL) {
        for (int64_t i = 0; i < n; ++i) cur[i] = h->alpha0 * (double)dinv[i] * b[i];
        for (int64_t i = 0; i < n; ++i) prev[i] = 0.0;
    } else {
        cheby_step(n, rp, ci, v, dinv, b, x0, x0, 1.0, 0.0, h->alpha0, cur, tmp);
        memcpy(prev, x0, sizeof(double) * (size_t)n);
    }
    for (int s = 1; s < h->degree; ++s) {
        /* out into prev (its row-local read precedes the write) */
        cheby_step(n, rp, ci, v, dinv, b, cur, prev, h->c_cur[s - 1], h->c_prev[s - 1],
                   h->alpha[s - 1], prev, tmp);
        double* t = cur;
        cur = prev;
        prev = t;
    }
    memcpy(out, cur, sizeof(double) * (size_t)n);
}

/* work: at least 6 * n_fine + 2 * n_coarse doubles, laid out per level by the caller-agnostic
 * offsets below. */
OR_API void or_vcycle_apply(const or_hier* h, const double* b, double* x, double* work) {
    if (h->n_levels == 0) {
        coarse_solve(h->n_coarse, h->coarse_l, b, x);
        return;
    }
    /* per level: bl (n), xl (n), rl (n), w1, w2, tmp (n) */
    double* lv_b[64];
    double* lv_x[64];
    double* lv_r[64];
    double* p = work;
    int64_t nmax = h->n[0];
    double* w1 = p;
    p += nmax;
    double* w2 = p;
    p += nmax;
    double* tmp = p;
    p += nmax;
    for (int l = 0; l < h->n_levels; ++l) {
        lv_b[l] = p;
        p += h->n[l];
        lv_x[l] = p;
        p += h->n[l];
        lv_r[l] = p;
        p += h->n[l];
    }
    double* cb = p;
    p += h->n_coarse;
    double* cx = p;
    memcpy(lv_b[0], b, sizeof(double) * (size_t)h->n[0]);
    for (int l = 0; l < h->n_levels; ++l) {
        const int64_t n = h->n[l];
        smooth(h, l, lv_b[l], NULL, lv_x[l], w1, w2, tmp);
        spmv_f(n, h->a_ptr[l], h->a_col[l], h->a_val[l], lv_x[l], tmp);
        for (int64_t i = 0; i < n; ++i) lv_r[l][i] = lv_b[l][i] - tmp[i];
        const int64_t nc = (l + 1 < h->n_levels) ? h->n[l + 1] : h->n_coarse;
        double* nb = (l + 1 < h->n_levels) ? lv_b[l + 1] : cb;
        spmv_f(nc, h->r_ptr[l], h->r_col[l], h->r_val[l], lv_r[l], nb);
    }
    coarse_solve(h->n_coarse, h->coarse_l, cb, cx);
    const double* xc = cx;
    for (int l = h->n_levels - 1; l >= 0; --l) {
        const int64_t n = h->n[l];
        spmv_f(n, h->p_ptr[l], h->p_col[l], h->p_val[l], xc, tmp);
        for (int64_t i = 0; i < n; ++i) lv_x[l][i] += tmp[i];
        smooth(h, l, lv_b[l], lv_x[l], (l == 0) ? x : lv_x[l], w1, w2, tmp);
        xc = lv_x[l];
    }
}

OR_API int64_t or_vcycle_work_size(const or_hier* h) {
    int64_t s = (h->n_levels ? 3 * (int64_t)h->n[0] : 0) + 2 * (int64_t)h->n_coarse;
    for (int l = 0; l < h->n_levels; ++l) s += 3 * (int64_t)h->n[l];
    return s;
}

/* Fixed-partition dot product: partials over 4096-row tiles summed in tile order. */
static double dot_fixed(int64_t n, const double* a, const double* b) {
    const int64_t tile = 4096;
    const int64_t nt = (n + tile - 1) / tile;
    double* part = (double*)malloc(sizeof(double) * (size_t)(nt > 0 ? nt : 1));
#pragma omp parallel for schedule(static)
    for (int64_t t = 0; t < nt; ++t) {
        double s = 0.0;
        const int64_t e = (t + 1) * tile < n ? (t + 1) * tile : n;
        for (int64_t i = t * tile; i < e; ++i) s += a[i] * b[i];
        part[t] = s;
    }
    double s = 0.0;
    for (int64_t t = 0; t < nt; ++t) s += part[t];
    free(part);
    return s;
}

static void spmv_d(int64_t n, const int32_t* rp, const int32_t* ci, const double* v,
                   const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int32_t j = rp[i]; j < rp[i + 1]; ++j) s += v[j] * x[ci[j]];
        y[i] = s;
    }
}

/* FP64 PCG, relative residual ||r|| / ||b|| <= tol (solver.cpp:237-246: measured against b).
 * x holds the initial guess on entry. Returns the iteration count; *rel the final residual. */
OR_API int32_t or_pcg(const or_hier* h, int64_t n, const int32_t* rp, const int32_t* ci,
                      const double* v, const double* b, double* x, double tol, int32_t max_iters,
                      double* rel_out) {
    double* r = (double*)malloc(sizeof(double) * (size_t)n);
    double* z = (double*)malloc(sizeof(double) * (size_t)n);
    double* p = (double*)malloc(sizeof(double) * (size_t)n);
    double* q = (double*)malloc(sizeof(double) * (size_t)n);
    double* work = (double*)malloc(sizeof(double) * (size_t)or_vcycle_work_size(h));
    const double bn = sqrt(dot_fixed(n, b, b));
    int32_t it = 0;
    double rel = 0.0;
    if (bn == 0.0) {
        for (int64_t i = 0; i < n; ++i) x[i] = 0.0;
        goto done;
    }
    spmv_d(n, rp, ci, v, x, q);
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] - q[i];
    rel = sqrt(dot_fixed(n, r, r)) / bn;
    if (rel <= tol) goto done;
    or_vcycle_apply(h, r, z, work);
    memcpy(p, z, sizeof(double) * (size_t)n);
    double rz = dot_fixed(n, r, z);
    for (it = 1; it <= max_iters; ++it) {
        spmv_d(n, rp, ci, v, p, q);
        const double alpha = rz / dot_fixed(n, p, q);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
        }
        rel = sqrt(dot_fixed(n, r, r)) / bn;
        if (rel <= tol) break;
        or_vcycle_apply(h, r, z, work);
        const double rz_new = dot_fixed(n, r, z);
        const double beta = rz_new / rz;
        rz = rz_new;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    if (it > max_iters) it = max_iters;
done:
    *rel_out = rel;
    free(r);
    free(z);
    free(p);
    free(q);
    free(work);
    return it;
}

OR_API void or_spmv_d(int64_t n, const int32_t* rp, const int32_t* ci, const double* v,
                      const double* x, double* y) {
    spmv_d(n, rp, ci, v, x, y);
}

/* ---------------------------------------------------------------------------------------------
 * Coil side. fem/placement.py:279-317 and solver/coil/dadt.cu:27-104: dA/dt at targets r from
 * dipoles (s_j, m_j): didt * mu0/4pi * sum_j m_j x (r - s_j) / |r - s_j|^3, evaluated directly in
 * FP64 (the closed form the reference's kernel approximates in FP32).
 * ------------------------------------------------------------------------------------------- */
OR_API void or_dadt(int64_t n_dip, const double* s, const double* m, int64_t n, const double* r,
                    double didt, double mu0_4pi, double* out) {
    const double scale = didt * mu0_4pi;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double ax = 0.0, ay = 0.0, az = 0.0;
        const double rx = r[3 * i], ry = r[3 * i + 1], rz = r[3 * i + 2];
        for (int64_t j = 0; j < n_dip; ++j) {
            const double dx = rx - s[3 * j], dy = ry - s[3 * j + 1], dz = rz - s[3 * j + 2];
            const double d2 = dx * dx + dy * dy + dz * dz;
            const double inv3 = 1.0 / (d2 * sqrt(d2));
            const double mx = m[3 * j], my = m[3 * j + 1], mz = m[3 * j + 2];
            ax += (my * dz - mz * dy) * inv3;
            ay += (mz * dx - mx * dz) * inv3;
            az += (mx * dy - my * dx) * inv3;
        }
        out[3 * i] = scale * ax;
        out[3 * i + 1] = scale * ay;
        out[3 * i + 2] = scale * az;
    }
}

/* Closest point on triangle (Ericson), solver/coil/place.cu:66-119. */
static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static double closest_on_tri(const double* p, const double* a, const double* b, const double* c,
                             double* q) {
    double ab[3], ac[3], ap[3], bp[3], cp[3];
    for (int i = 0; i < 3; ++i) {
        ab[i] = b[i] - a[i];
        ac[i] = c[i] - a[i];
        ap[i] = p[i] - a[i];
        bp[i] = p[i] - b[i];
        cp[i] = p[i] - c[i];
    }
    const double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    const double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
    const double d5 = dot3(ab, cp), d6 = dot3(ac, cp);
    const double vc = d1 * d4 - d3 * d2;
    const double vb = d5 * d2 - d1 * d6;
    const double va = d3 * d6 - d5 * d4;
    if (d1 <= 0.0 && d2 <= 0.0) {
        for (int i = 0; i < 3; ++i) q[i] = a[i];
    } else if (d3 >= 0.0 && d4 <= d3) {
        for (int i = 0; i < 3; ++i) q[i] = b[i];
    } else if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
        const double t = d1 / (d1 - d3);
        for (int i = 0; i < 3; ++i) q[i] = a[i] + t * ab[i];
    } else if (d6 >= 0.0 && d5 <= d6) {
        for (int i = 0; i < 3; ++i) q[i] = c[i];
    } else if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        const double t = d2 / (d2 - d6);
        for (int i = 0; i < 3; ++i) q[i] = a[i] + t * ac[i];
    } else if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
        const double t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        for (int i = 0; i < 3; ++i) q[i] = b[i] + t * (c[i] - b[i]);
    } else {
        const double denom = 1.0 / (va + vb + vc);
        const double v = vb * denom, w = vc * denom;
        for (int i = 0; i < 3; ++i) q[i] = a[i] + v * ab[i] + w * ac[i];
    }
    const double dx = q[0] - p[0], dy = q[1] - p[1], dz = q[2] - p[2];
    return dx * dx + dy * dy + dz * dz;
}

/* Coil frames, solver/coil/place.cu:122-230 + fem/placement.py:56-93: closest scalp triangle
 * under the total order (d^2, index), frame [x|y|z|c] with z = -normal, y the normalised in-plane
 * handle direction, x = y cross z, c = proj + distance * normal. handles may be NULL. */
OR_API void or_place(const double* centers, const double* handles, const double* dists,
                     int64_t n_pl, const double* ta, const double* tb, const double* tc,
                     const double* tn, int64_t n_tri, double* out, int32_t* degenerate) {
    for (int64_t p = 0; p < n_pl; ++p) {
        const double* ctr = centers + 3 * p;
        double best = INFINITY;
        int64_t tri = -1;
        double q[3];
        for (int64_t j = 0; j < n_tri; ++j) {
            const double d2 = closest_on_tri(ctr, ta + 3 * j, tb + 3 * j, tc + 3 * j, q);
            if (d2 < best) {
                best = d2;
                tri = j;
            }
        }
        double proj[3];
        closest_on_tri(ctr, ta + 3 * tri, tb + 3 * tri, tc + 3 * tri, proj);
        const double* nrm = tn + 3 * tri;
        const double z[3] = {-nrm[0], -nrm[1], -nrm[2]};
        double y[3];
        int from_handle = 0;
        if (handles != NULL) {
            for (int i = 0; i < 3; ++i) y[i] = handles[3 * p + i] - proj[i];
            if (dot3(y, y) > 0.0) {
                const double inv = 1.0 / sqrt(dot3(y, y));
                for (int i = 0; i < 3; ++i) y[i] *= inv;
                const double cz = dot3(y, z);
                if ((1.0 - cz * cz) > 1e-12) {
                    for (int i = 0; i < 3; ++i) y[i] -= z[i] * cz;
                    const double inv2 = 1.0 / sqrt(dot3(y, y));
                    for (int i = 0; i < 3; ++i) y[i] *= inv2;
                    from_handle = 1;
                }
            }
        }
        if (!from_handle) {
            int m = 0;
            if (fabs(z[1]) < fabs(z[m])) m = 1;
            if (fabs(z[2]) < fabs(z[m])) m = 2;
            double ref[3] = {0.0, 0.0, 0.0};
            ref[m] = 1.0;
            y[0] = z[1] * ref[2] - z[2] * ref[1];
            y[1] = z[2] * ref[0] - z[0] * ref[2];
            y[2] = z[0] * ref[1] - z[1] * ref[0];
            const double inv = 1.0 / sqrt(dot3(y, y));
            for (int i = 0; i < 3; ++i) y[i] *= inv;
        }
        if (degenerate) degenerate[p] = from_handle ? 0 : 1;
        const double x[3] = {y[1] * z[2] - y[2] * z[1], y[2] * z[0] - y[0] * z[2],
                             y[0] * z[1] - y[1] * z[0]};
        double* o = out + 16 * p;
        for (int i = 0; i < 16; ++i) o[i] = 0.0;
        for (int r = 0; r < 3; ++r) {
            o[4 * r + 0] = x[r];
            o[4 * r + 1] = y[r];
            o[4 * r + 2] = z[r];
            o[4 * r + 3] = proj[r] + dists[p] * nrm[r];
        }
        o[15] = 1.0;
    }
}

/* Smoothed skin normals, solver/fem/skin.cu:46-150: area-weighted face normals spread to nodes
 * in node2corner order and gathered back, two rounds, normalised. */
OR_API void or_skin_normals(const double* nodes, const int32_t* tris, int64_t n_tri, int64_t n_nodes,
                            const int32_t* ptr, const int32_t* idx, double* out) {
    double* face = (double*)malloc(sizeof(double) * (size_t)(3 * n_tri + 3));
    double* node = (double*)calloc((size_t)(3 * n_nodes + 3), sizeof(double));
    for (int64_t t = 0; t < n_tri; ++t) {
        const double *a = nodes + 3 * (int64_t)tris[3 * t], *b = nodes + 3 * (int64_t)tris[3 * t + 1],
                     *c = nodes + 3 * (int64_t)tris[3 * t + 2];
        double u[3], v[3];
        for (int i = 0; i < 3; ++i) {
            u[i] = b[i] - a[i];
            v[i] = c[i] - a[i];
        }
        face[3 * t] = u[1] * v[2] - u[2] * v[1];
        face[3 * t + 1] = u[2] * v[0] - u[0] * v[2];
        face[3 * t + 2] = u[0] * v[1] - u[1] * v[0];
    }
    for (int round = 0; round < 2; ++round) {
        if (round > 0) {
            for (int64_t t = 0; t < n_tri; ++t) {
                double acc[3];
                for (int i = 0; i < 3; ++i)
                    acc[i] = node[3 * (int64_t)tris[3 * t] + i] + node[3 * (int64_t)tris[3 * t + 1] + i] +
                             node[3 * (int64_t)tris[3 * t + 2] + i];
                const double nn = sqrt(acc[0] * acc[0] + acc[1] * acc[1] + acc[2] * acc[2]);
                if (nn > 0.0)
                    for (int i = 0; i < 3; ++i) acc[i] /= nn;
                for (int i = 0; i < 3; ++i) face[3 * t + i] = acc[i];
            }
        }
        for (int64_t nd = 0; nd < n_nodes; ++nd) {
            double acc[3] = {0.0, 0.0, 0.0};
            for (int32_t p = ptr[nd]; p < ptr[nd + 1]; ++p) {
                const int64_t t = idx[p] / 3;
                for (int i = 0; i < 3; ++i) acc[i] += face[3 * t + i];
            }
            for (int i = 0; i < 3; ++i) node[3 * nd + i] = acc[i];
        }
    }
    for (int64_t nd = 0; nd < n_nodes; ++nd) {
        double* v = node + 3 * nd;
        const double nn = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (nn > 0.0)
            for (int i = 0; i < 3; ++i) v[i] /= nn;
    }
    for (int64_t t = 0; t < n_tri; ++t) {
        double acc[3];
        for (int i = 0; i < 3; ++i)
            acc[i] = node[3 * (int64_t)tris[3 * t] + i] + node[3 * (int64_t)tris[3 * t + 1] + i] +
                     node[3 * (int64_t)tris[3 * t + 2] + i];
        const double nn = sqrt(acc[0] * acc[0] + acc[1] * acc[1] + acc[2] * acc[2]);
        if (nn > 0.0)
            for (int i = 0; i < 3; ++i) acc[i] /= nn;
        for (int i = 0; i < 3; ++i) out[3 * t + i] = acc[i];
    }
    free(face);
    free(node);
}

/* RHS, solver/fem/rhs.cu:13-104: b[node] = sum over the node's corners c (node2corner order) of
 * neg_vc[e] * (dadt_elm[e] . g[c]), neg_vc = -vol * sigma. FP64. */
OR_API void or_rhs(const double* dadt_elm, const double* g, const double* neg_vc, const int32_t* ptr,
                   const int32_t* idx, int64_t n, double* b) {
#pragma omp parallel for schedule(static)
    for (int64_t nd = 0; nd < n; ++nd) {
        double acc = 0.0;
        for (int32_t p = ptr[nd]; p < ptr[nd + 1]; ++p) {
            const int64_t c = idx[p], e = c >> 2;
            const double* d = dadt_elm + 3 * e;
            const double* gc = g + 3 * c;
            acc += neg_vc[e] * (d[0] * gc[0] + d[1] * gc[1] + d[2] * gc[2]);
        }
        b[nd] = acc;
    }
}

/* E = -grad v - dadt_elm per tetrahedron and |E|, solver/fem/reconstruct.cu:12-31. FP64. */
OR_API void or_reconstruct(const double* v, const int32_t* tets, const double* g,
                           const double* dadt_elm, int64_t m, double* e_out, double* magn) {
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m; ++e) {
        double gx = 0.0, gy = 0.0, gz = 0.0;
        for (int i = 0; i < 4; ++i) {
            const double vi = v[tets[4 * e + i]];
            gx += vi * g[12 * e + 3 * i];
            gy += vi * g[12 * e + 3 * i + 1];
            gz += vi * g[12 * e + 3 * i + 2];
        }
        const double ex = -gx - dadt_elm[3 * e], ey = -gy - dadt_elm[3 * e + 1],
                     ez = -gz - dadt_elm[3 * e + 2];
        e_out[3 * e] = ex;
        e_out[3 * e + 1] = ey;
        e_out[3 * e + 2] = ez;
        magn[e] = sqrt(ex * ex + ey * ey + ez * ez);
    }
}

OR_API int32_t or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

OR_API void or_set_threads(int32_t n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
