/*
 * cunibs_amd: C ABI of the MI355X TMS E-field FEM library.
 *
 * The reference binds its CUDA kernels through a nanobind module (cuNIBS
 * src/cunibs/solver/bindings.cpp:30-676) that takes CuPy device arrays. This ABI replaces it with
 * plain pointers and sizes, no exceptions across the boundary, and two layers:
 *
 *   1. A coarse layer the Python front end drives: a context built once per head mesh
 *      (replacing fem/solve.py:690-746, build_context) and a block solve of up to 32 placements
 *      (replacing fem/solve.py:865-979, solve_placements_block, and the metric reduction of
 *      simulation.py:330-376). Inputs and outputs are host arrays; every call is synchronous.
 *   2. Kernel-level entry points mirroring individual reference bindings, used by parity tests.
 *      They also take host arrays and copy in and out.
 *
 * Errors: every function returns a status code; cn_last_error() gives the message of the calling
 * thread's last failure. CN_EINVAL maps to Python ValueError, CN_EHIP to RuntimeError,
 * CN_ENOCONV to SolverConvergenceError (fem/solve.py:368-392).
 *
 * Threading: one thread at a time per context; contexts are independent.
 *
 * Indices are int32 and 0-based; coordinates are millimetres; arrays are C-contiguous.
 */
#ifndef CUNIBS_AMD_H
#define CUNIBS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CN_OK 0
#define CN_EINVAL 1
#define CN_EHIP 2
#define CN_ENOCONV 3
#define CN_EINTERNAL 4

/* Length of one placement's region-metric record, cn_solve_block's `metrics` output. Layout:
 * [0] peak |E|, [1..3] peak location mm, [4..6] |E|-weighted centre of gravity mm,
 * [7] region volume m^3, [8] focality at 0.5 x p99.9 (m^3), [9] focality at 0.75 x p99.9,
 * [10] volume-weighted mean, [11] std, [12..15] p50, p95, p99, p99.9 (metrics.py:209-243). */
#define CN_METRICS_LEN 16

typedef struct cn_context cn_context;

/* Message of this thread's last failed call ("" if none). */
const char* cn_last_error(void);
int32_t cn_version(void);
/* Compiled block widths, increasing (solver/__init__.py BLOCK_SIZES, bindings.cpp:31-35).
 * Writes up to cap entries, returns the count. */
int32_t cn_block_sizes(int32_t* out, int32_t cap);
/* Largest compiled block width (fem/solve.py:505 MAX_BLOCK). */
int32_t cn_max_block(void);

typedef struct {
    int32_t min_coarse_rows; /* fem/solve.py:146, default 128 */
    int32_t max_levels;      /* default 50 */
    int32_t rounds;          /* pairwise passes per level, default 2 */
    int32_t degree;          /* Chebyshev degree, default 2 (1..6) */
    float lower_ratio;       /* Chebyshev interval [1/lower_ratio, 1], default 12 */
    int32_t keep_stiffness;  /* keep the full FP64 stiffness for cn_context_stiffness */
    const double* conductivity; /* sigma by tag (NaN = unknown tag); NULL: fem/assembly.py:15-26 */
    int32_t n_conductivity;
    int32_t region_tag;      /* tag summarised by cn_solve_block's metrics; default 2 (GM) */
} cn_options;

void cn_default_options(cn_options* out);

/* Build the per-mesh state on the GPU: Morton layout, FP64 P1 assembly, grounding, AMG
 * hierarchy, skin normals (fem/solve.py:690-746). Rejects zero-volume elements and unknown tags
 * with CN_EINVAL. */
int32_t cn_context_create(const double* nodes_mm, int64_t n_nodes, const int32_t* tet_nodes,
                          const int32_t* tet_tags, int64_t n_tets, const int32_t* skin_tris,
                          int64_t n_skin, const cn_options* opts, cn_context** out);
void cn_context_destroy(cn_context* ctx);

/* Launch and smoother settings of a built context, changeable between solves (the autotuner's
 * knobs; the reference fixes them: Chebyshev degree 2 and ratio 12 at fem/solve.py:153-168,
 * threads per row by choose_tpr at solver/amg/vcycle.cu:104-120). Setting them drops the
 * captured iterations; the next solve of each width captures again. */
typedef struct {
    int32_t degree;     /* Chebyshev degree of the V-cycle smoother, 1..6 */
    float lower_ratio;  /* Chebyshev interval [1/lower_ratio, 1], > 1 */
    int32_t fine_sub;   /* lanes per 4-column group of the fine level's row kernels (the SpMM
                           tile): 0 = chosen from the level's shape, else a power of two <= 64,
                           clamped to one wavefront per row at wide blocks */
    int32_t coarse_sub; /* the same for the smoothing and residual kernels of every coarser
                           level (0 = chosen from each level's shape) */
    int32_t coil_blocks; /* cn_solve_sweep: workgroups of the next block's dA/dt while it overlaps
                            the current block's solve (0 = the full grid); fewer leave the solve
                            more of the device */
    int32_t layout;     /* finest level's operator in the row kernels: 0 = the sliced ELL copy
                           (built when every row has at most 24 entries) for blocks of up to 4
                           columns and CSR above, 1 = CSR, 2 = the ELL copy at every width; the
                           sums are the same */
    int32_t dense_tail; /* 1 (default): the V-cycle below the first level of at most 4096 rows
                           (that level's smoothing, the coarser levels and the coarse solve) is
                           formed once as a dense FP32 operator and applied as one product;
                           0: level by level */
} cn_tuning;

int32_t cn_context_get_tuning(const cn_context* ctx, cn_tuning* out);
int32_t cn_context_set_tuning(cn_context* ctx, const cn_tuning* tuning);

#define CN_MAX_LEVELS 32

typedef struct {
    int64_t n_nodes, n_tets, n_skin;
    int64_t n_reduced, nnz_reduced, nnz_full;
    int32_t ground_node;
    int32_t n_levels; /* smoothed levels; the coarsest operator follows them */
    int64_t level_rows[CN_MAX_LEVELS + 1];
    int64_t level_nnz[CN_MAX_LEVELS + 1];
    int64_t n_region;
    double region_volume_m3;
    int32_t n_cu;
    int64_t level_p_nnz[CN_MAX_LEVELS]; /* nonzeros of each smoothed level's prolongator */
    int32_t ell_width_full; /* slots per row of the finest operator's sliced ELL copy (0: none) */
    int32_t ell_width_off;  /* the same for the differenced off-diagonal operator */
    int32_t preconditioner_rebuilds; /* block solves whose missed columns were retried singly under a
                                        rebuilt hierarchy (fem/solve.py:508-552) */
    int32_t matched_preconditioners; /* throwaway preconditioners built for one column's own
                                        conductivities after it missed (uq/conductivity/run.py:333-342) */
    int32_t dense_tail_level; /* the level from which the V-cycle is applied as one dense product
                                 (cn_tuning.dense_tail), -1 when none */
} cn_info;

int32_t cn_context_info(const cn_context* ctx, cn_info* out);
/* Host wall time (ms) of the context build's phases, each ending at a wait on the library's
 * stream: ms10[0] argument checks; [1] uploads; [2] the Morton order; [3] gradients, volume and
 * conductivity checks; [4] node2corner and the stiffness pattern; [5] stiffness values,
 * grounding, resident arrays; [6] AMG hierarchy; [7] differenced operator, solver and ELL copies;
 * [8] skin surface and region slice; [9] the most recent recovery operator build (0 before the
 * first). A measurement aid with no counterpart in the reference. */
int32_t cn_context_setup_times(const cn_context* ctx, double* ms10);

/* The reordered mesh every result is indexed by (fem/solve.py:645-662), and the permutations
 * from the input order. Any pointer may be NULL. */
int32_t cn_context_mesh(const cn_context* ctx, double* nodes_mm, int32_t* tet_nodes,
                        int32_t* tet_tags, int32_t* skin_tris, int32_t* node_perm,
                        int32_t* tet_perm);
/* Per-element resident arrays: volumes (m^3, FP32), barycentres (mm, FP64), P1 gradients
 * (m, 4, 3) FP32, -vol*sigma FP32. Any pointer may be NULL. */
int32_t cn_context_geometry(const cn_context* ctx, float* vols, double* barycenters_mm, float* g,
                            float* neg_vc);
/* Full FP64 stiffness CSR (needs keep_stiffness): indptr (n+1), indices, data (nnz_full). */
int32_t cn_context_stiffness(const cn_context* ctx, int32_t* indptr, int32_t* indices, double* data);
/* Reduced (grounded) operator: ptr (n_red+1), col, FP64 values (nnz_reduced). */
int32_t cn_context_reduced(const cn_context* ctx, int32_t* ptr, int32_t* col, double* val);
/* One hierarchy level: operator A (FP32 CSR), l1 dinv, aggregates, prolongator P (FP32 CSR).
 * Call with NULL arrays first to read the sizes. level == n_levels gives the coarsest operator
 * (A only). */
int32_t cn_context_level(const cn_context* ctx, int32_t level, int64_t* n_rows, int64_t* nnz,
                         int64_t* n_coarse, int64_t* p_nnz, int32_t* a_ptr, int32_t* a_col,
                         float* a_val, float* dinv, int32_t* agg, int32_t* p_ptr, int32_t* p_col,
                         float* p_val);
/* Smoothed skin-triangle normals (n_skin, 3) FP64 (fem/solve.py:665-687). */
int32_t cn_context_skin_normals(const cn_context* ctx, double* out);

typedef struct {
    double tolerance;  /* relative true residual per column, default 1e-6 */
    int32_t max_iters; /* default 2000 */
    int32_t max_refine;
    double drop_tol;
    int32_t use_graph;
    int32_t check_every;
    int32_t recovery;  /* CN_RECOVERY_*: the field E, |E| and the metrics describe; default raw */
    int32_t warm_start; /* 1: Galerkin reduced-basis guess, 2: supplied coefficients (cn_warm_*); 0 */
    int32_t collect;    /* 1: append this block's solutions to the warm-start basis; default 0 */
    /* Per-placement conductivities (k x n_sigma, S/m by tag; NULL: the context's table for every
     * placement). Each placement then has its own operator K(sigma_c) = sum_t sigma_c,t K_t and
     * RHS (uq/conductivity/assembly.py:37-146 decomposes them the same way); the block runs as
     * independent CG chains under the context's AMG hierarchy. n_sigma must exceed every tag
     * present in the mesh. */
    const double* sigma;
    int32_t n_sigma;
    /* 1: the block runs as k independent CG chains in lockstep (the Gram matrices restricted to
     * their diagonals, each column with its own step lengths and its own freeze), the reference's
     * block solver (solver/amg/solver.cpp:84-198, fem/solve.py:508-552); 0 (default): the true
     * block method, whose shared Krylov space converges in fewer iterations. */
    int32_t independent;
    /* cn_solve_sweep: each full block starts from the Galerkin projection of its solution onto the
     * span of this many earlier full blocks' solutions (at most 2, and 32 vectors), the projection
     * method for successive right-hand sides; 0: every block from zero. Default 2. No reference
     * counterpart (its placements start from zero). */
    int32_t recycle;
} cn_solve_options;

void cn_default_solve_options(cn_solve_options* out);

/* Host outputs of a block solve, placement-major; NULL = not wanted. */
typedef struct {
    double* transforms;        /* (k, 4, 4) coil-to-head affine, mm */
    float* magnE;              /* (k, n_tets) */
    float* E;                  /* (k, n_tets, 3) */
    double* v;                 /* (k, n_nodes), ground node 0 */
    double* metrics;           /* (k, CN_METRICS_LEN), region_tag's summary */
    int32_t* iterations;       /* (k) inner PCG iterations until each column froze */
    double* relative_residual; /* (k) final FP64 true residual / ||b|| */
    float* dadt_elm;           /* (k, n_tets, 3) element dA/dt */
    float* dadt_nodes;         /* (k, n_nodes, 3) nodal dA/dt */
    float* rhs;                /* (k, n_nodes) right-hand side before grounding */
    int32_t* degenerate;       /* (k) 1 where the handle left the frame undefined */
    float* E_slots;            /* (k, n_slots, 3) recovered field on its slots (recovery != raw) */
    double* x0;                /* (k, n_reduced) the initial guess PCG started from (zeros if cold) */
    int32_t total_iterations;  /* out: inner iterations of the block */
    int32_t refinements;       /* out: FP64 refinement rounds */
    int32_t round_iterations[8]; /* out: cumulative inner iterations at the end of each round */
    double round_residual[8];    /* out: worst FP64 true relative residual after each round */
} cn_block_outputs;

/* Recovery modes (fem/recovery.py:77-100): the raw per-element field; superconvergent patch
 * recovery over each node's incident elements (spr_global) or each (node, tissue) pair's
 * (spr_tissue); or the harmonic-polynomial fit of the potential over same-tissue node patches
 * (harmonic, the reference's default). */
#define CN_RECOVERY_RAW 0
#define CN_RECOVERY_SPR_GLOBAL 1
#define CN_RECOVERY_SPR_TISSUE 2
#define CN_RECOVERY_HARMONIC 3

/* Build (once; cached on the context) the operator of a recovery mode (fem/recovery.py:385-453,
 * ensure_recovery). Reports the slot count and how many slots took the volume-weighted average.*/
int32_t cn_recovery_build(cn_context* ctx, int32_t mode, int64_t* n_slots, int32_t* n_fallback);
/* Copy a built operator back: ptr (n_slots+1), idx and w (4 n_tets: corner ids and weights; SPR
 * modes), slot_of_corner (4 n_tets), slot_node and slot_tag (n_slots; slot_tag for the tissue
 * modes), boundary (n_nodes; SPR modes), status (n_slots: SPR 0 fitted / 1 averaged; harmonic
 * 0 cubic / 1 quadratic / 2 linear / 3 none). Any pointer may be NULL. */
int32_t cn_recovery_operator(const cn_context* ctx, int32_t mode, int32_t* ptr, int32_t* idx, float* w,
                             int32_t* slot_of_corner, int32_t* slot_node, int32_t* slot_tag,
                             int32_t* boundary, int32_t* status);
/* The harmonic operator's node patches: pptr (n_slots+1), pidx (nnz) and the three gradient
 * weights per entry w3 (nnz, 3); call with NULL arrays first to read nnz. */
int32_t cn_recovery_patches(const cn_context* ctx, int64_t* nnz, int32_t* n_undetermined, int32_t* pptr,
                            int32_t* pidx, float* w3);

/* Reduced-basis warm start. No reference counterpart (its nearest analogue is the Galerkin
 * initial guess of uq/conductivity/run.py:154-200). Solves with opt->collect = 1 append their
 * solutions (FP32) to a snapshot basis W of up to 2048 columns. cn_warm_gram returns
 * G = W^T A W (count x count); the caller hands back M = V Lambda^-1/2 of its leading R <= 256
 * eigenpairs, and cn_warm_set_modes forms the modes Phi = W M (BF16 MFMA GEMM). cn_warm_mode_grams
 * returns Phi^T A Phi (R x R) and Phi^T A W (R x count); cn_warm_set_projector takes
 * P = (Phi^T A Phi)^-1. A solve with opt->warm_start = 1 starts from x0 = Phi P Phi^T b; with
 * warm_start = 2 from x0 = Phi c for the k x R coefficients given to cn_warm_set_coefficients
 * (e.g. predicted by the warm-start network). PCG iterates to the same tolerance either way. */
int32_t cn_warm_reset(cn_context* ctx, int32_t capacity);
int32_t cn_warm_count(const cn_context* ctx, int32_t* count);
int32_t cn_warm_gram(cn_context* ctx, double* G);
int32_t cn_warm_set_modes(cn_context* ctx, const double* M, int32_t count, int32_t R);
int32_t cn_warm_mode_grams(cn_context* ctx, double* G_modes, double* G_cross);
int32_t cn_warm_set_projector(cn_context* ctx, const double* P, int32_t R);
int32_t cn_warm_set_coefficients(cn_context* ctx, const double* c, int32_t k, int32_t R);
/* Phi^T b for the first k columns of the right-hand side of the most recent block solve (k x R
 * FP64, row-major): with P and Phi^T A Phi it gives the energy-norm error of each warm start's
 * initial guess. A measurement aid with no counterpart in the reference. */
int32_t cn_warm_project_last(cn_context* ctx, int32_t k, double* out);
/* The modes as BF16 hi + lo pairs: out holds (2, n_reduced, width) BF16 bits, the high halves
 * then the low halves, each row-major; width (a multiple of 16 >= R) is written to *width; call
 * with out = NULL to read it. A mode value is hi + lo. */
int32_t cn_warm_modes(const cn_context* ctx, uint16_t* out, int32_t* width);

/* Conductivity uncertainty (uq/conductivity of the reference): for one placement, the per-tissue
 * unit-conductivity operators K_t and right-hand sides b_t (assembly.py:37-146), so that a sample
 * sigma gives K(sigma) = sum_t sigma_t K_t and b(sigma) = sum_t sigma_t b_t on the shared pattern.
 * cn_uq_prepare returns the tissue tags (ascending, at most 16) that sigma's columns follow, and
 * builds the draws' initial guess (run.py:100-200): the nominal solution x_nom, the sensitivities
 * S_t (A_nom S_t = b_t - K_t x_nom, to 1e-3) and the Galerkin correction on their span.
 * cn_uq_solve solves k <= 16 samples (sigma: k x n_tissue) as independent CG chains under the
 * nominal AMG preconditioner, from x0 = x_nom + S E^+ S^T r0(sigma) when opt->warm_start != 0
 * (from zero otherwise), with the same outputs as cn_solve_block, and accumulates the |E|
 * moments that cn_uq_moments returns (per element mean and sample variance, FP64).
 * cn_uq_guess_info: out[0] iterations of the nominal solve, out[1] of the sensitivity block,
 * out[2] rank of the correction (0: x_nom alone). */
int32_t cn_uq_prepare(cn_context* ctx, const double* dip_pos_m, const double* dip_moment, int64_t n_dip,
                      const double* center_mm, const double* handle_mm, double distance_mm, double didt,
                      int32_t* n_tissue, int32_t* tags);
int32_t cn_uq_solve(cn_context* ctx, const double* sigma, int32_t k, const cn_solve_options* opt,
                    cn_block_outputs* out);
int32_t cn_uq_moments(const cn_context* ctx, double* mean, double* var, int64_t* count);
int32_t cn_uq_guess_info(const cn_context* ctx, int32_t* out3);

/* The warm-start network: n_in (<= 16) inputs -> hidden -> hidden -> n_out, GELU, trained on the
 * GPU in FP32 (Adam, mean squared error), inference with BF16 weights on MFMA. No reference
 * counterpart. Parameters are one flat FP32 array: W1 (16 x hidden, rows >= n_in zero), b1,
 * W2 (hidden x hidden), b2, W3 (hidden x n_out), b3. */
typedef struct cn_mlp cn_mlp;
int32_t cn_mlp_create(int32_t n_in, int32_t hidden, int32_t n_out, uint32_t seed, cn_mlp** out);
void cn_mlp_destroy(cn_mlp* net);
int32_t cn_mlp_n_params(const cn_mlp* net, int64_t* n);
int32_t cn_mlp_set_params(cn_mlp* net, const float* params);
int32_t cn_mlp_get_params(const cn_mlp* net, float* params);
/* Loss and gradients of one batch X (B x n_in), T (B x n_out). */
int32_t cn_mlp_gradients(cn_mlp* net, const float* X, const float* T, int32_t B, float* grads, double* loss);
int32_t cn_mlp_train(cn_mlp* net, const float* X, const float* T, int64_t N, int32_t epochs, int32_t batch,
                     float lr, uint32_t seed, double* loss);
int32_t cn_mlp_predict(cn_mlp* net, const float* X, int32_t B, float* Y);

/* Solve k <= 32 placements as one block (fem/solve.py:865-979): coil frames, dA/dt, RHS, block
 * PCG, E reconstruction, recovery (opt->recovery; built on first use), region metrics. Dipoles in metres / A m^2; centres and
 * handles (k, 3) mm; distances (k) mm; didt in A/s. A degenerate handle is CN_EINVAL naming the
 * placement (fem/placement.py:182-215). x0 (optional, (k, n_reduced) FP64) warm-starts PCG. */
int32_t cn_solve_block(cn_context* ctx, const double* dip_pos_m, const double* dip_moment,
                       int64_t n_dip, const double* centers_mm, const double* handles_mm,
                       const double* distances_mm, int32_t k, double didt,
                       const cn_solve_options* opt, const double* x0, cn_block_outputs* out);

/* A sweep of n_place placements in blocks of block_k (the loop of Subject.iter_simulate,
 * simulation.py, over fem/solve.py:865-979 per block), returning only the per-placement summaries:
 * transforms (n_place, 4, 4), metrics (n_place, CN_METRICS_LEN), iterations and relative_residual
 * (n_place), block_iterations (one per block); any may be NULL. Every block's numerics are cn_solve_block's for the same
 * placements; the library pipelines them, computing the next block's dA/dt and right-hand side on
 * a second stream while the current block solves. opt->sigma, opt->collect and warm_start = 2 are
 * not accepted here (CN_EINVAL); warm_start = 1 uses the context's modes. */
int32_t cn_solve_sweep(cn_context* ctx, const double* dip_pos_m, const double* dip_moment, int64_t n_dip,
                       const double* centers_mm, const double* handles_mm, const double* distances_mm,
                       int32_t n_place, int32_t block_k, double didt, const cn_solve_options* opt,
                       double* transforms, double* metrics, int32_t* iterations, double* relative_residual,
                       int32_t* block_iterations);

/* ---- kernel-level entry points (reference binding each one stands in for) ---- */

/* bindings.cpp:175-192 p1_gradients -> (m,4,3) gradients 1/m and volumes m^3, FP64. */
int32_t cn_k_p1_gradients(const double* nodes_mm, int64_t n_nodes, const int32_t* tet_nodes,
                          int64_t n_tets, double* g, double* vols);
/* fem/solve.py:315-323 Morton keys. */
int32_t cn_k_morton_keys(const double* points_mm, int64_t n, uint64_t* keys);
/* bindings.cpp:246-259 tet_lowest_node. */
int32_t cn_k_tet_lowest_node(const int32_t* inverse, int64_t n_nodes, const int32_t* tet_nodes,
                             int64_t n_tets, int32_t* lowest);
/* fem/assembly.py:142-162 build_node2corner: ptr (n_nodes+1), idx (n_elem*corners). */
int32_t cn_k_node2corner(const int32_t* elems, int64_t n_elem, int32_t corners, int64_t n_nodes,
                         int32_t* ptr, int32_t* idx);
/* bindings.cpp:354-370 l1_dinv (FP32 CSR). */
int32_t cn_k_l1_dinv(int64_t n, const int32_t* ptr, const int32_t* col, const float* val,
                     float* dinv);
/* bindings.cpp:85-104 select_size4: one SIZE_4 pass; returns the aggregate count in *n_coarse. */
int32_t cn_k_select_size4(int64_t n, const int32_t* ptr, const int32_t* col, const float* val,
                          int32_t* agg, int64_t* n_coarse);
/* Sparse product C = A B (FP32, fixed association), the Galerkin building block that replaces
 * cupyx SpGEMM (fem/solve.py:119-133). Call with c_col == NULL to read *c_nnz (c_ptr filled). */
int32_t cn_k_spgemm(int64_t a_rows, int64_t b_rows, int64_t b_cols, const int32_t* a_ptr,
                    const int32_t* a_col, const float* a_val, const int32_t* b_ptr,
                    const int32_t* b_col, const float* b_val, int32_t* c_ptr, int32_t* c_col,
                    float* c_val, int64_t* c_nnz);
/* Y = A X on the context's reduced FP32 operator; X, Y (n_reduced, K), K in the block sizes. */
int32_t cn_k_spmm(cn_context* ctx, const float* X, float* Y, int32_t K);
/* The same product in the block solver's differenced form, y_i = sum_{j!=i} a_ij (x_j - x_i)
 * + s_i x_i with s_i the FP64 row sum (no counterpart in the reference: it is what keeps the
 * FP32 recursive residual close to the FP64 true residual). */
int32_t cn_k_spmm_diff(cn_context* ctx, const float* X, float* Y, int32_t K);
/* FP64 Gram matrices of the block solver, G1 = A^T B1 and (B2 != NULL) G2 = A^T B2, for
 * A, B1, B2 (n_reduced, K) FP32: the products behind the search-block A-orthonormalisation,
 * the full-matrix form of the reference's column-wise dots (amg/block_cg.cu:145-160,371-379).
 * out: K*K, then K*K more for G2; row a, column c. */
int32_t cn_k_gram(cn_context* ctx, const float* A, const float* B1, const float* B2, double* out,
                  int32_t K);
/* X = M B, one zero-guess V-cycle (vcycle.cu:650-663 apply / apply_block). */
int32_t cn_k_vcycle(cn_context* ctx, const float* B, float* X, int32_t K);
/* bindings.cpp:194-206 dadt_nbody: targets r (n,3) metres relative to the coil origin, dipoles
 * s (D,3), packed moments mp (D,6) = [m | m x s], |s|^2 (D); out (n,3). FP32. */
int32_t cn_k_dadt(const float* s, const float* mp, const float* sn, int64_t n_dip, const float* r,
                  int64_t n, float didt, float mu0_4pi, float* out);
/* bindings.cpp:657-675 place_transforms on the context's scalp; handles may be NULL. */
int32_t cn_k_place(cn_context* ctx, const double* centers_mm, const double* handles_mm,
                   const double* distances_mm, int32_t n_pl, double* out, int32_t* degenerate);

/* Average device time (ms) of `reps` launches of the block solve's fine-level product at width K,
 * timed with HIP events on the library's stream. For roofline reporting and launch tuning.
 * mode 0: plain CSR SpMM Y = A X; 1: the differenced SpMM of the block iteration (S = A W);
 * 2: mode 1 followed by its Gram pass (W^T S, W^T R). sub = 0 picks the sub-lanes per column
 * group the solver uses; a power of two forces that count. */
int32_t cn_bench_spmm(cn_context* ctx, int32_t K, int32_t reps, int32_t mode, int32_t sub, double* ms);
/* Same for one V-cycle at width K. */
int32_t cn_bench_vcycle(cn_context* ctx, int32_t K, int32_t reps, double* ms);
/* In-solve kernel timing, a measurement aid with no counterpart in the reference. When on, every
 * block-PCG iteration (captured in its graph) times itself on the device wall clock: the
 * differenced SpMM's blocks stamp their start and end, one-thread stamp launches mark the V-cycle
 * and the iteration, and an accumulator launch at the end of the iteration adds the spans of
 * every iteration that ran in full (HIP events recorded inside a captured graph cannot be timed
 * on this runtime). Every call resets the sums; switching drops the captured iterations. */
int32_t cn_timing_enable(cn_context* ctx, int32_t on);
/* out[0] iterations sampled, out[1] mean ms of the differenced SpMM (earliest block start to
 * latest block end), out[2] of the V-cycle, out[3] of the whole iteration; out[4] placements whose
 * dA/dt was timed (HIP events around the N-body launch of each block, which runs outside the
 * graphs), out[5] mean ms of dA/dt per placement. */
int32_t cn_timing_read(const cn_context* ctx, double* out6);
/* Where the blocks solved while timing is on spent their device time, in ms summed over those
 * blocks (HIP events recorded on the library's stream between the phases of cn_solve_block, so
 * host work between the phases shows up in the phase it delays): out[0] coil frames, out[1] dA/dt,
 * out[2] element pass and RHS, out[3] initial guess, out[4] block PCG, out[5] field
 * reconstruction and recovery, out[6] outputs and region metrics; out[7] blocks, out[8]
 * placements. A measurement aid with no counterpart in the reference. */
int32_t cn_timing_phases(const cn_context* ctx, double* out9);
/* Average device time (ms) of the warm-start guess for a block of width K (<= 16) on the right-hand
 * sides of the last block solved at that width: ms3[0] the projection's block partials of
 * Phi^T B, ms3[1] their reduction and c = P Phi^T B, ms3[2] the expansion x0 = Phi c. Needs modes
 * (cn_warm_set_modes). A measurement aid with no counterpart in the reference. */
int32_t cn_bench_guess(cn_context* ctx, int32_t K, int32_t reps, double* ms3);
/* Average device time (ms) of the dA/dt N-body for k copies of one placement over the context's
 * nodes, repeated launches. A measurement aid with no counterpart in the reference. */
int32_t cn_bench_dadt(cn_context* ctx, const double* dip_pos_m, const double* dip_moment, int64_t n_dip,
                      const double* center_mm, const double* handle_mm, double distance_mm, int32_t k,
                      int32_t reps, double* ms);
/* Average device time (ms) of a streaming device-to-device copy of `bytes` bytes (a multiple of 16)
 * on the library's stream: the calibration launch for the HBM traffic counters and the practical
 * bandwidth ceiling. No counterpart in the reference (a measurement aid). */
int32_t cn_bench_copy(cn_context* ctx, int64_t bytes, int32_t reps, double* ms);
/* Average device time (ms) of the block stages around the solve, k copies of one placement,
 * repeated launches (HIP events): ms3[0] the coil stage (dA/dt N-body, element pass, node RHS),
 * ms3[1] the field on a zero potential (reconstruction, and the recovery named by `recovery`,
 * a CN_RECOVERY_* code), ms3[2] the region metrics (one host sync inside). The block PCG does not
 * run. A measurement aid with no counterpart in the reference. */
int32_t cn_bench_stages(cn_context* ctx, const double* dip_pos_m, const double* dip_moment, int64_t n_dip,
                        const double* center_mm, const double* handle_mm, double distance_mm, int32_t k,
                        int32_t recovery, int32_t reps, double* ms3);
/* G1 = A^T B1 and, with B2 non-null, G2 = A^T B2 for (n, K) row-major FP32 blocks A, B1, B2 in
 * device memory, written to `out` in device memory (K*K FP64 values, then K*K more with B2): the
 * kernels and the fixed-order reduction of cn_k_gram, launched on `stream` (a hipStream_t; null
 * for the default stream) without waiting for them. No counterpart in the reference: it is for
 * callers that cannot allocate (code under stream capture, tight serving loops). It calls no
 * hipMalloc, hipFree or allocating library routine: its partial sums live in a file-scope device
 * array of 512 x 2 x 32 x 32 doubles (8 MiB) that each GPU it runs on holds until the process
 * ends. K must be a compiled width (1, 2, 4, 8, 16, 32); any n >= 1. Every call on one GPU shares
 * that array, so calls that overlap on one GPU race. */
int32_t cn_k_gram_noalloc(int64_t n, int32_t K, const float* A, const float* B1, const float* B2, double* out,
                          void* stream);
/* Graph-capture failure handling (the reference's GraphCache falls back to direct execution when a
 * capture fails, solver/amg/solver.hpp:88-122): the next `n` captures of the block iteration
 * report failure, so the solves that needed them launch their iterations directly (and the next
 * solve tries to capture again). A test hook; the results are those of the captured path. */
int32_t cn_debug_fail_captures(cn_context* ctx, int32_t n);
/* Captured block-iteration executables the context holds (in use and retired): stays bounded
 * however many blocks of whichever operator mode are solved. A test hook. */
int32_t cn_graph_execs(const cn_context* ctx, int32_t* out);

#ifdef __cplusplus
}
#endif

#endif /* CUNIBS_AMD_H */
