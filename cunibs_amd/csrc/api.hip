// The C ABI (include/cunibs_amd.h): the per-mesh context, the block solve, and kernel-level
// entry points. No exception crosses the boundary; each call returns a status code and leaves
// its message in a thread-local buffer.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cunibs_amd.h"
#include "coil.hpp"
#include "prims.hpp"
#include "recovery.hpp"
#include "mlp.hpp"
#include "uq.hpp"
#include "warm.hpp"
#include "solve.hpp"

namespace {

thread_local std::string g_error;

template <typename F>
int32_t guarded(F&& f) {
    try {
        f();
        g_error.clear();
        return CN_OK;
    } catch (const cn::InvalidArgument& e) {
        g_error = e.what();
        return CN_EINVAL;
    } catch (const cn::NotConverged& e) {
        g_error = e.what();
        return CN_ENOCONV;
    } catch (const cn::HipError& e) {
        g_error = e.what();
        return CN_EHIP;
    } catch (const std::exception& e) {
        g_error = e.what();
        return CN_EINTERNAL;
    } catch (...) {
        g_error = "unknown error";
        return CN_EINTERNAL;
    }
}

// fem/assembly.py:15-26.
const double kDefaultSigma[11] = {NAN, 0.126, 0.275, 1.654, 0.010, 0.465, 0.5, 0.008, 0.025, 0.6, 0.16};
constexpr double kMu0Over4Pi = 1e-7;  // fem/placement.py:16
// Relative Schur pivot below which a direction of a sweep's Galerkin guess basis drops out: the
// basis's A-Gram comes from right-hand sides that match A X only to the solve tolerance.
constexpr double kGuessDropTol = 1e-6;

int device_cu_count() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 256;
    return cus > 0 ? cus : 256;
}

}  // namespace

struct cn_context {
    hipStream_t s = nullptr;
    cn_options opts{};
    std::int64_t n = 0, m = 0, n_skin = 0;
    int ground = 0;
    int n_cu = 256;
    cn::DeviceBuffer<double> nodes, bary, skin_a, skin_b, skin_c, skin_n;
    cn::DeviceBuffer<std::int32_t> tets, tags, skin, node_perm, tet_perm;
    cn::DeviceBuffer<float> g32, vols32, neg_vc;
    cn::DeviceBuffer<std::int32_t> n2c_ptr, n2c_idx;
    cn::DeviceBuffer<std::int32_t> corner_pos;  // position of corner 4 e + i in node2corner order
    // full stiffness (optional) and the reduced operator
    cn::DeviceBuffer<std::int32_t> indptr, indices;
    cn::DeviceBuffer<double> data;
    std::int64_t nnz_full = 0;
    cn::DeviceBuffer<std::int32_t> rptr, rcol;
    cn::DeviceBuffer<double> rval64;
    cn::DeviceBuffer<float> rval32;
    std::int64_t nnz_red = 0;
    // the reduced operator with its diagonal folded into FP64 row sums (differenced SpMM)
    cn::DeviceBuffer<std::int32_t> optr, ocol;
    cn::DeviceBuffer<float> oval, rowsum;
    cn::Hierarchy hier;
    std::unique_ptr<cn::BlockSolver> solver;
    // summary region
    cn::DeviceBuffer<std::int32_t> region_idx;
    std::int64_t n_region = 0;
    double region_volume = 0.0;
    // per-block scratch (grow-only)
    cn::DeviceBuffer<double> d_centers, d_handles, d_dists, d_tf, d_origins;
    cn::DeviceBuffer<std::int32_t> d_degen;
    cn::DeviceBuffer<float> d_dip, d_dadt_nodes, d_dadt_elm, d_q, d_B, d_E, d_magn, d_rhs, d_stage;
    cn::DeviceBuffer<double> d_X, d_v, d_x0;
    // Sliced ELL copies of the finest level's operator (with its diagonal, for the V-cycle) and of
    // the differenced off-diagonal operator (the block iteration's SpMM); width 0 when not built.
    cn::DeviceBuffer<int2> ell_full, ell_off;
    int ell_full_w = 0, ell_off_w = 0;
    // Missed-tolerance recovery counts: nominal hierarchy rebuilds, and preconditioners built for
    // one column's own operator (per-placement conductivities, conductivity draws).
    int precond_rebuilds = 0, matched_preconditioners = 0;
    double setup_ms[10] = {};  // host wall time of the setup phases and the last recovery build
    int layout = 0;  // cn_tuning.layout
    void set_layout(int l) {
        layout = l;
        const int max_k = l == 0 ? cn::kEllAutoMaxK : (l == 2 ? cn::kMaxBlock : 0);
        if (!hier.levels.empty()) {
            hier.levels[0].a.ell = max_k > 0 && ell_full_w > 0 ? ell_full.get() : nullptr;
            hier.levels[0].a.ell_w = ell_full_w;
            hier.levels[0].a.ell_max_k = max_k;
        }
        solver->set_off_ell(max_k > 0 && ell_off_w > 0 ? ell_off.get() : nullptr, ell_off_w, max_k);
    }
    // Pinned sources of the host-to-device copies a block issues without waiting on them: the
    // packed dipoles (coil stage, either stream) and the block's own small tables.
    cn::PinnedStage stage_coil, stage_blk;
    // Pipelined sweeps (cn_solve_sweep): the coil stages (dA/dt, element pass, RHS) of the next
    // two blocks run on s2 into their own buffer sets while the current block solves, recovers its
    // field and reduces its metrics on s.
    hipStream_t s2 = nullptr;
    int coil_blocks = 0;  // grid cap of the overlapped dA/dt (0: the full grid)
    hipEvent_t coil_done[3] = {nullptr, nullptr, nullptr};
    cn::DeviceBuffer<float> d_dadt_nodes2, d_B2, d_dadt_nodes3, d_B3;
    // ... and the field recovery and metrics of each solved block run on s3 (a worker thread drives
    // them) while the next block solves; the solution is double-buffered for that.
    hipStream_t s3 = nullptr;
    hipEvent_t solve_done[2] = {nullptr, nullptr};
    cn::DeviceBuffer<double> d_X2;
    // recovery operators, built on first use (index: CN_RECOVERY_*), and their block buffers
    std::unique_ptr<cn::RecoveryOp> recovery[4];
    cn::DeviceBuffer<float> d_slots, d_Eraw;
    int max_tag = 0;
    std::vector<double> lut;  // conductivity by tag (S/m)
    cn::WarmBasis warm;
    std::vector<double> warm_coef;  // coefficients for the next warm_start = 2 solve
    int warm_coef_k = 0;
    // conductivity uncertainty: per-tissue operators and RHS of one placement
    // Unit-conductivity decomposition of the reduced operator by tissue, K(sigma) = sum_t sigma_t K_t
    // (uq/conductivity/assembly.py:37-146), built on first use by the conductivity-UQ and the
    // per-placement-conductivity solves. Tags ascending, at most 16.
    struct Tissues {
        bool ready = false;
        std::vector<int> tags;
        cn::DeviceBuffer<double> vals, off, rowsum;  // (T, nnz_red), (T, nnz_off), (T, n_red) FP64
    } tissue;
    // Device buffers of the per-block region summary (grow-only, reused by every block).
    cn::MetricsWorkspace metrics_ws;
    // Phase timing of the blocks solved while in-solve timing is on (HIP events on the stream,
    // outside the graphs): coil frames, dA/dt, element pass + RHS, initial guess, block PCG, field
    // reconstruction, outputs + region metrics.
    static constexpr int kPhases = 7;
    hipEvent_t phase_ev[kPhases + 1] = {};
    double phase_ms[kPhases] = {};
    std::int64_t timed_blocks = 0, timed_placements = 0;
    void mark_phase(int i, hipStream_t st) {
        if (!solver->timing_on()) return;
        if (phase_ev[0] == nullptr)
            for (auto& e : phase_ev) cn::check(hipEventCreate(&e), "phase event");
        cn::check(hipEventRecord(phase_ev[i], st), "phase event");
    }
    // after the block's final sync: add the phase spans
    void close_phases(int k) {
        if (!solver->timing_on() || phase_ev[0] == nullptr) return;
        cn::check(hipEventSynchronize(phase_ev[kPhases]), "phase event");
        for (int i = 0; i < kPhases; ++i) {
            float ms = 0.f;
            cn::check(hipEventElapsedTime(&ms, phase_ev[i], phase_ev[i + 1]), "phase elapsed");
            phase_ms[i] += ms;
        }
        ++timed_blocks;
        timed_placements += k;
    }
    // Per-placement conductivities of the current block: (K, T) sigma, combined operator.
    cn::DeviceBuffer<double> blk_sigma;
    cn::DeviceBuffer<float> blk_sigma_tag, blk_off, blk_rowsum;
    cn::MultiOperator blk_mop;
    struct Uq {
        bool ready = false;
        cn::DeviceBuffer<float> rhs;                  // (T, n_red)
        cn::DeviceBuffer<float> dadt_nodes1;  // the placement's nodal dA/dt, one column
        cn::DeviceBuffer<double> sigma;               // (K, T)
        cn::DeviceBuffer<float> off_k, rowsum_k;      // (nnz_off, K), (n_red, K)
        cn::DeviceBuffer<double> sum, sumsq;          // |E| moments over every sample solved
        std::int64_t count = 0;
        // Initial guess of the draws: x_nom (n_red), sensitivities S (n_red, S_width) and the map
        // from a draw's sigma to its Galerkin coefficients, M (T x T, rows of dropped directions 0).
        cn::DeviceBuffer<double> x_nom, S;
        int S_width = 0;
        std::vector<double> guess_M;
        int nominal_iterations = 0, sensitivity_iterations = 0, guess_rank = 0;
        double transform[16] = {};
        cn::MultiOperator mop;
    } uq;

    cn::CsrView reduced_view() const {
        return {rptr.get(), rcol.get(), rval32.get(), n - 1, n - 1, nnz_red};
    }
    ~cn_context() {
        for (auto& e : phase_ev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : coil_done)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : solve_done)
            if (e) (void)hipEventDestroy(e);
        solver.reset();
        if (s3) (void)hipStreamDestroy(s3);
        if (s2) (void)hipStreamDestroy(s2);
        if (s) (void)hipStreamDestroy(s);
    }
};

namespace {

__global__ void region_index_kernel(const std::int32_t* __restrict__ tags, std::int64_t m, int tag,
                                    std::int32_t* __restrict__ flag) {
    const std::int64_t e = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < m) flag[e] = tags[e] == tag ? 1 : 0;
}

__global__ void region_compact_kernel(const std::int32_t* __restrict__ flag, const std::int32_t* __restrict__ pos,
                                      std::int64_t m, std::int32_t* __restrict__ idx) {
    const std::int64_t e = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e < m && flag[e]) idx[pos[e]] = static_cast<std::int32_t>(e);
}

__global__ void region_volume_kernel(const std::int32_t* __restrict__ idx, std::int64_t nr,
                                     const float* __restrict__ vols, double* __restrict__ out) {
    // One block, fixed order: thread strides then a tree.
    __shared__ double s[cn::kBlock];
    double acc = 0.0;
    for (std::int64_t j = threadIdx.x; j < nr; j += cn::kBlock) acc += static_cast<double>(vols[idx[j]]);
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = cn::kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}

// Row r of the differenced operator: its off-diagonal entries in order, and the FP64 row sum
// (diagonal included) rounded once to FP32.
__global__ void offdiag_kernel(const std::int32_t* __restrict__ rp, const std::int32_t* __restrict__ ci,
                               const double* __restrict__ v, std::int64_t n, std::int32_t* __restrict__ optr,
                               std::int32_t* __restrict__ ocol, float* __restrict__ oval, float* __restrict__ rowsum,
                               int* __restrict__ missing) {
    const std::int64_t r = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r > n) return;
    optr[r] = static_cast<std::int32_t>(rp[r] - r);
    if (r == n) return;
    std::int64_t at = rp[r] - r;
    const std::int64_t end = rp[r + 1] - (r + 1);
    double s = 0.0;
    int diag = 0;
    for (int p = rp[r]; p < rp[r + 1]; ++p) {
        s += v[p];
        if (ci[p] == r) {
            diag = 1;
            continue;
        }
        if (at < end) {  // bounded even for a row that lacks its diagonal (reported below)
            ocol[at] = ci[p];
            oval[at] = static_cast<float>(v[p]);
            ++at;
        }
    }
    rowsum[r] = static_cast<float>(s);
    if (!diag) atomicOr(missing, 1);
}

__global__ void product_kernel(const double* __restrict__ a, const double* __restrict__ b, std::int64_t n,
                               double* __restrict__ o) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) o[i] = a[i] * b[i];
}

// Column c of an (n, K) row-major block into a vector, and back.
template <typename T>
__global__ void column_get_kernel(const T* __restrict__ blk, std::int64_t n, int K, int c, T* __restrict__ v) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) v[i] = blk[i * K + c];
}

template <typename T>
__global__ void column_put_kernel(const T* __restrict__ v, std::int64_t n, int K, int c, T* __restrict__ blk) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) blk[i * K + c] = v[i];
}

// out64[i] = sum_t sigma[t] in[t * count + i] (FP64, tissues in order), out32 the same rounded.
__global__ void combine64_kernel(const double* __restrict__ in, std::int64_t count, int T,
                                 const double* __restrict__ sigma, double* __restrict__ out64,
                                 float* __restrict__ out32) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    double acc = 0.0;
    for (int q = 0; q < T; ++q) acc = fma(sigma[q], in[q * count + i], acc);
    out64[i] = acc;
    out32[i] = static_cast<float>(acc);
}

void build_context(cn_context& c, const double* nodes_mm, std::int64_t n, const std::int32_t* tet_nodes,
                   const std::int32_t* tet_tags, std::int64_t m, const std::int32_t* skin_tris, std::int64_t n_skin) {
    using namespace cn;
    const auto t_entry = std::chrono::steady_clock::now();
    require(n >= 2, "the mesh needs at least two nodes");
    require(m >= 1, "the mesh needs at least one tetrahedron");
    require(n < (1ll << 31) && 4 * m < (1ll << 31), "mesh too large for int32 indexing");
    // Branch-free scans (each folds its whole array, then one check).
    bool finite = true, tets_ok = true, skin_ok = true;
    for (std::int64_t i = 0; i < 3 * n; ++i) finite &= std::isfinite(nodes_mm[i]);
    const auto nn = static_cast<std::uint32_t>(n);
    for (std::int64_t i = 0; i < 4 * m; ++i) tets_ok &= static_cast<std::uint32_t>(tet_nodes[i]) < nn;
    for (std::int64_t i = 0; i < 3 * n_skin; ++i) skin_ok &= static_cast<std::uint32_t>(skin_tris[i]) < nn;
    require(finite, "node coordinates must be finite");
    require(tets_ok, "tetrahedron references a node outside the mesh");
    require(skin_ok, "skin triangle references a node outside the mesh");
    const hipStream_t s = c.s;
    auto t_phase = t_entry;
    int phase = 0;
    // Host wall time of each setup phase, ending at a wait on the stream (so it covers the phase's
    // device work): cn_context_setup_times.
    auto lap = [&] {
        sync(s, "setup phase");
        const auto now = std::chrono::steady_clock::now();
        c.setup_ms[phase++] = std::chrono::duration<double, std::milli>(now - t_phase).count();
        t_phase = now;
    };
    c.n = n;
    c.m = m;
    c.n_skin = n_skin;
    c.n_cu = device_cu_count();

    lap();  // 0: argument checks
    DeviceBuffer<double> nodes_in = upload(nodes_mm, 3 * n, s, "nodes");
    DeviceBuffer<std::int32_t> tets_in = upload(tet_nodes, 4 * m, s, "tets");
    DeviceBuffer<std::int32_t> tags_in = upload(tet_tags, m, s, "tags");
    c.skin.reset(std::max<std::int64_t>(3 * n_skin, 1), "skin");
    h2d(c.skin.get(), skin_tris, 3 * n_skin, s);
    lap();  // 1: uploads
    c.node_perm.reset(n, "node perm");
    c.tet_perm.reset(m, "tet perm");
    DeviceBuffer<std::int32_t> inverse(n, "inverse perm");
    spatial_order(nodes_in.get(), n, tets_in.get(), m, c.node_perm.get(), c.tet_perm.get(), inverse.get(), s);
    c.nodes.reset(3 * n, "nodes");
    c.tets.reset(4 * m, "tets");
    c.tags.reset(m, "tags");
    apply_spatial_order(nodes_in.get(), tets_in.get(), tags_in.get(), n, m, c.node_perm.get(), c.tet_perm.get(),
                        inverse.get(), c.nodes.get(), c.tets.get(), c.tags.get(), c.skin.get(), n_skin, s);
    nodes_in.release();
    tets_in.release();
    tags_in.release();

    lap();  // 2: Morton order
    DeviceBuffer<double> g64(12 * m, "g64"), vols64(m, "vols64"), cond(m, "cond");
    p1_gradients(c.nodes.get(), c.tets.get(), m, g64.get(), vols64.get(), s);
    const int bad = count_bad_volumes(vols64.get(), m, s);
    if (bad) {
        throw InvalidArgument(std::to_string(bad) + (bad == 1 ? " tetrahedron has" : " tetrahedra have") +
                              " zero or non-finite volume");
    }
    std::vector<double> lut;
    if (c.opts.conductivity != nullptr && c.opts.n_conductivity > 0)
        lut.assign(c.opts.conductivity, c.opts.conductivity + c.opts.n_conductivity);
    else
        lut.assign(kDefaultSigma, kDefaultSigma + 11);
    const int unknown = conductivity(c.tags.get(), m, lut.data(), static_cast<int>(lut.size()), cond.get(), s);
    if (unknown) throw InvalidArgument(std::to_string(unknown) + " tetrahedra have unknown tags");
    c.max_tag = static_cast<int>(lut.size()) - 1;  // every tag indexes the table
    c.opts.conductivity = nullptr;  // the caller's table need not outlive this call; keep a copy
    c.lut = lut;

    lap();  // 3: gradients, volume and conductivity checks
    c.n2c_ptr.reset(n + 1, "n2c ptr");
    c.n2c_idx.reset(4 * m, "n2c idx");
    node2corner(c.tets.get(), m, 4, n, c.n2c_ptr.get(), c.n2c_idx.get(), s);
    c.corner_pos.reset(4 * m, "corner positions");
    corner_positions(c.n2c_idx.get(), 4 * m, c.corner_pos.get(), s);
    DeviceBuffer<std::int32_t> indptr, indices;
    const std::int64_t nnz = stiffness_pattern(c.tets.get(), c.n2c_ptr.get(), c.n2c_idx.get(), n, indptr, indices, s);
    DeviceBuffer<double> data(std::max<std::int64_t>(nnz, 1), "stiffness");
    DeviceBuffer<double> scale(m, "vol*sigma");  // fem/assembly.py:131, FP64 product
    product_kernel<<<grid_for(m), kBlock, 0, s>>>(vols64.get(), cond.get(), m, scale.get());
    lap();  // 4: node2corner and the stiffness pattern
    stiffness_values(g64.get(), scale.get(), c.tets.get(), c.n2c_ptr.get(), c.n2c_idx.get(), n, indptr.get(),
                     indices.get(), data.get(), s);
    c.nnz_full = nnz;
    c.ground = lowest_z_node(c.nodes.get(), n, s);
    c.nnz_red = reduce_grounded(indptr.get(), indices.get(), data.get(), n, c.ground, c.rptr, c.rcol, c.rval64,
                                c.rval32, s);
    if (c.opts.keep_stiffness) {
        c.indptr = std::move(indptr);
        c.indices = std::move(indices);
        c.data = std::move(data);
    } else {
        indptr.release();
        indices.release();
        data.release();
    }

    c.g32.reset(12 * m, "g32");
    c.vols32.reset(m, "vols32");
    c.neg_vc.reset(m, "neg_vc");
    c.bary.reset(3 * m, "barycentres");
    resident_arrays(g64.get(), vols64.get(), cond.get(), c.nodes.get(), c.tets.get(), m, c.g32.get(), c.vols32.get(),
                    c.neg_vc.get(), c.bary.get(), s);
    g64.release();
    vols64.release();
    cond.release();
    scale.release();

    // AMG hierarchy on the FP32 reduced operator.
    AmgParams prm;
    prm.min_coarse_rows = c.opts.min_coarse_rows;
    prm.max_levels = c.opts.max_levels;
    prm.rounds = c.opts.rounds;
    prm.degree = c.opts.degree;
    prm.lower_ratio = c.opts.lower_ratio;
    lap();  // 5: stiffness values, grounding, resident arrays
    c.hier = build_hierarchy(c.reduced_view(), prm, s);
    lap();  // 6: AMG hierarchy
    // Differenced form: every reduced row keeps its diagonal (a node lies in its own elements), so
    // dropping it shifts row r's start by exactly r.
    const std::int64_t nr = n - 1;
    c.optr.reset(nr + 1, "offdiag ptr");
    c.ocol.reset(std::max<std::int64_t>(c.nnz_red - nr, 1), "offdiag col");
    c.oval.reset(std::max<std::int64_t>(c.nnz_red - nr, 1), "offdiag val");
    c.rowsum.reset(nr, "row sums");
    DeviceBuffer<int> missing(1, "missing diagonal");
    check(hipMemsetAsync(missing.get(), 0, sizeof(int), s), "memset");
    offdiag_kernel<<<grid_for(nr + 1), kBlock, 0, s>>>(c.rptr.get(), c.rcol.get(), c.rval64.get(), nr, c.optr.get(),
                                                       c.ocol.get(), c.oval.get(), c.rowsum.get(), missing.get());
    int hmiss = 0;
    d2h(&hmiss, missing.get(), 1, s);
    sync(s, "differenced operator");
    require(hmiss == 0, "a row of the reduced operator has no diagonal entry");
    Operator op;
    op.a32 = c.reduced_view();
    op.off = {c.optr.get(), c.ocol.get(), c.oval.get(), nr, nr, c.nnz_red - nr};
    op.rowsum = c.rowsum.get();
    op.val64 = c.rval64.get();
    c.solver = std::make_unique<BlockSolver>(op, c.hier, s);
    c.ell_full_w = build_ell(c.reduced_view(), c.ell_full, s);
    c.ell_off_w = build_ell(op.off, c.ell_off, s);
    sync(s, "ell copies");
    lap();  // 7: differenced operator, solver, ELL copies
    c.set_layout(0);

    // Skin surface for the placement scan.
    c.skin_a.reset(std::max<std::int64_t>(3 * n_skin, 1), "skin a");
    c.skin_b.reset(std::max<std::int64_t>(3 * n_skin, 1), "skin b");
    c.skin_c.reset(std::max<std::int64_t>(3 * n_skin, 1), "skin c");
    c.skin_n.reset(std::max<std::int64_t>(3 * n_skin, 1), "skin normals");
    if (n_skin > 0) {
        triangle_vertices(c.nodes.get(), c.skin.get(), n_skin, c.skin_a.get(), c.skin_b.get(), c.skin_c.get(), s);
        skin_normals(c.nodes.get(), c.skin.get(), n_skin, n, c.skin_n.get(), s);
    }

    // Summary region (metrics.py:246-259).
    DeviceBuffer<std::int32_t> flag(m + 1, "region flag"), pos(m + 1, "region pos");
    check(hipMemsetAsync(flag.get() + m, 0, sizeof(std::int32_t), s), "memset");
    region_index_kernel<<<grid_for(m), kBlock, 0, s>>>(c.tags.get(), m, c.opts.region_tag, flag.get());
    exclusive_scan_i32(flag.get(), pos.get(), m + 1, s);
    std::int32_t n_reg = 0;
    d2h(&n_reg, pos.get() + m, 1, s);
    sync(s, "region size");
    c.n_region = n_reg;
    c.region_idx.reset(std::max<std::int64_t>(n_reg, 1), "region idx");
    region_compact_kernel<<<grid_for(m), kBlock, 0, s>>>(flag.get(), pos.get(), m, c.region_idx.get());
    DeviceBuffer<double> vol(1, "region volume");
    region_volume_kernel<<<1, kBlock, 0, s>>>(c.region_idx.get(), n_reg, c.vols32.get(), vol.get());
    d2h(&c.region_volume, vol.get(), 1, s);
    sync(s, "context build");
    lap();  // 8: skin surface, region slice
}

// Dipoles of one placement in its coil frame, FP64 as fem/placement.py:294-313, then FP32 pack
// [s | m | m x s | |s|^2] centred at the coil origin.
// The direction every dipole moment lies along, if there is one (coil models of planar windings
// put all their moments on the coil normal): |m_j x d| <= 1e-12 |m_j| for all j. Writes the unit
// direction of the largest moment.
bool common_moment_direction(const double* mom, std::int64_t n_dip, double* d) {
    std::int64_t big = 0;
    double best = 0.0;
    for (std::int64_t j = 0; j < n_dip; ++j) {
        const double q = mom[3 * j] * mom[3 * j] + mom[3 * j + 1] * mom[3 * j + 1] + mom[3 * j + 2] * mom[3 * j + 2];
        if (q > best) {
            best = q;
            big = j;
        }
    }
    if (!(best > 0.0)) return false;
    const double nb = std::sqrt(best);
    for (int r = 0; r < 3; ++r) d[r] = mom[3 * big + r] / nb;
    for (std::int64_t j = 0; j < n_dip; ++j) {
        const double* m = mom + 3 * j;
        const double cx = m[1] * d[2] - m[2] * d[1], cy = m[2] * d[0] - m[0] * d[2], cz = m[0] * d[1] - m[1] * d[0];
        const double mm = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
        if (cx * cx + cy * cy + cz * cz > 1e-24 * mm) return false;
    }
    return true;
}

// Per placement, for the parallel-moment N-body: -2 s (3), |s|^2, mu, mu s (3) in FP32 with s
// relative to the coil origin (as pack_dipoles), and the rotated direction.
void pack_dipoles_parallel(const double* tf, const double* pos_m, const double* mom, std::int64_t n_dip,
                           const double* dir, float* out, double* origin, float* dir_out) {
    double rot[3][3], trans[3];
    for (int r = 0; r < 3; ++r) {
        for (int q = 0; q < 3; ++q) rot[r][q] = tf[4 * r + q];
        trans[r] = tf[4 * r + 3];
    }
    const double o[3] = {trans[0] * 1e-3, trans[1] * 1e-3, trans[2] * 1e-3};
    for (int r = 0; r < 3; ++r) origin[r] = o[r];
    for (int r = 0; r < 3; ++r) {
        double acc = 0.0;
        for (int q = 0; q < 3; ++q) acc += dir[q] * rot[r][q];
        dir_out[r] = static_cast<float>(acc);
    }
    for (std::int64_t j = 0; j < n_dip; ++j) {
        double sd[3];
        for (int r = 0; r < 3; ++r) {
            double acc = 0.0;
            for (int q = 0; q < 3; ++q) acc += pos_m[3 * j + q] * 1e3 * rot[r][q];
            sd[r] = (acc + trans[r]) * 1e-3 - o[r];
        }
        const double mu = mom[3 * j] * dir[0] + mom[3 * j + 1] * dir[1] + mom[3 * j + 2] * dir[2];
        float* p = out + 8 * j;
        const float sf[3] = {static_cast<float>(sd[0]), static_cast<float>(sd[1]), static_cast<float>(sd[2])};
        for (int r = 0; r < 3; ++r) {
            p[r] = -2.0f * sf[r];
            p[5 + r] = static_cast<float>(mu * sd[r]);
        }
        p[3] = sf[0] * sf[0] + sf[1] * sf[1] + sf[2] * sf[2];
        p[4] = static_cast<float>(mu);
    }
}

// dA/dt at the nodes for k placements (frames tf, k x 16) into out (n, 3, K): the parallel-moment
// kernel when the coil allows it, the general one otherwise.
void dadt_for_frames(cn_context& c, const double* tf, int k, int K, const double* pos_m, const double* mom,
                     std::int64_t n_dip, float scale, float* out, hipStream_t s, int max_blocks = 0);

void pack_dipoles(const double* tf, const double* pos_m, const double* mom, std::int64_t n_dip, float* out,
                  double* origin) {
    double rot[3][3], trans[3];
    for (int r = 0; r < 3; ++r) {
        for (int q = 0; q < 3; ++q) rot[r][q] = tf[4 * r + q];
        trans[r] = tf[4 * r + 3];
    }
    const double o[3] = {trans[0] * 1e-3, trans[1] * 1e-3, trans[2] * 1e-3};
    for (int r = 0; r < 3; ++r) origin[r] = o[r];
    for (std::int64_t j = 0; j < n_dip; ++j) {
        double s[3], mm[3];
        for (int r = 0; r < 3; ++r) {
            double acc = 0.0, am = 0.0;
            for (int q = 0; q < 3; ++q) {
                acc += pos_m[3 * j + q] * 1e3 * rot[r][q];
                am += mom[3 * j + q] * rot[r][q];
            }
            s[r] = (acc + trans[r]) * 1e-3 - o[r];
            mm[r] = am;
        }
        const double mxs[3] = {mm[1] * s[2] - mm[2] * s[1], mm[2] * s[0] - mm[0] * s[2], mm[0] * s[1] - mm[1] * s[0]};
        float* p = out + 10 * j;
        for (int r = 0; r < 3; ++r) {
            p[r] = static_cast<float>(s[r]);
            p[3 + r] = static_cast<float>(mm[r]);
            p[6 + r] = static_cast<float>(mxs[r]);
        }
        p[9] = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];  // from the FP32 positions, as the reference
    }
}

void dadt_for_frames(cn_context& c, const double* tf, int k, int K, const double* pos_m, const double* mom,
                     std::int64_t n_dip, float scale, float* out, hipStream_t s, int max_blocks) {
    using namespace cn;
    // The packed dipoles (and the moment directions) and the origins are built in the pinned
    // stage and copied from there: this function returns before the copies have run.
    double dir[3];
    const bool parallel = common_moment_direction(mom, n_dip, dir);
    const std::size_t n_pack = static_cast<std::size_t>((parallel ? 8 : 10) * n_dip * k);
    const std::size_t n_dirs = parallel ? 3 * static_cast<std::size_t>(k) : 0;
    const std::size_t f_bytes = (n_pack + n_dirs) * sizeof(float);
    const std::size_t o_off = (f_bytes + 15) / 16 * 16;
    char* host = static_cast<char*>(c.stage_coil.fill(o_off + 3 * sizeof(double) * k));
    float* dip = reinterpret_cast<float*>(host);
    double* origins = reinterpret_cast<double*>(host + o_off);
    for (int i = 0; i < k; ++i) {
        if (parallel)
            pack_dipoles_parallel(tf + 16 * i, pos_m, mom, n_dip, dir, dip + 8 * n_dip * i, origins + 3 * i,
                                  dip + n_pack + 3 * i);
        else
            pack_dipoles(tf + 16 * i, pos_m, mom, n_dip, dip + 10 * n_dip * i, origins + 3 * i);
    }
    c.d_dip.ensure(n_pack + n_dirs, "dipoles");
    c.d_origins.ensure(3 * k, "origins");
    c.stage_coil.send(c.d_dip.get(), 0, f_bytes, s);
    c.stage_coil.send(c.d_origins.get(), o_off, 3 * sizeof(double) * k, s);
    c.stage_coil.done(s);
    if (parallel)
        dadt_nbody_parallel(c.d_dip.get(), static_cast<int>(n_dip), c.nodes.get(), c.n, c.d_origins.get(),
                            c.d_dip.get() + n_pack, k, K, scale, out, s, max_blocks);
    else
        dadt_nbody(c.d_dip.get(), static_cast<int>(n_dip), c.nodes.get(), c.n, c.d_origins.get(), k, K, scale, out, s);
}

void place_block(cn_context& c, const double* centers, const double* handles, const double* dists, int k,
                 double* tf_host, std::int32_t* degen_host) {
    using namespace cn;
    const hipStream_t s = c.s;
    require(c.n_skin > 0, "the mesh has no scalp surface (tag 1005) to place a coil on");
    for (int i = 0; i < 3 * k; ++i)
        require(std::isfinite(centers[i]) && (handles == nullptr || std::isfinite(handles[i])),
                "center_mm and handle_mm must be finite 3-vectors");
    for (int i = 0; i < k; ++i) require(std::isfinite(dists[i]), "distance_mm must be finite");
    c.d_centers.ensure(3 * k, "centres");
    c.d_dists.ensure(k, "distances");
    c.d_tf.ensure(16 * k, "transforms");
    c.d_degen.ensure(k, "degenerate");
    h2d(c.d_centers.get(), centers, 3 * k, s);
    h2d(c.d_dists.get(), dists, k, s);
    const double* dh = nullptr;
    if (handles != nullptr) {
        c.d_handles.ensure(3 * k, "handles");
        h2d(c.d_handles.get(), handles, 3 * k, s);
        dh = c.d_handles.get();
    }
    place_transforms(c.d_centers.get(), dh, c.d_dists.get(), k, c.skin_a.get(), c.skin_b.get(), c.skin_c.get(),
                     c.skin_n.get(), static_cast<int>(c.n_skin), c.d_tf.get(), c.d_degen.get(), c.n_cu, s);
    d2h(tf_host, c.d_tf.get(), 16 * k, s);
    d2h(degen_host, c.d_degen.get(), k, s);
    sync(s, "placement");
}

// Per-tissue unit-conductivity operators on the reduced pattern (uq/conductivity/assembly.py:37-146):
// K(sigma) = sum_t sigma_t K_t, as FP64 values, differenced off-diagonals and row sums.
void ensure_tissues(cn_context& c) {
    using namespace cn;
    if (c.tissue.ready) return;
    const hipStream_t s = c.s;
    const std::int64_t n = c.n, m = c.m, nr = n - 1, nnz_off = c.nnz_red - nr;
    // tags present in the mesh, ascending
    std::vector<std::int32_t> tags(m);
    d2h(tags.data(), c.tags.get(), m, s);
    sync(s, "tags");
    std::vector<int> present;
    {
        std::vector<char> seen(c.max_tag + 1, 0);
        for (auto t : tags) seen[t] = 1;
        for (int t = 0; t <= c.max_tag; ++t)
            if (seen[t]) present.push_back(t);
    }
    require(static_cast<int>(present.size()) <= kMaxTissues, "more than 16 tissues");
    const int T = static_cast<int>(present.size());
    auto& u = c.tissue;
    u.tags = present;
    u.vals.reset(static_cast<std::size_t>(T) * c.nnz_red, "tissue values");
    u.off.reset(static_cast<std::size_t>(T) * std::max<std::int64_t>(nnz_off, 1), "tissue off-diagonals");
    u.rowsum.reset(static_cast<std::size_t>(T) * nr, "tissue row sums");
    // Stiffness per tissue: gradients and pattern again (the context keeps neither in FP64).
    DeviceBuffer<double> g64(12 * m, "g64"), vols64(m, "vols64"), scale(m, "scale");
    p1_gradients(c.nodes.get(), c.tets.get(), m, g64.get(), vols64.get(), s);
    DeviceBuffer<std::int32_t> indptr, indices;
    const std::int64_t nnz = stiffness_pattern(c.tets.get(), c.n2c_ptr.get(), c.n2c_idx.get(), n, indptr, indices, s);
    DeviceBuffer<double> data(std::max<std::int64_t>(nnz, 1), "tissue stiffness");
    DeviceBuffer<std::int32_t> rp, rc;
    DeviceBuffer<double> rv64;
    DeviceBuffer<float> rv32;
    for (int q = 0; q < T; ++q) {
        tissue_scale(vols64.get(), c.tags.get(), m, present[q], scale.get(), s);
        stiffness_values(g64.get(), scale.get(), c.tets.get(), c.n2c_ptr.get(), c.n2c_idx.get(), n, indptr.get(),
                         indices.get(), data.get(), s);
        const std::int64_t nz = reduce_grounded(indptr.get(), indices.get(), data.get(), n, c.ground, rp, rc, rv64, rv32, s);
        require(nz == c.nnz_red, "tissue operator pattern differs from the reduced operator");
        d2d(u.vals.get() + static_cast<std::int64_t>(q) * c.nnz_red, rv64.get(), c.nnz_red, s);
        split_offdiag(c.rptr.get(), c.rcol.get(), rv64.get(), nr, u.off.get() + static_cast<std::int64_t>(q) * nnz_off,
                      u.rowsum.get() + static_cast<std::int64_t>(q) * nr, s);
    }
    sync(s, "tissue operators");
    u.ready = true;
}

// Pseudo-inverse solve of a small SPD matrix by diagonally pivoted Cholesky on its unit-diagonal
// scaling: directions whose pivot falls below `drop` are dropped. Returns the kept rank and
// writes X (T x P) = E^+ C for C (T x P), zero rows for dropped directions.
int pivoted_solve(const std::vector<double>& E, int T, const std::vector<double>& C, int P, double drop,
                  std::vector<double>& X) {
    std::vector<double> d(T, 0.0), a(static_cast<std::size_t>(T) * T, 0.0);
    for (int i = 0; i < T; ++i) d[i] = E[i * T + i] > 0.0 ? 1.0 / std::sqrt(E[i * T + i]) : 0.0;
    for (int i = 0; i < T; ++i)
        for (int j = 0; j < T; ++j) a[i * T + j] = d[i] * E[i * T + j] * d[j];
    std::vector<int> order;
    std::vector<char> used(T, 0);
    std::vector<double> L(static_cast<std::size_t>(T) * T, 0.0);  // rows: kept directions in pivot order
    for (int step = 0; step < T; ++step) {
        int best = -1;
        double bv = drop;
        for (int i = 0; i < T; ++i) {
            if (used[i] || d[i] == 0.0) continue;
            double v = a[i * T + i];
            for (int r = 0; r < step; ++r) v -= L[r * T + i] * L[r * T + i];
            if (v > bv) {
                bv = v;
                best = i;
            }
        }
        if (best < 0) break;
        used[best] = 1;
        const double piv = std::sqrt(bv);
        for (int i = 0; i < T; ++i) {
            if (used[i] && i != best) continue;
            double v = a[best * T + i];
            for (int r = 0; r < step; ++r) v -= L[r * T + best] * L[r * T + i];
            L[step * T + i] = v / piv;
        }
        order.push_back(best);
    }
    const int R = static_cast<int>(order.size());
    X.assign(static_cast<std::size_t>(T) * P, 0.0);
    // Solve (L_k L_k^T) y = D C restricted to the kept directions, in pivot coordinates.
    for (int p = 0; p < P; ++p) {
        std::vector<double> y(R);
        for (int r = 0; r < R; ++r) {
            double v = d[order[r]] * C[order[r] * P + p];
            for (int q = 0; q < r; ++q) v -= L[q * T + order[r]] * y[q];
            y[r] = v / L[r * T + order[r]];
        }
        for (int r = R - 1; r >= 0; --r) {
            double v = y[r];
            for (int q = r + 1; q < R; ++q) v -= L[r * T + order[q]] * y[q];
            y[r] = v / L[r * T + order[r]];
        }
        for (int r = 0; r < R; ++r) X[order[r] * P + p] = d[order[r]] * y[r];
    }
    return R;
}

// The reference's initial guess of every draw (uq/conductivity/run.py:100-200): solve the nominal
// system once for x_nom; the sensitivities S_t solve A_nom S_t = b_t - K_t x_nom (tolerance 1e-3,
// run.py:46); a draw's residual at x_nom is r0 = sum_t sigma_t (b_t - K_t x_nom) since both the
// matrix and the RHS are linear in sigma, so the Galerkin correction on span(S) is
// x0 = x_nom + S E^+ (S^T R) sigma with E = S^T A_nom S. S^T R and E are formed once; a draw's
// coefficients are a T x T product on the host.
void uq_initial_guess(cn_context& c) {
    using namespace cn;
    auto& u = c.uq;
    const hipStream_t s = c.s;
    const std::int64_t nr = c.n - 1;
    const int T = static_cast<int>(c.tissue.tags.size());
    SolveOptions opt;
    // nominal solve, one column
    std::vector<double> sig0(T);
    for (int t = 0; t < T; ++t) sig0[t] = c.lut[c.tissue.tags[t]];
    DeviceBuffer<double> dsig(T, "nominal sigma");
    c.stage_blk.put(dsig.get(), sig0.data(), T, s);
    c.d_B.ensure(nr * 1, "rhs block");
    combine_tissues(u.rhs.get(), nr, T, dsig.get(), 1, 1, c.d_B.get(), s);
    c.d_X.ensure(nr * 1, "solution");
    SolveStats st = c.solver->solve(c.d_B.get(), c.d_X.get(), nullptr, 1, opt);
    u.nominal_iterations = st.iterations;
    u.x_nom.reset(nr, "nominal solution");
    d2d(u.x_nom.get(), c.d_X.get(), nr, s);
    // sensitivities
    const int K = padded_width(T);
    DeviceBuffer<double> R(static_cast<std::size_t>(nr) * K, "sensitivity rhs");
    c.d_B.ensure(nr * K, "rhs block");
    sensitivity_rhs(c.rptr.get(), c.rcol.get(), c.tissue.vals.get(), c.nnz_red, u.rhs.get(), u.x_nom.get(), nr, T, K, R.get(),
                    c.d_B.get(), s);
    std::vector<double> RR(static_cast<std::size_t>(T) * T);
    gram64(R.get(), R.get(), nr, T, K, RR.data(), s);
    u.S_width = K;
    u.guess_M.assign(static_cast<std::size_t>(T) * T, 0.0);
    u.guess_rank = 0;
    for (int t = 0; t < T; ++t)
        if (!(RR[t * T + t] > 0.0)) return;  // a tissue the field does not see: x_nom alone
    SolveOptions sopt;
    sopt.tolerance = 1e-3;
    c.d_X.ensure(nr * K, "solution");
    st = c.solver->solve(c.d_B.get(), c.d_X.get(), nullptr, T, sopt);
    u.sensitivity_iterations = st.iterations;
    u.S.reset(static_cast<std::size_t>(nr) * K, "sensitivities");
    d2d(u.S.get(), c.d_X.get(), static_cast<std::size_t>(nr) * K, s);
    DeviceBuffer<double> AS(static_cast<std::size_t>(nr) * K, "A S");
    spmm64(c.rptr.get(), c.rcol.get(), c.rval64.get(), nr, u.S.get(), T, K, AS.get(), s);
    std::vector<double> E(static_cast<std::size_t>(T) * T), C(static_cast<std::size_t>(T) * T);
    gram64(u.S.get(), AS.get(), nr, T, K, E.data(), s);
    gram64(u.S.get(), R.get(), nr, T, K, C.data(), s);
    for (int i = 0; i < T; ++i)
        for (int j = 0; j < i; ++j) E[i * T + j] = E[j * T + i] = 0.5 * (E[i * T + j] + E[j * T + i]);
    u.guess_rank = pivoted_solve(E, T, C, T, 1e-12, u.guess_M);
}

void uq_prepare(cn_context& c, const double* pos_m, const double* mom, std::int64_t n_dip, const double* center,
                const double* handle, double dist, double didt) {
    using namespace cn;
    const hipStream_t s = c.s;
    const std::int64_t n = c.n, m = c.m, nr = n - 1, nnz_off = c.nnz_red - nr;
    ensure_tissues(c);
    const std::vector<int>& present = c.tissue.tags;
    const int T = static_cast<int>(present.size());
    auto& u = c.uq;
    u = cn_context::Uq{};
    u.rhs.reset(static_cast<std::size_t>(T) * nr, "tissue rhs");
    // The placement's dA/dt (one column), then the RHS of each tissue with unit conductivity.
    std::vector<double> tf(16);
    std::vector<std::int32_t> degen(1);
    place_block(c, center, handle, &dist, 1, tf.data(), degen.data());
    require(handle == nullptr || degen[0] == 0, "degenerate coil placement (handle along the scalp normal)");
    std::copy(tf.begin(), tf.end(), u.transform);
    std::vector<float> dip(static_cast<std::size_t>(10 * n_dip));
    std::vector<double> origin(3);
    pack_dipoles(tf.data(), pos_m, mom, n_dip, dip.data(), origin.data());
    DeviceBuffer<float> ddip(dip.size(), "dipoles");
    DeviceBuffer<double> dorg(3, "origin");
    h2d(ddip.get(), dip.data(), dip.size(), s);
    h2d(dorg.get(), origin.data(), 3, s);
    sync(s, "dipole upload");  // the sources are function-local
    u.dadt_nodes1.reset(3 * n, "uq dadt nodes");
    dadt_nbody(ddip.get(), static_cast<int>(n_dip), c.nodes.get(), n, dorg.get(), 1, 1,
               static_cast<float>(didt) * static_cast<float>(kMu0Over4Pi), u.dadt_nodes1.get(), s);
    DeviceBuffer<float> negvc(m, "tissue neg_vc"), q(4 * m, "corner terms");
    for (int t = 0; t < T; ++t) {
        tissue_neg_vc(c.vols32.get(), c.tags.get(), m, present[t], negvc.get(), s);
        element_pass(u.dadt_nodes1.get(), c.tets.get(), c.g32.get(), negvc.get(), m, 1, 1, nullptr, q.get(),
                     c.corner_pos.get(), s);
        node_rhs(q.get(), c.n2c_ptr.get(), n, 1, 1, c.ground, u.rhs.get() + static_cast<std::int64_t>(t) * nr, nullptr,
                 s);
    }
    u.sum.reset(m, "|E| sum");
    u.sumsq.reset(m, "|E| sum of squares");
    check(hipMemsetAsync(u.sum.get(), 0, sizeof(double) * m, s), "moments clear");
    check(hipMemsetAsync(u.sumsq.get(), 0, sizeof(double) * m, s), "moments clear");
    sync(s, "uq prepare");
    uq_initial_guess(c);
    u.ready = true;
}

const cn::RecoveryOp* ensure_recovery(cn_context& c, int mode);

cn::AmgParams amg_params(const cn_context& c) {
    cn::AmgParams prm;
    prm.min_coarse_rows = c.opts.min_coarse_rows;
    prm.max_levels = c.opts.max_levels;
    prm.rounds = c.opts.rounds;
    prm.degree = c.hier.cheb.degree;
    prm.lower_ratio = c.opts.lower_ratio;
    return prm;
}

// A preconditioner matched to one column's operator K(sigma) = sum_t sigma_t K_t: its FP64, FP32
// and differenced values on the reduced pattern, the nominal hierarchy's aggregates with the
// prolongators, Galerkin products and coarse inverse recomputed from it, and a solver over them.
struct MatchedSolve {
    cn::DeviceBuffer<double> v64;
    cn::DeviceBuffer<float> v32, off32, rs32;
    cn::Hierarchy h;
    std::unique_ptr<cn::BlockSolver> solver;
};

std::unique_ptr<MatchedSolve> matched_solve(cn_context& c, const double* sigma_row) {
    using namespace cn;
    const hipStream_t s = c.s;
    const std::int64_t nr = c.n - 1, nnz_off = c.nnz_red - nr;
    const int T = static_cast<int>(c.tissue.tags.size());
    auto ms = std::make_unique<MatchedSolve>();
    ms->v64.reset(c.nnz_red, "matched values");
    ms->v32.reset(c.nnz_red, "matched values");
    ms->off32.reset(std::max<std::int64_t>(nnz_off, 1), "matched off-diagonals");
    ms->rs32.reset(nr, "matched row sums");
    combine64_kernel<<<grid_for(c.nnz_red), kBlock, 0, s>>>(c.tissue.vals.get(), c.nnz_red, T, sigma_row,
                                                            ms->v64.get(), ms->v32.get());
    combine_tissues(c.tissue.off.get(), nnz_off, T, sigma_row, 1, 1, ms->off32.get(), s);
    combine_tissues(c.tissue.rowsum.get(), nr, T, sigma_row, 1, 1, ms->rs32.get(), s);
    const CsrView a32{c.rptr.get(), c.rcol.get(), ms->v32.get(), nr, nr, c.nnz_red};
    ms->h = build_hierarchy_like(a32, c.hier, amg_params(c), s);
    Operator op;
    op.a32 = a32;
    op.off = {c.optr.get(), c.ocol.get(), ms->off32.get(), nr, nr, nnz_off};
    op.rowsum = ms->rs32.get();
    op.val64 = ms->v64.get();
    ms->solver = std::make_unique<BlockSolver>(op, ms->h, s);
    sync(s, "matched preconditioner");
    return ms;
}

// Columns of a block solve that missed tolerance are solved again one at a time from their own
// iterate (fem/solve.py:508-552): under a rebuilt nominal hierarchy (sigma == nullptr), or, when
// the columns carry their own conductivities ((K, T) device `sigma`), each under a preconditioner
// built for its own operator and then discarded (uq/conductivity/run.py:333-342). The block's X
// and the columns' statistics are updated in place; a column that misses again stays missed.
void retry_missed(cn_context& c, cn::SolveStats& st, const cn::SolveOptions& opt, const float* B, int k, int K,
                  const double* sigma, double* X) {
    using namespace cn;
    std::vector<int> missed;
    for (int i = 0; i < k; ++i)
        if (!(st.relative_residual[i] <= opt.tolerance)) missed.push_back(i);
    if (missed.empty()) return;
    const hipStream_t s = c.s;
    const std::int64_t nr = c.n - 1;
    if (sigma == nullptr) {
        // Rebuild the V-cycle from the operator's current values (fem/solve.py:429-442).
        sync(s, "before the rebuild");
        c.hier = build_hierarchy(c.reduced_view(), amg_params(c), s);
        c.set_layout(c.layout);
        c.solver->invalidate();
        ++c.precond_rebuilds;
    }
    DeviceBuffer<float> b1(nr, "retry rhs");
    DeviceBuffer<double> x0(nr, "retry guess"), x1(nr, "retry solution");
    const int T = static_cast<int>(c.tissue.tags.size());
    for (int col : missed) {
        column_get_kernel<float><<<grid_for(nr), kBlock, 0, s>>>(B, nr, K, col, b1.get());
        column_get_kernel<double><<<grid_for(nr), kBlock, 0, s>>>(X, nr, K, col, x0.get());
        SolveStats r;
        if (sigma != nullptr) {
            auto ms = matched_solve(c, sigma + static_cast<std::int64_t>(col) * T);
            ++c.matched_preconditioners;
            r = ms->solver->solve(b1.get(), x1.get(), x0.get(), 1, opt);
        } else {
            r = c.solver->solve(b1.get(), x1.get(), x0.get(), 1, opt);
        }
        column_put_kernel<double><<<grid_for(nr), kBlock, 0, s>>>(x1.get(), nr, K, col, X);
        st.relative_residual[col] = r.relative_residual[0];
        st.column_iterations[col] = (st.column_iterations[col] ? st.column_iterations[col] : st.iterations) +
                                    r.iterations;
    }
    sync(s, "retried columns");
}

// k <= 16 conductivity samples (sigma: k x T, T in the order of the tags uq_prepare found) of the
// prepared placement, as K independent CG chains under the nominal preconditioner.
void uq_solve(cn_context& c, const double* sigma, int k, const cn_solve_options* sopt, cn_block_outputs* out) {
    using namespace cn;
    auto& u = c.uq;
    require(u.ready, "call cn_uq_prepare first");
    require(k >= 1 && k <= 16, "k must be in [1, 16]");
    const int T = static_cast<int>(c.tissue.tags.size());
    for (int i = 0; i < k * T; ++i) require(std::isfinite(sigma[i]) && sigma[i] > 0.0, "conductivities must be positive");
    const hipStream_t s = c.s;
    const int K = padded_width(k);
    const std::int64_t n = c.n, m = c.m, nr = n - 1, nnz_off = c.nnz_red - nr;
    std::vector<double> sg(static_cast<std::size_t>(K) * T, 0.0);
    std::copy(sigma, sigma + static_cast<std::size_t>(k) * T, sg.begin());
    u.sigma.ensure(sg.size(), "uq sigma");
    c.stage_blk.put(u.sigma.get(), sg.data(), sg.size(), s);
    u.off_k.ensure(static_cast<std::size_t>(std::max<std::int64_t>(nnz_off, 1)) * K, "sample off-diagonals");
    u.rowsum_k.ensure(static_cast<std::size_t>(nr) * K, "sample row sums");
    c.d_B.ensure(nr * K, "rhs block");
    combine_tissues(c.tissue.off.get(), nnz_off, T, u.sigma.get(), k, K, u.off_k.get(), s);
    combine_tissues(c.tissue.rowsum.get(), nr, T, u.sigma.get(), k, K, u.rowsum_k.get(), s);
    combine_tissues(u.rhs.get(), nr, T, u.sigma.get(), k, K, c.d_B.get(), s);
    u.mop.width = K;
    u.mop.off_vals = u.off_k.get();
    u.mop.rowsum = u.rowsum_k.get();
    u.mop.tissue_vals = c.tissue.vals.get();
    u.mop.n_tissue = T;
    u.mop.sigma = u.sigma.get();
    SolveOptions opt;
    if (sopt) {
        opt.tolerance = sopt->tolerance;
        opt.max_iters = sopt->max_iters;
        opt.max_refine = sopt->max_refine;
        opt.drop_tol = sopt->drop_tol;
        opt.use_graph = sopt->use_graph != 0;
        opt.check_every = std::max(1, sopt->check_every);
        opt.independent = sopt->independent != 0;
    }
    c.d_X.ensure(nr * K, "solution");
    const double* dx0 = nullptr;
    if (sopt && sopt->warm_start != 0) {
        // x0 = x_nom + S M sigma per draw (M zero when the field sees some tissue not at all).
        std::vector<double> coef(static_cast<std::size_t>(T) * K, 0.0);
        for (int q = 0; q < T; ++q)
            for (int col = 0; col < k; ++col) {
                double v = 0.0;
                for (int t = 0; t < T; ++t) v += u.guess_M[q * T + t] * sigma[col * T + t];
                coef[q * K + col] = v;
            }
        DeviceBuffer<double> dcoef(coef.size(), "guess coefficients");
        c.stage_blk.put(dcoef.get(), coef.data(), coef.size(), s);
        c.d_x0.ensure(nr * K, "initial guess");
        expand_guess(u.x_nom.get(), u.guess_rank > 0 ? u.S.get() : u.x_nom.get(), nr, u.guess_rank > 0 ? T : 0,
                     u.S_width, dcoef.get(), k, K, c.d_x0.get(), s);
        sync(s, "uq initial guess");
        dx0 = c.d_x0.get();
    }
    c.solver->set_multi(&u.mop);
    SolveStats st;
    try {
        st = c.solver->solve(c.d_B.get(), c.d_X.get(), dx0, k, opt);
    } catch (...) {
        c.solver->set_multi(nullptr);
        throw;
    }
    c.solver->set_multi(nullptr);
    retry_missed(c, st, opt, c.d_B.get(), k, K, u.sigma.get(), c.d_X.get());
    // |E| per sample (dA/dt is the same for every sample), recovered before any statistic is
    // taken (uq/conductivity/run.py:340-352).
    const int mode = sopt ? sopt->recovery : CN_RECOVERY_RAW;
    const RecoveryOp* rec = mode == CN_RECOVERY_RAW ? nullptr : ensure_recovery(c, mode);
    c.d_magn.ensure(m * K, "magnE");
    float* dE = nullptr;
    if (out && out->E) {
        c.d_E.ensure(3 * m * K, "E");
        dE = c.d_E.get();
    }
    if (rec == nullptr) {
        reconstruct(c.d_X.get(), c.tets.get(), c.g32.get(), u.dadt_nodes1.get(), 1, m, k, K, c.ground, dE,
                    c.d_magn.get(), s);
    } else if (rec->mode == kRecoveryHarmonic) {
        c.d_slots.ensure(3 * rec->n_slots * K, "recovered slots");
        apply_harmonic(*rec, c.d_X.get(), c.ground, u.dadt_nodes1.get(), 1, m, k, K, c.d_slots.get(), dE,
                       c.d_magn.get(), s);
    } else {
        c.d_Eraw.ensure(3 * m * K, "raw E");
        c.d_slots.ensure(3 * rec->n_slots * K, "recovered slots");
        reconstruct(c.d_X.get(), c.tets.get(), c.g32.get(), u.dadt_nodes1.get(), 1, m, k, K, c.ground, c.d_Eraw.get(),
                    c.d_magn.get(), s);
        apply_spr(*rec, c.d_Eraw.get(), m, k, K, c.d_slots.get(), dE, c.d_magn.get(), s);
    }
    accumulate_moments(c.d_magn.get(), m, k, K, u.sum.get(), u.sumsq.get(), s);
    u.count += k;
    if (out) {
        auto fetch = [&](float* host, const float* dev, std::int64_t rows, int inner) {
            c.d_stage.ensure(rows * inner * k, "output staging");
            unpack(dev, rows, inner, k, K, c.d_stage.get(), s);
            d2h(host, c.d_stage.get(), rows * inner * k, s);
            sync(s, "output copy");
        };
        if (out->transforms)
            for (int i = 0; i < k; ++i) std::memcpy(out->transforms + 16 * i, u.transform, sizeof(double) * 16);
        if (out->magnE) fetch(out->magnE, c.d_magn.get(), m, 1);
        if (out->E) fetch(out->E, c.d_E.get(), m, 3);
        if (out->v) {
            c.d_v.ensure(n * k, "potential");
            scatter_potential(c.d_X.get(), n, k, K, c.ground, c.d_v.get(), s);
            d2h(out->v, c.d_v.get(), n * k, s);
        }
        if (out->metrics)
            region_metrics(c.d_magn.get(), m, k, K, c.region_idx.get(), c.n_region, c.vols32.get(), c.bary.get(),
                           c.metrics_ws, out->metrics, s);
        for (int i = 0; i < k; ++i) {
            if (out->iterations) out->iterations[i] = st.column_iterations[i];
            if (out->relative_residual) out->relative_residual[i] = st.relative_residual[i];
        }
        out->total_iterations = st.iterations;
        out->refinements = st.refinements;
    }
    sync(s, "uq outputs");
    for (int i = 0; i < k; ++i)
        if (!(st.relative_residual[i] <= opt.tolerance))
            throw NotConverged("PCG did not converge (column " + std::to_string(i) + " of the sample block)");
}

const cn::RecoveryOp* ensure_recovery(cn_context& c, int mode) {
    using namespace cn;
    require(mode >= CN_RECOVERY_RAW && mode <= CN_RECOVERY_HARMONIC, "unknown recovery mode");
    require(mode != CN_RECOVERY_RAW, "the raw mode has no recovery operator");
    if (!c.recovery[mode]) {
        const auto t0 = std::chrono::steady_clock::now();
        auto op = std::make_unique<RecoveryOp>();
        if (mode == CN_RECOVERY_HARMONIC)
            build_harmonic(*op, c.nodes.get(), c.tets.get(), c.tags.get(), c.n, c.m, c.max_tag, c.s);
        else
            build_spr(*op, mode, c.nodes.get(), c.tets.get(), c.tags.get(), c.vols32.get(), c.n2c_ptr.get(),
                      c.n2c_idx.get(), c.n, c.m, c.max_tag, c.s);
        sync(c.s, "recovery operator");
        c.setup_ms[9] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        c.recovery[mode] = std::move(op);
    }
    return c.recovery[mode].get();
}

void solve_block(cn_context& c, const double* pos_m, const double* mom, std::int64_t n_dip, const double* centers,
                 const double* handles, const double* dists, int k, double didt, const cn_solve_options* sopt,
                 const double* x0, cn_block_outputs* out) {
    using namespace cn;
    require(k >= 1 && k <= kMaxBlock, "k must be in [1, 32]");
    require(n_dip >= 1, "the coil needs at least one dipole");
    require(std::isfinite(didt), "didt must be finite");
    for (int i = 0; i < 3 * k; ++i) {
        require(std::isfinite(centers[i]) && (handles == nullptr || std::isfinite(handles[i])),
                "center_mm and handle_mm must be finite 3-vectors");
    }
    const hipStream_t s = c.s;
    const int K = padded_width(k);
    const std::int64_t n = c.n, m = c.m, nr = n - 1;

    c.mark_phase(0, s);
    // 1. coil frames (place.cu), degenerate handles rejected with the placement's index.
    std::vector<double> tf(16 * k);
    std::vector<std::int32_t> degen(k);
    place_block(c, centers, handles, dists, k, tf.data(), degen.data());
    if (out && out->transforms) std::memcpy(out->transforms, tf.data(), sizeof(double) * 16 * k);
    if (out && out->degenerate) std::memcpy(out->degenerate, degen.data(), sizeof(std::int32_t) * k);
    if (handles != nullptr) {
        for (int i = 0; i < k; ++i) {
            if (degen[i]) {
                throw InvalidArgument("Degenerate coil placement at index " + std::to_string(i) + " of " +
                                      std::to_string(k) +
                                      ": handle_mm lies on the outward normal through the scalp projection of "
                                      "center_mm, so the coil's rotation about the normal is undefined.");
            }
        }
    }
    // 2. dA/dt at nodes, element average, RHS.
    c.d_dadt_nodes.ensure(3 * n * K, "dadt nodes");
    c.d_q.ensure(4 * m * K, "corner terms");
    c.d_B.ensure(nr * K, "rhs block");
    const float scale = static_cast<float>(didt) * static_cast<float>(kMu0Over4Pi);
    c.mark_phase(1, s);
    dadt_for_frames(c, tf.data(), k, K, pos_m, mom, n_dip, scale, c.d_dadt_nodes.get(), s);
    c.mark_phase(2, s);
    // Per-placement conductivities: the tissue decomposition of the operator and a (K, n_sigma)
    // table for the RHS; padded columns take the context's table.
    const bool per_sigma = sopt != nullptr && sopt->sigma != nullptr;
    if (per_sigma) {
        ensure_tissues(c);
        const auto& tg = c.tissue.tags;
        const int T = static_cast<int>(tg.size());
        const int ns = sopt->n_sigma;
        require(ns > tg.back(), "sigma rows must cover every tissue tag of the mesh (n_sigma > largest tag)");
        std::vector<float> tab(static_cast<std::size_t>(K) * ns, 0.0f);
        std::vector<double> sg(static_cast<std::size_t>(K) * T, 0.0);
        for (int col = 0; col < K; ++col) {
            for (int t = 0; t < T; ++t) {
                const double v = col < k ? sopt->sigma[static_cast<std::size_t>(col) * ns + tg[t]] : c.lut[tg[t]];
                require(std::isfinite(v) && v > 0.0, "conductivities of the tissues present must be positive");
                sg[static_cast<std::size_t>(col) * T + t] = v;
                tab[static_cast<std::size_t>(col) * ns + tg[t]] = static_cast<float>(v);
            }
        }
        c.blk_sigma.ensure(sg.size(), "block sigma");
        c.blk_sigma_tag.ensure(tab.size(), "block sigma by tag");
        const std::size_t sg_bytes = sg.size() * sizeof(double);
        char* host = static_cast<char*>(c.stage_blk.fill(sg_bytes + tab.size() * sizeof(float)));
        std::memcpy(host, sg.data(), sg_bytes);
        std::memcpy(host + sg_bytes, tab.data(), tab.size() * sizeof(float));
        c.stage_blk.send(c.blk_sigma.get(), 0, sg_bytes, s);
        c.stage_blk.send(c.blk_sigma_tag.get(), sg_bytes, tab.size() * sizeof(float), s);
        c.stage_blk.done(s);
        const std::int64_t nnz_off = c.nnz_red - nr;
        c.blk_off.ensure(static_cast<std::size_t>(std::max<std::int64_t>(nnz_off, 1)) * K, "block off-diagonals");
        c.blk_rowsum.ensure(static_cast<std::size_t>(nr) * K, "block row sums");
        combine_tissues(c.tissue.off.get(), nnz_off, T, c.blk_sigma.get(), k, K, c.blk_off.get(), s);
        combine_tissues(c.tissue.rowsum.get(), nr, T, c.blk_sigma.get(), k, K, c.blk_rowsum.get(), s);
        c.blk_mop.width = K;
        c.blk_mop.off_vals = c.blk_off.get();
        c.blk_mop.rowsum = c.blk_rowsum.get();
        c.blk_mop.tissue_vals = c.tissue.vals.get();
        c.blk_mop.n_tissue = T;
        c.blk_mop.sigma = c.blk_sigma.get();
    }
    // The element dA/dt is stored only when it is an output; the field pass rebuilds it from the nodes.
    float* elm_out = nullptr;
    if (out && out->dadt_elm) {
        c.d_dadt_elm.ensure(3 * m * K, "dadt elements");
        elm_out = c.d_dadt_elm.get();
    }
    element_pass(c.d_dadt_nodes.get(), c.tets.get(), c.g32.get(), c.neg_vc.get(), m, k, K, elm_out,
                 c.d_q.get(), c.corner_pos.get(), s, c.vols32.get(), c.tags.get(),
                 per_sigma ? c.blk_sigma_tag.get() : nullptr, per_sigma ? sopt->n_sigma : 0);
    float* rhs_full = nullptr;
    if (out && out->rhs) {
        c.d_rhs.ensure(n * k, "rhs full");
        rhs_full = c.d_rhs.get();
    }
    node_rhs(c.d_q.get(), c.n2c_ptr.get(), n, k, K, c.ground, c.d_B.get(), rhs_full, s);
    c.mark_phase(3, s);
    // 3. block PCG.
    SolveOptions opt;
    if (sopt) {
        opt.tolerance = sopt->tolerance;
        opt.max_iters = sopt->max_iters;
        opt.max_refine = sopt->max_refine;
        opt.drop_tol = sopt->drop_tol;
        opt.use_graph = sopt->use_graph != 0;
        opt.check_every = std::max(1, sopt->check_every);
        opt.independent = sopt->independent != 0;
    }
    c.d_X.ensure(nr * K, "solution");
    const double* dx0 = nullptr;
    if (x0 != nullptr) {
        // (k, n_red) placement-major host guess -> (n_red, K) row-major device block
        double* blk = static_cast<double*>(c.stage_blk.fill(sizeof(double) * nr * K));
        std::fill(blk, blk + nr * K, 0.0);
        for (int ci = 0; ci < k; ++ci)
            for (std::int64_t r = 0; r < nr; ++r) blk[r * K + ci] = x0[ci * nr + r];
        c.d_x0.ensure(nr * K, "x0");
        c.stage_blk.send(c.d_x0.get(), 0, sizeof(double) * nr * K, s);
        c.stage_blk.done(s);
        dx0 = c.d_x0.get();
    } else if (sopt && sopt->warm_start == 1) {
        c.d_x0.ensure(nr * K, "x0");
        c.warm.initial_guess(c.d_B.get(), k, K, c.d_x0.get(), s);
        dx0 = c.d_x0.get();
    } else if (sopt && sopt->warm_start == 2) {
        require(c.warm_coef_k == k, "warm_start = 2 needs cn_warm_set_coefficients for exactly this block's k");
        c.d_x0.ensure(nr * K, "x0");
        c.warm.guess_from_coefficients(c.warm_coef.data(), k, K, c.d_x0.get(), s);
        dx0 = c.d_x0.get();
        c.warm_coef_k = 0;
    }
    if (out && out->x0) {
        // (n_red, K) device block -> (k, n_red) host, zeros for a cold start
        std::vector<double> blk(static_cast<std::size_t>(nr * K), 0.0);
        if (dx0) {
            d2h(blk.data(), dx0, blk.size(), s);
            sync(s, "x0 copy");
        }
        for (int ci = 0; ci < k; ++ci)
            for (std::int64_t r = 0; r < nr; ++r) out->x0[ci * nr + r] = blk[r * K + ci];
    }
    SolveStats st;
    c.mark_phase(4, s);
    if (per_sigma) {
        c.solver->set_multi(&c.blk_mop);
        try {
            st = c.solver->solve(c.d_B.get(), c.d_X.get(), dx0, k, opt);
        } catch (...) {
            c.solver->set_multi(nullptr);
            throw;
        }
        c.solver->set_multi(nullptr);
        retry_missed(c, st, opt, c.d_B.get(), k, K, c.blk_sigma.get(), c.d_X.get());
    } else {
        st = c.solver->solve(c.d_B.get(), c.d_X.get(), dx0, k, opt);
        retry_missed(c, st, opt, c.d_B.get(), k, K, nullptr, c.d_X.get());
    }
    if (sopt && sopt->collect) c.warm.append(c.d_X.get(), k, K, s);
    c.mark_phase(5, s);
    // 4. reconstruction, recovery and metrics.
    const int mode = sopt ? sopt->recovery : CN_RECOVERY_RAW;
    const RecoveryOp* rec = mode == CN_RECOVERY_RAW ? nullptr : ensure_recovery(c, mode);
    c.d_magn.ensure(m * K, "magnE");
    float* dE = nullptr;
    if (out && out->E) {
        c.d_E.ensure(3 * m * K, "E");
        dE = c.d_E.get();
    }
    if (rec == nullptr) {
        reconstruct(c.d_X.get(), c.tets.get(), c.g32.get(), c.d_dadt_nodes.get(), K, m, k, K, c.ground, dE,
                    c.d_magn.get(), s);
    } else if (rec->mode == kRecoveryHarmonic) {
        // The potential and the exact nodal dA/dt feed the fits; |E| and E describe the result.
        c.d_slots.ensure(3 * rec->n_slots * K, "recovered slots");
        apply_harmonic(*rec, c.d_X.get(), c.ground, c.d_dadt_nodes.get(), K, m, k, K, c.d_slots.get(), dE,
                       c.d_magn.get(), s);
    } else {
        // The raw field feeds the patch fits; |E| and E then describe the recovered field.
        c.d_Eraw.ensure(3 * m * K, "raw E");
        c.d_slots.ensure(3 * rec->n_slots * K, "recovered slots");
        reconstruct(c.d_X.get(), c.tets.get(), c.g32.get(), c.d_dadt_nodes.get(), K, m, k, K, c.ground,
                    c.d_Eraw.get(), c.d_magn.get(), s);
        apply_spr(*rec, c.d_Eraw.get(), m, k, K, c.d_slots.get(), dE, c.d_magn.get(), s);
    }
    c.mark_phase(6, s);
    if (out) {
        // Placement-fastest device arrays -> placement-major host arrays through one staging buffer.
        auto fetch = [&](float* host, const float* dev, std::int64_t rows, int inner) {
            c.d_stage.ensure(rows * inner * k, "output staging");
            unpack(dev, rows, inner, k, K, c.d_stage.get(), s);
            d2h(host, c.d_stage.get(), rows * inner * k, s);
            sync(s, "output copy");
        };
        if (out->magnE) fetch(out->magnE, c.d_magn.get(), m, 1);
        if (out->E) fetch(out->E, c.d_E.get(), m, 3);
        if (out->E_slots) {
            require(rec != nullptr, "E_slots needs a recovery mode");
            fetch(out->E_slots, c.d_slots.get(), rec->n_slots, 3);
        }
        if (out->dadt_elm) fetch(out->dadt_elm, c.d_dadt_elm.get(), m, 3);
        if (out->dadt_nodes) fetch(out->dadt_nodes, c.d_dadt_nodes.get(), n, 3);
        if (out->rhs) d2h(out->rhs, c.d_rhs.get(), n * k, s);
        if (out->v) {
            c.d_v.ensure(n * k, "potential");
            scatter_potential(c.d_X.get(), n, k, K, c.ground, c.d_v.get(), s);
            d2h(out->v, c.d_v.get(), n * k, s);
        }
        if (out->metrics)
            region_metrics(c.d_magn.get(), m, k, K, c.region_idx.get(), c.n_region, c.vols32.get(), c.bary.get(),
                           c.metrics_ws, out->metrics, s);
        for (int i = 0; i < k; ++i) {
            if (out->iterations) out->iterations[i] = st.column_iterations[i];
            if (out->relative_residual) out->relative_residual[i] = st.relative_residual[i];
        }
        out->total_iterations = st.iterations;
        out->refinements = st.refinements;
        for (int r = 0; r < 8; ++r) {
            out->round_iterations[r] = r < cn::kMaxRounds ? st.round_iterations[r] : 0;
            out->round_residual[r] = r < cn::kMaxRounds ? st.round_residual[r] : 0.0;
        }
    }
    c.mark_phase(7, s);
    sync(s, "block outputs");
    c.close_phases(k);
    for (int i = 0; i < k; ++i) {
        if (!(st.relative_residual[i] <= opt.tolerance)) {
            char buf[256];
            std::snprintf(buf, sizeof buf,
                          "PCG did not converge (column %d of the block): %d iterations, relative residual %.3e > "
                          "tolerance %.3e",
                          i, st.column_iterations[i] ? st.column_iterations[i] : st.iterations,
                          st.relative_residual[i], opt.tolerance);
            throw NotConverged(buf);
        }
    }
}

// A sweep of n placements in blocks of block_k, pipelined over two streams: block b + 1's coil
// stage (dA/dt at the nodes, the element pass and the node RHS, into buffer set (b + 1) % 2) runs on
// s2 while block b solves on s (initial guess, block PCG, field, region metrics). The two overlap
// on the device: the N-body keeps the vector ALUs busy while the solve streams the operator. The
// frames of every placement are found first, in one pass. Each block's numerics are those of
// cn_solve_block for the same block.
void solve_sweep(cn_context& c, const double* pos_m, const double* mom, std::int64_t n_dip, const double* centers,
                 const double* handles, const double* dists, int n_place, int block_k, double didt,
                 const cn_solve_options* sopt, double* transforms_out, double* metrics_out, std::int32_t* iters_out,
                 double* resid_out, std::int32_t* block_iters_out) {
    using namespace cn;
    require(n_place >= 1, "no placements");
    require(block_k >= 1 && block_k <= kMaxBlock, "block_k must be in [1, 32]");
    require(n_dip >= 1, "the coil needs at least one dipole");
    require(std::isfinite(didt), "didt must be finite");
    require(sopt == nullptr || (sopt->sigma == nullptr && sopt->collect == 0 && sopt->warm_start != 2),
            "a sweep takes the context's conductivities, no collection and no supplied coefficients");
    const std::int64_t n = c.n, m = c.m, nr = n - 1;
    const int K = padded_width(block_k);
    // frames of every placement (place.cu), degenerate handles rejected with the placement's index
    std::vector<double> tf(static_cast<std::size_t>(16) * n_place);
    std::vector<std::int32_t> degen(n_place);
    place_block(c, centers, handles, dists, n_place, tf.data(), degen.data());
    if (handles != nullptr)
        for (int i = 0; i < n_place; ++i)
            if (degen[i])
                throw InvalidArgument("Degenerate coil placement at index " + std::to_string(i) + " of " +
                                      std::to_string(n_place) +
                                      ": handle_mm lies on the outward normal through the scalp projection of "
                                      "center_mm, so the coil's rotation about the normal is undefined.");
    if (transforms_out) std::memcpy(transforms_out, tf.data(), sizeof(double) * tf.size());
    if (c.s2 == nullptr) check(hipStreamCreateWithFlags(&c.s2, hipStreamNonBlocking), "create stream");
    for (auto& e : c.coil_done)
        if (e == nullptr) check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "coil event");
    c.d_dadt_nodes.ensure(3 * n * K, "dadt nodes");
    c.d_dadt_nodes2.ensure(3 * n * K, "dadt nodes");
    c.d_dadt_nodes3.ensure(3 * n * K, "dadt nodes");
    c.d_B.ensure(nr * K, "rhs block");
    c.d_B2.ensure(nr * K, "rhs block");
    c.d_B3.ensure(nr * K, "rhs block");
    c.d_q.ensure(4 * m * K, "corner terms");
    // Block b uses buffer set b % 3: the coil stage of block b + 2 fills its set while block b
    // still reads its own (the nodal dA/dt in the field recovery) and block b + 1's waits.
    float* const nodal[3] = {c.d_dadt_nodes.get(), c.d_dadt_nodes2.get(), c.d_dadt_nodes3.get()};
    float* const rhs[3] = {c.d_B.get(), c.d_B2.get(), c.d_B3.get()};
    const float scale = static_cast<float>(didt) * static_cast<float>(kMu0Over4Pi);
    const int n_blocks = (n_place + block_k - 1) / block_k;
    // Each block is laid out at its own padded width (a ragged last block may be narrower), as
    // the solver takes it. The first block's N-body overlaps nothing (the first solve waits for
    // it), so it takes the full grid; later ones take the capped grid beside a solve.
    auto coil_stage = [&](int b) {
        const int first = b * block_k, k = std::min(block_k, n_place - first), buf = b % 3;
        const int Kb = padded_width(k);
        dadt_for_frames(c, tf.data() + 16 * static_cast<std::size_t>(first), k, Kb, pos_m, mom, n_dip, scale,
                        nodal[buf], c.s2, b == 0 ? 0 : c.coil_blocks);
        element_pass(nodal[buf], c.tets.get(), c.g32.get(), c.neg_vc.get(), m, k, Kb, nullptr, c.d_q.get(),
                     c.corner_pos.get(), c.s2);
        node_rhs(c.d_q.get(), c.n2c_ptr.get(), n, k, Kb, c.ground, rhs[buf], nullptr, c.s2);
        check(hipEventRecord(c.coil_done[buf], c.s2), "coil event");
    };
    SolveOptions opt;
    if (sopt) {
        opt.tolerance = sopt->tolerance;
        opt.max_iters = sopt->max_iters;
        opt.max_refine = sopt->max_refine;
        opt.drop_tol = sopt->drop_tol;
        opt.use_graph = sopt->use_graph != 0;
        opt.check_every = std::max(1, sopt->check_every);
        opt.independent = sopt->independent != 0;
    }
    const bool warm = sopt && sopt->warm_start == 1;
    // Blocks of earlier solutions in each full block's initial guess (none with a warm-start basis,
    // and at most 32 basis vectors).
    const int recycle = warm ? 0 : std::min(sopt ? std::max(sopt->recycle, 0) : 2, std::min(2, 32 / K));
    require(sopt == nullptr || sopt->recycle >= 0, "recycle must be >= 0");
    const int mode = sopt ? sopt->recovery : CN_RECOVERY_RAW;
    const RecoveryOp* rec = mode == CN_RECOVERY_RAW ? nullptr : ensure_recovery(c, mode);
    c.d_X.ensure(nr * K, "solution");
    c.d_X2.ensure(nr * K, "solution");
    if (recycle > 0) c.solver->guess_reset(recycle, K);
    double* const sol[2] = {c.d_X.get(), c.d_X2.get()};
    // The block outputs run on s3, driven by a worker thread, while the next block solves: every
    // buffer they write is theirs alone, allocated here before any of them starts.
    c.d_magn.ensure(m * K, "magnE");
    if (rec != nullptr) c.d_slots.ensure(3 * rec->n_slots * K, "recovered slots");
    if (rec != nullptr && rec->mode != kRecoveryHarmonic) c.d_Eraw.ensure(3 * m * K, "raw E");
    if (c.s3 == nullptr) check(hipStreamCreateWithFlags(&c.s3, hipStreamNonBlocking), "create stream");
    for (auto& e : c.solve_done)
        if (e == nullptr) check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "solve event");
    const hipStream_t s = c.s, s3 = c.s3;
    // Block b's outputs: its field from its solution and nodal dA/dt, then its region metrics, on
    // s3 after its solve; the worker waits for them before it returns.
    int device = 0;
    check(hipGetDevice(&device), "device");
    // The caller orders s3 after block b's solve before it starts the worker: a stream wait on an
    // event of the solve stream is refused while that stream captures a graph, which the next
    // block's solve may be doing by the time the worker runs.
    auto outputs = [&c, rec, m, s3, metrics_out, &nodal, &sol, device](int b, int first, int k, int Kb, int buf) {
        check(hipSetDevice(device), "device");  // the worker thread works on the context's device
        const double* X = sol[b % 2];
        if (rec == nullptr) {
            reconstruct(X, c.tets.get(), c.g32.get(), nodal[buf], Kb, m, k, Kb, c.ground, nullptr, c.d_magn.get(), s3);
        } else if (rec->mode == kRecoveryHarmonic) {
            apply_harmonic(*rec, X, c.ground, nodal[buf], Kb, m, k, Kb, c.d_slots.get(), nullptr, c.d_magn.get(), s3);
        } else {
            reconstruct(X, c.tets.get(), c.g32.get(), nodal[buf], Kb, m, k, Kb, c.ground, c.d_Eraw.get(),
                        c.d_magn.get(), s3);
            apply_spr(*rec, c.d_Eraw.get(), m, k, Kb, c.d_slots.get(), nullptr, c.d_magn.get(), s3);
        }
        if (metrics_out)
            region_metrics(c.d_magn.get(), m, k, Kb, c.region_idx.get(), c.n_region, c.vols32.get(), c.bary.get(),
                           c.metrics_ws, metrics_out + static_cast<std::size_t>(kMetricsLen) * first, s3);
        sync(s3, "block outputs");
    };
    std::future<void> pending;  // the previous block's outputs
    auto join = [&pending] {
        if (pending.valid()) pending.get();  // rethrows the worker's error
    };
    try {
        // Two coil stages ahead. Block b + 2's is queued after block b has solved and block b - 1's
        // outputs are done, which frees its buffer set; block b's solve writes the solution buffer
        // block b - 2's outputs read, done by then too.
        coil_stage(0);
        if (n_blocks > 1) coil_stage(1);
        for (int b = 0; b < n_blocks; ++b) {
            const int first = b * block_k, k = std::min(block_k, n_place - first), buf = b % 3;
            const int Kb = padded_width(k);
            check(hipStreamWaitEvent(s, c.coil_done[buf], 0), "wait for the coil stage");
            const double* dx0 = nullptr;
            if (warm) {
                c.d_x0.ensure(nr * K, "x0");
                c.warm.initial_guess(rhs[buf], k, Kb, c.d_x0.get(), s);
                dx0 = c.d_x0.get();
            } else if (recycle > 0 && Kb == K && c.solver->guess_blocks() > 0) {
                c.d_x0.ensure(nr * K, "x0");
                c.solver->guess(rhs[buf], K, kGuessDropTol, c.d_x0.get());
                dx0 = c.d_x0.get();
            }
            SolveStats st = c.solver->solve(rhs[buf], sol[b % 2], dx0, k, opt);
            retry_missed(c, st, opt, rhs[buf], k, Kb, nullptr, sol[b % 2]);
            if (recycle > 0 && Kb == K) c.solver->guess_push(sol[b % 2], rhs[buf], K);
            check(hipEventRecord(c.solve_done[b % 2], s), "solve event");
            join();
            if (b + 2 < n_blocks) coil_stage(b + 2);
            check(hipStreamWaitEvent(s3, c.solve_done[b % 2], 0), "wait for the solve");
            pending = std::async(std::launch::async, outputs, b, first, k, Kb, buf);
            for (int i = 0; i < k; ++i) {
                if (iters_out) iters_out[first + i] = st.column_iterations[i];
                if (resid_out) resid_out[first + i] = st.relative_residual[i];
            }
            if (block_iters_out) block_iters_out[b] = st.iterations;
            for (int i = 0; i < k; ++i) {
                if (!(st.relative_residual[i] <= opt.tolerance)) {
                    char msg[256];
                    std::snprintf(msg, sizeof msg,
                                  "PCG did not converge (placement %d of the sweep): %d iterations, relative residual "
                                  "%.3e > tolerance %.3e",
                                  first + i, st.column_iterations[i] ? st.column_iterations[i] : st.iterations,
                                  st.relative_residual[i], opt.tolerance);
                    throw NotConverged(msg);
                }
            }
        }
        join();
    } catch (...) {
        if (pending.valid()) pending.wait();  // the outputs in flight read this sweep's buffers
        (void)hipStreamSynchronize(c.s2);     // the next coil stages may still be writing theirs
        (void)hipStreamSynchronize(s);
        (void)hipStreamSynchronize(s3);
        throw;
    }
}

}  // namespace

extern "C" {

const char* cn_last_error(void) { return g_error.c_str(); }
int32_t cn_version(void) { return 1; }

int32_t cn_block_sizes(int32_t* out, int32_t cap) {
    for (int i = 0; i < cn::kNumBlockWidths && i < cap; ++i) out[i] = cn::kBlockWidths[i];
    return cn::kNumBlockWidths;
}

int32_t cn_max_block(void) { return cn::kMaxBlock; }

void cn_default_options(cn_options* o) {
    std::memset(o, 0, sizeof(*o));
    o->min_coarse_rows = 128;
    o->max_levels = 50;
    o->rounds = 2;
    o->degree = 2;
    o->lower_ratio = 12.0f;
    o->keep_stiffness = 0;
    o->conductivity = nullptr;
    o->n_conductivity = 0;
    o->region_tag = 2;
}

void cn_default_solve_options(cn_solve_options* o) {
    o->recovery = CN_RECOVERY_RAW;
    o->warm_start = 0;
    o->collect = 0;
    o->sigma = nullptr;
    o->n_sigma = 0;
    o->tolerance = 1e-6;
    o->max_iters = 2000;
    o->max_refine = 8;
    o->drop_tol = 1e-6;
    o->use_graph = 1;
    o->check_every = 4;
    o->independent = 0;
    o->recycle = 2;
}

int32_t cn_context_create(const double* nodes_mm, int64_t n_nodes, const int32_t* tet_nodes, const int32_t* tet_tags,
                          int64_t n_tets, const int32_t* skin_tris, int64_t n_skin, const cn_options* opts,
                          cn_context** out) {
    return guarded([&] {
        cn::require(out != nullptr && nodes_mm != nullptr && tet_nodes != nullptr && tet_tags != nullptr,
                    "null argument");
        cn::require(n_skin == 0 || skin_tris != nullptr, "null skin_tris with n_skin > 0");
        auto c = std::make_unique<cn_context>();
        if (opts)
            c->opts = *opts;
        else
            cn_default_options(&c->opts);
        cn::require(c->opts.degree >= 1 && c->opts.degree <= cn::kMaxDegree, "degree must be in [1, 6]");
        cn::require(c->opts.lower_ratio > 1.0f, "lower_ratio must exceed 1");
        cn::check(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking), "create stream");
        build_context(*c, nodes_mm, n_nodes, tet_nodes, tet_tags, n_tets, skin_tris, n_skin);
        *out = c.release();
    });
}

void cn_context_destroy(cn_context* ctx) {
    if (ctx) {
        (void)hipStreamSynchronize(ctx->s);
        delete ctx;
    }
}

int32_t cn_context_get_tuning(const cn_context* c, cn_tuning* o) {
    return guarded([&] {
        cn::require(c != nullptr && o != nullptr, "null argument");
        o->degree = c->hier.cheb.degree;
        o->lower_ratio = c->opts.lower_ratio;
        o->fine_sub = c->solver->fine_sub_setting();
        o->coarse_sub = c->solver->coarse_sub_setting();
        o->coil_blocks = c->coil_blocks;
        o->layout = c->layout;
        o->dense_tail = c->solver->dense_tail_setting() ? 1 : 0;
    });
}

int32_t cn_context_set_tuning(cn_context* c, const cn_tuning* t) {
    return guarded([&] {
        cn::require(c != nullptr && t != nullptr, "null argument");
        cn::require(t->degree >= 1 && t->degree <= cn::kMaxDegree, "degree must be in [1, 6]");
        cn::require(t->lower_ratio > 1.0f, "lower_ratio must exceed 1");
        cn::require(t->fine_sub >= 0 && t->fine_sub <= 64 && (t->fine_sub & (t->fine_sub - 1)) == 0,
                    "fine_sub must be 0 or a power of two <= 64");
        cn::require(t->coarse_sub >= 0 && t->coarse_sub <= 64 && (t->coarse_sub & (t->coarse_sub - 1)) == 0,
                    "coarse_sub must be 0 or a power of two <= 64");
        cn::require(t->coil_blocks >= 0, "coil_blocks must be >= 0");
        cn::require(t->layout >= 0 && t->layout <= 2, "layout must be 0 (automatic), 1 (CSR) or 2 (sliced ELL)");
        cn::require(t->dense_tail == 0 || t->dense_tail == 1, "dense_tail must be 0 or 1");
        cn::check(hipStreamSynchronize(c->s), "sync before retuning");
        c->coil_blocks = t->coil_blocks;
        c->set_layout(t->layout);
        c->hier.cheb = cn::chebyshev(t->degree, t->lower_ratio);
        c->opts.degree = t->degree;
        c->opts.lower_ratio = t->lower_ratio;
        c->solver->set_fine_sub(t->fine_sub);
        c->solver->set_coarse_sub(t->coarse_sub);
        c->solver->set_dense_tail(t->dense_tail != 0);
        c->solver->invalidate();
    });
}

int32_t cn_context_info(const cn_context* c, cn_info* o) {
    return guarded([&] {
        cn::require(c && o, "null argument");
        std::memset(o, 0, sizeof(*o));
        o->n_nodes = c->n;
        o->n_tets = c->m;
        o->n_skin = c->n_skin;
        o->n_reduced = c->n - 1;
        o->nnz_reduced = c->nnz_red;
        o->nnz_full = c->nnz_full;
        o->ground_node = c->ground;
        o->n_levels = static_cast<int32_t>(std::min<std::size_t>(c->hier.levels.size(), CN_MAX_LEVELS));
        for (int i = 0; i < o->n_levels; ++i) {
            o->level_rows[i] = c->hier.levels[i].a.n;
            o->level_nnz[i] = c->hier.levels[i].a.nnz;
            o->level_p_nnz[i] = c->hier.levels[i].p.nnz;
        }
        o->level_rows[o->n_levels] = c->hier.coarse.n;
        o->level_nnz[o->n_levels] = c->hier.coarse.nnz;
        o->n_region = c->n_region;
        o->region_volume_m3 = c->region_volume;
        o->ell_width_full = c->ell_full_w;
        o->ell_width_off = c->ell_off_w;
        o->preconditioner_rebuilds = c->precond_rebuilds;
        o->matched_preconditioners = c->matched_preconditioners;
        o->dense_tail_level = -1;
        if (c->solver->dense_tail_setting()) {
            for (std::size_t i = 0; i < c->hier.levels.size(); ++i) {
                if (c->hier.levels[i].a.n <= cn::kDenseTailMaxRows) {
                    o->dense_tail_level = static_cast<int32_t>(i);
                    break;
                }
            }
        }
        o->n_cu = c->n_cu;
    });
}

int32_t cn_context_mesh(const cn_context* c, double* nodes_mm, int32_t* tets, int32_t* tags, int32_t* skin,
                        int32_t* node_perm, int32_t* tet_perm) {
    return guarded([&] {
        cn::require(c != nullptr, "null context");
        if (nodes_mm) cn::d2h(nodes_mm, c->nodes.get(), 3 * c->n, c->s);
        if (tets) cn::d2h(tets, c->tets.get(), 4 * c->m, c->s);
        if (tags) cn::d2h(tags, c->tags.get(), c->m, c->s);
        if (skin) cn::d2h(skin, c->skin.get(), 3 * c->n_skin, c->s);
        if (node_perm) cn::d2h(node_perm, c->node_perm.get(), c->n, c->s);
        if (tet_perm) cn::d2h(tet_perm, c->tet_perm.get(), c->m, c->s);
        cn::sync(c->s, "mesh readback");
    });
}

int32_t cn_context_geometry(const cn_context* c, float* vols, double* bary, float* g, float* neg_vc) {
    return guarded([&] {
        cn::require(c != nullptr, "null context");
        if (vols) cn::d2h(vols, c->vols32.get(), c->m, c->s);
        if (bary) cn::d2h(bary, c->bary.get(), 3 * c->m, c->s);
        if (g) cn::d2h(g, c->g32.get(), 12 * c->m, c->s);
        if (neg_vc) cn::d2h(neg_vc, c->neg_vc.get(), c->m, c->s);
        cn::sync(c->s, "geometry readback");
    });
}

int32_t cn_context_stiffness(const cn_context* c, int32_t* indptr, int32_t* indices, double* data) {
    return guarded([&] {
        cn::require(c != nullptr, "null context");
        cn::require(static_cast<bool>(c->indptr), "the context was built without keep_stiffness");
        if (indptr) cn::d2h(indptr, c->indptr.get(), c->n + 1, c->s);
        if (indices) cn::d2h(indices, c->indices.get(), c->nnz_full, c->s);
        if (data) cn::d2h(data, c->data.get(), c->nnz_full, c->s);
        cn::sync(c->s, "stiffness readback");
    });
}

int32_t cn_context_reduced(const cn_context* c, int32_t* ptr, int32_t* col, double* val) {
    return guarded([&] {
        cn::require(c != nullptr, "null context");
        if (ptr) cn::d2h(ptr, c->rptr.get(), c->n, c->s);
        if (col) cn::d2h(col, c->rcol.get(), c->nnz_red, c->s);
        if (val) cn::d2h(val, c->rval64.get(), c->nnz_red, c->s);
        cn::sync(c->s, "reduced readback");
    });
}

int32_t cn_context_level(const cn_context* c, int32_t level, int64_t* n_rows, int64_t* nnz, int64_t* n_coarse,
                         int64_t* p_nnz, int32_t* a_ptr, int32_t* a_col, float* a_val, float* dinv, int32_t* agg,
                         int32_t* p_ptr, int32_t* p_col, float* p_val) {
    return guarded([&] {
        cn::require(c != nullptr, "null context");
        const int nl = static_cast<int>(c->hier.levels.size());
        cn::require(level >= 0 && level <= nl, "level out of range");
        const hipStream_t s = c->s;
        if (level == nl) {
            const cn::CsrF& a = c->hier.coarse;
            if (n_rows) *n_rows = a.n;
            if (nnz) *nnz = a.nnz;
            if (n_coarse) *n_coarse = 0;
            if (p_nnz) *p_nnz = 0;
            if (a_ptr) cn::d2h(a_ptr, a.ptr.get(), a.n + 1, s);
            if (a_col) cn::d2h(a_col, a.col.get(), a.nnz, s);
            if (a_val) cn::d2h(a_val, a.val.get(), a.nnz, s);
        } else {
            const cn::Level& L = c->hier.levels[level];
            if (n_rows) *n_rows = L.a.n;
            if (nnz) *nnz = L.a.nnz;
            if (n_coarse) *n_coarse = L.n_coarse;
            if (p_nnz) *p_nnz = L.p.nnz;
            if (a_ptr) cn::d2h(a_ptr, L.a.ptr, L.a.n + 1, s);
            if (a_col) cn::d2h(a_col, L.a.col, L.a.nnz, s);
            if (a_val) cn::d2h(a_val, L.a.val, L.a.nnz, s);
            if (dinv) cn::d2h(dinv, L.dinv.get(), L.a.n, s);
            if (agg) cn::d2h(agg, L.agg.get(), L.a.n, s);
            if (p_ptr) cn::d2h(p_ptr, L.p.ptr.get(), L.p.n + 1, s);
            if (p_col) cn::d2h(p_col, L.p.col.get(), L.p.nnz, s);
            if (p_val) cn::d2h(p_val, L.p.val.get(), L.p.nnz, s);
        }
        cn::sync(s, "level readback");
    });
}

int32_t cn_context_skin_normals(const cn_context* c, double* out) {
    return guarded([&] {
        cn::require(c != nullptr && out != nullptr, "null argument");
        cn::d2h(out, c->skin_n.get(), 3 * c->n_skin, c->s);
        cn::sync(c->s, "skin normals readback");
    });
}

int32_t cn_solve_sweep(cn_context* c, const double* dip_pos_m, const double* dip_moment, int64_t n_dip,
                       const double* centers_mm, const double* handles_mm, const double* distances_mm,
                       int32_t n_place, int32_t block_k, double didt, const cn_solve_options* opt,
                       double* transforms, double* metrics, int32_t* iterations, double* relative_residual,
                       int32_t* block_iterations) {
    return guarded([&] {
        cn::require(c != nullptr && dip_pos_m && dip_moment && centers_mm && distances_mm, "null argument");
        solve_sweep(*c, dip_pos_m, dip_moment, n_dip, centers_mm, handles_mm, distances_mm, n_place, block_k, didt,
                    opt, transforms, metrics, iterations, relative_residual, block_iterations);
    });
}

int32_t cn_solve_block(cn_context* c, const double* dip_pos_m, const double* dip_moment, int64_t n_dip,
                       const double* centers_mm, const double* handles_mm, const double* distances_mm, int32_t k,
                       double didt, const cn_solve_options* opt, const double* x0, cn_block_outputs* out) {
    return guarded([&] {
        cn::require(c && dip_pos_m && dip_moment && centers_mm && distances_mm, "null argument");
        solve_block(*c, dip_pos_m, dip_moment, n_dip, centers_mm, handles_mm, distances_mm, k, didt, opt, x0, out);
    });
}

// ---- kernel-level entry points ----

int32_t cn_k_p1_gradients(const double* nodes_mm, int64_t n, const int32_t* tets, int64_t m, double* g, double* vols) {
    return guarded([&] {
        hipStream_t s = nullptr;
        auto dn = cn::upload(nodes_mm, 3 * n, s, "nodes");
        auto dt = cn::upload(tets, 4 * m, s, "tets");
        cn::DeviceBuffer<double> dg(12 * m, "g"), dv(m, "vols");
        cn::p1_gradients(dn.get(), dt.get(), m, dg.get(), dv.get(), s);
        cn::d2h(g, dg.get(), 12 * m, s);
        cn::d2h(vols, dv.get(), m, s);
        cn::sync(s, "gradients");
    });
}

int32_t cn_k_morton_keys(const double* pts, int64_t n, uint64_t* keys) {
    return guarded([&] {
        hipStream_t s = nullptr;
        auto dp = cn::upload(pts, 3 * n, s, "points");
        cn::DeviceBuffer<std::uint64_t> dk(std::max<int64_t>(n, 1), "keys");
        cn::morton_keys(dp.get(), n, dk.get(), s);
        cn::d2h(keys, dk.get(), n, s);
        cn::sync(s, "morton");
    });
}

int32_t cn_k_tet_lowest_node(const int32_t* inverse, int64_t n, const int32_t* tets, int64_t m, int32_t* lowest) {
    return guarded([&] {
        hipStream_t s = nullptr;
        auto di = cn::upload(inverse, n, s, "inverse");
        auto dt = cn::upload(tets, 4 * m, s, "tets");
        cn::DeviceBuffer<std::uint32_t> dl(std::max<int64_t>(m, 1), "lowest");
        cn::tet_lowest_node(di.get(), dt.get(), m, dl.get(), s);
        cn::d2h(reinterpret_cast<std::uint32_t*>(lowest), dl.get(), m, s);
        cn::sync(s, "tet lowest");
    });
}

int32_t cn_k_node2corner(const int32_t* elems, int64_t n_elem, int32_t corners, int64_t n_nodes, int32_t* ptr,
                         int32_t* idx) {
    return guarded([&] {
        hipStream_t s = nullptr;
        auto de = cn::upload(elems, n_elem * corners, s, "elements");
        cn::DeviceBuffer<std::int32_t> dp(n_nodes + 1, "ptr"), di(std::max<int64_t>(n_elem * corners, 1), "idx");
        cn::node2corner(de.get(), n_elem, corners, n_nodes, dp.get(), di.get(), s);
        cn::d2h(ptr, dp.get(), n_nodes + 1, s);
        cn::d2h(idx, di.get(), n_elem * corners, s);
        cn::sync(s, "node2corner");
    });
}

static cn::CsrF upload_csr(int64_t n, int64_t n_cols, const int32_t* ptr, const int32_t* col, const float* val,
                           hipStream_t s) {
    cn::CsrF a;
    a.n = n;
    a.n_cols = n_cols;
    a.nnz = ptr[n];
    a.ptr = cn::upload(ptr, n + 1, s, "ptr");
    a.col.reset(std::max<int64_t>(a.nnz, 1), "col");
    a.val.reset(std::max<int64_t>(a.nnz, 1), "val");
    cn::h2d(a.col.get(), col, a.nnz, s);
    cn::h2d(a.val.get(), val, a.nnz, s);
    return a;
}

int32_t cn_k_l1_dinv(int64_t n, const int32_t* ptr, const int32_t* col, const float* val, float* dinv) {
    return guarded([&] {
        hipStream_t s = nullptr;
        cn::CsrF a = upload_csr(n, n, ptr, col, val, s);
        cn::DeviceBuffer<float> d(std::max<int64_t>(n, 1), "dinv");
        cn::l1_dinv(a.view(), d.get(), s);
        cn::d2h(dinv, d.get(), n, s);
        cn::sync(s, "l1");
    });
}

int32_t cn_k_select_size4(int64_t n, const int32_t* ptr, const int32_t* col, const float* val, int32_t* agg,
                          int64_t* n_coarse) {
    return guarded([&] {
        hipStream_t s = nullptr;
        cn::CsrF a = upload_csr(n, n, ptr, col, val, s);
        cn::DeviceBuffer<std::int32_t> d(std::max<int64_t>(n, 1), "agg");
        *n_coarse = cn::select_size4(a.view(), d.get(), s);
        cn::d2h(agg, d.get(), n, s);
        cn::sync(s, "select_size4");
    });
}

int32_t cn_k_spgemm(int64_t a_rows, int64_t b_rows, int64_t b_cols, const int32_t* a_ptr, const int32_t* a_col,
                    const float* a_val, const int32_t* b_ptr, const int32_t* b_col, const float* b_val, int32_t* c_ptr,
                    int32_t* c_col, float* c_val, int64_t* c_nnz) {
    return guarded([&] {
        hipStream_t s = nullptr;
        cn::CsrF a = upload_csr(a_rows, b_rows, a_ptr, a_col, a_val, s);
        cn::CsrF b = upload_csr(b_rows, b_cols, b_ptr, b_col, b_val, s);
        cn::CsrF c = cn::spgemm(a.view(), b.view(), s);
        *c_nnz = c.nnz;
        cn::d2h(c_ptr, c.ptr.get(), a_rows + 1, s);
        if (c_col) cn::d2h(c_col, c.col.get(), c.nnz, s);
        if (c_val) cn::d2h(c_val, c.val.get(), c.nnz, s);
        cn::sync(s, "spgemm");
    });
}

int32_t cn_k_spmm(cn_context* c, const float* X, float* Y, int32_t K) {
    return guarded([&] {
        cn::require(c != nullptr, "null context");
        cn::require(cn::padded_width(K) == K, "K must be a compiled block width");
        const int64_t nk = (c->n - 1) * K;
        auto dx = cn::upload(X, nk, c->s, "X");
        cn::DeviceBuffer<float> dy(nk, "Y");
        c->solver->spmm(dx.get(), dy.get(), K);
        cn::check(hipGetLastError(), "spmm launch");
        cn::d2h(Y, dy.get(), nk, c->s);
        cn::sync(c->s, "spmm");
    });
}

int32_t cn_k_spmm_diff(cn_context* c, const float* X, float* Y, int32_t K) {
    return guarded([&] {
        cn::require(c != nullptr, "null context");
        cn::require(cn::padded_width(K) == K, "K must be a compiled block width");
        const int64_t nk = (c->n - 1) * K;
        auto dx = cn::upload(X, nk, c->s, "X");
        cn::DeviceBuffer<float> dy(nk, "Y");
        c->solver->spmm_diff(dx.get(), dy.get(), K);
        cn::check(hipGetLastError(), "spmm launch");
        cn::d2h(Y, dy.get(), nk, c->s);
        cn::sync(c->s, "spmm");
    });
}

int32_t cn_k_gram(cn_context* c, const float* A, const float* B1, const float* B2, double* out, int32_t K) {
    return guarded([&] {
        cn::require(c != nullptr, "null context");
        cn::require(A != nullptr && B1 != nullptr && out != nullptr, "null operand");
        cn::require(cn::padded_width(K) == K, "K must be a compiled block width");
        const int64_t nk = (c->n - 1) * K;
        auto da = cn::upload(A, nk, c->s, "A");
        auto db1 = cn::upload(B1, nk, c->s, "B1");
        cn::DeviceBuffer<float> db2;
        if (B2 != nullptr) db2 = cn::upload(B2, nk, c->s, "B2");
        c->solver->gram(da.get(), db1.get(), B2 != nullptr ? db2.get() : nullptr, out, K);
    });
}

int32_t cn_k_vcycle(cn_context* c, const float* B, float* X, int32_t K) {
    return guarded([&] {
        cn::require(c != nullptr, "null context");
        cn::require(cn::padded_width(K) == K, "K must be a compiled block width");
        const int64_t nk = (c->n - 1) * K;
        auto db = cn::upload(B, nk, c->s, "B");
        cn::DeviceBuffer<float> dx(nk, "X");
        c->solver->vcycle(db.get(), dx.get(), K);
        cn::check(hipGetLastError(), "vcycle launch");
        cn::d2h(X, dx.get(), nk, c->s);
        cn::sync(c->s, "vcycle");
    });
}

int32_t cn_k_dadt(const float* s_, const float* mp, const float* sn, int64_t n_dip, const float* r, int64_t n,
                  float didt, float mu0_4pi, float* out) {
    return guarded([&] {
        hipStream_t s = nullptr;
        std::vector<float> pack(static_cast<std::size_t>(10 * n_dip));
        for (int64_t j = 0; j < n_dip; ++j) {
            for (int q = 0; q < 3; ++q) pack[10 * j + q] = s_[3 * j + q];
            for (int q = 0; q < 6; ++q) pack[10 * j + 3 + q] = mp[6 * j + q];
            pack[10 * j + 9] = sn[j];
        }
        std::vector<double> nodes_mm(static_cast<std::size_t>(3 * n));
        for (int64_t i = 0; i < 3 * n; ++i) nodes_mm[i] = static_cast<double>(r[i]) * 1e3;
        auto dd = cn::upload(pack.data(), pack.size(), s, "dipoles");
        auto dn = cn::upload(nodes_mm.data(), nodes_mm.size(), s, "targets");
        const double zero[3] = {0.0, 0.0, 0.0};
        auto dorig = cn::upload(zero, 3, s, "origin");
        cn::DeviceBuffer<float> dout(std::max<int64_t>(3 * n, 1), "out");
        cn::dadt_nbody(dd.get(), static_cast<int>(n_dip), dn.get(), n, dorig.get(), 1, 1, didt * mu0_4pi, dout.get(), s);
        cn::d2h(out, dout.get(), 3 * n, s);
        cn::sync(s, "dadt");
    });
}

int32_t cn_k_place(cn_context* c, const double* centers, const double* handles, const double* dists, int32_t n_pl,
                   double* out, int32_t* degenerate) {
    return guarded([&] {
        cn::require(c != nullptr && n_pl >= 1, "bad argument");
        std::vector<std::int32_t> deg(n_pl);
        place_block(*c, centers, handles, dists, n_pl, out, deg.data());
        if (degenerate) std::memcpy(degenerate, deg.data(), sizeof(std::int32_t) * n_pl);
    });
}

int32_t cn_bench_spmm(cn_context* c, int32_t K, int32_t reps, int32_t mode, int32_t sub, double* ms) {
    return guarded([&] {
        cn::require(c != nullptr && reps >= 1 && ms != nullptr, "bad argument");
        *ms = c->solver->bench_spmm(K, reps, mode, sub);
    });
}

int32_t cn_bench_vcycle(cn_context* c, int32_t K, int32_t reps, double* ms) {
    return guarded([&] {
        cn::require(c != nullptr && reps >= 1, "bad argument");
        *ms = c->solver->bench_vcycle(K, reps);
    });
}

int32_t cn_bench_guess(cn_context* c, int32_t K, int32_t reps, double* ms3) {
    return guarded([&] {
        cn::require(c != nullptr && ms3 != nullptr && reps >= 1, "bad argument");
        cn::require(K >= 1 && K <= 16 && cn::padded_width(K) == K, "K must be a compiled block width up to 16");
        cn::require(c->d_B.size() >= static_cast<std::size_t>((c->n - 1) * K),
                    "solve a block of this width first (its right-hand sides are the projected block)");
        c->d_x0.ensure((c->n - 1) * K, "x0");
        c->warm.bench(c->d_B.get(), K, K, c->d_x0.get(), reps, ms3, c->s);
    });
}

int32_t cn_bench_stages(cn_context* c, const double* dip_pos_m, const double* dip_moment, int64_t n_dip,
                        const double* center_mm, const double* handle_mm, double distance_mm, int32_t k,
                        int32_t recovery, int32_t reps, double* ms3) {
    return guarded([&] {
        using namespace cn;
        require(c != nullptr && dip_pos_m && dip_moment && center_mm && ms3 && n_dip >= 1, "bad argument");
        require(k >= 1 && k <= kMaxBlock && reps >= 1, "bad argument");
        const hipStream_t s = c->s;
        const int K = padded_width(k);
        const std::int64_t n = c->n, m = c->m, nr = n - 1;
        std::vector<double> tf(16);
        std::vector<std::int32_t> degen(1);
        place_block(*c, center_mm, handle_mm, &distance_mm, 1, tf.data(), degen.data());
        std::vector<double> frames(16 * static_cast<std::size_t>(k));
        for (int i = 0; i < k; ++i) std::copy(tf.begin(), tf.end(), frames.begin() + 16 * i);
        c->d_dadt_nodes.ensure(3 * n * K, "dadt nodes");
        c->d_q.ensure(4 * m * K, "corner terms");
        c->d_B.ensure(nr * K, "rhs block");
        c->d_X.ensure(nr * K, "solution");
        c->d_magn.ensure(m * K, "magnE");
        check(hipMemsetAsync(c->d_X.get(), 0, sizeof(double) * nr * K, s), "potential clear");
        const RecoveryOp* rec = recovery == CN_RECOVERY_RAW ? nullptr : ensure_recovery(*c, recovery);
        if (rec) c->d_slots.ensure(3 * rec->n_slots * K, "recovered slots");
        if (rec && rec->mode != kRecoveryHarmonic) c->d_Eraw.ensure(3 * m * K, "raw E");
        std::vector<double> met(static_cast<std::size_t>(kMetricsLen) * k);
        auto coil = [&] {
            dadt_for_frames(*c, frames.data(), k, K, dip_pos_m, dip_moment, n_dip, 0.1f, c->d_dadt_nodes.get(), s);
            element_pass(c->d_dadt_nodes.get(), c->tets.get(), c->g32.get(), c->neg_vc.get(), m, k, K, nullptr,
                         c->d_q.get(), c->corner_pos.get(), s, c->vols32.get(), c->tags.get(), nullptr, 0);
            node_rhs(c->d_q.get(), c->n2c_ptr.get(), n, k, K, c->ground, c->d_B.get(), nullptr, s);
        };
        auto field = [&] {
            if (rec == nullptr)
                reconstruct(c->d_X.get(), c->tets.get(), c->g32.get(), c->d_dadt_nodes.get(), K, m, k, K, c->ground,
                            nullptr, c->d_magn.get(), s);
            else if (rec->mode == kRecoveryHarmonic)
                apply_harmonic(*rec, c->d_X.get(), c->ground, c->d_dadt_nodes.get(), K, m, k, K, c->d_slots.get(),
                               nullptr, c->d_magn.get(), s);
            else {
                reconstruct(c->d_X.get(), c->tets.get(), c->g32.get(), c->d_dadt_nodes.get(), K, m, k, K, c->ground,
                            c->d_Eraw.get(), c->d_magn.get(), s);
                apply_spr(*rec, c->d_Eraw.get(), m, k, K, c->d_slots.get(), nullptr, c->d_magn.get(), s);
            }
        };
        auto metrics = [&] {
            region_metrics(c->d_magn.get(), m, k, K, c->region_idx.get(), c->n_region, c->vols32.get(), c->bary.get(),
                           c->metrics_ws, met.data(), s);
        };
        hipEvent_t e[2];
        for (auto& x : e) check(hipEventCreate(&x), "event");
        auto timed = [&](auto&& f) {
            f();  // untimed first run
            sync(s, "stage warmup");
            check(hipEventRecord(e[0], s), "event");
            for (int r = 0; r < reps; ++r) f();
            check(hipEventRecord(e[1], s), "event");
            check(hipEventSynchronize(e[1]), "event");
            float t = 0.f;
            check(hipEventElapsedTime(&t, e[0], e[1]), "elapsed");
            return static_cast<double>(t) / reps;
        };
        ms3[0] = timed(coil);
        ms3[1] = timed(field);
        ms3[2] = timed(metrics);
        for (auto& x : e) (void)hipEventDestroy(x);
    });
}

int32_t cn_k_gram_noalloc(int64_t n, int32_t K, const float* A, const float* B1, const float* B2, double* out,
                          void* stream) {
    return guarded([&] { cn::gram_noalloc(n, K, A, B1, B2, out, static_cast<hipStream_t>(stream)); });
}

int32_t cn_context_setup_times(const cn_context* c, double* ms10) {
    return guarded([&] {
        cn::require(c != nullptr && ms10 != nullptr, "null argument");
        for (int i = 0; i < 10; ++i) ms10[i] = c->setup_ms[i];
    });
}

int32_t cn_debug_fail_captures(cn_context* c, int32_t n) {
    return guarded([&] {
        cn::require(c != nullptr && n >= 0, "bad argument");
        c->solver->fail_next_captures(n);
    });
}

int32_t cn_graph_execs(const cn_context* c, int32_t* out) {
    return guarded([&] {
        cn::require(c != nullptr && out != nullptr, "bad argument");
        *out = c->solver->graph_execs();
    });
}

int32_t cn_timing_enable(cn_context* c, int32_t on) {
    return guarded([&] {
        cn::require(c != nullptr, "null context");
        c->solver->set_timing(on != 0);
        for (auto& v : c->phase_ms) v = 0.0;
        c->timed_blocks = 0;
        c->timed_placements = 0;
    });
}

int32_t cn_timing_read(const cn_context* c, double* out) {
    return guarded([&] {
        cn::require(c != nullptr && out != nullptr, "null argument");
        c->solver->timing(out);
        out[4] = static_cast<double>(c->timed_placements);
        out[5] = c->timed_placements ? c->phase_ms[1] / static_cast<double>(c->timed_placements) : 0.0;
    });
}

int32_t cn_timing_phases(const cn_context* c, double* out) {
    return guarded([&] {
        cn::require(c != nullptr && out != nullptr, "null argument");
        for (int i = 0; i < cn_context::kPhases; ++i) out[i] = c->phase_ms[i];
        out[7] = static_cast<double>(c->timed_blocks);
        out[8] = static_cast<double>(c->timed_placements);
    });
}

int32_t cn_bench_dadt(cn_context* c, const double* dip_pos_m, const double* dip_moment, int64_t n_dip,
                      const double* center_mm, const double* handle_mm, double distance_mm, int32_t k,
                      int32_t reps, double* ms) {
    return guarded([&] {
        cn::require(c != nullptr && dip_pos_m && dip_moment && center_mm && ms && n_dip >= 1, "bad argument");
        cn::require(k >= 1 && k <= cn::kMaxBlock && reps >= 1, "bad argument");
        const hipStream_t s = c->s;
        std::vector<double> tf(16);
        std::vector<std::int32_t> degen(1);
        place_block(*c, center_mm, handle_mm, &distance_mm, 1, tf.data(), degen.data());
        std::vector<double> frames(16 * static_cast<std::size_t>(k));
        for (int i = 0; i < k; ++i) std::copy(tf.begin(), tf.end(), frames.begin() + 16 * i);
        const int K = cn::padded_width(k);
        cn::DeviceBuffer<float> out(3 * c->n * K, "dadt");
        auto run = [&] {
            dadt_for_frames(*c, frames.data(), k, K, dip_pos_m, dip_moment, n_dip, 1e-7f, out.get(), s);
        };
        run();
        hipEvent_t e0, e1;
        cn::check(hipEventCreate(&e0), "event");
        cn::check(hipEventCreate(&e1), "event");
        cn::check(hipEventRecord(e0, s), "event");
        for (int r = 0; r < reps; ++r) run();
        cn::check(hipEventRecord(e1, s), "event");
        cn::check(hipEventSynchronize(e1), "event");
        float t = 0.f;
        cn::check(hipEventElapsedTime(&t, e0, e1), "elapsed");
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms = static_cast<double>(t) / reps;
    });
}

int32_t cn_bench_copy(cn_context* c, int64_t bytes, int32_t reps, double* ms) {
    return guarded([&] {
        cn::require(c != nullptr && reps >= 1 && ms != nullptr, "bad argument");
        *ms = c->solver->bench_copy(bytes, reps);
    });
}


int32_t cn_recovery_build(cn_context* c, int32_t mode, int64_t* n_slots, int32_t* n_fallback) {
    return guarded([&] {
        cn::require(c != nullptr, "null context");
        const cn::RecoveryOp* op = ensure_recovery(*c, mode);
        if (n_slots) *n_slots = op->n_slots;
        if (n_fallback) *n_fallback = op->n_fallback;
    });
}

int32_t cn_recovery_operator(const cn_context* c, int32_t mode, int32_t* ptr, int32_t* idx, float* w,
                             int32_t* slot_of_corner, int32_t* slot_node, int32_t* slot_tag, int32_t* boundary,
                             int32_t* status) {
    return guarded([&] {
        cn::require(c != nullptr && mode >= 1 && mode <= 3 && c->recovery[mode] != nullptr,
                    "recovery operator not built (call cn_recovery_build first)");
        const cn::RecoveryOp& op = *c->recovery[mode];
        const hipStream_t s = c->s;
        const std::int64_t nc = 4 * c->m;
        if (ptr) cn::d2h(ptr, op.ptr_view, op.n_slots + 1, s);
        if (idx) cn::d2h(idx, op.idx_view, nc, s);
        if (w) cn::d2h(w, op.w.get(), nc, s);
        if (slot_of_corner) cn::d2h(slot_of_corner, op.soc_view, nc, s);
        if (slot_node) cn::d2h(slot_node, op.slot_node.get(), op.n_slots, s);
        if (slot_tag && mode != CN_RECOVERY_SPR_GLOBAL) cn::d2h(slot_tag, op.slot_tag.get(), op.n_slots, s);
        if (boundary && mode != CN_RECOVERY_HARMONIC) cn::d2h(boundary, op.boundary.get(), c->n, s);
        if (status) cn::d2h(status, op.status.get(), op.n_slots, s);
        cn::sync(s, "recovery operator copy");
    });
}

int32_t cn_warm_reset(cn_context* c, int32_t capacity) {
    return guarded([&] {
        cn::require(c != nullptr, "null context");
        c->warm.reset(c->n - 1, capacity);
    });
}

int32_t cn_warm_count(const cn_context* c, int32_t* count) {
    return guarded([&] {
        cn::require(c != nullptr && count != nullptr, "null argument");
        *count = c->warm.count;
    });
}

int32_t cn_warm_gram(cn_context* c, double* G) {
    return guarded([&] {
        cn::require(c != nullptr && G != nullptr, "null argument");
        const int S = c->warm.count;
        cn::require(S >= 1, "the warm-start basis is empty");
        const std::int64_t nr = c->n - 1;
        const int chunks = (S + 31) / 32;
        std::vector<cn::DeviceBuffer<float>> w(chunks);
        for (int i = 0; i < chunks; ++i) {
            w[i].reset(static_cast<std::size_t>(nr) * 32, "basis chunk");
            c->warm.chunk(32 * i, w[i].get(), c->s);
        }
        cn::DeviceBuffer<float> aw(static_cast<std::size_t>(nr) * 32, "A W chunk");
        std::vector<double> tile(32 * 32);
        // G is symmetric: tiles i <= j are computed, the others mirrored.
        for (int j = 0; j < chunks; ++j) {
            c->solver->spmm_diff(w[j].get(), aw.get(), 32);
            for (int i = 0; i <= j; ++i) {
                c->solver->cross_gram32(w[i].get(), aw.get(), tile.data());
                for (int a = 0; a < 32 && 32 * i + a < S; ++a)
                    for (int b = 0; b < 32 && 32 * j + b < S; ++b) {
                        G[static_cast<std::int64_t>(32 * i + a) * S + 32 * j + b] = tile[a * 32 + b];
                        G[static_cast<std::int64_t>(32 * j + b) * S + 32 * i + a] = tile[a * 32 + b];
                    }
            }
        }
    });
}

int32_t cn_warm_set_modes(cn_context* c, const double* M, int32_t count, int32_t R) {
    return guarded([&] {
        cn::require(c != nullptr && M != nullptr, "null argument");
        c->warm.set_modes(M, count, R, c->s);
    });
}

// Phi^T A Phi and Phi^T A W through 32-column chunks: differenced SpMM, then FP64 cross Grams.
int32_t cn_warm_mode_grams(cn_context* c, double* G_modes, double* G_cross) {
    return guarded([&] {
        cn::require(c != nullptr && c->warm.ready, "warm-start modes not set");
        const int R = c->warm.used_modes, S = c->warm.count;
        const std::int64_t nr = c->n - 1;
        const int rc = (R + 31) / 32, sc = (S + 31) / 32;
        std::vector<cn::DeviceBuffer<float>> phi(rc);
        for (int i = 0; i < rc; ++i) {
            phi[i].reset(static_cast<std::size_t>(nr) * 32, "mode chunk");
            c->warm.mode_chunk(32 * i, phi[i].get(), c->s);
        }
        cn::DeviceBuffer<float> x(static_cast<std::size_t>(nr) * 32, "chunk"), ax(static_cast<std::size_t>(nr) * 32, "A chunk");
        std::vector<double> tile(32 * 32);
        auto fill = [&](double* out, int cols_total, int rows_total, int ci, int cj) {
            for (int a = 0; a < 32 && 32 * ci + a < rows_total; ++a)
                for (int b = 0; b < 32 && 32 * cj + b < cols_total; ++b)
                    out[static_cast<std::int64_t>(32 * ci + a) * cols_total + 32 * cj + b] = tile[a * 32 + b];
        };
        if (G_modes) {
            for (int j = 0; j < rc; ++j) {
                c->solver->spmm_diff(phi[j].get(), ax.get(), 32);
                for (int i = 0; i < rc; ++i) {
                    c->solver->cross_gram32(phi[i].get(), ax.get(), tile.data());
                    fill(G_modes, R, R, i, j);
                }
            }
        }
        if (G_cross) {
            for (int j = 0; j < sc; ++j) {
                c->warm.chunk(32 * j, x.get(), c->s);
                c->solver->spmm_diff(x.get(), ax.get(), 32);
                for (int i = 0; i < rc; ++i) {
                    c->solver->cross_gram32(phi[i].get(), ax.get(), tile.data());
                    fill(G_cross, S, R, i, j);
                }
            }
        }
    });
}

int32_t cn_warm_set_projector(cn_context* c, const double* P, int32_t R) {
    return guarded([&] {
        cn::require(c != nullptr && P != nullptr, "null argument");
        c->warm.set_projector(P, R, c->s);
    });
}

int32_t cn_warm_set_coefficients(cn_context* c, const double* coef, int32_t k, int32_t R) {
    return guarded([&] {
        cn::require(c != nullptr && coef != nullptr, "null argument");
        cn::require(c->warm.ready && R == c->warm.used_modes, "coefficients must match the current modes");
        cn::require(k >= 1 && k <= cn::kMaxBlock, "k must be in [1, 32]");
        c->warm_coef.assign(coef, coef + static_cast<std::size_t>(k) * R);
        c->warm_coef_k = k;
    });
}

int32_t cn_warm_project_last(cn_context* c, int32_t k, double* out) {
    return guarded([&] {
        cn::require(c != nullptr && out != nullptr, "null argument");
        cn::require(c->warm.ready, "warm-start modes not set");
        cn::require(k >= 1 && k <= cn::kMaxBlock, "k must be in [1, 32]");
        const int K = cn::padded_width(k);
        cn::require(c->d_B.size() >= static_cast<std::size_t>((c->n - 1) * K),
                    "no block of that width has been solved");
        c->warm.project_host(c->d_B.get(), k, K, out, c->s);
    });
}

int32_t cn_warm_modes(const cn_context* c, uint16_t* out, int32_t* width) {
    return guarded([&] {
        cn::require(c != nullptr && c->warm.ready, "warm-start modes not set");
        if (width) *width = c->warm.n_modes;
        if (out) {
            cn::d2h(out, c->warm.modes.get(), static_cast<std::size_t>(2) * (c->n - 1) * c->warm.n_modes, c->s);
            cn::sync(c->s, "modes copy");
        }
    });
}

struct cn_mlp {
    hipStream_t s = nullptr;
    std::unique_ptr<cn::Mlp> net;
    ~cn_mlp() {
        net.reset();
        if (s) (void)hipStreamDestroy(s);
    }
};

int32_t cn_mlp_create(int32_t n_in, int32_t hidden, int32_t n_out, uint32_t seed, cn_mlp** out) {
    return guarded([&] {
        cn::require(out != nullptr, "null argument");
        auto m = std::make_unique<cn_mlp>();
        cn::check(hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking), "create stream");
        m->net = std::make_unique<cn::Mlp>(n_in, hidden, n_out, seed, m->s);
        *out = m.release();
    });
}

void cn_mlp_destroy(cn_mlp* net) { delete net; }

int32_t cn_mlp_n_params(const cn_mlp* net, int64_t* n) {
    return guarded([&] {
        cn::require(net != nullptr && n != nullptr, "null argument");
        *n = net->net->n_params();
    });
}

int32_t cn_mlp_set_params(cn_mlp* net, const float* params) {
    return guarded([&] {
        cn::require(net != nullptr && params != nullptr, "null argument");
        net->net->set_params(params);
    });
}

int32_t cn_mlp_get_params(const cn_mlp* net, float* params) {
    return guarded([&] {
        cn::require(net != nullptr && params != nullptr, "null argument");
        net->net->get_params(params);
    });
}

int32_t cn_mlp_gradients(cn_mlp* net, const float* X, const float* T, int32_t B, float* grads, double* loss) {
    return guarded([&] {
        cn::require(net != nullptr && X != nullptr && T != nullptr && grads != nullptr && B >= 1, "bad argument");
        const double l = net->net->gradients(X, T, B, grads);
        if (loss) *loss = l;
    });
}

int32_t cn_mlp_train(cn_mlp* net, const float* X, const float* T, int64_t N, int32_t epochs, int32_t batch, float lr,
                     uint32_t seed, double* loss) {
    return guarded([&] {
        cn::require(net != nullptr && X != nullptr && T != nullptr, "null argument");
        const double l = net->net->train(X, T, N, epochs, batch, lr, seed);
        if (loss) *loss = l;
    });
}

int32_t cn_mlp_predict(cn_mlp* net, const float* X, int32_t B, float* Y) {
    return guarded([&] {
        cn::require(net != nullptr && X != nullptr && Y != nullptr && B >= 1, "bad argument");
        net->net->predict(X, B, Y);
    });
}

int32_t cn_recovery_patches(const cn_context* c, int64_t* nnz, int32_t* n_undetermined, int32_t* pptr,
                            int32_t* pidx, float* w3) {
    return guarded([&] {
        cn::require(c != nullptr && c->recovery[CN_RECOVERY_HARMONIC] != nullptr,
                    "harmonic operator not built (call cn_recovery_build first)");
        const cn::RecoveryOp& op = *c->recovery[CN_RECOVERY_HARMONIC];
        if (nnz) *nnz = op.nnz_patch;
        if (n_undetermined) *n_undetermined = op.n_undetermined;
        if (pptr) cn::d2h(pptr, op.pptr.get(), op.n_slots + 1, c->s);
        if (pidx) cn::d2h(pidx, op.pidx.get(), op.nnz_patch, c->s);
        if (w3) cn::d2h(w3, op.w3.get(), 3 * op.nnz_patch, c->s);
        cn::sync(c->s, "patch copy");
    });
}

int32_t cn_uq_guess_info(const cn_context* c, int32_t* out) {
    return guarded([&] {
        cn::require(c != nullptr && out != nullptr, "null argument");
        cn::require(c->uq.ready, "call cn_uq_prepare first");
        out[0] = c->uq.nominal_iterations;
        out[1] = c->uq.sensitivity_iterations;
        out[2] = c->uq.guess_rank;
    });
}

int32_t cn_uq_prepare(cn_context* c, const double* pos_m, const double* mom, int64_t n_dip, const double* center,
                      const double* handle, double distance, double didt, int32_t* n_tissue, int32_t* tags) {
    return guarded([&] {
        cn::require(c != nullptr && pos_m != nullptr && mom != nullptr && center != nullptr && n_dip >= 1,
                    "bad argument");
        uq_prepare(*c, pos_m, mom, n_dip, center, handle, distance, didt);
        if (n_tissue) *n_tissue = static_cast<int32_t>(c->tissue.tags.size());
        if (tags)
            for (std::size_t i = 0; i < c->tissue.tags.size(); ++i) tags[i] = c->tissue.tags[i];
    });
}

int32_t cn_uq_solve(cn_context* c, const double* sigma, int32_t k, const cn_solve_options* opt, cn_block_outputs* out) {
    return guarded([&] {
        cn::require(c != nullptr && sigma != nullptr, "null argument");
        uq_solve(*c, sigma, k, opt, out);
    });
}

int32_t cn_uq_moments(const cn_context* c, double* mean, double* var, int64_t* count) {
    return guarded([&] {
        cn::require(c != nullptr && c->uq.ready, "call cn_uq_prepare first");
        const std::int64_t m = c->m, cnt = c->uq.count;
        if (count) *count = cnt;
        if (!mean && !var) return;
        cn::require(cnt >= 1, "no samples solved yet");
        std::vector<double> s1(m), s2(m);
        cn::d2h(s1.data(), c->uq.sum.get(), m, c->s);
        cn::d2h(s2.data(), c->uq.sumsq.get(), m, c->s);
        cn::sync(c->s, "moments copy");
        for (std::int64_t e = 0; e < m; ++e) {
            const double mu = s1[e] / static_cast<double>(cnt);
            if (mean) mean[e] = mu;
            if (var) var[e] = cnt > 1 ? std::max(0.0, (s2[e] - static_cast<double>(cnt) * mu * mu) / (cnt - 1)) : 0.0;
        }
    });
}

}  // extern "C"
