// Device data structures and host launchers of the library's kernels.
#pragma once

#include <climits>
#include <cmath>
#include <utility>
#include <vector>

#include "common.hpp"

namespace cn {

// Minimum over the lanes of a wavefront, returned to every lane.
__device__ __forceinline__ int wave_min_i32(int v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, kWave));
    return v;
}

// Non-owning CSR (int32 indices, FP32 values) handed to kernels.
struct CsrView {
    const std::int32_t* ptr = nullptr;
    const std::int32_t* col = nullptr;
    const float* val = nullptr;
    std::int64_t n = 0;       // rows
    std::int64_t n_cols = 0;  // columns
    std::int64_t nnz = 0;
    // Optional copy of the same rows in sliced ELL form: slices of kEllSlice consecutive rows,
    // each row padded to ell_w slots, slot j of the slice's rows contiguous, (column, value bits)
    // pairs; a padding slot names the row itself with value 0. Entry (row r, slot j) is at
    // (r / kEllSlice) * kEllSlice * ell_w + j * kEllSlice + r % kEllSlice.
    const int2* ell = nullptr;
    int ell_w = 0;
    int ell_max_k = 0;  // the row kernels take the ELL copy for block widths up to this
};

constexpr int kEllSlice = 64;
// Widest block for which the sliced ELL copy is the default: at 1, 2 and 4 columns the row
// kernels are bound by their index and value loads, which the ELL copy issues without waiting on
// row pointers (the 1.5 mm head's plain solve: SpMM 34 -> 23 us, V-cycle 252 -> 209 us); at 8 and
// 16 the gathers of whole W rows dominate and the CSR arrays measured as fast or faster.
constexpr int kEllAutoMaxK = 4;
constexpr int kEllMaxWidth = 24;  // rows longer than this keep the CSR kernels

// Owning CSR.
struct CsrF {
    DeviceBuffer<std::int32_t> ptr, col;
    DeviceBuffer<float> val;
    std::int64_t n = 0, n_cols = 0, nnz = 0;
    CsrView view() const { return {ptr.get(), col.get(), val.get(), n, n_cols, nnz}; }
};

struct AmgParams {
    int min_coarse_rows = 128;  // fem/solve.py:146
    int max_levels = 50;        // fem/solve.py:147
    int rounds = 2;             // fem/solve.py:150
    int degree = 2;             // fem/solve.py:167
    float lower_ratio = 12.0f;  // fem/solve.py:168
};

constexpr int kMaxDegree = 6;

// Chebyshev x-recurrence coefficients over [1/lower_ratio, 1] (solver/amg/vcycle.cu:434-463).
struct Cheby {
    int degree = 2;
    float alpha0 = 0.f;
    float c_cur[kMaxDegree] = {};
    float c_prev[kMaxDegree] = {};
    float alpha[kMaxDegree] = {};
};

Cheby chebyshev(int degree, float lower_ratio);

struct Level {
    CsrView a;     // this level's operator (level 0: the context's reduced stiffness)
    CsrF a_store;  // owns `a` on the coarser levels
    DeviceBuffer<float> dinv;
    DeviceBuffer<std::int32_t> agg;
    CsrF p, r;
    std::int64_t n_coarse = 0;
};

struct Hierarchy {
    std::vector<Level> levels;
    CsrF coarse;
    DeviceBuffer<double> coarse_inv;  // FP64 dense inverse of the coarsest operator
    Cheby cheb;
    std::int64_t n_coarse() const { return coarse.n; }
};

// ---------------- mesh.hip ----------------
void morton_keys(const double* pts, std::int64_t n, std::uint64_t* keys, hipStream_t s);
void tet_lowest_node(const std::int32_t* inverse, const std::int32_t* tets, std::int64_t m,
                     std::uint32_t* lowest, hipStream_t s);
void spatial_order(const double* nodes, std::int64_t n, const std::int32_t* tets, std::int64_t m,
                   std::int32_t* node_perm, std::int32_t* tet_perm, std::int32_t* inverse, hipStream_t s);
void apply_spatial_order(const double* nodes_in, const std::int32_t* tets_in, const std::int32_t* tags_in,
                         std::int64_t n, std::int64_t m, const std::int32_t* node_perm,
                         const std::int32_t* tet_perm, const std::int32_t* inverse, double* nodes_out,
                         std::int32_t* tets_out, std::int32_t* tags_out, std::int32_t* skin,
                         std::int64_t n_skin, hipStream_t s);
void p1_gradients(const double* nodes, const std::int32_t* tets, std::int64_t m, double* g, double* vols,
                  hipStream_t s);
int count_bad_volumes(const double* vols, std::int64_t m, hipStream_t s);
int conductivity(const std::int32_t* tags, std::int64_t m, const double* lut_host, int lut_len, double* cond,
                 hipStream_t s);
void node2corner(const std::int32_t* elems, std::int64_t n_elem, int corners, std::int64_t n_nodes,
                 std::int32_t* ptr, std::int32_t* idx, hipStream_t s);
std::int64_t stiffness_pattern(const std::int32_t* tets, const std::int32_t* ptr, const std::int32_t* idx,
                               std::int64_t n, DeviceBuffer<std::int32_t>& indptr,
                               DeviceBuffer<std::int32_t>& indices, hipStream_t s);
void stiffness_values(const double* g, const double* scale, const std::int32_t* tets, const std::int32_t* ptr,
                      const std::int32_t* idx, std::int64_t n, const std::int32_t* indptr,
                      const std::int32_t* indices, double* data, hipStream_t s);
int lowest_z_node(const double* nodes, std::int64_t n, hipStream_t s);
std::int64_t reduce_grounded(const std::int32_t* indptr, const std::int32_t* indices, const double* data,
                             std::int64_t n, int ground, DeviceBuffer<std::int32_t>& rptr,
                             DeviceBuffer<std::int32_t>& rcol, DeviceBuffer<double>& rval64,
                             DeviceBuffer<float>& rval32, hipStream_t s);
void resident_arrays(const double* g64, const double* vols64, const double* cond, const double* nodes,
                     const std::int32_t* tets, std::int64_t m, float* g32, float* vols32, float* neg_vc,
                     double* bary, hipStream_t s);
void skin_normals(const double* nodes, const std::int32_t* tris, std::int64_t n_tri, std::int64_t n_nodes,
                  double* out, hipStream_t s);
void triangle_vertices(const double* nodes, const std::int32_t* tris, std::int64_t n_tri, double* a,
                       double* b, double* c, hipStream_t s);

// ---------------- amg.hip ----------------
void l1_dinv(const CsrView& a, float* dinv, hipStream_t s);
void agg_edge_weights(const CsrView& a, float* w, hipStream_t s);
std::int64_t select_size4(const CsrView& a, std::int32_t* agg, hipStream_t s);
void compose_aggregates(std::int32_t* agg, std::int64_t n, const std::int32_t* step, hipStream_t s);
CsrF spgemm(const CsrView& a, const CsrView& b, hipStream_t s);
CsrF transpose(const CsrView& a, hipStream_t s);
CsrF tentative_prolongator(const std::int32_t* agg, std::int64_t n, std::int64_t n_coarse, hipStream_t s);
CsrF smoothed_prolongator(const CsrView& a, const std::int32_t* agg, std::int64_t n_coarse, const float* dinv,
                          hipStream_t s);
CsrF galerkin(const CsrView& a, const CsrView& p, hipStream_t s);
void coarse_inverse(const CsrView& a, DeviceBuffer<double>& inv, hipStream_t s);
std::pair<DeviceBuffer<std::int32_t>, std::int64_t> aggregate_level(const CsrView& a, const AmgParams& prm,
                                                                     hipStream_t s);
Hierarchy build_hierarchy(const CsrView& a, const AmgParams& prm, hipStream_t s);
// The same levels as `like` (its aggregates and coarse sizes) for another operator on the same
// pattern: l1 diagonals, smoothed prolongators, Galerkin products and the coarse inverse recomputed
// from `a`, no aggregation pass. The result keeps no aggregates.
Hierarchy build_hierarchy_like(const CsrView& a, const Hierarchy& like, const AmgParams& prm, hipStream_t s);

}  // namespace cn
