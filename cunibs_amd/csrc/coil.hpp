// Coil-side and post-processing launchers.
//
// Per-placement device arrays of a block are laid out placement-fastest with the padded block width
// K: nodal dA/dt (n, 3, K), element dA/dt (m, 3, K), corner terms (4m, K), E (m, 3, K), |E| (m, K).
// The K lanes that serve one node or element then touch contiguous memory. unpack() turns such an
// array into the placement-major (k, rows, inner) layout the C ABI hands out.
#pragma once
#include "kernels.hpp"

namespace cn {

// coil.hip
void place_transforms(const double* centers, const double* handles, const double* dists, int n_pl, const double* ta,
                      const double* tb, const double* tc, const double* tn, int n_tri, double* out, int* degenerate,
                      int n_cu, hipStream_t s);
// dip: per placement n_dip x [s(3) m(3) m x s(3) |s|^2] FP32, centred at that placement's origin.
void dadt_nbody(const float* dip, int n_dip, const double* nodes_mm, std::int64_t n, const double* origins, int k,
                int K, float scale, float* out, hipStream_t s);
// All moments along one direction per placement (dirs: k x 3): dip8 per placement n_dip x
// [-2 s(3) |s|^2 mu mu s(3)] FP32.
// max_blocks > 0 caps the grid (all k placements together) and the blocks walk the nodes: a
// smaller footprint for running beside other work.
void dadt_nbody_parallel(const float* dip8, int n_dip, const double* nodes_mm, std::int64_t n, const double* origins,
                         const float* dirs, int k, int K, float scale, float* out, hipStream_t s, int max_blocks = 0);
// Element average of the nodal dA/dt and the RHS corner terms; sigma_tag (K x n_sigma FP32, with
// vols and tags) gives each placement its own conductivities instead of neg_vc. The corner term of
// corner 4 e + i is stored at row corner_pos[4 e + i] of q: its position in node2corner order.
void element_pass(const float* nodal, const std::int32_t* tets, const float* g, const float* neg_vc, std::int64_t m,
                  int k, int K, float* elm, float* q, const std::int32_t* corner_pos, hipStream_t s,
                  const float* vols = nullptr, const std::int32_t* tags = nullptr, const float* sigma_tag = nullptr,
                  int n_sigma = 0);
// corner_pos[idx[p]] = p for the 4m corners of a node2corner list.
void corner_positions(const std::int32_t* idx, std::int64_t count, std::int32_t* corner_pos, hipStream_t s);
// b[node] = sum of the node's corner terms, q in node2corner order (rows ptr[node] .. ptr[node+1]).
void node_rhs(const float* q, const std::int32_t* ptr, std::int64_t n, int k, int K, int ground, float* B,
              float* b_full, hipStream_t s);

// post.hip
void scatter_potential(const double* X, std::int64_t n, int k, int K, int ground, double* v, hipStream_t s);
void reconstruct(const double* X, const std::int32_t* tets, const float* g, const float* nodal, int Kn,
                 std::int64_t m, int k, int K, int ground, float* E, float* magn, hipStream_t s);
// (rows, inner, K) -> (k, rows, inner)
void unpack(const float* in, std::int64_t rows, int inner, int k, int K, float* out, hipStream_t s);

constexpr int kMetricsLen = 16;
// Device buffers of the region summary, kept by the caller across blocks (grow-only).
struct MetricsWorkspace {
    DeviceBuffer<float> vals;                  // (k, n_region) gathered |E|
    DeviceBuffer<std::uint64_t> keys, sorted;  // placement << 32 | bits of |E|
    DeviceBuffer<std::int32_t> pos, order;     // volume bits (the sort payload), in sorted order
    DeviceBuffer<double> w, cw, totals;        // sorted volumes, their prefix sums, scan tile totals
    DeviceBuffer<double> part, small;          // block partials; per-placement scalars and results
    DeviceBuffer<unsigned char> scratch;       // radix sort temporary storage
};
// Region summary per placement (metrics.py:209-243) from |E| laid out (m, K), all k placements of
// the block in one pass.
void region_metrics(const float* magn, std::int64_t m, int k, int K, const std::int32_t* region_idx,
                    std::int64_t n_region, const float* vols, const double* bary, MetricsWorkspace& ws,
                    double* out_host, hipStream_t s);

}  // namespace cn
