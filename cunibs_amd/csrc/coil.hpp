// Coil-side and post-processing launchers.
#pragma once
#include "kernels.hpp"

namespace cn {

// coil.hip
void place_transforms(const double* centers, const double* handles, const double* dists, int n_pl, const double* ta,
                      const double* tb, const double* tc, const double* tn, int n_tri, double* out, int* degenerate,
                      int n_cu, hipStream_t s);
// dip: per placement n_dip x [s(3) m(3) m x s(3) |s|^2] FP32, centred at that placement's origin.
void dadt_nbody(const float* dip, int n_dip, const double* nodes_mm, std::int64_t n, const double* origins, int k,
                float scale, float* out, hipStream_t s);
void element_average(const float* nodal, const std::int32_t* tets, std::int64_t n, std::int64_t m, int k, float* elm,
                     hipStream_t s);
void assemble_rhs(const float* elm, const float* g, const float* neg_vc, const std::int32_t* ptr,
                  const std::int32_t* idx, std::int64_t n, std::int64_t m, int k, int K, int ground, float* B,
                  float* b_full, hipStream_t s);

// post.hip
void scatter_potential(const double* X, std::int64_t n, int k, int K, int ground, double* v, hipStream_t s);
void reconstruct(const double* X, const std::int32_t* tets, const float* g, const float* elm, std::int64_t n,
                 std::int64_t m, int k, int K, int ground, float* E, float* magn, hipStream_t s);

constexpr int kMetricsLen = 16;
// Gray-matter summary per placement (metrics.py:209-243): see include/cunibs_amd.h for the layout.
void region_metrics(const float* magn, std::int64_t m, int k, const std::int32_t* region_idx, std::int64_t n_region,
                    const float* vols, const double* bary, double* out_host, hipStream_t s);

}  // namespace cn
