// Patch recovery operators (recovery.hip).
#pragma once
#include <vector>

#include "kernels.hpp"

namespace cn {

// Mode numbers of the C ABI (CN_RECOVERY_*).
constexpr int kRecoveryRaw = 0;
constexpr int kRecoverySprGlobal = 1;
constexpr int kRecoverySprTissue = 2;
constexpr int kRecoveryHarmonic = 3;

// One SPR operator: a corner CSR over slots, a weight per corner, and the maps back to the mesh.
// For "spr_global" the CSR and the corner-to-slot map are the context's node2corner and
// connectivity (views, not copies).
struct RecoveryOp {
    int mode = kRecoveryRaw;
    std::int64_t n_slots = 0;
    int n_fallback = 0;  // slots that took the volume-weighted average
    DeviceBuffer<std::int32_t> ptr, idx, soc, slot_node, slot_tag, boundary, status;
    DeviceBuffer<float> w;
    std::uint64_t stride = 0;  // tissue slot key = node * stride + tag
    // harmonic mode: node patch per slot and three gradient weights per patch entry
    DeviceBuffer<std::int32_t> pptr, pidx;
    DeviceBuffer<float> w3;
    std::int64_t nnz_patch = 0;
    int n_undetermined = 0;
    const std::int32_t* ptr_view = nullptr;
    const std::int32_t* idx_view = nullptr;
    const std::int32_t* soc_view = nullptr;
};

// mark[node] = 1 on the outer boundary of the tetrahedral volume (faces with a single owner).
void outer_boundary_nodes(const std::int32_t* tets, std::int64_t m, const std::int32_t* n2c_ptr,
                          const std::int32_t* n2c_idx, std::int64_t n, std::int32_t* mark, hipStream_t s);
void build_spr(RecoveryOp& op, int mode, const double* nodes, const std::int32_t* tets, const std::int32_t* tags,
               const float* vols, const std::int32_t* n2c_ptr, const std::int32_t* n2c_idx, std::int64_t n,
               std::int64_t m, int max_tag, hipStream_t s);
void tissue_slots(RecoveryOp& op, const std::int32_t* tets, const std::int32_t* tags, std::int64_t n, std::int64_t m,
                  int max_tag, hipStream_t s);
// Harmonic recovery of the potential (fem/recovery.py:269-383, solver/fem/recovery.cu:290-610).
void build_harmonic(RecoveryOp& op, const double* nodes, const std::int32_t* tets, const std::int32_t* tags,
                    std::int64_t n, std::int64_t m, int max_tag, hipStream_t s);
// X (n-1, K) FP64 reduced potential with `ground` removed; dadt_nodes (n, 3, K) FP32.
// dadt_nodes is (n, 3, Kn): Kn = K gives each column its own dA/dt, Kn = 1 shares one.
void apply_harmonic(const RecoveryOp& op, const double* X, int ground, const float* dadt_nodes, int Kn, std::int64_t m,
                    int k, int K, float* slots, float* E_out, float* magn, hipStream_t s);
// E (m, 3, K) -> slots (n_slots, 3, K), E_out (m, 3, K) (may be null), magn (m, K).
void apply_spr(const RecoveryOp& op, const float* E, std::int64_t m, int k, int K, float* slots, float* E_out,
               float* magn, hipStream_t s);

}  // namespace cn
