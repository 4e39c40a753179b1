// Superconvergent patch recovery (SPR) of the per-element field, the "spr_global" and
// "spr_tissue" modes of fem/recovery.py (operator build fem/recovery.py:385-433, kernels
// solver/fem/recovery.cu:67-280 and 620-687 of the reference).
//
// A slot is what a patch is fitted around: a node ("spr_global") or a (node, tissue) pair
// ("spr_tissue"). Each slot owns a segment of a corner CSR (corner c = 4e + i). The fit of
// p(x) = [1, (x - x_n)/h] over the patch's barycentres, evaluated at the node, is linear in the
// element field, so the operator is one FP32 weight per corner:
//     E*[s] = sum_{c in s} w[c] E[c >> 2],        E_out[e] = mean of its four corner slots.
// Slots on the boundary of their patch volume (outer boundary; for tissue slots also any node
// carrying more than one tag), slots with fewer than four elements, singular fits and fits whose
// Lebesgue constant sum|w| exceeds 8 take the volume-weighted average instead.
//
// Device layout: per-placement arrays placement-fastest, E (m, 3, K), slots (n_slots, 3, K), the
// same layout as the rest of the block pipeline. Every sum walks its CSR segment in order.
#include "recovery.hpp"

#include <algorithm>
#include <vector>

#include "prims.hpp"

namespace cn {

namespace {

constexpr double kMaxLebesgue = 8.0;

// One thread per face (4 per tet). A face is on the outer boundary when exactly one tetrahedron
// owns it; every owner holds each of the face's nodes, so the owners are found among the
// corners of whichever of the three nodes has the fewest. Marks are idempotent stores of 1.
__global__ void outer_boundary_kernel(const std::int32_t* __restrict__ tets, std::int64_t m,
                                      const std::int32_t* __restrict__ ptr, const std::int32_t* __restrict__ idx,
                                      std::int32_t* __restrict__ mark) {
    const std::int64_t f = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (f >= 4 * m) return;
    const std::int64_t e = f >> 2;
    const int skip = static_cast<int>(f & 3);
    int a[3];
    int q = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        if (i != skip) a[q++] = tets[4 * e + i];
    int pivot = a[0], o1 = a[1], o2 = a[2];
    int len = ptr[a[0] + 1] - ptr[a[0]];
    const int l1 = ptr[a[1] + 1] - ptr[a[1]], l2 = ptr[a[2] + 1] - ptr[a[2]];
    if (l1 < len) {
        len = l1;
        pivot = a[1];
        o1 = a[2];
        o2 = a[0];
    }
    if (l2 < len) {
        pivot = a[2];
        o1 = a[0];
        o2 = a[1];
    }
    int owners = 0;
    for (int j = ptr[pivot]; j < ptr[pivot + 1] && owners < 2; ++j) {
        const int4 t = *reinterpret_cast<const int4*>(tets + 4 * static_cast<std::int64_t>(idx[j] >> 2));
        const bool h1 = t.x == o1 || t.y == o1 || t.z == o1 || t.w == o1;
        const bool h2 = t.x == o2 || t.y == o2 || t.z == o2 || t.w == o2;
        owners += (h1 && h2) ? 1 : 0;
    }
    if (owners == 1) {
        mark[a[0]] = 1;
        mark[a[1]] = 1;
        mark[a[2]] = 1;
    }
}

__global__ void tissue_keys_kernel(const std::int32_t* __restrict__ tets, const std::int32_t* __restrict__ tags,
                                   std::int64_t m, std::uint64_t stride, std::uint64_t* __restrict__ keys) {
    const std::int64_t c = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c < 4 * m) keys[c] = static_cast<std::uint64_t>(tets[c]) * stride + static_cast<std::uint64_t>(tags[c >> 2]);
}

__global__ void head_flags_kernel(const std::uint64_t* __restrict__ keys, std::int64_t count,
                                  std::int32_t* __restrict__ flag) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

// pos = exclusive scan of the head flags: slot of sorted entry i is pos[i] + flag[i] - 1.
__global__ void slots_fill_kernel(const std::uint64_t* __restrict__ keys, const std::int32_t* __restrict__ flag,
                                  const std::int32_t* __restrict__ pos, const std::int32_t* __restrict__ corner,
                                  std::int64_t count, std::uint64_t stride, std::int32_t* __restrict__ ptr,
                                  std::int32_t* __restrict__ slot_of_corner, std::int32_t* __restrict__ slot_node,
                                  std::int32_t* __restrict__ slot_tag) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int slot = pos[i] + flag[i] - 1;
    slot_of_corner[corner[i]] = slot;
    if (flag[i]) {
        ptr[slot] = static_cast<std::int32_t>(i);
        slot_node[slot] = static_cast<std::int32_t>(keys[i] / stride);
        slot_tag[slot] = static_cast<std::int32_t>(keys[i] % stride);
    }
    if (i == count - 1) ptr[slot + 1] = static_cast<std::int32_t>(count);
}

// A node with slots in more than one tissue lies on each tissue's boundary.
__global__ void mixed_nodes_kernel(const std::int32_t* __restrict__ slot_node, std::int64_t n_slots,
                                   std::int32_t* __restrict__ mark) {
    const std::int64_t s = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s + 1 < n_slots && slot_node[s] == slot_node[s + 1]) mark[slot_node[s]] = 1;
}

__device__ __forceinline__ void barycentre(const double* __restrict__ nodes, const std::int32_t* __restrict__ tets,
                                           std::int64_t e, double b[3]) {
    b[0] = b[1] = b[2] = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const std::int64_t v = tets[4 * e + i];
        b[0] += nodes[3 * v];
        b[1] += nodes[3 * v + 1];
        b[2] += nodes[3 * v + 2];
    }
    b[0] *= 0.25;
    b[1] *= 0.25;
    b[2] *= 0.25;
}

// One thread per slot, FP64. Passes over the patch: the scale h = max |b - x_n| (component-wise),
// the 4 x 4 normal matrix in the centred, scaled basis, the solve A z = e_0 (Cholesky; a pivot at
// or below 1e-12 A_00 rejects the fit), the Lebesgue constant, then the weights
// w_c = z . [1, (b_c - x_n)/h]. status: 0 fitted, 1 averaged.
__global__ void spr_weights_kernel(const double* __restrict__ nodes, const std::int32_t* __restrict__ tets,
                                   const float* __restrict__ vols, const std::int32_t* __restrict__ ptr,
                                   const std::int32_t* __restrict__ idx, const std::int32_t* __restrict__ slot_node,
                                   const std::int32_t* __restrict__ boundary, std::int64_t n_slots,
                                   float* __restrict__ w, std::int32_t* __restrict__ status) {
    const std::int64_t s = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s >= n_slots) return;
    const std::int64_t node = slot_node ? slot_node[s] : s;
    const int begin = ptr[s], end = ptr[s + 1];
    const int count = end - begin;
    if (count <= 0) {
        status[s] = 1;
        return;
    }
    const double xn[3] = {nodes[3 * node], nodes[3 * node + 1], nodes[3 * node + 2]};
    bool fit = count >= 4 && boundary[node] == 0;
    double h = 0.0;
    if (fit) {
        for (int p = begin; p < end; ++p) {
            double b[3];
            barycentre(nodes, tets, idx[p] >> 2, b);
            for (int d = 0; d < 3; ++d) h = fmax(h, fabs(b[d] - xn[d]));
        }
        fit = h > 0.0;
    }
    double z[4] = {1.0, 0.0, 0.0, 0.0};
    if (fit) {
        double a[4][4];
        for (int i = 0; i < 4; ++i)
            for (int j = 0; j < 4; ++j) a[i][j] = 0.0;
        for (int p = begin; p < end; ++p) {
            double b[3];
            barycentre(nodes, tets, idx[p] >> 2, b);
            const double q[4] = {1.0, (b[0] - xn[0]) / h, (b[1] - xn[1]) / h, (b[2] - xn[2]) / h};
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) a[i][j] = fma(q[i], q[j], a[i][j]);
        }
        // Cholesky A = L L^T, then L y = e_0, L^T z = y.
        double l[4][4] = {};
        const double floor_pivot = 1e-12 * a[0][0];
        for (int j = 0; j < 4 && fit; ++j) {
            double d = a[j][j];
            for (int k = 0; k < j; ++k) d = fma(-l[j][k], l[j][k], d);
            if (!(d > floor_pivot)) {
                fit = false;
                break;
            }
            l[j][j] = sqrt(d);
            for (int i = j + 1; i < 4; ++i) {
                double t = a[i][j];
                for (int k = 0; k < j; ++k) t = fma(-l[i][k], l[j][k], t);
                l[i][j] = t / l[j][j];
            }
        }
        if (fit) {
            double y[4];
            for (int i = 0; i < 4; ++i) {
                double t = (i == 0) ? 1.0 : 0.0;
                for (int k = 0; k < i; ++k) t = fma(-l[i][k], y[k], t);
                y[i] = t / l[i][i];
            }
            for (int i = 3; i >= 0; --i) {
                double t = y[i];
                for (int k = i + 1; k < 4; ++k) t = fma(-l[k][i], z[k], t);
                z[i] = t / l[i][i];
            }
            double leb = 0.0;
            for (int p = begin; p < end; ++p) {
                double b[3];
                barycentre(nodes, tets, idx[p] >> 2, b);
                const double wc = z[0] + z[1] * (b[0] - xn[0]) / h + z[2] * (b[1] - xn[1]) / h +
                                  z[3] * (b[2] - xn[2]) / h;
                leb += fabs(wc);
            }
            fit = leb <= kMaxLebesgue;
        }
    }
    if (fit) {
        for (int p = begin; p < end; ++p) {
            double b[3];
            barycentre(nodes, tets, idx[p] >> 2, b);
            const double wc =
                z[0] + z[1] * (b[0] - xn[0]) / h + z[2] * (b[1] - xn[1]) / h + z[3] * (b[2] - xn[2]) / h;
            w[p] = static_cast<float>(wc);
        }
        status[s] = 0;
        return;
    }
    double total = 0.0;
    for (int p = begin; p < end; ++p) total += static_cast<double>(vols[idx[p] >> 2]);
    const bool by_volume = total > 0.0;
    const double inv = by_volume ? 1.0 / total : 1.0 / static_cast<double>(count);
    for (int p = begin; p < end; ++p) {
        const double num = by_volume ? static_cast<double>(vols[idx[p] >> 2]) : 1.0;
        w[p] = static_cast<float>(num * inv);
    }
    status[s] = 1;
}

// slots[s, d, c] = sum over the slot's corners (in order) of w * E[e, d, c]; one thread per
// (slot, placement lane), lanes of a slot adjacent.
__global__ void recover_slots_kernel(const float* __restrict__ E, const float* __restrict__ w,
                                     const std::int32_t* __restrict__ ptr, const std::int32_t* __restrict__ idx,
                                     std::int64_t n_slots, int k, int K, float* __restrict__ slots) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n_slots * K) return;
    const std::int64_t s = t / K;
    const int c = static_cast<int>(t % K);
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    if (c < k) {
        for (int p = ptr[s]; p < ptr[s + 1]; ++p) {
            const float wp = w[p];
            const float* src = E + static_cast<std::int64_t>(idx[p] >> 2) * 3 * K + c;
            a0 = fmaf(wp, src[0], a0);
            a1 = fmaf(wp, src[K], a1);
            a2 = fmaf(wp, src[2 * K], a2);
        }
    }
    float* dst = slots + s * 3 * K + c;
    dst[0] = a0;
    dst[K] = a1;
    dst[2 * K] = a2;
}

// E_out[e] = mean of the element's four corner slots (FP64 sum), |E_out|.
__global__ void recover_elements_kernel(const float* __restrict__ slots, const std::int32_t* __restrict__ soc,
                                        std::int64_t m, int k, int K, float* __restrict__ E_out,
                                        float* __restrict__ magn) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= m * K) return;
    const std::int64_t e = t / K;
    const int c = static_cast<int>(t % K);
    double v[3] = {0.0, 0.0, 0.0};
    if (c < k) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float* src = slots + static_cast<std::int64_t>(soc[4 * e + i]) * 3 * K + c;
            v[0] += static_cast<double>(src[0]);
            v[1] += static_cast<double>(src[K]);
            v[2] += static_cast<double>(src[2 * K]);
        }
        v[0] *= 0.25;
        v[1] *= 0.25;
        v[2] *= 0.25;
    }
    if (E_out) {
        float* dst = E_out + e * 3 * K + c;
        dst[0] = static_cast<float>(v[0]);
        dst[K] = static_cast<float>(v[1]);
        dst[2 * K] = static_cast<float>(v[2]);
    }
    magn[e * K + c] = static_cast<float>(sqrt(fma(v[2], v[2], fma(v[1], v[1], v[0] * v[0]))));
}

}  // namespace

void outer_boundary_nodes(const std::int32_t* tets, std::int64_t m, const std::int32_t* n2c_ptr,
                          const std::int32_t* n2c_idx, std::int64_t n, std::int32_t* mark, hipStream_t s) {
    check(hipMemsetAsync(mark, 0, sizeof(std::int32_t) * n, s), "boundary clear");
    if (const unsigned g = grid_for(4 * m)) outer_boundary_kernel<<<g, kBlock, 0, s>>>(tets, m, n2c_ptr, n2c_idx, mark);
}

// Corners grouped by (node, tag), stable: the node2corner order refined by tissue
// (fem/recovery.py:187-228). Fills ptr/idx/soc/slot_node/slot_tag and the views.
void tissue_slots(RecoveryOp& op, const std::int32_t* tets, const std::int32_t* tags, std::int64_t n, std::int64_t m,
                  int max_tag, hipStream_t s) {
    const std::int64_t nc = 4 * m;
    const std::uint64_t stride = static_cast<std::uint64_t>(max_tag) + 1;
    op.stride = stride;
    DeviceBuffer<std::uint64_t> keys(nc, "slot keys"), skeys(nc, "sorted slot keys");
    DeviceBuffer<std::int32_t> corner(nc, "corners"), flag(nc, "heads"), pos(nc, "head scan");
    op.idx.reset(nc, "slot corners");
    op.soc.reset(nc, "slot of corner");
    tissue_keys_kernel<<<grid_for(nc), kBlock, 0, s>>>(tets, tags, m, stride, keys.get());
    iota_i32(corner.get(), nc, s);
    int bits = 1;
    while (bits < 64 && ((static_cast<std::uint64_t>(n) * stride) >> bits) != 0) ++bits;
    sort_pairs_u64(keys.get(), skeys.get(), corner.get(), op.idx.get(), nc, bits, s);
    head_flags_kernel<<<grid_for(nc), kBlock, 0, s>>>(skeys.get(), nc, flag.get());
    exclusive_scan_i32(flag.get(), pos.get(), nc, s);
    std::int32_t last_pos = 0, last_flag = 0;
    d2h(&last_pos, pos.get() + (nc - 1), 1, s);
    d2h(&last_flag, flag.get() + (nc - 1), 1, s);
    sync(s, "slot count");
    op.n_slots = last_pos + last_flag;
    op.ptr.reset(op.n_slots + 1, "slot ptr");
    op.slot_node.reset(op.n_slots, "slot node");
    op.slot_tag.reset(op.n_slots, "slot tag");
    slots_fill_kernel<<<grid_for(nc), kBlock, 0, s>>>(skeys.get(), flag.get(), pos.get(), op.idx.get(), nc, stride,
                                                      op.ptr.get(), op.soc.get(), op.slot_node.get(),
                                                      op.slot_tag.get());
    op.ptr_view = op.ptr.get();
    op.idx_view = op.idx.get();
    op.soc_view = op.soc.get();
    sync(s, "tissue slots");  // the scratch above is freed on return
}

void build_spr(RecoveryOp& op, int mode, const double* nodes, const std::int32_t* tets, const std::int32_t* tags,
               const float* vols, const std::int32_t* n2c_ptr, const std::int32_t* n2c_idx, std::int64_t n,
               std::int64_t m, int max_tag, hipStream_t s) {
    require(mode == kRecoverySprGlobal || mode == kRecoverySprTissue, "not an SPR recovery mode");
    op.mode = mode;
    op.boundary.reset(static_cast<std::size_t>(std::max<std::int64_t>(n, 1)), "recovery boundary");
    outer_boundary_nodes(tets, m, n2c_ptr, n2c_idx, n, op.boundary.get(), s);
    const std::int64_t nc = 4 * m;
    if (mode == kRecoverySprGlobal) {
        op.n_slots = n;
        op.ptr_view = n2c_ptr;
        op.idx_view = n2c_idx;
        op.soc_view = tets;  // a corner reads back from its own node
        op.slot_node.reset(static_cast<std::size_t>(n), "slot node");
        iota_i32(op.slot_node.get(), n, s);
    } else {
        tissue_slots(op, tets, tags, n, m, max_tag, s);
        mixed_nodes_kernel<<<grid_for(op.n_slots), kBlock, 0, s>>>(op.slot_node.get(), op.n_slots, op.boundary.get());
    }
    op.w.reset(nc, "recovery weights");
    op.status.reset(static_cast<std::size_t>(op.n_slots), "recovery status");
    spr_weights_kernel<<<grid_for(op.n_slots, 128), 128, 0, s>>>(
        nodes, tets, vols, op.ptr_view, op.idx_view, mode == kRecoverySprGlobal ? nullptr : op.slot_node.get(),
        op.boundary.get(), op.n_slots, op.w.get(), op.status.get());
    std::vector<std::int32_t> st(static_cast<std::size_t>(op.n_slots));
    d2h(st.data(), op.status.get(), op.n_slots, s);
    sync(s, "recovery weights");
    op.n_fallback = 0;
    for (auto v : st) op.n_fallback += v;
}

void apply_spr(const RecoveryOp& op, const float* E, std::int64_t m, int k, int K, float* slots, float* E_out,
               float* magn, hipStream_t s) {
    if (const unsigned g = grid_for(op.n_slots * K))
        recover_slots_kernel<<<g, kBlock, 0, s>>>(E, op.w.get(), op.ptr_view, op.idx_view, op.n_slots, k, K, slots);
    if (const unsigned g = grid_for(m * K))
        recover_elements_kernel<<<g, kBlock, 0, s>>>(slots, op.soc_view, m, k, K, E_out, magn);
}

// ---------------------------------------------------------------------------------------------
// Harmonic recovery (fem/recovery.py:23-50,269-383; solver/fem/recovery.cu:290-610 of the
// reference). The potential is fitted over a same-tissue node patch around each (node, tissue)
// slot in the space of harmonic polynomials -- degree <= 3 (16 terms), else degree <= 2 (9), else
// linear (4) -- in the node-centred frame scaled by the patch half-width h; the gradient at the
// node is the linear coefficients / h, so each patch node m gets three weights and
//     E[s] = -sum_m w[s, m] v[m] - dA/dt(x_node).
// A rung is rejected when its normal matrix has a pivot <= 1e-12 A_00, or when its gradient can
// amplify the potential by more than 50 (h max_c sum_m |w[m, c]|); status 0/1/2 is the rung taken,
// 3 no gradient at all (zero weights, E = -dA/dt there).
// ---------------------------------------------------------------------------------------------
namespace {

constexpr double kMaxGradientAmplification = 50.0;
constexpr int kMinPatchNodes = 16;  // below this a slot's patch grows to its neighbours' rings
constexpr int kFitWaves = 4;
constexpr int kFitChunk = 64;
constexpr int kMaxTerms = 16;

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Harmonic polynomials up to degree 3 in (u, v, t): 1, the linear terms, the five quadratic and
// the seven cubic solid harmonics (each satisfies Laplace's equation).
template <int N>
__device__ __forceinline__ void harmonic_basis(double u, double v, double t, double* p) {
    p[0] = 1.0;
    p[1] = u;
    p[2] = v;
    p[3] = t;
    if constexpr (N >= 9) {
        p[4] = u * u - v * v;
        p[5] = v * v - t * t;
        p[6] = u * v;
        p[7] = u * t;
        p[8] = v * t;
    }
    if constexpr (N >= 16) {
        const double uu = u * u, vv = v * v, tt = t * t;
        p[9] = t * (2.0 * tt - 3.0 * uu - 3.0 * vv);
        p[10] = u * (4.0 * tt - uu - vv);
        p[11] = v * (4.0 * tt - uu - vv);
        p[12] = t * (uu - vv);
        p[13] = u * v * t;
        p[14] = u * (uu - 3.0 * vv);
        p[15] = v * (3.0 * uu - vv);
    }
}

template <int N>
constexpr int min_patch() {
    return N == 16 ? 20 : (N == 9 ? 9 : 4);
}

// One rung of the fit for one slot, one wavefront. Normal-matrix entries (lower triangle) are
// dealt round the lanes and summed over the patch in order from an LDS-staged chunk of basis
// values; the lanes factor the matrix together and three of them solve for rows 1..3 of the
// inverse; the lanes then write the weights of their patch entries and the amplification sums are
// folded by a fixed xor tree.
template <int N>
__device__ bool fit_rung(const double* __restrict__ nodes, const std::int32_t* __restrict__ pidx, int begin, int end,
                         const double (&xn)[3], double h, float* __restrict__ w3, double* q, double* a, double* z,
                         int lane) {
    if (end - begin < min_patch<N>()) return false;
    constexpr int kEntries = N * (N + 1) / 2;
    constexpr int kPer = (kEntries + kWave - 1) / kWave;
    int ri[kPer], ci[kPer];
    double acc[kPer];
#pragma unroll
    for (int e = 0; e < kPer; ++e) {
        const int k = lane + kWave * e;
        int i = 0;
        while ((i + 1) * (i + 2) / 2 <= k) ++i;
        ri[e] = i;
        ci[e] = k - i * (i + 1) / 2;
        acc[e] = 0.0;
    }
    for (int base = begin; base < end; base += kFitChunk) {
        const int cnt = min(kFitChunk, end - base);
        if (lane < cnt) {
            const std::int64_t mm = pidx[base + lane];
            harmonic_basis<N>((nodes[3 * mm] * 1e-3 - xn[0]) / h, (nodes[3 * mm + 1] * 1e-3 - xn[1]) / h,
                              (nodes[3 * mm + 2] * 1e-3 - xn[2]) / h, q + lane * N);
        }
        wave_sync();
        for (int p = 0; p < cnt; ++p) {
#pragma unroll
            for (int e = 0; e < kPer; ++e)
                if (lane + kWave * e < kEntries) acc[e] = fma(q[p * N + ri[e]], q[p * N + ci[e]], acc[e]);
        }
        wave_sync();
    }
#pragma unroll
    for (int e = 0; e < kPer; ++e)
        if (lane + kWave * e < kEntries) a[ri[e] * N + ci[e]] = acc[e];
    wave_sync();
    // Cholesky factor L (lower triangle of a), right-looking across the lanes: step j takes the
    // pivot, scales column j, then folds column j into the trailing triangle. Entry (i, k) thus
    // receives fma(-l_ij, l_kj, .) for j = 0, 1, ... in order and is divided by its pivot last,
    // the operations of the serial left-looking factorization in the same order. A pivot not above
    // 1e-12 A_00 rejects the rung.
    const double floor_pivot = 1e-12 * a[0];
    for (int j = 0; j < N; ++j) {
        const double d = a[j * N + j];
        if (!(d > floor_pivot)) return false;  // wave-uniform: every lane read the same pivot
        const double ljj = sqrt(d);
        wave_sync();
        if (lane == 0) a[j * N + j] = ljj;
        if (lane > j && lane < N) a[lane * N + j] = a[lane * N + j] / ljj;
        wave_sync();
        const int t = N - 1 - j;  // trailing rows j+1 .. N-1: t (t + 1) / 2 entries
        for (int q = lane; q < t * (t + 1) / 2; q += kWave) {
            int r = 0;
            while ((r + 1) * (r + 2) / 2 <= q) ++r;
            const int i = j + 1 + r, k = j + 1 + (q - r * (r + 1) / 2);
            a[i * N + k] = fma(-a[i * N + j], a[k * N + j], a[i * N + k]);
        }
        wave_sync();
    }
    // Rows 1..3 of the inverse (z_c = A^-1 e_{c+1}): L y = e_{c+1}, then L^T z = y in place in row c
    // of z. The forward solve runs by columns across the lanes (lane 16 c + r holds row r of
    // right-hand side c and folds in y_i, i ascending, as each is found: the multiply-adds of the
    // row-by-row form in the same order); the backward solve runs on lane c.
    {
        static_assert(3 * 16 <= kWave, "three right-hand sides of up to 16 rows fit a wavefront");
        const int c = lane / 16, r = lane % 16;
        const bool mine = c < 3 && r < N;
        double v = (mine && r == c + 1) ? 1.0 : 0.0;
        for (int i = 0; i < N; ++i) {
            if (mine && r == i) z[c * N + i] = v / a[i * N + i];
            wave_sync();
            if (mine && r > i) v = fma(-a[r * N + i], z[c * N + i], v);
        }
        wave_sync();
    }
    if (lane < 3) {
        const int c = lane;
        double* zc = z + c * N;
#pragma unroll 1
        for (int i = N - 1; i >= 0; --i) {
            double v = zc[i];
            for (int k = i + 1; k < N; ++k) v = fma(-a[k * N + i], zc[k], v);
            zc[i] = v / a[i * N + i];
        }
    }
    wave_sync();
    double sum[3] = {0.0, 0.0, 0.0};
    for (int p = begin + lane; p < end; p += kWave) {
        const std::int64_t mm = pidx[p];
        double qq[N];
        harmonic_basis<N>((nodes[3 * mm] * 1e-3 - xn[0]) / h, (nodes[3 * mm + 1] * 1e-3 - xn[1]) / h,
                          (nodes[3 * mm + 2] * 1e-3 - xn[2]) / h, qq);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double g = 0.0;
#pragma unroll
            for (int i = 0; i < N; ++i) g = fma(z[c * N + i], qq[i], g);
            sum[c] += fabs(g);
            w3[3 * static_cast<std::int64_t>(p) + c] = static_cast<float>(g / h);
        }
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1)
        for (int c = 0; c < 3; ++c) sum[c] += __shfl_xor(sum[c], off, kWave);
    return fmax(sum[0], fmax(sum[1], sum[2])) <= kMaxGradientAmplification;
}

// Three waves per SIMD: the LDS of three blocks fills a CU, and the register budget that leaves
// holds the kernel without spills.
__global__ void __launch_bounds__(kFitWaves * kWave) __attribute__((amdgpu_waves_per_eu(3))) hpr_fit_kernel(
    const double* __restrict__ nodes, const std::int32_t* __restrict__ pptr, const std::int32_t* __restrict__ pidx,
    const std::int32_t* __restrict__ slot_node, std::int64_t n_slots, float* __restrict__ w3,
    std::int32_t* __restrict__ status) {
    __shared__ double q_s[kFitWaves][kFitChunk * kMaxTerms];
    __shared__ double a_s[kFitWaves][kMaxTerms * kMaxTerms];
    __shared__ double z_s[kFitWaves][3 * kMaxTerms];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & (kWave - 1);
    const std::int64_t s = static_cast<std::int64_t>(blockIdx.x) * kFitWaves + wave;
    if (s >= n_slots) return;  // wave-uniform: no block barrier follows
    const int begin = pptr[s], end = pptr[s + 1];
    const std::int64_t node = slot_node[s];
    // In metres: the weights differentiate, so E comes out in V/m.
    const double xn[3] = {nodes[3 * node] * 1e-3, nodes[3 * node + 1] * 1e-3, nodes[3 * node + 2] * 1e-3};
    double h = 0.0;
    for (int p = begin + lane; p < end; p += kWave) {
        const std::int64_t mm = pidx[p];
        for (int c = 0; c < 3; ++c) h = fmax(h, fabs(nodes[3 * mm + c] * 1e-3 - xn[c]));
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) h = fmax(h, __shfl_xor(h, off, kWave));
    int rung = 3;
    if (h > 0.0) {
        double* q = q_s[wave];
        double* a = a_s[wave];
        double* z = z_s[wave];
        if (fit_rung<16>(nodes, pidx, begin, end, xn, h, w3, q, a, z, lane))
            rung = 0;
        else if (fit_rung<9>(nodes, pidx, begin, end, xn, h, w3, q, a, z, lane))
            rung = 1;
        else if (fit_rung<4>(nodes, pidx, begin, end, xn, h, w3, q, a, z, lane))
            rung = 2;
    }
    if (rung == 3) {
        for (int p = begin + lane; p < end; p += kWave)
            for (int c = 0; c < 3; ++c) w3[3 * static_cast<std::int64_t>(p) + c] = 0.0f;
    }
    if (lane == 0) status[s] = rung;
}

// E_slots[s, d, c] = -sum_p w3[p, d] v[pidx[p], c] - dA/dt[node, d, c], FP64 accumulation in patch
// order; v is the reduced potential with the ground node's zero restored.
constexpr int kGradBatch = 8;
constexpr int kGradStage = 2048;  // patch entries a block stages in LDS (32 KB with their weights)

// a0..a2 += sum over the patch entries [b, e) of w3[p, d] v[pidx[p], c], in order. Entries go in
// batches: the batch's indices and weights, then its potentials, are all loaded before the first
// use (an entry past the patch's end reads the last entry again and is not added), so a patch
// costs a few memory round trips instead of one per entry. `idx` and `w` are either the global
// arrays (base 0) or the block's LDS copy of its entries from `base` on.
__device__ __forceinline__ void hpr_accumulate(const double* __restrict__ X, int ground, int K, int c, int b, int e,
                                               int base, const std::int32_t* idx, const float* w, double& a0,
                                               double& a1, double& a2) {
    for (int p = b; p < e; p += kGradBatch) {
        int mm[kGradBatch];
        float wv[kGradBatch][3];
#pragma unroll
        for (int u = 0; u < kGradBatch; ++u) {
            const std::int64_t pp = min(p + u, e - 1) - base;
            mm[u] = idx[pp];
            wv[u][0] = w[3 * pp];
            wv[u][1] = w[3 * pp + 1];
            wv[u][2] = w[3 * pp + 2];
        }
        double vm[kGradBatch];
#pragma unroll
        for (int u = 0; u < kGradBatch; ++u) {
            const int r = mm[u] - (mm[u] > ground);
            const double v = X[static_cast<std::int64_t>(mm[u] == ground ? 0 : r) * K + c];
            vm[u] = (mm[u] == ground) ? 0.0 : v;
        }
#pragma unroll
        for (int u = 0; u < kGradBatch; ++u) {
            if (p + u < e) {
                a0 = fma(static_cast<double>(wv[u][0]), vm[u], a0);
                a1 = fma(static_cast<double>(wv[u][1]), vm[u], a1);
                a2 = fma(static_cast<double>(wv[u][2]), vm[u], a2);
            }
        }
    }
}

// One thread per (slot, column); a block takes kBlock / K consecutive slots, whose patches are one
// contiguous run of entries. The block copies that run's indices and weights into LDS once
// (coalesced) when it fits, instead of each of a slot's K threads loading every entry.
__global__ void __launch_bounds__(kBlock) hpr_grad_kernel(
    const double* __restrict__ X, int ground, const float* __restrict__ w3, const std::int32_t* __restrict__ pptr,
    const std::int32_t* __restrict__ pidx, const std::int32_t* __restrict__ slot_node, const float* __restrict__ dadt,
    int Kn, std::int64_t n_slots, int K, float* __restrict__ slots) {
    __shared__ std::int32_t sidx[kGradStage];
    __shared__ float sw[3 * kGradStage];
    const int per = kBlock / K;  // K divides kBlock
    const std::int64_t s0 = static_cast<std::int64_t>(blockIdx.x) * per;
    const std::int64_t s1 = min(s0 + per, n_slots);
    const int pb = pptr[s0], pe = pptr[s1];
    const bool staged = pe - pb <= kGradStage;  // block-uniform
    if (staged) {
        for (int i = threadIdx.x; i < pe - pb; i += kBlock) sidx[i] = pidx[pb + i];
        for (int i = threadIdx.x; i < 3 * (pe - pb); i += kBlock) sw[i] = w3[3 * static_cast<std::int64_t>(pb) + i];
    }
    __syncthreads();
    const std::int64_t s = s0 + threadIdx.x / K;
    const int c = static_cast<int>(threadIdx.x % K);
    if (s >= n_slots) return;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    const int b = pptr[s], e = pptr[s + 1];
    if (staged)
        hpr_accumulate(X, ground, K, c, b, e, pb, sidx, sw, a0, a1, a2);
    else
        hpr_accumulate(X, ground, K, c, b, e, 0, pidx, w3, a0, a1, a2);
    const std::int64_t node = slot_node[s];
    const float* da = dadt + node * 3 * Kn + (Kn == 1 ? 0 : c);
    float* dst = slots + s * 3 * K + c;
    dst[0] = static_cast<float>(-a0 - static_cast<double>(da[0]));
    dst[K] = static_cast<float>(-a1 - static_cast<double>(da[Kn]));
    dst[2 * K] = static_cast<float>(-a2 - static_cast<double>(da[2 * Kn]));
}

// Slots per node (slots are ordered by node then tag, so node v's slots are the range
// [node_slots[v], node_slots[v + 1]) of the exclusive scan of these counts).
__global__ void slots_per_node_kernel(const std::int32_t* __restrict__ slot_node, std::int64_t n_slots,
                                      std::int32_t* __restrict__ counts) {
    const std::int64_t s = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (s < n_slots) atomicAdd(&counts[slot_node[s]], 1);
}

// Slot of (node, tag); -1 if there is none.
__device__ int find_slot(const std::int32_t* __restrict__ node_slots, const std::int32_t* __restrict__ slot_tag,
                         int node, int tag) {
    for (int s = node_slots[node]; s < node_slots[node + 1]; ++s)
        if (slot_tag[s] == tag) return s;
    return -1;
}

static_assert(kMinPatchNodes <= kWave, "a short first ring must fit the lanes of one wavefront");

// Patch of each slot (fem/recovery.py:231-293): its first ring (the sorted distinct nodes of its
// own tetrahedra) when that has at least kMinPatchNodes nodes, else the union of the first rings
// of the same-tissue slots of its ring's nodes (the slot is its own neighbour), sorted and
// distinct. One wavefront per slot. The union is a merge of sorted rings: lane j keeps a cursor
// into the ring of the slot's j-th ring node, and each step takes the wavefront minimum of the
// lanes' heads and advances the lanes that held it, one distinct node per step. The rings are first
// copied into the wavefront's LDS stage (each lane its own ring, at an offset from a prefix sum of
// the ring sizes, all loads independent), so the merge steps wait on LDS, not on global memory;
// rings that do not fit the stage are merged from global memory.
constexpr int kPatchStage = 1024;  // ints of staged rings per wavefront

template <bool kFill>
__global__ void __launch_bounds__(kBlock) patch_kernel(
    const std::int32_t* __restrict__ rp, const std::int32_t* __restrict__ ri,
    const std::int32_t* __restrict__ node_slots, const std::int32_t* __restrict__ slot_tag, std::int64_t n_slots,
    std::int32_t* __restrict__ counts_or_ptr, std::int32_t* __restrict__ pidx, int* __restrict__ missing) {
    const int lane = threadIdx.x & (kWave - 1);
    const std::int64_t slot = static_cast<std::int64_t>(blockIdx.x) * (kBlock / kWave) +
                              __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    if (slot >= n_slots) return;  // wave-uniform
    const int b = rp[slot], len = rp[slot + 1] - b;
    const int at = kFill ? counts_or_ptr[slot] : 0;
    if (len >= kMinPatchNodes) {
        if (kFill)
            for (int p = lane; p < len; p += kWave) pidx[at + p] = ri[b + p];
        else if (lane == 0)
            counts_or_ptr[slot] = len;
        return;
    }
    __shared__ int stage[kBlock / kWave][kPatchStage];
    int* const st = stage[threadIdx.x / kWave];
    int cur = 0, end = 0;
    if (lane < len) {
        const int nb = find_slot(node_slots, slot_tag, ri[b + lane], slot_tag[slot]);
        if (nb < 0) {
            if (!kFill) atomicOr(missing, 1);
        } else {
            cur = rp[nb];
            end = rp[nb + 1];
        }
    }
    // Offsets of the rings in the stage: an exclusive prefix sum of their sizes over the lanes.
    const int size = end - cur;
    int incl = size;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const int t = __shfl_up(incl, d, kWave);
        if (lane >= d) incl += t;
    }
    const int total = __shfl(incl, kWave - 1, kWave);
    const bool staged = total <= kPatchStage;  // wave-uniform
    const int* src = ri;
    if (staged) {
        const int off = incl - size;
        for (int q = 0; q < size; ++q) st[off + q] = ri[cur + q];
        wave_sync();
        src = st;
        end = off + size;
        cur = off;
    }
    int head = cur < end ? src[cur] : INT_MAX;
    int k = 0, held = 0;
    while (true) {
        const int v = wave_min_i32(head);
        if (v == INT_MAX) break;
        if (kFill) {
            if (lane == (k & (kWave - 1))) held = v;
            if ((k & (kWave - 1)) == kWave - 1) pidx[at + (k & ~(kWave - 1)) + lane] = held;
        }
        ++k;
        if (head == v) {
            ++cur;
            head = cur < end ? src[cur] : INT_MAX;
        }
    }
    if (kFill) {
        if (lane < (k & (kWave - 1))) pidx[at + (k & ~(kWave - 1)) + lane] = held;
    } else if (lane == 0) {
        counts_or_ptr[slot] = k;
    }
}

}  // namespace

void build_harmonic(RecoveryOp& op, const double* nodes, const std::int32_t* tets, const std::int32_t* tags,
                    std::int64_t n, std::int64_t m, int max_tag, hipStream_t s) {
    op.mode = kRecoveryHarmonic;
    tissue_slots(op, tets, tags, n, m, max_tag, s);
    const std::int64_t ns = op.n_slots;
    // First ring of each slot: the distinct nodes of its own tetrahedra, sorted.
    DeviceBuffer<std::int32_t> r1_ptr, r1_idx;
    stiffness_pattern(tets, op.ptr.get(), op.idx.get(), ns, r1_ptr, r1_idx, s);
    DeviceBuffer<std::int32_t> per_node(static_cast<std::size_t>(n + 1), "slots per node"),
        node_slots(static_cast<std::size_t>(n + 1), "node slot ranges");
    check(hipMemsetAsync(per_node.get(), 0, sizeof(std::int32_t) * (n + 1), s), "memset");
    if (const unsigned g = grid_for(ns)) slots_per_node_kernel<<<g, kBlock, 0, s>>>(op.slot_node.get(), ns, per_node.get());
    exclusive_scan_i32(per_node.get(), node_slots.get(), n + 1, s);
    DeviceBuffer<std::int32_t> counts(static_cast<std::size_t>(ns + 1), "patch sizes");
    DeviceBuffer<int> missing(1, "missing neighbour slot");
    check(hipMemsetAsync(counts.get() + ns, 0, sizeof(std::int32_t), s), "memset");
    check(hipMemsetAsync(missing.get(), 0, sizeof(int), s), "memset");
    const unsigned g_slots = grid_for(ns, kBlock / kWave);
    if (g_slots)
        patch_kernel<false><<<g_slots, kBlock, 0, s>>>(r1_ptr.get(), r1_idx.get(), node_slots.get(),
                                                       op.slot_tag.get(), ns, counts.get(), nullptr, missing.get());
    op.pptr.reset(static_cast<std::size_t>(ns + 1), "patch ptr");
    exclusive_scan_i32(counts.get(), op.pptr.get(), ns + 1, s);
    std::int32_t nnz = 0;
    int hmiss = 0;
    d2h(&nnz, op.pptr.get() + ns, 1, s);
    d2h(&hmiss, missing.get(), 1, s);
    sync(s, "patch sizes");
    if (hmiss) throw std::runtime_error("harmonic patch: missing neighbour slot");
    op.nnz_patch = nnz;
    op.pidx.reset(static_cast<std::size_t>(std::max(nnz, 1)), "patch nodes");
    if (g_slots)
        patch_kernel<true><<<g_slots, kBlock, 0, s>>>(r1_ptr.get(), r1_idx.get(), node_slots.get(),
                                                      op.slot_tag.get(), ns, op.pptr.get(), op.pidx.get(), nullptr);
    op.w3.reset(static_cast<std::size_t>(3 * std::max(nnz, 1)), "harmonic weights");
    op.status.reset(static_cast<std::size_t>(std::max<std::int64_t>(ns, 1)), "harmonic status");
    if (const unsigned g = grid_for(ns, kFitWaves))
        hpr_fit_kernel<<<g, kFitWaves * kWave, 0, s>>>(nodes, op.pptr.get(), op.pidx.get(), op.slot_node.get(), ns,
                                                       op.w3.get(), op.status.get());
    std::vector<std::int32_t> status(static_cast<std::size_t>(ns));
    d2h(status.data(), op.status.get(), ns, s);
    sync(s, "harmonic weights");
    op.n_fallback = 0;
    op.n_undetermined = 0;
    for (auto v : status) {
        op.n_fallback += v >= 2 ? 1 : 0;
        op.n_undetermined += v == 3 ? 1 : 0;
    }
}

void apply_harmonic(const RecoveryOp& op, const double* X, int ground, const float* dadt_nodes, int Kn, std::int64_t m,
                    int k, int K, float* slots, float* E_out, float* magn, hipStream_t s) {
    require(Kn == 1 || Kn == K, "dA/dt column stride must be 1 or the block width");
    if (const unsigned g = grid_for(op.n_slots * K))
        hpr_grad_kernel<<<g, kBlock, 0, s>>>(X, ground, op.w3.get(), op.pptr.get(), op.pidx.get(), op.slot_node.get(),
                                             dadt_nodes, Kn, op.n_slots, K, slots);
    if (const unsigned g = grid_for(m * K))
        recover_elements_kernel<<<g, kBlock, 0, s>>>(slots, op.soc_view, m, k, K, E_out, magn);
}

}  // namespace cn
