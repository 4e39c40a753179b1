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
        const std::uint64_t stride = static_cast<std::uint64_t>(max_tag) + 1;
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
        mixed_nodes_kernel<<<grid_for(op.n_slots), kBlock, 0, s>>>(op.slot_node.get(), op.n_slots, op.boundary.get());
        op.ptr_view = op.ptr.get();
        op.idx_view = op.idx.get();
        op.soc_view = op.soc.get();
        sync(s, "tissue slots");  // the scratch above is freed on return
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

}  // namespace cn
