// Mesh layout and P1 assembly on the device: Morton order, P1 gradients, node-to-corner CSR,
// stiffness sparsity pattern and FP64 values, grounding, skin normals.
//
// Every kernel whose result must be bit-identical to the CPU oracle spells its fused
// multiply-adds out with fma(); the library is compiled with -ffp-contract=off so the compiler
// adds none of its own.
#include "kernels.hpp"
#include "prims.hpp"

namespace cn {

namespace {

__device__ __forceinline__ std::uint64_t spread3(std::uint64_t v) {
    v &= 0x1FFFFFull;
    v = (v | (v << 32)) & 0x1F00000000FFFFull;
    v = (v | (v << 16)) & 0x1F0000FF0000FFull;
    v = (v | (v << 8)) & 0x100F00F00F00F00Full;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
    return (v | (v << 2)) & 0x1249249249249249ull;
}

// Per-block min/max of each axis; the host folds the block results (order-free: min/max).
__global__ void bbox_kernel(const double* __restrict__ pts, std::int64_t n, double* __restrict__ out) {
    __shared__ double s[6][kBlock];
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<std::int64_t>(gridDim.x) * blockDim.x) {
        for (int a = 0; a < 3; ++a) {
            const double v = pts[3 * i + a];
            lo[a] = fmin(lo[a], v);
            hi[a] = fmax(hi[a], v);
        }
    }
    for (int a = 0; a < 3; ++a) {
        s[a][threadIdx.x] = lo[a];
        s[3 + a][threadIdx.x] = hi[a];
    }
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            for (int a = 0; a < 3; ++a) {
                s[a][threadIdx.x] = fmin(s[a][threadIdx.x], s[a][threadIdx.x + w]);
                s[3 + a][threadIdx.x] = fmax(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + w]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) out[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

// fem/solve.py:315-323: q = (p - lo) / span * (2^21 - 1) truncated, keys interleave x, y, z.
__global__ void morton_kernel(const double* __restrict__ pts, std::int64_t n, double lx, double ly,
                              double lz, double sx, double sy, double sz,
                              std::uint64_t* __restrict__ keys) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double scale = 2097151.0;
    const double lo[3] = {lx, ly, lz}, sp[3] = {sx, sy, sz};
    std::uint64_t key = 0;
    for (int a = 0; a < 3; ++a) {
        const double q = (pts[3 * i + a] - lo[a]) / sp[a] * scale;
        key |= spread3(static_cast<std::uint64_t>(q)) << a;
    }
    keys[i] = key;
}

__global__ void invert_perm_kernel(const std::int32_t* __restrict__ perm, std::int64_t n,
                                   std::int32_t* __restrict__ inverse) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) inverse[perm[i]] = static_cast<std::int32_t>(i);
}

// reorder.cu:14-24: the smallest renumbered node of each tetrahedron, the tet sort key.
__global__ void tet_lowest_kernel(const std::int32_t* __restrict__ inverse,
                                  const std::int32_t* __restrict__ tets, std::int64_t m,
                                  std::uint32_t* __restrict__ lowest) {
    const std::int64_t e = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= m) return;
    const int4 t = reinterpret_cast<const int4*>(tets)[e];
    const int a = min(inverse[t.x], inverse[t.y]);
    const int b = min(inverse[t.z], inverse[t.w]);
    lowest[e] = static_cast<std::uint32_t>(min(a, b));
}

__global__ void gather_nodes_kernel(const double* __restrict__ src, const std::int32_t* __restrict__ perm,
                                    std::int64_t n, double* __restrict__ dst) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const std::int64_t j = perm[i];
    dst[3 * i] = src[3 * j];
    dst[3 * i + 1] = src[3 * j + 1];
    dst[3 * i + 2] = src[3 * j + 2];
}

__global__ void gather_tets_kernel(const std::int32_t* __restrict__ tets, const std::int32_t* __restrict__ tags,
                                   const std::int32_t* __restrict__ tet_perm,
                                   const std::int32_t* __restrict__ inverse, std::int64_t m,
                                   std::int32_t* __restrict__ tets_out, std::int32_t* __restrict__ tags_out) {
    const std::int64_t e = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= m) return;
    const std::int64_t src = tet_perm[e];
    const int4 t = reinterpret_cast<const int4*>(tets)[src];
    reinterpret_cast<int4*>(tets_out)[e] =
        make_int4(inverse[t.x], inverse[t.y], inverse[t.z], inverse[t.w]);
    tags_out[e] = tags[src];
}

__global__ void renumber_kernel(std::int32_t* __restrict__ ids, std::int64_t count,
                                const std::int32_t* __restrict__ inverse) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) ids[i] = inverse[ids[i]];
}

__device__ __forceinline__ void cross3(const double* a, const double* b, double* o) {
    o[0] = fma(a[1], b[2], -(a[2] * b[1]));
    o[1] = fma(a[2], b[0], -(a[0] * b[2]));
    o[2] = fma(a[0], b[1], -(a[1] * b[0]));
}

// gradient.cu:29-66. One thread per tetrahedron; FP64.
__global__ void p1_gradients_kernel(const double* __restrict__ nodes, const std::int32_t* __restrict__ tets,
                                    std::int64_t m, double* __restrict__ g, double* __restrict__ vols) {
    const std::int64_t e = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= m) return;
    const int4 t4 = reinterpret_cast<const int4*>(tets)[e];
    const int ids[4] = {t4.x, t4.y, t4.z, t4.w};
    const double* p0 = nodes + 3 * static_cast<std::int64_t>(ids[0]);
    double t[3][3];
    for (int i = 0; i < 3; ++i) {
        const double* pi = nodes + 3 * static_cast<std::int64_t>(ids[i + 1]);
        for (int c = 0; c < 3; ++c) t[i][c] = pi[c] - p0[c];
    }
    double adj[3][3];
    cross3(t[1], t[2], adj[0]);
    cross3(t[2], t[0], adj[1]);
    cross3(t[0], t[1], adj[2]);
    const double det = fma(t[0][2], adj[0][2], fma(t[0][1], adj[0][1], t[0][0] * adj[0][0]));
    const double inv = 1e3 / det;
    double* ge = g + 12 * e;
    for (int c = 0; c < 3; ++c) {
        const double g1 = adj[0][c] * inv, g2 = adj[1][c] * inv, g3 = adj[2][c] * inv;
        ge[0 + c] = -((g1 + g2) + g3);
        ge[3 + c] = g1;
        ge[6 + c] = g2;
        ge[9 + c] = g3;
    }
    vols[e] = fabs(det) * (1e-9 / 6.0);
}

// Counts tetrahedra whose volume is not positive and finite (fem/solve.py:702-706).
__global__ void count_bad_volume_kernel(const double* __restrict__ vols, std::int64_t m,
                                        int* __restrict__ count) {
    const std::int64_t e = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    int bad = 0;
    if (e < m) {
        const double v = vols[e];
        bad = !(v > 0.0) || !isfinite(v);
    }
    const unsigned long long mask = __ballot(bad);
    if ((threadIdx.x & (kWave - 1)) == 0 && mask) atomicAdd(count, __popcll(mask));
}

// fem/assembly.py:31-42: tag -> sigma by table, an unknown tag counted.
__global__ void conductivity_kernel(const std::int32_t* __restrict__ tags, std::int64_t m,
                                    const double* __restrict__ lut, int lut_len,
                                    double* __restrict__ cond, int* __restrict__ unknown) {
    const std::int64_t e = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= m) return;
    const int t = tags[e];
    double s = NAN;
    if (t >= 0 && t < lut_len) s = lut[t];
    if (isnan(s)) atomicAdd(unknown, 1);
    cond[e] = s;
}

__global__ void count_keys_kernel(const std::int32_t* __restrict__ keys, std::int64_t count,
                                  std::int32_t* __restrict__ hist) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) atomicAdd(&hist[keys[i]], 1);
}

// Row i of the stiffness pattern: the sorted distinct nodes of the tetrahedra in node i's
// node2corner segment (pattern.cu:65-87). One wavefront per row: lane l holds the four nodes of
// the row's tetrahedra l, l + 64, ... (the first in registers, any further ones re-read), and
// each step takes the wavefront minimum over the lanes' smallest node above the last one taken.
// The row comes out sorted and distinct in as many steps as it has entries, each tetrahedron is
// read once, and the entries are written 64 at a time by the lanes that took them.
__device__ __forceinline__ int min_above(int4 t, int last, int best) {
    if (t.x > last && t.x < best) best = t.x;
    if (t.y > last && t.y < best) best = t.y;
    if (t.z > last && t.z < best) best = t.z;
    if (t.w > last && t.w < best) best = t.w;
    return best;
}

template <bool kFill>
__global__ void __launch_bounds__(kBlock) pattern_wave_kernel(const std::int32_t* __restrict__ tets,
                                                              const std::int32_t* __restrict__ ptr,
                                                              const std::int32_t* __restrict__ idx, std::int64_t n,
                                                              std::int32_t* __restrict__ counts_or_indptr,
                                                              std::int32_t* __restrict__ indices) {
    const int lane = threadIdx.x & (kWave - 1);
    const std::int64_t row = static_cast<std::int64_t>(blockIdx.x) * (kBlock / kWave) +
                             __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    if (row >= n) return;  // wave-uniform
    const int4* t4 = reinterpret_cast<const int4*>(tets);
    const int b = ptr[row], e = ptr[row + 1];
    const int4 none = {INT_MAX, INT_MAX, INT_MAX, INT_MAX};
    const int4 own = b + lane < e ? t4[idx[b + lane] >> 2] : none;
    const bool wide = e - b > kWave;
    const int at = kFill ? counts_or_indptr[row] : 0;
    int k = 0, last = -1, held = 0;
    while (true) {
        int best = min_above(own, last, INT_MAX);
        if (wide)
            for (int p = b + kWave + lane; p < e; p += kWave) best = min_above(t4[idx[p] >> 2], last, best);
        best = wave_min_i32(best);
        if (best == INT_MAX) break;
        if (kFill) {
            if (lane == (k & (kWave - 1))) held = best;
            if ((k & (kWave - 1)) == kWave - 1) indices[at + (k & ~(kWave - 1)) + lane] = held;
        }
        ++k;
        last = best;
    }
    if (kFill) {
        if (lane < (k & (kWave - 1))) indices[at + (k & ~(kWave - 1)) + lane] = held;
    } else if (lane == 0) {
        counts_or_indptr[row] = k;
    }
}

// stiffness.cu:32-66: entry (i, j) sums vol*sigma (grad_i . grad_j) over the corners of node i in
// node2corner order, then local column ascending, in FP64. One wavefront per row: lane l owns the
// row's entries l, l + 64, ..., every lane walks the row's corners in that order (the same
// addresses across the wavefront) and folds each contribution into the entry it owns, so every
// entry sees the fused multiply-adds of a serial walk in the same order, and the row is written
// coalesced without read-modify-write.
__global__ void __launch_bounds__(kBlock) stiffness_wave_kernel(
    const double* __restrict__ g, const double* __restrict__ scale, const std::int32_t* __restrict__ tets,
    const std::int32_t* __restrict__ ptr, const std::int32_t* __restrict__ idx, std::int64_t n,
    const std::int32_t* __restrict__ indptr, const std::int32_t* __restrict__ indices, double* __restrict__ data) {
    const int lane = threadIdx.x & (kWave - 1);
    const std::int64_t row = static_cast<std::int64_t>(blockIdx.x) * (kBlock / kWave) +
                             __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    if (row >= n) return;  // wave-uniform
    const int rb = indptr[row], re = indptr[row + 1];
    const int cb = ptr[row], ce = ptr[row + 1];
    for (int base = rb; base < re; base += kWave) {
        const int mine = base + lane < re ? indices[base + lane] : -1;
        double acc = 0.0;
        for (int p = cb; p < ce; ++p) {
            const std::int64_t corner = idx[p];
            const std::int64_t e = corner >> 2;
            const double s = scale[e];
            const double gi0 = g[3 * corner], gi1 = g[3 * corner + 1], gi2 = g[3 * corner + 2];
            const int4 t = reinterpret_cast<const int4*>(tets)[e];
            const int cols[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const double* gj = g + 3 * (4 * e + j);
                const double dot = fma(gi2, gj[2], fma(gi1, gj[1], gi0 * gj[0]));
                if (cols[j] == mine) acc = fma(s, dot, acc);
            }
        }
        if (base + lane < re) data[base + lane] = acc;
    }
}

// Grounding (fem/solve.py:292-296,361-365): reduced row r is mesh row r + (r >= ground), minus
// the ground column; columns above the ground shift down by one.
__global__ void reduced_count_kernel(const std::int32_t* __restrict__ indptr,
                                     const std::int32_t* __restrict__ indices, std::int64_t n_red,
                                     int ground, std::int32_t* __restrict__ counts) {
    const std::int64_t r = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= n_red) return;
    const std::int64_t row = r + (r >= ground);
    int c = 0;
    for (int p = indptr[row]; p < indptr[row + 1]; ++p) c += indices[p] != ground;
    counts[r] = c;
}

__global__ void reduced_fill_kernel(const std::int32_t* __restrict__ indptr,
                                    const std::int32_t* __restrict__ indices,
                                    const double* __restrict__ data, std::int64_t n_red, int ground,
                                    const std::int32_t* __restrict__ rptr, std::int32_t* __restrict__ rcol,
                                    double* __restrict__ rval64, float* __restrict__ rval32) {
    const std::int64_t r = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= n_red) return;
    const std::int64_t row = r + (r >= ground);
    int at = rptr[r];
    for (int p = indptr[row]; p < indptr[row + 1]; ++p) {
        const int c = indices[p];
        if (c == ground) continue;
        rcol[at] = c - (c > ground);
        rval64[at] = data[p];
        rval32[at] = static_cast<float>(data[p]);
        ++at;
    }
}

__global__ void lowest_z_kernel(const double* __restrict__ nodes, std::int64_t n,
                                double* __restrict__ best_z, std::int32_t* __restrict__ best_i) {
    // Per-block argmin of z with the lowest index on ties; the host folds the blocks.
    __shared__ double sz[kBlock];
    __shared__ std::int32_t si[kBlock];
    double z = INFINITY;
    std::int32_t id = INT_MAX;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<std::int64_t>(gridDim.x) * blockDim.x) {
        const double v = nodes[3 * i + 2];
        if (v < z) {
            z = v;
            id = static_cast<std::int32_t>(i);
        }
    }
    sz[threadIdx.x] = z;
    si[threadIdx.x] = id;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            const double oz = sz[threadIdx.x + w];
            const std::int32_t oi = si[threadIdx.x + w];
            if (oz < sz[threadIdx.x] || (oz == sz[threadIdx.x] && oi < si[threadIdx.x])) {
                sz[threadIdx.x] = oz;
                si[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        best_z[blockIdx.x] = sz[0];
        best_i[blockIdx.x] = si[0];
    }
}

__global__ void resident_arrays_kernel(const double* __restrict__ g64, const double* __restrict__ vols64,
                                       const double* __restrict__ cond, const double* __restrict__ nodes,
                                       const std::int32_t* __restrict__ tets, std::int64_t m,
                                       float* __restrict__ g32, float* __restrict__ vols32,
                                       float* __restrict__ neg_vc, double* __restrict__ bary) {
    const std::int64_t e = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= m) return;
    for (int i = 0; i < 12; ++i) g32[12 * e + i] = static_cast<float>(g64[12 * e + i]);
    const float v = static_cast<float>(vols64[e]);
    vols32[e] = v;
    neg_vc[e] = -(v * static_cast<float>(cond[e]));  // fem/solve.py:714, FP32 product
    const int4 t = reinterpret_cast<const int4*>(tets)[e];
    for (int c = 0; c < 3; ++c) {
        // cupy mean over axis 1: ((a + b) + c + d) / 4
        const double s = ((nodes[3 * (std::int64_t)t.x + c] + nodes[3 * (std::int64_t)t.y + c]) +
                          nodes[3 * (std::int64_t)t.z + c]) + nodes[3 * (std::int64_t)t.w + c];
        bary[3 * e + c] = s / 4.0;
    }
}

// skin.cu:46-150. Face normal (unnormalised, twice the area) per triangle.
__global__ void face_normals_kernel(const double* __restrict__ nodes, const std::int32_t* __restrict__ tris,
                                    std::int64_t n_tri, double* __restrict__ out) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n_tri) return;
    const double* a = nodes + 3 * static_cast<std::int64_t>(tris[3 * t]);
    const double* b = nodes + 3 * static_cast<std::int64_t>(tris[3 * t + 1]);
    const double* c = nodes + 3 * static_cast<std::int64_t>(tris[3 * t + 2]);
    double u[3], v[3];
    for (int i = 0; i < 3; ++i) {
        u[i] = b[i] - a[i];
        v[i] = c[i] - a[i];
    }
    out[3 * t] = u[1] * v[2] - u[2] * v[1];
    out[3 * t + 1] = u[2] * v[0] - u[0] * v[2];
    out[3 * t + 2] = u[0] * v[1] - u[1] * v[0];
}

__global__ void spread_to_nodes_kernel(const double* __restrict__ face, const std::int32_t* __restrict__ ptr,
                                       const std::int32_t* __restrict__ idx, std::int64_t n,
                                       double* __restrict__ out) {
    const std::int64_t nd = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (nd >= n) return;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int p = ptr[nd]; p < ptr[nd + 1]; ++p) {
        const std::int64_t t = idx[p] / 3;
        a0 += face[3 * t];
        a1 += face[3 * t + 1];
        a2 += face[3 * t + 2];
    }
    out[3 * nd] = a0;
    out[3 * nd + 1] = a1;
    out[3 * nd + 2] = a2;
}

__device__ __forceinline__ void normalise3(double* v) {
    const double nn = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (nn > 0.0) {
        v[0] /= nn;
        v[1] /= nn;
        v[2] /= nn;
    }
}

__global__ void gather_to_faces_kernel(const double* __restrict__ node, const std::int32_t* __restrict__ tris,
                                       std::int64_t n_tri, double* __restrict__ out) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n_tri) return;
    double acc[3];
    for (int i = 0; i < 3; ++i)
        acc[i] = node[3 * (std::int64_t)tris[3 * t] + i] + node[3 * (std::int64_t)tris[3 * t + 1] + i] +
                 node[3 * (std::int64_t)tris[3 * t + 2] + i];
    normalise3(acc);
    for (int i = 0; i < 3; ++i) out[3 * t + i] = acc[i];
}

__global__ void normalise_rows_kernel(double* __restrict__ v, std::int64_t n) {
    const std::int64_t r = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r < n) normalise3(v + 3 * r);
}

__global__ void tri_vertices_kernel(const double* __restrict__ nodes, const std::int32_t* __restrict__ tris,
                                    std::int64_t n_tri, double* __restrict__ a, double* __restrict__ b,
                                    double* __restrict__ c) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n_tri) return;
    for (int i = 0; i < 3; ++i) {
        a[3 * t + i] = nodes[3 * (std::int64_t)tris[3 * t] + i];
        b[3 * t + i] = nodes[3 * (std::int64_t)tris[3 * t + 1] + i];
        c[3 * t + i] = nodes[3 * (std::int64_t)tris[3 * t + 2] + i];
    }
}

}  // namespace

// ----------------------------------------------------------------------------------------------
// Host launchers
// ----------------------------------------------------------------------------------------------

void morton_keys(const double* pts, std::int64_t n, std::uint64_t* keys, hipStream_t s) {
    if (n <= 0) return;
    const int blocks = static_cast<int>(std::min<std::int64_t>(grid_for(n), 1024));
    DeviceBuffer<double> part(6 * static_cast<std::size_t>(blocks), "bbox partials");
    bbox_kernel<<<blocks, kBlock, 0, s>>>(pts, n, part.get());
    std::vector<double> h(6 * static_cast<std::size_t>(blocks));
    d2h(h.data(), part.get(), h.size(), s);
    sync(s, "bbox");
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int b = 0; b < blocks; ++b)
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], h[6 * b + a]);
            hi[a] = std::max(hi[a], h[6 * b + 3 + a]);
        }
    double sp[3];
    for (int a = 0; a < 3; ++a) {
        sp[a] = hi[a] - lo[a];
        if (!(sp[a] > 1e-12)) sp[a] = 1e-12;
    }
    morton_kernel<<<grid_for(n), kBlock, 0, s>>>(pts, n, lo[0], lo[1], lo[2], sp[0], sp[1], sp[2], keys);
}

void tet_lowest_node(const std::int32_t* inverse, const std::int32_t* tets, std::int64_t m,
                     std::uint32_t* lowest, hipStream_t s) {
    if (const unsigned g = grid_for(m)) tet_lowest_kernel<<<g, kBlock, 0, s>>>(inverse, tets, m, lowest);
}

void spatial_order(const double* nodes, std::int64_t n, const std::int32_t* tets, std::int64_t m,
                   std::int32_t* node_perm, std::int32_t* tet_perm, std::int32_t* inverse, hipStream_t s) {
    DeviceBuffer<std::uint64_t> keys(std::max<std::int64_t>(n, 1), "morton keys");
    DeviceBuffer<std::uint64_t> keys_out(std::max<std::int64_t>(n, 1), "morton keys sorted");
    DeviceBuffer<std::int32_t> iota(std::max<std::int64_t>(std::max(n, m), 1), "iota");
    morton_keys(nodes, n, keys.get(), s);
    iota_i32(iota.get(), std::max(n, m), s);
    sort_pairs_u64(keys.get(), keys_out.get(), iota.get(), node_perm, n, 63, s);
    if (const unsigned g = grid_for(n)) invert_perm_kernel<<<g, kBlock, 0, s>>>(node_perm, n, inverse);
    DeviceBuffer<std::uint32_t> low(std::max<std::int64_t>(m, 1), "tet lowest");
    DeviceBuffer<std::uint32_t> low_out(std::max<std::int64_t>(m, 1), "tet lowest sorted");
    tet_lowest_node(inverse, tets, m, low.get(), s);
    int bits = 1;
    while (bits < 32 && (1ll << bits) < n) ++bits;
    sort_pairs_u32(low.get(), low_out.get(), iota.get(), tet_perm, m, bits, s);
}

void apply_spatial_order(const double* nodes_in, const std::int32_t* tets_in, const std::int32_t* tags_in,
                         std::int64_t n, std::int64_t m, const std::int32_t* node_perm,
                         const std::int32_t* tet_perm, const std::int32_t* inverse, double* nodes_out,
                         std::int32_t* tets_out, std::int32_t* tags_out, std::int32_t* skin,
                         std::int64_t n_skin, hipStream_t s) {
    if (const unsigned g = grid_for(n)) gather_nodes_kernel<<<g, kBlock, 0, s>>>(nodes_in, node_perm, n, nodes_out);
    if (const unsigned g = grid_for(m))
        gather_tets_kernel<<<g, kBlock, 0, s>>>(tets_in, tags_in, tet_perm, inverse, m, tets_out, tags_out);
    if (const unsigned g = grid_for(3 * n_skin)) renumber_kernel<<<g, kBlock, 0, s>>>(skin, 3 * n_skin, inverse);
}

void p1_gradients(const double* nodes, const std::int32_t* tets, std::int64_t m, double* g,
                  double* vols, hipStream_t s) {
    if (const unsigned gr = grid_for(m)) p1_gradients_kernel<<<gr, kBlock, 0, s>>>(nodes, tets, m, g, vols);
}

int count_bad_volumes(const double* vols, std::int64_t m, hipStream_t s) {
    DeviceBuffer<int> c(1, "bad volume count");
    check(hipMemsetAsync(c.get(), 0, sizeof(int), s), "memset");
    if (const unsigned g = grid_for(m)) count_bad_volume_kernel<<<g, kBlock, 0, s>>>(vols, m, c.get());
    int h = 0;
    d2h(&h, c.get(), 1, s);
    sync(s, "volume check");
    return h;
}

int conductivity(const std::int32_t* tags, std::int64_t m, const double* lut_host, int lut_len,
                 double* cond, hipStream_t s) {
    DeviceBuffer<double> lut = upload(lut_host, static_cast<std::size_t>(lut_len), s, "sigma table");
    DeviceBuffer<int> c(1, "unknown tag count");
    check(hipMemsetAsync(c.get(), 0, sizeof(int), s), "memset");
    if (const unsigned g = grid_for(m))
        conductivity_kernel<<<g, kBlock, 0, s>>>(tags, m, lut.get(), lut_len, cond, c.get());
    int h = 0;
    d2h(&h, c.get(), 1, s);
    sync(s, "conductivity");
    return h;
}

void node2corner(const std::int32_t* elems, std::int64_t n_elem, int corners, std::int64_t n_nodes,
                 std::int32_t* ptr, std::int32_t* idx, hipStream_t s) {
    const std::int64_t count = n_elem * corners;
    DeviceBuffer<std::int32_t> hist(static_cast<std::size_t>(n_nodes + 1), "node2corner counts");
    check(hipMemsetAsync(hist.get(), 0, sizeof(std::int32_t) * (n_nodes + 1), s), "memset");
    if (count > 0) {
        if (const unsigned g = grid_for(count)) count_keys_kernel<<<g, kBlock, 0, s>>>(elems, count, hist.get());
        DeviceBuffer<std::int32_t> iota(static_cast<std::size_t>(count), "corner ids");
        DeviceBuffer<std::uint32_t> ksorted(static_cast<std::size_t>(count), "corner keys");
        iota_i32(iota.get(), count, s);
        int bits = 1;
        while (bits < 32 && (1ll << bits) < n_nodes) ++bits;
        sort_pairs_u32(reinterpret_cast<const std::uint32_t*>(elems), ksorted.get(), iota.get(), idx,
                       count, bits, s);
    }
    exclusive_scan_i32(hist.get(), ptr, n_nodes + 1, s);
}

std::int64_t stiffness_pattern(const std::int32_t* tets, const std::int32_t* ptr, const std::int32_t* idx,
                               std::int64_t n, DeviceBuffer<std::int32_t>& indptr,
                               DeviceBuffer<std::int32_t>& indices, hipStream_t s) {
    DeviceBuffer<std::int32_t> counts(static_cast<std::size_t>(n + 1), "pattern counts");
    check(hipMemsetAsync(counts.get() + n, 0, sizeof(std::int32_t), s), "memset");
    const unsigned rows = grid_for(n, kBlock / kWave);
    if (rows) pattern_wave_kernel<false><<<rows, kBlock, 0, s>>>(tets, ptr, idx, n, counts.get(), nullptr);
    indptr.reset(static_cast<std::size_t>(n + 1), "indptr");
    exclusive_scan_i32(counts.get(), indptr.get(), n + 1, s);
    std::int32_t nnz = 0;
    d2h(&nnz, indptr.get() + n, 1, s);
    sync(s, "pattern nnz");
    indices.reset(static_cast<std::size_t>(std::max(nnz, 1)), "indices");
    if (rows) pattern_wave_kernel<true><<<rows, kBlock, 0, s>>>(tets, ptr, idx, n, indptr.get(), indices.get());
    return nnz;
}

void stiffness_values(const double* g, const double* scale, const std::int32_t* tets,
                      const std::int32_t* ptr, const std::int32_t* idx, std::int64_t n,
                      const std::int32_t* indptr, const std::int32_t* indices, double* data, hipStream_t s) {
    if (const unsigned gr = grid_for(n, kBlock / kWave))
        stiffness_wave_kernel<<<gr, kBlock, 0, s>>>(g, scale, tets, ptr, idx, n, indptr, indices, data);
}

int lowest_z_node(const double* nodes, std::int64_t n, hipStream_t s) {
    const int blocks = static_cast<int>(std::min<std::int64_t>(grid_for(n), 1024));
    DeviceBuffer<double> bz(static_cast<std::size_t>(blocks), "lowest z");
    DeviceBuffer<std::int32_t> bi(static_cast<std::size_t>(blocks), "lowest z index");
    lowest_z_kernel<<<blocks, kBlock, 0, s>>>(nodes, n, bz.get(), bi.get());
    std::vector<double> hz(blocks);
    std::vector<std::int32_t> hi(blocks);
    d2h(hz.data(), bz.get(), blocks, s);
    d2h(hi.data(), bi.get(), blocks, s);
    sync(s, "lowest z");
    double z = INFINITY;
    std::int32_t id = INT_MAX;
    for (int b = 0; b < blocks; ++b)
        if (hz[b] < z || (hz[b] == z && hi[b] < id)) {
            z = hz[b];
            id = hi[b];
        }
    return id;
}

std::int64_t reduce_grounded(const std::int32_t* indptr, const std::int32_t* indices, const double* data,
                             std::int64_t n, int ground, DeviceBuffer<std::int32_t>& rptr,
                             DeviceBuffer<std::int32_t>& rcol, DeviceBuffer<double>& rval64,
                             DeviceBuffer<float>& rval32, hipStream_t s) {
    const std::int64_t n_red = n - 1;
    DeviceBuffer<std::int32_t> counts(static_cast<std::size_t>(n_red + 1), "reduced counts");
    check(hipMemsetAsync(counts.get() + n_red, 0, sizeof(std::int32_t), s), "memset");
    if (const unsigned g = grid_for(n_red))
        reduced_count_kernel<<<g, kBlock, 0, s>>>(indptr, indices, n_red, ground, counts.get());
    rptr.reset(static_cast<std::size_t>(n_red + 1), "reduced ptr");
    exclusive_scan_i32(counts.get(), rptr.get(), n_red + 1, s);
    std::int32_t nnz = 0;
    d2h(&nnz, rptr.get() + n_red, 1, s);
    sync(s, "reduced nnz");
    rcol.reset(static_cast<std::size_t>(std::max(nnz, 1)), "reduced col");
    rval64.reset(static_cast<std::size_t>(std::max(nnz, 1)), "reduced val64");
    rval32.reset(static_cast<std::size_t>(std::max(nnz, 1)), "reduced val32");
    if (const unsigned g = grid_for(n_red))
        reduced_fill_kernel<<<g, kBlock, 0, s>>>(indptr, indices, data, n_red, ground, rptr.get(),
                                                 rcol.get(), rval64.get(), rval32.get());
    return nnz;
}

void resident_arrays(const double* g64, const double* vols64, const double* cond, const double* nodes,
                     const std::int32_t* tets, std::int64_t m, float* g32, float* vols32, float* neg_vc,
                     double* bary, hipStream_t s) {
    if (const unsigned g = grid_for(m))
        resident_arrays_kernel<<<g, kBlock, 0, s>>>(g64, vols64, cond, nodes, tets, m, g32, vols32, neg_vc, bary);
}

void skin_normals(const double* nodes, const std::int32_t* tris, std::int64_t n_tri, std::int64_t n_nodes,
                  double* out, hipStream_t s) {
    if (n_tri <= 0) return;
    DeviceBuffer<std::int32_t> ptr(static_cast<std::size_t>(n_nodes + 1), "skin ptr");
    DeviceBuffer<std::int32_t> idx(static_cast<std::size_t>(3 * n_tri), "skin idx");
    node2corner(tris, n_tri, 3, n_nodes, ptr.get(), idx.get(), s);
    DeviceBuffer<double> face(static_cast<std::size_t>(3 * n_tri), "face normals");
    DeviceBuffer<double> node(static_cast<std::size_t>(3 * n_nodes), "node normals");
    const unsigned tg = grid_for(n_tri), ng = grid_for(n_nodes);
    face_normals_kernel<<<tg, kBlock, 0, s>>>(nodes, tris, n_tri, face.get());
    spread_to_nodes_kernel<<<ng, kBlock, 0, s>>>(face.get(), ptr.get(), idx.get(), n_nodes, node.get());
    gather_to_faces_kernel<<<tg, kBlock, 0, s>>>(node.get(), tris, n_tri, face.get());
    spread_to_nodes_kernel<<<ng, kBlock, 0, s>>>(face.get(), ptr.get(), idx.get(), n_nodes, node.get());
    normalise_rows_kernel<<<ng, kBlock, 0, s>>>(node.get(), n_nodes);
    gather_to_faces_kernel<<<tg, kBlock, 0, s>>>(node.get(), tris, n_tri, out);
    sync(s, "skin normals");
}

void triangle_vertices(const double* nodes, const std::int32_t* tris, std::int64_t n_tri, double* a,
                       double* b, double* c, hipStream_t s) {
    if (const unsigned g = grid_for(n_tri)) tri_vertices_kernel<<<g, kBlock, 0, s>>>(nodes, tris, n_tri, a, b, c);
}

}  // namespace cn
