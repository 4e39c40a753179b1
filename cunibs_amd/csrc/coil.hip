// Coil side of a block of placements: scalp projection and coil frames, the dipole dA/dt N-body,
// the node-to-element average and the right-hand side.
//
// Layouts. Per-placement arrays are placement-major: dA/dt at nodes (k, n, 3), at elements
// (k, m, 3). The right-hand side goes straight into the solver's (n_red, K) row-major block with
// the grounded node dropped.
#include <cfloat>

#include "coil.hpp"

namespace cn {

namespace {

struct Candidate {
    double d2;
    int tri;
};

__device__ __forceinline__ bool closer(const Candidate& x, const Candidate& y) {
    return x.d2 < y.d2 || (x.d2 == y.d2 && x.tri < y.tri);
}

__device__ __forceinline__ double dot3(const double* a, const double* b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// Closest point on triangle (a, b, c) to p (Ericson), solver/coil/place.cu:66-119.
__device__ double closest_on_tri(const double* p, const double* a, const double* b, const double* c, double* q) {
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

// place.cu:122-150: one block per (placement, triangle chunk); the minimum over the total order
// (d^2, triangle index) within the chunk.
__global__ void __launch_bounds__(kBlock) place_scan_kernel(const double* __restrict__ centers,
                                                            const double* __restrict__ ta,
                                                            const double* __restrict__ tb,
                                                            const double* __restrict__ tc, int n_tri,
                                                            int n_chunks, Candidate* __restrict__ cand) {
    __shared__ double sd[kBlock];
    __shared__ int st[kBlock];
    const int p = blockIdx.x, chunk = blockIdx.y;
    const int per = (n_tri + n_chunks - 1) / n_chunks;
    const int lo = chunk * per, hi = min(lo + per, n_tri);
    const double ctr[3] = {centers[3 * p], centers[3 * p + 1], centers[3 * p + 2]};
    Candidate best{DBL_MAX, INT_MAX};
    double q[3];
    for (int j = lo + static_cast<int>(threadIdx.x); j < hi; j += kBlock) {
        const double d2 = closest_on_tri(ctr, ta + 3 * static_cast<std::int64_t>(j), tb + 3 * static_cast<std::int64_t>(j),
                                         tc + 3 * static_cast<std::int64_t>(j), q);
        if (d2 < best.d2) best = {d2, j};
    }
    sd[threadIdx.x] = best.d2;
    st[threadIdx.x] = best.tri;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            const Candidate o{sd[threadIdx.x + w], st[threadIdx.x + w]};
            const Candidate m{sd[threadIdx.x], st[threadIdx.x]};
            if (closer(o, m)) {
                sd[threadIdx.x] = o.d2;
                st[threadIdx.x] = o.tri;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) cand[static_cast<std::int64_t>(p) * n_chunks + chunk] = {sd[0], st[0]};
}

// place.cu:153-230: frame [x | y | z | c] per placement, z the inward normal, y the in-plane
// handle direction, c the projection lifted `distance` along the outward normal.
__global__ void place_frame_kernel(const double* __restrict__ centers, const double* __restrict__ handles,
                                   const double* __restrict__ dists, const double* __restrict__ ta,
                                   const double* __restrict__ tb, const double* __restrict__ tc,
                                   const double* __restrict__ tn, const Candidate* __restrict__ cand,
                                   int n_pl, int n_chunks, int n_tri, double* __restrict__ out,
                                   int* __restrict__ degenerate) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_pl) return;
    const Candidate* mine = cand + static_cast<std::int64_t>(p) * n_chunks;
    Candidate best = mine[0];
    for (int i = 1; i < n_chunks; ++i)
        if (closer(mine[i], best)) best = mine[i];
    // Every distance was NaN only if the inputs were; the host rejects non-finite centres, and a
    // frame built on triangle 0 then stays in bounds rather than reading past the surface.
    const std::int64_t tri = (best.tri >= 0 && best.tri < n_tri) ? best.tri : 0;
    const double ctr[3] = {centers[3 * p], centers[3 * p + 1], centers[3 * p + 2]};
    double proj[3];
    closest_on_tri(ctr, ta + 3 * tri, tb + 3 * tri, tc + 3 * tri, proj);
    const double nrm[3] = {tn[3 * tri], tn[3 * tri + 1], tn[3 * tri + 2]};
    const double z[3] = {-nrm[0], -nrm[1], -nrm[2]};
    double y[3];
    bool from_handle = false;
    if (handles != nullptr) {
        for (int i = 0; i < 3; ++i) y[i] = handles[3 * p + i] - proj[i];
        if (dot3(y, y) > 0.0) {
            const double inv = 1.0 / sqrt(dot3(y, y));
            for (int i = 0; i < 3; ++i) y[i] *= inv;
            const double cz = dot3(y, z);
            if ((1.0 - cz * cz) > 1e-12) {
                for (int i = 0; i < 3; ++i) y[i] -= z[i] * cz;
                const double inv2 = 1.0 / sqrt(dot3(y, y));
                for (int i = 0; i < 3; ++i) y[i] *= inv2;
                from_handle = true;
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
    if (degenerate != nullptr) degenerate[p] = from_handle ? 0 : 1;
    const double x[3] = {y[1] * z[2] - y[2] * z[1], y[2] * z[0] - y[0] * z[2], y[0] * z[1] - y[1] * z[0]};
    double* o = out + 16 * static_cast<std::int64_t>(p);
    for (int i = 0; i < 16; ++i) o[i] = 0.0;
    for (int r = 0; r < 3; ++r) {
        o[4 * r + 0] = x[r];
        o[4 * r + 1] = y[r];
        o[4 * r + 2] = z[r];
        o[4 * r + 3] = proj[r] + dists[p] * nrm[r];
    }
    o[15] = 1.0;
}

constexpr int kDadtBlock = 256;
constexpr int kDipTile = 256;
constexpr int kNodesPerThread = 8;

// dadt.cu:27-104 restated for CDNA: dipole tiles staged in LDS, 8 nodes per thread in registers,
// one placement per blockIdx.y. w_j = |r - s_j|^-3 from rn + sn - 2 r.s, FP32, on the vector ALUs.
// The reciprocal square root is the bare v_rsq_f32 (rsq_normal): rsqrtf wraps the same instruction
// in a denormal rescale (compare, two multiplies, two selects per pair), and a squared coil-to-node
// distance is never denormal, so the results are the same bits with a third fewer vector
// instructions per pair. (Matrix-core versions, with the distances as 16x16x4 FP32 MFMAs and the
// sums either as MFMAs or on the vector ALUs, measured 0.82-1.35 ms per placement at 1.0M nodes x
// 2400 dipoles against 0.78 ms for this kernel with rsqrtf, and were dropped.)
// Output (n, 3, K): the K placements of a node component are contiguous.

// v_rsq_f32 without the denormal rescale; for normal positive inputs it equals rsqrtf bit for bit.
__device__ __forceinline__ float rsq_normal(float x) { return __builtin_amdgcn_rsqf(x); }
__global__ void __launch_bounds__(kDadtBlock) dadt_kernel(const float* __restrict__ dip, int n_dip,
                                                          const double* __restrict__ nodes_mm, std::int64_t n,
                                                          const double* __restrict__ origin, float scale, int K,
                                                          float* __restrict__ out) {
    __shared__ float s_s[kDipTile * 3];
    __shared__ float s_mp[kDipTile * 6];
    __shared__ float s_sn[kDipTile];
    const int c = blockIdx.y;
    const float* dp = dip + static_cast<std::int64_t>(c) * n_dip * 10;  // per placement: s(3) mp(6) sn(1)
    const double ox = origin[3 * c], oy = origin[3 * c + 1], oz = origin[3 * c + 2];
    const std::int64_t T = static_cast<std::int64_t>(gridDim.x) * blockDim.x;
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    float rx[kNodesPerThread], ry[kNodesPerThread], rz[kNodesPerThread], rn[kNodesPerThread];
    float wx[kNodesPerThread], wy[kNodesPerThread], wz[kNodesPerThread];
    float px[kNodesPerThread], py[kNodesPerThread], pz[kNodesPerThread];
#pragma unroll
    for (int k = 0; k < kNodesPerThread; ++k) {
        const std::int64_t node = t + k * T;
        rx[k] = ry[k] = rz[k] = rn[k] = 0.f;
        if (node < n) {
            rx[k] = static_cast<float>(nodes_mm[3 * node] * 1e-3 - ox);
            ry[k] = static_cast<float>(nodes_mm[3 * node + 1] * 1e-3 - oy);
            rz[k] = static_cast<float>(nodes_mm[3 * node + 2] * 1e-3 - oz);
            rn[k] = fmaf(rz[k], rz[k], fmaf(ry[k], ry[k], rx[k] * rx[k]));
        }
        wx[k] = wy[k] = wz[k] = px[k] = py[k] = pz[k] = 0.f;
    }
    for (int base = 0; base < n_dip; base += kDipTile) {
        const int tile = min(kDipTile, n_dip - base);
        __syncthreads();
        for (int i = threadIdx.x; i < tile; i += kDadtBlock) {
            const float* d = dp + static_cast<std::int64_t>(base + i) * 10;
            s_s[3 * i] = d[0];
            s_s[3 * i + 1] = d[1];
            s_s[3 * i + 2] = d[2];
            for (int q = 0; q < 6; ++q) s_mp[6 * i + q] = d[3 + q];
            s_sn[i] = d[9];
        }
        __syncthreads();
#pragma unroll 4
        for (int j = 0; j < tile; ++j) {
            const float sx = s_s[3 * j], sy = s_s[3 * j + 1], sz = s_s[3 * j + 2], snj = s_sn[j];
            const float m0 = s_mp[6 * j], m1 = s_mp[6 * j + 1], m2 = s_mp[6 * j + 2];
            const float m3 = s_mp[6 * j + 3], m4 = s_mp[6 * j + 4], m5 = s_mp[6 * j + 5];
#pragma unroll
            for (int k = 0; k < kNodesPerThread; ++k) {
                const float rs = fmaf(rz[k], sz, fmaf(ry[k], sy, rx[k] * sx));
                const float d2 = fmaf(-2.f, rs, rn[k] + snj);
                const float inv = rsq_normal(d2);
                const float w = inv * inv * inv;
                wx[k] = fmaf(w, m0, wx[k]);
                wy[k] = fmaf(w, m1, wy[k]);
                wz[k] = fmaf(w, m2, wz[k]);
                px[k] = fmaf(w, m3, px[k]);
                py[k] = fmaf(w, m4, py[k]);
                pz[k] = fmaf(w, m5, pz[k]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kNodesPerThread; ++k) {
        const std::int64_t node = t + k * T;
        if (node < n) {
            float* o = out + node * 3 * K + c;
            o[0] = scale * ((wy[k] * rz[k] - wz[k] * ry[k]) - px[k]);
            o[K] = scale * ((wz[k] * rx[k] - wx[k] * rz[k]) - py[k]);
            o[2 * K] = scale * ((wx[k] * ry[k] - wy[k] * rx[k]) - pz[k]);
        }
    }
}

// The same sum when every dipole moment lies along one direction d (planar windings: the moments
// are the coil normal scaled by mu_j). Then W = d S0 and P = d x S with S0 = sum_j w_j mu_j and
// S = sum_j w_j mu_j s_j, so dA/dt = scale d x (S0 r - S): four multiply-adds per pair instead of
// six. Per dipole (FP32): -2 s (3), |s|^2, mu, mu s (3); d per placement. The distance keeps the
// reference's rn + sn - 2 r.s form.
__global__ void __launch_bounds__(kDadtBlock) dadt_parallel_kernel(const float* __restrict__ dip, int n_dip,
                                                                   const double* __restrict__ nodes_mm, std::int64_t n,
                                                                   const double* __restrict__ origin,
                                                                   const float* __restrict__ dirs, float scale, int K,
                                                                   float* __restrict__ out) {
    __shared__ float4 s_a[kDipTile];  // -2 sx, -2 sy, -2 sz, |s|^2
    __shared__ float4 s_m[kDipTile];  // mu, mu sx, mu sy, mu sz
    const int c = blockIdx.y;
    const float* dp = dip + static_cast<std::int64_t>(c) * n_dip * 8;
    const double ox = origin[3 * c], oy = origin[3 * c + 1], oz = origin[3 * c + 2];
    // A thread's nodes are t, t + NT, t + 2 NT, ...; a grid smaller than NT threads walks the
    // slots in block-uniform steps (so every thread meets the same barriers).
    const std::int64_t NT = (n + kNodesPerThread - 1) / kNodesPerThread;
    const float dx = dirs[3 * c], dy = dirs[3 * c + 1], dz = dirs[3 * c + 2];
    for (std::int64_t slot0 = static_cast<std::int64_t>(blockIdx.x) * blockDim.x; slot0 < NT;
         slot0 += static_cast<std::int64_t>(gridDim.x) * blockDim.x) {
        const std::int64_t t = slot0 + threadIdx.x;
        const std::int64_t T = NT;
        float rx[kNodesPerThread], ry[kNodesPerThread], rz[kNodesPerThread], rn[kNodesPerThread];
        float s0[kNodesPerThread], sx[kNodesPerThread], sy[kNodesPerThread], sz[kNodesPerThread];
#pragma unroll
        for (int k = 0; k < kNodesPerThread; ++k) {
            const std::int64_t node = t + k * T;
            rx[k] = ry[k] = rz[k] = rn[k] = 0.f;
            if (node < n) {
                rx[k] = static_cast<float>(nodes_mm[3 * node] * 1e-3 - ox);
                ry[k] = static_cast<float>(nodes_mm[3 * node + 1] * 1e-3 - oy);
                rz[k] = static_cast<float>(nodes_mm[3 * node + 2] * 1e-3 - oz);
                rn[k] = fmaf(rz[k], rz[k], fmaf(ry[k], ry[k], rx[k] * rx[k]));
            }
            s0[k] = sx[k] = sy[k] = sz[k] = 0.f;
        }
        for (int base = 0; base < n_dip; base += kDipTile) {
            const int tile = min(kDipTile, n_dip - base);
            __syncthreads();
            for (int i = threadIdx.x; i < tile; i += kDadtBlock) {
                const float* d = dp + static_cast<std::int64_t>(base + i) * 8;
                s_a[i] = make_float4(d[0], d[1], d[2], d[3]);
                s_m[i] = make_float4(d[4], d[5], d[6], d[7]);
            }
            __syncthreads();
#pragma unroll 4
            for (int j = 0; j < tile; ++j) {
                const float4 a = s_a[j], m = s_m[j];
#pragma unroll
                for (int k = 0; k < kNodesPerThread; ++k) {
                    const float q = fmaf(rz[k], a.z, fmaf(ry[k], a.y, fmaf(rx[k], a.x, a.w)));
                    const float inv = rsq_normal(q + rn[k]);
                    const float w = inv * inv * inv;
                    s0[k] = fmaf(w, m.x, s0[k]);
                    sx[k] = fmaf(w, m.y, sx[k]);
                    sy[k] = fmaf(w, m.z, sy[k]);
                    sz[k] = fmaf(w, m.w, sz[k]);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kNodesPerThread; ++k) {
            const std::int64_t node = t + k * T;
            if (node < n) {
                const float fx = s0[k] * rx[k] - sx[k], fy = s0[k] * ry[k] - sy[k], fz = s0[k] * rz[k] - sz[k];
                float* o = out + node * 3 * K + c;
                o[0] = scale * (dy * fz - dz * fy);
                o[K] = scale * (dz * fx - dx * fz);
                o[2 * K] = scale * (dx * fy - dy * fx);
            }
        }
    }
}

// Element pass over (element, placement), placements fastest so the K lanes of an element read
// and write contiguous memory:
//   dadt_elm = mean of the four nodal vectors       (dadt_element.cu:6-21)
//   q[corner] = dadt_elm . (g[corner] * neg_vc)     (rhs.cu:65-80, the staged corner terms, FP32)
// Placements beyond k (block padding) get zeros, so their right-hand-side columns are zero.
// With `sigma_tag` (K x n_sigma, FP32), placement c weighs element e by -(vol_e * sigma_c[tag_e])
// instead of neg_vc[e] (the same FP32 product the context's neg_vc holds for its own table).
__global__ void element_pass_kernel(const float* __restrict__ nodal, const std::int32_t* __restrict__ tets,
                                    const float* __restrict__ g, const float* __restrict__ neg_vc, std::int64_t m,
                                    int k, int K, float* __restrict__ elm, float* __restrict__ q,
                                    const std::int32_t* __restrict__ corner_pos, const float* __restrict__ vols,
                                    const std::int32_t* __restrict__ tags, const float* __restrict__ sigma_tag,
                                    int n_sigma) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= m * K) return;
    const std::int64_t e = t / K;
    const int c = static_cast<int>(t % K);
    float d[3] = {0.f, 0.f, 0.f};
    if (c < k) {
        const int4 tt = reinterpret_cast<const int4*>(tets)[e];
        const std::int64_t stride = 3 * static_cast<std::int64_t>(K);
        for (int a = 0; a < 3; ++a) {
            const float v0 = nodal[tt.x * stride + a * K + c], v1 = nodal[tt.y * stride + a * K + c];
            const float v2 = nodal[tt.z * stride + a * K + c], v3 = nodal[tt.w * stride + a * K + c];
            d[a] = 0.25f * (((v0 + v1) + v2) + v3);
        }
    }
    if (elm != nullptr)
        for (int a = 0; a < 3; ++a) elm[(e * 3 + a) * K + c] = d[a];
    const float w = sigma_tag != nullptr ? -(vols[e] * sigma_tag[c * n_sigma + tags[e]]) : neg_vc[e];
    for (int i = 0; i < 4; ++i) {
        const float* gc = g + (4 * e + i) * 3;
        const float w0 = gc[0] * w, w1 = gc[1] * w, w2 = gc[2] * w;
        q[static_cast<std::int64_t>(corner_pos[4 * e + i]) * K + c] = fmaf(d[2], w2, fmaf(d[1], w1, d[0] * w0));
    }
}

// Node gather (rhs.cu:84-104): b[node] = sum of its corners' q in node2corner order, written to
// the solver block with the ground row dropped, and optionally to (k, n) full-node storage. The
// element pass stored the corner terms in that order, so a node's terms are one contiguous run.
__global__ void node_rhs_kernel(const float* __restrict__ q, const std::int32_t* __restrict__ ptr, std::int64_t n,
                                int k, int K, int ground, float* __restrict__ B, float* __restrict__ b_full) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n * K) return;
    const std::int64_t node = t / K;
    const int c = static_cast<int>(t % K);
    float acc = 0.f;
    for (int p = ptr[node]; p < ptr[node + 1]; ++p) acc += q[static_cast<std::int64_t>(p) * K + c];
    if (b_full != nullptr && c < k) b_full[c * n + node] = acc;
    if (node != ground) B[(node - (node > ground)) * K + c] = acc;
}

__global__ void corner_pos_kernel(const std::int32_t* __restrict__ idx, std::int64_t count,
                                  std::int32_t* __restrict__ corner_pos) {
    const std::int64_t p = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p < count) corner_pos[idx[p]] = static_cast<std::int32_t>(p);
}

}  // namespace

void place_transforms(const double* centers, const double* handles, const double* dists, int n_pl, const double* ta,
                      const double* tb, const double* tc, const double* tn, int n_tri, double* out, int* degenerate,
                      int n_cu, hipStream_t s) {
    if (n_pl <= 0 || n_tri <= 0) return;
    const int for_occ = (n_cu + n_pl - 1) / n_pl;
    const int for_size = (n_tri + 1023) / 1024;
    const int n_chunks = std::max(1, std::min(for_occ, for_size));
    DeviceBuffer<Candidate> cand(static_cast<std::size_t>(n_pl) * n_chunks, "placement candidates");
    place_scan_kernel<<<dim3(n_pl, n_chunks), kBlock, 0, s>>>(centers, ta, tb, tc, n_tri, n_chunks, cand.get());
    place_frame_kernel<<<grid_for(n_pl), kBlock, 0, s>>>(centers, handles, dists, ta, tb, tc, tn, cand.get(), n_pl,
                                                         n_chunks, n_tri, out, degenerate);
    sync(s, "placement");
}

void dadt_nbody(const float* dip, int n_dip, const double* nodes_mm, std::int64_t n, const double* origins, int k,
                int K, float scale, float* out, hipStream_t s) {
    if (n <= 0 || k <= 0) return;
    const std::int64_t threads = (n + kNodesPerThread - 1) / kNodesPerThread;
    const unsigned gx = grid_for(threads, kDadtBlock);
    dadt_kernel<<<dim3(gx, k), kDadtBlock, 0, s>>>(dip, n_dip, nodes_mm, n, origins, scale, K, out);
}

void dadt_nbody_parallel(const float* dip8, int n_dip, const double* nodes_mm, std::int64_t n, const double* origins,
                         const float* dirs, int k, int K, float scale, float* out, hipStream_t s, int max_blocks) {
    if (n <= 0 || k <= 0) return;
    const std::int64_t threads = (n + kNodesPerThread - 1) / kNodesPerThread;
    unsigned gx = grid_for(threads, kDadtBlock);
    if (max_blocks > 0) gx = std::max(1u, std::min(gx, static_cast<unsigned>((max_blocks + k - 1) / k)));
    dadt_parallel_kernel<<<dim3(gx, k), kDadtBlock, 0, s>>>(dip8, n_dip, nodes_mm, n, origins, dirs, scale, K, out);
}

void element_pass(const float* nodal, const std::int32_t* tets, const float* g, const float* neg_vc, std::int64_t m,
                  int k, int K, float* elm, float* q, const std::int32_t* corner_pos, hipStream_t s, const float* vols,
                  const std::int32_t* tags, const float* sigma_tag, int n_sigma) {
    if (const unsigned gr = grid_for(m * K))
        element_pass_kernel<<<gr, kBlock, 0, s>>>(nodal, tets, g, neg_vc, m, k, K, elm, q, corner_pos, vols, tags,
                                                  sigma_tag, n_sigma);
}

void corner_positions(const std::int32_t* idx, std::int64_t count, std::int32_t* corner_pos, hipStream_t s) {
    if (const unsigned gr = grid_for(count)) corner_pos_kernel<<<gr, kBlock, 0, s>>>(idx, count, corner_pos);
}

void node_rhs(const float* q, const std::int32_t* ptr, std::int64_t n, int k, int K, int ground, float* B,
              float* b_full, hipStream_t s) {
    if (const unsigned gr = grid_for(n * K)) node_rhs_kernel<<<gr, kBlock, 0, s>>>(q, ptr, n, k, K, ground, B, b_full);
}

}  // namespace cn
