// Post-processing of a solved block: potential scatter, E reconstruction and the region metrics.
#include <algorithm>
#include <cstring>

#include "coil.hpp"
#include "prims.hpp"

namespace cn {

namespace {

__device__ __forceinline__ double potential(const double* X, std::int64_t node, int c, int K, int ground) {
    if (node == ground) return 0.0;
    return X[(node - (node > ground)) * K + c];
}

__global__ void scatter_kernel(const double* __restrict__ X, std::int64_t n, int k, int K, int ground,
                               double* __restrict__ v) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n * k) return;
    const std::int64_t c = t / n, node = t % n;
    v[t] = potential(X, node, static_cast<int>(c), K, ground);
}

// reconstruct.cu:36-78: E = -sum_i v_i g_i - dadt_elm with the gradient in FP64, |E| per element.
// Threads over (element, placement), placements fastest; E (m, 3, K), |E| (m, K).
__global__ void reconstruct_kernel(const double* __restrict__ X, const std::int32_t* __restrict__ tets,
                                   const float* __restrict__ g, const float* __restrict__ elm, std::int64_t m, int k,
                                   int K, int ground, float* __restrict__ E, float* __restrict__ magn) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= m * K) return;
    const std::int64_t e = t / K;
    const int c = static_cast<int>(t % K);
    if (c >= k) return;
    const int4 tt = reinterpret_cast<const int4*>(tets)[e];
    const int nodes[4] = {tt.x, tt.y, tt.z, tt.w};
    const float* ge = g + 12 * e;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    for (int i = 0; i < 4; ++i) {
        const double vi = potential(X, nodes[i], c, K, ground);
        gx = fma(vi, static_cast<double>(ge[3 * i]), gx);
        gy = fma(vi, static_cast<double>(ge[3 * i + 1]), gy);
        gz = fma(vi, static_cast<double>(ge[3 * i + 2]), gz);
    }
    const double ex = -gx - static_cast<double>(elm[(e * 3) * K + c]);
    const double ey = -gy - static_cast<double>(elm[(e * 3 + 1) * K + c]);
    const double ez = -gz - static_cast<double>(elm[(e * 3 + 2) * K + c]);
    if (E != nullptr) {
        E[(e * 3) * K + c] = static_cast<float>(ex);
        E[(e * 3 + 1) * K + c] = static_cast<float>(ey);
        E[(e * 3 + 2) * K + c] = static_cast<float>(ez);
    }
    magn[e * K + c] = static_cast<float>(sqrt(fma(ez, ez, fma(ey, ey, ex * ex))));
}

__global__ void unpack_kernel(const float* __restrict__ in, std::int64_t rows, int inner, int k, int K,
                              float* __restrict__ out) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const std::int64_t per = rows * inner;
    if (t >= per * k) return;
    const std::int64_t c = t / per, ri = t % per;  // ri = r * inner + d
    out[t] = in[ri * K + c];
}

// ------------------------------------------------------------------------------------------------
// Region metrics (metrics.py:82-135,209-243). Sums are FP64 with a fixed association: per-block
// partials over fixed contiguous ranges, folded on the host in block order.
// ------------------------------------------------------------------------------------------------
constexpr int kStatBlocks = 256;

__global__ void gather_region_kernel(const float* __restrict__ magn, int K, const std::int32_t* __restrict__ idx,
                                     std::int64_t nr, int c, float* __restrict__ vals, std::uint32_t* __restrict__ keys) {
    const std::int64_t j = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= nr) return;
    const float v = magn[static_cast<std::int64_t>(idx[j]) * K + c];
    vals[j] = v;
    keys[j] = __float_as_uint(v);  // |E| >= 0: the bit pattern orders like the value
}

// Per block: [max, argmax, sum w, sum w m, sum w m bx, by, bz, unused]
__global__ void stats_kernel(const float* __restrict__ vals, const std::int32_t* __restrict__ idx,
                             const float* __restrict__ vols, const double* __restrict__ bary, std::int64_t nr,
                             double* __restrict__ out) {
    __shared__ double s[6][kBlock];
    __shared__ float smax[kBlock];
    __shared__ std::int64_t sarg[kBlock];
    const std::int64_t per = (nr + kStatBlocks - 1) / kStatBlocks;
    const std::int64_t lo = blockIdx.x * per, hi = min(lo + per, nr);
    double acc[6] = {0, 0, 0, 0, 0, 0};
    float mx = -INFINITY;
    std::int64_t arg = INT64_MAX;
    for (std::int64_t j = lo + threadIdx.x; j < hi; j += kBlock) {
        const float v = vals[j];
        if (v > mx) {
            mx = v;
            arg = j;
        }
        const std::int64_t e = idx[j];
        const double w = vols[e];
        const double wm = w * static_cast<double>(v);
        acc[0] += w;
        acc[1] += wm;
        acc[2] += wm * bary[3 * e];
        acc[3] += wm * bary[3 * e + 1];
        acc[4] += wm * bary[3 * e + 2];
    }
    for (int q = 0; q < 6; ++q) s[q][threadIdx.x] = acc[q];
    smax[threadIdx.x] = mx;
    sarg[threadIdx.x] = arg;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            for (int q = 0; q < 6; ++q) s[q][threadIdx.x] += s[q][threadIdx.x + w];
            const float om = smax[threadIdx.x + w];
            const std::int64_t oa = sarg[threadIdx.x + w];
            if (om > smax[threadIdx.x] || (om == smax[threadIdx.x] && oa < sarg[threadIdx.x])) {
                smax[threadIdx.x] = om;
                sarg[threadIdx.x] = oa;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double* o = out + 8 * static_cast<std::int64_t>(blockIdx.x);
        o[0] = smax[0];
        o[1] = static_cast<double>(sarg[0]);
        for (int q = 0; q < 5; ++q) o[2 + q] = s[q][0];
        o[7] = 0.0;
    }
}

__global__ void variance_kernel(const float* __restrict__ vals, const std::int32_t* __restrict__ idx,
                                const float* __restrict__ vols, std::int64_t nr, double mean,
                                double* __restrict__ out) {
    __shared__ double s[kBlock];
    const std::int64_t per = (nr + kStatBlocks - 1) / kStatBlocks;
    const std::int64_t lo = blockIdx.x * per, hi = min(lo + per, nr);
    double acc = 0.0;
    for (std::int64_t j = lo + threadIdx.x; j < hi; j += kBlock) {
        const double d = static_cast<double>(vals[j]) - mean;
        acc += static_cast<double>(vols[idx[j]]) * d * d;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

__global__ void sorted_weights_kernel(const std::int32_t* __restrict__ order, const std::int32_t* __restrict__ idx,
                                      const float* __restrict__ vols, std::int64_t nr, double* __restrict__ w) {
    const std::int64_t j = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j < nr) w[j] = static_cast<double>(vols[idx[order[j]]]);
}

// Fixed-order inclusive scan of one 1024-element tile per block: 4 sequential items per thread,
// a Hillis-Steele pass over the 256 thread totals, then the tile total to `totals`.
constexpr int kScanItems = 4;
constexpr int kScanTile = kBlock * kScanItems;

__global__ void scan_tiles_kernel(double* __restrict__ a, std::int64_t n, double* __restrict__ totals) {
    __shared__ double s[kBlock];
    const std::int64_t base = static_cast<std::int64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
    double v[kScanItems];
    double run = 0.0;
    for (int i = 0; i < kScanItems; ++i) {
        const std::int64_t j = base + i;
        run += (j < n) ? a[j] : 0.0;
        v[i] = run;
    }
    s[threadIdx.x] = run;
    __syncthreads();
    for (int off = 1; off < kBlock; off <<= 1) {
        const double add = threadIdx.x >= off ? s[threadIdx.x - off] : 0.0;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    const double before = threadIdx.x > 0 ? s[threadIdx.x - 1] : 0.0;
    for (int i = 0; i < kScanItems; ++i) {
        const std::int64_t j = base + i;
        if (j < n) a[j] = before + v[i];
    }
    if (threadIdx.x == kBlock - 1) totals[blockIdx.x] = s[kBlock - 1];
}

__global__ void add_offsets_kernel(double* __restrict__ a, std::int64_t n, const double* __restrict__ scanned_totals) {
    const std::int64_t j = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const std::int64_t tile = j / kScanTile;
    if (tile > 0) a[j] += scanned_totals[tile - 1];
}

// np.interp(q, pos, v) with pos = (cw - w/2) / cw[-1] (metrics.py:82-95), q in FP64 from the FP32
// value the reference passes.
__global__ void quantile_kernel(const std::uint32_t* __restrict__ sorted_keys, const double* __restrict__ w,
                                const double* __restrict__ cw, std::int64_t nr, const double* __restrict__ qs, int nq,
                                double* __restrict__ out) {
    const int t = threadIdx.x;
    if (t >= nq) return;
    const double total = cw[nr - 1];
    const double x = qs[t];
    auto pos = [&](std::int64_t j) { return (cw[j] - 0.5 * w[j]) / total; };
    auto val = [&](std::int64_t j) { return static_cast<double>(__uint_as_float(sorted_keys[j])); };
    if (nr == 1 || x <= pos(0)) {  // np.interp clamps below the first sample point
        out[t] = val(0);
        return;
    }
    if (x >= pos(nr - 1)) {
        out[t] = val(nr - 1);
        return;
    }
    std::int64_t lo = 0, hi = nr - 1;  // pos(lo) < x < pos(hi) invariant (x strictly inside)
    while (hi - lo > 1) {
        const std::int64_t mid = lo + (hi - lo) / 2;
        if (pos(mid) <= x)
            lo = mid;
        else
            hi = mid;
    }
    const double x0 = pos(lo), x1 = pos(lo + 1);
    const double f0 = val(lo), f1 = val(lo + 1);
    const double slope = (f1 - f0) / (x1 - x0);
    out[t] = slope * (x - x0) + f0;
}

// Volume with |E| >= thr: the first sorted position at or above thr, then total minus prefix.
__global__ void focality_kernel(const std::uint32_t* __restrict__ sorted_keys, const double* __restrict__ cw,
                                std::int64_t nr, const float* __restrict__ thr, int nf, double* __restrict__ out) {
    const int t = threadIdx.x;
    if (t >= nf) return;
    const float th = thr[t];
    std::int64_t lo = 0, hi = nr;
    while (lo < hi) {
        const std::int64_t mid = lo + (hi - lo) / 2;
        if (__uint_as_float(sorted_keys[mid]) >= th)
            hi = mid;
        else
            lo = mid + 1;
    }
    out[t] = cw[nr - 1] - (lo > 0 ? cw[lo - 1] : 0.0);
}

void prefix_sum_fixed(double* a, std::int64_t n, hipStream_t s) {
    if (n <= 0) return;
    const std::int64_t tiles = (n + kScanTile - 1) / kScanTile;
    DeviceBuffer<double> totals(static_cast<std::size_t>(tiles), "scan totals");
    scan_tiles_kernel<<<static_cast<unsigned>(tiles), kBlock, 0, s>>>(a, n, totals.get());
    if (tiles > 1) {
        prefix_sum_fixed(totals.get(), tiles, s);
        add_offsets_kernel<<<grid_for(n), kBlock, 0, s>>>(a, n, totals.get());
    }
    sync(s, "prefix sum");
}

}  // namespace

void scatter_potential(const double* X, std::int64_t n, int k, int K, int ground, double* v, hipStream_t s) {
    if (const unsigned g = grid_for(n * k)) scatter_kernel<<<g, kBlock, 0, s>>>(X, n, k, K, ground, v);
}

void reconstruct(const double* X, const std::int32_t* tets, const float* g, const float* elm, std::int64_t m, int k,
                 int K, int ground, float* E, float* magn, hipStream_t s) {
    if (const unsigned gr = grid_for(m * K))
        reconstruct_kernel<<<gr, kBlock, 0, s>>>(X, tets, g, elm, m, k, K, ground, E, magn);
}

void unpack(const float* in, std::int64_t rows, int inner, int k, int K, float* out, hipStream_t s) {
    if (const unsigned gr = grid_for(rows * inner * k)) unpack_kernel<<<gr, kBlock, 0, s>>>(in, rows, inner, k, K, out);
}

void region_metrics(const float* magn, std::int64_t m, int k, int K, const std::int32_t* idx, std::int64_t nr,
                    const float* vols, const double* bary, double* out_host, hipStream_t s) {
    (void)m;
    std::memset(out_host, 0, sizeof(double) * kMetricsLen * k);
    if (nr <= 0) return;
    DeviceBuffer<float> vals(static_cast<std::size_t>(nr), "region values");
    DeviceBuffer<std::uint32_t> keys(static_cast<std::size_t>(nr), "region keys"), sorted(static_cast<std::size_t>(nr), "sorted keys");
    DeviceBuffer<std::int32_t> iota(static_cast<std::size_t>(nr), "iota"), order(static_cast<std::size_t>(nr), "order");
    DeviceBuffer<double> w(static_cast<std::size_t>(nr), "sorted w"), cw(static_cast<std::size_t>(nr), "cum w");
    DeviceBuffer<double> part(8 * kStatBlocks, "stat partials"), q_out(8, "quantiles");
    DeviceBuffer<double> qd(4, "q"), foc(2, "focality");
    DeviceBuffer<float> thr(2, "thresholds");
    // metrics.py:131 passes the percentiles through the values' dtype (FP32), then FP64.
    const double qs[4] = {static_cast<double>(0.5f), static_cast<double>(0.95f), static_cast<double>(0.99f),
                          static_cast<double>(0.999f)};
    h2d(qd.get(), qs, 4, s);
    iota_i32(iota.get(), nr, s);
    std::vector<double> hp(8 * kStatBlocks), hv(kStatBlocks);
    for (int c = 0; c < k; ++c) {
        double* o = out_host + kMetricsLen * c;
        gather_region_kernel<<<grid_for(nr), kBlock, 0, s>>>(magn, K, idx, nr, c, vals.get(), keys.get());
        stats_kernel<<<kStatBlocks, kBlock, 0, s>>>(vals.get(), idx, vols, bary, nr, part.get());
        d2h(hp.data(), part.get(), hp.size(), s);
        sync(s, "region stats");
        double mx = -INFINITY, sw = 0, swm = 0, sb[3] = {0, 0, 0};
        std::int64_t arg = -1;
        for (int b = 0; b < kStatBlocks; ++b) {
            const double* p = hp.data() + 8 * b;
            if (p[0] > mx || (p[0] == mx && static_cast<std::int64_t>(p[1]) < arg)) {
                mx = p[0];
                arg = static_cast<std::int64_t>(p[1]);
            }
            sw += p[2];
            swm += p[3];
            for (int d = 0; d < 3; ++d) sb[d] += p[4 + d];
        }
        const double mean = swm / sw;
        variance_kernel<<<kStatBlocks, kBlock, 0, s>>>(vals.get(), idx, vols, nr, mean, part.get());
        d2h(hv.data(), part.get(), hv.size(), s);
        // peak location (NaN when no value compared greater than -inf, i.e. all NaN)
        std::int32_t e_peak = -1;
        if (arg >= 0 && arg < nr) d2h(&e_peak, idx + arg, 1, s);
        sync(s, "region variance");
        double var = 0;
        for (int b = 0; b < kStatBlocks; ++b) var += hv[b];
        var /= sw;
        double loc[3] = {NAN, NAN, NAN};
        if (e_peak >= 0) d2h(loc, bary + 3 * static_cast<std::int64_t>(e_peak), 3, s);
        // distribution quantiles
        sort_pairs_u32(keys.get(), sorted.get(), iota.get(), order.get(), nr, 32, s);
        sorted_weights_kernel<<<grid_for(nr), kBlock, 0, s>>>(order.get(), idx, vols, nr, w.get());
        d2d(cw.get(), w.get(), static_cast<std::size_t>(nr), s);
        prefix_sum_fixed(cw.get(), nr, s);
        quantile_kernel<<<1, 64, 0, s>>>(sorted.get(), w.get(), cw.get(), nr, qd.get(), 4, q_out.get());
        double qv[4];
        d2h(qv, q_out.get(), 4, s);
        sync(s, "quantiles");
        const float th[2] = {static_cast<float>(0.5 * qv[3]), static_cast<float>(0.75 * qv[3])};
        h2d(thr.get(), th, 2, s);
        focality_kernel<<<1, 64, 0, s>>>(sorted.get(), cw.get(), nr, thr.get(), 2, foc.get());
        double fv[2];
        d2h(fv, foc.get(), 2, s);
        sync(s, "focality");
        o[0] = mx;
        o[1] = loc[0];
        o[2] = loc[1];
        o[3] = loc[2];
        for (int d = 0; d < 3; ++d) o[4 + d] = sb[d] / swm;
        o[7] = sw;
        o[8] = fv[0];
        o[9] = fv[1];
        o[10] = mean;
        o[11] = std::sqrt(var);
        for (int q = 0; q < 4; ++q) o[12 + q] = qv[q];
    }
}

}  // namespace cn
