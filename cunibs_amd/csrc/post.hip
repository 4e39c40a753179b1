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
// Threads over (element, placement), placements fastest; E (m, 3, K), |E| (m, K). The element
// dA/dt is the corner average of the nodal dA/dt (n, 3, Kn), formed here exactly as the element
// pass forms it (dadt_element.cu); Kn = 1 serves every column from one placement.
__global__ void reconstruct_kernel(const double* __restrict__ X, const std::int32_t* __restrict__ tets,
                                   const float* __restrict__ g, const float* __restrict__ nodal, int Kn,
                                   std::int64_t m, int k, int K, int ground, float* __restrict__ E,
                                   float* __restrict__ magn) {
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
    const int cn = Kn == 1 ? 0 : c;
    const std::int64_t stride = 3 * static_cast<std::int64_t>(Kn);
    float d[3];
    for (int a = 0; a < 3; ++a) {
        const float v0 = nodal[tt.x * stride + a * Kn + cn], v1 = nodal[tt.y * stride + a * Kn + cn];
        const float v2 = nodal[tt.z * stride + a * Kn + cn], v3 = nodal[tt.w * stride + a * Kn + cn];
        d[a] = 0.25f * (((v0 + v1) + v2) + v3);
    }
    const double ex = -gx - static_cast<double>(d[0]);
    const double ey = -gy - static_cast<double>(d[1]);
    const double ez = -gz - static_cast<double>(d[2]);
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

// Every placement of the block at once: thread j reads its element's K-wide |E| row and writes
// vals[c nr + j], the sort key c << 32 | bits(|E|) (|E| >= 0, so the bits order like the values
// and the placement index keeps the placements apart) and, as the sort's payload, the bits of the
// element's volume.
__global__ void gather_region_kernel(const float* __restrict__ magn, int K, int k, const std::int32_t* __restrict__ idx,
                                     const float* __restrict__ vols, std::int64_t nr, float* __restrict__ vals,
                                     std::uint64_t* __restrict__ keys, std::int32_t* __restrict__ vol_bits) {
    const std::int64_t j = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= nr) return;
    const std::int64_t e = idx[j];
    const float* row = magn + e * K;
    const std::int32_t vb = __float_as_int(vols[e]);
    for (int c = 0; c < k; ++c) {
        const float v = row[c];
        const std::int64_t t = c * nr + j;
        vals[t] = v;
        keys[t] = (static_cast<std::uint64_t>(c) << 32) | __float_as_uint(v);
        vol_bits[t] = vb;
    }
}

// Per block of placement blockIdx.y: [max, argmax, sum w, sum w m, sum w m bx, by, bz, unused]
__global__ void stats_kernel(const float* __restrict__ vals_all, const std::int32_t* __restrict__ idx,
                             const float* __restrict__ vols, const double* __restrict__ bary, std::int64_t nr,
                             double* __restrict__ out_all) {
    __shared__ double s[6][kBlock];
    __shared__ float smax[kBlock];
    __shared__ std::int64_t sarg[kBlock];
    const float* vals = vals_all + blockIdx.y * nr;
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
        double* o = out_all + 8 * (static_cast<std::int64_t>(blockIdx.y) * kStatBlocks + blockIdx.x);
        o[0] = smax[0];
        o[1] = static_cast<double>(sarg[0]);
        for (int q = 0; q < 5; ++q) o[2 + q] = s[q][0];
        o[7] = 0.0;
    }
}

// sum w (m - mean_c)^2 per block of placement blockIdx.y.
__global__ void variance_kernel(const float* __restrict__ vals_all, const std::int32_t* __restrict__ idx,
                                const float* __restrict__ vols, std::int64_t nr, const double* __restrict__ means,
                                double* __restrict__ out_all) {
    __shared__ double s[kBlock];
    const float* vals = vals_all + blockIdx.y * nr;
    const double mean = means[blockIdx.y];
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
    if (threadIdx.x == 0) out_all[static_cast<std::int64_t>(blockIdx.y) * kStatBlocks + blockIdx.x] = s[0];
}

// Barycentre of each placement's peak element (region position args[c]; NaN when < 0).
__global__ void peak_location_kernel(const double* __restrict__ args, const std::int32_t* __restrict__ idx,
                                     const double* __restrict__ bary, int k, double* __restrict__ loc) {
    const int c = threadIdx.x;
    if (c >= k) return;
    const double a = args[c];
    const std::int64_t e = a >= 0.0 ? static_cast<std::int64_t>(idx[static_cast<std::int64_t>(a)]) : -1;
    for (int d = 0; d < 3; ++d) loc[3 * c + d] = e >= 0 ? bary[3 * e + d] : NAN;
}

// w[t] = cw[t] = volume of the element at sorted position t (every placement's segment at once);
// cw is then scanned in place.
__global__ void sorted_weights_kernel(const std::int32_t* __restrict__ vol_bits, std::int64_t total,
                                      double* __restrict__ w, double* __restrict__ cw) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < total) w[t] = cw[t] = static_cast<double>(__int_as_float(vol_bits[t]));
}

// Fixed-order inclusive scan, per segment (one placement each: `len` elements at stride `stride`),
// of one 1024-element tile per block: 4 sequential items per thread, a Hillis-Steele pass over
// the 256 thread totals, then the tile total to `totals` (segment-major, `tiles` per segment).
// The association depends on the segment length alone.
constexpr int kScanItems = 4;
constexpr int kScanTile = kBlock * kScanItems;

__global__ void scan_tiles_kernel(double* __restrict__ a, std::int64_t len, std::int64_t stride, std::int64_t tiles,
                                  double* __restrict__ totals) {
    __shared__ double s[kBlock];
    const std::int64_t seg = blockIdx.x / tiles, tile = blockIdx.x % tiles;
    double* const sa = a + seg * stride;
    const std::int64_t base = tile * kScanTile + threadIdx.x * kScanItems;
    double v[kScanItems];
    double run = 0.0;
    for (int i = 0; i < kScanItems; ++i) {
        const std::int64_t j = base + i;
        run += (j < len) ? sa[j] : 0.0;
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
        if (j < len) sa[j] = before + v[i];
    }
    if (threadIdx.x == kBlock - 1) totals[blockIdx.x] = s[kBlock - 1];
}

__global__ void add_offsets_kernel(double* __restrict__ a, std::int64_t len, std::int64_t stride, std::int64_t nseg,
                                   std::int64_t tiles, const double* __restrict__ scanned_totals) {
    const std::int64_t t = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= nseg * len) return;
    const std::int64_t seg = t / len, j = t % len;
    const std::int64_t tile = j / kScanTile;
    if (tile > 0) a[seg * stride + j] += scanned_totals[seg * tiles + tile - 1];
}

// Scratch doubles the segmented scan of `len` elements per segment needs for its tile totals.
std::int64_t scan_scratch(std::int64_t len, std::int64_t nseg) {
    std::int64_t need = 0;
    while (len > 0) {
        const std::int64_t tiles = (len + kScanTile - 1) / kScanTile;
        need += tiles * nseg;
        if (tiles <= 1) break;
        len = tiles;
    }
    return need;
}

void prefix_sum_fixed(double* a, std::int64_t len, std::int64_t stride, std::int64_t nseg, double* scratch,
                      hipStream_t s) {
    if (len <= 0 || nseg <= 0) return;
    const std::int64_t tiles = (len + kScanTile - 1) / kScanTile;
    double* const totals = scratch;
    scan_tiles_kernel<<<static_cast<unsigned>(tiles * nseg), kBlock, 0, s>>>(a, len, stride, tiles, totals);
    if (tiles > 1) {
        prefix_sum_fixed(totals, tiles, tiles, nseg, scratch + tiles * nseg, s);
        add_offsets_kernel<<<grid_for(nseg * len), kBlock, 0, s>>>(a, len, stride, nseg, tiles, totals);
    }
}

// np.interp(q, pos, v) with pos = (cw - w/2) / cw[-1] (metrics.py:82-95), q in FP64 from the FP32
// value the reference passes; one block per placement.
__global__ void quantile_kernel(const std::uint64_t* __restrict__ sorted_all, const double* __restrict__ w_all,
                                const double* __restrict__ cw_all, std::int64_t nr, const double* __restrict__ qs,
                                int nq, double* __restrict__ out_all) {
    const int t = threadIdx.x;
    if (t >= nq) return;
    const std::int64_t off = static_cast<std::int64_t>(blockIdx.x) * nr;
    const std::uint64_t* sorted_keys = sorted_all + off;
    const double* w = w_all + off;
    const double* cw = cw_all + off;
    double* out = out_all + static_cast<std::int64_t>(blockIdx.x) * nq;
    const double total = cw[nr - 1];
    const double x = qs[t];
    auto pos = [&](std::int64_t j) { return (cw[j] - 0.5 * w[j]) / total; };
    auto val = [&](std::int64_t j) {
        return static_cast<double>(__uint_as_float(static_cast<std::uint32_t>(sorted_keys[j])));
    };
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

// Volume with |E| >= frac * P99.9 (the anchor percentile rounded to FP32, as the reference's
// threshold is): the first sorted position at or above it, then total minus prefix. One block per
// placement; q holds nq quantiles per placement, the anchor last.
__global__ void focality_kernel(const std::uint64_t* __restrict__ sorted_all, const double* __restrict__ cw_all,
                                std::int64_t nr, const double* __restrict__ q, int nq, const double* __restrict__ fracs,
                                int nf, double* __restrict__ out_all) {
    const int t = threadIdx.x;
    if (t >= nf) return;
    const std::int64_t off = static_cast<std::int64_t>(blockIdx.x) * nr;
    const std::uint64_t* sorted_keys = sorted_all + off;
    const double* cw = cw_all + off;
    const float th = static_cast<float>(fracs[t] * q[static_cast<std::int64_t>(blockIdx.x) * nq + nq - 1]);
    std::int64_t lo = 0, hi = nr;
    while (lo < hi) {
        const std::int64_t mid = lo + (hi - lo) / 2;
        if (__uint_as_float(static_cast<std::uint32_t>(sorted_keys[mid])) >= th)
            hi = mid;
        else
            lo = mid + 1;
    }
    out_all[static_cast<std::int64_t>(blockIdx.x) * nf + t] = cw[nr - 1] - (lo > 0 ? cw[lo - 1] : 0.0);
}

}  // namespace

void scatter_potential(const double* X, std::int64_t n, int k, int K, int ground, double* v, hipStream_t s) {
    if (const unsigned g = grid_for(n * k)) scatter_kernel<<<g, kBlock, 0, s>>>(X, n, k, K, ground, v);
}

void reconstruct(const double* X, const std::int32_t* tets, const float* g, const float* nodal, int Kn,
                 std::int64_t m, int k, int K, int ground, float* E, float* magn, hipStream_t s) {
    require(Kn == 1 || Kn == K, "nodal dA/dt must have one column or the block's width");
    if (const unsigned gr = grid_for(m * K))
        reconstruct_kernel<<<gr, kBlock, 0, s>>>(X, tets, g, nodal, Kn, m, k, K, ground, E, magn);
}

void unpack(const float* in, std::int64_t rows, int inner, int k, int K, float* out, hipStream_t s) {
    if (const unsigned gr = grid_for(rows * inner * k)) unpack_kernel<<<gr, kBlock, 0, s>>>(in, rows, inner, k, K, out);
}

void region_metrics(const float* magn, std::int64_t m, int k, int K, const std::int32_t* idx, std::int64_t nr,
                    const float* vols, const double* bary, MetricsWorkspace& ws, double* out_host, hipStream_t s) {
    (void)m;
    std::memset(out_host, 0, sizeof(double) * kMetricsLen * k);
    if (nr <= 0 || k <= 0) return;
    const std::int64_t total = nr * k;
    ws.vals.ensure(static_cast<std::size_t>(total), "region values");
    ws.keys.ensure(static_cast<std::size_t>(total), "region keys");
    ws.sorted.ensure(static_cast<std::size_t>(total), "sorted keys");
    ws.pos.ensure(static_cast<std::size_t>(total), "region positions");
    ws.order.ensure(static_cast<std::size_t>(total), "sorted positions");
    ws.w.ensure(static_cast<std::size_t>(total), "sorted volumes");
    ws.cw.ensure(static_cast<std::size_t>(total), "cumulative volumes");
    ws.totals.ensure(static_cast<std::size_t>(std::max<std::int64_t>(scan_scratch(nr, k), 1)), "scan totals");
    ws.part.ensure(static_cast<std::size_t>(8) * kStatBlocks * k, "stat partials");
    // small: qs[4] fracs[2] | means[k] args[k] | q[4k] foc[2k] loc[3k]
    ws.small.ensure(static_cast<std::size_t>(6 + 11 * k), "metric scalars");
    double* const qd = ws.small.get();
    double* const fracs = qd + 4;
    double* const means = qd + 6;
    double* const args = means + k;
    double* const q_out = args + k;
    double* const foc = q_out + 4 * k;
    double* const loc = foc + 2 * k;
    // metrics.py:131 passes the percentiles through the values' dtype (FP32), then FP64.
    const double consts[6] = {static_cast<double>(0.5f), static_cast<double>(0.95f), static_cast<double>(0.99f),
                              static_cast<double>(0.999f), 0.5, 0.75};
    h2d(qd, consts, 6, s);
    gather_region_kernel<<<grid_for(nr), kBlock, 0, s>>>(magn, K, k, idx, vols, nr, ws.vals.get(), ws.keys.get(),
                                                         ws.pos.get());
    stats_kernel<<<dim3(kStatBlocks, k), kBlock, 0, s>>>(ws.vals.get(), idx, vols, bary, nr, ws.part.get());
    // distribution quantiles and focality: a stable sort by (placement, |E|) carrying the volumes,
    // their per-placement fixed-order prefix sums
    int bits = 32;
    while ((1 << (bits - 32)) < k) ++bits;
    sort_pairs_u64_ws(ws.keys.get(), ws.sorted.get(), ws.pos.get(), ws.order.get(), total, bits, ws.scratch, s);
    sorted_weights_kernel<<<grid_for(total), kBlock, 0, s>>>(ws.order.get(), total, ws.w.get(), ws.cw.get());
    prefix_sum_fixed(ws.cw.get(), nr, nr, k, ws.totals.get(), s);
    quantile_kernel<<<k, 64, 0, s>>>(ws.sorted.get(), ws.w.get(), ws.cw.get(), nr, qd, 4, q_out);
    focality_kernel<<<k, 64, 0, s>>>(ws.sorted.get(), ws.cw.get(), nr, q_out, 4, fracs, 2, foc);
    std::vector<double> hp(static_cast<std::size_t>(8) * kStatBlocks * k);
    d2h(hp.data(), ws.part.get(), hp.size(), s);
    sync(s, "region stats");
    std::vector<double> mean(k), hargs(k), sw(k), swm(k), mx(k), sb(3 * k);
    for (int c = 0; c < k; ++c) {
        double best = -INFINITY, a = -1.0, w = 0, wm = 0, b[3] = {0, 0, 0};
        std::int64_t arg = -1;
        for (int blk = 0; blk < kStatBlocks; ++blk) {
            const double* p = hp.data() + 8 * (static_cast<std::size_t>(c) * kStatBlocks + blk);
            if (p[0] > best || (p[0] == best && static_cast<std::int64_t>(p[1]) < arg)) {
                best = p[0];
                arg = static_cast<std::int64_t>(p[1]);
            }
            w += p[2];
            wm += p[3];
            for (int d = 0; d < 3; ++d) b[d] += p[4 + d];
        }
        // peak location is NaN when no value compared greater than -inf (all NaN)
        a = (arg >= 0 && arg < nr) ? static_cast<double>(arg) : -1.0;
        mx[c] = best;
        hargs[c] = a;
        sw[c] = w;
        swm[c] = wm;
        for (int d = 0; d < 3; ++d) sb[3 * c + d] = b[d];
        mean[c] = wm / w;
    }
    h2d(means, mean.data(), k, s);
    h2d(args, hargs.data(), k, s);
    variance_kernel<<<dim3(kStatBlocks, k), kBlock, 0, s>>>(ws.vals.get(), idx, vols, nr, means, ws.part.get());
    peak_location_kernel<<<1, 64, 0, s>>>(args, idx, bary, k, loc);
    std::vector<double> hv(static_cast<std::size_t>(kStatBlocks) * k), res(9 * static_cast<std::size_t>(k));
    d2h(hv.data(), ws.part.get(), hv.size(), s);
    d2h(res.data(), q_out, res.size(), s);
    sync(s, "region variance");
    const double* qv = res.data();
    const double* fv = qv + 4 * k;
    const double* lv = fv + 2 * k;
    for (int c = 0; c < k; ++c) {
        double var = 0;
        for (int blk = 0; blk < kStatBlocks; ++blk) var += hv[static_cast<std::size_t>(c) * kStatBlocks + blk];
        var /= sw[c];
        double* o = out_host + kMetricsLen * c;
        o[0] = mx[c];
        for (int d = 0; d < 3; ++d) o[1 + d] = lv[3 * c + d];
        for (int d = 0; d < 3; ++d) o[4 + d] = sb[3 * c + d] / swm[c];
        o[7] = sw[c];
        o[8] = fv[2 * c];
        o[9] = fv[2 * c + 1];
        o[10] = mean[c];
        o[11] = std::sqrt(var);
        for (int qi = 0; qi < 4; ++qi) o[12 + qi] = qv[4 * c + qi];
    }
}

}  // namespace cn
