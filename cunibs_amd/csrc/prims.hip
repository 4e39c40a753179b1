// Setup-time primitives: stable radix sorts and integer scans over rocPRIM's device algorithms.
// These are exact by definition (integer keys, stable order), so which library runs them cannot
// change a result; every floating-point reduction in the library is written out by hand instead.
#include <algorithm>
#include <climits>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "prims.hpp"

namespace cn {

template <typename K>
static void sort_pairs_impl(const K* keys_in, K* keys_out, const std::int32_t* vals_in,
                            std::int32_t* vals_out, std::int64_t n, int end_bit, hipStream_t s) {
    if (n <= 0) return;
    std::size_t bytes = 0;
    check(rocprim::radix_sort_pairs(nullptr, bytes, keys_in, keys_out, vals_in, vals_out,
                                    static_cast<std::size_t>(n), 0, end_bit, s),
          "radix sort (size)");
    DeviceBuffer<unsigned char> tmp(bytes ? bytes : 1, "radix sort scratch");
    check(rocprim::radix_sort_pairs(tmp.get(), bytes, keys_in, keys_out, vals_in, vals_out,
                                    static_cast<std::size_t>(n), 0, end_bit, s),
          "radix sort");
    sync(s, "radix sort");  // the scratch dies with this frame
}

void sort_pairs_u64(const std::uint64_t* keys_in, std::uint64_t* keys_out,
                    const std::int32_t* vals_in, std::int32_t* vals_out, std::int64_t n,
                    int end_bit, hipStream_t s) {
    sort_pairs_impl(keys_in, keys_out, vals_in, vals_out, n, end_bit, s);
}

void sort_pairs_u64_ws(const std::uint64_t* keys_in, std::uint64_t* keys_out, const std::int32_t* vals_in,
                       std::int32_t* vals_out, std::int64_t n, int end_bit, DeviceBuffer<unsigned char>& scratch,
                       hipStream_t s) {
    if (n <= 0) return;
    std::size_t bytes = 0;
    check(rocprim::radix_sort_pairs(nullptr, bytes, keys_in, keys_out, vals_in, vals_out,
                                    static_cast<std::size_t>(n), 0, end_bit, s),
          "radix sort (size)");
    scratch.ensure(bytes ? bytes : 1, "radix sort scratch");
    check(rocprim::radix_sort_pairs(scratch.get(), bytes, keys_in, keys_out, vals_in, vals_out,
                                    static_cast<std::size_t>(n), 0, end_bit, s),
          "radix sort");
}

void sort_pairs_u32(const std::uint32_t* keys_in, std::uint32_t* keys_out,
                    const std::int32_t* vals_in, std::int32_t* vals_out, std::int64_t n,
                    int end_bit, hipStream_t s) {
    sort_pairs_impl(keys_in, keys_out, vals_in, vals_out, n, end_bit, s);
}

void exclusive_scan_i32(const std::int32_t* in, std::int32_t* out, std::int64_t n, hipStream_t s) {
    if (n <= 0) return;
    std::size_t bytes = 0;
    check(rocprim::exclusive_scan(nullptr, bytes, in, out, std::int32_t(0),
                                  static_cast<std::size_t>(n), rocprim::plus<std::int32_t>(), s),
          "scan (size)");
    DeviceBuffer<unsigned char> tmp(bytes ? bytes : 1, "scan scratch");
    check(rocprim::exclusive_scan(tmp.get(), bytes, in, out, std::int32_t(0),
                                  static_cast<std::size_t>(n), rocprim::plus<std::int32_t>(), s),
          "scan");
    sync(s, "scan");
}

// Integer maximum: a wavefront max by shuffles, then one atomicMax per wavefront (order-free).
__global__ void max_i32_kernel(const std::int32_t* __restrict__ in, std::int64_t n, int* __restrict__ out) {
    int m = INT_MIN;
    for (std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<std::int64_t>(gridDim.x) * blockDim.x)
        m = max(m, in[i]);
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0) atomicMax(out, m);
}

int max_i32(const std::int32_t* in, std::int64_t n, hipStream_t s) {
    if (n <= 0) return INT_MIN;
    DeviceBuffer<int> d(1, "max");
    const int lowest = INT_MIN;
    h2d(d.get(), &lowest, 1, s);
    const unsigned g = static_cast<unsigned>(std::min<std::int64_t>(grid_for(n), 1024));
    max_i32_kernel<<<g, kBlock, 0, s>>>(in, n, d.get());
    int h = 0;
    d2h(&h, d.get(), 1, s);
    sync(s, "max");
    return h;
}

__global__ void iota_kernel(std::int32_t* out, std::int64_t n) {
    const std::int64_t i = static_cast<std::int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = static_cast<std::int32_t>(i);
}

void iota_i32(std::int32_t* out, std::int64_t n, hipStream_t s) {
    if (const unsigned g = grid_for(n)) iota_kernel<<<g, kBlock, 0, s>>>(out, n);
}

}  // namespace cn
