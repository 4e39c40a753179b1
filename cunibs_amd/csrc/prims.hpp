// Setup-time sort / scan primitives (see prims.hip).
#pragma once
#include "common.hpp"

namespace cn {

void sort_pairs_u64(const std::uint64_t* keys_in, std::uint64_t* keys_out,
                    const std::int32_t* vals_in, std::int32_t* vals_out, std::int64_t n,
                    int end_bit, hipStream_t s);
// The same sort with the temporary storage in a caller-held, grow-only buffer and no stream sync.
void sort_pairs_u64_ws(const std::uint64_t* keys_in, std::uint64_t* keys_out, const std::int32_t* vals_in,
                       std::int32_t* vals_out, std::int64_t n, int end_bit, DeviceBuffer<unsigned char>& scratch,
                       hipStream_t s);
void sort_pairs_u32(const std::uint32_t* keys_in, std::uint32_t* keys_out,
                    const std::int32_t* vals_in, std::int32_t* vals_out, std::int64_t n,
                    int end_bit, hipStream_t s);
// out[i] = sum(in[0..i)); n elements.
void exclusive_scan_i32(const std::int32_t* in, std::int32_t* out, std::int64_t n, hipStream_t s);
void iota_i32(std::int32_t* out, std::int64_t n, hipStream_t s);
// max(in[0..n)) read back to the host (INT_MIN for n = 0); waits on s.
int max_i32(const std::int32_t* in, std::int64_t n, hipStream_t s);

}  // namespace cn
