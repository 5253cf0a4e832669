#pragma once
#include "common.h"

namespace cube2 {
// Stable sort of (key, value) pairs of 32-bit words by the key's bits
// [0, end_bit) (sort.hip).  Full keys and values are carried to kout/vout;
// kin/vin are left unchanged.  tmp and scan_tmp are grow-only scratch.
int radix_sort_pairs_u32(const uint32_t* kin, uint32_t* kout, const uint32_t* vin,
                         uint32_t* vout, int64_t n, int end_bit, DevBuf& tmp, DevBuf& scan_tmp,
                         hipStream_t s);
}  // namespace cube2
