#pragma once
#include "common.h"

namespace cube2 {
// In-place exclusive scan of d[0..n); d must have n+1 entries and d[n]
// receives the total.  tmp is a grow-only scratch buffer.
int exclusive_scan(int* d, int64_t n, DevBuf& tmp, hipStream_t s);
size_t scan_tmp_ints(int64_t n);
}  // namespace cube2
