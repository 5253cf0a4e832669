// Device-wide exclusive prefix sum of int32 counts (cell lists).
#include "common.h"
#include "scan.h"

namespace cube2 {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ inline int wave_inclusive_scan(int x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}

// Block-wide exclusive scan of one value per thread; returns the block total.
__device__ inline int block_exclusive_scan(int v, int* total) {
    __shared__ int wsum[kScanThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int inc = wave_inclusive_scan(v);
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    int off = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kScanThreads / 64; w++) {
        int s = wsum[w];
        if (w < wave) off += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

__global__ __launch_bounds__(kScanThreads) void k_block_sums(const int* __restrict__ d, int64_t n,
                                                             int* __restrict__ sums) {
    int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int s = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; q++)
        if (base + q < n) s += d[base + q];
    int tot;
    (void)block_exclusive_scan(s, &tot);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// Exclusive scan of each tile, offset by offs[blockIdx.x] (nullptr: 0);
// writes the grand total to d[n] from the last block.
__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(int* __restrict__ d, int64_t n,
                                                             const int* __restrict__ offs) {
    int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; q++) {
        v[q] = base + q < n ? d[base + q] : 0;
        s += v[q];
    }
    int tot;
    int pre = block_exclusive_scan(s, &tot);
    int o = offs ? offs[blockIdx.x] : 0;
    int run = o + pre;
#pragma unroll
    for (int q = 0; q < kScanItems; q++) {
        if (base + q < n) d[base + q] = run;
        run += v[q];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kScanThreads - 1) d[n] = o + tot;
}

static int scan_rec(int* d, int64_t n, int* tmp, hipStream_t s) {
    const int64_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb <= 1) {
        k_scan_tiles<<<1, kScanThreads, 0, s>>>(d, n, nullptr);
        C2_HIP(hipGetLastError());
        return CUBE2_OK;
    }
    int* sums = tmp;
    k_block_sums<<<(unsigned)nb, kScanThreads, 0, s>>>(d, n, sums);
    C2_HIP(hipGetLastError());
    C2_TRY(scan_rec(sums, nb, tmp + nb + 1, s));
    k_scan_tiles<<<(unsigned)nb, kScanThreads, 0, s>>>(d, n, sums);
    C2_HIP(hipGetLastError());
    return CUBE2_OK;
}

size_t scan_tmp_ints(int64_t n) {
    size_t t = 0;
    int64_t m = n;
    while (m > kScanTile) {
        m = (m + kScanTile - 1) / kScanTile;
        t += (size_t)m + 1;
    }
    return t + 16;
}

__global__ void k_fill_zero(uint32_t* __restrict__ d, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = 0u;
}

// Zeroes bytes (a multiple of 4) with a kernel: recordable into HIP graphs,
// where memset nodes are avoided.
int fill_zero(void* d, size_t bytes, hipStream_t s) {
    const size_t n = bytes / 4;
    if (!n) return CUBE2_OK;
    k_fill_zero<<<ceil_div(n, 256), 256, 0, s>>>(static_cast<uint32_t*>(d), n);
    C2_HIP(hipGetLastError());
    return CUBE2_OK;
}

int exclusive_scan(int* d, int64_t n, DevBuf& tmp, hipStream_t s) {
    if (n <= 0) return fill_zero(d, sizeof(int), s);
    C2_TRY(tmp.reserve(scan_tmp_ints(n) * sizeof(int)));
    return scan_rec(d, n, tmp.as<int>(), s);
}

}  // namespace cube2
