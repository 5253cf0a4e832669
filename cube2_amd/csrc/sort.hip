// Stable LSD radix sort of 32-bit (key, value) pairs: ceil(end_bit / 11)
// passes of equal digit widths (at most 11 bits, 2048 digits).
//
// A pass splits the n pairs into kSortBlocks contiguous segments, one per
// block.  k_sort_hist counts each segment's digits in LDS and stores the
// counts digit-major ([digit][block]), so that one exclusive scan of them
// gives every (digit, block) its first output position: all pairs of a
// smaller digit first, and within a digit the segments in order.  The
// scatter (k_sort_scatter) walks its segment 256 pairs at a time in input
// order; a pair's place is its (digit, block) base, the pairs of its digit in
// the earlier wavefronts of the step, and its rank among its own
// wavefront's lanes of that digit, found with one ballot per digit bit.
// Equal digits thus keep their input order (stability), which makes the
// passes compose into a sort by the whole key.  Only the first lane of each
// digit's group in a wavefront touches the LDS counts, so nearly sorted
// input (the drift's) does not serialise on one address, and a step costs
// the same whatever the digit width.
#include "sort.h"

#include <algorithm>

#include "scan.h"

namespace cube2 {

constexpr int kSortThreads = 256;
constexpr int kSortMaxBits = 11;
constexpr int kSortMaxBins = 1 << kSortMaxBits;
constexpr int kSortBlocks = 1024;

__device__ inline void sort_segment(int n, int& beg, int& end) {
    const int seg = (n + kSortBlocks - 1) / kSortBlocks;
    beg = min(n, (int)blockIdx.x * seg);
    end = min(n, beg + seg);
}

// Lanes of the calling wavefront that are valid and hold digit d (bits
// below `bits`).
__device__ inline uint64_t digit_peers(bool valid, int d, int bits) {
    uint64_t peers = __ballot(valid);
    for (int q = 0; q < bits; q++) {
        const uint64_t m = __ballot(valid && ((d >> q) & 1));
        peers &= ((d >> q) & 1) ? m : ~m;
    }
    return peers;
}

__global__ __launch_bounds__(kSortThreads) void k_sort_hist(const uint32_t* __restrict__ keys,
                                                         int n, int shift, int bits,
                                                         int* __restrict__ hist) {
    __shared__ int h[kSortMaxBins];
    int beg, end;
    sort_segment(n, beg, end);
    const int t = threadIdx.x, lane = t & 63, nbin = 1 << bits;
    const uint32_t mask = (uint32_t)nbin - 1u;
    for (int q = t; q < nbin; q += kSortThreads) h[q] = 0;
    __syncthreads();
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (int t0 = beg; t0 < end; t0 += kSortThreads) {  // block-uniform trip count
        const int i = t0 + t;
        const bool valid = i < end;
        const int d = valid ? (int)((keys[i] >> shift) & mask) : 0;
        const uint64_t peers = digit_peers(valid, d, bits);
        if (valid && __popcll(peers & lt) == 0) atomicAdd(&h[d], __popcll(peers));
    }
    __syncthreads();
    for (int q = t; q < nbin; q += kSortThreads) hist[q * kSortBlocks + blockIdx.x] = h[q];
}

__global__ __launch_bounds__(kSortThreads) void k_sort_scatter(const uint32_t* __restrict__ kin,
                                                            const uint32_t* __restrict__ vin,
                                                            int n, int shift, int bits,
                                                            const int* __restrict__ hist,
                                                            uint32_t* __restrict__ kout,
                                                            uint32_t* __restrict__ vout) {
    __shared__ int base[kSortMaxBins];
    __shared__ int wc[4][kSortMaxBins];
    int beg, end;
    sort_segment(n, beg, end);
    const int t = threadIdx.x, wave = t >> 6, lane = t & 63, nbin = 1 << bits;
    const uint32_t mask = (uint32_t)nbin - 1u;
    for (int q = t; q < nbin; q += kSortThreads) {
        base[q] = hist[q * kSortBlocks + blockIdx.x];
        wc[0][q] = wc[1][q] = wc[2][q] = wc[3][q] = 0;
    }
    __syncthreads();
    const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (int t0 = beg; t0 < end; t0 += kSortThreads) {  // block-uniform trip count
        const int i = t0 + t;
        const bool valid = i < end;
        const uint32_t k = valid ? kin[i] : 0u;
        const int d = (int)((k >> shift) & mask);
        const uint64_t peers = digit_peers(valid, d, bits);
        const int rank = __popcll(peers & lt), cnt = __popcll(peers);
        const bool leader = valid && rank == 0;
        if (leader) wc[wave][d] = cnt;  // the other wavefronts' entries for d are 0
        __syncthreads();
        if (valid) {
            int pos = base[d] + rank;
            for (int w = 0; w < wave; w++) pos += wc[w][d];
            kout[pos] = k;
            vout[pos] = vin[i];
        }
        __syncthreads();
        if (leader) {
            atomicAdd(&base[d], cnt);
            wc[wave][d] = 0;
        }
        __syncthreads();
    }
}

int radix_sort_pairs_u32(const uint32_t* kin, uint32_t* kout, const uint32_t* vin,
                         uint32_t* vout, int64_t n, int end_bit, DevBuf& tmp, DevBuf& scan_tmp,
                         hipStream_t s) {
    if (n < 0 || n >= ((int64_t)1 << 31) || end_bit < 0 || end_bit > 32) {
        set_error("radix sort: %lld pairs, end_bit %d out of range", (long long)n, end_bit);
        return CUBE2_ERR_ARG;
    }
    if (n == 0) return CUBE2_OK;
    if (end_bit == 0) {  // every key equal: the input order
        C2_HIP(hipMemcpyAsync(kout, kin, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
        C2_HIP(hipMemcpyAsync(vout, vin, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
        return CUBE2_OK;
    }
    const int passes = std::max(1, (end_bit + kSortMaxBits - 1) / kSortMaxBits);
    const int width = std::max(1, (end_bit + passes - 1) / passes);
    const size_t nh = (size_t)kSortMaxBins * kSortBlocks + 1;
    const size_t pair_bytes = ((size_t)n * 4 + 255) / 256 * 256;
    C2_TRY(tmp.reserve(2 * pair_bytes + nh * sizeof(int)));
    uint32_t* kt = tmp.as<uint32_t>();
    uint32_t* vt = reinterpret_cast<uint32_t*>(static_cast<char*>(tmp.ptr) + pair_bytes);
    int* hist = reinterpret_cast<int*>(static_cast<char*>(tmp.ptr) + 2 * pair_bytes);
    // the last pass writes kout/vout: an odd pass count starts there
    const uint32_t* ks = kin;
    const uint32_t* vs = vin;
    for (int p = 0; p < passes; p++) {
        const bool to_out = ((passes - 1 - p) & 1) == 0;
        uint32_t* kd = to_out ? kout : kt;
        uint32_t* vd = to_out ? vout : vt;
        // digit bits [p width, min((p + 1) width, end_bit)): bits at or
        // above end_bit are ignored
        const int shift = p * width;
        const int bits = std::max(1, std::min(width, end_bit - shift));
        k_sort_hist<<<kSortBlocks, kSortThreads, 0, s>>>(ks, (int)n, shift, bits, hist);
        C2_HIP(hipGetLastError());
        C2_TRY(exclusive_scan(hist, (int64_t)(1 << bits) * kSortBlocks, scan_tmp, s));
        k_sort_scatter<<<kSortBlocks, kSortThreads, 0, s>>>(ks, vs, (int)n, shift, bits, hist, kd,
                                                            vd);
        C2_HIP(hipGetLastError());
        ks = kd;
        vs = vd;
    }
    return CUBE2_OK;
}

}  // namespace cube2
